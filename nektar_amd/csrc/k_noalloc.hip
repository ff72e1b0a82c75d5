// PhysDeriv operator()(dir, in, out) for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops that forbid
// hipMalloc): the two derivative directions not asked for go to a
// file-scope device array instead of a workspace the library allocates.
//
// The array is a module global: every GPU the library runs on holds its own
// copy, for the rest of the process.  Host code never passes its name to a
// kernel (on the host that name is a stand-in, not a device address): the
// device address is looked up once per device with hipGetSymbolAddress,
// with that device current, and cached.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "common.hpp"
#include "handles.hpp"

namespace nektar_amd {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per device

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's copy of `scratch` (looked up on the
// first call with the device current, then cached).
void *scratch_on_current_device()
{
    int dev = -1;
    NMF_HIP(hipGetDevice(&dev));
    NMF_REQUIRE(dev >= 0 && dev < kMaxDevices, NMF_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        NMF_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        NMF_REQUIRE(ptr, NMF_ERR_DEVICE, "hipGetSymbolAddress returned a null address for the scratch array");
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

}  // namespace
}  // namespace nektar_amd

using namespace nektar_amd;

extern "C" int nmf_operator_apply_physderiv_dir_noalloc(nmf_operator op, int dir, const void *in, void *out,
                                                        void *stream)
{
    return guard([&] {
        NMF_REQUIRE(op && op->impl, NMF_ERR_INVALID, "null operator handle");
        MFOperator &o = *op->impl;
        NMF_REQUIRE(o.op == OpType::PhysDeriv, NMF_ERR_INVALID, "not a PhysDeriv operator");
        NMF_REQUIRE(dir >= 0 && dir < 3, NMF_ERR_INVALID, "PhysDeriv direction must be 0, 1 or 2");
        const size_t need = o.wsp_size() * o.value_bytes();
        NMF_REQUIRE(need <= kScratchBytes, NMF_ERR_UNSUPPORTED,
                    "PhysDeriv workspace of " + std::to_string(need) + " bytes (2 * nelmt * nqtot values) exceeds "
                        "the " + std::to_string(kScratchBytes) + "-byte static scratch of the no-allocation form");
        // the first apply packs and uploads the geometry (an allocation);
        // this form only runs once that has happened
        NMF_REQUIRE(o.geometry_current(), NMF_ERR_STATE,
                    "geometry not yet on the device: SetDF was called after the last apply; run one "
                    "nmf_operator_apply_physderiv first");
        if (o.nelmt == 0) return;
        o.apply_physderiv_dir(dir, in, out, scratch_on_current_device(), static_cast<hipStream_t>(stream));
    });
}
