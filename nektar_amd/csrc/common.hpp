// Shared host-side utilities: error propagation across the C ABI and RAII
// device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/nektar_mf.h"

namespace nektar_amd {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);
void clear_last_error();

#define NMF_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t nmf_e_ = (call);                                                            \
        if (nmf_e_ != hipSuccess)                                                              \
            throw ::nektar_amd::Error(NMF_ERR_DEVICE, std::string(#call) + ": " +              \
                                                          hipGetErrorString(nmf_e_));          \
    } while (0)

#define NMF_REQUIRE(cond, code, msg)                                                           \
    do {                                                                                       \
        if (!(cond)) throw ::nektar_amd::Error((code), (msg));                                 \
    } while (0)

template <typename F>
int guard(F &&f)
{
    try {
        clear_last_error();
        f();
        return NMF_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &e) {
        set_last_error(std::string("host allocation failed: ") + e.what());
        return NMF_ERR_MEMORY;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return NMF_ERR_INVALID;
    }
}

// Owning device allocation.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes) { allocate(bytes); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    DeviceBuffer(DeviceBuffer &&o) noexcept : p_(o.p_), bytes_(o.bytes_) { o.p_ = nullptr; o.bytes_ = 0; }
    DeviceBuffer &operator=(DeviceBuffer &&o) noexcept
    {
        if (this != &o) {
            release();
            p_ = o.p_;
            bytes_ = o.bytes_;
            o.p_ = nullptr;
            o.bytes_ = 0;
        }
        return *this;
    }
    ~DeviceBuffer() { release(); }

    void allocate(size_t bytes)
    {
        if (bytes == bytes_ && p_) return;
        release();
        if (bytes == 0) return;
        hipError_t e = hipMalloc(&p_, bytes);
        if (e != hipSuccess) {
            p_ = nullptr;
            throw Error(NMF_ERR_MEMORY, "hipMalloc(" + std::to_string(bytes) + " bytes): " + hipGetErrorString(e));
        }
        bytes_ = bytes;
    }
    void release()
    {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        bytes_ = 0;
    }
    template <typename T>
    T *as() const { return static_cast<T *>(p_); }
    void *get() const { return p_; }
    size_t bytes() const { return bytes_; }

    // Upload host doubles, converting to float when `fp32`.
    void upload(const std::vector<double> &h, bool fp32)
    {
        if (fp32) {
            std::vector<float> f(h.begin(), h.end());
            allocate(f.size() * sizeof(float));
            if (!f.empty()) NMF_HIP(hipMemcpy(p_, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
        } else {
            allocate(h.size() * sizeof(double));
            if (!h.empty()) NMF_HIP(hipMemcpy(p_, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
        }
    }

private:
    void *p_ = nullptr;
    size_t bytes_ = 0;
};

}  // namespace nektar_amd
