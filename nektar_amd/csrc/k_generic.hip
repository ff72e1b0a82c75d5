// Runtime-size operators: any modes and points per direction (anisotropic
// orders, orders and point counts without a compile-time kernel).  This is
// the path the reference takes when its compile-time switch finds no
// instantiation (MatrixFreeOps/SwitchNodesPoints.h:302-343 -> the
// runtime-size operator3D of Helmholtz.h:374-464, BwdTrans.h:364-494,
// IProduct.h:410-565, PhysDeriv.h:478-573, IProductWRTDerivBase.h and
// LinearAdvectionDiffusionReaction.h).
//
// One workgroup per element.  Every stage is a set of independent outputs,
// one dot product each, spread over the workgroup's work-items; the
// element's intermediate tensors live in four buffers of `buf` values, in
// LDS when they fit one workgroup's 64 KiB, otherwise in a per-workgroup
// slice of a global workspace (grid of at most GenArgs::groups workgroups
// walking the elements).
//
// Arithmetic (the same as the compile-time kernels): the transform to the
// quadrature points, collocation derivatives, the pointwise metric /
// Jacobian / weights, then IProduct w.r.t. dbdata_d = D_d bdata_d written
// as IProduct(bdata, D_d^T v) so that the four inner products of the
// reference share one transpose transform.  Tets use the collapsed
// coordinates with the singular vertex/edge corrections of
// BwdTransKernels.hpp:296-324 and IProductKernels.hpp:456-513, evaluated
// from the p = 0 / p = 1 partial sums.
#include <hip/hip_runtime.h>

#include "generic_kernels.hpp"

namespace nektar_amd {
namespace gen {

constexpr int kThreads = 256;

// Intermediates in LDS: one element per workgroup.
template <typename T, int SH>
__global__ __launch_bounds__(kThreads) void lds_kernel(GenArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    T *X = reinterpret_cast<T *>(smem);
    element<T, SH>(a, blockIdx.x, X, X + a.buf, X + 2 * a.buf, X + 3 * a.buf);
}

// Intermediates in the global workspace: a.groups workgroups walk the
// elements, each with its own four buffers.
template <typename T, int SH>
__global__ __launch_bounds__(kThreads) void global_kernel(GenArgs a)
{
    T *X = static_cast<T *>(a.ws) + (size_t)blockIdx.x * 4 * a.buf;
    for (int e = blockIdx.x; e < a.nelmt; e += gridDim.x) {
        element<T, SH>(a, e, X, X + a.buf, X + 2 * a.buf, X + 3 * a.buf);
        __syncthreads();  // the buffers are reused by the next element
    }
}

template <typename T, int SH>
hipError_t launch_t(const GenArgs &a)
{
    if (a.nelmt == 0) return hipSuccess;
    const size_t lds = lds_bytes(a);
    if (lds <= kMaxLds) {
        hipLaunchKernelGGL((lds_kernel<T, SH>), dim3(a.nelmt), dim3(kThreads), lds, a.stream, a);
    } else {
        if (!a.ws) return hipErrorInvalidValue;
        hipLaunchKernelGGL((global_kernel<T, SH>), dim3(a.groups), dim3(kThreads), 0, a.stream, a);
    }
    return hipGetLastError();
}

}  // namespace gen

hipError_t launch_generic(const GenArgs &a)
{
    if (a.shape == Shape::Hex)
        return a.fp32 ? gen::launch_t<float, gen::SH_HEX>(a) : gen::launch_t<double, gen::SH_HEX>(a);
    if (a.shape == Shape::Prism)
        return a.fp32 ? gen::launch_t<float, gen::SH_PRISM>(a) : gen::launch_t<double, gen::SH_PRISM>(a);
    if (a.shape == Shape::Pyr)
        return a.fp32 ? gen::launch_t<float, gen::SH_PYR>(a) : gen::launch_t<double, gen::SH_PYR>(a);
    return a.fp32 ? gen::launch_t<float, gen::SH_TET>(a) : gen::launch_t<double, gen::SH_TET>(a);
}

}  // namespace nektar_amd
