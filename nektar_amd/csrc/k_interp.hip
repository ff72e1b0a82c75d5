// PhysInterp1DScaled on gfx950: point values of an element at the
// quadrature points -> values at the points of the same families, scaled
// in number (PhysInterp1DScaled.h:360-430, PhysInterp1DScaledKernels.hpp
// :151-217).  The numbers of points come from a run-time scaling factor,
// so the kernel takes its sizes at run time.  One element per 256-thread
// workgroup; the three interpolation matrices and two ping-pong buffers
// live in dynamic LDS:
//   stage 0: direction 0, A (input, [r][q][p]) -> B [r][q][i]
//   stage 1: direction 1, B -> A [r][j][i]
//   stage 2: direction 2, A -> global output [k][j][i]
// Work-items take the output index with i fastest, so table reads are
// consecutive and the global stores coalesce.  Summation order as the
// reference: first term a product, then FMAs in increasing input index.
// The operation is HBM-bound (it writes (factor)^3 times what it reads).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "launch.hpp"

namespace nektar_amd {
namespace {

constexpr int kInterpThreads = 256;

template <typename T>
__global__ __launch_bounds__(kInterpThreads) void phys_interp_kernel(int ni0, int ni1, int ni2, int no0, int no1,
                                                                     int no2, int sizeA, const T *__restrict__ tab,
                                                                     const T *__restrict__ in, T *__restrict__ out)
{
    extern __shared__ __align__(16) unsigned char smem[];
    T *I0 = reinterpret_cast<T *>(smem);
    T *I1 = I0 + ni0 * no0;
    T *I2 = I1 + ni1 * no1;
    T *A = I2 + ni2 * no2;
    T *B = A + sizeA;
    const int tid = threadIdx.x;
    const int nin = ni0 * ni1 * ni2, nout = no0 * no1 * no2;
    const int ntab = ni0 * no0 + ni1 * no1 + ni2 * no2;
    for (int x = tid; x < ntab; x += kInterpThreads) I0[x] = tab[x];
    const T *xin = in + (size_t)blockIdx.x * nin;
    for (int x = tid; x < nin; x += kInterpThreads) A[x] = xin[x];
    __syncthreads();
    // direction 0: B[(r*ni1 + q)*no0 + i] = sum_p A[(r*ni1 + q)*ni0 + p] I0[p*no0 + i]
    for (int x = tid; x < no0 * ni1 * ni2; x += kInterpThreads) {
        const int i = x % no0, rq = x / no0;
        const T *a = A + rq * ni0;
        T s = a[0] * I0[i];
        for (int p = 1; p < ni0; ++p) s = fma(a[p], I0[p * no0 + i], s);
        B[x] = s;
    }
    __syncthreads();
    // direction 1: A[(r*no1 + j)*no0 + i] = sum_q B[(r*ni1 + q)*no0 + i] I1[q*no1 + j]
    for (int x = tid; x < no0 * no1 * ni2; x += kInterpThreads) {
        const int i = x % no0, j = (x / no0) % no1, r = x / (no0 * no1);
        const T *b = B + r * ni1 * no0 + i;
        T s = b[0] * I1[j];
        for (int q = 1; q < ni1; ++q) s = fma(b[q * no0], I1[q * no1 + j], s);
        A[x] = s;
    }
    __syncthreads();
    // direction 2: out[(k*no1 + j)*no0 + i] = sum_r A[(r*no1 + j)*no0 + i] I2[r*no2 + k]
    T *y = out + (size_t)blockIdx.x * nout;
    const int plane = no0 * no1;
    for (int x = tid; x < nout; x += kInterpThreads) {
        const int ji = x % plane, k = x / plane;
        T s = A[ji] * I2[k];
        for (int r = 1; r < ni2; ++r) s = fma(A[r * plane + ji], I2[r * no2 + k], s);
        y[x] = s;
    }
}

int interp_size_a(const int ni[3], const int no[3])
{
    return std::max(ni[0] * ni[1] * ni[2], no[0] * no[1] * ni[2]);
}

}  // namespace

size_t interp_lds_bytes(const int ni[3], const int no[3], bool fp32)
{
    const size_t n = (size_t)ni[0] * no[0] + (size_t)ni[1] * no[1] + (size_t)ni[2] * no[2] + interp_size_a(ni, no) +
                     (size_t)no[0] * ni[1] * ni[2];
    return n * (fp32 ? sizeof(float) : sizeof(double));
}

hipError_t launch_phys_interp(const InterpArgs &a)
{
    if (a.nelmt <= 0) return hipSuccess;
    const size_t lds = interp_lds_bytes(a.ni, a.no, a.fp32);
    if (lds > kInterpMaxLds) return hipErrorInvalidValue;
    const int sizeA = interp_size_a(a.ni, a.no);
    const dim3 grid(a.nelmt), block(kInterpThreads);
    if (a.fp32)
        hipLaunchKernelGGL(phys_interp_kernel<float>, grid, block, lds, a.stream, a.ni[0], a.ni[1], a.ni[2], a.no[0],
                           a.no[1], a.no[2], sizeA, static_cast<const float *>(a.tab),
                           static_cast<const float *>(a.in), static_cast<float *>(a.out));
    else
        hipLaunchKernelGGL(phys_interp_kernel<double>, grid, block, lds, a.stream, a.ni[0], a.ni[1], a.ni[2],
                           a.no[0], a.no[1], a.no[2], sizeA, static_cast<const double *>(a.tab),
                           static_cast<const double *>(a.in), static_cast<double *>(a.out));
    return hipGetLastError();
}

}  // namespace nektar_amd
