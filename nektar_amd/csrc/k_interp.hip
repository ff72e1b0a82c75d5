// PhysInterp1DScaled on gfx950: point values of an element at the
// quadrature points -> values at the points of the same families, scaled
// in number (PhysInterp1DScaled.h:360-430, PhysInterp1DScaledKernels.hpp
// :151-217).  The numbers of points come from a run-time scaling factor,
// so the kernel takes its sizes at run time (inputs up to kMaxIn points
// per direction).  One element per 256-thread workgroup, one work-item per
// 1-D line of each stage:
//   stage 0: direction 0, lines (q,r) of the input (staged in A) -> B [r][q][i]
//   stage 1: direction 1, lines (i,r) of B -> A [r][j][i]
//   stage 2: direction 2, lines (i,j) of A -> global output [k][j][i]
// A work-item reads its line once into registers and forms every output of
// the line from it; the interpolation matrices (stored transposed, one row
// per output) are indexed by the loop counters only, so each row is a
// wave-uniform scalar load, issued one output ahead, and takes no LDS
// bandwidth.  Summation order as the reference: first term a product,
// then FMAs in increasing input index.  The operation is HBM-bound (it
// writes factor^3 times what it reads).
#include <hip/hip_runtime.h>

#include "launch.hpp"

namespace nektar_amd {
namespace {

constexpr int kInterpThreads = 256;
constexpr int kMaxIn = 12;  // input points per direction (nummodes <= 10, Q <= nummodes + 2)
// The kernel is instantiated for at most MI = 4, 8 and 12 input points per
// direction and launched with the smallest that holds the operator's, so
// that the register rows and the guarded FMA chains are no longer than
// needed.

// One row of a transposed interpolation matrix (the weights of one
// output), wave-uniform: kMaxIn unconditional loads that merge into wide
// scalar loads (the table buffer carries kMaxIn values of padding, and
// only the first n weights are used).
template <typename T, int MI>
__device__ __forceinline__ void table_row(const T *__restrict__ row, T (&w)[MI])
{
#pragma unroll
    for (int p = 0; p < MI; ++p) w[p] = row[p];
}

// s = sum_p v[p] w[p], first term a product (the reference's order)
template <typename T, int MI>
__device__ __forceinline__ T dot_row(const T (&v)[MI], const T (&w)[MI], int n)
{
    T s = v[0] * w[0];
#pragma unroll
    for (int p = 1; p < MI; ++p)
        if (p < n) s = fma(v[p], w[p], s);
    return s;
}

template <typename T, int MI>
__global__ __launch_bounds__(kInterpThreads) void phys_interp_kernel(int ni0, int ni1, int ni2, int no0, int no1,
                                                                     int no2, int sizeA, const T *__restrict__ tab,
                                                                     const T *__restrict__ in, T *__restrict__ out)
{
    extern __shared__ __align__(16) unsigned char smem[];
    T *A = reinterpret_cast<T *>(smem);
    T *B = A + sizeA;
    // transposed tables: ITd[i*ni_d + p] = weight of input p for output i
    const T *__restrict__ IT0 = tab;
    const T *__restrict__ IT1 = IT0 + ni0 * no0;
    const T *__restrict__ IT2 = IT1 + ni1 * no1;
    const int tid = threadIdx.x;
    const int nin = ni0 * ni1 * ni2, nout = no0 * no1 * no2;
    const T *xin = in + (size_t)blockIdx.x * nin;
    for (int x = tid; x < nin; x += kInterpThreads) A[x] = xin[x];
    __syncthreads();
    T v[MI], w[MI], wn[MI];
    // direction 0: B[(r*ni1 + q)*no0 + i] = sum_p A[(r*ni1 + q)*ni0 + p] I0(p, i)
    for (int l = tid; l < ni1 * ni2; l += kInterpThreads) {
#pragma unroll
        for (int p = 0; p < MI; ++p) v[p] = p < ni0 ? A[l * ni0 + p] : T(0);
        table_row(IT0, w);
        for (int i = 0; i < no0; ++i) {
            table_row(IT0 + min(i + 1, no0 - 1) * ni0, wn);  // next row in flight
            B[l * no0 + i] = dot_row(v, w, ni0);
#pragma unroll
            for (int p = 0; p < MI; ++p) w[p] = wn[p];
        }
    }
    __syncthreads();
    // direction 1: A[(r*no1 + j)*no0 + i] = sum_q B[(r*ni1 + q)*no0 + i] I1(q, j)
    for (int l = tid; l < no0 * ni2; l += kInterpThreads) {
        const int i = l % no0, r = l / no0;
        const T *b = B + r * ni1 * no0 + i;
#pragma unroll
        for (int q = 0; q < MI; ++q) v[q] = q < ni1 ? b[q * no0] : T(0);
        T *a = A + r * no1 * no0 + i;
        table_row(IT1, w);
        for (int j = 0; j < no1; ++j) {
            table_row(IT1 + min(j + 1, no1 - 1) * ni1, wn);
            a[j * no0] = dot_row(v, w, ni1);
#pragma unroll
            for (int q = 0; q < MI; ++q) w[q] = wn[q];
        }
    }
    __syncthreads();
    // direction 2: out[(k*no1 + j)*no0 + i] = sum_r A[(r*no1 + j)*no0 + i] I2(r, k)
    T *y = out + (size_t)blockIdx.x * nout;
    const int plane = no0 * no1;
    for (int l = tid; l < plane; l += kInterpThreads) {
#pragma unroll
        for (int r = 0; r < MI; ++r) v[r] = r < ni2 ? A[r * plane + l] : T(0);
        table_row(IT2, w);
        for (int k = 0; k < no2; ++k) {
            table_row(IT2 + min(k + 1, no2 - 1) * ni2, wn);
            y[k * plane + l] = dot_row(v, w, ni2);
#pragma unroll
            for (int r = 0; r < MI; ++r) w[r] = wn[r];
        }
    }
}

int interp_size_a(const int ni[3], const int no[3])
{
    const int a = ni[0] * ni[1] * ni[2], b = no[0] * no[1] * ni[2];
    return a > b ? a : b;
}

}  // namespace

size_t interp_lds_bytes(const int ni[3], const int no[3], bool fp32)
{
    const size_t n = (size_t)interp_size_a(ni, no) + (size_t)no[0] * ni[1] * ni[2];
    return n * (fp32 ? sizeof(float) : sizeof(double));
}

namespace {

template <typename T, int MI>
hipError_t launch_mi(const InterpArgs &a, size_t lds)
{
    const int sizeA = interp_size_a(a.ni, a.no);
    const dim3 grid(a.nelmt), block(kInterpThreads);
    hipLaunchKernelGGL((phys_interp_kernel<T, MI>), grid, block, lds, a.stream, a.ni[0], a.ni[1], a.ni[2], a.no[0],
                       a.no[1], a.no[2], sizeA, static_cast<const T *>(a.tab), static_cast<const T *>(a.in),
                       static_cast<T *>(a.out));
    return hipGetLastError();
}

template <typename T>
hipError_t launch_t(const InterpArgs &a, size_t lds)
{
    const int m = a.ni[0] > a.ni[1] ? (a.ni[0] > a.ni[2] ? a.ni[0] : a.ni[2]) : (a.ni[1] > a.ni[2] ? a.ni[1] : a.ni[2]);
    if (m <= 4) return launch_mi<T, 4>(a, lds);
    if (m <= 8) return launch_mi<T, 8>(a, lds);
    return launch_mi<T, kMaxIn>(a, lds);
}

}  // namespace

hipError_t launch_phys_interp(const InterpArgs &a)
{
    if (a.nelmt <= 0) return hipSuccess;
    const size_t lds = interp_lds_bytes(a.ni, a.no, a.fp32);
    if (lds > kInterpMaxLds || a.ni[0] > kMaxIn || a.ni[1] > kMaxIn || a.ni[2] > kMaxIn) return hipErrorInvalidValue;
    return a.fp32 ? launch_t<float>(a, lds) : launch_t<double>(a, lds);
}

}  // namespace nektar_amd
