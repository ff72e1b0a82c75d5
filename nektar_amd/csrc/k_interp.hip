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
//
// The sizes of the default factor 1.5 (hex and tet, Q = 3..12) and of
// factor 2 (Q = 3..11) also have a compile-time-size kernel,
// phys_interp_fixed: one element per 128-thread
// workgroup, stage 0 reading its input line straight from global memory
// into registers, every loop unrolled, the same layouts and the same
// summation order (so its results are bitwise those of the run-time
// kernel).  At hex Q = 8 -> 12 over 32768 elements it takes 105 us against
// 250 us for the run-time kernel (5.6 against 2.3 TB/s); two or four
// elements per workgroup, stage 1 writing over stage 0's buffer, or stages
// 1 and 2 fused per output column all measured 2-12% slower.
//
// Where an element's output is a whole number of 128-byte lines (hex
// Q = 8 -> 12 in FP64 and FP32, among others) the compile-time kernel
// writes it with non-temporal stores: 88 against 105 us at hex FP64
// (6.7 TB/s), 49 against 58 us in FP32.  Where it is not (the tet sizes:
// 11,616 bytes per element at Q = 8), partial lines reach memory unmerged
// and the same stores measured slower (tet FP64 124 against 97 us), so
// those keep ordinary stores, as does the run-time kernel.  Non-temporal
// loads of the input measured slower as well (119-125 against 105 us).
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

constexpr int kFixedThreads = 128;

// s = sum_p v[p] w[p], w a row of a table at a compile-time offset
template <typename T, int N>
__device__ __forceinline__ T dot_fixed(const T (&v)[N], const T *__restrict__ w)
{
    T s = v[0] * w[0];
#pragma unroll
    for (int p = 1; p < N; ++p) s = fma(v[p], w[p], s);
    return s;
}

// I0..I2 input and O0..O2 output points per direction; the stages, layouts
// and tables of phys_interp_kernel with B = [r][q][i] (O0 I1 I2 values) and
// A = [r][j][i] (O0 O1 I2 values) in static LDS
template <typename T, int I0, int I1, int I2, int O0, int O1, int O2>
__global__ __launch_bounds__(kFixedThreads) void phys_interp_fixed(const T *__restrict__ tab, const T *__restrict__ in,
                                                                   T *__restrict__ out)
{
    constexpr int NIN = I0 * I1 * I2, NOUT = O0 * O1 * O2;
    constexpr int L0 = I1 * I2, L1 = O0 * I2, L2 = O0 * O1;
    // non-temporal output stores where an element's output is whole 128-byte lines
    constexpr bool kStream = (NOUT * sizeof(T)) % 128 == 0;
    __shared__ T B[O0 * I1 * I2];
    __shared__ T A[O0 * O1 * I2];
    const T *__restrict__ IT0 = tab;
    const T *__restrict__ IT1 = IT0 + O0 * I0;
    const T *__restrict__ IT2 = IT1 + O1 * I1;
    const int tid = threadIdx.x;
    // direction 0: B[l*O0 + i] = sum_p in[l*I0 + p] I0(p, i), l = r*I1 + q
    const T *xin = in + (size_t)blockIdx.x * NIN;
#pragma unroll
    for (int u0 = 0; u0 < L0; u0 += kFixedThreads) {
        const int l = u0 + tid;
        if (l < L0) {
            T v[I0];
#pragma unroll
            for (int p = 0; p < I0; ++p) v[p] = xin[l * I0 + p];
#pragma unroll
            for (int i = 0; i < O0; ++i) B[l * O0 + i] = dot_fixed(v, IT0 + i * I0);
        }
    }
    __syncthreads();
    // direction 1: A[(r*O1 + j)*O0 + i] = sum_q B[(r*I1 + q)*O0 + i] I1(q, j)
#pragma unroll
    for (int u0 = 0; u0 < L1; u0 += kFixedThreads) {
        const int l = u0 + tid;
        if (l < L1) {
            const int i = l % O0, r = l / O0;
            T v[I1];
#pragma unroll
            for (int q = 0; q < I1; ++q) v[q] = B[(r * I1 + q) * O0 + i];
#pragma unroll
            for (int j = 0; j < O1; ++j) A[(r * O1 + j) * O0 + i] = dot_fixed(v, IT1 + j * I1);
        }
    }
    __syncthreads();
    // direction 2: out[k*L2 + l] = sum_r A[r*L2 + l] I2(r, k)
    T *y = out + (size_t)blockIdx.x * NOUT;
#pragma unroll
    for (int u0 = 0; u0 < L2; u0 += kFixedThreads) {
        const int l = u0 + tid;
        if (l < L2) {
            T v[I2];
#pragma unroll
            for (int r = 0; r < I2; ++r) v[r] = A[r * L2 + l];
#pragma unroll
            for (int k = 0; k < O2; ++k) {
                if constexpr (kStream)
                    __builtin_nontemporal_store(dot_fixed(v, IT2 + k * I2), y + k * L2 + l);
                else
                    y[k * L2 + l] = dot_fixed(v, IT2 + k * I2);
            }
        }
    }
}

template <typename T, int I0, int I1, int I2, int O0, int O1, int O2>
bool try_fixed(const InterpArgs &a, hipError_t &e)
{
    if (a.ni[0] != I0 || a.ni[1] != I1 || a.ni[2] != I2 || a.no[0] != O0 || a.no[1] != O1 || a.no[2] != O2)
        return false;
    hipLaunchKernelGGL((phys_interp_fixed<T, I0, I1, I2, O0, O1, O2>), dim3(a.nelmt), dim3(kFixedThreads), 0, a.stream,
                       static_cast<const T *>(a.tab), static_cast<const T *>(a.in), static_cast<T *>(a.out));
    e = hipGetLastError();
    return true;
}

// the sizes of factor 1.5 (Operator.hpp:568-575): hex Q^3 -> O^3 and tet
// (Q, Q-1, Q-1) -> (O, O-1, O-1), O = floor(1.5 Q)
#define NMF_INTERP_HEX(Q, O) try_fixed<T, Q, Q, Q, O, O, O>(a, e)
#define NMF_INTERP_TET(Q, O) try_fixed<T, Q, Q - 1, Q - 1, O, O - 1, O - 1>(a, e)
template <typename T>
bool launch_fixed(const InterpArgs &a, hipError_t &e)
{
    return NMF_INTERP_HEX(3, 4) || NMF_INTERP_HEX(4, 6) || NMF_INTERP_HEX(5, 7) || NMF_INTERP_HEX(6, 9) ||
           NMF_INTERP_HEX(7, 10) || NMF_INTERP_HEX(8, 12) || NMF_INTERP_HEX(9, 13) || NMF_INTERP_HEX(10, 15) ||
           NMF_INTERP_HEX(11, 16) || NMF_INTERP_HEX(12, 18) || NMF_INTERP_TET(3, 4) || NMF_INTERP_TET(4, 6) ||
           NMF_INTERP_TET(5, 7) || NMF_INTERP_TET(6, 9) || NMF_INTERP_TET(7, 10) || NMF_INTERP_TET(8, 12) ||
           NMF_INTERP_TET(9, 13) || NMF_INTERP_TET(10, 15) || NMF_INTERP_TET(11, 16) || NMF_INTERP_TET(12, 18);
}
// factor 2: hex Q^3 -> (2Q)^3, tet (Q, Q-1, Q-1) -> (2Q, 2Q-1, 2Q-1), for
// Q <= 11 (Q = 12 exceeds the 64 KiB staging and is rejected on the host)
template <typename T>
bool launch_fixed_f2(const InterpArgs &a, hipError_t &e)
{
    return NMF_INTERP_HEX(3, 6) || NMF_INTERP_HEX(4, 8) || NMF_INTERP_HEX(5, 10) || NMF_INTERP_HEX(6, 12) ||
           NMF_INTERP_HEX(7, 14) || NMF_INTERP_HEX(8, 16) || NMF_INTERP_HEX(9, 18) || NMF_INTERP_HEX(10, 20) ||
           NMF_INTERP_HEX(11, 22) || NMF_INTERP_TET(3, 6) || NMF_INTERP_TET(4, 8) || NMF_INTERP_TET(5, 10) ||
           NMF_INTERP_TET(6, 12) || NMF_INTERP_TET(7, 14) || NMF_INTERP_TET(8, 16) || NMF_INTERP_TET(9, 18) ||
           NMF_INTERP_TET(10, 20) || NMF_INTERP_TET(11, 22);
}
#undef NMF_INTERP_HEX
#undef NMF_INTERP_TET

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
    hipError_t e = hipSuccess;
    if (launch_fixed<T>(a, e) || launch_fixed_f2<T>(a, e)) return e;
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
