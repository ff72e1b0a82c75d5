// Static-condensation diagonal (Preconditioner.cpp:185-236): for every
// element the diagonal of the Schur complement S = A_bb - A_bi A_ii^-1 A_ib
// of its Helmholtz matrix.  The element matrices come from operator applies
// to unit modes (all elements of a collection at once); these kernels
// gather their columns, factor A_ii = L L^T per element and form
// S_kk = A_kk - |L^-1 a_k|^2 for batches of boundary modes k.
//
// Layouts (per element, column-major, leading dimension ni):
//   A_ii / L : ni x ni (lower triangle holds L after the factorisation)
//   A_ib     : ni x kb (the interior rows of kb boundary columns)
//   A_kk     : kb
#include <hip/hip_runtime.h>

#include "schur.hpp"

namespace nektar_amd {
namespace schur {

namespace {

constexpr int kThreads = 256;

// dst[(e - e0) * stride + col * ld + r] = (acc ? dst : 0) + alpha y[e * nc + rows[r]]
// for the elements e0 .. e0 + ne - 1 (rows == nullptr: the single row `row`)
__global__ void take_rows_kernel(int ne, int e0, int nc, int nrow, const int *__restrict__ rows, int row,
                                 const double *__restrict__ y, double *__restrict__ dst, size_t stride, int col,
                                 int ld, double alpha, int acc)
{
    const size_t n = (size_t)ne * nrow;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int e = (int)(i / nrow), r = (int)(i % nrow);
        const int src = rows ? rows[r] : row;
        double *d = dst + (size_t)e * stride + (size_t)col * ld + r;
        const double v = alpha * y[(size_t)(e0 + e) * nc + src];
        *d = acc ? *d + v : v;
    }
}

// y[(e0 + e) * nc + rows[r]] = alpha src[e * stride + col * ld + r] (the
// inverse of take_rows; the other entries of y are left as they are)
__global__ void put_rows_kernel(int ne, int e0, int nc, int nrow, const int *__restrict__ rows,
                                const double *__restrict__ src, size_t stride, int col, int ld, double alpha,
                                double *__restrict__ y)
{
    const size_t n = (size_t)ne * nrow;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int e = (int)(i / nrow), r = (int)(i % nrow);
        y[(size_t)(e0 + e) * nc + rows[r]] = alpha * src[(size_t)e * stride + (size_t)col * ld + r];
    }
}

// B <- A_ii^-1 B = L^-T L^-1 B for kb columns (leading dimension ni, kbld
// columns per element) of one element per workgroup, in LDS
__global__ __launch_bounds__(kThreads) void chol_solve_kernel(int ni, int kb, int kbld, const double *__restrict__ L,
                                                              double *__restrict__ B)
{
    extern __shared__ double Z[];  // kb * ni
    const int tid = threadIdx.x, e = blockIdx.x;
    const double *Le = L + (size_t)e * ni * ni;
    double *Be = B + (size_t)e * ni * kbld;
    for (int idx = tid; idx < kb * ni; idx += kThreads) Z[idx] = Be[idx];
    __syncthreads();
    for (int k = 0; k < ni; ++k) {  // L z = b, column-oriented
        const double lkk = Le[(size_t)k * ni + k];
        if (tid < kb) Z[tid * ni + k] /= lkk;
        __syncthreads();
        const int rows = ni - k - 1;
        for (int idx = tid; idx < rows * kb; idx += kThreads) {
            const int r = k + 1 + idx % rows, c = idx / rows;
            Z[c * ni + r] -= Le[(size_t)k * ni + r] * Z[c * ni + k];
        }
        __syncthreads();
    }
    for (int k = ni - 1; k >= 0; --k) {  // L^T x = z: x_k, then z_r -= L(k, r) x_k for r < k
        if (tid < kb) Z[tid * ni + k] /= Le[(size_t)k * ni + k];
        __syncthreads();
        for (int idx = tid; idx < k * kb; idx += kThreads) {
            const int r = idx % k, c = idx / k;
            Z[c * ni + r] -= Le[(size_t)r * ni + k] * Z[c * ni + k];
        }
        __syncthreads();
    }
    for (int idx = tid; idx < kb * ni; idx += kThreads) Be[idx] = Z[idx];
}

// y_e = M_e x_e for dense column-major n x n blocks (the interior
// inverses A_ii^-1, m_invD of GlobalLinSysStaticCond), `per` elements per
// workgroup (n <= 256 / per), one row per work-item; HBM-bound on M
__global__ __launch_bounds__(kThreads) void block_gemv_kernel(int ne, int n, int per, const double *__restrict__ M,
                                                              const double *__restrict__ x, double *__restrict__ y)
{
    extern __shared__ double xs[];  // per * n
    const int tpe = kThreads / per, slot = threadIdx.x / tpe, lt = threadIdx.x % tpe;
    const int e = blockIdx.x * per + slot;
    double *xe = xs + slot * n;
    if (e < ne)
        for (int c = lt; c < n; c += tpe) xe[c] = x[(size_t)e * n + c];
    __syncthreads();
    if (e >= ne) return;
    const double *Me = M + (size_t)e * n * n;
    for (int r = lt; r < n; r += tpe) {
        double acc = 0.0;
        for (int c = 0; c < n; ++c) acc = fma(Me[(size_t)c * n + r], xe[c], acc);
        y[(size_t)e * n + r] = acc;
    }
}

// A_e = I (n x n) for ne elements
__global__ void identity_blocks_kernel(int ne, int n, double *__restrict__ A)
{
    const size_t tot = (size_t)ne * n * n;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
        const size_t k = i % ((size_t)n * n);
        A[i] = (k / n) == (k % n) ? 1.0 : 0.0;
    }
}

// A = L L^T in place, one workgroup per element: left-looking in panels of
// NB columns held in LDS (rows J0..ni-1 of the panel); the columns left of
// the panel are read back from global memory, where earlier panels stored
// them.  info[e] = 1 when a pivot is not positive (the values are then
// not a factorisation).
template <int NB>
__global__ __launch_bounds__(kThreads) void chol_kernel(int ni, double *__restrict__ A, int *__restrict__ info)
{
    extern __shared__ double P[];  // NB * (ni - J0)
    __shared__ int bad;
    const int tid = threadIdx.x;
    double *a = A + (size_t)blockIdx.x * ni * ni;
    if (tid == 0) bad = 0;
    for (int J0 = 0; J0 < ni; J0 += NB) {
        const int w = min(NB, ni - J0), rows = ni - J0;
        for (int idx = tid; idx < rows * w; idx += kThreads) {
            const int r = idx % rows, c = idx / rows;
            double v = a[(size_t)(J0 + c) * ni + J0 + r];
            if (r >= c)
                for (int k = 0; k < J0; ++k) v -= a[(size_t)k * ni + J0 + r] * a[(size_t)k * ni + J0 + c];
            P[c * rows + r] = v;
        }
        __syncthreads();
        for (int c = 0; c < w; ++c) {
            const double piv = P[c * rows + c];
            if (!(piv > 0.0) && tid == 0) bad = 1;
            const double dcc = sqrt(piv);
            __syncthreads();  // every thread has read the pivot
            for (int r = c + tid; r < rows; r += kThreads) P[c * rows + r] = r == c ? dcc : P[c * rows + r] / dcc;
            __syncthreads();
            const int nc2 = w - c - 1;
            for (int idx = tid; idx < rows * nc2; idx += kThreads) {
                const int r = idx % rows, c2 = c + 1 + idx / rows;
                if (r >= c2) P[c2 * rows + r] -= P[c * rows + r] * P[c * rows + c2];
            }
            __syncthreads();
        }
        for (int idx = tid; idx < rows * w; idx += kThreads) {
            const int r = idx % rows, c = idx / rows;
            if (r >= c) a[(size_t)(J0 + c) * ni + J0 + r] = P[c * rows + r];
        }
        __syncthreads();
    }
    if (tid == 0) info[blockIdx.x] = bad;
}

// S_kk = A_kk - |L^-1 a_k|^2 for kb boundary columns of one element per
// workgroup (forward substitution with the columns in LDS), written to
// sloc[(e0 + e) * nc + bmode[c]] (the element-local layout)
__global__ __launch_bounds__(kThreads) void schur_diag_kernel(int ni, int kb, int kbld, const double *__restrict__ L,
                                                              const double *__restrict__ Ab,
                                                              const double *__restrict__ dkk,
                                                              const int *__restrict__ bmode, int nc, int e0,
                                                              double *__restrict__ sloc)
{
    extern __shared__ double Z[];  // kb * ni
    const int tid = threadIdx.x, e = blockIdx.x;
    const double *Le = L + (size_t)e * ni * ni;
    const double *Ae = Ab + (size_t)e * ni * kbld;
    for (int idx = tid; idx < kb * ni; idx += kThreads) Z[idx] = Ae[idx];
    __syncthreads();
    for (int k = 0; k < ni; ++k) {
        const double lkk = Le[(size_t)k * ni + k];
        if (tid < kb) Z[tid * ni + k] /= lkk;
        __syncthreads();
        const int rows = ni - k - 1;
        for (int idx = tid; idx < rows * kb; idx += kThreads) {
            const int r = k + 1 + idx % rows, c = idx / rows;
            Z[c * ni + r] -= Le[(size_t)k * ni + r] * Z[c * ni + k];
        }
        __syncthreads();
    }
    if (tid < kb) {
        double s = 0.0;
        for (int r = 0; r < ni; ++r) s += Z[tid * ni + r] * Z[tid * ni + r];
        sloc[(size_t)(e0 + e) * nc + bmode[tid]] = dkk[(size_t)e * kbld + tid] - s;
    }
}

int grid_for(size_t n)
{
    const size_t g = (n + kThreads - 1) / kThreads;
    return (int)(g < 4096 ? (g > 0 ? g : 1) : 4096);
}

}  // namespace

hipError_t launch_take_rows(hipStream_t stream, int ne, int e0, int nc, int nrow, const int *rows, int row,
                            const double *y, double *dst, size_t stride, int col, int ld, double alpha, bool acc)
{
    if (ne <= 0 || nrow <= 0) return hipSuccess;
    hipLaunchKernelGGL(take_rows_kernel, dim3(grid_for((size_t)ne * nrow)), dim3(kThreads), 0, stream, ne, e0, nc,
                       nrow, rows, row, y, dst, stride, col, ld, alpha, acc ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_put_rows(hipStream_t stream, int ne, int e0, int nc, int nrow, const int *rows, const double *src,
                           size_t stride, int col, int ld, double alpha, double *y)
{
    if (ne <= 0 || nrow <= 0) return hipSuccess;
    hipLaunchKernelGGL(put_rows_kernel, dim3(grid_for((size_t)ne * nrow)), dim3(kThreads), 0, stream, ne, e0, nc,
                       nrow, rows, src, stride, col, ld, alpha, y);
    return hipGetLastError();
}

hipError_t launch_chol_solve(hipStream_t stream, int ne, int ni, int kb, int kbld, const double *L, double *B)
{
    if (ne <= 0 || ni <= 0 || kb <= 0) return hipSuccess;
    if (ni > kMaxInterior || kb > kbld || kb > kMaxBatch || (size_t)kb * ni * sizeof(double) > kLdsBytes)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(chol_solve_kernel, dim3(ne), dim3(kThreads), (size_t)kb * ni * sizeof(double), stream, ni, kb,
                       kbld, L, B);
    return hipGetLastError();
}

hipError_t launch_block_gemv(hipStream_t stream, int ne, int n, const double *M, const double *x, double *y)
{
    if (ne <= 0 || n <= 0) return hipSuccess;
    if (n > kMaxInterior) return hipErrorInvalidValue;
    const int per = n <= 64 ? 4 : n <= 128 ? 2 : 1;
    hipLaunchKernelGGL(block_gemv_kernel, dim3((ne + per - 1) / per), dim3(kThreads), (size_t)per * n * sizeof(double),
                       stream, ne, n, per, M, x, y);
    return hipGetLastError();
}

hipError_t launch_identity_blocks(hipStream_t stream, int ne, int n, double *A)
{
    if (ne <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(identity_blocks_kernel, dim3(grid_for((size_t)ne * n * n)), dim3(kThreads), 0, stream, ne, n,
                       A);
    return hipGetLastError();
}

hipError_t launch_cholesky(hipStream_t stream, int ne, int ni, double *A, int *info)
{
    if (ne <= 0 || ni <= 0) return hipSuccess;
    if (ni > kMaxInterior) return hipErrorInvalidValue;
    const size_t lds = (size_t)kPanel * ni * sizeof(double);
    hipLaunchKernelGGL(chol_kernel<kPanel>, dim3(ne), dim3(kThreads), lds, stream, ni, A, info);
    return hipGetLastError();
}

hipError_t launch_schur_diag(hipStream_t stream, int ne, int ni, int kb, int kbld, const double *L, const double *Ab,
                             const double *dkk, const int *bmode, int nc, int e0, double *sloc)
{
    if (ne <= 0 || kb <= 0) return hipSuccess;
    if (ni > kMaxInterior || kb > kbld || kbld > kMaxBatch || (size_t)kb * ni * sizeof(double) > kLdsBytes)
        return hipErrorInvalidValue;
    const size_t lds = (size_t)kb * ni * sizeof(double);
    hipLaunchKernelGGL(schur_diag_kernel, dim3(ne), dim3(kThreads), lds, stream, ni, kb, kbld, L, Ab, dkk, bmode, nc,
                       e0, sloc);
    return hipGetLastError();
}

}  // namespace schur
}  // namespace nektar_amd
