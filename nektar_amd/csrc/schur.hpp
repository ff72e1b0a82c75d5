// Static-condensation diagonal kernels (k_schur.hip): the diagonal of each
// element's Schur complement A_bb - A_bi A_ii^-1 A_ib from the columns of
// its matrix (Preconditioner.cpp:185-236).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "basis.hpp"

namespace nektar_amd {
namespace schur {

constexpr int kPanel = 16;          // columns per LDS panel of the Cholesky factorisation
constexpr int kMaxBatch = 16;       // columns per substitution pass, at most
constexpr int kLdsBytes = 65536;    // LDS of one substitution pass: columns * ni doubles

// columns per substitution pass for ni interior modes.  The passes are
// latency-bound (one barrier per row); 16 columns (16 ni doubles of LDS)
// keep more elements resident per CU than 64 columns, which measured
// slower (hex P=6: 167 against 127 ms for the Schur diagonal, 92 against
// 77 ms for the interior inverses)
inline int batch_columns(int ni)
{
    const int b = ni > 0 ? kLdsBytes / (8 * ni) : kMaxBatch;
    return b < 1 ? 1 : (b > kMaxBatch ? kMaxBatch : b);
}
constexpr int kMaxInterior = 512;   // interior modes per element (hex nm = 10): 64 KiB of LDS per pass

// the interior modes of a shape's coefficient order: hex p, q, r >= 2 (p
// fastest); tet p >= 2, q >= 1, r >= 1 with r fastest
// (StdTetExp.cpp:1156-1190); the rest are boundary modes
std::vector<int> interior_modes(Shape shape, int nm);
std::vector<int> boundary_modes(Shape shape, int nm);

// dst[(e - e0) * stride + col * ld + r] = (acc ? dst : 0) + alpha y[e * nc + rows[r]]
// (rows null: the single row `row`) for ne elements from e0
hipError_t launch_take_rows(hipStream_t, int ne, int e0, int nc, int nrow, const int *rows, int row, const double *y,
                            double *dst, size_t stride, int col, int ld, double alpha = 1.0, bool acc = false);
// y[(e0 + e) * nc + rows[r]] = alpha src[e * stride + col * ld + r]
hipError_t launch_put_rows(hipStream_t, int ne, int e0, int nc, int nrow, const int *rows, const double *src,
                           size_t stride, int col, int ld, double alpha, double *y);
// B_e <- A_ii,e^-1 B_e with the Cholesky factors L (kb columns of kbld)
hipError_t launch_chol_solve(hipStream_t, int ne, int ni, int kb, int kbld, const double *L, double *B);
// y_e = M_e x_e for ne dense column-major n x n blocks (n <= kMaxInterior)
hipError_t launch_block_gemv(hipStream_t, int ne, int n, const double *M, const double *x, double *y);
// A_e = I for ne n x n blocks
hipError_t launch_identity_blocks(hipStream_t, int ne, int n, double *A);
// in-place Cholesky of ne column-major ni x ni blocks; info[e] = 1 if not
// positive definite
hipError_t launch_cholesky(hipStream_t, int ne, int ni, double *A, int *info);
// sloc[(e0 + e) * nc + bmode[c]] = dkk[e * kbld + c] - |L_e^-1 Ab_e[:, c]|^2
hipError_t launch_schur_diag(hipStream_t, int ne, int ni, int kb, int kbld, const double *L, const double *Ab,
                             const double *dkk, const int *bmode, int nc, int e0, double *sloc);

}  // namespace schur
}  // namespace nektar_amd
