// Element-level stages of the runtime-size operators (k_generic.hip), in a
// header that also compiles as plain C++ for the host-side check of their
// indexing (tests/cpu_generic): there one "work-item" runs every stage
// sequentially and the barriers are no-ops.
#pragma once

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define NMF_GD __device__
#define NMF_GFI __device__ __forceinline__
namespace nektar_amd {
namespace gen {
__device__ __forceinline__ int gtid() { return threadIdx.x; }
__device__ __forceinline__ int gnt() { return blockDim.x; }
__device__ __forceinline__ void gsync() { __syncthreads(); }
}  // namespace gen
}  // namespace nektar_amd
#else
#include <cmath>
#define NMF_GD
#define NMF_GFI inline
namespace nektar_amd {
namespace gen {
inline int gtid() { return 0; }
inline int gnt() { return 1; }
inline void gsync() {}
using std::fma;
}  // namespace gen
}  // namespace nektar_amd
#endif

#include "generic.hpp"

namespace nektar_amd {
namespace gen {

// element shapes of the stages below (Shape's values)
constexpr int SH_HEX = 0, SH_TET = 1, SH_PRISM = 2, SH_PYR = 3;

template <typename T>
struct Ctx {
    const GenArgs &a;
    const T *tab;
    NMF_GD T B(int d, int m, int i) const { return tab[a.oB[d] + m * a.nq[d] + i]; }
    NMF_GD T D(int d, int l, int i) const { return tab[a.oD[d] + l * a.nq[d] + i]; }
    NMF_GD T W(int d, int i) const { return tab[a.oW[d] + i]; }
    NMF_GD T H(int c, int i) const { return tab[a.oH[c] + i]; }
};

// ---- collocation derivatives of X0 (u at the points) into X1, X2, X3
template <typename T>
NMF_GD void derivatives(const Ctx<T> &c, const T *X0, T *X1, T *X2, T *X3)
{
    const int n0 = c.a.nq[0], n1 = c.a.nq[1], n2 = c.a.nq[2], nt = c.a.nqtot;
    for (int q = gtid(); q < nt; q += gnt()) {
        const int i = q % n0, j = (q / n0) % n1, k = q / (n0 * n1);
        T s0 = T(0), s1 = T(0), s2 = T(0);
        for (int l = 0; l < n0; ++l) s0 = fma(c.D(0, l, i), X0[(k * n1 + j) * n0 + l], s0);
        for (int l = 0; l < n1; ++l) s1 = fma(c.D(1, l, j), X0[(k * n1 + l) * n0 + i], s1);
        for (int l = 0; l < n2; ++l) s2 = fma(c.D(2, l, k), X0[(l * n1 + j) * n0 + i], s2);
        X1[q] = s0;
        X2[q] = s1;
        X3[q] = s2;
    }
}

// ---- Y = X0 + D0^T X1 + D1^T X2 + D2^T X3, in place in X0
template <typename T>
NMF_GD void transposed_derivatives(const Ctx<T> &c, T *X0, const T *X1, const T *X2, const T *X3)
{
    const int n0 = c.a.nq[0], n1 = c.a.nq[1], n2 = c.a.nq[2], nt = c.a.nqtot;
    for (int q = gtid(); q < nt; q += gnt()) {
        const int i = q % n0, j = (q / n0) % n1, k = q / (n0 * n1);
        T s = X0[q];
        for (int l = 0; l < n0; ++l) s = fma(c.D(0, i, l), X1[(k * n1 + j) * n0 + l], s);
        for (int l = 0; l < n1; ++l) s = fma(c.D(1, j, l), X2[(k * n1 + l) * n0 + i], s);
        for (int l = 0; l < n2; ++l) s = fma(c.D(2, k, l), X3[(l * n1 + j) * n0 + i], s);
        X0[q] = s;
    }
}

// ---- hex transforms (BwdTransKernels.hpp:160-227, IProductKernels.hpp:270-350)
// coefficients x (p fastest) -> u at the points (dst)
template <typename T>
NMF_GD void hex_forward(const Ctx<T> &c, const T *__restrict__ x, T *X1, T *X2, T *dst)
{
    const int m0 = c.a.nm[0], m1 = c.a.nm[1], m2 = c.a.nm[2];
    const int n0 = c.a.nq[0], n1 = c.a.nq[1], n2 = c.a.nq[2];
    for (int o = gtid(); o < m2 * m1 * n0; o += gnt()) {
        const int i = o % n0, rq = o / n0;
        T s = T(0);
        for (int p = 0; p < m0; ++p) s = fma(c.B(0, p, i), x[rq * m0 + p], s);
        X1[o] = s;  // [(r*m1 + q)*n0 + i]
    }
    gsync();
    for (int o = gtid(); o < m2 * n1 * n0; o += gnt()) {
        const int i = o % n0, j = (o / n0) % n1, r = o / (n0 * n1);
        T s = T(0);
        for (int q = 0; q < m1; ++q) s = fma(c.B(1, q, j), X1[(r * m1 + q) * n0 + i], s);
        X2[o] = s;  // [(r*n1 + j)*n0 + i]
    }
    gsync();
    for (int o = gtid(); o < n2 * n1 * n0; o += gnt()) {
        const int ij = o % (n0 * n1), k = o / (n0 * n1);
        T s = T(0);
        for (int r = 0; r < m2; ++r) s = fma(c.B(2, r, k), X2[r * n0 * n1 + ij], s);
        dst[o] = s;
    }
}

// y at the points (already weighted) -> coefficients (unweighted transpose)
template <typename T>
NMF_GD void hex_reverse(const Ctx<T> &c, const T *Y, T *X1, T *X2, T *__restrict__ out)
{
    const int m0 = c.a.nm[0], m1 = c.a.nm[1], m2 = c.a.nm[2];
    const int n0 = c.a.nq[0], n1 = c.a.nq[1], n2 = c.a.nq[2];
    for (int o = gtid(); o < m2 * n1 * n0; o += gnt()) {
        const int ij = o % (n0 * n1), r = o / (n0 * n1);
        T s = T(0);
        for (int k = 0; k < n2; ++k) s = fma(c.B(2, r, k), Y[k * n0 * n1 + ij], s);
        X1[o] = s;  // [(r*n1 + j)*n0 + i]
    }
    gsync();
    for (int o = gtid(); o < m2 * m1 * n0; o += gnt()) {
        const int i = o % n0, q = (o / n0) % m1, r = o / (n0 * m1);
        T s = T(0);
        for (int j = 0; j < n1; ++j) s = fma(c.B(1, q, j), X1[(r * n1 + j) * n0 + i], s);
        X2[o] = s;  // [(r*m1 + q)*n0 + i]
    }
    gsync();
    for (int o = gtid(); o < m2 * m1 * m0; o += gnt()) {
        const int p = o % m0, rq = o / m0;
        T s = T(0);
        for (int i = 0; i < n0; ++i) s = fma(c.B(0, p, i), X2[rq * n0 + i], s);
        out[o] = s;
    }
}

// ---- tet transforms (BwdTransKernels.hpp:231-330, IProductKernels.hpp:354-514)
// Index tables (GenArgs::itab): per coefficient c its p and (p,q) pair;
// per pair its first coefficient, first Modified_C mode, r count and p.
struct TetIdx {
    const int *cP = nullptr, *cPair = nullptr, *cM2 = nullptr, *pC0 = nullptr, *pM20 = nullptr, *pNR = nullptr,
              *pP = nullptr;
    NMF_GD TetIdx() {}
    NMF_GD TetIdx(const int *it, int nc, int npq)
        : cP(it), cPair(it + nc), cM2(it + 2 * nc), pC0(it + 3 * nc), pM20(it + 3 * nc + npq),
          pNR(it + 3 * nc + 2 * npq), pP(it + 3 * nc + 3 * npq)
    {
    }
};

template <typename T>
NMF_GD void tet_forward(const Ctx<T> &c, const TetIdx &ix, const T *__restrict__ x, T *X1, T *X2, T *dst)
{
    const int m0 = c.a.nm[0], m1 = c.a.nm[1], m2 = c.a.nm[2], npq = c.a.npq;
    const int n0 = c.a.nq[0], n1 = c.a.nq[1], n2 = c.a.nq[2];
    // fpq(k) = sum_r B2(mode(p,q,r), k) x(p,q,r)
    for (int o = gtid(); o < npq * n2; o += gnt()) {
        const int k = o % n2, pq = o / n2;
        const int c0 = ix.pC0[pq], mc = ix.pM20[pq], nr = ix.pNR[pq];
        T s = T(0);
        for (int r = 0; r < nr; ++r) s = fma(c.B(2, mc + r, k), x[c0 + r], s);
        X1[o] = s;  // [pq*n2 + k]
    }
    gsync();
    // fp(j, k) = sum_q B1(pair(p,q), j) fpq(k)
    for (int o = gtid(); o < m0 * n2 * n1; o += gnt()) {
        const int j = o % n1, k = (o / n1) % n2, p = o / (n1 * n2);
        const int pq0 = p * m1 - p * (p - 1) / 2;
        T s = T(0);
        for (int q = 0; q < m1 - p; ++q) s = fma(c.B(1, pq0 + q, j), X1[(pq0 + q) * n2 + k], s);
        X2[o] = s;  // [(p*n2 + k)*n1 + j]
    }
    gsync();
    const T xtop = x[1], xbot = x[m2];
    for (int o = gtid(); o < n2 * n1 * n0; o += gnt()) {
        const int i = o % n0, j = (o / n0) % n1, k = o / (n0 * n1);
        T s = T(0);
        for (int p = 0; p < m0; ++p) s = fma(c.B(0, p, i), X2[(p * n2 + k) * n1 + j], s);
        // collapsed vertex and edge modes (BwdTransKernels.hpp:296-324)
        const T b00 = c.B(0, 0, i), b01 = c.B(0, 1, i), b10 = c.B(1, 0, j), b11 = c.B(1, 1, j);
        T t = b00 * b11;
        t = fma(b01, b10, t);
        t = fma(b01, b11, t);
        s = fma(xtop, t * c.B(2, 1, k), s);
        const T e = b01 * b11;
        s = fma(xbot, e * c.B(2, 0, k), s);
        for (int r = 1; r < m2 - 1; ++r) s = fma(x[m2 + r], e * c.B(2, r + 1, k), s);
        dst[o] = s;
    }
}

template <typename T>
NMF_GD void tet_reverse(const Ctx<T> &c, const TetIdx &ix, const T *Y, T *X1, T *X2, T *X3,
                            T *__restrict__ out)
{
    const int m0 = c.a.nm[0], m2 = c.a.nm[2], npq = c.a.npq, nc = c.a.ncoeff;
    const int n0 = c.a.nq[0], n1 = c.a.nq[1], n2 = c.a.nq[2];
    // A_p(j, k) = sum_i B0(p, i) y(i, j, k)
    for (int o = gtid(); o < m0 * n2 * n1; o += gnt()) {
        const int j = o % n1, k = (o / n1) % n2, p = o / (n1 * n2);
        T s = T(0);
        for (int i = 0; i < n0; ++i) s = fma(c.B(0, p, i), Y[(k * n1 + j) * n0 + i], s);
        X1[o] = s;  // [(p*n2 + k)*n1 + j]
    }
    gsync();
    // A_pq(k) = sum_j B1(pair, j) A_p(j, k); the two correction sums per k
    for (int o = gtid(); o < npq * n2 + 2 * n2; o += gnt()) {
        if (o < npq * n2) {
            const int k = o % n2, pq = o / n2, p = ix.pP[pq];
            T s = T(0);
            for (int j = 0; j < n1; ++j) s = fma(c.B(1, pq, j), X1[(p * n2 + k) * n1 + j], s);
            X2[o] = s;  // [pq*n2 + k]
        } else {
            const int w = o - npq * n2, k = w % n2;
            const T *a0 = X1 + (0 * n2 + k) * n1, *a1 = X1 + (1 * n2 + k) * n1;
            T s = T(0);
            if (w < n2) {  // top vertex: B1(1) A_0 + B1(0) A_1 + B1(1) A_1
                for (int j = 0; j < n1; ++j) {
                    s = fma(c.B(1, 1, j), a0[j], s);
                    s = fma(c.B(1, 0, j), a1[j], s);
                    s = fma(c.B(1, 1, j), a1[j], s);
                }
            } else {  // bottom vertex and singular edge: B1(1) A_1
                for (int j = 0; j < n1; ++j) s = fma(c.B(1, 1, j), a1[j], s);
            }
            X3[w] = s;  // [0, n2): top, [n2, 2 n2): bottom/edge
        }
    }
    gsync();
    for (int o = gtid(); o < nc; o += gnt()) {
        const int pq = ix.cPair[o], mc = ix.cM2[o];
        T s = T(0);
        for (int k = 0; k < n2; ++k) s = fma(c.B(2, mc, k), X2[pq * n2 + k], s);
        if (o == 1) {
            for (int k = 0; k < n2; ++k) s = fma(c.B(2, 1, k), X3[k], s);
        } else if (o >= m2 && o < 2 * m2 - 1) {
            // o = m2: bottom vertex (B2 mode 0); o = m2 + r: edge (mode r + 1)
            const int mode = o == m2 ? 0 : o - m2 + 1;
            for (int k = 0; k < n2; ++k) s = fma(c.B(2, mode, k), X3[n2 + k], s);
        }
        out[o] = s;
    }
}

// ---- prism transforms (BwdTransKernels.hpp:334-405, IProductKernels.hpp:519-655)
// The same index tables as the tet (TetIdx): columns (p, q) in coefficient
// order with their first coefficient, first Modified_B mode (p, r = 0),
// r count nm2 - p and p; per coefficient its p, column and Modified_B mode.
// The collapsed-edge modes (0, q, 1) carry an extra B0_1 B1_q B2_1 term
// (the reference's `correct`), which enters column (1, q) with B2 mode 1.
template <typename T>
NMF_GD void prism_forward(const Ctx<T> &c, const TetIdx &ix, const T *__restrict__ x, T *X1, T *X2, T *dst)
{
    const int m0 = c.a.nm[0], m1 = c.a.nm[1], m2 = c.a.nm[2], npq = c.a.npq;
    const int n0 = c.a.nq[0], n1 = c.a.nq[1], n2 = c.a.nq[2];
    // fpq(k) = sum_r B2(mode(p, r), k) x(p, q, r) (+ the (0, q, 1) term for p = 1)
    for (int o = gtid(); o < npq * n2; o += gnt()) {
        const int k = o % n2, col = o / n2;
        const int c0 = ix.pC0[col], mc = ix.pM20[col], nr = ix.pNR[col], p = ix.pP[col];
        T s = T(0);
        for (int r = 0; r < nr; ++r) s = fma(c.B(2, mc + r, k), x[c0 + r], s);
        if (p == 1) s = fma(c.B(2, 1, k), x[(col - m1) * m2 + 1], s);
        X1[o] = s;  // [col*n2 + k]
    }
    gsync();
    // fp(j, k) = sum_q B1(q, j) fpq(k)
    for (int o = gtid(); o < m0 * n2 * n1; o += gnt()) {
        const int j = o % n1, k = (o / n1) % n2, p = o / (n1 * n2);
        T s = T(0);
        for (int q = 0; q < m1; ++q) s = fma(c.B(1, q, j), X1[(p * m1 + q) * n2 + k], s);
        X2[o] = s;  // [(p*n2 + k)*n1 + j]
    }
    gsync();
    for (int o = gtid(); o < n2 * n1 * n0; o += gnt()) {
        const int i = o % n0, j = (o / n0) % n1, k = o / (n0 * n1);
        T s = T(0);
        for (int p = 0; p < m0; ++p) s = fma(c.B(0, p, i), X2[(p * n2 + k) * n1 + j], s);
        dst[o] = s;
    }
}

template <typename T>
NMF_GD void prism_reverse(const Ctx<T> &c, const TetIdx &ix, const T *Y, T *X1, T *X2, T *__restrict__ out)
{
    const int m0 = c.a.nm[0], m1 = c.a.nm[1], npq = c.a.npq, nc = c.a.ncoeff;
    const int n0 = c.a.nq[0], n1 = c.a.nq[1], n2 = c.a.nq[2];
    // A_p(j, k) = sum_i B0(p, i) y(i, j, k)
    for (int o = gtid(); o < m0 * n2 * n1; o += gnt()) {
        const int j = o % n1, k = (o / n1) % n2, p = o / (n1 * n2);
        T s = T(0);
        for (int i = 0; i < n0; ++i) s = fma(c.B(0, p, i), Y[(k * n1 + j) * n0 + i], s);
        X1[o] = s;  // [(p*n2 + k)*n1 + j]
    }
    gsync();
    // A_pq(k) = sum_j B1(q, j) A_p(j, k)
    for (int o = gtid(); o < npq * n2; o += gnt()) {
        const int k = o % n2, col = o / n2, p = ix.pP[col], q = col - p * m1;
        T s = T(0);
        for (int j = 0; j < n1; ++j) s = fma(c.B(1, q, j), X1[(p * n2 + k) * n1 + j], s);
        X2[o] = s;  // [col*n2 + k]
    }
    gsync();
    // out(p, q, r) = sum_k B2(mode(p, r), k) A_pq(k); (0, q, 1) also takes
    // sum_k B2(1, k) A_1q(k)
    for (int o = gtid(); o < nc; o += gnt()) {
        const int col = ix.cPair[o], mc = ix.cM2[o];
        T s = T(0);
        for (int k = 0; k < n2; ++k) s = fma(c.B(2, mc, k), X2[col * n2 + k], s);
        if (ix.cP[o] == 0 && mc == 1)
            for (int k = 0; k < n2; ++k) s = fma(c.B(2, 1, k), X2[(m1 + col) * n2 + k], s);
        out[o] = s;
    }
}

// ---- pyramid transforms (BwdTransKernels.hpp:407-489, IProductKernels.hpp:660-815)
// Index tables as the tet (TetIdx): columns (p, q), p-major, with their
// first coefficient, first ModifiedPyr_C mode, r count nm2 - max(p, q) and
// p; per coefficient its p, column and mode.  Directions 0 and 1 are
// Modified_A (indexed by p and q).  The top vertex (coefficient 1) also
// carries (B0_0 B1_1 + B0_1 B1_0 + B0_1 B1_1) B2_1 (the reference's
// `correct`).
template <typename T>
NMF_GD void pyr_forward(const Ctx<T> &c, const TetIdx &ix, const T *__restrict__ x, T *X1, T *X2, T *dst)
{
    const int m0 = c.a.nm[0], m1 = c.a.nm[1], npq = c.a.npq;
    const int n0 = c.a.nq[0], n1 = c.a.nq[1], n2 = c.a.nq[2];
    // fpq(k) = sum_r B2(mode(p, q, r), k) x(p, q, r)
    for (int o = gtid(); o < npq * n2; o += gnt()) {
        const int k = o % n2, col = o / n2;
        const int c0 = ix.pC0[col], mc = ix.pM20[col], nr = ix.pNR[col];
        T s = T(0);
        for (int r = 0; r < nr; ++r) s = fma(c.B(2, mc + r, k), x[c0 + r], s);
        X1[o] = s;  // [col*n2 + k]
    }
    gsync();
    // fp(j, k) = sum_q B1(q, j) fpq(k)
    for (int o = gtid(); o < m0 * n2 * n1; o += gnt()) {
        const int j = o % n1, k = (o / n1) % n2, p = o / (n1 * n2);
        T s = T(0);
        for (int q = 0; q < m1; ++q) s = fma(c.B(1, q, j), X1[(p * m1 + q) * n2 + k], s);
        X2[o] = s;  // [(p*n2 + k)*n1 + j]
    }
    gsync();
    const T xtop = x[1];
    for (int o = gtid(); o < n2 * n1 * n0; o += gnt()) {
        const int i = o % n0, j = (o / n0) % n1, k = o / (n0 * n1);
        T s = T(0);
        for (int p = 0; p < m0; ++p) s = fma(c.B(0, p, i), X2[(p * n2 + k) * n1 + j], s);
        const T b00 = c.B(0, 0, i), b01 = c.B(0, 1, i), b10 = c.B(1, 0, j), b11 = c.B(1, 1, j);
        T t = b00 * b11;
        t = fma(b01, b10, t);
        t = fma(b01, b11, t);
        s = fma(xtop, t * c.B(2, 1, k), s);
        dst[o] = s;
    }
}

template <typename T>
NMF_GD void pyr_reverse(const Ctx<T> &c, const TetIdx &ix, const T *Y, T *X1, T *X2, T *X3, T *__restrict__ out)
{
    const int m0 = c.a.nm[0], m1 = c.a.nm[1], npq = c.a.npq, nc = c.a.ncoeff;
    const int n0 = c.a.nq[0], n1 = c.a.nq[1], n2 = c.a.nq[2];
    // A_p(j, k) = sum_i B0(p, i) y(i, j, k)
    for (int o = gtid(); o < m0 * n2 * n1; o += gnt()) {
        const int j = o % n1, k = (o / n1) % n2, p = o / (n1 * n2);
        T s = T(0);
        for (int i = 0; i < n0; ++i) s = fma(c.B(0, p, i), Y[(k * n1 + j) * n0 + i], s);
        X1[o] = s;  // [(p*n2 + k)*n1 + j]
    }
    gsync();
    // A_pq(k) = sum_j B1(q, j) A_p(j, k); the top-vertex sum per k
    for (int o = gtid(); o < npq * n2 + n2; o += gnt()) {
        if (o < npq * n2) {
            const int k = o % n2, col = o / n2, p = ix.pP[col], q = col - p * m1;
            T s = T(0);
            for (int j = 0; j < n1; ++j) s = fma(c.B(1, q, j), X1[(p * n2 + k) * n1 + j], s);
            X2[o] = s;  // [col*n2 + k]
        } else {
            // B1(1) A_0 + B1(0) A_1 + B1(1) A_1
            const int k = o - npq * n2;
            const T *a0 = X1 + (0 * n2 + k) * n1, *a1 = X1 + (1 * n2 + k) * n1;
            T s = T(0);
            for (int j = 0; j < n1; ++j) {
                s = fma(c.B(1, 1, j), a0[j], s);
                s = fma(c.B(1, 0, j), a1[j], s);
                s = fma(c.B(1, 1, j), a1[j], s);
            }
            X3[k] = s;
        }
    }
    gsync();
    for (int o = gtid(); o < nc; o += gnt()) {
        const int col = ix.cPair[o], mc = ix.cM2[o];
        T s = T(0);
        for (int k = 0; k < n2; ++k) s = fma(c.B(2, mc, k), X2[col * n2 + k], s);
        if (o == 1)
            for (int k = 0; k < n2; ++k) s = fma(c.B(2, 1, k), X3[k], s);
        out[o] = s;
    }
}

// ---- pointwise stages -------------------------------------------------
// geometry of point q of element e: regular g[e*G + c], deformed
// g[(e*G + c)*nqtot + q]
template <typename T>
NMF_GFI T geo_at(const GenArgs &a, const T *g, int G, int e, int cc, int q)
{
    return a.deformed ? g[((size_t)e * G + cc) * a.nqtot + q] : g[(size_t)e * G + cc];
}

// tet half-space factors (HelmholtzKernels.hpp:1368-1389)
template <typename T>
struct Half {
    T h0, h1, h2, h3;
};
template <typename T>
NMF_GFI Half<T> half_at(const Ctx<T> &c, int i, int j, int k)
{
    return {c.H(0, i), c.H(1, j), c.H(2, j), c.H(3, k)};
}

// Helmholtz / LinearAdvectionDiffusionReaction pointwise stage: in X0 = u,
// X1..X3 = derivatives along the reference (collapsed) directions; out
// X0 = weighted mass integrand, X1..X3 = weighted flux integrands.
template <typename T, int SH>
NMF_GD void helmholtz_pointwise(const Ctx<T> &c, int e, T *X0, T *X1, T *X2, T *X3)
{
    const GenArgs &a = c.a;
    const T *g = static_cast<const T *>(a.geo);
    const T *adv = static_cast<const T *>(a.adv);
    const T lambda = T(a.lambda);
    const int n0 = a.nq[0], n1 = a.nq[1];
    for (int q = gtid(); q < a.nqtot; q += gnt()) {
        const int i = q % n0, j = (q / n0) % n1, k = q / (n0 * n1);
        const T w = c.W(0, i) * c.W(1, j) * c.W(2, k);
        const T a0 = X1[q], a1 = X2[q], a2 = X3[q];
        T m = X0[q];
        if (a.op == OpType::AdvDiffReact) {
            const T *ac = adv + (size_t)e * 3 * a.nqtot + q;
            m = fma(ac[0], a0, m);
            m = fma(ac[a.nqtot], a1, m);
            m = fma(ac[2 * a.nqtot], a2, m);
        }
        T t0, t1, t2, mass;
        if constexpr (SH == SH_HEX) {
            // packed J m_de (regular, weights applied here) or W J m_de
            T gg[7];
            for (int cc = 0; cc < 7; ++cc) gg[cc] = geo_at(a, g, 7, e, cc, q);
            t0 = gg[0] * a0;
            t0 = fma(gg[1], a1, t0);
            t0 = fma(gg[2], a2, t0);
            t1 = gg[1] * a0;
            t1 = fma(gg[3], a1, t1);
            t1 = fma(gg[4], a2, t1);
            t2 = gg[2] * a0;
            t2 = fma(gg[4], a1, t2);
            t2 = fma(gg[5], a2, t2);
            if (!a.deformed) {
                t0 *= w;
                t1 *= w;
                t2 *= w;
                mass = (lambda * gg[6]) * (w * m);
            } else {
                mass = (lambda * gg[6]) * m;
            }
        } else if constexpr (SH == SH_PYR) {
            // DiffusionCoeffPyrKernel (HelmholtzKernels.hpp:1119-1326):
            // s0..2 = h2 df[0,3,6] + h0 h2 df[2,5,8], s3..5 = h2 df[1,4,7] +
            // h1 h2 df[2,5,8], h0 = (1+z0)/2, h1 = (1+z1)/2, h2 = 2/(1-z2)
            T f[9];
            for (int cc = 0; cc < 9; ++cc) f[cc] = geo_at(a, g, 10, e, cc, q);
            const T jw = geo_at(a, g, 10, e, 9, q) * w;
            const T h2 = c.H(3, k), h0h2 = c.H(0, i) * h2, h1h2 = c.H(1, j) * h2;
            const T s0 = fma(h0h2, f[2], h2 * f[0]), s1 = fma(h0h2, f[5], h2 * f[3]),
                    s2 = fma(h0h2, f[8], h2 * f[6]);
            const T s3 = fma(h1h2, f[2], h2 * f[1]), s4 = fma(h1h2, f[5], h2 * f[4]),
                    s5 = fma(h1h2, f[8], h2 * f[7]);
            T g0 = s0 * s0;
            g0 = fma(s1, s1, g0);
            g0 = fma(s2, s2, g0);
            T g1 = s3 * s3;
            g1 = fma(s4, s4, g1);
            g1 = fma(s5, s5, g1);
            T g2 = f[2] * f[2];
            g2 = fma(f[5], f[5], g2);
            g2 = fma(f[8], f[8], g2);
            T g3 = s0 * s3;
            g3 = fma(s1, s4, g3);
            g3 = fma(s2, s5, g3);
            T g4 = f[2] * s0;
            g4 = fma(f[5], s1, g4);
            g4 = fma(f[8], s2, g4);
            T g5 = f[2] * s3;
            g5 = fma(f[5], s4, g5);
            g5 = fma(f[8], s5, g5);
            t0 = g0 * a0;
            t0 = fma(g3, a1, t0);
            t0 = fma(g4, a2, t0);
            t1 = g3 * a0;
            t1 = fma(g1, a1, t1);
            t1 = fma(g5, a2, t1);
            t2 = g4 * a0;
            t2 = fma(g5, a1, t2);
            t2 = fma(g2, a2, t2);
            t0 *= jw;
            t1 *= jw;
            t2 *= jw;
            mass = lambda * (jw * m);
        } else if constexpr (SH == SH_PRISM) {
            // DiffusionCoeffPrismKernel (HelmholtzKernels.hpp:918-1119):
            // s_c = h1 (h0 df[3c+2] + df[3c]), h0 = (1+z0)/2, h1 = 2/(1-z2)
            T f[9];
            for (int cc = 0; cc < 9; ++cc) f[cc] = geo_at(a, g, 10, e, cc, q);
            const T jw = geo_at(a, g, 10, e, 9, q) * w;
            const T h0 = c.H(0, i), h1 = c.H(3, k);
            const T s1 = h1 * fma(h0, f[2], f[0]);
            const T s2 = h1 * fma(h0, f[5], f[3]);
            const T s3 = h1 * fma(h0, f[8], f[6]);
            T g0 = s1 * s1;
            g0 = fma(s2, s2, g0);
            g0 = fma(s3, s3, g0);
            T g3 = f[1] * s1;
            g3 = fma(f[4], s2, g3);
            g3 = fma(f[7], s3, g3);
            T g4 = f[2] * s1;
            g4 = fma(f[5], s2, g4);
            g4 = fma(f[8], s3, g4);
            T g1 = f[1] * f[1];
            g1 = fma(f[4], f[4], g1);
            g1 = fma(f[7], f[7], g1);
            T g2 = f[2] * f[2];
            g2 = fma(f[5], f[5], g2);
            g2 = fma(f[8], f[8], g2);
            T g5 = f[1] * f[2];
            g5 = fma(f[4], f[5], g5);
            g5 = fma(f[7], f[8], g5);
            t0 = g0 * a0;
            t0 = fma(g3, a1, t0);
            t0 = fma(g4, a2, t0);
            t1 = g3 * a0;
            t1 = fma(g1, a1, t1);
            t1 = fma(g5, a2, t1);
            t2 = g4 * a0;
            t2 = fma(g5, a1, t2);
            t2 = fma(g2, a2, t2);
            t0 *= jw;
            t1 *= jw;
            t2 *= jw;
            mass = lambda * (jw * m);
        } else {
            // DiffusionCoeffTetKernel (HelmholtzKernels.hpp:700-913), metric
            // g0..g5 of the collapsed derivatives
            T f[9];
            for (int cc = 0; cc < 9; ++cc) f[cc] = geo_at(a, g, 10, e, cc, q);
            const T jw = geo_at(a, g, 10, e, 9, q) * w;
            const Half<T> h = half_at(c, i, j, k);
            const T h2h3 = h.h2 * h.h3, h1h3 = h.h1 * h.h3, h0h2h3 = h.h0 * h2h3;
            T s1 = (f[1] + f[2]) * h0h2h3;
            s1 = fma(f[0], h2h3, s1);
            T s2 = (f[4] + f[5]) * h0h2h3;
            s2 = fma(f[3], h2h3, s2);
            T s3 = (f[7] + f[8]) * h0h2h3;
            s3 = fma(f[6], h2h3, s3);
            T s4 = f[1] * h.h3;
            s4 = fma(f[2], h1h3, s4);
            T s5 = f[4] * h.h3;
            s5 = fma(f[5], h1h3, s5);
            T s6 = f[7] * h.h3;
            s6 = fma(f[8], h1h3, s6);
            T g0 = s1 * s1;
            g0 = fma(s2, s2, g0);
            g0 = fma(s3, s3, g0);
            T g4 = f[2] * s1;
            g4 = fma(f[5], s2, g4);
            g4 = fma(f[8], s3, g4);
            T g3 = s1 * s4;
            g3 = fma(s2, s5, g3);
            g3 = fma(s3, s6, g3);
            T g1 = s4 * s4;
            g1 = fma(s5, s5, g1);
            g1 = fma(s6, s6, g1);
            T g5 = f[2] * s4;
            g5 = fma(f[5], s5, g5);
            g5 = fma(f[8], s6, g5);
            T g2 = f[2] * f[2];
            g2 = fma(f[5], f[5], g2);
            g2 = fma(f[8], f[8], g2);
            t0 = g0 * a0;
            t0 = fma(g3, a1, t0);
            t0 = fma(g4, a2, t0);
            t1 = g3 * a0;
            t1 = fma(g1, a1, t1);
            t1 = fma(g5, a2, t1);
            t2 = g4 * a0;
            t2 = fma(g5, a1, t2);
            t2 = fma(g2, a2, t2);
            t0 *= jw;
            t1 *= jw;
            t2 *= jw;
            mass = lambda * (jw * m);
        }
        X0[q] = mass;
        X1[q] = t0;
        X2[q] = t1;
        X3[q] = t2;
    }
}

// PhysDeriv pointwise stage: collapsed chain rule for tets
// (PhysDerivKernels.hpp:198-259) and d/dx_j = sum_d df[j*3+d] d/dxi_d
// (:326-340); writes the three outputs of element e.
template <typename T, int SH>
NMF_GD void physderiv_pointwise(const Ctx<T> &c, int e, const T *X1, const T *X2, const T *X3)
{
    const GenArgs &a = c.a;
    const T *g = static_cast<const T *>(a.geo);
    const int n0 = a.nq[0], n1 = a.nq[1];
    T *o0 = static_cast<T *>(a.out0) + (size_t)e * a.nqtot;
    T *o1 = static_cast<T *>(a.out1) + (size_t)e * a.nqtot;
    T *o2 = static_cast<T *>(a.out2) + (size_t)e * a.nqtot;
    for (int q = gtid(); q < a.nqtot; q += gnt()) {
        T d0 = X1[q], d1 = X2[q], d2 = X3[q];
        if constexpr (SH == SH_PRISM) {
            // PhysDerivKernels.hpp:293-301: d0 *= 2/(1-z2), d2 += (1+z0)/2 d0
            const int i = q % n0, k = q / (n0 * n1);
            d0 = d0 * c.H(3, k);
            d2 = fma(c.H(0, i), d0, d2);
        }
        if constexpr (SH == SH_PYR) {
            // PhysDerivKernels.hpp:301-311: d0, d1 *= 2/(1-z2),
            // d2 += (1+z0)/2 d0 + (1+z1)/2 d1
            const int i = q % n0, j = (q / n0) % n1, k = q / (n0 * n1);
            d0 = d0 * c.H(3, k);
            d1 = d1 * c.H(3, k);
            d2 = fma(c.H(0, i), d0, d2);
            d2 = fma(c.H(1, j), d1, d2);
        }
        if constexpr (SH == SH_TET) {
            const int i = q % n0, j = (q / n0) % n1, k = q / (n0 * n1);
            const Half<T> h = half_at(c, i, j, k);
            d0 = d0 * (h.h2 * h.h3);
            const T w0 = d0 * h.h0;
            const T w1 = d1 * h.h3;
            d1 = w0 + w1;
            d2 = fma(w1, h.h1, w0) + d2;
        }
        T f[9];
        for (int cc = 0; cc < 9; ++cc) f[cc] = geo_at(a, g, 9, e, cc, q);
        T r0 = d0 * f[0];
        r0 = fma(d1, f[1], r0);
        r0 = fma(d2, f[2], r0);
        T r1 = d0 * f[3];
        r1 = fma(d1, f[4], r1);
        r1 = fma(d2, f[5], r1);
        T r2 = d0 * f[6];
        r2 = fma(d1, f[7], r2);
        r2 = fma(d2, f[8], r2);
        o0[q] = r0;
        o1[q] = r1;
        o2[q] = r2;
    }
}

// IProductWRTDerivBase pointwise stage (IProductWRTDerivBaseKernels.hpp:
// 194-335): t_d = sum_j df[j*3+d] in_j, mapped to the collapsed directions
// for tets, weighted; X0 = 0.
template <typename T, int SH>
NMF_GD void iprod_deriv_pointwise(const Ctx<T> &c, int e, T *X0, T *X1, T *X2, T *X3)
{
    const GenArgs &a = c.a;
    const T *g = static_cast<const T *>(a.geo);
    const int n0 = a.nq[0], n1 = a.nq[1];
    const size_t base = (size_t)e * a.nqtot;
    const T *in0 = static_cast<const T *>(a.in0) + base, *in1 = static_cast<const T *>(a.in1) + base,
            *in2 = static_cast<const T *>(a.in2) + base;
    for (int q = gtid(); q < a.nqtot; q += gnt()) {
        const int i = q % n0, j = (q / n0) % n1, k = q / (n0 * n1);
        const T w = c.W(0, i) * c.W(1, j) * c.W(2, k);
        const T f0 = in0[q], f1 = in1[q], f2 = in2[q];
        T t0, t1, t2;
        if constexpr (SH == SH_HEX) {
            // packed J df (regular, weights here) or W J df
            T gg[9];
            for (int cc = 0; cc < 9; ++cc) gg[cc] = geo_at(a, g, 9, e, cc, q);
            t0 = gg[0] * f0;
            t0 = fma(gg[3], f1, t0);
            t0 = fma(gg[6], f2, t0);
            t1 = gg[1] * f0;
            t1 = fma(gg[4], f1, t1);
            t1 = fma(gg[7], f2, t1);
            t2 = gg[2] * f0;
            t2 = fma(gg[5], f1, t2);
            t2 = fma(gg[8], f2, t2);
            if (!a.deformed) {
                t0 *= w;
                t1 *= w;
                t2 *= w;
            }
        } else {
            T d[9];
            for (int cc = 0; cc < 9; ++cc) d[cc] = geo_at(a, g, 10, e, cc, q);
            const T jw = geo_at(a, g, 10, e, 9, q) * w;
            t0 = d[0] * f0 + d[3] * f1 + d[6] * f2;
            t1 = d[1] * f0 + d[4] * f1 + d[7] * f2;
            t2 = d[2] * f0 + d[5] * f1 + d[8] * f2;
            if constexpr (SH == SH_PRISM) {
                // IProductWRTDerivBaseKernels.hpp:391-403: t0 <- h1 (t0 + h0 t2)
                t0 = c.H(3, k) * fma(c.H(0, i), t2, t0);
            } else if constexpr (SH == SH_PYR) {
                // IProductWRTDerivBaseKernels.hpp:469-488: t0 <- h2 (t0 + h0 t2),
                // t1 <- h2 (t1 + h1 t2)
                t0 = c.H(3, k) * fma(c.H(0, i), t2, t0);
                t1 = c.H(3, k) * fma(c.H(1, j), t2, t1);
            } else {
                const Half<T> h = half_at(c, i, j, k);
                t0 = fma(t1 + t2, h.h0, t0);
                t0 = t0 * (h.h2 * h.h3);
                t1 = fma(t2, h.h1, t1);
                t1 = t1 * h.h3;
            }
            t0 *= jw;
            t1 *= jw;
            t2 *= jw;
        }
        X0[q] = T(0);
        X1[q] = t0;
        X2[q] = t1;
        X3[q] = t2;
    }
}

// IProductWRTBase pointwise stage: y = W J f
template <typename T, int SH>
NMF_GD void iprod_pointwise(const Ctx<T> &c, int e, T *X0)
{
    const GenArgs &a = c.a;
    const T *jac = static_cast<const T *>(a.geo);
    const T *f = static_cast<const T *>(a.in0) + (size_t)e * a.nqtot;
    const int n0 = a.nq[0], n1 = a.nq[1];
    const T jv = a.deformed ? T(0) : jac[e];
    for (int q = gtid(); q < a.nqtot; q += gnt()) {
        const int i = q % n0, j = (q / n0) % n1, k = q / (n0 * n1);
        const T w = c.W(0, i) * c.W(1, j) * c.W(2, k);
        const T jq = a.deformed ? jac[(size_t)e * a.nqtot + q] : jv;
        X0[q] = f[q] * jq * w;
    }
}

// ---- one element -------------------------------------------------------
template <typename T, int SH>
NMF_GD void element(const GenArgs &a, int e, T *X0, T *X1, T *X2, T *X3)
{
    const Ctx<T> c{a, static_cast<const T *>(a.tab)};
    const TetIdx ix = SH != SH_HEX ? TetIdx(a.itab, a.ncoeff, a.npq) : TetIdx();
    const OpType op = a.op;
    const bool coeff_in = op == OpType::BwdTrans || op == OpType::Helmholtz || op == OpType::AdvDiffReact;
    if (coeff_in) {
        const T *x = static_cast<const T *>(a.in0) + (size_t)e * a.ncoeff;
        T *dst = op == OpType::BwdTrans ? static_cast<T *>(a.out0) + (size_t)e * a.nqtot : X0;
        if constexpr (SH == SH_TET) tet_forward(c, ix, x, X1, X2, dst);
        else if constexpr (SH == SH_PRISM) prism_forward(c, ix, x, X1, X2, dst);
        else if constexpr (SH == SH_PYR) pyr_forward(c, ix, x, X1, X2, dst);
        else hex_forward(c, x, X1, X2, dst);
        if (op == OpType::BwdTrans) return;
    } else if (op == OpType::PhysDeriv) {
        const T *f = static_cast<const T *>(a.in0) + (size_t)e * a.nqtot;
        for (int q = gtid(); q < a.nqtot; q += gnt()) X0[q] = f[q];
    }
    gsync();
    if (op == OpType::IProduct) {
        iprod_pointwise<T, SH>(c, e, X0);
    } else if (op == OpType::IProductDeriv) {
        iprod_deriv_pointwise<T, SH>(c, e, X0, X1, X2, X3);
        gsync();
        transposed_derivatives(c, X0, X1, X2, X3);
    } else {
        derivatives(c, X0, X1, X2, X3);
        gsync();
        if (op == OpType::PhysDeriv) {
            physderiv_pointwise<T, SH>(c, e, X1, X2, X3);
            return;
        }
        helmholtz_pointwise<T, SH>(c, e, X0, X1, X2, X3);
        gsync();
        transposed_derivatives(c, X0, X1, X2, X3);
    }
    gsync();
    T *out = static_cast<T *>(a.out0) + (size_t)e * a.ncoeff;
    if constexpr (SH == SH_TET) tet_reverse(c, ix, X0, X1, X2, X3, out);
    else if constexpr (SH == SH_PRISM) prism_reverse(c, ix, X0, X1, X2, out);
    else if constexpr (SH == SH_PYR) pyr_reverse(c, ix, X0, X1, X2, X3, out);
    else hex_reverse(c, X0, X1, X2, out);
}

}  // namespace gen
}  // namespace nektar_amd
