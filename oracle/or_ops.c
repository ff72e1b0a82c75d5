/*
 * or_ops.c -- restatement of Nektar++'s MatrixFreeOps kernels for hex and
 * tet elements: BwdTrans, IProductWRTBase, PhysDeriv and the fused
 * Helmholtz operator, in FP64, vectorised across elements exactly like the
 * reference (one SIMD lane = one element, fused multiply-add wherever the
 * reference calls vec_t::fma, which the AVX2/AVX-512 NektarSIMD builds fuse:
 * SimdLib/avx2.hpp:503-508).
 *
 * TEST INFRASTRUCTURE ONLY (parity oracle and CPU baseline).
 *
 * Sources restated (library/MatrixFreeOps/):
 *   BwdTransKernels.hpp:160-227 (hex), :231-330 (tet), :334-405 (prism)
 *   IProductKernels.hpp:44-64 (ScaleAppend), :270-350 (hex), :354-514 (tet),
 *                       :519-655 (prism)
 *   HelmholtzKernels.hpp:918-1119 (prism metric), :1390-1403 (prism h0, h1)
 *   IProductWRTDerivBaseKernels.hpp:338-411 (prism),
 *   LinearAdvectionDiffusionReactionKernels.hpp:384-470 (prism)
 *   PhysDerivKernels.hpp:184-345 (PhysDeriv3DKernel), :436-507 (tensor)
 *   HelmholtzKernels.hpp:389-696 (hex metric), :700-913 (tet metric),
 *                         :1354-1425 (tet half-space factors h0..h3)
 *   Helmholtz.h:549-599 (fusedKernel3D), :374-464 (block loop, padding)
 *   Operator.hpp:499-674 (Helper tables, Radau weight pre-scaling)
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define VW 4 /* SIMD lanes (elements) per block: AVX2 width for doubles */

typedef struct { double v[VW]; } vd;

static inline vd vz(void) { vd r; for (int l = 0; l < VW; ++l) r.v[l] = 0.0; return r; }
static inline vd vs(double a) { vd r; for (int l = 0; l < VW; ++l) r.v[l] = a; return r; }
static inline vd vmul(vd a, vd b) { vd r; for (int l = 0; l < VW; ++l) r.v[l] = a.v[l] * b.v[l]; return r; }
static inline vd vmuls(vd a, double b) { vd r; for (int l = 0; l < VW; ++l) r.v[l] = a.v[l] * b; return r; }
static inline vd vadd(vd a, vd b) { vd r; for (int l = 0; l < VW; ++l) r.v[l] = a.v[l] + b.v[l]; return r; }
/* acc + a*b fused (tinysimd fma: "this = this + a * b") */
static inline vd vfma(vd acc, vd a, vd b) { vd r; for (int l = 0; l < VW; ++l) r.v[l] = fma(a.v[l], b.v[l], acc.v[l]); return r; }
static inline vd vfmas(vd acc, vd a, double b) { vd r; for (int l = 0; l < VW; ++l) r.v[l] = fma(a.v[l], b, acc.v[l]); return r; }

/* ------------------------------------------------------------------ */
/* Per-operator tables: MatrixFree::Helper (Operator.hpp:529-641), with
 * modes nmd[d] and points nq[d] per direction (the runtime-size operators,
 * Helmholtz.h:374-464, take any of them; the MODIFIED defaults have equal
 * modes and nq = {Q, Q, Q} (hex) or {Q, Q-1, Q-1} (tet)).  nm = nmd[0]. */
typedef struct {
    int shape, nm, nmd[3], nq[3], nqtot, ncoeff;
    double *b[3], *db[3], *D[3], *w[3], *Z[3];
    double h0[64], h1[64], h2[64], h3[64];
} tables_t;

/* Coefficients of a tet with nm0 <= nm1 <= nm2 modes: p < nm0, q < nm1 - p,
 * r < nm2 - p - q (the packing BwdTransKernels.hpp:244-266 walks). */
static int tet_ncoeffs3(int nm0, int nm1, int nm2)
{
    int n = 0;
    for (int p = 0; p < nm0; ++p)
        for (int q = 0; q < nm1 - p; ++q) n += nm2 - p - q;
    return n;
}

/* Coefficients of a prism with nm0 <= nm2: p < nm0, q < nm1, r < nm2 - p
 * (StdPrismExp::GetMode, StdPrismExp.cpp:1947-1956; the order
 * BwdTransKernels.hpp:345-360 walks). */
static int prism_ncoeffs3(int nm0, int nm1, int nm2)
{
    int n = 0;
    for (int p = 0; p < nm0; ++p) n += nm1 * (nm2 - p);
    return n;
}

/* Coefficients of a pyramid: p < nm0, q < nm1, r < nm2 - max(p, q)
 * (StdPyrExp, the order BwdTransKernels.hpp:419-446 walks) */
static int pyr_ncoeffs3(int nm0, int nm1, int nm2)
{
    int n = 0;
    for (int p = 0; p < nm0; ++p)
        for (int q = 0; q < nm1; ++q) n += nm2 - (p > q ? p : q);
    return n;
}

int or_ncoeffs_aniso(int shape, const int nm[3])
{
    return shape == OR_HEX     ? nm[0] * nm[1] * nm[2]
           : shape == OR_TET   ? tet_ncoeffs3(nm[0], nm[1], nm[2])
           : shape == OR_PRISM ? prism_ncoeffs3(nm[0], nm[1], nm[2])
                               : pyr_ncoeffs3(nm[0], nm[1], nm[2]);
}

static void tables_init_full(tables_t *t, int shape, const int nm[3], const int nq[3])
{
    t->shape = shape;
    t->nm = nm[0];
    for (int d = 0; d < 3; ++d) {
        t->nmd[d] = nm[d];
        t->nq[d] = nq[d];
    }
    t->nqtot = t->nq[0] * t->nq[1] * t->nq[2];
    t->ncoeff = or_ncoeffs_aniso(shape, nm);
    /* prism: Modified_A, Modified_A, Modified_B on GLL, GLL, Radau(1,0);
     * pyramid: Modified_A, Modified_A, ModifiedPyr_C on GLL, GLL, Radau(2,0)
     * (MeshGraph.cpp:1754-1770) */
    static const int prism_bt[3] = {OR_MOD_A, OR_MOD_A, OR_MOD_B}, prism_pt[3] = {OR_GLL, OR_GLL, OR_RADAU_A1};
    static const int pyr_bt[3] = {OR_MOD_A, OR_MOD_A, OR_MOD_PYR_C}, pyr_pt[3] = {OR_GLL, OR_GLL, OR_RADAU_A2};
    for (int d = 0; d < 3; ++d) {
        int bt = shape == OR_HEX ? OR_MOD_A : shape == OR_TET ? d : shape == OR_PRISM ? prism_bt[d] : pyr_bt[d];
        int pt = shape == OR_HEX ? OR_GLL : shape == OR_TET ? d : shape == OR_PRISM ? prism_pt[d] : pyr_pt[d];
        int nq = t->nq[d];
        int nmodes = or_basis_nmodes(bt, nm[d]);
        t->b[d] = (double *)malloc((size_t)nmodes * nq * sizeof(double));
        t->db[d] = (double *)malloc((size_t)nmodes * nq * sizeof(double));
        t->D[d] = (double *)malloc((size_t)nq * nq * sizeof(double));
        t->w[d] = (double *)malloc((size_t)nq * sizeof(double));
        t->Z[d] = (double *)malloc((size_t)nq * sizeof(double));
        or_basis(bt, nm[d], nq, t->b[d], t->db[d]);
        or_points(pt, nq, t->Z[d], t->w[d], t->D[d]);
        /* Operator.hpp:600-616: Radau(1,0) weights x0.5, Radau(2,0) x0.25 */
        double fac = pt == OR_RADAU_A1 ? 0.5 : pt == OR_RADAU_A2 ? 0.25 : 1.0;
        for (int i = 0; i < nq; ++i) t->w[d][i] *= fac;
    }
    if (shape == OR_TET) {
        /* HelmholtzKernels.hpp:1368-1389 (GetHelmholtz3DHalfSpace, tet) */
        for (int i = 0; i < t->nq[0]; ++i) t->h0[i] = 0.5 * (1 + t->Z[0][i]);
        for (int j = 0; j < t->nq[1]; ++j) {
            t->h1[j] = 0.5 * (1 + t->Z[1][j]);
            t->h2[j] = 2.0 / (1 - t->Z[1][j]);
        }
        for (int k = 0; k < t->nq[2]; ++k) t->h3[k] = 2.0 / (1 - t->Z[2][k]);
    }
    if (shape == OR_PRISM) {
        /* HelmholtzKernels.hpp:1390-1403 (prism): h0[i] = (1 + z0)/2 and
         * h1[k] = 2/(1 - z2), kept here in h0 and h3 */
        for (int i = 0; i < t->nq[0]; ++i) t->h0[i] = 0.5 * (1 + t->Z[0][i]);
        for (int k = 0; k < t->nq[2]; ++k) t->h3[k] = 2.0 / (1 - t->Z[2][k]);
    }
    if (shape == OR_PYR) {
        /* HelmholtzKernels.hpp:1405-1423 (pyramid): h0[i] = (1 + z0)/2,
         * h1[j] = (1 + z1)/2, h2[k] = 2/(1 - z2), kept here in h0, h1, h3 */
        for (int i = 0; i < t->nq[0]; ++i) t->h0[i] = 0.5 * (1 + t->Z[0][i]);
        for (int j = 0; j < t->nq[1]; ++j) t->h1[j] = 0.5 * (1 + t->Z[1][j]);
        for (int k = 0; k < t->nq[2]; ++k) t->h3[k] = 2.0 / (1 - t->Z[2][k]);
    }
}

static void tables_free(tables_t *t)
{
    for (int d = 0; d < 3; ++d) {
        free(t->b[d]); free(t->db[d]); free(t->D[d]); free(t->w[d]); free(t->Z[d]);
    }
}

/* ------------------------------------------------------------------ */
/* BwdTransKernels.hpp:160-227 */
static void bwd_hex(int nm0, int nm1, int nm2, int nq0, int nq1, int nq2,
                    const vd *in, const double *b0, const double *b1,
                    const double *b2, vd *sum_irq, vd *sum_jir, vd *out)
{
    for (int i = 0, cnt_irq = 0; i < nq0; ++i)
        for (int r = 0, cnt_rqp = 0; r < nm2; ++r)
            for (int q = 0; q < nm1; ++q, ++cnt_irq) {
                vd tmp = vmuls(in[cnt_rqp], b0[i]);
                ++cnt_rqp;
                for (int p = 1; p < nm0; ++p, ++cnt_rqp)
                    tmp = vfmas(tmp, in[cnt_rqp], b0[p * nq0 + i]);
                sum_irq[cnt_irq] = tmp;
            }
    for (int j = 0, cnt_jir = 0; j < nq1; ++j)
        for (int i = 0, cnt_irq = 0; i < nq0; ++i)
            for (int r = 0; r < nm2; ++r, ++cnt_jir) {
                vd tmp = vmuls(sum_irq[cnt_irq], b1[j]);
                ++cnt_irq;
                for (int q = 1; q < nm1; ++q)
                    tmp = vfmas(tmp, sum_irq[cnt_irq++], b1[q * nq1 + j]);
                sum_jir[cnt_jir] = tmp;
            }
    for (int k = 0, cnt_kji = 0; k < nq2; ++k)
        for (int j = 0, cnt_jir = 0; j < nq1; ++j)
            for (int i = 0; i < nq0; ++i, ++cnt_kji) {
                vd tmp = vmuls(sum_jir[cnt_jir], b2[k]);
                ++cnt_jir;
                for (int r = 1; r < nm2; ++r)
                    tmp = vfmas(tmp, sum_jir[cnt_jir++], b2[r * nq2 + k]);
                out[cnt_kji] = tmp;
            }
}

/* BwdTransKernels.hpp:231-330 */
static void bwd_tet(int nm0, int nm1, int nm2, int nq0, int nq1, int nq2,
                    int correct, const vd *in, const double *b0,
                    const double *b1, const double *b2, vd *fpq, vd *fp,
                    vd *out)
{
    for (int k = 0, cnt_kji = 0; k < nq2; ++k) {
        int cnt_pq = 0, cnt_pqr = 0, mode = 0;
        for (int p = 0; p < nm0; ++p) {
            for (int q = 0; q < nm1 - p; ++q, ++cnt_pq) {
                vd prod = vmuls(in[cnt_pqr], b2[k + nq2 * mode]);
                ++mode;
                ++cnt_pqr;
                for (int r = 1; r < nm2 - p - q; ++r, ++mode, ++cnt_pqr)
                    prod = vfmas(prod, in[cnt_pqr], b2[k + nq2 * mode]);
                fpq[cnt_pq] = prod;
            }
            for (int q = nm1 - p; q < nm2 - p; ++q) mode += nm2 - p - q;
        }
        for (int j = 0; j < nq1; ++j) {
            mode = cnt_pq = 0;
            for (int p = 0; p < nm0; ++p) {
                vd prod = vmuls(fpq[cnt_pq], b1[mode * nq1 + j]);
                ++cnt_pq;
                for (int q = 1; q < nm1 - p; ++q, ++cnt_pq)
                    prod = vfmas(prod, fpq[cnt_pq], b1[(mode + q) * nq1 + j]);
                fp[p] = prod;
                mode += nm1 - p;
            }
            for (int i = 0; i < nq0; ++i, ++cnt_kji) {
                vd tmp = vmuls(fp[0], b0[i]);
                for (int p = 1; p < nm0; ++p) tmp = vfmas(tmp, fp[p], b0[p * nq0 + i]);
                if (correct) {
                    /* top vertex */
                    double t1 = b0[i] * b1[nq1 + j];
                    t1 = fma(b0[nq0 + i], b1[j], t1);
                    t1 = fma(b0[nq0 + i], b1[nq1 + j], t1);
                    t1 = t1 * b2[nq2 + k];
                    tmp = vfmas(tmp, in[1], t1);
                    /* bottom vertex */
                    t1 = b0[nq0 + i] * b1[nq1 + j];
                    t1 = t1 * b2[k];
                    tmp = vfmas(tmp, in[nm2], t1);
                    /* singular edge */
                    for (int r = 1; r < nm2 - 1; ++r) {
                        t1 = b1[nq1 + j] * b0[nq0 + i];
                        t1 = t1 * b2[(r + 1) * nq2 + k];
                        tmp = vfmas(tmp, in[nm2 + r], t1);
                    }
                }
                out[cnt_kji] = tmp;
            }
        }
    }
}

/* BwdTransKernels.hpp:334-405 (BwdTransPrismKernel): B2 indexed by the
 * (p, r) pair, the collapsed-edge correction adds mode (0, q, 1) times
 * B0_1 B1_q B2_1 */
static void bwd_prism(int nm0, int nm1, int nm2, int nq0, int nq1, int nq2,
                      int correct, const vd *in, const double *b0,
                      const double *b1, const double *b2, vd *fpq, vd *fp,
                      vd *out)
{
    for (int k = 0, cnt_kji = 0; k < nq2; ++k) {
        int mode_pqr = 0, mode_pq = 0, mode_pr = 0;
        for (int p = 0; p < nm0; ++p) {
            for (int q = 0; q < nm1; ++q, ++mode_pq) {
                vd prod = vz();
                for (int r = 0; r < nm2 - p; ++r, ++mode_pqr)
                    prod = vfmas(prod, in[mode_pqr], b2[(mode_pr + r) * nq2 + k]);
                fpq[mode_pq] = prod;
            }
            mode_pr += nm2 - p;
        }
        for (int j = 0; j < nq1; ++j) {
            mode_pq = 0;
            for (int p = 0; p < nm0; ++p) {
                vd prod = vz();
                for (int q = 0; q < nm1; ++q, ++mode_pq) prod = vfmas(prod, fpq[mode_pq], b1[q * nq1 + j]);
                fp[p] = prod;
            }
            for (int i = 0; i < nq0; ++i, ++cnt_kji) {
                vd val = vz();
                for (int p = 0; p < nm0; ++p) val = vfmas(val, fp[p], b0[p * nq0 + i]);
                if (correct) {
                    const double basis_2 = b2[nq2 + k], basis_0 = b0[nq0 + i];
                    for (int q = 0; q < nm1; ++q) {
                        const double basis_1 = b1[q * nq1 + j];
                        val = vfma(val, vs(basis_2 * basis_1), vmuls(in[q * nm2 + 1], basis_0));
                    }
                }
                out[cnt_kji] = val;
            }
        }
    }
}

/* BwdTransKernels.hpp:407-489 (BwdTransPyrKernel): B2 indexed by the mode
 * (p, q, r); the top vertex (coefficient 1) also carries
 * (B0_0 B1_1 + B0_1 B1_0 + B0_1 B1_1) B2_1 */
static void bwd_pyr(int nm0, int nm1, int nm2, int nq0, int nq1, int nq2,
                    int correct, const vd *in, const double *b0,
                    const double *b1, const double *b2, vd *fpq, vd *fp,
                    vd *out)
{
    for (int k = 0, cnt_kji = 0; k < nq2; ++k) {
        int cnt_pqr = 0, mode_pqr = 0, mode_pq = 0;
        for (int p = 0; p < nm0; ++p) {
            for (int q = 0; q < p; ++q, ++mode_pq) {
                vd prod = vz();
                for (int r = 0; r < nm2 - p; ++r, ++mode_pqr, ++cnt_pqr)
                    prod = vfmas(prod, in[cnt_pqr], b2[mode_pqr * nq2 + k]);
                fpq[mode_pq] = prod;
            }
            for (int q = p; q < nm1; ++q, ++mode_pq) {
                vd prod = vz();
                for (int r = 0; r < nm2 - q; ++r, ++mode_pqr, ++cnt_pqr)
                    prod = vfmas(prod, in[cnt_pqr], b2[mode_pqr * nq2 + k]);
                fpq[mode_pq] = prod;
            }
            for (int q = nm1; q < nm2; ++q) mode_pqr += nm2 - q;
        }
        for (int j = 0; j < nq1; ++j) {
            mode_pq = 0;
            for (int p = 0; p < nm0; ++p) {
                vd prod = vz();
                for (int q = 0; q < nm1; ++q, ++mode_pq) prod = vfmas(prod, fpq[mode_pq], b1[q * nq1 + j]);
                fp[p] = prod;
            }
            for (int i = 0; i < nq0; ++i, ++cnt_kji) {
                vd val = vz();
                for (int p = 0; p < nm0; ++p) val = vfmas(val, fp[p], b0[p * nq0 + i]);
                if (correct) {
                    double t = b0[i] * b1[nq1 + j];
                    t = fma(b0[nq0 + i], b1[j], t);
                    t = fma(b0[nq0 + i], b1[nq1 + j], t);
                    t = t * b2[nq2 + k];
                    val = vfmas(val, in[1], t);
                }
                out[cnt_kji] = val;
            }
        }
    }
}

/* IProductKernels.hpp:44-64 (ScaleAppend) */
static inline void scale_append(int SCALE, int APPEND, vd *store, vd pos,
                                double scale)
{
    if (SCALE && APPEND) *store = vfmas(*store, pos, scale);
    else if (APPEND) *store = vadd(*store, pos);
    else if (SCALE) *store = vmuls(pos, scale);
    else *store = pos;
}

/* IProductKernels.hpp:270-350 */
static void iprod_hex(int SCALE, int APPEND, int DEF, int nm0, int nm1,
                      int nm2, int nq0, int nq1, int nq2, const vd *in,
                      const double *b0, const double *b1, const double *b2,
                      const double *w0, const double *w1, const double *w2,
                      const vd *jac, vd *sums_kj, vd *sums_k, vd *out,
                      double scale)
{
    for (int p = 0; p < nm0; ++p) {
        int cnt_kji = 0, cnt_kj = 0;
        for (int k = 0; k < nq2; ++k)
            for (int j = 0; j < nq1; ++j, ++cnt_kj) {
                vd sum_kj = vz();
                for (int i = 0; i < nq0; ++i, ++cnt_kji) {
                    vd jv = DEF ? jac[nq0 * nq1 * k + nq0 * j + i] : jac[0];
                    vd prod = vmul(vmuls(in[cnt_kji], b0[i + nq0 * p]), jv);
                    sum_kj = vfmas(sum_kj, prod, w0[i]);
                }
                sums_kj[cnt_kj] = sum_kj;
            }
        for (int q = 0; q < nm1; ++q) {
            cnt_kj = 0;
            for (int k = 0; k < nq2; ++k) {
                vd sum_k = vz();
                for (int j = 0; j < nq1; ++j, ++cnt_kj) {
                    vd prod = vmuls(sums_kj[cnt_kj], b1[q * nq1 + j]);
                    sum_k = vfmas(sum_k, prod, w1[j]);
                }
                sums_k[k] = sum_k;
            }
            for (int r = 0; r < nm2; ++r) {
                vd sum = vz();
                for (int k = 0; k < nq2; ++k) {
                    vd prod = vmuls(sums_k[k], b2[r * nq2 + k]);
                    sum = vfmas(sum, prod, w2[k]);
                }
                scale_append(SCALE, APPEND, &out[r * nm0 * nm1 + q * nm0 + p], sum, scale);
            }
        }
    }
}

/* IProductKernels.hpp:354-514 */
static void iprod_tet(int SCALE, int APPEND, int DEF, int nm0, int nm1,
                      int nm2, int nq0, int nq1, int nq2, int correct,
                      const vd *in, const double *b0, const double *b1,
                      const double *b2, const double *w0, const double *w1,
                      const double *w2, const vd *jac, vd *sums_kj,
                      vd *sums_k, vd *out, double scale)
{
    for (int p = 0, mode = 0, mode2 = 0, cnt_pqr = 0; p < nm0; ++p) {
        int cnt_kji = 0;
        for (int k = 0, cnt_kj = 0; k < nq2; ++k)
            for (int j = 0; j < nq1; ++j, ++cnt_kj) {
                vd jv = DEF ? jac[nq0 * nq1 * k + nq0 * j] : jac[0];
                vd sum_kj = vmuls(vmul(vmuls(in[cnt_kji], b0[nq0 * p]), jv), w0[0]);
                ++cnt_kji;
                for (int i = 1; i < nq0; ++i, ++cnt_kji) {
                    jv = DEF ? jac[nq0 * nq1 * k + nq0 * j + i] : jac[0];
                    vd inxmm = vmul(vmuls(in[cnt_kji], b0[i + nq0 * p]), jv);
                    sum_kj = vfmas(sum_kj, inxmm, w0[i]);
                }
                sums_kj[cnt_kj] = sum_kj;
            }
        for (int q = 0; q < nm1 - p; ++q, ++mode) {
            int cnt_kj = 0;
            for (int k = 0; k < nq2; ++k) {
                vd sum_k = vmuls(vmuls(sums_kj[cnt_kj], b1[mode * nq1]), w1[0]);
                ++cnt_kj;
                for (int j = 1; j < nq1; ++j, ++cnt_kj) {
                    vd tmp2 = vmuls(sums_kj[cnt_kj], b1[mode * nq1 + j]);
                    sum_k = vfmas(sum_k, tmp2, w1[j]);
                }
                sums_k[k] = sum_k;
            }
            for (int r = 0; r < nm2 - p - q; ++r, ++mode2, ++cnt_pqr) {
                vd tmp = vmuls(vmuls(sums_k[0], b2[mode2 * nq2]), w2[0]);
                for (int k = 1; k < nq2; ++k) {
                    vd tmp2 = vmuls(sums_k[k], b2[mode2 * nq2 + k]);
                    tmp = vfmas(tmp, tmp2, w2[k]);
                }
                scale_append(SCALE, APPEND, &out[cnt_pqr], tmp, scale);
            }
        }
        for (int q = nm1 - p; q < nm2 - p; ++q) mode2 += nm2 - p - q;
    }
    if (correct) {
        for (int k = 0, cnt = 0; k < nq2; ++k) {
            vd tQ2 = vs(w2[k]);
            if (!DEF) tQ2 = vmul(tQ2, jac[0]);
            for (int j = 0; j < nq1; ++j) {
                vd tQ1 = vmuls(tQ2, w1[j]);
                for (int i = 0; i < nq0; ++i, ++cnt) {
                    vd tQ = vmuls(tQ1, w0[i]);
                    vd tin = in[cnt];
                    if (DEF) tQ = vmul(tQ, jac[k * nq0 * nq1 + j * nq0 + i]);
                    /* top vertex */
                    double t = b0[i] * b1[nq1 + j];
                    t = fma(b0[nq0 + i], b1[j], t);
                    t = fma(b0[nq0 + i], b1[nq1 + j], t);
                    t = t * b2[nq2 + k];
                    vd tv = vmuls(tin, t);
                    scale_append(SCALE, 1, &out[1], vmul(tv, tQ), scale);
                    /* bottom vertex */
                    vd tb = vmul(vs(b0[nq0 + i] * b1[nq1 + j] * b2[k]), tin);
                    scale_append(SCALE, 1, &out[nm2], vmul(tb, tQ), scale);
                    /* singular edge */
                    for (int r = 1; r < nm2 - 1; ++r) {
                        vd te = vmul(vs(b2[(r + 1) * nq2 + k] * b1[nq1 + j] * b0[nq0 + i]), tin);
                        scale_append(SCALE, 1, &out[nm2 + r], vmul(te, tQ), scale);
                    }
                }
            }
        }
    }
}

/* IProductKernels.hpp:519-655 (IProductPrismKernel) */
static void iprod_prism(int SCALE, int APPEND, int DEF, int nm0, int nm1,
                        int nm2, int nq0, int nq1, int nq2, int correct,
                        const vd *in, const double *b0, const double *b1,
                        const double *b2, const double *w0, const double *w1,
                        const double *w2, const vd *jac, vd *sums_kj,
                        vd *sums_k, vd *out, double scale)
{
    int mode_pr = 0, mode_pqr = 0;
    for (int p = 0; p < nm0; ++p) {
        int cnt_kji = 0, cnt_kj = 0;
        for (int k = 0; k < nq2; ++k)
            for (int j = 0; j < nq1; ++j, ++cnt_kj) {
                vd sum_kj = vz();
                for (int i = 0; i < nq0; ++i, ++cnt_kji) {
                    vd jv = DEF ? jac[nq0 * nq1 * k + nq0 * j + i] : jac[0];
                    vd prod = vmuls(vmuls(jv, b0[nq0 * p + i]), w0[i]);
                    sum_kj = vfma(sum_kj, prod, in[cnt_kji]);
                }
                sums_kj[cnt_kj] = sum_kj;
            }
        for (int q = 0; q < nm1; ++q) {
            cnt_kj = 0;
            for (int k = 0; k < nq2; ++k) {
                vd sum_k = vz();
                for (int j = 0; j < nq1; ++j, ++cnt_kj)
                    sum_k = vfmas(sum_k, sums_kj[cnt_kj], b1[q * nq1 + j] * w1[j]);
                sums_k[k] = sum_k;
            }
            for (int r = 0; r < nm2 - p; ++r, ++mode_pqr) {
                vd sum_k = vz();
                for (int k = 0; k < nq2; ++k) sum_k = vfmas(sum_k, sums_k[k], b2[(mode_pr + r) * nq2 + k] * w2[k]);
                scale_append(SCALE, APPEND, &out[mode_pqr], sum_k, scale);
            }
        }
        mode_pr += nm2 - p;
    }
    if (correct) {
        /* singular edge: phi_1q1 added to phi_0q1 */
        vd corr[64];
        for (int q = 0; q < nm1; ++q) corr[q] = vz();
        for (int k = 0, cnt = 0; k < nq2; ++k) {
            vd kw = vs(w2[k]);
            if (!DEF) kw = vmul(kw, jac[0]);
            for (int j = 0; j < nq1; ++j) {
                vd kjw = vmuls(kw, w1[j]);
                for (int i = 0; i < nq0; ++i, ++cnt) {
                    vd prod = vmul(vmuls(kjw, w0[i]), in[cnt]);
                    if (DEF) prod = vmul(prod, jac[k * nq1 * nq0 + j * nq0 + i]);
                    const double basis_2 = b2[nq2 + k], basis_0 = b0[nq0 + i];
                    for (int q = 0; q < nm1; ++q)
                        corr[q] = vfma(corr[q], vs(basis_2 * b1[q * nq1 + j]), vmuls(prod, basis_0));
                }
            }
        }
        for (int q = 0; q < nm1; ++q) scale_append(SCALE, 1, &out[nm2 * q + 1], corr[q], scale);
    }
}

/* IProductKernels.hpp:660-815 (IProductPyrKernel) */
static void iprod_pyr(int SCALE, int APPEND, int DEF, int nm0, int nm1,
                      int nm2, int nq0, int nq1, int nq2, int correct,
                      const vd *in, const double *b0, const double *b1,
                      const double *b2, const double *w0, const double *w1,
                      const double *w2, const vd *jac, vd *sums_kj,
                      vd *sums_k, vd *out, double scale)
{
    int mode2 = 0, mode_pqr = 0;
    for (int p = 0; p < nm0; ++p) {
        int cnt_kji = 0, cnt_kj = 0;
        for (int k = 0; k < nq2; ++k)
            for (int j = 0; j < nq1; ++j, ++cnt_kj) {
                vd sum_kj = vz();
                for (int i = 0; i < nq0; ++i, ++cnt_kji) {
                    vd jv = DEF ? jac[nq0 * nq1 * k + nq0 * j + i] : jac[0];
                    vd prod = vmuls(vmuls(jv, b0[nq0 * p + i]), w0[i]);
                    sum_kj = vfma(sum_kj, prod, in[cnt_kji]);
                }
                sums_kj[cnt_kj] = sum_kj;
            }
        for (int q = 0; q < nm1; ++q) {
            cnt_kj = 0;
            for (int k = 0; k < nq2; ++k) {
                vd sum_k = vz();
                for (int j = 0; j < nq1; ++j, ++cnt_kj)
                    sum_k = vfmas(sum_k, sums_kj[cnt_kj], b1[q * nq1 + j] * w1[j]);
                sums_k[k] = sum_k;
            }
            const int nr = nm2 - (q < p ? p : q);
            for (int r = 0; r < nr; ++r, ++mode2, ++mode_pqr) {
                vd sum_k = vz();
                for (int k = 0; k < nq2; ++k) sum_k = vfmas(sum_k, sums_k[k], b2[mode2 * nq2 + k] * w2[k]);
                scale_append(SCALE, APPEND, &out[mode_pqr], sum_k, scale);
            }
        }
        for (int q = nm1; q < nm2; ++q) mode2 += nm2 - q;
    }
    if (correct) {
        /* top vertex, appended point by point as the reference does */
        for (int k = 0, cnt = 0; k < nq2; ++k) {
            vd tq2 = vs(w2[k]);
            if (!DEF) tq2 = vmul(tq2, jac[0]);
            for (int j = 0; j < nq1; ++j) {
                vd tq1 = vmuls(tq2, w1[j]);
                for (int i = 0; i < nq0; ++i, ++cnt) {
                    vd tq = vmuls(tq1, w0[i]);
                    if (DEF) tq = vmul(tq, jac[k * nq0 * nq1 + j * nq0 + i]);
                    double t = b0[i] * b1[nq1 + j];
                    t = fma(b0[nq0 + i], b1[j], t);
                    t = fma(b0[nq0 + i], b1[nq1 + j], t);
                    t = t * b2[nq2 + k];
                    scale_append(SCALE, 1, &out[1], vmul(vmuls(in[cnt], t), tq), scale);
                }
            }
        }
    }
}

/* PhysDerivKernels.hpp:436-507 (column-major D: D[k*nq+i]) */
static void deriv_tensor(int nq0, int nq1, int nq2, const vd *in,
                         const double *D0, const double *D1, const double *D2,
                         vd *d0, vd *d1, vd *d2)
{
    for (int i = 0; i < nq0; ++i)
        for (int j = 0; j < nq1 * nq2; ++j) {
            vd s = vz();
            for (int k = 0; k < nq0; ++k) s = vfmas(s, in[j * nq0 + k], D0[k * nq0 + i]);
            d0[j * nq0 + i] = s;
        }
    for (int blk = 0; blk < nq2; ++blk) {
        int start = blk * nq0 * nq1;
        for (int i = 0; i < nq0; ++i)
            for (int j = 0; j < nq1; ++j) {
                vd s = vz();
                for (int k = 0; k < nq1; ++k) s = vfmas(s, in[start + k * nq0 + i], D1[k * nq1 + j]);
                d1[start + j * nq0 + i] = s;
            }
    }
    for (int i = 0; i < nq0 * nq1; ++i)
        for (int j = 0; j < nq2; ++j) {
            vd s = vz();
            for (int k = 0; k < nq2; ++k) s = vfmas(s, in[k * nq0 * nq1 + i], D2[k * nq2 + j]);
            d2[j * nq0 * nq1 + i] = s;
        }
}

/* PhysDerivKernels.hpp:184-345 (PhysDeriv3DKernel, hex and tet) */
static void physderiv_chain(int shape, int DEF, int nq0, int nq1, int nq2,
                            const double *Z0, const double *Z1,
                            const double *Z2, const vd *df, vd *wsp0,
                            vd *wsp1, vd *d0, vd *d1, vd *d2)
{
    if (shape == OR_TET) {
        for (int k = 0, e = 0; k < nq2; ++k) {
            double x2 = 2.0 / (1.0 - Z2[k]);
            for (int j = 0; j < nq1; ++j) {
                double x1 = 2.0 / (1.0 - Z1[j]);
                double xf = x1 * x2;
                for (int i = 0; i < nq0; ++i, ++e) {
                    vd v = vmuls(d0[e], xf);
                    d0[e] = v;
                    wsp0[e] = v;
                }
            }
        }
        for (int k = 0, e = 0; k < nq2; ++k) {
            double x2 = 2.0 / (1.0 - Z2[k]);
            for (int j = 0; j < nq1; ++j)
                for (int i = 0; i < nq0; ++i, ++e) {
                    double x0 = 0.5 * (1.0 + Z0[i]);
                    vd o0 = vmuls(wsp0[e], x0);
                    wsp0[e] = o0;
                    vd v1 = vmuls(d1[e], x2);
                    d1[e] = vadd(o0, v1);
                    wsp1[e] = v1;
                }
        }
        for (int k = 0, e = 0; k < nq2; ++k)
            for (int j = 0; j < nq1; ++j) {
                double x1 = 0.5 * (1.0 + Z1[j]);
                for (int i = 0; i < nq0; ++i, ++e) {
                    vd o = vfmas(wsp0[e], wsp1[e], x1);
                    d2[e] = vadd(o, d2[e]);
                }
            }
    }
    if (shape == OR_PYR) {
        /* PhysDerivKernels.hpp:301-311: d0, d1 *= 2/(1 - z2),
         * d2 += (1 + z0)/2 d0 + (1 + z1)/2 d1 */
        for (int k = 0, e = 0; k < nq2; ++k) {
            const double x2 = 2.0 / (1.0 - Z2[k]);
            for (int j = 0; j < nq1; ++j) {
                const double x1 = 0.5 * (1.0 + Z1[j]);
                for (int i = 0; i < nq0; ++i, ++e) {
                    d0[e] = vmuls(d0[e], x2);
                    d1[e] = vmuls(d1[e], x2);
                    d2[e] = vfmas(d2[e], d0[e], 0.5 * (1.0 + Z0[i]));
                    d2[e] = vfmas(d2[e], d1[e], x1);
                }
            }
        }
    }
    if (shape == OR_PRISM) {
        /* PhysDerivKernels.hpp:293-301: d0 *= 2/(1 - z2), d2 += (1 + z0)/2 d0 */
        for (int k = 0, e = 0; k < nq2; ++k) {
            const double x2 = 2.0 / (1.0 - Z2[k]);
            for (int j = 0; j < nq1; ++j)
                for (int i = 0; i < nq0; ++i, ++e) {
                    d0[e] = vmuls(d0[e], x2);
                    d2[e] = vfmas(d2[e], d0[e], 0.5 * (1.0 + Z0[i]));
                }
        }
    }
    vd f[9];
    if (!DEF) for (int d = 0; d < 9; ++d) f[d] = df[d];
    const int nqt = nq0 * nq1 * nq2;
    for (int c = 0; c < nqt; ++c) {
        vd a0 = d0[c], a1 = d1[c], a2 = d2[c];
        if (DEF) for (int d = 0; d < 9; ++d) f[d] = df[c * 9 + d];
        vd o0 = vmul(a0, f[0]); o0 = vfma(o0, a1, f[1]); o0 = vfma(o0, a2, f[2]);
        vd o1 = vmul(a0, f[3]); o1 = vfma(o1, a1, f[4]); o1 = vfma(o1, a2, f[5]);
        vd o2 = vmul(a0, f[6]); o2 = vfma(o2, a1, f[7]); o2 = vfma(o2, a2, f[8]);
        d0[c] = o0; d1[c] = o1; d2[c] = o2;
    }
}

/* metric m_de from df, optionally through a constant diffusion tensor:
 * HelmholtzKernels.hpp:440-525 */
static void hex_metric(const vd *f, int cvd, const double *dd, vd m[6])
{
    if (!cvd) {
        m[0] = vmul(f[0], f[0]); m[0] = vfma(m[0], f[3], f[3]); m[0] = vfma(m[0], f[6], f[6]);
        m[1] = vmul(f[0], f[1]); m[1] = vfma(m[1], f[3], f[4]); m[1] = vfma(m[1], f[6], f[7]);
        m[2] = vmul(f[0], f[2]); m[2] = vfma(m[2], f[3], f[5]); m[2] = vfma(m[2], f[6], f[8]);
        m[3] = vmul(f[1], f[1]); m[3] = vfma(m[3], f[4], f[4]); m[3] = vfma(m[3], f[7], f[7]);
        m[4] = vmul(f[1], f[2]); m[4] = vfma(m[4], f[4], f[5]); m[4] = vfma(m[4], f[7], f[8]);
        m[5] = vmul(f[2], f[2]); m[5] = vfma(m[5], f[5], f[5]); m[5] = vfma(m[5], f[8], f[8]);
        return;
    }
    const double d00 = dd[0], d01 = dd[1], d11 = dd[2], d02 = dd[3], d12 = dd[4], d22 = dd[5];
    vd t0 = vmuls(f[0], d00); t0 = vfmas(t0, f[3], d01); t0 = vfmas(t0, f[6], d02);
    vd t1 = vmuls(f[0], d01); t1 = vfmas(t1, f[3], d11); t1 = vfmas(t1, f[6], d12);
    vd t2 = vmuls(f[0], d02); t2 = vfmas(t2, f[3], d12); t2 = vfmas(t2, f[6], d22);
    vd t3 = vmuls(f[1], d00); t3 = vfmas(t3, f[4], d01); t3 = vfmas(t3, f[7], d02);
    vd t4 = vmuls(f[1], d01); t4 = vfmas(t4, f[4], d11); t4 = vfmas(t4, f[7], d12);
    vd t5 = vmuls(f[1], d02); t5 = vfmas(t5, f[4], d12); t5 = vfmas(t5, f[7], d22);
    vd t6 = vmuls(f[2], d00); t6 = vfmas(t6, f[5], d01); t6 = vfmas(t6, f[8], d02);
    vd t7 = vmuls(f[2], d01); t7 = vfmas(t7, f[5], d11); t7 = vfmas(t7, f[8], d12);
    vd t8 = vmuls(f[2], d02); t8 = vfmas(t8, f[5], d12); t8 = vfmas(t8, f[8], d22);
    m[0] = vmul(t0, f[0]); m[0] = vfma(m[0], t1, f[3]); m[0] = vfma(m[0], t2, f[6]);
    m[1] = vmul(t0, f[1]); m[1] = vfma(m[1], t1, f[4]); m[1] = vfma(m[1], t2, f[7]);
    m[2] = vmul(t0, f[2]); m[2] = vfma(m[2], t1, f[5]); m[2] = vfma(m[2], t2, f[8]);
    m[3] = vmul(t3, f[1]); m[3] = vfma(m[3], t4, f[4]); m[3] = vfma(m[3], t5, f[7]);
    m[4] = vmul(t3, f[2]); m[4] = vfma(m[4], t4, f[5]); m[4] = vfma(m[4], t5, f[8]);
    m[5] = vmul(t6, f[2]); m[5] = vfma(m[5], t7, f[5]); m[5] = vfma(m[5], t8, f[8]);
}

/* HelmholtzKernels.hpp:389-696 (DiffusionCoeffHexKernel, no variable
 * diffusion).  m = {00,01,02,11,12,22}. */
static void diffusion_hex(int DEF, int nqt, int cvd, const double *dd,
                          const vd *df, vd *d0, vd *d1, vd *d2)
{
    vd m[6], f[9];
    if (!DEF) {
        for (int d = 0; d < 9; ++d) f[d] = df[d];
        hex_metric(f, cvd, dd, m);
    }
    for (int c = 0; c < nqt; ++c) {
        if (DEF) {
            for (int d = 0; d < 9; ++d) f[d] = df[c * 9 + d];
            hex_metric(f, cvd, dd, m);
        }
        vd a0 = d0[c], a1 = d1[c], a2 = d2[c];
        vd t = vmul(m[0], a0); t = vfma(t, m[1], a1); t = vfma(t, m[2], a2); d0[c] = t;
        t = vmul(m[1], a0); t = vfma(t, m[3], a1); t = vfma(t, m[4], a2); d1[c] = t;
        t = vmul(m[2], a0); t = vfma(t, m[4], a1); t = vfma(t, m[5], a2); d2[c] = t;
    }
}

/* HelmholtzKernels.hpp:700-913 (DiffusionCoeffTetKernel, no variable
 * diffusion) */
static void diffusion_tet(int DEF, const tables_t *T, int cvd,
                          const double *dd, const vd *df, vd *d0, vd *d1,
                          vd *d2)
{
    const int nq0 = T->nq[0], nq1 = T->nq[1], nq2 = T->nq[2];
    double d00 = 1, d01 = 0, d11 = 1, d02 = 0, d12 = 0, d22 = 1;
    if (cvd) { d00 = dd[0]; d01 = dd[1]; d11 = dd[2]; d02 = dd[3]; d12 = dd[4]; d22 = dd[5]; }
    vd f[9];
    if (!DEF) for (int d = 0; d < 9; ++d) f[d] = df[d];
    for (int k = 0, cnt = 0; k < nq2; ++k) {
        double h3 = T->h3[k];
        for (int j = 0; j < nq1; ++j) {
            double h1 = T->h1[j], h2 = T->h2[j];
            double h2h3 = h2 * h3, h1h3 = h1 * h3;
            for (int i = 0; i < nq0; ++i, ++cnt) {
                if (DEF) for (int d = 0; d < 9; ++d) f[d] = df[cnt * 9 + d];
                double h0h2h3 = T->h0[i] * h2h3;
                vd t1 = vmuls(vadd(f[1], f[2]), h0h2h3); t1 = vfmas(t1, f[0], h2h3);
                vd t2 = vmuls(vadd(f[4], f[5]), h0h2h3); t2 = vfmas(t2, f[3], h2h3);
                vd t3 = vmuls(vadd(f[7], f[8]), h0h2h3); t3 = vfmas(t3, f[6], h2h3);
                vd t4 = vmuls(f[1], h3); t4 = vfmas(t4, f[2], h1h3);
                vd t5 = vmuls(f[4], h3); t5 = vfmas(t5, f[5], h1h3);
                vd t6 = vmuls(f[7], h3); t6 = vfmas(t6, f[8], h1h3);
                vd g0, g1, g2, g3, g4, g5;
                if (!cvd) {
                    g0 = vmul(t1, t1); g0 = vfma(g0, t2, t2); g0 = vfma(g0, t3, t3);
                    g4 = vmul(f[2], t1); g4 = vfma(g4, f[5], t2); g4 = vfma(g4, f[8], t3);
                    g3 = vmul(t1, t4); g3 = vfma(g3, t2, t5); g3 = vfma(g3, t3, t6);
                    g1 = vmul(t4, t4); g1 = vfma(g1, t5, t5); g1 = vfma(g1, t6, t6);
                    g5 = vmul(f[2], t4); g5 = vfma(g5, f[5], t5); g5 = vfma(g5, f[8], t6);
                    g2 = vmul(f[2], f[2]); g2 = vfma(g2, f[5], f[5]); g2 = vfma(g2, f[8], f[8]);
                } else {
                    vd s0 = vmuls(t1, d00); s0 = vfmas(s0, t2, d01); s0 = vfmas(s0, t3, d02);
                    vd s1 = vmuls(t1, d01); s1 = vfmas(s1, t2, d11); s1 = vfmas(s1, t3, d12);
                    vd s2 = vmuls(t1, d02); s2 = vfmas(s2, t2, d12); s2 = vfmas(s2, t3, d22);
                    vd s3 = vmuls(t4, d00); s3 = vfmas(s3, t5, d01); s3 = vfmas(s3, t6, d02);
                    vd s4 = vmuls(t4, d01); s4 = vfmas(s4, t5, d11); s4 = vfmas(s4, t6, d12);
                    vd s5 = vmuls(t4, d02); s5 = vfmas(s5, t5, d12); s5 = vfmas(s5, t6, d22);
                    vd s6 = vmuls(f[2], d00); s6 = vfmas(s6, f[5], d01); s6 = vfmas(s6, f[8], d02);
                    vd s7 = vmuls(f[2], d01); s7 = vfmas(s7, f[5], d11); s7 = vfmas(s7, f[8], d12);
                    vd s8 = vmuls(f[2], d02); s8 = vfmas(s8, f[5], d12); s8 = vfmas(s8, f[8], d22);
                    g0 = vmul(s0, t1); g0 = vfma(g0, s1, t2); g0 = vfma(g0, s2, t3);
                    g3 = vmul(s0, t4); g3 = vfma(g3, s1, t5); g3 = vfma(g3, s2, t6);
                    g4 = vmul(s0, f[2]); g4 = vfma(g4, s1, f[5]); g4 = vfma(g4, s2, f[8]);
                    g1 = vmul(s3, t4); g1 = vfma(g1, s4, t5); g1 = vfma(g1, s5, t6);
                    g5 = vmul(s3, f[2]); g5 = vfma(g5, s4, f[5]); g5 = vfma(g5, s5, f[8]);
                    g2 = vmul(s6, f[2]); g2 = vfma(g2, s7, f[5]); g2 = vfma(g2, s8, f[8]);
                }
                vd a0 = d0[cnt], a1 = d1[cnt], a2 = d2[cnt];
                vd o = vmul(g0, a0); o = vfma(o, g3, a1); o = vfma(o, g4, a2); d0[cnt] = o;
                o = vmul(g3, a0); o = vfma(o, g1, a1); o = vfma(o, g5, a2); d1[cnt] = o;
                o = vmul(g4, a0); o = vfma(o, g5, a1); o = vfma(o, g2, a2); d2[cnt] = o;
            }
        }
    }
}

/* HelmholtzKernels.hpp:918-1119 (DiffusionCoeffPrismKernel, no variable
 * diffusion): tmp_c = h1 (h0 df[3c+2] + df[3c]) */
static void diffusion_prism(int DEF, const tables_t *T, int cvd,
                            const double *dd, const vd *df, vd *d0, vd *d1,
                            vd *d2)
{
    const int nq0 = T->nq[0], nq1 = T->nq[1], nq2 = T->nq[2];
    double d00 = 1, d01 = 0, d11 = 1, d02 = 0, d12 = 0, d22 = 1;
    if (cvd) { d00 = dd[0]; d01 = dd[1]; d11 = dd[2]; d02 = dd[3]; d12 = dd[4]; d22 = dd[5]; }
    vd f[9];
    if (!DEF) for (int d = 0; d < 9; ++d) f[d] = df[d];
    for (int k = 0, cnt = 0; k < nq2; ++k) {
        const double h1 = T->h3[k];
        for (int j = 0; j < nq1; ++j)
            for (int i = 0; i < nq0; ++i, ++cnt) {
                const double h0 = T->h0[i];
                if (DEF) for (int d = 0; d < 9; ++d) f[d] = df[cnt * 9 + d];
                vd t1 = vmuls(vfmas(f[0], f[2], h0), h1);
                vd t2 = vmuls(vfmas(f[3], f[5], h0), h1);
                vd t3 = vmuls(vfmas(f[6], f[8], h0), h1);
                vd g0, g1, g2, g3, g4, g5;
                if (!cvd) {
                    g0 = vmul(t1, t1); g0 = vfma(g0, t2, t2); g0 = vfma(g0, t3, t3);
                    g3 = vmul(f[1], t1); g3 = vfma(g3, f[4], t2); g3 = vfma(g3, f[7], t3);
                    g4 = vmul(f[2], t1); g4 = vfma(g4, f[5], t2); g4 = vfma(g4, f[8], t3);
                    g1 = vmul(f[1], f[1]); g1 = vfma(g1, f[4], f[4]); g1 = vfma(g1, f[7], f[7]);
                    g2 = vmul(f[2], f[2]); g2 = vfma(g2, f[5], f[5]); g2 = vfma(g2, f[8], f[8]);
                    g5 = vmul(f[1], f[2]); g5 = vfma(g5, f[4], f[5]); g5 = vfma(g5, f[7], f[8]);
                } else {
                    vd s0 = vmuls(t1, d00); s0 = vfmas(s0, t2, d01); s0 = vfmas(s0, t3, d02);
                    vd s1 = vmuls(t1, d01); s1 = vfmas(s1, t2, d11); s1 = vfmas(s1, t3, d12);
                    vd s2 = vmuls(t1, d02); s2 = vfmas(s2, t2, d12); s2 = vfmas(s2, t3, d22);
                    vd s3 = vmuls(f[1], d00); s3 = vfmas(s3, f[4], d01); s3 = vfmas(s3, f[7], d02);
                    vd s4 = vmuls(f[1], d01); s4 = vfmas(s4, f[4], d11); s4 = vfmas(s4, f[7], d12);
                    vd s5 = vmuls(f[1], d02); s5 = vfmas(s5, f[4], d12); s5 = vfmas(s5, f[7], d22);
                    vd s6 = vmuls(f[2], d00); s6 = vfmas(s6, f[5], d01); s6 = vfmas(s6, f[8], d02);
                    vd s7 = vmuls(f[2], d01); s7 = vfmas(s7, f[5], d11); s7 = vfmas(s7, f[8], d12);
                    vd s8 = vmuls(f[2], d02); s8 = vfmas(s8, f[5], d12); s8 = vfmas(s8, f[8], d22);
                    g0 = vmul(s0, t1); g0 = vfma(g0, s1, t2); g0 = vfma(g0, s2, t3);
                    g3 = vmul(s0, f[1]); g3 = vfma(g3, s1, f[4]); g3 = vfma(g3, s2, f[7]);
                    g4 = vmul(s0, f[2]); g4 = vfma(g4, s1, f[5]); g4 = vfma(g4, s2, f[8]);
                    g1 = vmul(s3, f[1]); g1 = vfma(g1, s4, f[4]); g1 = vfma(g1, s5, f[7]);
                    g5 = vmul(s3, f[2]); g5 = vfma(g5, s4, f[5]); g5 = vfma(g5, s5, f[8]);
                    g2 = vmul(s6, f[2]); g2 = vfma(g2, s7, f[5]); g2 = vfma(g2, s8, f[8]);
                }
                vd a0 = d0[cnt], a1 = d1[cnt], a2 = d2[cnt];
                vd o = vmul(g0, a0); o = vfma(o, g3, a1); o = vfma(o, g4, a2); d0[cnt] = o;
                o = vmul(g3, a0); o = vfma(o, g1, a1); o = vfma(o, g5, a2); d1[cnt] = o;
                o = vmul(g4, a0); o = vfma(o, g5, a1); o = vfma(o, g2, a2); d2[cnt] = o;
            }
    }
}

/* HelmholtzKernels.hpp:1119-1326 (DiffusionCoeffPyrKernel, no variable
 * diffusion): tmp0..2 = h2 df[0,3,6] + h0 h2 df[2,5,8],
 * tmp3..5 = h2 df[1,4,7] + h1 h2 df[2,5,8] */
static void diffusion_pyr(int DEF, const tables_t *T, int cvd,
                          const double *dd, const vd *df, vd *d0, vd *d1,
                          vd *d2)
{
    const int nq0 = T->nq[0], nq1 = T->nq[1], nq2 = T->nq[2];
    double d00 = 1, d01 = 0, d11 = 1, d02 = 0, d12 = 0, d22 = 1;
    if (cvd) { d00 = dd[0]; d01 = dd[1]; d11 = dd[2]; d02 = dd[3]; d12 = dd[4]; d22 = dd[5]; }
    vd f[9];
    if (!DEF) for (int d = 0; d < 9; ++d) f[d] = df[d];
    for (int k = 0, cnt = 0; k < nq2; ++k) {
        const double h2 = T->h3[k];
        for (int j = 0; j < nq1; ++j) {
            const double h1 = T->h1[j], h1h2 = h1 * h2;
            for (int i = 0; i < nq0; ++i, ++cnt) {
                const double h0 = T->h0[i], h0h2 = h0 * h2;
                if (DEF) for (int d = 0; d < 9; ++d) f[d] = df[cnt * 9 + d];
                vd t0 = vfmas(vmuls(f[0], h2), f[2], h0h2);
                vd t1 = vfmas(vmuls(f[3], h2), f[5], h0h2);
                vd t2 = vfmas(vmuls(f[6], h2), f[8], h0h2);
                vd t3 = vfmas(vmuls(f[1], h2), f[2], h1h2);
                vd t4 = vfmas(vmuls(f[4], h2), f[5], h1h2);
                vd t5 = vfmas(vmuls(f[7], h2), f[8], h1h2);
                vd g0, g1, g2, g3, g4, g5;
                if (!cvd) {
                    g0 = vmul(t0, t0); g0 = vfma(g0, t1, t1); g0 = vfma(g0, t2, t2);
                    g1 = vmul(t3, t3); g1 = vfma(g1, t4, t4); g1 = vfma(g1, t5, t5);
                    g2 = vmul(f[2], f[2]); g2 = vfma(g2, f[5], f[5]); g2 = vfma(g2, f[8], f[8]);
                    g3 = vmul(t0, t3); g3 = vfma(g3, t1, t4); g3 = vfma(g3, t2, t5);
                    g4 = vmul(f[2], t0); g4 = vfma(g4, f[5], t1); g4 = vfma(g4, f[8], t2);
                    g5 = vmul(f[2], t3); g5 = vfma(g5, f[5], t4); g5 = vfma(g5, f[8], t5);
                } else {
                    vd s0 = vmuls(t0, d00); s0 = vfmas(s0, t1, d01); s0 = vfmas(s0, t2, d02);
                    vd s1 = vmuls(t0, d01); s1 = vfmas(s1, t1, d11); s1 = vfmas(s1, t2, d12);
                    vd s2 = vmuls(t0, d02); s2 = vfmas(s2, t1, d12); s2 = vfmas(s2, t2, d22);
                    vd s3 = vmuls(t3, d00); s3 = vfmas(s3, t4, d01); s3 = vfmas(s3, t5, d02);
                    vd s4 = vmuls(t3, d01); s4 = vfmas(s4, t4, d11); s4 = vfmas(s4, t5, d12);
                    vd s5 = vmuls(t3, d02); s5 = vfmas(s5, t4, d12); s5 = vfmas(s5, t5, d22);
                    vd s6 = vmuls(f[2], d00); s6 = vfmas(s6, f[5], d01); s6 = vfmas(s6, f[8], d02);
                    vd s7 = vmuls(f[2], d01); s7 = vfmas(s7, f[5], d11); s7 = vfmas(s7, f[8], d12);
                    vd s8 = vmuls(f[2], d02); s8 = vfmas(s8, f[5], d12); s8 = vfmas(s8, f[8], d22);
                    g0 = vmul(s0, t0); g0 = vfma(g0, s1, t1); g0 = vfma(g0, s2, t2);
                    g3 = vmul(s0, t3); g3 = vfma(g3, s1, t4); g3 = vfma(g3, s2, t5);
                    g4 = vmul(s0, f[2]); g4 = vfma(g4, s1, f[5]); g4 = vfma(g4, s2, f[8]);
                    g1 = vmul(s3, t3); g1 = vfma(g1, s4, t4); g1 = vfma(g1, s5, t5);
                    g5 = vmul(s3, f[2]); g5 = vfma(g5, s4, f[5]); g5 = vfma(g5, s5, f[8]);
                    g2 = vmul(s6, f[2]); g2 = vfma(g2, s7, f[5]); g2 = vfma(g2, s8, f[8]);
                }
                vd a0 = d0[cnt], a1 = d1[cnt], a2 = d2[cnt];
                vd o = vmul(g0, a0); o = vfma(o, g3, a1); o = vfma(o, g4, a2); d0[cnt] = o;
                o = vmul(g3, a0); o = vfma(o, g1, a1); o = vfma(o, g5, a2); d1[cnt] = o;
                o = vmul(g4, a0); o = vfma(o, g5, a1); o = vfma(o, g2, a2); d2[cnt] = o;
            }
        }
    }
}

static void diffusion(int DEF, const tables_t *T, int cvd, const double *dd, const vd *df, vd *d0, vd *d1,
                      vd *d2)
{
    if (T->shape == OR_HEX)
        diffusion_hex(DEF, T->nqtot, cvd, dd, df, d0, d1, d2);
    else if (T->shape == OR_TET)
        diffusion_tet(DEF, T, cvd, dd, df, d0, d1, d2);
    else if (T->shape == OR_PRISM)
        diffusion_prism(DEF, T, cvd, dd, df, d0, d1, d2);
    else
        diffusion_pyr(DEF, T, cvd, dd, df, d0, d1, d2);
}

/* ------------------------------------------------------------------ */
/* Block workspace and per-block dispatch                              */
typedef struct {
    vd *in, *out, *bwd, *d0, *d1, *d2, *w0, *w1, *w2, *jac, *df;
} ws_t;

static void ws_alloc(ws_t *w, const tables_t *T, int deformed)
{
    const int nqt = T->nqtot, nc = T->ncoeff;
    const size_t m0 = T->nmd[0], m1 = T->nmd[1], m2 = T->nmd[2];
    const size_t q0 = T->nq[0], q1 = T->nq[1], q2 = T->nq[2];
    size_t big = (size_t)(nqt > nc ? nqt : nc);
    /* bwd stages (q0 m1 m2, q1 q0 m2), iprod sums (q1 q2), tet pairs (m0 m1) */
    size_t w0s = q0 * m1 * m2;
    if (q0 * q1 * m2 > w0s) w0s = q0 * q1 * m2;
    if (q1 * q2 > w0s) w0s = q1 * q2;
    if (m0 * m1 > w0s) w0s = m0 * m1;
    if (big > w0s) w0s = big;
    w->in = (vd *)calloc(big, sizeof(vd));
    w->out = (vd *)calloc(big, sizeof(vd));
    w->bwd = (vd *)calloc(nqt, sizeof(vd));
    w->d0 = (vd *)calloc(nqt, sizeof(vd));
    w->d1 = (vd *)calloc(nqt, sizeof(vd));
    w->d2 = (vd *)calloc(nqt, sizeof(vd));
    w->w0 = (vd *)calloc(w0s, sizeof(vd));
    w->w1 = (vd *)calloc(w0s, sizeof(vd));
    w->w2 = (vd *)calloc(w0s, sizeof(vd));
    w->jac = (vd *)calloc(deformed ? nqt : 1, sizeof(vd));
    w->df = (vd *)calloc(9 * (size_t)(deformed ? nqt : 1), sizeof(vd));
}

static void ws_free(ws_t *w)
{
    free(w->in); free(w->out); free(w->bwd); free(w->d0); free(w->d1);
    free(w->d2); free(w->w0); free(w->w1); free(w->w2); free(w->jac); free(w->df);
}

/* load_unalign_interleave with zero padding of the last block
 * (tinysimd.hpp:766; Helmholtz.h:448-461) */
static void load_block(const double *src, int nelmt, int b, int n, vd *dst)
{
    for (int l = 0; l < VW; ++l) {
        int e = b * VW + l;
        for (int m = 0; m < n; ++m) dst[m].v[l] = e < nelmt ? src[(size_t)e * n + m] : 0.0;
    }
}

static void store_block(const vd *src, int nelmt, int b, int n, double *dst)
{
    for (int l = 0; l < VW; ++l) {
        int e = b * VW + l;
        if (e >= nelmt) break;
        for (int m = 0; m < n; ++m) dst[(size_t)e * n + m] = src[m].v[l];
    }
}

static void bwd_block(const tables_t *T, const vd *in, ws_t *w, vd *out)
{
    const int m0 = T->nmd[0], m1 = T->nmd[1], m2 = T->nmd[2];
    if (T->shape == OR_HEX)
        bwd_hex(m0, m1, m2, T->nq[0], T->nq[1], T->nq[2], in, T->b[0], T->b[1], T->b[2], w->w0, w->w1, out);
    else if (T->shape == OR_TET)
        bwd_tet(m0, m1, m2, T->nq[0], T->nq[1], T->nq[2], 1, in, T->b[0], T->b[1], T->b[2], w->w0, w->w1, out);
    else if (T->shape == OR_PRISM)
        bwd_prism(m0, m1, m2, T->nq[0], T->nq[1], T->nq[2], 1, in, T->b[0], T->b[1], T->b[2], w->w0, w->w1,
                  out);
    else
        bwd_pyr(m0, m1, m2, T->nq[0], T->nq[1], T->nq[2], 1, in, T->b[0], T->b[1], T->b[2], w->w0, w->w1, out);
}

static void iprod_block(const tables_t *T, int SCALE, int APPEND, int DEF,
                        const vd *in, const double *b0, const double *b1,
                        const double *b2, const vd *jac, ws_t *w, vd *out,
                        double scale)
{
    const int m0 = T->nmd[0], m1 = T->nmd[1], m2 = T->nmd[2];
    if (T->shape == OR_HEX)
        iprod_hex(SCALE, APPEND, DEF, m0, m1, m2, T->nq[0], T->nq[1], T->nq[2], in, b0, b1, b2,
                  T->w[0], T->w[1], T->w[2], jac, w->w0, w->w1, out, scale);
    else if (T->shape == OR_TET)
        iprod_tet(SCALE, APPEND, DEF, m0, m1, m2, T->nq[0], T->nq[1], T->nq[2], 1, in, b0, b1, b2,
                  T->w[0], T->w[1], T->w[2], jac, w->w0, w->w1, out, scale);
    else if (T->shape == OR_PRISM)
        /* the correction also runs with derivative tables: it then adds the
         * derivative of the collapsed-edge term (Helmholtz.h:564-599 passes
         * correct with every table set) */
        iprod_prism(SCALE, APPEND, DEF, m0, m1, m2, T->nq[0], T->nq[1], T->nq[2], 1, in, b0, b1, b2, T->w[0],
                    T->w[1], T->w[2], jac, w->w0, w->w1, out, scale);
    else
        iprod_pyr(SCALE, APPEND, DEF, m0, m1, m2, T->nq[0], T->nq[1], T->nq[2], 1, in, b0, b1, b2, T->w[0],
                  T->w[1], T->w[2], jac, w->w0, w->w1, out, scale);
}

static int g_threads = 1;
void or_set_threads(int n) { g_threads = n > 0 ? n : 1; }
int or_get_threads(void) { return g_threads; }

static void load_geom(const tables_t *T, int nelmt, int b, int deformed,
                      const double *jac, const double *df, ws_t *w)
{
    const int gq = deformed ? T->nqtot : 1;
    if (jac) load_block(jac, nelmt, b, gq, w->jac);
    if (df) load_block(df, nelmt, b, 9 * gq, w->df);
}

void or_bwdtrans_aniso(int shape, const int *nm, const int *nq, int nelmt, const double *in,
                 double *out)
{
    tables_t T;
    tables_init_full(&T, shape, nm, nq);
    const int nblk = (nelmt + VW - 1) / VW;
#pragma omp parallel num_threads(g_threads)
    {
        ws_t w;
        ws_alloc(&w, &T, 0);
#pragma omp for schedule(static)
        for (int b = 0; b < nblk; ++b) {
            load_block(in, nelmt, b, T.ncoeff, w.in);
            bwd_block(&T, w.in, &w, w.bwd);
            store_block(w.bwd, nelmt, b, T.nqtot, out);
        }
        ws_free(&w);
    }
    tables_free(&T);
}

void or_iproduct_aniso(int shape, const int *nm, const int *nq, int nelmt, int deformed,
                 const double *jac, const double *in, double *out)
{
    tables_t T;
    tables_init_full(&T, shape, nm, nq);
    const int nblk = (nelmt + VW - 1) / VW;
#pragma omp parallel num_threads(g_threads)
    {
        ws_t w;
        ws_alloc(&w, &T, deformed);
#pragma omp for schedule(static)
        for (int b = 0; b < nblk; ++b) {
            load_block(in, nelmt, b, T.nqtot, w.in);
            load_geom(&T, nelmt, b, deformed, jac, NULL, &w);
            iprod_block(&T, 0, 0, deformed, w.in, T.b[0], T.b[1], T.b[2], w.jac, &w, w.out, 1.0);
            store_block(w.out, nelmt, b, T.ncoeff, out);
        }
        ws_free(&w);
    }
    tables_free(&T);
}

void or_physderiv_aniso(int shape, const int *nm, const int *nq, int nelmt, int deformed,
                  const double *df, const double *in, double *out0,
                  double *out1, double *out2)
{
    tables_t T;
    tables_init_full(&T, shape, nm, nq);
    const int nblk = (nelmt + VW - 1) / VW;
#pragma omp parallel num_threads(g_threads)
    {
        ws_t w;
        ws_alloc(&w, &T, deformed);
#pragma omp for schedule(static)
        for (int b = 0; b < nblk; ++b) {
            load_block(in, nelmt, b, T.nqtot, w.in);
            load_geom(&T, nelmt, b, deformed, NULL, df, &w);
            deriv_tensor(T.nq[0], T.nq[1], T.nq[2], w.in, T.D[0], T.D[1], T.D[2], w.d0, w.d1, w.d2);
            physderiv_chain(shape, deformed, T.nq[0], T.nq[1], T.nq[2], T.Z[0], T.Z[1], T.Z[2], w.df,
                            w.w0, w.w1, w.d0, w.d1, w.d2);
            store_block(w.d0, nelmt, b, T.nqtot, out0);
            store_block(w.d1, nelmt, b, T.nqtot, out1);
            store_block(w.d2, nelmt, b, T.nqtot, out2);
        }
        ws_free(&w);
    }
    tables_free(&T);
}

/* Helmholtz.h:549-599 (fusedKernel3D) on one block */
static void helm_block(const tables_t *T, int deformed, double lambda,
                       const double *diff6, ws_t *w)
{
    bwd_block(T, w->in, w, w->bwd);
    iprod_block(T, 1, 0, deformed, w->bwd, T->b[0], T->b[1], T->b[2], w->jac, w, w->out, lambda);
    deriv_tensor(T->nq[0], T->nq[1], T->nq[2], w->bwd, T->D[0], T->D[1], T->D[2], w->d0, w->d1, w->d2);
    diffusion(deformed, T, diff6 != NULL, diff6, w->df, w->d0, w->d1, w->d2);
    iprod_block(T, 0, 1, deformed, w->d0, T->db[0], T->b[1], T->b[2], w->jac, w, w->out, 1.0);
    iprod_block(T, 0, 1, deformed, w->d1, T->b[0], T->db[1], T->b[2], w->jac, w, w->out, 1.0);
    iprod_block(T, 0, 1, deformed, w->d2, T->b[0], T->b[1], T->db[2], w->jac, w, w->out, 1.0);
}

void or_helmholtz_aniso(int shape, const int *nm, const int *nq, int nelmt, int deformed,
                  const double *jac, const double *df, double lambda,
                  const double *diff6, const double *in, double *out)
{
    tables_t T;
    tables_init_full(&T, shape, nm, nq);
    const int nblk = (nelmt + VW - 1) / VW;
#pragma omp parallel num_threads(g_threads)
    {
        ws_t w;
        ws_alloc(&w, &T, deformed);
#pragma omp for schedule(static)
        for (int b = 0; b < nblk; ++b) {
            load_block(in, nelmt, b, T.ncoeff, w.in);
            load_geom(&T, nelmt, b, deformed, jac, df, &w);
            helm_block(&T, deformed, lambda, diff6, &w);
            store_block(w.out, nelmt, b, T.ncoeff, out);
        }
        ws_free(&w);
    }
    tables_free(&T);
}

/* LinearAdvectionDiffusionReactionKernels.hpp:193-277 (hex) and :279-386
 * (tet): bwd += V . grad_x u at every point, with the velocity per point
 * (vel[cnt*3 + j] of the element) and grad_x u = df^T (collapsed ->
 * standard derivatives); the same operation order as the reference. */
static void advection_pointwise(const tables_t *T, int DEF, const vd *df, const vd *vel,
                                const vd *d0, const vd *d1, const vd *d2, vd *bwd)
{
    const int nq0 = T->nq[0], nq1 = T->nq[1], nq2 = T->nq[2];
    vd f[9];
    if (!DEF)
        for (int c = 0; c < 9; ++c) f[c] = df[c];
    const int tet = T->shape == OR_TET;  /* the h factors exist for tets only */
    for (int k = 0, cnt = 0; k < nq2; ++k) {
        const double h3 = tet ? 2.0 / (1.0 - T->Z[2][k]) : 0.0;
        for (int j = 0; j < nq1; ++j) {
            const double z1 = T->Z[1][j];
            const double h1 = 0.5 * (1.0 + z1), h2 = tet ? 2.0 / (1.0 - z1) : 0.0;
            const double h2h3 = h2 * h3, h1h3 = h1 * h3;
            for (int i = 0; i < nq0; ++i, ++cnt) {
                if (DEF)
                    for (int c = 0; c < 9; ++c) f[c] = df[(size_t)cnt * 9 + c];
                const vd vx = vel[cnt * 3], vy = vel[cnt * 3 + 1], vzz = vel[cnt * 3 + 2];
                vd m0 = d0[cnt], m1 = d1[cnt], m2 = d2[cnt];
                if (T->shape == OR_PYR) {
                    /* AdvectionPyrKernel (LinearAdvectionDiffusionReactionKernels.hpp:
                     * 475-560): metric0 = h2 d0, metric1 = h2 d1,
                     * metric2 = d2 + h0 h2 d0 + h1 h2 d1 */
                    const double hy2 = T->h3[k], hy0h2 = T->h0[i] * hy2, hy1h2 = T->h1[j] * hy2;
                    m0 = vmuls(d0[cnt], hy2);
                    m1 = vmuls(d1[cnt], hy2);
                    m2 = vfmas(d2[cnt], d0[cnt], hy0h2);
                    m2 = vfmas(m2, d1[cnt], hy1h2);
                }
                if (T->shape == OR_PRISM) {
                    /* AdvectionPrismKernel (LinearAdvectionDiffusionReactionKernels.hpp:
                     * 384-470): metric0 = h1 d0, metric2 = h0 metric0 + d2 */
                    const double hp0 = T->h0[i], hp1 = T->h3[k];
                    m0 = vmuls(d0[cnt], hp1);
                    m2 = vadd(vmuls(m0, hp0), d2[cnt]);
                }
                if (tet) {
                    const double h0 = 0.5 * (1.0 + T->Z[0][i]);
                    const vd metric0 = vmuls(d0[cnt], h2h3);
                    vd metric1 = vmuls(metric0, h0);
                    vd metric2 = vadd(metric1, d2[cnt]);
                    metric2 = vfmas(metric2, d1[cnt], h1h3);
                    metric1 = vfmas(metric1, d1[cnt], h3);
                    m0 = metric0;
                    m1 = metric1;
                    m2 = metric2;
                }
                vd tmp = vmul(f[0], m0);
                tmp = vfma(tmp, f[1], m1);
                tmp = vfma(tmp, f[2], m2);
                tmp = vmul(tmp, vx);
                vd tmp2 = vmul(f[3], m0);
                tmp2 = vfma(tmp2, f[4], m1);
                tmp2 = vfma(tmp2, f[5], m2);
                tmp = vfma(tmp, tmp2, vy);
                tmp2 = vmul(f[6], m0);
                tmp2 = vfma(tmp2, f[7], m1);
                tmp2 = vfma(tmp2, f[8], m2);
                tmp = vfma(tmp, tmp2, vzz);
                bwd[cnt] = vadd(bwd[cnt], tmp);
            }
        }
    }
}

/* LinearAdvectionDiffusionReaction.h:600-658 (fusedKernel3D): BwdTrans,
 * collocation derivatives, bwd += V . grad u, out = lambda IProduct(bwd),
 * diffusion metric, out += IProduct w.r.t. the three derivative bases. */
void or_adv_diff_react_aniso(int shape, const int *nm, const int *nq, int nelmt, int deformed,
                       const double *jac, const double *df, const double *vel,
                       double lambda, const double *diff6, const double *in,
                       double *out)
{
    tables_t T;
    tables_init_full(&T, shape, nm, nq);
    const int nblk = (nelmt + VW - 1) / VW;
#pragma omp parallel num_threads(g_threads)
    {
        ws_t w;
        ws_alloc(&w, &T, deformed);
        vd *v = (vd *)calloc(3 * (size_t)T.nqtot, sizeof(vd));
#pragma omp for schedule(static)
        for (int b = 0; b < nblk; ++b) {
            load_block(in, nelmt, b, T.ncoeff, w.in);
            load_geom(&T, nelmt, b, deformed, jac, df, &w);
            load_block(vel, nelmt, b, 3 * T.nqtot, v);
            bwd_block(&T, w.in, &w, w.bwd);
            deriv_tensor(T.nq[0], T.nq[1], T.nq[2], w.bwd, T.D[0], T.D[1], T.D[2], w.d0, w.d1, w.d2);
            advection_pointwise(&T, deformed, w.df, v, w.d0, w.d1, w.d2, w.bwd);
            iprod_block(&T, 1, 0, deformed, w.bwd, T.b[0], T.b[1], T.b[2], w.jac, &w, w.out, lambda);
            diffusion(deformed, &T, diff6 != NULL, diff6, w.df, w.d0, w.d1, w.d2);
            iprod_block(&T, 0, 1, deformed, w.d0, T.db[0], T.b[1], T.b[2], w.jac, &w, w.out, 1.0);
            iprod_block(&T, 0, 1, deformed, w.d1, T.b[0], T.db[1], T.b[2], w.jac, &w, w.out, 1.0);
            iprod_block(&T, 0, 1, deformed, w.d2, T.b[0], T.b[1], T.db[2], w.jac, &w, w.out, 1.0);
            store_block(w.out, nelmt, b, T.ncoeff, out);
        }
        free(v);
        ws_free(&w);
    }
    tables_free(&T);
}

/* IProductWRTDerivBaseKernels.hpp:194-247 (hex) and :249-335 (tet): the
 * integrand of direction d is t_d = sum_j df[j*3+d] in_j at every point;
 * for tets it is then mapped to the collapsed directions:
 * t0 <- f0 (t0 + f1 (t1 + t2)), t1 <- f2 (t1 + f3 t2). */
static void iprod_deriv_pointwise(const tables_t *T, int DEF, const vd *df,
                                  const vd *in0, const vd *in1, const vd *in2,
                                  vd *o0, vd *o1, vd *o2)
{
    const int nq0 = T->nq[0], nq1 = T->nq[1], nq2 = T->nq[2];
    vd d[9];
    if (!DEF)
        for (int c = 0; c < 9; ++c) d[c] = df[c];
    for (int k = 0, cnt = 0; k < nq2; ++k) {
        const double f2 = 2.0 / (1.0 - T->Z[2][k]);
        for (int j = 0; j < nq1; ++j) {
            const double z1 = T->Z[1][j];
            const double f3 = 0.5 * (1.0 + z1);
            const double f0 = 2.0 * f2 / (1.0 - z1);
            for (int i = 0; i < nq0; ++i, ++cnt) {
                if (DEF)
                    for (int c = 0; c < 9; ++c) d[c] = df[(size_t)cnt * 9 + c];
                const vd a = in0[cnt], b = in1[cnt], e = in2[cnt];
                vd t0 = vadd(vadd(vmul(d[0], a), vmul(d[3], b)), vmul(d[6], e));
                vd t1 = vadd(vadd(vmul(d[1], a), vmul(d[4], b)), vmul(d[7], e));
                vd t2 = vadd(vadd(vmul(d[2], a), vmul(d[5], b)), vmul(d[8], e));
                if (T->shape == OR_TET) {
                    const double f1 = 0.5 * (1.0 + T->Z[0][i]);
                    t0 = vfmas(t0, vadd(t1, t2), f1);
                    t0 = vmuls(t0, f0);
                    t1 = vfmas(t1, t2, f3);
                    t1 = vmuls(t1, f2);
                } else if (T->shape == OR_PYR) {
                    /* IProductWRTDerivBaseKernels.hpp:469-488: t0 <- f t0 + f (h0 t2),
                     * t1 <- f t1 + f (h1 t2), f = 2/(1 - z2) */
                    const double hf1 = 0.5 * (1.0 + T->Z[0][i]), hf2 = 0.5 * (1.0 + z1);
                    t0 = vmuls(t0, f2);
                    t0 = vfmas(t0, vmuls(t2, hf1), f2);
                    t1 = vmuls(t1, f2);
                    t1 = vfmas(t1, vmuls(t2, hf2), f2);
                } else if (T->shape == OR_PRISM) {
                    /* IProductWRTDerivBaseKernels.hpp:391-403: t0 <- f t0 + f (h0 t2),
                     * f = 2/(1 - z2) */
                    const double hf1 = 0.5 * (1.0 + T->Z[0][i]);
                    t0 = vmuls(t0, f2);
                    t0 = vfmas(t0, vmuls(t2, hf1), f2);
                }
                o0[cnt] = t0;
                o1[cnt] = t1;
                o2[cnt] = t2;
            }
        }
    }
}

/* IProductWRTDerivBase.h:751-788 (fusedKernel3D): pointwise step, then
 * IProduct w.r.t. (db0 b1 b2), += (b0 db1 b2), += (b0 b1 db2). */
void or_iproduct_deriv_aniso(int shape, const int *nm, const int *nq, int nelmt, int deformed,
                       const double *jac, const double *df, const double *in0,
                       const double *in1, const double *in2, double *out)
{
    tables_t T;
    tables_init_full(&T, shape, nm, nq);
    const int nblk = (nelmt + VW - 1) / VW;
#pragma omp parallel num_threads(g_threads)
    {
        ws_t w;
        ws_alloc(&w, &T, deformed);
        vd *a1 = (vd *)calloc(T.nqtot, sizeof(vd)), *a2 = (vd *)calloc(T.nqtot, sizeof(vd));
#pragma omp for schedule(static)
        for (int b = 0; b < nblk; ++b) {
            load_block(in0, nelmt, b, T.nqtot, w.in);
            load_block(in1, nelmt, b, T.nqtot, a1);
            load_block(in2, nelmt, b, T.nqtot, a2);
            load_geom(&T, nelmt, b, deformed, jac, df, &w);
            iprod_deriv_pointwise(&T, deformed, w.df, w.in, a1, a2, w.d0, w.d1, w.d2);
            iprod_block(&T, 0, 0, deformed, w.d0, T.db[0], T.b[1], T.b[2], w.jac, &w, w.out, 1.0);
            iprod_block(&T, 0, 1, deformed, w.d1, T.b[0], T.db[1], T.b[2], w.jac, &w, w.out, 1.0);
            iprod_block(&T, 0, 1, deformed, w.d2, T.b[0], T.b[1], T.db[2], w.jac, &w, w.out, 1.0);
            store_block(w.out, nelmt, b, T.ncoeff, out);
        }
        free(a1);
        free(a2);
        ws_free(&w);
    }
    tables_free(&T);
}

void or_helmholtz_diag_bruteforce_aniso(int shape, const int *nm, const int *nq, int nelmt,
                                  int deformed, const double *jac,
                                  const double *df, double lambda,
                                  double *diag)
{
    tables_t T;
    tables_init_full(&T, shape, nm, nq);
    const int nblk = (nelmt + VW - 1) / VW;
    const int nc = T.ncoeff;
#pragma omp parallel num_threads(g_threads)
    {
        ws_t w;
        ws_alloc(&w, &T, deformed);
#pragma omp for schedule(static)
        for (int b = 0; b < nblk; ++b) {
            load_geom(&T, nelmt, b, deformed, jac, df, &w);
            for (int m = 0; m < nc; ++m) {
                for (int c = 0; c < nc; ++c) w.in[c] = vs(c == m ? 1.0 : 0.0);
                helm_block(&T, deformed, lambda, NULL, &w);
                for (int l = 0; l < VW; ++l) {
                    int e = b * VW + l;
                    if (e < nelmt) diag[(size_t)e * nc + m] = w.out[m].v[l];
                }
            }
        }
        ws_free(&w);
    }
    tables_free(&T);
}

/* ---- PhysInterp1DScaled -------------------------------------------- */

/* Points per direction after scaling (Operator.hpp:568-575): a direction
 * with one point fewer than direction 0 (the tet's collapsed directions)
 * gets (int)(nq0*factor - 1), every other (int)(nq*factor). */
void or_interp_nq(int shape, int nq0, double factor, int out[3])
{
    int nq[3];
    or_default_nq(shape, nq0, nq);
    for (int d = 0; d < 3; ++d)
        out[d] = (nq0 - nq[d] == 1) ? (int)(nq0 * factor - 1) : (int)(nq[d] * factor);
}

/* Interpolation matrix from nq_in points of one family to nq_out points of
 * the same family, I[p*nq_out + i] = h_p(zo_i), the Lagrange polynomial of
 * input point p at output point i (GaussPoints.cpp:283-380 via Polylib
 * Imglj / Imgrjm, Polylib.cpp:1085-1160).  Evaluated here as the product
 * formula of the same polynomial, h_p(x) = prod_{k != p} (x - z_k)/(z_p -
 * z_k), and exactly 1 / 0 where an output point coincides with an input
 * point. */
void or_interp_matrix(int ptype, int nq_in, int nq_out, double *I)
{
    double *zi = (double *)malloc(sizeof(double) * nq_in), *wi = (double *)malloc(sizeof(double) * nq_in);
    double *Di = (double *)malloc(sizeof(double) * nq_in * nq_in);
    double *zo = (double *)malloc(sizeof(double) * nq_out), *wo = (double *)malloc(sizeof(double) * nq_out);
    double *Do = (double *)malloc(sizeof(double) * nq_out * nq_out);
    or_points(ptype, nq_in, zi, wi, Di);
    or_points(ptype, nq_out, zo, wo, Do);
    for (int p = 0; p < nq_in; ++p)
        for (int i = 0; i < nq_out; ++i) {
            double h = 1.0;
            int hit = -1;
            for (int k = 0; k < nq_in; ++k)
                if (fabs(zo[i] - zi[k]) < 1e-15) hit = k;
            if (hit >= 0) {
                h = hit == p ? 1.0 : 0.0;
            } else {
                for (int k = 0; k < nq_in; ++k)
                    if (k != p) h *= (zo[i] - zi[k]) / (zi[p] - zi[k]);
            }
            I[p * nq_out + i] = h;
        }
    free(zi); free(wi); free(Di); free(zo); free(wo); free(Do);
}

/* PhysInterp1DScaled3DKernel (PhysInterp1DScaledKernels.hpp:151-217): the
 * three 1-D interpolations in turn, directions 0, 1, 2, in the reference's
 * loop and summation order. */
void or_phys_interp1d_scaled(int shape, int nq0, int nelmt, double factor,
                             const double *in, double *out)
{
    int ni[3], no[3];
    or_default_nq(shape, nq0, ni);
    or_interp_nq(shape, nq0, factor, no);
    const int pt[3] = {OR_GLL, shape == OR_TET ? OR_RADAU_A1 : OR_GLL, shape == OR_TET ? OR_RADAU_A2 : OR_GLL};
    double *I[3];
    for (int d = 0; d < 3; ++d) {
        I[d] = (double *)malloc(sizeof(double) * ni[d] * no[d]);
        or_interp_matrix(pt[d], ni[d], no[d], I[d]);
    }
    const int nin = ni[0] * ni[1] * ni[2], nout = no[0] * no[1] * no[2];
#pragma omp parallel num_threads(g_threads)
    {
        double *s_irq = (double *)malloc(sizeof(double) * no[0] * ni[1] * ni[2]);
        double *s_jir = (double *)malloc(sizeof(double) * no[1] * no[0] * ni[2]);
#pragma omp for schedule(static)
        for (int e = 0; e < nelmt; ++e) {
            const double *x = in + (size_t)e * nin;
            double *y = out + (size_t)e * nout;
            for (int i = 0, c = 0; i < no[0]; ++i)
                for (int r = 0, rqp = 0; r < ni[2]; ++r)
                    for (int q = 0; q < ni[1]; ++q, ++c) {
                        double t = x[rqp++] * I[0][i];
                        for (int p = 1; p < ni[0]; ++p, ++rqp) t = fma(x[rqp], I[0][p * no[0] + i], t);
                        s_irq[c] = t;
                    }
            for (int j = 0, c = 0; j < no[1]; ++j)
                for (int i = 0, irq = 0; i < no[0]; ++i)
                    for (int r = 0; r < ni[2]; ++r, ++c) {
                        double t = s_irq[irq++] * I[1][j];
                        for (int q = 1; q < ni[1]; ++q) t = fma(s_irq[irq++], I[1][q * no[1] + j], t);
                        s_jir[c] = t;
                    }
            for (int k = 0, c = 0; k < no[2]; ++k)
                for (int j = 0, jir = 0; j < no[1]; ++j)
                    for (int i = 0; i < no[0]; ++i, ++c) {
                        double t = s_jir[jir++] * I[2][k];
                        for (int r = 1; r < ni[2]; ++r) t = fma(s_jir[jir++], I[2][r * no[2] + k], t);
                        y[c] = t;
                    }
        }
        free(s_irq);
        free(s_jir);
    }
    for (int d = 0; d < 3; ++d) free(I[d]);
}

/* The MODIFIED defaults: nm modes in every direction, nq0 points in
 * direction 0 (or_default_nq for the others). */
void or_bwdtrans(int shape, int nm, int nq0, int nelmt, const double *in,
                 double *out)
{
    const int m[3] = {nm, nm, nm};
    int q[3];
    or_default_nq(shape, nq0, q);
    or_bwdtrans_aniso(shape, m, q, nelmt, in, out);
}

void or_iproduct(int shape, int nm, int nq0, int nelmt, int deformed,
                 const double *jac, const double *in, double *out)
{
    const int m[3] = {nm, nm, nm};
    int q[3];
    or_default_nq(shape, nq0, q);
    or_iproduct_aniso(shape, m, q, nelmt, deformed, jac, in, out);
}

void or_physderiv(int shape, int nm, int nq0, int nelmt, int deformed,
                  const double *df, const double *in, double *out0,
                  double *out1, double *out2)
{
    const int m[3] = {nm, nm, nm};
    int q[3];
    or_default_nq(shape, nq0, q);
    or_physderiv_aniso(shape, m, q, nelmt, deformed, df, in, out0, out1, out2);
}

void or_helmholtz(int shape, int nm, int nq0, int nelmt, int deformed,
                  const double *jac, const double *df, double lambda,
                  const double *diff6, const double *in, double *out)
{
    const int m[3] = {nm, nm, nm};
    int q[3];
    or_default_nq(shape, nq0, q);
    or_helmholtz_aniso(shape, m, q, nelmt, deformed, jac, df, lambda, diff6, in, out);
}

void or_adv_diff_react(int shape, int nm, int nq0, int nelmt, int deformed,
                       const double *jac, const double *df, const double *vel,
                       double lambda, const double *diff6, const double *in,
                       double *out)
{
    const int m[3] = {nm, nm, nm};
    int q[3];
    or_default_nq(shape, nq0, q);
    or_adv_diff_react_aniso(shape, m, q, nelmt, deformed, jac, df, vel, lambda, diff6, in, out);
}

void or_iproduct_deriv(int shape, int nm, int nq0, int nelmt, int deformed,
                       const double *jac, const double *df, const double *in0,
                       const double *in1, const double *in2, double *out)
{
    const int m[3] = {nm, nm, nm};
    int q[3];
    or_default_nq(shape, nq0, q);
    or_iproduct_deriv_aniso(shape, m, q, nelmt, deformed, jac, df, in0, in1, in2, out);
}

void or_helmholtz_diag_bruteforce(int shape, int nm, int nq0, int nelmt,
                                  int deformed, const double *jac,
                                  const double *df, double lambda,
                                  double *diag)
{
    const int m[3] = {nm, nm, nm};
    int q[3];
    or_default_nq(shape, nq0, q);
    or_helmholtz_diag_bruteforce_aniso(shape, m, q, nelmt, deformed, jac, df, lambda, diag);
}
