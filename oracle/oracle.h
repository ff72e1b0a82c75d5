/*
 * oracle.h -- CPU FP64 restatement of Nektar++'s matrix-free spectral/hp
 * operator path (quadrature, modified bases, sum-factorised BwdTrans /
 * IProductWRTBase / PhysDeriv / Helmholtz on hex and tet, geometric factors,
 * C0 assembly, Jacobi diagonal and the reference's single-reduction PCG).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in nektar_amd/ and the CPU baseline timed by bench.py.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it;
 * the product path never links, loads or falls back to it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the Nektar++ 5.10.0 source tree).  Parity pins (what checks each function):
 *   quadrature  : oracle/_ref (the reference Polylib.cpp compiled as-is),
 *                 monomial exactness, closed-form GLL diagonal entries;
 *   bases       : closed forms via scipy Jacobi polynomials (tests/);
 *   hex ops     : dense Kronecker brute force, exact polynomial integrals;
 *   tet ops     : Duffy-transformed Gauss-Legendre brute force of the same
 *                 basis definitions, partition-of-unity / symmetry identities;
 *   geometry    : closed-form trilinear/affine maps vs finite differences;
 *   CG          : dense numpy solve on small SPD systems, manufactured
 *                 solution convergence.
 */
#ifndef NEKTAR_ORACLE_H
#define NEKTAR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- shapes / point types ------------------------------------------- */
enum { OR_HEX = 0, OR_TET = 1, OR_PRISM = 2, OR_PYR = 3 };
enum { OR_GLL = 0, OR_RADAU_A1 = 1, OR_RADAU_A2 = 2 };
enum { OR_MOD_A = 0, OR_MOD_B = 1, OR_MOD_C = 2, OR_MOD_PYR_C = 3 };

/* ---- Polylib restatement (or_polylib.c) ----------------------------- */
void or_jacobfd(int np, const double *z, double *poly, double *polyd, int n,
                double alpha, double beta);
void or_jacobd(int np, const double *z, double *polyd, int n, double alpha,
               double beta);
void or_jacobz(int n, double *z, double alpha, double beta);
void or_zwglj(double *z, double *w, int np, double alpha, double beta);
void or_zwgrjm(double *z, double *w, int np, double alpha, double beta);
void or_dglj(double *D, const double *z, int np, double alpha, double beta);
void or_dgrjm(double *D, const double *z, int np, double alpha, double beta);

/* ---- points and bases (or_basis.c) ---------------------------------- */
/* z[nq], w[nq], D[nq*nq] with D[k*nq+i] = l_k'(z_i). */
void or_points(int ptype, int nq, double *z, double *w, double *D);
int or_basis_nmodes(int btype, int nm);
/* bdata/dbdata: [mode*nq + i]; sizes or_basis_nmodes(btype,nm)*nq. */
void or_basis(int btype, int nm, int nq, double *bdata, double *dbdata);
int or_ncoeffs(int shape, int nm);
/* Number of quadrature points per direction for a shape's default
 * (MODIFIED) expansion with nq0 points in direction 0. */
void or_default_nq(int shape, int nq0, int nq[3]);

/* ---- operators (or_ops.c) ------------------------------------------- *
 * Element-major, unpadded in/out exactly like the Collections interface.
 * Geometry (per element, not interleaved):
 *   regular : jac[e], df[e*9 + d]
 *   deformed: jac[e*nqtot + q], df[(e*nqtot + q)*9 + d]
 * df[j*3 + k] = d xi_k / d x_j.                                          */
void or_set_threads(int n);
int or_get_threads(void);
void or_bwdtrans(int shape, int nm, int nq0, int nelmt, const double *in,
                 double *out);
void or_iproduct(int shape, int nm, int nq0, int nelmt, int deformed,
                 const double *jac, const double *in, double *out);
void or_physderiv(int shape, int nm, int nq0, int nelmt, int deformed,
                  const double *df, const double *in, double *out0,
                  double *out1, double *out2);
/* diff6 = {d00,d01,d11,d02,d12,d22} or NULL for the identity tensor. */
/* IProductWRTDerivBase: out = sum_d IProduct(dphi/dxi_d, sum_j df[j*3+d] in_j)
 * (IProductWRTDerivBase.h:592-825, IProductWRTDerivBaseKernels.hpp:194-335). */
void or_iproduct_deriv(int shape, int nm, int nq0, int nelmt, int deformed,
                       const double *jac, const double *df, const double *in0,
                       const double *in1, const double *in2, double *out);
void or_helmholtz(int shape, int nm, int nq0, int nelmt, int deformed,
                  const double *jac, const double *df, double lambda,
                  const double *diff6, const double *in, double *out);
/* LinearAdvectionDiffusionReaction (fused, LinearAdvectionDiffusionReaction.h
 * :600-658): out = lambda IProduct(u + V.grad u) + L u, u = BwdTrans(in),
 * vel[(e*nqtot + q)*3 + j] per point. */
void or_adv_diff_react(int shape, int nm, int nq0, int nelmt, int deformed,
                       const double *jac, const double *df, const double *vel,
                       double lambda, const double *diff6, const double *in,
                       double *out);

/* The same operators with per-direction modes nm[3] and points nq[3] (the
 * runtime-size operators, MatrixFreeOps/Helmholtz.h:374-464 and
 * SwitchNodesPoints.h:302-343).  Hex: Modified_A on GLL in every
 * direction; tet: Modified_A/GLL, Modified_B/Radau(1,0), Modified_C/
 * Radau(2,0) with nm[0] <= nm[1] <= nm[2]; coefficients p < nm0, q < nm1-p,
 * r < nm2-p-q. */
int or_ncoeffs_aniso(int shape, const int nm[3]);
void or_bwdtrans_aniso(int shape, const int *nm, const int *nq, int nelmt,
                       const double *in, double *out);
void or_iproduct_aniso(int shape, const int *nm, const int *nq, int nelmt,
                       int deformed, const double *jac, const double *in,
                       double *out);
void or_physderiv_aniso(int shape, const int *nm, const int *nq, int nelmt,
                        int deformed, const double *df, const double *in,
                        double *out0, double *out1, double *out2);
void or_iproduct_deriv_aniso(int shape, const int *nm, const int *nq,
                             int nelmt, int deformed, const double *jac,
                             const double *df, const double *in0,
                             const double *in1, const double *in2,
                             double *out);
void or_helmholtz_aniso(int shape, const int *nm, const int *nq, int nelmt,
                        int deformed, const double *jac, const double *df,
                        double lambda, const double *diff6, const double *in,
                        double *out);
void or_adv_diff_react_aniso(int shape, const int *nm, const int *nq,
                             int nelmt, int deformed, const double *jac,
                             const double *df, const double *vel,
                             double lambda, const double *diff6,
                             const double *in, double *out);
void or_helmholtz_diag_bruteforce_aniso(int shape, const int *nm,
                                        const int *nq, int nelmt,
                                        int deformed, const double *jac,
                                        const double *df, double lambda,
                                        double *diag);

/* PhysInterp1DScaled (PhysInterp1DScaled.h:360-430): point values of
 * nq_in points per direction -> or_interp_nq points per direction. */
void or_interp_nq(int shape, int nq0, double factor, int out[3]);
void or_interp_matrix(int ptype, int nq_in, int nq_out, double *I);
void or_phys_interp1d_scaled(int shape, int nq0, int nelmt, double factor,
                             const double *in, double *out);
/* Diagonal of each element's Helmholtz matrix (ncoeff per element),
 * obtained by applying the element operator to unit vectors (the same
 * numbers as the diagonal of the reference's dense local matrices). */
void or_helmholtz_diag_bruteforce(int shape, int nm, int nq0, int nelmt,
                                  int deformed, const double *jac,
                                  const double *df, double lambda,
                                  double *diag);

/* ---- geometry (or_geom.c) ------------------------------------------- */
/* verts: 8 vertices x 3 coords per element in Nektar hex vertex order. */
int or_hex_is_regular(const double *verts8x3);
void or_hex_geom(const double *verts, int nelmt, int nq, int deformed,
                 double *jac, double *df);
/* verts: 4 vertices x 3 per element; affine (regular) factors only. */
void or_tet_geom(const double *verts, int nelmt, double *jac, double *df);
/* Physical coordinates of the quadrature points, [e][q][3]. */
void or_hex_coords(const double *verts, int nelmt, int nq, double *xyz);
/* Structured n^3 hex mesh on [0,1]^3, interior vertices perturbed by
 * U[-amp*h, amp*h] per coordinate from SplitMix64(seed).  verts out:
 * nelmt*8*3.  Element e = (ex + n*ey + n*n*ez). */
void or_hex_mesh(int n, double amp, uint64_t seed, double *verts);

/* ---- C0 assembly for the structured hex mesh (or_geom.c) ----------- */
/* map[e*ncoeff + m] = global id; Dirichlet flags per global id. */
int or_hex_nglobal(int n, int nm);
void or_hex_l2g(int n, int nm, int *map, uint8_t *dirichlet);
void or_assemble(int nloc, const int *map, const double *loc, int nglob,
                 double *glob);
void or_global_to_local(int nloc, const int *map, const double *glob,
                        double *loc);

/* ---- CG (or_cg.c): NekLinSysIterCG::DoConjugateGradient ------------- */
typedef struct {
    int iterations;
    double final_error;   /* sqrt(eps / rhs_magnitude) */
    double rhs_magnitude; /* Set_Rhs_Magnitude value */
} or_cg_info;
/* Solve A x = b for the Helmholtz operator on the structured mesh with
 * homogeneous Dirichlet data.  b_glob has nglob entries (Dirichlet rows
 * included; they enter rhs_magnitude as in the reference), x_glob out.
 * inv_diag: Jacobi inverse diagonal per global id (NULL: identity). */
int or_cg_helmholtz_hex(int n, int nm, int nq, int deformed,
                        const double *jac, const double *df, double lambda,
                        const double *b_glob, const double *inv_diag,
                        double tol, int maxit, double *x_glob,
                        or_cg_info *info);
/* Generic reference CG on a dense SPD matrix (test use, small n). */
int or_cg_dense(int n, int ndir, const double *A, const double *b,
                const double *inv_diag, double tol, int maxit, double *x,
                or_cg_info *info);
/* The same with the rhs magnitude set by the caller (rhs_mag > 0), as
 * GlobalLinSysIterativeStaticCond::v_PreSolve sets it from the
 * uncondensed boundary forcing (GlobalLinSysIterativeStaticCond.cpp:436-492). */
int or_cg_dense_mag(int n, int ndir, const double *A, const double *b,
                    const double *inv_diag, double tol, int maxit,
                    double rhs_mag, double *x, or_cg_info *info);
/* Element-local CG (NekLinSysIterCGLoc.cpp:114-308): Krylov vectors in
 * the element-local layout, operator = nblk dense element blocks (sizes
 * boff[k+1]-boff[k], row-major, consecutive in Ablk), signed map with the
 * ndir Dirichlet DOFs first, Jacobi preconditioner in local form
 * (inv_diag: nglob entries, 0 on Dirichlet rows).  x_loc: boff[nblk]. */
int or_cg_loc(int nblk, const int *boff, const double *Ablk, const int *map,
              const double *sign, int nglob, int ndir, const double *inv_diag,
              const double *b_loc, double tol, int maxit, double *x_loc,
              or_cg_info *info);

/* Static-condensation diagonal (Preconditioner.cpp:185-236): per block the
 * Schur complement of its interior entries (interior[i] != 0), summed over
 * boundary pairs with equal global id into d (nglob; rows < ndir and
 * interior-only rows 0).  Returns 0, or 1 + a block whose interior block is
 * not positive definite. */
int or_sc_diagonal(int nblk, const int *boff, const double *Ablk,
                   const uint8_t *interior, const int *map,
                   const double *sign, int nglob, int ndir, double *d);

#ifdef __cplusplus
}
#endif
#endif
