/*
 * or_cg.c -- the reference's single-reduction preconditioned CG
 * (LibUtilities/LinearAlgebra/NekLinSysIterCG.cpp:118-372, Demmel, Heath &
 * van der Vorst 1993 reordering) with the Jacobi preconditioner
 * (MultiRegions/PreconditionerDiagonal.cpp:93-125 build, :159-189 apply) and
 * the C0 matrix-vector product of GlobalLinSysIterativeFull::v_DoMatrixMultiply
 * (GlobalLinSysIterativeFull.cpp:180-241: GlobalToLocal -> Helmholtz ->
 * Assemble).  Rhs magnitude: NekLinSysIter::Set_Rhs_Magnitude
 * (NekLinSysIter.cpp:96-117) over all global rows, floored to 1 below 1e-6.
 *
 * TEST INFRASTRUCTURE ONLY (parity oracle and CPU baseline).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef void (*matvec_fn)(void *ctx, const double *w_glob, double *s_glob);

/* Inner products are summed pairwise (runs of up to 128 terms added in
 * order, then halves combined recursively): the rounding error grows as
 * log n instead of n.  The reference's Vmath::Dot adds in order on each
 * rank; at the 6,967,871 unknowns of the 32^3 P=6 box that sequential
 * FP64 sum alone costs the single-reduction CG 28 iterations (164 against
 * 136 with extended-precision sums), so the order of summation is not the
 * reference's algorithm but a property of one build of it. */
static double dot(int n, const double *a, const double *b)
{
    if (n <= 128) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += a[i] * b[i];
        return s;
    }
    const int h = n / 2;
    return dot(h, a, b) + dot(n - h, a + h, b + h);
}

/* nd[]: the nNonDir non-Dirichlet global ids in ascending order. */
static int cg_core_mag(int nglob, int nnd, const int *nd, matvec_fn A, void *ctx,
                       const double *b, const double *inv_diag, double tol,
                       int maxit, double rhs_mag_set, double *x, or_cg_info *info);

static int cg_core(int nglob, int nnd, const int *nd, matvec_fn A, void *ctx,
                   const double *b, const double *inv_diag, double tol,
                   int maxit, double *x, or_cg_info *info)
{
    return cg_core_mag(nglob, nnd, nd, A, ctx, b, inv_diag, tol, maxit, -1.0, x, info);
}

/* rhs_mag_set > 0: the rhs magnitude the caller set (SetRhsMagnitude,
 * e.g. GlobalLinSysIterativeStaticCond::v_PreSolve); else Set_Rhs_Magnitude */
static int cg_core_mag(int nglob, int nnd, const int *nd, matvec_fn A, void *ctx,
                       const double *b, const double *inv_diag, double tol,
                       int maxit, double rhs_mag_set, double *x, or_cg_info *info)
{
    double *w = (double *)calloc(nglob, sizeof(double));
    double *s = (double *)calloc(nglob, sizeof(double));
    double *p = (double *)calloc(nnd, sizeof(double));
    double *r = (double *)calloc(nnd, sizeof(double));
    double *q = (double *)calloc(nnd, sizeof(double));
    double *wn = (double *)calloc(nnd, sizeof(double));
    double *sn = (double *)calloc(nnd, sizeof(double));
    tol = fmax(tol, 1.0e-16); /* NekLinSysIter.cpp:57 */

    for (int i = 0; i < nnd; ++i) { r[i] = b[nd[i]]; x[nd[i]] = 0.0; }
    double eps = dot(nnd, r, r);
    double rhs_mag = dot(nglob, b, b);
    rhs_mag = rhs_mag > 1.0e-6 ? rhs_mag : 1.0;
    if (rhs_mag_set > 0.0) rhs_mag = rhs_mag_set;
    int iters = 0;
    if (eps < tol * tol * rhs_mag) goto done;

#define PRECON()                                                              \
    for (int i = 0; i < nnd; ++i) {                                           \
        wn[i] = inv_diag ? r[i] * inv_diag[nd[i]] : r[i];                     \
        w[nd[i]] = wn[i];                                                     \
    }
#define MATVEC()                                                              \
    A(ctx, w, s);                                                             \
    for (int i = 0; i < nnd; ++i) sn[i] = s[nd[i]];

    PRECON();
    MATVEC();
    double rho = dot(nnd, r, wn), mu = dot(nnd, sn, wn);
    double rho_star = 0.0, beta = 0.0, alpha = rho / mu, rho_new;
    iters = 1;
    while (1) {
        if (iters > maxit) break;
        for (int i = 0; i < nnd; ++i) p[i] = beta * p[i] + wn[i];
        for (int i = 0; i < nnd; ++i) q[i] = beta * q[i] + sn[i];
        for (int i = 0; i < nnd; ++i) x[nd[i]] = alpha * p[i] + x[nd[i]];
        for (int i = 0; i < nnd; ++i) r[i] = -alpha * q[i] + r[i];
        PRECON();
        MATVEC();
        rho_new = dot(nnd, r, wn);
        mu = dot(nnd, sn, wn);
        eps = dot(nnd, r, r);
        iters++;
        if (eps < tol * tol * rhs_mag) break;
        beta = (rho_new - rho_star) / rho;
        alpha = rho_new / (mu - rho_new * beta / alpha);
        rho = rho_new;
    }
#undef PRECON
#undef MATVEC
done:
    if (info) {
        info->iterations = iters;
        info->final_error = sqrt(eps / rhs_mag);
        info->rhs_magnitude = rhs_mag;
    }
    free(w); free(s); free(p); free(r); free(q); free(wn); free(sn);
    return iters;
}

/* ---- dense operator (tests) ---------------------------------------- */
typedef struct { int n; const double *A; } dense_ctx;

static void dense_mv(void *c, const double *w, double *s)
{
    const dense_ctx *d = (const dense_ctx *)c;
    for (int i = 0; i < d->n; ++i) {
        double t = 0.0;
        for (int j = 0; j < d->n; ++j) t += d->A[(size_t)i * d->n + j] * w[j];
        s[i] = t;
    }
}

int or_cg_dense(int n, int ndir, const double *A, const double *b,
                const double *inv_diag, double tol, int maxit, double *x,
                or_cg_info *info)
{
    int *nd = (int *)malloc((size_t)n * sizeof(int));
    for (int i = ndir; i < n; ++i) nd[i - ndir] = i;
    dense_ctx c = {n, A};
    int it = cg_core(n, n - ndir, nd, dense_mv, &c, b, inv_diag, tol, maxit, x, info);
    free(nd);
    return it;
}

int or_cg_dense_mag(int n, int ndir, const double *A, const double *b,
                    const double *inv_diag, double tol, int maxit,
                    double rhs_mag, double *x, or_cg_info *info)
{
    int *nd = (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    for (int i = ndir; i < n; ++i) nd[i - ndir] = i;
    dense_ctx c = {n, A};
    int it = cg_core_mag(n, n - ndir, nd, dense_mv, &c, b, inv_diag, tol, maxit, rhs_mag, x, info);
    free(nd);
    return it;
}

/* ---- Helmholtz on the structured hex mesh ---------------------------- */
typedef struct {
    int n, nm, nq, deformed, nelmt, nloc, nglob;
    const double *jac, *df;
    double lambda;
    const int *map;
    double *loc_in, *loc_out;
} helm_ctx;

static void helm_mv(void *c, const double *w, double *s)
{
    helm_ctx *h = (helm_ctx *)c;
    or_global_to_local(h->nloc, h->map, w, h->loc_in);
    or_helmholtz(OR_HEX, h->nm, h->nq, h->nelmt, h->deformed, h->jac, h->df, h->lambda, NULL,
                 h->loc_in, h->loc_out);
    or_assemble(h->nloc, h->map, h->loc_out, h->nglob, s);
}

int or_cg_helmholtz_hex(int n, int nm, int nq, int deformed,
                        const double *jac, const double *df, double lambda,
                        const double *b_glob, const double *inv_diag,
                        double tol, int maxit, double *x_glob,
                        or_cg_info *info)
{
    helm_ctx h;
    h.n = n; h.nm = nm; h.nq = nq; h.deformed = deformed;
    h.nelmt = n * n * n;
    h.nloc = h.nelmt * nm * nm * nm;
    h.nglob = or_hex_nglobal(n, nm);
    h.jac = jac; h.df = df; h.lambda = lambda;
    int *map = (int *)malloc((size_t)h.nloc * sizeof(int));
    uint8_t *dir = (uint8_t *)malloc((size_t)h.nglob);
    or_hex_l2g(n, nm, map, dir);
    h.map = map;
    h.loc_in = (double *)malloc((size_t)h.nloc * sizeof(double));
    h.loc_out = (double *)malloc((size_t)h.nloc * sizeof(double));
    int *nd = (int *)malloc((size_t)h.nglob * sizeof(int));
    int nnd = 0;
    for (int g = 0; g < h.nglob; ++g)
        if (!dir[g]) nd[nnd++] = g;
        else x_glob[g] = 0.0;
    int it = cg_core(h.nglob, nnd, nd, helm_mv, &h, b_glob, inv_diag, tol, maxit, x_glob, info);
    free(map); free(dir); free(h.loc_in); free(h.loc_out); free(nd);
    return it;
}

/* ---- element-local CG (NekLinSysIterCGLoc.cpp:114-308) ---------------
 * Every Krylov vector lives in the element-local layout (nloc entries, the
 * concatenated element coefficients).  The operator is the block-diagonal
 * element matrix (GlobalLinSysIterativeFull.cpp:192-198: GeneralMatrixOp
 * without assembly); the preconditioner assembles, scales the
 * non-Dirichlet rows by the inverse diagonal, zeroes the Dirichlet rows and
 * scatters back (PreconditionerDiagonal.cpp:174-184); the residual norm is
 * <AssembleLoc(r), r> with the Dirichlet rows zeroed
 * (Preconditioner.cpp:89-102, NekLinSysIterCGLoc.cpp:154-156,:248-252) and
 * the rhs magnitude <AssembleLoc(b), b> over all rows
 * (NekLinSysIter.cpp:104-109).  Blocks: nblk dense row-major matrices of
 * sizes nk = boff[k+1]-boff[k], stored one after the other in Ablk. */
typedef struct {
    int nloc, nglob, ndir;
    const int *map;
    const double *sign;
    double *wk;
} loc_map;

static void assemble_loc(const loc_map *m, const double *loc, double *glob)
{
    memset(glob, 0, (size_t)m->nglob * sizeof(double));
    for (int i = 0; i < m->nloc; ++i) glob[m->map[i]] += (m->sign ? m->sign[i] : 1.0) * loc[i];
}

static void to_local(const loc_map *m, const double *glob, double *loc)
{
    for (int i = 0; i < m->nloc; ++i) loc[i] = (m->sign ? m->sign[i] : 1.0) * glob[m->map[i]];
}

/* DoAssembleLoc(in, out, zero_dir) */
static void assemble_scatter(const loc_map *m, const double *in, double *out, int zero_dir)
{
    assemble_loc(m, in, m->wk);
    if (zero_dir) memset(m->wk, 0, (size_t)m->ndir * sizeof(double));
    to_local(m, m->wk, out);
}

static void precon_loc(const loc_map *m, const double *inv_diag, const double *in, double *out)
{
    assemble_loc(m, in, m->wk);
    for (int g = m->ndir; g < m->nglob; ++g) m->wk[g] *= inv_diag ? inv_diag[g] : 1.0;
    memset(m->wk, 0, (size_t)m->ndir * sizeof(double));
    to_local(m, m->wk, out);
}

static void block_apply(int nblk, const int *boff, const double *Ablk, const double *x, double *y)
{
    size_t off = 0;
    for (int k = 0; k < nblk; ++k) {
        const int b0 = boff[k], nk = boff[k + 1] - boff[k];
        for (int i = 0; i < nk; ++i) {
            double t = 0.0;
            for (int j = 0; j < nk; ++j) t += Ablk[off + (size_t)i * nk + j] * x[b0 + j];
            y[b0 + i] = t;
        }
        off += (size_t)nk * nk;
    }
}

int or_cg_loc(int nblk, const int *boff, const double *Ablk, const int *map, const double *sign, int nglob,
              int ndir, const double *inv_diag, const double *b_loc, double tol, int maxit, double *x_loc,
              or_cg_info *info)
{
    const int n = boff[nblk];
    loc_map m = {n, nglob, ndir, map, sign, (double *)calloc(nglob > 0 ? nglob : 1, sizeof(double))};
    double *w = (double *)calloc(n, sizeof(double));
    double *s = (double *)calloc(n, sizeof(double));
    double *p = (double *)calloc(n, sizeof(double));
    double *r = (double *)calloc(n, sizeof(double));
    double *q = (double *)calloc(n, sizeof(double));
    double *wk = (double *)calloc(n, sizeof(double));
    tol = fmax(tol, 1.0e-16); /* NekLinSysIter.cpp:57 */

    memcpy(r, b_loc, (size_t)n * sizeof(double));
    memset(x_loc, 0, (size_t)n * sizeof(double));
    assemble_scatter(&m, r, wk, 1);
    double eps = dot(n, wk, r);
    assemble_scatter(&m, b_loc, wk, 0);
    double rhs_mag = dot(n, wk, b_loc);
    rhs_mag = rhs_mag > 1.0e-6 ? rhs_mag : 1.0;
    int iters = 0;
    if (eps >= tol * tol * rhs_mag) {
        precon_loc(&m, inv_diag, r, w);
        block_apply(nblk, boff, Ablk, w, s);
        double rho = dot(n, r, w), mu = dot(n, s, w);
        double beta = 0.0, alpha = rho / mu, rho_new;
        iters = 1;
        while (iters <= maxit) {
            for (int i = 0; i < n; ++i) p[i] = beta * p[i] + w[i];
            for (int i = 0; i < n; ++i) q[i] = beta * q[i] + s[i];
            for (int i = 0; i < n; ++i) x_loc[i] = alpha * p[i] + x_loc[i];
            for (int i = 0; i < n; ++i) r[i] = -alpha * q[i] + r[i];
            precon_loc(&m, inv_diag, r, w);
            block_apply(nblk, boff, Ablk, w, s);
            rho_new = dot(n, r, w);
            mu = dot(n, s, w);
            assemble_scatter(&m, r, wk, 1);
            eps = dot(n, wk, r);
            iters++;
            if (eps < tol * tol * rhs_mag) break;
            beta = rho_new / rho;
            alpha = rho_new / (mu - rho_new * beta / alpha);
            rho = rho_new;
        }
    }
    if (info) {
        info->iterations = iters;
        info->final_error = sqrt(eps / rhs_mag);
        info->rhs_magnitude = rhs_mag;
    }
    free(m.wk); free(w); free(s); free(p); free(r); free(q); free(wk);
    return iters;
}

/* ---- static-condensation diagonal ------------------------------------
 * Preconditioner::AssembleStaticCondGlobalDiagonals (Preconditioner.cpp:
 * 185-236) with PreconditionerDiagonal::StaticCondDiagonalPreconditionerSum
 * (PreconditionerDiagonal.cpp:132-154) before the inversion: for every
 * element the Schur complement S = A_bb - A_bi A_ii^-1 A_ib of its matrix
 * (the boundary block of the statically condensed matrix), and
 * d[g] = sum over the element's boundary pairs (i, j) with map[i] = map[j]
 * = g of sign_i sign_j S_ij, for the non-Dirichlet g (Dirichlet and
 * interior-only rows stay 0).  interior[i] flags the interior local
 * entries.  A_ii^-1 by a dense Cholesky factorisation.  Returns 0, or
 * 1 + the block whose interior block is not positive definite. */
int or_sc_diagonal(int nblk, const int *boff, const double *Ablk, const uint8_t *interior, const int *map,
                   const double *sign, int nglob, int ndir, double *d)
{
    memset(d, 0, (size_t)nglob * sizeof(double));
    size_t off = 0;
    int ret = 0;
    for (int k = 0; k < nblk && !ret; ++k) {
        const int b0 = boff[k], n = boff[k + 1] - boff[k];
        const double *A = Ablk + off;
        off += (size_t)n * n;
        int *ii = (int *)malloc((size_t)n * sizeof(int)), *bb = (int *)malloc((size_t)n * sizeof(int));
        int ni = 0, nb = 0;
        for (int i = 0; i < n; ++i) {
            if (interior[b0 + i]) ii[ni++] = i;
            else bb[nb++] = i;
        }
        /* L L^T = A_ii (row-major lower triangle) */
        double *L = (double *)calloc((size_t)(ni > 0 ? ni : 1) * (ni > 0 ? ni : 1), sizeof(double));
        for (int c = 0; c < ni && !ret; ++c) {
            double s = A[(size_t)ii[c] * n + ii[c]];
            for (int j = 0; j < c; ++j) s -= L[(size_t)c * ni + j] * L[(size_t)c * ni + j];
            if (!(s > 0.0)) { ret = k + 1; break; }
            const double dcc = sqrt(s);
            L[(size_t)c * ni + c] = dcc;
            for (int r = c + 1; r < ni; ++r) {
                double t = A[(size_t)ii[r] * n + ii[c]];
                for (int j = 0; j < c; ++j) t -= L[(size_t)r * ni + j] * L[(size_t)c * ni + j];
                L[(size_t)r * ni + c] = t / dcc;
            }
        }
        /* Y[:, j] = L^-1 a_j for every boundary j; S_ij = A_ij - Y_i . Y_j */
        double *Y = (double *)calloc((size_t)(ni > 0 ? ni : 1) * (nb > 0 ? nb : 1), sizeof(double));
        for (int j = 0; j < nb && !ret; ++j) {
            for (int r = 0; r < ni; ++r) {
                double t = A[(size_t)ii[r] * n + bb[j]];
                for (int c = 0; c < r; ++c) t -= L[(size_t)r * ni + c] * Y[(size_t)c * nb + j];
                Y[(size_t)r * nb + j] = t / L[(size_t)r * ni + r];
            }
        }
        for (int i = 0; i < nb && !ret; ++i) {
            const int gi = map[b0 + bb[i]];
            if (gi < ndir) continue;
            for (int j = 0; j < nb; ++j) {
                if (map[b0 + bb[j]] != gi) continue;
                double sij = A[(size_t)bb[i] * n + bb[j]];
                for (int r = 0; r < ni; ++r) sij -= Y[(size_t)r * nb + i] * Y[(size_t)r * nb + j];
                const double si = sign ? sign[b0 + bb[i]] : 1.0, sj = sign ? sign[b0 + bb[j]] : 1.0;
                d[gi] += si * sj * sij;
            }
        }
        free(ii); free(bb); free(L); free(Y);
    }
    return ret;
}
