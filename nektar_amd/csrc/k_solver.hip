// Order-independent kernels of the structured-hex C0 Helmholtz solver.
//
// Global vectors hold every C0 degree of freedom of the n^3 mesh (N^3 with
// N = n*(nm-1)+1), Dirichlet rows included and kept at zero, so the
// local->global map is pure index arithmetic and never stored.
//
// CG (NekLinSysIterCG.cpp:118-372): every iteration is four launches -
// update (p, q, x, r, w and the partial dots r.w, r.r), gathered operator
// apply, assembly (with the partial dot s.w) and a one-block scalar step that
// finishes the reductions in a fixed order, tests ||r||^2 < tol^2 ||b||^2 and
// forms beta and alpha.  All state (alpha, beta, rho, ..., iteration count,
// done flag) lives in device memory, so a fixed batch of iterations can be
// launched (or replayed from a HIP graph) without host round trips; kernels
// of a finished solve return immediately.  Reductions use a fixed grid and a
// fixed combination order, so results are bitwise reproducible.
#include "solver.hpp"

#include <cmath>

namespace nektar_amd {
namespace solver {

namespace {

__device__ __forceinline__ bool on_boundary(int gx, int gy, int gz, int N)
{
    return gx == 0 || gy == 0 || gz == 0 || gx == N - 1 || gy == N - 1 || gz == N - 1;
}

// Block-wide deterministic sum of K values per thread; result in out[] of
// thread 0.  Each 64-lane wave folds its values with a fixed xor butterfly
// (cross-lane moves, no LDS round trips), lane 0 of every wave parks the
// wave's sums in LDS, and thread 0 adds the kBlock/64 wave sums in wave
// order: one barrier instead of log2(kBlock) + 1.  The summation order is
// fixed, so repeated runs give identical bits.  A caller that runs two
// block_sum<K> of the same K back to back separates them with a barrier.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double *out)
{
    constexpr int W = 64, NW = kBlock / W;
    static_assert(kBlock % W == 0, "block must be whole waves");
#pragma unroll
    for (int c = 0; c < K; ++c) {
#pragma unroll
        for (int o = W / 2; o > 0; o >>= 1) v[c] += __shfl_xor(v[c], o, W);
    }
    __shared__ double red[K][NW];
    const int lane = threadIdx.x % W, wv = threadIdx.x / W;
    if (lane == 0) {
#pragma unroll
        for (int c = 0; c < K; ++c) red[c][wv] = v[c];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < K; ++c) {
            double s = red[c][0];
#pragma unroll
            for (int w = 1; w < NW; ++w) s += red[c][w];
            out[c] = s;
        }
    }
}

// Entries of 1-D global node g: element index and local mode (<= 2
// entries).  NM is a compile-time constant, so the divisions by NM-1 are
// multiplications.
template <int NM>
__device__ __forceinline__ int node_contrib(int g, int n, int (&e)[2], int (&p)[2])
{
    constexpr int s = NM - 1;
    const int v = g / s, m = g - v * s;
    if (m == 0) {
        int c = 0;
        if (v > 0) {
            e[c] = v - 1;
            p[c] = 1;
            ++c;
        }
        if (v < n) {
            e[c] = v;
            p[c] = 0;
            ++c;
        }
        return c;
    }
    e[0] = v;
    p[0] = m + 1;
    return 1;
}

__device__ __forceinline__ void trilinear_deriv(const double *__restrict__ X, int n, int ex, int ey, int ez,
                                                double x0, double x1, double x2, double M[3][3], double *pos)
{
    const int nv = n + 1;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) M[a][b] = 0.0;
    if (pos) pos[0] = pos[1] = pos[2] = 0.0;
    for (int v = 0; v < 8; ++v) {
        const int sx = (v & 1) ? 1 : -1, sy = (v & 2) ? 1 : -1, sz = (v & 4) ? 1 : -1;
        const int ix = ex + (v & 1), iy = ey + ((v >> 1) & 1), iz = ez + ((v >> 2) & 1);
        const double *P = X + ((size_t)(iz * nv + iy) * nv + ix) * 3;
        const double f0 = 1 + sx * x0, f1 = 1 + sy * x1, f2 = 1 + sz * x2;
        const double dn[3] = {sx * f1 * f2 * 0.125, f0 * sy * f2 * 0.125, f0 * f1 * sz * 0.125};
        for (int k = 0; k < 3; ++k) {
            for (int i = 0; i < 3; ++i) M[k][i] += dn[i] * P[k];  // M[k][i] = dx_k / dxi_i
            if (pos) pos[k] += f0 * f1 * f2 * 0.125 * P[k];
        }
    }
}

// J, and metric m_de = sum_j df[j*3+d] df[j*3+e] with df = (dx/dxi)^{-T}
__device__ __forceinline__ void metric_from(const double M[3][3], double &J, double m[6])
{
    const double c00 = M[1][1] * M[2][2] - M[1][2] * M[2][1];
    const double c01 = M[1][2] * M[2][0] - M[1][0] * M[2][2];
    const double c02 = M[1][0] * M[2][1] - M[1][1] * M[2][0];
    const double det = M[0][0] * c00 + M[0][1] * c01 + M[0][2] * c02;
    const double id = 1.0 / det;
    // inverse Minv[i][k] = d xi_i / d x_k ; df[j*3+d] = Minv[d][j]
    double Mi[3][3];
    Mi[0][0] = c00 * id;
    Mi[1][0] = c01 * id;
    Mi[2][0] = c02 * id;
    Mi[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) * id;
    Mi[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) * id;
    Mi[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) * id;
    Mi[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) * id;
    Mi[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) * id;
    Mi[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) * id;
    const int idx[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    for (int c = 0; c < 6; ++c) {
        const int d = idx[c][0], e = idx[c][1];
        m[c] = Mi[d][0] * Mi[e][0] + Mi[d][1] * Mi[e][1] + Mi[d][2] * Mi[e][2];
    }
    J = fabs(det);
}

// ---- general C0 assembly (AssemblyMapCG.cpp:2804-2897) for any mesh:
// the caller's local-to-global map with signs; Dirichlet DOFs first.

// Signed indices: an entry m >= 0 is index m with sign +1, m < 0 is index
// ~m with sign -1 (the orientation sign of AssemblyMapCG folded into the
// 32-bit index, so the gathers read no separate sign array).
__device__ __forceinline__ int sidx(int m) { return m < 0 ? ~m : m; }
__device__ __forceinline__ double sapply(int m, double v) { return m < 0 ? -v : v; }

// loc[i] = sign[i] glob[map[i]]   (GlobalToLocal, Vmath::Gathr with signs)
__global__ void gather_signed_kernel(size_t nloc, const int *__restrict__ menc, const double *__restrict__ glob,
                                     double *__restrict__ loc)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nloc; i += (size_t)gridDim.x * blockDim.x) {
        const int m = menc[i];
        loc[i] = sapply(m, glob[sidx(m)]);
    }
}

// glob[g] = sum over the local entries of g, in increasing local index, of
// sign loc (Assemble, Vmath::Assmb): the transpose of the map as CSR rows,
// so every global value is one work-item's fixed-order sum.  Rows g < ndir
// are set to zero when zero_dir; with wdot, the partial sums of glob.wdot
// (the CG's s.w) per block.  sq: sum sign^2 loc instead (diagonals).
__global__ void assemble_csr_kernel(int nglob, int ndir, const int *__restrict__ row, const int *__restrict__ cenc,
                                    const double *__restrict__ loc, double *__restrict__ glob,
                                    const double *__restrict__ wdot, double *__restrict__ partials, int zero_dir,
                                    int sq)
{
    double acc[1] = {0.0};
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < nglob; g += gridDim.x * blockDim.x) {
        double v = 0.0;
        if (!(zero_dir && g < ndir)) {
            for (int k = row[g]; k < row[g + 1]; ++k) {
                const int c = cenc[k];
                const double lv = loc[sidx(c)];
                v += sq ? lv : sapply(c, lv);
            }
        }
        glob[g] = v;
        if (wdot) acc[0] += v * wdot[g];
    }
    if (partials) block_sum<1>(acc, partials + blockIdx.x);
}

// x_e = unit vector m of every element of a collection (ncoeff nc)
__global__ void unit_mode_kernel(size_t n, int nc, int m, double *__restrict__ x)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        x[i] = (int)(i % nc) == m ? 1.0 : 0.0;
}

// d[e*nc + m] = y[e*nc + m] (the diagonal entry m of every element matrix)
__global__ void take_mode_kernel(size_t nelmt, int nc, int m, const double *__restrict__ y, double *__restrict__ d)
{
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nelmt; e += (size_t)gridDim.x * blockDim.x)
        d[e * nc + m] = y[e * nc + m];
}

// 1/d on rows g >= ndir, 0 on the Dirichlet rows
__global__ void invert_diag_dir_kernel(int nglob, int ndir, double *__restrict__ d)
{
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < nglob; g += gridDim.x * blockDim.x)
        d[g] = g < ndir ? 0.0 : 1.0 / d[g];
}

// NekLinSysIterCG.cpp:118-213 with the Dirichlet DOFs first: r = b on rows
// g >= ndir (0 on the Dirichlet rows), w = M^-1 r, x = p = q = 0; partial
// sums r.r, r.w and b.b (over all rows, Set_Rhs_Magnitude).
__global__ void cg_init_dirfirst_kernel(int nglob, int ndir, const double *__restrict__ b,
                                        const double *__restrict__ inv_diag, double *__restrict__ x,
                                        double *__restrict__ r, double *__restrict__ p, double *__restrict__ q,
                                        double *__restrict__ w, double *__restrict__ partials)
{
    double acc[3] = {0.0, 0.0, 0.0};
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < nglob; g += gridDim.x * blockDim.x) {
        const double bv = b[g];
        acc[2] += bv * bv;
        const double rv = g < ndir ? 0.0 : bv;
        const double wv = inv_diag[g] * rv;
        x[g] = 0.0;
        p[g] = 0.0;
        q[g] = 0.0;
        r[g] = rv;
        w[g] = wv;
        acc[0] += rv * rv;  // the order cg_scalar_init reads: r.r, r.w, b.b
        acc[1] += rv * wv;
    }
    block_sum<3>(acc, partials + 3 * blockIdx.x);
}

// ---- element-local CG (NekLinSysIterCGLoc.cpp:114-308): the Krylov
// vectors are element-local; only the preconditioner assembles.

// PreconditionerDiagonal.cpp:174-184 with Preconditioner::DoAssembleLoc
// (Preconditioner.cpp:89-102) folded in: v_g = Assemble(r)_g (CSR rows,
// fixed order), wg_g = inv_diag_g v_g off the Dirichlet rows (0 on them),
// to be scattered back by the signed gather.  The local dots follow from
// the assembled values: <r, GlobalToLocal(wg)> = sum_g v_g wg_g (r.w) and
// <r, AssembleLoc(r, zero Dirichlet)> = sum_{g >= ndir} v_g^2 (eps).
// INIT: partials {eps, r.w, b.b (all rows: Set_Rhs_Magnitude,
// NekLinSysIter.cpp:104-109)} in cg_scalar_init's order; else {r.w, eps}
// in cg_scalar's.
template <int INIT>
__global__ void cgloc_precon_kernel(int nglob, int ndir, const int *__restrict__ row, const int *__restrict__ cenc,
                                    const double *__restrict__ r, const double *__restrict__ inv_diag,
                                    double *__restrict__ wg, double *__restrict__ partials)
{
    double acc[INIT ? 3 : 2] = {};
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < nglob; g += gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int k = row[g]; k < row[g + 1]; ++k) {
            const int c = cenc[k];
            v += sapply(c, r[sidx(c)]);
        }
        const bool dir = g < ndir;
        const double wv = dir ? 0.0 : inv_diag[g] * v;
        const double e = dir ? 0.0 : v * v;
        wg[g] = wv;
        if constexpr (INIT) {
            acc[0] += e;
            acc[1] += v * wv;
            acc[2] += v * v;
        } else {
            acc[0] += v * wv;
            acc[1] += e;
        }
    }
    block_sum<INIT ? 3 : 2>(acc, partials + (INIT ? 3 : 2) * blockIdx.x);
}

// NekLinSysIterCGLoc.cpp:210-221 on the local vectors: search directions,
// solution and residual
__global__ void cgloc_update_kernel(size_t n, const CGState *__restrict__ st, double *__restrict__ p,
                                    double *__restrict__ q, double *__restrict__ x, double *__restrict__ r,
                                    const double *__restrict__ w, const double *__restrict__ s)
{
    if (st->done) return;
    const double alpha = st->alpha, beta = st->beta;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double pv = beta * p[i] + w[i];
        const double qv = beta * q[i] + s[i];
        x[i] = alpha * pv + x[i];
        r[i] = -alpha * qv + r[i];
        p[i] = pv;
        q[i] = qv;
    }
}

// partial sums of a.b over a grid-stride loop (one per block)
__global__ void dot_kernel(size_t n, const double *__restrict__ a, const double *__restrict__ b,
                           double *__restrict__ partials)
{
    double acc[1] = {0.0};
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        acc[0] += a[i] * b[i];
    block_sum<1>(acc, partials + blockIdx.x);
}

// dst[idx[i]] = src[i]
__global__ void scatter_index_kernel(int n, const int *__restrict__ idx, const double *__restrict__ src,
                                     double *__restrict__ dst)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[idx[i]] = src[i];
}

// glob[g] = sign loc of the first local entry of g (0 for a DOF no element
// holds): the global values of a continuous local vector
__global__ void csr_first_kernel(int nglob, const int *__restrict__ row, const int *__restrict__ cenc,
                                 const double *__restrict__ loc, double *__restrict__ glob)
{
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < nglob; g += gridDim.x * blockDim.x) {
        const int k = row[g];
        glob[g] = k < row[g + 1] ? sapply(cenc[k], loc[sidx(cenc[k])]) : 0.0;
    }
}

}  // namespace

// Packed Helmholtz geometry of the structured mesh from its vertices
// (GeomFactors.cpp:150-477 for trilinear maps).
__global__ void geom_kernel(int n, int NQ, const double *__restrict__ X, const double *__restrict__ z,
                            const double *__restrict__ w, int deformed, double *__restrict__ geo)
{
    const size_t nelmt = (size_t)n * n * n, NQ3 = (size_t)NQ * NQ * NQ;
    const size_t total = deformed ? nelmt * NQ3 : nelmt;
    for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < total; id += (size_t)gridDim.x * blockDim.x) {
        const size_t e = deformed ? id / NQ3 : id;
        const int q = deformed ? (int)(id % NQ3) : 0;
        const int ex = (int)(e % n), ey = (int)((e / n) % n), ez = (int)(e / ((size_t)n * n));
        const int i = q % NQ, j = (q / NQ) % NQ, k = q / (NQ * NQ);
        double M[3][3], J, m[6];
        trilinear_deriv(X, n, ex, ey, ez, deformed ? z[i] : 0.0, deformed ? z[j] : 0.0, deformed ? z[k] : 0.0, M,
                        nullptr);
        metric_from(M, J, m);
        if (deformed) {
            const double wj = w[i] * w[j] * w[k] * J;
            for (int c = 0; c < 6; ++c) geo[(e * 7 + c) * NQ3 + q] = wj * m[c];
            geo[(e * 7 + 6) * NQ3 + q] = wj;
        } else {
            for (int c = 0; c < 6; ++c) geo[e * 7 + c] = J * m[c];
            geo[e * 7 + 6] = J;
        }
    }
}

// f = (3 k^2 pi^2 + lambda) prod_d sin(k pi x_d) at the quadrature points:
// the forcing of -lap(u) + lambda u = f with u = prod sin(k pi x_d).
__global__ void forcing_kernel(int n, int NQ, const double *__restrict__ X, const double *__restrict__ z,
                               int kw, double lambda, double *__restrict__ f)
{
    const size_t NQ3 = (size_t)NQ * NQ * NQ, total = (size_t)n * n * n * NQ3;
    const double kp = kw * M_PI, amp = 3.0 * kp * kp + lambda;
    for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < total; id += (size_t)gridDim.x * blockDim.x) {
        const size_t e = id / NQ3;
        const int q = (int)(id % NQ3);
        const int ex = (int)(e % n), ey = (int)((e / n) % n), ez = (int)(e / ((size_t)n * n));
        double M[3][3], x[3];
        trilinear_deriv(X, n, ex, ey, ez, z[q % NQ], z[(q / NQ) % NQ], z[q / (NQ * NQ)], M, x);
        f[id] = amp * sin(kp * x[0]) * sin(kp * x[1]) * sin(kp * x[2]);
    }
}

// sum_e sum_q W J (u - u_exact)^2, u_exact = prod sin(k pi x_d); partials
// per block (deterministic).
__global__ void l2err_kernel(int n, int NQ, const double *__restrict__ X, const double *__restrict__ z,
                             const double *__restrict__ w, const double *__restrict__ geo, int deformed, int kw,
                             const double *__restrict__ u, double *__restrict__ partials)
{
    const size_t NQ3 = (size_t)NQ * NQ * NQ, total = (size_t)n * n * n * NQ3;
    const double kp = kw * M_PI;
    double acc[1] = {0.0};
    for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < total; id += (size_t)gridDim.x * blockDim.x) {
        const size_t e = id / NQ3;
        const int q = (int)(id % NQ3);
        const int i = q % NQ, j = (q / NQ) % NQ, k = q / (NQ * NQ);
        const int ex = (int)(e % n), ey = (int)((e / n) % n), ez = (int)(e / ((size_t)n * n));
        double M[3][3], x[3];
        trilinear_deriv(X, n, ex, ey, ez, z[i], z[j], z[k], M, x);
        const double ue = sin(kp * x[0]) * sin(kp * x[1]) * sin(kp * x[2]);
        const double wj = deformed ? geo[(e * 7 + 6) * NQ3 + q] : geo[e * 7 + 6] * w[i] * w[j] * w[k];
        const double d = u[id] - ue;
        acc[0] += wj * d * d;
    }
    block_sum<1>(acc, partials + blockIdx.x);
}

// Direct-stiffness summation of element results into the global vector
// (AssemblyMapCG::Assemble, AssemblyMapCG.cpp:2860-2886) as a gather: each
// global row sums its <= 8 element contributions in a fixed order.
// Dirichlet rows are set to 0.  Optionally accumulates s . wdot.
template <typename T, int NM>
__global__ __launch_bounds__(kBlock) void assemble_kernel(int n, int N, const T *__restrict__ sloc,
                                                          T *__restrict__ sg, const T *__restrict__ wdot,
                                                          double *__restrict__ partials,
                                                          const int *__restrict__ done, int keep_dirichlet)
{
    if (done && *done) return;
    // One wave per grid line (gy, gz) of global nodes: the y and z
    // contributions and the element-row offsets are wave-uniform, lanes
    // run along gx (consecutive gx are consecutive modes of one element
    // row, so the reads of sloc are nearly contiguous).
    constexpr int WAVES = kBlock / 64;
    constexpr size_t NC = (size_t)NM * NM * NM;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const long lines = (long)N * N;
    double acc[1] = {0.0};
    for (long L = (long)blockIdx.x * WAVES + wave; L < lines; L += (long)gridDim.x * WAVES) {
        const int gz = (int)(L / N), gy = (int)(L - (long)gz * N);
        int ey[2], py[2], ez[2], pz[2];
        const int cy = node_contrib<NM>(gy, n, ey, py), cz = node_contrib<NM>(gz, n, ez, pz);
        const bool edge_line = gy == 0 || gz == 0 || gy == N - 1 || gz == N - 1;
        size_t base[4] = {0, 0, 0, 0};
        int nrow = 0;
        for (int c = 0; c < cz; ++c)
            for (int b = 0; b < cy; ++b)
                base[nrow++] = ((size_t)ez[c] * n + ey[b]) * n * NC + ((size_t)pz[c] * NM + py[b]) * NM;
        T *out = sg + (size_t)L * N;
        const T *wl = wdot ? wdot + (size_t)L * N : nullptr;
        // CH chunks of 64 nodes at a time with every load in its own
        // register (the row loop has a wave-uniform bound; the second x
        // contribution exists only at element vertices), so all reads of
        // the CH chunks are in flight together; sums in the fixed order
        // rows (z, y) then x, absent terms 0 (adding 0 is exact)
        constexpr int CH = sizeof(T) == 8 ? 3 : 2;  // measured best per type
        for (int gx0 = lane; gx0 < N; gx0 += 64 * CH) {
            T v[CH][4][2];
            bool live[CH];
#pragma unroll
            for (int h = 0; h < CH; ++h) {
                const int gx = gx0 + 64 * h;
                live[h] = gx < N && (keep_dirichlet || !(edge_line || gx == 0 || gx == N - 1));
                int ex[2] = {0, 0}, px[2] = {0, 0};
                const int cx = live[h] ? node_contrib<NM>(gx, n, ex, px) : 0;
                // offsets within one element row of the line (< 2^31 for
                // n <= 255 elements per direction at NM <= 10)
                const int o0 = ex[0] * (int)NC + px[0], o1 = ex[1] * (int)NC + px[1];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    v[h][r][0] = v[h][r][1] = T(0);
                    if (r < nrow) {  // wave-uniform
                        const T *row = sloc + base[r];
                        if (cx > 0) v[h][r][0] = row[o0];
                        if (cx > 1) v[h][r][1] = row[o1];
                    }
                }
            }
#pragma unroll
            for (int h = 0; h < CH; ++h) {
                const int gx = gx0 + 64 * h;
                if (gx >= N) continue;
                double s = 0.0;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    s += (double)v[h][r][0];
                    s += (double)v[h][r][1];
                }
                out[gx] = T(s);
                if (wl && live[h]) acc[0] += (double)T(s) * (double)wl[gx];
            }
        }
    }
    if (partials) block_sum<1>(acc, partials + blockIdx.x);
}

// r = b (non-Dirichlet rows), x = p = q = 0, w = M^{-1} r; partials
// {r.r, r.w, b.b(all rows)}.
template <typename T>
__global__ void cg_init_kernel(int N, const double *__restrict__ b, const double *__restrict__ inv_diag,
                               T *__restrict__ x, T *__restrict__ r, T *__restrict__ p, T *__restrict__ q,
                               T *__restrict__ w, double *__restrict__ partials)
{
    const size_t NG = (size_t)N * N * N;
    double acc[3] = {0.0, 0.0, 0.0};
    for (size_t G = blockIdx.x * (size_t)blockDim.x + threadIdx.x; G < NG; G += (size_t)gridDim.x * blockDim.x) {
        const int gx = (int)(G % N), gy = (int)((G / N) % N), gz = (int)(G / ((size_t)N * N));
        const double bv = b[G];
        acc[2] += bv * bv;
        const T rv = on_boundary(gx, gy, gz, N) ? T(0) : T(bv);
        const T wv = T(inv_diag[G] * (double)rv);
        x[G] = T(0);
        p[G] = T(0);
        q[G] = T(0);
        r[G] = rv;
        w[G] = wv;
        acc[0] += (double)rv * (double)rv;
        acc[1] += (double)rv * (double)wv;
    }
    block_sum<3>(acc, partials + 3 * blockIdx.x);
}

// NekLinSysIterCG.cpp:240-249 (search directions, solution, residual) and
// :264-307 (preconditioner, r.w, r.r).
template <typename T>
__global__ void cg_update_kernel(size_t NG, const CGState *__restrict__ st, T *__restrict__ p, T *__restrict__ q,
                                 T *__restrict__ x, T *__restrict__ r, T *__restrict__ w, const T *__restrict__ s,
                                 const double *__restrict__ inv_diag, double *__restrict__ partials)
{
    if (st->done) return;
    const T alpha = T(st->alpha), beta = T(st->beta);
    double acc[2] = {0.0, 0.0};
    for (size_t G = blockIdx.x * (size_t)blockDim.x + threadIdx.x; G < NG; G += (size_t)gridDim.x * blockDim.x) {
        const T pv = beta * p[G] + w[G];
        const T qv = beta * q[G] + s[G];
        x[G] = alpha * pv + x[G];
        const T rv = -alpha * qv + r[G];
        const T wv = T(inv_diag[G] * (double)rv);
        p[G] = pv;
        q[G] = qv;
        r[G] = rv;
        w[G] = wv;
        acc[0] += (double)rv * (double)wv;
        acc[1] += (double)rv * (double)rv;
    }
    block_sum<2>(acc, partials + 2 * blockIdx.x);
}

__device__ __forceinline__ void reduce_partials(const double *__restrict__ pa, int stride, int comp, int nb,
                                                double *out)
{
    double v[1] = {0.0};
    for (int b = threadIdx.x; b < nb; b += kBlock) v[0] += pa[(size_t)b * stride + comp];
    block_sum<1>(v, out);
}

__global__ void cg_scalar_init_kernel(CGState *st, const double *__restrict__ pa, const double *__restrict__ pmu,
                                      int nb)
{
    __shared__ double res[4];
    reduce_partials(pa, 3, 0, nb, &res[0]);
    __syncthreads();
    reduce_partials(pa, 3, 1, nb, &res[1]);
    __syncthreads();
    reduce_partials(pa, 3, 2, nb, &res[2]);
    __syncthreads();
    reduce_partials(pmu, 1, 0, nb, &res[3]);
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double eps = res[0], rho = res[1], bb = res[2], mu = res[3];
    st->rhs_mag = st->rhs_mag_override > 0.0 ? st->rhs_mag_override : (bb > 1.0e-6 ? bb : 1.0);  // NekLinSysIter.cpp:116
    st->eps = eps;
    st->iter = 0;
    st->converged = 0;
    st->done = 0;
    st->final_error = sqrt(eps / st->rhs_mag);
    if (eps < st->tol2 * st->rhs_mag) {  // NekLinSysIterCG.cpp:169
        st->done = 1;
        st->converged = 1;
        return;
    }
    st->rho = rho;
    st->mu = mu;
    st->beta = 0.0;
    st->alpha = rho / mu;
    st->iter = 1;
}

__global__ void cg_scalar_kernel(CGState *st, const double *__restrict__ pa, const double *__restrict__ pmu, int nb)
{
    if (st->done) return;
    __shared__ double res[3];
    reduce_partials(pa, 2, 0, nb, &res[0]);
    __syncthreads();
    reduce_partials(pa, 2, 1, nb, &res[1]);
    __syncthreads();
    reduce_partials(pmu, 1, 0, nb, &res[2]);
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double rho_new = res[0], eps = res[1], mu = res[2];
    st->iter += 1;
    st->eps = eps;
    st->final_error = sqrt(eps / st->rhs_mag);
    if (eps < st->tol2 * st->rhs_mag) {  // NekLinSysIterCG.cpp:340
        st->done = 1;
        st->converged = 1;
        return;
    }
    const double beta = rho_new / st->rho;  // :358 (rho_star = 0)
    st->alpha = rho_new / (mu - rho_new * beta / st->alpha);
    st->beta = beta;
    st->rho = rho_new;
    st->mu = mu;
    if (st->iter > st->maxit) st->done = 1;  // :215 (checked at the loop top)
}

// ---- Hestenes-Stiefel PCG ------------------------------------------
// r = b (non-Dirichlet), x = 0, p = z = M^{-1} r; partials {r.z, r.r, b.b}
template <typename T>
__global__ void pcg_init_kernel(int N, const double *__restrict__ b, const T *__restrict__ inv_diag,
                                T *__restrict__ x, T *__restrict__ r, T *__restrict__ p, double *__restrict__ partials)
{
    const size_t NG = (size_t)N * N * N;
    double acc[3] = {0.0, 0.0, 0.0};
    for (size_t G = blockIdx.x * (size_t)blockDim.x + threadIdx.x; G < NG; G += (size_t)gridDim.x * blockDim.x) {
        const int gx = (int)(G % N), gy = (int)((G / N) % N), gz = (int)(G / ((size_t)N * N));
        const double bv = b[G];
        acc[2] += bv * bv;
        const T rv = on_boundary(gx, gy, gz, N) ? T(0) : T(bv);
        const double zv = (double)inv_diag[G] * (double)rv;
        x[G] = T(0);
        r[G] = rv;
        p[G] = T(zv);
        acc[0] += (double)rv * zv;
        acc[1] += (double)rv * (double)rv;
    }
    block_sum<3>(acc, partials + 3 * blockIdx.x);
}

__global__ void pcg_scalar_init_kernel(CGState *st, const double *__restrict__ pa, int nb)
{
    __shared__ double res[3];
    reduce_partials(pa, 3, 0, nb, &res[0]);
    __syncthreads();
    reduce_partials(pa, 3, 1, nb, &res[1]);
    __syncthreads();
    reduce_partials(pa, 3, 2, nb, &res[2]);
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double rz = res[0], eps = res[1], bb = res[2];
    st->rhs_mag = st->rhs_mag_override > 0.0 ? st->rhs_mag_override : (bb > 1.0e-6 ? bb : 1.0);
    st->eps = eps;
    st->rho = rz;
    st->iter = 0;
    st->converged = 0;
    st->done = 0;
    st->final_error = sqrt(eps / st->rhs_mag);
    if (eps < st->tol2 * st->rhs_mag) {
        st->done = 1;
        st->converged = 1;
    }
}

__global__ void pcg_alpha_kernel(CGState *st, const double *__restrict__ pmu, int nb)
{
    if (st->done || st->pause) return;
    __shared__ double res[1];
    reduce_partials(pmu, 1, 0, nb, &res[0]);
    __syncthreads();
    if (threadIdx.x != 0) return;
    st->mu = res[0];
    st->alpha = st->rho / res[0];
}

// x += alpha p, r -= alpha q, partials {r.z, r.r} with z = M^{-1} r
template <typename T>
__global__ void pcg_update_xr_kernel(size_t NG, const CGState *__restrict__ st, T *__restrict__ x, T *__restrict__ r,
                                     const T *__restrict__ p, const T *__restrict__ q,
                                     const T *__restrict__ inv_diag, double *__restrict__ partials)
{
    if (st->done || st->pause) return;
    const T alpha = T(st->alpha);
    double acc[2] = {0.0, 0.0};
    for (size_t G = blockIdx.x * (size_t)blockDim.x + threadIdx.x; G < NG; G += (size_t)gridDim.x * blockDim.x) {
        x[G] = alpha * p[G] + x[G];
        const T rv = -alpha * q[G] + r[G];
        r[G] = rv;
        acc[0] += (double)rv * ((double)inv_diag[G] * (double)rv);
        acc[1] += (double)rv * (double)rv;
    }
    block_sum<2>(acc, partials + 2 * blockIdx.x);
}

__global__ void pcg_beta_kernel(CGState *st, const double *__restrict__ pa, int nb)
{
    if (st->done || st->pause) return;
    __shared__ double res[2];
    reduce_partials(pa, 2, 0, nb, &res[0]);
    __syncthreads();
    reduce_partials(pa, 2, 1, nb, &res[1]);
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double rz = res[0], eps = res[1];
    st->iter += 1;
    st->eps = eps;
    st->final_error = sqrt(eps / st->rhs_mag);
    if (st->delta2 > 0.0) {
        // reliable-update mode: convergence is only declared on a true
        // residual, so a small recurrence residual suspends the iterations
        if (eps < st->tol2 * st->rhs_mag || eps < st->delta2 * st->rr_ref || st->iter >= st->maxit) {
            st->pause = 1;  // beta and rho are then formed from the true residual
        } else {
            st->beta = rz / st->rho;
            st->rho = rz;
        }
        return;
    }
    if (eps < st->tol2 * st->rhs_mag) {
        st->done = 1;
        st->converged = 1;
        return;
    }
    st->beta = rz / st->rho;
    st->rho = rz;
    if (st->iter >= st->maxit) st->done = 1;
}

// p = M^{-1} r + beta p
template <typename T>
__global__ void pcg_update_p_kernel(size_t NG, const CGState *__restrict__ st, const T *__restrict__ r,
                                    T *__restrict__ p, const T *__restrict__ inv_diag)
{
    if (st->done || st->pause) return;
    const double beta = st->beta;
    for (size_t G = blockIdx.x * (size_t)blockDim.x + threadIdx.x; G < NG; G += (size_t)gridDim.x * blockDim.x)
        p[G] = T((double)inv_diag[G] * (double)r[G] + beta * (double)p[G]);
}

// reliable update, part 1: r32 = r64 on non-Dirichlet rows; partial r.z
__global__ void ru_resume_kernel(int N, const double *__restrict__ r64, float *__restrict__ r32,
                                 const float *__restrict__ inv_diag, double *__restrict__ partials)
{
    const size_t NG = (size_t)N * N * N;
    double acc[1] = {0.0};
    for (size_t G = blockIdx.x * (size_t)blockDim.x + threadIdx.x; G < NG; G += (size_t)gridDim.x * blockDim.x) {
        const float rv = (float)r64[G];  // r64 is 0 on Dirichlet rows
        r32[G] = rv;
        acc[0] += (double)rv * ((double)inv_diag[G] * (double)rv);
    }
    block_sum<1>(acc, partials + blockIdx.x);
}

// reliable update, part 2: true residual test; beta restarted from r.z of
// the true residual (p = z + beta p follows in pcg_update_p).
__global__ void ru_scalar_kernel(CGState *st, const double *__restrict__ pa, const double *__restrict__ prr, int nb)
{
    __shared__ double res[2];
    reduce_partials(pa, 1, 0, nb, &res[0]);
    __syncthreads();
    reduce_partials(prr, 1, 0, nb, &res[1]);
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double rz = res[0], rr = res[1];
    st->rr_true = rr;
    st->rr_ref = rr;
    st->eps = rr;
    st->updates += 1;
    st->final_error = sqrt(rr / st->rhs_mag);
    if (rr < st->tol2 * st->rhs_mag) {
        st->done = 1;
        st->converged = 1;
        st->pause = 0;
        return;
    }
    if (st->iter >= st->maxit) {
        st->done = 1;
        st->pause = 0;
        return;
    }
    st->beta = rz / st->rho;
    st->rho = rz;
    st->pause = 0;
}

template <typename T>
__global__ void zero_kernel(size_t n, T *__restrict__ a)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        a[i] = T(0);
}

// r64 = b - s on non-Dirichlet rows (0 elsewhere); partial r.r
__global__ void residual_kernel(int N, const double *__restrict__ b, const double *__restrict__ s,
                                double *__restrict__ r, double *__restrict__ partials)
{
    const size_t NG = (size_t)N * N * N;
    double acc[1] = {0.0};
    for (size_t G = blockIdx.x * (size_t)blockDim.x + threadIdx.x; G < NG; G += (size_t)gridDim.x * blockDim.x) {
        const int gx = (int)(G % N), gy = (int)((G / N) % N), gz = (int)(G / ((size_t)N * N));
        const double v = on_boundary(gx, gy, gz, N) ? 0.0 : b[G] - s[G];
        r[G] = v;
        acc[0] += v * v;
    }
    block_sum<1>(acc, partials + blockIdx.x);
}

// x64 += d
template <typename T>
__global__ void add_kernel(size_t NG, double *__restrict__ x, const T *__restrict__ d)
{
    for (size_t G = blockIdx.x * (size_t)blockDim.x + threadIdx.x; G < NG; G += (size_t)gridDim.x * blockDim.x)
        x[G] += (double)d[G];
}

// out = a - b (out may alias a or b)
__global__ void sub_kernel(size_t n, const double *a, const double *b, double *out)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = a[i] - b[i];
}

// Robin boundary mass rows: y[rows[k]] += sum_m vals[m] x(cols[m]), with
// x local, or global through the signed map (entry c >= 0: +x[c], c < 0:
// -x[~c]); one work-item per row (rows distinct, so no two write one entry)
__global__ void robin_kernel(int64_t nrows, const int64_t *__restrict__ rows, const int64_t *__restrict__ rowptr,
                             const int64_t *__restrict__ cols, const double *__restrict__ vals,
                             const int *__restrict__ smap, const double *__restrict__ x, double *__restrict__ y)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nrows; k += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int64_t m = rowptr[k]; m < rowptr[k + 1]; ++m) {
            double xv;
            if (smap) {
                const int g = smap[cols[m]];
                xv = g >= 0 ? x[g] : -x[~g];
            } else {
                xv = x[cols[m]];
            }
            s = fma(vals[m], xv, s);
        }
        y[rows[k]] += s;
    }
}

// d[rows[k]] += the row's diagonal entry (cols[m] == rows[k])
__global__ void robin_diag_kernel(int64_t nrows, const int64_t *__restrict__ rows,
                                  const int64_t *__restrict__ rowptr, const int64_t *__restrict__ cols,
                                  const double *__restrict__ vals, double *__restrict__ d)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nrows; k += (int64_t)gridDim.x * blockDim.x)
        for (int64_t m = rowptr[k]; m < rowptr[k + 1]; ++m)
            if (cols[m] == rows[k]) d[rows[k]] += vals[m];
}

// inv_diag = 1/diag on non-Dirichlet rows, 0 on Dirichlet rows
__global__ void invert_diag_kernel(int N, double *__restrict__ d)
{
    const size_t NG = (size_t)N * N * N;
    for (size_t G = blockIdx.x * (size_t)blockDim.x + threadIdx.x; G < NG; G += (size_t)gridDim.x * blockDim.x) {
        const int gx = (int)(G % N), gy = (int)((G / N) % N), gz = (int)(G / ((size_t)N * N));
        d[G] = on_boundary(gx, gy, gz, N) ? 0.0 : 1.0 / d[G];
    }
}

__global__ void sum_partials_kernel(const double *__restrict__ pa, int nb, double *out)
{
    reduce_partials(pa, 1, 0, nb, out);
}

// Initial state of the mixed solve: relative tolerance against ||b||^2
// with the rhs-magnitude floor of NekLinSysIter::Set_Rhs_Magnitude
// (NekLinSysIter.cpp:96-117), reliable updates against ||r0||^2.
__global__ void mixed_state_init_kernel(CGState *st, const double *__restrict__ dsc, double tol2, int maxit,
                                        double delta2)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    CGState s{};
    s.tol2 = tol2;
    s.maxit = maxit;
    s.rhs_mag_override = dsc[0] > 1.0e-6 ? dsc[0] : 1.0;
    s.delta2 = delta2;
    s.rr_ref = dsc[1];
    *st = s;
}

__global__ void sumsq_kernel(size_t n, const double *__restrict__ a, double *__restrict__ partials)
{
    double acc[1] = {0.0};
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        acc[0] += a[i] * a[i];
    block_sum<1>(acc, partials + blockIdx.x);
}

__global__ void cast_kernel(size_t n, const double *__restrict__ a, float *__restrict__ b)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

// ---------------------------------------------------------------- launchers
#define NMF_LAUNCH(kernel, grid, ...) \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), 0, stream, __VA_ARGS__)

static int grid_for(size_t n)
{
    size_t g = (n + kBlock - 1) / kBlock;
    return (int)(g < (size_t)kRedBlocks ? (g > 0 ? g : 1) : kRedBlocks);
}

hipError_t launch_geom(hipStream_t stream, int n, int NQ, const double *X, const double *z, const double *w,
                       bool deformed, double *geo)
{
    const size_t total = (size_t)n * n * n * (deformed ? (size_t)NQ * NQ * NQ : 1);
    NMF_LAUNCH(geom_kernel, grid_for(total), n, NQ, X, z, w, deformed ? 1 : 0, geo);
    return hipGetLastError();
}

hipError_t launch_forcing(hipStream_t stream, int n, int NQ, const double *X, const double *z, int k, double lambda,
                          double *f)
{
    NMF_LAUNCH(forcing_kernel, grid_for((size_t)n * n * n * NQ * NQ * NQ), n, NQ, X, z, k, lambda, f);
    return hipGetLastError();
}

hipError_t launch_l2err(hipStream_t stream, int n, int NQ, const double *X, const double *z, const double *w,
                        const double *geo, bool deformed, int k, const double *u, double *partials, int *nb)
{
    *nb = grid_for((size_t)n * n * n * NQ * NQ * NQ);
    NMF_LAUNCH(l2err_kernel, *nb, n, NQ, X, z, w, geo, deformed ? 1 : 0, k, u, partials);
    return hipGetLastError();
}

int reduction_blocks(size_t NG) { return grid_for(NG); }

template <typename T>
hipError_t launch_assemble(hipStream_t stream, int n, int nm, int N, const T *sloc, T *sg, const T *wdot,
                           double *partials, const int *done, bool keep_dirichlet)
{
    const int grid = grid_for((size_t)N * N * N), keep = keep_dirichlet ? 1 : 0;
    switch (nm) {
#define NMF_ASM(M) \
    case M: NMF_LAUNCH((assemble_kernel<T, M>), grid, n, N, sloc, sg, wdot, partials, done, keep); break;
        NMF_ASM(2) NMF_ASM(3) NMF_ASM(4) NMF_ASM(5) NMF_ASM(6) NMF_ASM(7) NMF_ASM(8) NMF_ASM(9) NMF_ASM(10)
#undef NMF_ASM
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_cg_init(hipStream_t stream, int N, const double *b, const double *inv_diag, T *x, T *r, T *p, T *q,
                          T *w, double *partials)
{
    NMF_LAUNCH((cg_init_kernel<T>), grid_for((size_t)N * N * N), N, b, inv_diag, x, r, p, q, w, partials);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_cg_update(hipStream_t stream, int N, const CGState *st, T *p, T *q, T *x, T *r, T *w, const T *s,
                            const double *inv_diag, double *partials)
{
    const size_t NG = (size_t)N * N * N;
    NMF_LAUNCH((cg_update_kernel<T>), grid_for(NG), NG, st, p, q, x, r, w, s, inv_diag, partials);
    return hipGetLastError();
}

hipError_t launch_cg_scalar_init(hipStream_t stream, CGState *st, const double *pa, const double *pmu, int nb)
{
    NMF_LAUNCH(cg_scalar_init_kernel, 1, st, pa, pmu, nb);
    return hipGetLastError();
}

hipError_t launch_cg_scalar(hipStream_t stream, CGState *st, const double *pa, const double *pmu, int nb)
{
    NMF_LAUNCH(cg_scalar_kernel, 1, st, pa, pmu, nb);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_init(hipStream_t stream, int N, const double *b, const T *inv_diag, T *x, T *r, T *p,
                           double *partials)
{
    NMF_LAUNCH((pcg_init_kernel<T>), grid_for((size_t)N * N * N), N, b, inv_diag, x, r, p, partials);
    return hipGetLastError();
}

hipError_t launch_pcg_scalar_init(hipStream_t stream, CGState *st, const double *pa, int nb)
{
    NMF_LAUNCH(pcg_scalar_init_kernel, 1, st, pa, nb);
    return hipGetLastError();
}

hipError_t launch_pcg_alpha(hipStream_t stream, CGState *st, const double *pmu, int nb)
{
    NMF_LAUNCH(pcg_alpha_kernel, 1, st, pmu, nb);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_update_xr(hipStream_t stream, int N, const CGState *st, T *x, T *r, const T *p, const T *q,
                                const T *inv_diag, double *partials)
{
    const size_t NG = (size_t)N * N * N;
    NMF_LAUNCH((pcg_update_xr_kernel<T>), grid_for(NG), NG, st, x, r, p, q, inv_diag, partials);
    return hipGetLastError();
}

hipError_t launch_pcg_beta(hipStream_t stream, CGState *st, const double *pa, int nb)
{
    NMF_LAUNCH(pcg_beta_kernel, 1, st, pa, nb);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_update_p(hipStream_t stream, int N, const CGState *st, const T *r, T *p, const T *inv_diag)
{
    const size_t NG = (size_t)N * N * N;
    NMF_LAUNCH((pcg_update_p_kernel<T>), grid_for(NG), NG, st, r, p, inv_diag);
    return hipGetLastError();
}

template hipError_t launch_pcg_init<float>(hipStream_t, int, const double *, const float *, float *, float *, float *,
                                           double *);
template hipError_t launch_pcg_init<double>(hipStream_t, int, const double *, const double *, double *, double *,
                                            double *, double *);
template hipError_t launch_pcg_update_xr<float>(hipStream_t, int, const CGState *, float *, float *, const float *,
                                                const float *, const float *, double *);
template hipError_t launch_pcg_update_xr<double>(hipStream_t, int, const CGState *, double *, double *, const double *,
                                                 const double *, const double *, double *);
template hipError_t launch_pcg_update_p<float>(hipStream_t, int, const CGState *, const float *, float *,
                                               const float *);
template hipError_t launch_pcg_update_p<double>(hipStream_t, int, const CGState *, const double *, double *,
                                                const double *);

hipError_t launch_ru_resume(hipStream_t stream, int N, const double *r64, float *r32, const float *inv_diag,
                            double *partials)
{
    NMF_LAUNCH(ru_resume_kernel, grid_for((size_t)N * N * N), N, r64, r32, inv_diag, partials);
    return hipGetLastError();
}

hipError_t launch_ru_scalar(hipStream_t stream, CGState *st, const double *pa, const double *prr, int nb)
{
    NMF_LAUNCH(ru_scalar_kernel, 1, st, pa, prr, nb);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_zero(hipStream_t stream, size_t n, T *a)
{
    NMF_LAUNCH((zero_kernel<T>), grid_for(n), n, a);
    return hipGetLastError();
}
template hipError_t launch_zero<float>(hipStream_t, size_t, float *);
template hipError_t launch_zero<double>(hipStream_t, size_t, double *);

hipError_t launch_residual(hipStream_t stream, int N, const double *b, const double *s, double *r, double *partials)
{
    NMF_LAUNCH(residual_kernel, grid_for((size_t)N * N * N), N, b, s, r, partials);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_add(hipStream_t stream, size_t NG, double *x, const T *d)
{
    NMF_LAUNCH((add_kernel<T>), grid_for(NG), NG, x, d);
    return hipGetLastError();
}

hipError_t launch_sub(hipStream_t stream, size_t n, const double *a, const double *b, double *out)
{
    NMF_LAUNCH(sub_kernel, grid_for(n), n, a, b, out);
    return hipGetLastError();
}

hipError_t launch_robin(hipStream_t stream, int64_t nrows, const int64_t *rows, const int64_t *rowptr,
                        const int64_t *cols, const double *vals, const int *smap, const double *x, double *y)
{
    if (nrows <= 0) return hipSuccess;
    NMF_LAUNCH(robin_kernel, grid_for((size_t)nrows), nrows, rows, rowptr, cols, vals, smap, x, y);
    return hipGetLastError();
}

hipError_t launch_robin_diag(hipStream_t stream, int64_t nrows, const int64_t *rows, const int64_t *rowptr,
                             const int64_t *cols, const double *vals, double *d)
{
    if (nrows <= 0) return hipSuccess;
    NMF_LAUNCH(robin_diag_kernel, grid_for((size_t)nrows), nrows, rows, rowptr, cols, vals, d);
    return hipGetLastError();
}

hipError_t launch_invert_diag(hipStream_t stream, int N, double *d)
{
    NMF_LAUNCH(invert_diag_kernel, grid_for((size_t)N * N * N), N, d);
    return hipGetLastError();
}

hipError_t launch_sum_partials(hipStream_t stream, const double *pa, int nb, double *out)
{
    NMF_LAUNCH(sum_partials_kernel, 1, pa, nb, out);
    return hipGetLastError();
}

hipError_t launch_mixed_state_init(hipStream_t stream, CGState *st, const double *dsc, double tol2, int maxit,
                                  double delta2)
{
    NMF_LAUNCH(mixed_state_init_kernel, 1, st, dsc, tol2, maxit, delta2);
    return hipGetLastError();
}

hipError_t launch_sumsq(hipStream_t stream, size_t n, const double *a, double *partials)
{
    NMF_LAUNCH(sumsq_kernel, grid_for(n), n, a, partials);
    return hipGetLastError();
}

hipError_t launch_cast(hipStream_t stream, size_t n, const double *a, float *b)
{
    NMF_LAUNCH(cast_kernel, grid_for(n), n, a, b);
    return hipGetLastError();
}

template hipError_t launch_assemble<double>(hipStream_t, int, int, int, const double *, double *, const double *,
                                            double *, const int *, bool);
template hipError_t launch_assemble<float>(hipStream_t, int, int, int, const float *, float *, const float *, double *,
                                           const int *, bool);
template hipError_t launch_cg_init<double>(hipStream_t, int, const double *, const double *, double *, double *,
                                           double *, double *, double *, double *);
template hipError_t launch_cg_init<float>(hipStream_t, int, const double *, const double *, float *, float *, float *,
                                          float *, float *, double *);
hipError_t launch_gather_signed(hipStream_t stream, size_t nloc, const int *menc, const double *glob, double *loc)
{
    NMF_LAUNCH(gather_signed_kernel, grid_for(nloc), nloc, menc, glob, loc);
    return hipGetLastError();
}

hipError_t launch_assemble_csr(hipStream_t stream, int nglob, int ndir, const int *row, const int *cenc,
                               const double *loc, double *glob, const double *wdot, double *partials, bool zero_dir,
                               bool sq)
{
    NMF_LAUNCH(assemble_csr_kernel, grid_for((size_t)nglob), nglob, ndir, row, cenc, loc, glob, wdot, partials,
               zero_dir ? 1 : 0, sq ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_unit_mode(hipStream_t stream, size_t n, int nc, int m, double *x)
{
    NMF_LAUNCH(unit_mode_kernel, grid_for(n), n, nc, m, x);
    return hipGetLastError();
}

hipError_t launch_take_mode(hipStream_t stream, size_t nelmt, int nc, int m, const double *y, double *d)
{
    NMF_LAUNCH(take_mode_kernel, grid_for(nelmt), nelmt, nc, m, y, d);
    return hipGetLastError();
}

hipError_t launch_invert_diag_dir(hipStream_t stream, int nglob, int ndir, double *d)
{
    NMF_LAUNCH(invert_diag_dir_kernel, grid_for((size_t)nglob), nglob, ndir, d);
    return hipGetLastError();
}

hipError_t launch_cg_init_dirfirst(hipStream_t stream, int nglob, int ndir, const double *b, const double *inv_diag,
                                   double *x, double *r, double *p, double *q, double *w, double *partials)
{
    NMF_LAUNCH(cg_init_dirfirst_kernel, grid_for((size_t)nglob), nglob, ndir, b, inv_diag, x, r, p, q, w, partials);
    return hipGetLastError();
}

hipError_t launch_cg_update_n(hipStream_t stream, size_t NG, const CGState *st, double *p, double *q, double *x,
                              double *r, double *w, const double *s, const double *inv_diag, double *partials)
{
    NMF_LAUNCH((cg_update_kernel<double>), grid_for(NG), NG, st, p, q, x, r, w, s, inv_diag, partials);
    return hipGetLastError();
}

hipError_t launch_cgloc_precon(hipStream_t stream, bool init, int nglob, int ndir, const int *row, const int *cenc,
                               const double *r, const double *inv_diag, double *wg, double *partials)
{
    const int grid = grid_for((size_t)nglob);
    if (init)
        NMF_LAUNCH(cgloc_precon_kernel<1>, grid, nglob, ndir, row, cenc, r, inv_diag, wg, partials);
    else
        NMF_LAUNCH(cgloc_precon_kernel<0>, grid, nglob, ndir, row, cenc, r, inv_diag, wg, partials);
    return hipGetLastError();
}

hipError_t launch_cgloc_update(hipStream_t stream, size_t n, const CGState *st, double *p, double *q, double *x,
                               double *r, const double *w, const double *s)
{
    NMF_LAUNCH(cgloc_update_kernel, grid_for(n), n, st, p, q, x, r, w, s);
    return hipGetLastError();
}

hipError_t launch_dot(hipStream_t stream, size_t n, int nb, const double *a, const double *b, double *partials)
{
    if (nb < 1) return hipErrorInvalidValue;
    NMF_LAUNCH(dot_kernel, nb, n, a, b, partials);
    return hipGetLastError();
}

hipError_t launch_scatter_index(hipStream_t stream, int n, const int *idx, const double *src, double *dst)
{
    if (n <= 0) return hipSuccess;
    NMF_LAUNCH(scatter_index_kernel, grid_for((size_t)n), n, idx, src, dst);
    return hipGetLastError();
}

hipError_t launch_csr_first(hipStream_t stream, int nglob, const int *row, const int *cenc, const double *loc,
                            double *glob)
{
    NMF_LAUNCH(csr_first_kernel, grid_for((size_t)nglob), nglob, row, cenc, loc, glob);
    return hipGetLastError();
}

template hipError_t launch_cg_update<double>(hipStream_t, int, const CGState *, double *, double *, double *, double *,
                                             double *, const double *, const double *, double *);
template hipError_t launch_cg_update<float>(hipStream_t, int, const CGState *, float *, float *, float *, float *,
                                            float *, const float *, const double *, double *);
template hipError_t launch_add<float>(hipStream_t, size_t, double *, const float *);
template hipError_t launch_add<double>(hipStream_t, size_t, double *, const double *);

}  // namespace solver
}  // namespace nektar_amd
