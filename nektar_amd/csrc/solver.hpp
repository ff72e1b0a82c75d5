// C0 Helmholtz solver on structured hex meshes: the GPU counterpart of
// ContField::HelmSolve -> GlobalLinSysIterativeFull -> NekLinSysIterCG with
// the Diagonal (Jacobi) preconditioner, plus a mixed-precision
// iterative-refinement variant (FP32 inner CG, FP64 residuals, dots and
// diagonal).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "launch.hpp"
#include "tune.hpp"

namespace nektar_amd {

// Held while a solver captures its iterations into a hipGraph, and by
// anything that builds a solver on another thread: allocations,
// synchronous copies and stream synchronisation elsewhere in the process
// invalidate a capture in progress.
std::mutex &graph_capture_mutex();
namespace solver {

constexpr int kBlock = 256;
constexpr int kRedBlocks = 2048;

struct CGState {
    double alpha, beta, rho, mu, eps, rhs_mag, tol2, final_error;
    double rhs_mag_override;  // > 0: use instead of ||b||^2
    double delta2;            // reliable update when ||r||^2 < delta2 * rr_ref (0: never)
    double rr_ref;            // ||r||^2 of the true residual at the last update
    double rr_true;           // scratch: true residual norm^2 from the FP64 recomputation
    int iter, maxit, done, converged;
    int pause;                // 1: iterations suspended for a reliable update
    int updates;              // number of reliable updates performed
};

int reduction_blocks(size_t NG);
hipError_t launch_geom(hipStream_t, int n, int NQ, const double *X, const double *z, const double *w, bool deformed,
                       double *geo);
hipError_t launch_forcing(hipStream_t, int n, int NQ, const double *X, const double *z, int k, double lambda,
                          double *f);
hipError_t launch_l2err(hipStream_t, int n, int NQ, const double *X, const double *z, const double *w,
                        const double *geo, bool deformed, int k, const double *u, double *partials, int *nb);
template <typename T>
hipError_t launch_assemble(hipStream_t, int n, int nm, int N, const T *sloc, T *sg, const T *wdot, double *partials,
                           const int *done, bool keep_dirichlet = false);
template <typename T>
hipError_t launch_cg_init(hipStream_t, int N, const double *b, const double *inv_diag, T *x, T *r, T *p, T *q, T *w,
                          double *partials);
template <typename T>
hipError_t launch_cg_update(hipStream_t, int N, const CGState *st, T *p, T *q, T *x, T *r, T *w, const T *s,
                            const double *inv_diag, double *partials);
hipError_t launch_cg_scalar_init(hipStream_t, CGState *st, const double *pa, const double *pmu, int nb);
hipError_t launch_cg_scalar(hipStream_t, CGState *st, const double *pa, const double *pmu, int nb);
hipError_t launch_residual(hipStream_t, int N, const double *b, const double *s, double *r, double *partials);
// Hestenes-Stiefel PCG (used for the FP32 inner solve of the mixed path)
template <typename T>
hipError_t launch_pcg_init(hipStream_t, int N, const double *b, const T *inv_diag, T *x, T *r, T *p,
                           double *partials);
hipError_t launch_pcg_scalar_init(hipStream_t, CGState *st, const double *pa, int nb);
hipError_t launch_pcg_alpha(hipStream_t, CGState *st, const double *pmu, int nb);
template <typename T>
hipError_t launch_pcg_update_xr(hipStream_t, int N, const CGState *st, T *x, T *r, const T *p, const T *q,
                                const T *inv_diag, double *partials);
hipError_t launch_pcg_beta(hipStream_t, CGState *st, const double *pa, int nb);
template <typename T>
hipError_t launch_pcg_update_p(hipStream_t, int N, const CGState *st, const T *r, T *p, const T *inv_diag);
// reliable update: r32 = r64 (non-Dirichlet), partial r.z; then the scalar
// step that tests the true residual and restarts beta from it
hipError_t launch_ru_resume(hipStream_t, int N, const double *r64, float *r32, const float *inv_diag,
                            double *partials);
hipError_t launch_ru_scalar(hipStream_t, CGState *st, const double *pa, const double *prr, int nb);
template <typename T>
hipError_t launch_zero(hipStream_t, size_t n, T *a);
template <typename T>
hipError_t launch_add(hipStream_t, size_t NG, double *x, const T *d);
// out = a - b
hipError_t launch_sub(hipStream_t, size_t n, const double *a, const double *b, double *out);
// Robin boundary mass rows (k_solver.hip robin_kernel): y[rows[k]] +=
// sum vals[m] x(cols[m]), x local (smap NULL) or global through the signed
// map; and their diagonal entries added to d
hipError_t launch_robin(hipStream_t, int64_t nrows, const int64_t *rows, const int64_t *rowptr, const int64_t *cols,
                        const double *vals, const int *smap, const double *x, double *y);
hipError_t launch_robin_diag(hipStream_t, int64_t nrows, const int64_t *rows, const int64_t *rowptr,
                             const int64_t *cols, const double *vals, double *d);
hipError_t launch_invert_diag(hipStream_t, int N, double *d);
hipError_t launch_sum_partials(hipStream_t, const double *pa, int nb, double *out);
// state of a mixed solve from dsc[0] = ||b||^2 and dsc[1] = ||r0||^2
hipError_t launch_mixed_state_init(hipStream_t, CGState *st, const double *dsc, double tol2, int maxit, double delta2);
// partial sums of a[i]^2 over i < n (grid of reduction_blocks(n))
hipError_t launch_sumsq(hipStream_t, size_t n, const double *a, double *partials);
hipError_t launch_cast(hipStream_t, size_t n, const double *a, float *b);

// General C0 assembly with the caller's map and signs, Dirichlet DOFs
// first (assembly.hpp); reductions use reduction_blocks(nglob) blocks.
// Indices are signed: m >= 0 is index m with sign +1, m < 0 index ~m with
// sign -1 (menc: local -> global, cenc: the CSR columns, global -> local).
hipError_t launch_gather_signed(hipStream_t, size_t nloc, const int *menc, const double *glob, double *loc);
hipError_t launch_assemble_csr(hipStream_t, int nglob, int ndir, const int *row, const int *cenc, const double *loc,
                               double *glob, const double *wdot, double *partials, bool zero_dir, bool sq);
hipError_t launch_unit_mode(hipStream_t, size_t n, int nc, int m, double *x);
hipError_t launch_take_mode(hipStream_t, size_t nelmt, int nc, int m, const double *y, double *d);
hipError_t launch_invert_diag_dir(hipStream_t, int nglob, int ndir, double *d);
hipError_t launch_cg_init_dirfirst(hipStream_t, int nglob, int ndir, const double *b, const double *inv_diag,
                                   double *x, double *r, double *p, double *q, double *w, double *partials);
hipError_t launch_cg_update_n(hipStream_t, size_t NG, const CGState *st, double *p, double *q, double *x, double *r,
                              double *w, const double *s, const double *inv_diag, double *partials);
// element-local CG (NekLinSysIterCGLoc): preconditioner with the assembled
// dots (reduction_blocks(nglob) blocks; init: {eps, r.w, b.b}, else
// {r.w, eps}), local vector update, local dot over nb blocks, and the
// global values of a continuous local vector
hipError_t launch_cgloc_precon(hipStream_t, bool init, int nglob, int ndir, const int *row, const int *cenc,
                               const double *r, const double *inv_diag, double *wg, double *partials);
hipError_t launch_cgloc_update(hipStream_t, size_t n, const CGState *st, double *p, double *q, double *x, double *r,
                               const double *w, const double *s);
hipError_t launch_dot(hipStream_t, size_t n, int nb, const double *a, const double *b, double *partials);
hipError_t launch_csr_first(hipStream_t, int nglob, const int *row, const int *cenc, const double *loc,
                            double *glob);
// dst[idx[i]] = src[i], i < n
hipError_t launch_scatter_index(hipStream_t, int n, const int *idx, const double *src, double *dst);

}  // namespace solver

struct SolveInfo {
    int iterations = 0;        // total CG iterations (inner iterations for mixed)
    int outer_iterations = 0;  // refinement steps (1 for plain FP64)
    int converged = 0;
    double final_error = 0.0;  // sqrt(||r||^2 / ||b||^2) of the FP64 residual
    double rhs_magnitude = 0.0;
    double solve_ms = 0.0;
};

// Structured n^3 hex mesh on [0,1]^3 with per-vertex coordinates.
struct HexBoxMesh {
    int n = 0;
    std::vector<double> X;  // (n+1)^3 * 3, vertex (ix,iy,iz) at ((iz*(n+1)+iy)*(n+1)+ix)*3
    bool deformed = false;
    static HexBoxMesh make(int n, double perturb, uint64_t seed);
};

class HexSolver {
public:
    HexSolver(const HexBoxMesh &mesh, int nummodes, int numpoints, double lambda, hipStream_t stream);
    ~HexSolver();
    HexSolver(const HexSolver &) = delete;
    HexSolver &operator=(const HexSolver &) = delete;

    // b = Assemble(IProductWRTBase(f)), f = (3 k^2 pi^2 + lambda) prod sin(k pi x)
    void set_forcing_manufactured(int k);
    // b = Assemble(IProductWRTBase(f)) for f at quadrature points (device, FP64)
    void set_forcing_phys(const double *f_dev);
    // mixed: FP32 PCG with FP64 reliable updates (residual, accumulation)
    SolveInfo solve(bool mixed, double tol, int maxit);
    // The same solve as a non-blocking sequence, for serving loops that
    // interleave many solves: begin_solve enqueues the set-up; advance
    // enqueues the next 8 iterations (or a reliable update) once the state
    // of the previous chunk is on the host and returns true when the solve
    // has finished (wait = block for the chunk in flight instead of
    // returning false); finish returns the result.
    void begin_solve(bool mixed, double tol, int maxit);
    bool advance(bool wait);
    SolveInfo finish();
    bool solving() const { return job_.running; }
    // New lambda: rebuilds the diagonal and drops captured iterations; the
    // forcing must be set again afterwards (it depends on lambda).
    void set_lambda(double lambda);
    // L2 error of the current solution against prod sin(k pi x)
    double l2_error_manufactured(int k);
    // ||x|| over all global rows of the current solution
    double solution_norm();
    // y = A x on global vectors (device, FP64), Dirichlet rows zeroed
    void apply_global(const double *x, double *y);
    // Choose the workgroup tile of the gather-apply kernel in FP64 and
    // FP32 by timing each candidate on this mesh; the choice is recorded
    // for later solvers of the same kind.  Captured iteration graphs are
    // rebuilt on the next solve.
    void autotune(int *tile64, int *tile32, std::vector<double> *times64 = nullptr,
                  std::vector<double> *times32 = nullptr);
    int tile(bool fp32) const { return fp32 ? tile32_ : tile64_; }

    size_t nglobal() const { return NG_; }
    size_t nlocal() const { return (size_t)nelmt_ * nc_; }
    size_t ndirichlet() const;
    int nelmt() const { return nelmt_; }
    int nm() const { return nm_; }
    int nq() const { return nq_; }
    bool deformed() const { return deformed_; }
    double lambda() const { return lambda_; }
    hipStream_t stream() const { return stream_; }
    const double *solution() const { return x_.as<double>(); }
    const double *rhs() const { return b_.as<double>(); }
    const double *inv_diag() const { return invd_.as<double>(); }
    size_t device_bytes() const;
    static size_t estimate_bytes(int n, int nm, int nq, bool deformed, bool mixed);


private:
    void build_geometry();
    void build_diagonal();
    void apply_local(bool fp32, const void *xg, void *sloc, const int *done);
    void apply_global_fp64(const double *x, double *y);  // no state check (used inside solves)
    void capture_iterations(bool fp32);
    void assemble_rhs(const double *f_dev);
    void enqueue_state_read();
    void enqueue_reliable_update();
    void begin_solve_enqueue(bool mixed, double tol, int maxit);

    struct Job {
        bool running = false, mixed = false, pending = false, finalizing = false, finished = false;
        double tol = 0.0;
        int max_rounds = 0, rounds = 0, reported_updates = 0;
        SolveInfo info;
    };
    Job job_;
    hipEvent_t ev_state_ = nullptr, ev_t0_ = nullptr, ev_t1_ = nullptr;
    // reliable update once ||r||^2 drops below kReliableDelta^2 * ||r_ref||^2
    static constexpr double kReliableDelta = 0.1;

    int n_, nm_, nq_, N_, nelmt_, nc_;
    size_t NG_;
    bool deformed_;
    double lambda_;
    hipStream_t stream_;
    bool own_stream_;
    int nb_;  // reduction blocks
    int tile64_ = 0, tile32_ = 0;  // gather-apply workgroup tiles (0 = default)

    DeviceBuffer X_, z_, w_, tab64_, tab32_, dtab_, geo64_, geo32_;
    DeviceBuffer b_, x_, invd_, sloc64_, tmp64_, r64_;
    DeviceBuffer invd32_;  // FP32 copy of invd_ for the FP32 iterations of the mixed solve
    // CG work vectors per precision
    DeviceBuffer cg64_[6], cg32_[6], sloc32_;  // x r p q w s
    DeviceBuffer state_, pa_, pmu_, scalar_;
    solver::CGState *host_state_ = nullptr;
    double *host_scalar_ = nullptr;

    hipGraphExec_t graph_[3] = {nullptr, nullptr, nullptr};
    void capture_pcg32();
    void ensure_fp32();
    static constexpr int kItersPerGraph = 8;
};

}  // namespace nektar_amd
