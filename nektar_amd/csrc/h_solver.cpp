// Host side of the structured-hex C0 Helmholtz solver.
//
// Reference flow (ContField::v_HelmSolve, ContField.cpp:797-925 ->
// GlobalLinSysIterativeFull::v_Solve, :96-178 -> NekLinSysIterCG):
//   rhs = IProductWRTBase(f) assembled; homogeneous Dirichlet rows held at 0;
//   Jacobi preconditioner from the assembled diagonal; single-reduction PCG
//   stopping at ||r||^2 < tol^2 ||b||^2 (b over all global rows).
// The mixed-precision variant wraps an FP32 CG (FP32 operator, vectors and
// geometry; FP64 dots and diagonal) in FP64 iterative refinement with the
// residual recomputed by the FP64 operator (NekLinSysIterCG::v_DoIterate,
// NekLinSysIterCG.cpp:85-116, is the reference's one-step analogue).
#include "solver.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "basis.hpp"
#include "mf_operator.hpp"

namespace nektar_amd {

using solver::CGState;

namespace {

// NMF_VERBOSE=1 prints CG progress like the reference's --verbose
// (NekLinSysIterCG.cpp:178-190, 323-355).
bool verbose()
{
    static const bool v = [] {
        const char *e = std::getenv("NMF_VERBOSE");
        return e && *e && *e != '0';
    }();
    return v;
}

uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

double uniform_pm1(uint64_t &s) { return 2.0 * (double)(splitmix64(s) >> 11) * 0x1.0p-53 - 1.0; }

}  // namespace

HexBoxMesh HexBoxMesh::make(int n, double perturb, uint64_t seed)
{
    NMF_REQUIRE(n >= 1 && n <= 1024, NMF_ERR_INVALID, "mesh size n must be in [1, 1024]");
    HexBoxMesh m;
    m.n = n;
    m.deformed = perturb != 0.0 && n > 1;
    const int nv = n + 1;
    const double h = 1.0 / n;
    m.X.resize((size_t)nv * nv * nv * 3);
    uint64_t st = seed;
    for (int iz = 0; iz < nv; ++iz)
        for (int iy = 0; iy < nv; ++iy)
            for (int ix = 0; ix < nv; ++ix) {
                double *x = &m.X[((size_t)(iz * nv + iy) * nv + ix) * 3];
                x[0] = ix * h;
                x[1] = iy * h;
                x[2] = iz * h;
                const bool interior = ix > 0 && ix < n && iy > 0 && iy < n && iz > 0 && iz < n;
                if (interior && perturb != 0.0)
                    for (int d = 0; d < 3; ++d) x[d] += perturb * h * uniform_pm1(st);
            }
    return m;
}

size_t HexSolver::estimate_bytes(int n, int nm, int nq, bool deformed, bool mixed)
{
    const size_t N = (size_t)n * (nm - 1) + 1, NG = N * N * N, nel = (size_t)n * n * n;
    const size_t nloc = nel * nm * nm * nm;
    const size_t geo = nel * 7 * (deformed ? (size_t)nq * nq * nq : 1);
    size_t b = 8 * (NG * (5 + 6) + nloc + geo);
    if (mixed) b += 4 * (NG * 7 + nloc + geo);
    b += 8 * nel * (size_t)nq * nq * nq;  // forcing / quadrature scratch
    // the small buffers, so that the estimate bounds device_bytes(): mesh
    // vertices, points and weights, FP64 + FP32 tables, the diagonal's
    // tables, CG state and the partial sums of the reductions
    const size_t tab = (size_t)2 * nq * nq + 2 * (size_t)nm * nq + nq;
    b += 24 * (size_t)(n + 1) * (n + 1) * (n + 1) + 16 * (size_t)nq + 12 * tab + 24 * (size_t)nm * nq;
    b += sizeof(CGState) + 4 * (size_t)solver::kRedBlocks * sizeof(double) + 8 * sizeof(double);
    return b;
}

HexSolver::HexSolver(const HexBoxMesh &mesh, int nummodes, int numpoints, double lambda, hipStream_t stream)
    : n_(mesh.n), nm_(nummodes), nq_(numpoints), deformed_(mesh.deformed), lambda_(lambda)
{
    NMF_REQUIRE(hex_supported(nummodes, numpoints), NMF_ERR_UNSUPPORTED,
                "solver order (nummodes, numpoints) not instantiated");
    NMF_REQUIRE(mesh.X.size() == (size_t)(n_ + 1) * (n_ + 1) * (n_ + 1) * 3, NMF_ERR_INVALID, "bad mesh vertices");
    N_ = n_ * (nm_ - 1) + 1;
    NG_ = (size_t)N_ * N_ * N_;
    nelmt_ = n_ * n_ * n_;
    nc_ = nm_ * nm_ * nm_;
    own_stream_ = stream == nullptr;
    if (own_stream_) NMF_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    else stream_ = stream;
    nb_ = solver::reduction_blocks(NG_);

    X_.upload(mesh.X, false);
    Points1D P = make_points(PointsType::GLL, nq_);
    z_.upload(P.z, false);
    w_.upload(P.w, false);
    const std::vector<double> tab = MFOperator::build_tables(Shape::Hex, nm_, nq_);
    tab64_.upload(tab, false);
    tab32_.upload(tab, true);
    Basis1D B = make_basis(BasisType::ModifiedA, nm_, nq_);
    std::vector<double> dt(3 * (size_t)nm_ * nq_);
    for (size_t i = 0; i < (size_t)nm_ * nq_; ++i) {
        dt[i] = B.bdata[i] * B.bdata[i];
        dt[(size_t)nm_ * nq_ + i] = B.bdata[i] * B.dbdata[i];
        dt[2 * (size_t)nm_ * nq_ + i] = B.dbdata[i] * B.dbdata[i];
    }
    dtab_.upload(dt, false);

    const size_t nloc = (size_t)nelmt_ * nc_;
    for (auto *buf : {&b_, &x_, &invd_, &tmp64_, &r64_}) buf->allocate(NG_ * sizeof(double));
    sloc64_.allocate(nloc * sizeof(double));
    for (auto &v : cg64_) v.allocate(NG_ * sizeof(double));
    state_.allocate(sizeof(CGState));
    // partial-sum buffers sized for the largest reduction grid any kernel uses
    pa_.allocate(3 * (size_t)solver::kRedBlocks * sizeof(double));
    pmu_.allocate((size_t)solver::kRedBlocks * sizeof(double));
    scalar_.allocate(8 * sizeof(double));
    NMF_HIP(hipHostMalloc((void **)&host_state_, sizeof(CGState), hipHostMallocDefault));
    NMF_HIP(hipHostMalloc((void **)&host_scalar_, 8 * sizeof(double), hipHostMallocDefault));
    NMF_HIP(hipMemsetAsync(b_.get(), 0, NG_ * sizeof(double), stream_));
    NMF_HIP(hipMemsetAsync(x_.get(), 0, NG_ * sizeof(double), stream_));

    NMF_HIP(hipEventCreateWithFlags(&ev_state_, hipEventDisableTiming));
    NMF_HIP(hipEventCreate(&ev_t0_));
    NMF_HIP(hipEventCreate(&ev_t1_));
    build_geometry();
    build_diagonal();
    tile64_ = tuned_tile(TuneKey{(int)Shape::Hex, (int)OpType::Helmholtz, nm_, nq_, 0, deformed_ ? 1 : 0, 1});
    tile32_ = tuned_tile(TuneKey{(int)Shape::Hex, (int)OpType::Helmholtz, nm_, nq_, 1, deformed_ ? 1 : 0, 1});
    NMF_HIP(hipStreamSynchronize(stream_));
}

HexSolver::~HexSolver()
{
    for (auto &g : graph_)
        if (g) (void)hipGraphExecDestroy(g);
    if (own_stream_ || job_.running) (void)hipStreamSynchronize(stream_);
    for (auto ev : {ev_state_, ev_t0_, ev_t1_})
        if (ev) (void)hipEventDestroy(ev);
    if (host_state_) (void)hipHostFree(host_state_);
    if (host_scalar_) (void)hipHostFree(host_scalar_);
    if (own_stream_ && stream_) {
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
    }
}

size_t HexSolver::ndirichlet() const
{
    const size_t Ni = (size_t)N_ - 2;
    return NG_ - Ni * Ni * Ni;
}

size_t HexSolver::device_bytes() const
{
    size_t s = 0;
    for (const DeviceBuffer *b : {&X_, &z_, &w_, &tab64_, &tab32_, &dtab_, &geo64_, &geo32_, &b_, &x_, &invd_,
                                  &invd32_, &sloc64_, &tmp64_, &r64_, &sloc32_, &state_, &pa_, &pmu_, &scalar_})
        s += b->bytes();
    for (const auto &v : cg64_) s += v.bytes();
    for (const auto &v : cg32_) s += v.bytes();
    return s;
}

void HexSolver::build_geometry()
{
    const size_t g = (size_t)nelmt_ * 7 * (deformed_ ? (size_t)nq_ * nq_ * nq_ : 1);
    geo64_.allocate(g * sizeof(double));
    NMF_HIP(solver::launch_geom(stream_, n_, nq_, X_.as<double>(), z_.as<double>(), w_.as<double>(), deformed_,
                                geo64_.as<double>()));
}

void HexSolver::build_diagonal()
{
    SolverArgs a{SolverKernel::Diag, false, deformed_, n_, N_, tab64_.get(), dtab_.get(), geo64_.get(), lambda_,
                 nullptr, sloc64_.get(), nullptr, stream_};
    NMF_HIP(launch_hex_solver(nm_, nq_, a));
    NMF_HIP(solver::launch_assemble<double>(stream_, n_, nm_, N_, sloc64_.as<double>(), invd_.as<double>(), nullptr,
                                            nullptr, nullptr));
    NMF_HIP(solver::launch_invert_diag(stream_, N_, invd_.as<double>()));
}

void HexSolver::set_forcing_phys(const double *f_dev)
{
    NMF_REQUIRE(!job_.running, NMF_ERR_STATE, "set_forcing during a solve");
    assemble_rhs(f_dev);
    NMF_HIP(hipStreamSynchronize(stream_));  // f_dev is the caller's
}

void HexSolver::set_forcing_manufactured(int k)
{
    NMF_REQUIRE(!job_.running, NMF_ERR_STATE, "set_forcing during a solve");
    // stream-ordered scratch: no device-wide synchronisation, so solves on
    // other streams keep running while a new request is set up
    double *f = nullptr;
    const size_t bytes = (size_t)nelmt_ * nq_ * nq_ * nq_ * sizeof(double);
    hipError_t e = hipMallocAsync((void **)&f, bytes, stream_);
    NMF_REQUIRE(e == hipSuccess, NMF_ERR_MEMORY, std::string("hipMallocAsync: ") + hipGetErrorString(e));
    NMF_HIP(solver::launch_forcing(stream_, n_, nq_, X_.as<double>(), z_.as<double>(), k, lambda_, f));
    assemble_rhs(f);
    NMF_HIP(hipFreeAsync(f, stream_));
}

void HexSolver::assemble_rhs(const double *f_dev)
{
    SolverArgs a{SolverKernel::Rhs, false, deformed_, n_, N_, tab64_.get(), nullptr, geo64_.get(), lambda_, f_dev,
                 sloc64_.get(), nullptr, stream_};
    NMF_HIP(launch_hex_solver(nm_, nq_, a));
    // Full assembly (Dirichlet rows kept: they enter ||b||^2 as in the
    // reference's Set_Rhs_Magnitude over all global rows).
    NMF_HIP(solver::launch_assemble<double>(stream_, n_, nm_, N_, sloc64_.as<double>(), b_.as<double>(), nullptr,
                                            nullptr, nullptr, true));
}

void HexSolver::apply_local(bool fp32, const void *xg, void *sloc, const int *done)
{
    SolverArgs a{SolverKernel::Apply, fp32, deformed_, n_, N_, fp32 ? tab32_.get() : tab64_.get(), nullptr,
                 fp32 ? geo32_.get() : geo64_.get(), lambda_, xg, sloc, done, stream_, fp32 ? tile32_ : tile64_};
    NMF_HIP(launch_hex_solver(nm_, nq_, a));
}

void HexSolver::autotune(int *tile64, int *tile32, std::vector<double> *times64, std::vector<double> *times32)
{
    NMF_REQUIRE(!job_.running, NMF_ERR_STATE, "autotune during a solve");
    ensure_fp32();
    for (int prec = 0; prec < 2; ++prec) {
        const bool fp32 = prec == 1;
        int buf[8];
        const int nc = hex_tile_candidates(nm_, nq_, fp32, buf, 8);
        std::vector<int> cand(buf, buf + nc);
        // a scratch global vector of non-zero normal values (every byte
        // 0x3F) serves as input, so the timed arithmetic draws real power
        void *xg = fp32 ? cg32_[2].get() : tmp64_.get();
        NMF_HIP(hipMemsetAsync(xg, 0x3F, NG_ * (fp32 ? sizeof(float) : sizeof(double)), stream_));
        void *sloc = fp32 ? sloc32_.get() : sloc64_.get();
        SolverArgs a{SolverKernel::Apply, fp32, deformed_, n_, N_, fp32 ? tab32_.get() : tab64_.get(), nullptr,
                     fp32 ? geo32_.get() : geo64_.get(), lambda_, xg, sloc, nullptr, stream_, 0};
        const int best = time_tiles(
            stream_, cand,
            [&](int t) {
                a.tile = t;
                return launch_hex_solver(nm_, nq_, a);
            },
            10, fp32 ? times32 : times64);
        (fp32 ? tile32_ : tile64_) = best;
        record_tile(TuneKey{(int)Shape::Hex, (int)OpType::Helmholtz, nm_, nq_, prec, deformed_ ? 1 : 0, 1}, best);
    }
    for (auto &g : graph_)
        if (g) {
            NMF_HIP(hipGraphExecDestroy(g));
            g = nullptr;
        }
    if (tile64) *tile64 = tile64_;
    if (tile32) *tile32 = tile32_;
}

void HexSolver::apply_global(const double *x, double *y)
{
    NMF_REQUIRE(!job_.running, NMF_ERR_STATE, "apply_global during a solve (it shares the solve's work buffers)");
    apply_global_fp64(x, y);
}

void HexSolver::apply_global_fp64(const double *x, double *y)
{
    apply_local(false, x, sloc64_.get(), nullptr);
    NMF_HIP(solver::launch_assemble<double>(stream_, n_, nm_, N_, sloc64_.as<double>(), y, nullptr, nullptr, nullptr));
}

std::mutex &graph_capture_mutex()
{
    static std::mutex m;
    return m;
}

void HexSolver::capture_iterations(bool fp32)
{
    const int slot = fp32 ? 1 : 0;
    if (graph_[slot]) return;
    std::lock_guard<std::mutex> lock(graph_capture_mutex());
    CGState *st = state_.as<CGState>();
    const int *done = &st->done;
    hipGraph_t g;
    NMF_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    for (int it = 0; it < kItersPerGraph; ++it) {
        if (fp32) {
            float *x = cg32_[0].as<float>(), *r = cg32_[1].as<float>(), *p = cg32_[2].as<float>(),
                  *q = cg32_[3].as<float>(), *w = cg32_[4].as<float>(), *s = cg32_[5].as<float>();
            NMF_HIP(solver::launch_cg_update<float>(stream_, N_, st, p, q, x, r, w, s, invd_.as<double>(),
                                                    pa_.as<double>()));
            apply_local(true, w, sloc32_.get(), done);
            NMF_HIP(solver::launch_assemble<float>(stream_, n_, nm_, N_, sloc32_.as<float>(), s, w, pmu_.as<double>(),
                                                   done));
        } else {
            double *x = cg64_[0].as<double>(), *r = cg64_[1].as<double>(), *p = cg64_[2].as<double>(),
                   *q = cg64_[3].as<double>(), *w = cg64_[4].as<double>(), *s = cg64_[5].as<double>();
            NMF_HIP(solver::launch_cg_update<double>(stream_, N_, st, p, q, x, r, w, s, invd_.as<double>(),
                                                     pa_.as<double>()));
            apply_local(false, w, sloc64_.get(), done);
            NMF_HIP(solver::launch_assemble<double>(stream_, n_, nm_, N_, sloc64_.as<double>(), s, w,
                                                    pmu_.as<double>(), done));
        }
        NMF_HIP(solver::launch_cg_scalar(stream_, st, pa_.as<double>(), pmu_.as<double>(), nb_));
    }
    NMF_HIP(hipStreamEndCapture(stream_, &g));
    hipError_t e = hipGraphInstantiate(&graph_[slot], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    NMF_REQUIRE(e == hipSuccess, NMF_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
}

void HexSolver::ensure_fp32()
{
    if (!geo32_.get()) {
        const size_t g = geo64_.bytes() / sizeof(double);
        geo32_.allocate(g * sizeof(float));
        NMF_HIP(solver::launch_cast(stream_, g, geo64_.as<double>(), geo32_.as<float>()));
    }
    if (!cg32_[0].get()) {
        for (auto &v : cg32_) v.allocate(NG_ * sizeof(float));
        sloc32_.allocate((size_t)nelmt_ * nc_ * sizeof(float));
        invd32_.allocate(NG_ * sizeof(float));
    }
}

void HexSolver::capture_pcg32()
{
    if (graph_[2]) return;
    std::lock_guard<std::mutex> lock(graph_capture_mutex());
    CGState *st = state_.as<CGState>();
    const int *done = &st->done;
    float *x = cg32_[0].as<float>(), *r = cg32_[1].as<float>(), *p = cg32_[2].as<float>(), *q = cg32_[3].as<float>();
    hipGraph_t g;
    NMF_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    for (int it = 0; it < kItersPerGraph; ++it) {
        apply_local(true, p, sloc32_.get(), done);
        NMF_HIP(solver::launch_assemble<float>(stream_, n_, nm_, N_, sloc32_.as<float>(), q, p, pmu_.as<double>(),
                                               done));
        NMF_HIP(solver::launch_pcg_alpha(stream_, st, pmu_.as<double>(), nb_));
        NMF_HIP(solver::launch_pcg_update_xr<float>(stream_, N_, st, x, r, p, q, invd32_.as<float>(),
                                                    pa_.as<double>()));
        NMF_HIP(solver::launch_pcg_beta(stream_, st, pa_.as<double>(), nb_));
        NMF_HIP(solver::launch_pcg_update_p<float>(stream_, N_, st, r, p, invd32_.as<float>()));
    }
    NMF_HIP(hipStreamEndCapture(stream_, &g));
    hipError_t e = hipGraphInstantiate(&graph_[2], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    NMF_REQUIRE(e == hipSuccess, NMF_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
}

void HexSolver::enqueue_state_read()
{
    NMF_HIP(hipMemcpyAsync(host_state_, state_.get(), sizeof(CGState), hipMemcpyDeviceToHost, stream_));
    NMF_HIP(hipEventRecord(ev_state_, stream_));
    job_.pending = true;
}

void HexSolver::begin_solve(bool mixed, double tol, int maxit)
{
    NMF_REQUIRE(!job_.running, NMF_ERR_STATE, "a solve is already in progress on this solver");
    NMF_REQUIRE(maxit >= 0, NMF_ERR_INVALID, "maxit must be >= 0");
    job_ = Job{};
    job_.running = true;
    job_.mixed = mixed;
    tol = std::max(tol, 1.0e-16);  // NekLinSysIter.cpp:57
    job_.tol = tol;
    try {
        begin_solve_enqueue(mixed, tol, maxit);
    } catch (...) {
        // a failed set-up (graph instantiation, FP32 buffers, a launch)
        // leaves the solver usable: drain what was enqueued, forget the job
        (void)hipStreamSynchronize(stream_);
        job_ = Job{};
        throw;
    }
}

void HexSolver::begin_solve_enqueue(bool mixed, double tol, int maxit)
{
    NMF_HIP(hipEventRecord(ev_t0_, stream_));
    CGState *st = state_.as<CGState>();
    if (!mixed) {
        CGState init{};
        init.tol2 = tol * tol;
        init.maxit = maxit;
        *host_state_ = init;
        NMF_HIP(hipMemcpyAsync(state_.get(), host_state_, sizeof(CGState), hipMemcpyHostToDevice, stream_));
        double *x = cg64_[0].as<double>(), *r = cg64_[1].as<double>(), *p = cg64_[2].as<double>(),
               *q = cg64_[3].as<double>(), *w = cg64_[4].as<double>(), *s = cg64_[5].as<double>();
        NMF_HIP(solver::launch_cg_init<double>(stream_, N_, b_.as<double>(), invd_.as<double>(), x, r, p, q, w,
                                               pa_.as<double>()));
        apply_local(false, w, sloc64_.get(), nullptr);
        NMF_HIP(solver::launch_assemble<double>(stream_, n_, nm_, N_, sloc64_.as<double>(), s, w, pmu_.as<double>(),
                                                nullptr));
        NMF_HIP(solver::launch_cg_scalar_init(stream_, st, pa_.as<double>(), pmu_.as<double>(), nb_));
        capture_iterations(false);
        job_.max_rounds = maxit / kItersPerGraph + 3;
    } else {
        ensure_fp32();
        float *y = cg32_[0].as<float>(), *r = cg32_[1].as<float>(), *p = cg32_[2].as<float>();
        double *dsc = scalar_.as<double>();
        // x64 = 0, r64 = b on non-Dirichlet rows; ||b||^2 over all global
        // rows (Set_Rhs_Magnitude, NekLinSysIter.cpp:96-117) and ||r64||^2
        // go to the device state without a host round trip.
        NMF_HIP(solver::launch_zero<double>(stream_, NG_, x_.as<double>()));
        NMF_HIP(solver::launch_zero<double>(stream_, NG_, tmp64_.as<double>()));
        NMF_HIP(solver::launch_sumsq(stream_, NG_, b_.as<double>(), pmu_.as<double>()));
        NMF_HIP(solver::launch_sum_partials(stream_, pmu_.as<double>(), nb_, dsc));
        NMF_HIP(solver::launch_residual(stream_, N_, b_.as<double>(), tmp64_.as<double>(), r64_.as<double>(),
                                        pmu_.as<double>()));
        NMF_HIP(solver::launch_sum_partials(stream_, pmu_.as<double>(), nb_, dsc + 1));
        NMF_HIP(solver::launch_mixed_state_init(stream_, st, dsc, tol * tol, maxit, kReliableDelta * kReliableDelta));
        // the FP32 iterations precondition with an FP32 copy of the diagonal
        // (4 bytes less per row in each of their two vector updates), taken
        // here so that it follows every rebuild of the FP64 one
        NMF_HIP(solver::launch_cast(stream_, NG_, invd_.as<double>(), invd32_.as<float>()));
        NMF_HIP(solver::launch_pcg_init<float>(stream_, N_, r64_.as<double>(), invd32_.as<float>(), y, r, p,
                                               pa_.as<double>()));
        NMF_HIP(solver::launch_pcg_scalar_init(stream_, st, pa_.as<double>(), nb_));
        capture_pcg32();
        job_.max_rounds = maxit / kItersPerGraph + 64 + maxit;
    }
    enqueue_state_read();
}

void HexSolver::enqueue_reliable_update()
{
    // accumulate the FP32 correction in FP64, recompute the true residual
    // with the FP64 operator, restart the recurrence residual from it
    CGState *st = state_.as<CGState>();
    float *y = cg32_[0].as<float>(), *r = cg32_[1].as<float>(), *p = cg32_[2].as<float>();
    NMF_HIP(solver::launch_add<float>(stream_, NG_, x_.as<double>(), y));
    NMF_HIP(solver::launch_zero<float>(stream_, NG_, y));
    apply_global_fp64(x_.as<double>(), tmp64_.as<double>());
    NMF_HIP(solver::launch_residual(stream_, N_, b_.as<double>(), tmp64_.as<double>(), r64_.as<double>(),
                                    pmu_.as<double>()));
    NMF_HIP(solver::launch_ru_resume(stream_, N_, r64_.as<double>(), r, invd32_.as<float>(), pa_.as<double>()));
    NMF_HIP(solver::launch_ru_scalar(stream_, st, pa_.as<double>(), pmu_.as<double>(), nb_));
    NMF_HIP(solver::launch_pcg_update_p<float>(stream_, N_, st, r, p, invd32_.as<float>()));
}

bool HexSolver::advance(bool wait)
{
    NMF_REQUIRE(job_.running, NMF_ERR_STATE, "no solve in progress on this solver");
    if (job_.finished) return true;
    if (job_.finalizing) {
        if (wait) {
            NMF_HIP(hipEventSynchronize(ev_t1_));
        } else {
            const hipError_t q = hipEventQuery(ev_t1_);
            if (q == hipErrorNotReady) return false;
            NMF_HIP(q);
        }
        float ms = 0.f;
        NMF_HIP(hipEventElapsedTime(&ms, ev_t0_, ev_t1_));
        job_.info.solve_ms = ms;
        job_.finished = true;
        return true;
    }
    if (job_.pending) {
        if (wait) {
            NMF_HIP(hipEventSynchronize(ev_state_));
        } else {
            const hipError_t q = hipEventQuery(ev_state_);
            if (q == hipErrorNotReady) return false;
            NMF_HIP(q);
        }
        job_.pending = false;
    }
    const CGState st = *host_state_;
    if (job_.mixed && verbose() && st.updates > job_.reported_updates) {
        job_.reported_updates = st.updates;
        std::fprintf(stderr, "mixed CG reliable update %d at iteration %d: FP64 error = %.6e\n", st.updates, st.iter,
                     st.final_error);
    }
    if (st.done || job_.rounds >= job_.max_rounds) {
        SolveInfo &info = job_.info;
        info.iterations = st.iter;
        info.converged = st.converged;
        info.final_error = st.final_error;
        if (job_.mixed) {
            info.outer_iterations = st.updates;
            info.rhs_magnitude = st.rhs_mag_override;
        } else {
            info.outer_iterations = 1;
            info.rhs_magnitude = st.rhs_mag;
            NMF_HIP(hipMemcpyAsync(x_.get(), cg64_[0].get(), NG_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
            if (verbose())
                std::fprintf(stderr, "CG iterations made = %d using tolerance of %.1e (error = %.6e, rhs_mag = %.6e)\n",
                             st.iter, job_.tol, st.final_error, std::sqrt(st.rhs_mag));
        }
        NMF_HIP(hipEventRecord(ev_t1_, stream_));
        job_.finalizing = true;
        return advance(wait);
    }
    if (job_.mixed && st.pause)
        enqueue_reliable_update();
    else
        NMF_HIP(hipGraphLaunch(graph_[job_.mixed ? 2 : 0], stream_));
    ++job_.rounds;
    enqueue_state_read();
    return false;
}

SolveInfo HexSolver::finish()
{
    NMF_REQUIRE(job_.running && job_.finished, NMF_ERR_STATE, "solve not finished");
    job_.running = false;
    return job_.info;
}

SolveInfo HexSolver::solve(bool mixed, double tol, int maxit)
{
    begin_solve(mixed, tol, maxit);
    try {
        while (!advance(true)) {
        }
    } catch (...) {
        (void)hipStreamSynchronize(stream_);
        job_ = Job{};
        throw;
    }
    return finish();
}

void HexSolver::set_lambda(double lambda)
{
    NMF_REQUIRE(!job_.running, NMF_ERR_STATE, "set_lambda during a solve");
    lambda_ = lambda;
    build_diagonal();
    // the captured iterations carry lambda as a kernel argument
    for (auto &g : graph_)
        if (g) {
            NMF_HIP(hipGraphExecDestroy(g));
            g = nullptr;
        }
}

double HexSolver::solution_norm()
{
    NMF_REQUIRE(!job_.running, NMF_ERR_STATE, "solution_norm during a solve");
    NMF_HIP(solver::launch_sumsq(stream_, NG_, x_.as<double>(), pmu_.as<double>()));
    NMF_HIP(solver::launch_sum_partials(stream_, pmu_.as<double>(), nb_, scalar_.as<double>()));
    NMF_HIP(hipMemcpyAsync(host_scalar_, scalar_.get(), sizeof(double), hipMemcpyDeviceToHost, stream_));
    NMF_HIP(hipStreamSynchronize(stream_));
    return std::sqrt(host_scalar_[0]);
}

double HexSolver::l2_error_manufactured(int k)
{
    NMF_REQUIRE(!job_.running, NMF_ERR_STATE, "l2_error_manufactured during a solve");
    const size_t nq3 = (size_t)nq_ * nq_ * nq_;
    DeviceBuffer u((size_t)nelmt_ * nq3 * sizeof(double));
    SolverArgs a{SolverKernel::BwdGather, false, deformed_, n_, N_, tab64_.get(), nullptr, nullptr, lambda_,
                 x_.get(), u.get(), nullptr, stream_};
    NMF_HIP(launch_hex_solver(nm_, nq_, a));
    int nb = 0;
    NMF_HIP(solver::launch_l2err(stream_, n_, nq_, X_.as<double>(), z_.as<double>(), w_.as<double>(),
                                 geo64_.as<double>(), deformed_, k, u.as<double>(), pmu_.as<double>(), &nb));
    NMF_HIP(solver::launch_sum_partials(stream_, pmu_.as<double>(), nb, scalar_.as<double>()));
    NMF_HIP(hipMemcpyAsync(host_scalar_, scalar_.get(), sizeof(double), hipMemcpyDeviceToHost, stream_));
    NMF_HIP(hipStreamSynchronize(stream_));
    return std::sqrt(host_scalar_[0]);
}

}  // namespace nektar_amd
