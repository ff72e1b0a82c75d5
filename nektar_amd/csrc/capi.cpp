// C ABI entry points (include/nektar_mf.h).
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "basis.hpp"
#include "batch.hpp"
#include "common.hpp"
#include "assembly.hpp"
#include "mf_operator.hpp"
#include "server.hpp"
#include "solver.hpp"
#include "handles.hpp"

namespace nektar_amd {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string &msg) { g_last_error = msg; }
void clear_last_error() { g_last_error.clear(); }
}  // namespace nektar_amd

using namespace nektar_amd;

static_assert(kMfmaTile == NMF_TILE_MFMA, "matrix-core tile code differs from the header's");

extern "C" {

const char *nmf_last_error(void) { return g_last_error.c_str(); }

int nmf_version(void) { return 100; }

int nmf_device_malloc(void **ptr, size_t bytes)
{
    return guard([&] {
        NMF_REQUIRE(ptr, NMF_ERR_INVALID, "null output pointer");
        *ptr = nullptr;
        if (bytes == 0) return;
        hipError_t e = hipMalloc(ptr, bytes);
        NMF_REQUIRE(e == hipSuccess, NMF_ERR_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    });
}

int nmf_device_free(void *ptr)
{
    return guard([&] {
        if (ptr) NMF_HIP(hipFree(ptr));
    });
}

int nmf_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    return guard([&] {
        if (bytes) NMF_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    });
}

int nmf_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    return guard([&] {
        if (bytes) NMF_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    });
}

int nmf_device_synchronize(void)
{
    return guard([&] { NMF_HIP(hipDeviceSynchronize()); });
}

int nmf_basis_tables(int shape, int dir, int nummodes, int numpoints0, int *nq_dir, int *nmodes_dir,
                     double *bdata, double *dbdata, double *z, double *w, double *D)
{
    return guard([&] {
        NMF_REQUIRE(shape == NMF_SHAPE_HEX || shape == NMF_SHAPE_TET || shape == NMF_SHAPE_PRISM ||
                        shape == NMF_SHAPE_PYR,
                    NMF_ERR_UNSUPPORTED, "shape must be hex, tet, prism or pyramid");
        NMF_REQUIRE(dir >= 0 && dir < 3, NMF_ERR_INVALID, "dir must be 0..2");
        const Shape s = static_cast<Shape>(shape);
        int nq[3];
        default_points(s, numpoints0, nq);
        Basis1D b = make_basis(default_basis(s, dir), nummodes, nq[dir]);
        if (nq_dir) *nq_dir = nq[dir];
        if (nmodes_dir) *nmodes_dir = b.nmodes;
        if (bdata) std::memcpy(bdata, b.bdata.data(), b.bdata.size() * sizeof(double));
        if (dbdata) std::memcpy(dbdata, b.dbdata.data(), b.dbdata.size() * sizeof(double));
        if (z) std::memcpy(z, b.pts.z.data(), b.pts.z.size() * sizeof(double));
        if (w) std::memcpy(w, b.pts.w.data(), b.pts.w.size() * sizeof(double));
        if (D) std::memcpy(D, b.pts.D.data(), b.pts.D.size() * sizeof(double));
    });
}

int nmf_operator_create(const char *key, int nummodes, int numpoints0, int nelmt, int precision, nmf_operator *out)
{
    return guard([&] {
        NMF_REQUIRE(key && out, NMF_ERR_INVALID, "null key or output handle");
        *out = nullptr;
        auto h = std::make_unique<nmf_operator_s>();
        h->impl = std::make_unique<MFOperator>(key, nummodes, numpoints0, nelmt, precision);
        *out = h.release();
    });
}

int nmf_operator_create_basis(const char *key, const int *nummodes, const int *numpoints, int nelmt, int precision,
                              nmf_operator *out)
{
    return guard([&] {
        NMF_REQUIRE(key && out && nummodes && numpoints, NMF_ERR_INVALID, "null key, modes, points or output handle");
        *out = nullptr;
        auto h = std::make_unique<nmf_operator_s>();
        h->impl = std::make_unique<MFOperator>(key, nummodes, numpoints, nelmt, precision);
        *out = h.release();
    });
}

int nmf_operator_runtime_size(nmf_operator op, int *generic)
{
    return guard([&] {
        NMF_REQUIRE(op && op->impl, NMF_ERR_INVALID, "null operator handle");
        NMF_REQUIRE(generic, NMF_ERR_INVALID, "null output");
        *generic = op->impl->generic ? 1 : 0;
    });
}

int nmf_operator_destroy(nmf_operator op)
{
    return guard([&] { delete op; });
}

#define NMF_OP_CHECK() NMF_REQUIRE(op && op->impl, NMF_ERR_INVALID, "null operator handle")

int nmf_operator_info(nmf_operator op, int *ncoeff, int *nqtot, int *nelmt, int *deformed, int *precision)
{
    return guard([&] {
        NMF_OP_CHECK();
        const MFOperator &o = *op->impl;
        if (ncoeff) *ncoeff = o.ncoeff;
        if (nqtot) *nqtot = o.nqtot;
        if (nelmt) *nelmt = o.nelmt;
        if (deformed) *deformed = o.deformed;
        if (precision) *precision = o.fp32 ? NMF_FP32 : NMF_FP64;
    });
}

int nmf_operator_set_jac(nmf_operator op, const double *jac)
{
    return guard([&] {
        NMF_OP_CHECK();
        op->impl->set_jac(jac);
    });
}

int nmf_operator_set_df(nmf_operator op, const double *df)
{
    return guard([&] {
        NMF_OP_CHECK();
        op->impl->set_df(df);
    });
}

int nmf_operator_set_lambda(nmf_operator op, double lambda)
{
    return guard([&] {
        NMF_OP_CHECK();
        NMF_REQUIRE(op->impl->op == OpType::Helmholtz || op->impl->op == OpType::AdvDiffReact, NMF_ERR_INVALID,
                    "SetLambda: not a Helmholtz or LinearAdvectionDiffusionReaction operator");
        op->impl->set_lambda(lambda);
    });
}

int nmf_operator_set_const_var_diffusion(nmf_operator op, const double *d6)
{
    return guard([&] {
        NMF_OP_CHECK();
        NMF_REQUIRE(op->impl->op == OpType::Helmholtz || op->impl->op == OpType::AdvDiffReact, NMF_ERR_INVALID,
                    "SetConstVarDiffusion: not a Helmholtz or LinearAdvectionDiffusionReaction operator");
        op->impl->set_diffusion(d6);
    });
}

int nmf_operator_set_scaling_factor(nmf_operator op, double factor)
{
    return guard([&] {
        NMF_OP_CHECK();
        op->impl->set_scaling_factor(factor);
    });
}

int nmf_operator_io_size(nmf_operator op, int64_t *in_size, int64_t *out_size)
{
    return guard([&] {
        NMF_OP_CHECK();
        const MFOperator &o = *op->impl;
        if (in_size) *in_size = (int64_t)o.nelmt * (int64_t)o.in_per_elmt();
        if (out_size) *out_size = (int64_t)o.nelmt * (int64_t)o.out_per_elmt();
    });
}

int nmf_operator_set_advection(nmf_operator op, const double *vel)
{
    return guard([&] {
        NMF_OP_CHECK();
        op->impl->set_advection(vel);
    });
}

int nmf_operator_apply(nmf_operator op, const void *in, void *out, void *stream)
{
    return guard([&] {
        NMF_OP_CHECK();
        NMF_REQUIRE(op->impl->op != OpType::PhysDeriv, NMF_ERR_INVALID,
                    "PhysDeriv has three outputs: use nmf_operator_apply_physderiv");
        NMF_REQUIRE(op->impl->op != OpType::IProductDeriv, NMF_ERR_INVALID,
                    "IProductWRTDerivBase has three inputs: use nmf_operator_apply_iproduct_deriv");
        op->impl->apply(in, out, nullptr, nullptr, static_cast<hipStream_t>(stream));
    });
}

int nmf_operator_apply_physderiv(nmf_operator op, const void *in, void *out0, void *out1, void *out2, void *stream)
{
    return guard([&] {
        NMF_OP_CHECK();
        NMF_REQUIRE(op->impl->op == OpType::PhysDeriv, NMF_ERR_INVALID, "not a PhysDeriv operator");
        op->impl->apply(in, out0, out1, out2, static_cast<hipStream_t>(stream));
    });
}

int nmf_operator_wsp_size(nmf_operator op, int64_t *wsp_size)
{
    return guard([&] {
        NMF_OP_CHECK();
        NMF_REQUIRE(wsp_size, NMF_ERR_INVALID, "null output");
        *wsp_size = (int64_t)op->impl->wsp_size();
    });
}

int nmf_operator_apply_physderiv_dir(nmf_operator op, int dir, const void *in, void *out, void *wsp, void *stream)
{
    return guard([&] {
        NMF_OP_CHECK();
        MFOperator &o = *op->impl;
        NMF_REQUIRE(o.op == OpType::PhysDeriv, NMF_ERR_INVALID, "not a PhysDeriv operator");
        const hipStream_t s = static_cast<hipStream_t>(stream);
        if (wsp || o.nelmt == 0) {
            o.apply_physderiv_dir(dir, in, out, wsp, s);
            return;
        }
        // no workspace given: a stream-ordered one for this call, as the
        // reference's Collection::ApplyOperator allocates one per call
        void *w = nullptr;
        const hipError_t e = hipMallocAsync(&w, o.wsp_size() * o.value_bytes(), s);
        NMF_REQUIRE(e == hipSuccess, NMF_ERR_MEMORY, std::string("hipMallocAsync: ") + hipGetErrorString(e));
        try {
            o.apply_physderiv_dir(dir, in, out, w, s);
        } catch (...) {
            (void)hipFreeAsync(w, s);
            throw;
        }
        NMF_HIP(hipFreeAsync(w, s));
    });
}

namespace {
size_t in_count(const MFOperator &o) { return (size_t)o.nelmt * o.in_per_elmt(); }
size_t out_count(const MFOperator &o) { return (size_t)o.nelmt * o.out_per_elmt(); }
void to_device(DeviceBuffer &b, const double *h, size_t n, bool fp32)
{
    std::vector<double> v(h, h + n);
    b.upload(v, fp32);
}
void to_host(const DeviceBuffer &b, double *h, size_t n, bool fp32)
{
    if (n == 0) return;
    if (fp32) {
        std::vector<float> f(n);
        NMF_HIP(hipMemcpy(f.data(), b.get(), n * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) h[i] = f[i];
    } else {
        NMF_HIP(hipMemcpy(h, b.get(), n * sizeof(double), hipMemcpyDeviceToHost));
    }
}
}  // namespace

int nmf_operator_apply_host(nmf_operator op, const double *in, double *out)
{
    return guard([&] {
        NMF_OP_CHECK();
        MFOperator &o = *op->impl;
        NMF_REQUIRE(o.op != OpType::PhysDeriv, NMF_ERR_INVALID, "use nmf_operator_apply_physderiv_host");
        NMF_REQUIRE(o.op != OpType::IProductDeriv, NMF_ERR_INVALID, "use nmf_operator_apply_iproduct_deriv_host");
        const size_t ni = in_count(o), no = out_count(o);
        DeviceBuffer di, dout;
        to_device(di, in, ni, o.fp32);
        dout.allocate(no * (o.fp32 ? sizeof(float) : sizeof(double)));
        o.apply(di.get(), dout.get(), nullptr, nullptr, nullptr);
        NMF_HIP(hipDeviceSynchronize());
        to_host(dout, out, no, o.fp32);
    });
}

int nmf_operator_apply_iproduct_deriv(nmf_operator op, const void *in0, const void *in1, const void *in2, void *out,
                                      void *stream)
{
    return guard([&] {
        NMF_OP_CHECK();
        op->impl->apply_iproduct_deriv(in0, in1, in2, out, static_cast<hipStream_t>(stream));
    });
}

int nmf_operator_apply_iproduct_deriv_host(nmf_operator op, const double *in0, const double *in1,
                                           const double *in2, double *out)
{
    return guard([&] {
        NMF_OP_CHECK();
        MFOperator &o = *op->impl;
        NMF_REQUIRE(o.op == OpType::IProductDeriv, NMF_ERR_INVALID, "not an IProductWRTDerivBase operator");
        const size_t n = (size_t)o.nelmt * o.nqtot, no = out_count(o), sz = o.fp32 ? sizeof(float) : sizeof(double);
        DeviceBuffer d0, d1, d2, dout(no * sz);
        to_device(d0, in0, n, o.fp32);
        to_device(d1, in1, n, o.fp32);
        to_device(d2, in2, n, o.fp32);
        o.apply_iproduct_deriv(d0.get(), d1.get(), d2.get(), dout.get(), nullptr);
        NMF_HIP(hipDeviceSynchronize());
        to_host(dout, out, no, o.fp32);
    });
}

int nmf_operator_apply_physderiv_host(nmf_operator op, const double *in, double *out0, double *out1, double *out2)
{
    return guard([&] {
        NMF_OP_CHECK();
        MFOperator &o = *op->impl;
        NMF_REQUIRE(o.op == OpType::PhysDeriv, NMF_ERR_INVALID, "not a PhysDeriv operator");
        const size_t n = (size_t)o.nelmt * o.nqtot, sz = o.fp32 ? sizeof(float) : sizeof(double);
        DeviceBuffer di, d0(n * sz), d1(n * sz), d2(n * sz);
        to_device(di, in, n, o.fp32);
        o.apply(di.get(), d0.get(), d1.get(), d2.get(), nullptr);
        NMF_HIP(hipDeviceSynchronize());
        to_host(d0, out0, n, o.fp32);
        to_host(d1, out1, n, o.fp32);
        to_host(d2, out2, n, o.fp32);
    });
}

#define NMF_SOLVER_CHECK() NMF_REQUIRE(s && s->impl, NMF_ERR_INVALID, "null solver handle")

int nmf_operator_tile_candidates(nmf_operator op, int *tiles, int max, int *count)
{
    return guard([&] {
        NMF_OP_CHECK();
        NMF_REQUIRE(count, NMF_ERR_INVALID, "null count");
        NMF_REQUIRE(max >= 0 && (tiles || max == 0), NMF_ERR_INVALID, "null tiles array");
        const auto c = op->impl->tile_candidates();
        *count = (int)c.size();
        for (int i = 0; i < (int)c.size() && i < max; ++i) tiles[i] = c[i];
    });
}

int nmf_operator_set_tile(nmf_operator op, int tile)
{
    return guard([&] {
        NMF_OP_CHECK();
        op->impl->set_tile(tile);
    });
}

int nmf_operator_get_tile(nmf_operator op, int *tile)
{
    return guard([&] {
        NMF_OP_CHECK();
        NMF_REQUIRE(tile, NMF_ERR_INVALID, "null tile");
        *tile = op->impl->tile;
    });
}

int nmf_operator_autotune(nmf_operator op, void *stream, int *tile, double *times_us, int max)
{
    return guard([&] {
        NMF_OP_CHECK();
        NMF_REQUIRE(!times_us || max >= 0, NMF_ERR_INVALID, "negative max");
        std::vector<double> t;
        const int best = op->impl->autotune(static_cast<hipStream_t>(stream), &t);
        if (tile) *tile = best;
        if (times_us)
            for (int i = 0; i < (int)t.size() && i < max; ++i) times_us[i] = t[i];
    });
}

int nmf_batch_create(nmf_batch *out)
{
    return guard([&] {
        NMF_REQUIRE(out, NMF_ERR_INVALID, "null output handle");
        *out = nullptr;
        auto h = std::make_unique<nmf_batch_s>();
        h->impl = std::make_unique<Batch>();
        *out = h.release();
    });
}

int nmf_batch_destroy(nmf_batch b)
{
    return guard([&] { delete b; });
}

#define NMF_BATCH_CHECK() NMF_REQUIRE(b && b->impl, NMF_ERR_INVALID, "null batch handle")

int nmf_batch_add(nmf_batch b, nmf_operator op, const void *in, void *out)
{
    return guard([&] {
        NMF_BATCH_CHECK();
        NMF_OP_CHECK();
        NMF_REQUIRE(op->impl->op != OpType::PhysDeriv, NMF_ERR_INVALID,
                    "PhysDeriv has three outputs: use nmf_batch_add_physderiv");
        NMF_REQUIRE(op->impl->op != OpType::IProductDeriv, NMF_ERR_UNSUPPORTED,
                    "IProductWRTDerivBase operators are not batched");
        b->impl->add(op->impl.get(), in, out, nullptr, nullptr);
    });
}

int nmf_batch_add_physderiv(nmf_batch b, nmf_operator op, const void *in, void *out0, void *out1, void *out2)
{
    return guard([&] {
        NMF_BATCH_CHECK();
        NMF_OP_CHECK();
        NMF_REQUIRE(op->impl->op == OpType::PhysDeriv, NMF_ERR_INVALID, "not a PhysDeriv operator");
        b->impl->add(op->impl.get(), in, out0, out1, out2);
    });
}

int nmf_batch_set_mode(nmf_batch b, int mode)
{
    return guard([&] {
        NMF_BATCH_CHECK();
        b->impl->set_mode(mode);
    });
}

int nmf_batch_set_streams(nmf_batch b, int streams)
{
    return guard([&] {
        NMF_BATCH_CHECK();
        b->impl->set_streams(streams);
    });
}

int nmf_batch_autotune(nmf_batch b, void *stream)
{
    return guard([&] {
        NMF_BATCH_CHECK();
        b->impl->autotune(static_cast<hipStream_t>(stream));
    });
}

int nmf_batch_run(nmf_batch b, void *stream)
{
    return guard([&] {
        NMF_BATCH_CHECK();
        b->impl->run(static_cast<hipStream_t>(stream));
    });
}

int nmf_batch_info(nmf_batch b, int *items, int *groups, int64_t *dofs)
{
    return guard([&] {
        NMF_BATCH_CHECK();
        if (items) *items = b->impl->items();
        if (groups) *groups = b->impl->groups();
        if (dofs) *dofs = b->impl->dofs();
    });
}

int nmf_solver_create_hex_box(int n, int nummodes, int numpoints0, double perturb, uint64_t seed, double lambda,
                              void *stream, nmf_solver *out)
{
    return guard([&] {
        NMF_REQUIRE(out, NMF_ERR_INVALID, "null output handle");
        *out = nullptr;
        NMF_REQUIRE(perturb >= 0.0 && perturb < 0.5, NMF_ERR_INVALID, "perturb must be in [0, 0.5)");
        auto mesh = HexBoxMesh::make(n, perturb, seed);
        auto h = std::make_unique<nmf_solver_s>();
        h->impl = std::make_unique<HexSolver>(mesh, nummodes, numpoints0, lambda, static_cast<hipStream_t>(stream));
        *out = h.release();
    });
}

int nmf_solver_destroy(nmf_solver s)
{
    return guard([&] { delete s; });
}

int nmf_solver_sizes(nmf_solver s, int64_t *nglobal, int64_t *nlocal, int64_t *ndirichlet, int64_t *device_bytes)
{
    return guard([&] {
        NMF_SOLVER_CHECK();
        if (nglobal) *nglobal = (int64_t)s->impl->nglobal();
        if (nlocal) *nlocal = (int64_t)s->impl->nlocal();
        if (ndirichlet) *ndirichlet = (int64_t)s->impl->ndirichlet();
        if (device_bytes) *device_bytes = (int64_t)s->impl->device_bytes();
    });
}

int64_t nmf_solver_estimate_bytes(int n, int nummodes, int numpoints0, int deformed, int mixed)
{
    if (n < 1 || nummodes < 2 || numpoints0 < nummodes) return -1;
    return (int64_t)HexSolver::estimate_bytes(n, nummodes, numpoints0, deformed != 0, mixed != 0);
}

int nmf_solver_set_forcing_manufactured(nmf_solver s, int k)
{
    return guard([&] {
        NMF_SOLVER_CHECK();
        NMF_REQUIRE(k >= 1, NMF_ERR_INVALID, "frequency k must be >= 1");
        s->impl->set_forcing_manufactured(k);
    });
}

int nmf_solver_set_forcing_phys(nmf_solver s, const double *f_dev)
{
    return guard([&] {
        NMF_SOLVER_CHECK();
        NMF_REQUIRE(f_dev, NMF_ERR_INVALID, "null forcing pointer");
        s->impl->set_forcing_phys(f_dev);
    });
}

int nmf_solver_solve(nmf_solver s, int mixed, double tol, int maxit, nmf_solve_info *info)
{
    return guard([&] {
        NMF_SOLVER_CHECK();
        NMF_REQUIRE(tol > 0.0 && maxit > 0, NMF_ERR_INVALID, "tol and maxit must be positive");
        SolveInfo r = s->impl->solve(mixed != 0, tol, maxit);
        if (info) {
            info->iterations = r.iterations;
            info->outer_iterations = r.outer_iterations;
            info->converged = r.converged;
            info->final_error = r.final_error;
            info->rhs_magnitude = r.rhs_magnitude;
            info->solve_ms = r.solve_ms;
        }
    });
}

int nmf_solver_begin(nmf_solver s, int mixed, double tol, int maxit)
{
    return guard([&] {
        NMF_SOLVER_CHECK();
        NMF_REQUIRE(tol > 0.0 && maxit > 0, NMF_ERR_INVALID, "tol and maxit must be positive");
        s->impl->begin_solve(mixed != 0, tol, maxit);
    });
}

int nmf_solver_advance(nmf_solver s, int wait, int *done)
{
    return guard([&] {
        NMF_SOLVER_CHECK();
        NMF_REQUIRE(done, NMF_ERR_INVALID, "null done flag");
        *done = s->impl->advance(wait != 0) ? 1 : 0;
    });
}

int nmf_solver_finish(nmf_solver s, nmf_solve_info *info)
{
    return guard([&] {
        NMF_SOLVER_CHECK();
        const SolveInfo r = s->impl->finish();
        if (info) {
            info->iterations = r.iterations;
            info->outer_iterations = r.outer_iterations;
            info->converged = r.converged;
            info->final_error = r.final_error;
            info->rhs_magnitude = r.rhs_magnitude;
            info->solve_ms = r.solve_ms;
        }
    });
}

int nmf_solver_l2_error_manufactured(nmf_solver s, int k, double *l2)
{
    return guard([&] {
        NMF_SOLVER_CHECK();
        NMF_REQUIRE(l2, NMF_ERR_INVALID, "null output");
        *l2 = s->impl->l2_error_manufactured(k);
    });
}

namespace {
void copy_global(nmf_solver s, const double *src, double *dst)
{
    NMF_REQUIRE(dst, NMF_ERR_INVALID, "null output");
    NMF_HIP(hipStreamSynchronize(s->impl->stream()));
    NMF_HIP(hipMemcpy(dst, src, s->impl->nglobal() * sizeof(double), hipMemcpyDeviceToHost));
}
}  // namespace

int nmf_solver_get_solution(nmf_solver s, double *x_host)
{
    return guard([&] {
        NMF_SOLVER_CHECK();
        copy_global(s, s->impl->solution(), x_host);
    });
}

int nmf_solver_get_rhs(nmf_solver s, double *b_host)
{
    return guard([&] {
        NMF_SOLVER_CHECK();
        copy_global(s, s->impl->rhs(), b_host);
    });
}

int nmf_solver_get_inv_diag(nmf_solver s, double *d_host)
{
    return guard([&] {
        NMF_SOLVER_CHECK();
        copy_global(s, s->impl->inv_diag(), d_host);
    });
}

int nmf_solver_apply_global(nmf_solver s, const double *x_dev, double *y_dev)
{
    return guard([&] {
        NMF_SOLVER_CHECK();
        NMF_REQUIRE(x_dev && y_dev, NMF_ERR_INVALID, "null vector");
        s->impl->apply_global(x_dev, y_dev);
        NMF_HIP(hipStreamSynchronize(s->impl->stream()));
    });
}

int nmf_solver_autotune(nmf_solver s, int *tile64, int *tile32)
{
    return guard([&] {
        NMF_REQUIRE(s && s->impl, NMF_ERR_INVALID, "null solver handle");
        s->impl->autotune(tile64, tile32);
    });
}

int nmf_clear_tuning(void)
{
    return guard([&] { clear_tuned_tiles(); });
}

int nmf_server_create(int64_t hbm_budget_bytes, int streams, int max_active, nmf_server *out)
{
    return guard([&] {
        NMF_REQUIRE(out, NMF_ERR_INVALID, "null output handle");
        *out = nullptr;
        auto h = std::make_unique<nmf_server_s>();
        h->impl = std::make_unique<Server>(hbm_budget_bytes, streams, max_active);
        *out = h.release();
    });
}

int nmf_server_destroy(nmf_server s)
{
    return guard([&] { delete s; });
}

int nmf_server_run(nmf_server s, const nmf_request *reqs, int count, int validate, nmf_result *results,
                   nmf_server_stats *stats)
{
    return guard([&] {
        NMF_REQUIRE(s && s->impl, NMF_ERR_INVALID, "null server handle");
        NMF_REQUIRE(count >= 0 && (count == 0 || (reqs && results)), NMF_ERR_INVALID, "null request/result array");
        std::vector<ServeRequest> rq(count);
        for (int i = 0; i < count; ++i) {
            const nmf_request &r = reqs[i];
            rq[i] = ServeRequest{r.n, r.nummodes, r.numpoints, r.k, r.mixed, r.maxit, r.lambda, r.tol, r.arrival_s};
        }
        ServeStats st;
        const std::vector<ServeResult> out = s->impl->run(rq, validate != 0, &st);
        for (int i = 0; i < count; ++i) {
            const ServeResult &o = out[i];
            nmf_result &r = results[i];
            r.unknowns = o.unknowns;
            r.iterations = o.iterations;
            r.outer_iterations = o.outer_iterations;
            r.converged = o.converged;
            r.reused_solver = o.reused_solver;
            r.stream = o.stream;
            r.final_error = o.final_error;
            r.l2_error = o.l2_error;
            r.solution_norm = o.solution_norm;
            r.arrival_s = o.arrival_s;
            r.admit_s = o.admit_s;
            r.finish_s = o.finish_s;
            r.latency_ms = o.latency_ms;
            r.queue_ms = o.queue_ms;
            r.device_ms = o.device_ms;
        }
        if (stats) {
            stats->peak_bytes = st.peak_bytes;
            stats->peak_active = st.peak_active;
            stats->solvers_created = st.solvers_created;
            stats->solvers_reused = st.solvers_reused;
            stats->evictions = st.evictions;
            stats->makespan_s = st.makespan_s;
        }
    });
}

int nmf_hex_box_vertices(int n, double perturb, uint64_t seed, double *xyz)
{
    return guard([&] {
        NMF_REQUIRE(xyz, NMF_ERR_INVALID, "null output");
        auto m = HexBoxMesh::make(n, perturb, seed);
        std::memcpy(xyz, m.X.data(), m.X.size() * sizeof(double));
    });
}

}  // extern "C"

/* ---- general C0 assembly and iterative solve ---------------------- */

int nmf_assembly_create(const int *map, const double *sign, int64_t nloc, int nglobal, int ndir, nmf_assembly *out)
{
    return guard([&] {
        NMF_REQUIRE(out, NMF_ERR_INVALID, "null output handle");
        *out = nullptr;
        auto h = std::make_unique<nmf_assembly_s>();
        h->impl = std::make_unique<AssemblyMap>(map, sign, nloc, nglobal, ndir);
        *out = h.release();
    });
}

int nmf_assembly_destroy(nmf_assembly a)
{
    return guard([&] { delete a; });
}

#define NMF_ASM_CHECK() NMF_REQUIRE(a && a->impl, NMF_ERR_INVALID, "null assembly handle")

int nmf_assembly_global_to_local(nmf_assembly a, const double *glob, double *loc, void *stream)
{
    return guard([&] {
        NMF_ASM_CHECK();
        NMF_REQUIRE(a->impl->nloc() == 0 || (glob && loc), NMF_ERR_INVALID, "null vector");
        a->impl->global_to_local(glob, loc, static_cast<hipStream_t>(stream));
    });
}

int nmf_assembly_assemble(nmf_assembly a, const double *loc, double *glob, void *stream)
{
    return guard([&] {
        NMF_ASM_CHECK();
        NMF_REQUIRE(a->impl->nglobal() == 0 || (glob && (loc || a->impl->nloc() == 0)), NMF_ERR_INVALID,
                    "null vector");
        a->impl->assemble(loc, glob, static_cast<hipStream_t>(stream));
    });
}

int nmf_linsys_create(nmf_assembly a, const nmf_operator *ops, int nops, void *stream, nmf_linsys *out)
{
    return nmf_linsys_create_solver(a, ops, nops, "ConjugateGradient", stream, out);
}

int nmf_linsys_create_solver(nmf_assembly a, const nmf_operator *ops, int nops, const char *solver, void *stream,
                             nmf_linsys *out)
{
    return guard([&] {
        NMF_ASM_CHECK();
        NMF_REQUIRE(solver, NMF_ERR_INVALID, "null solver name");
        const std::string name(solver);
        NMF_REQUIRE(name == "ConjugateGradient" || name == "ConjugateGradientLoc" || name == "IterativeStaticCond",
                    NMF_ERR_UNSUPPORTED,
                    "solver '" + name +
                        "': this library has ConjugateGradient, ConjugateGradientLoc and IterativeStaticCond");
        const CGSpace space = name == "ConjugateGradientLoc"  ? CGSpace::Local
                              : name == "IterativeStaticCond" ? CGSpace::StaticCond
                                                              : CGSpace::Global;
        NMF_REQUIRE(out, NMF_ERR_INVALID, "null output handle");
        *out = nullptr;
        NMF_REQUIRE(nops >= 1 && ops, NMF_ERR_INVALID, "at least one operator");
        std::vector<MFOperator *> v;
        for (int i = 0; i < nops; ++i) {
            NMF_REQUIRE(ops[i] && ops[i]->impl, NMF_ERR_INVALID, "null operator handle");
            v.push_back(ops[i]->impl.get());
        }
        auto h = std::make_unique<nmf_linsys_s>();
        h->impl = std::make_unique<IterativeSolver>(*a->impl, v, static_cast<hipStream_t>(stream), space);
        *out = h.release();
    });
}

int nmf_linsys_destroy(nmf_linsys l)
{
    return guard([&] { delete l; });
}

#define NMF_LS_CHECK() NMF_REQUIRE(l && l->impl, NMF_ERR_INVALID, "null linear-system handle")

int nmf_linsys_set_rhs_local(nmf_linsys l, const double *bloc)
{
    return guard([&] {
        NMF_LS_CHECK();
        l->impl->set_rhs_local(bloc);
    });
}

int nmf_linsys_set_dirichlet_local(nmf_linsys l, const double *dloc)
{
    return guard([&] {
        NMF_LS_CHECK();
        l->impl->set_dirichlet_local(dloc);
    });
}

int nmf_linsys_set_robin_local(nmf_linsys l, int64_t nrows, const int64_t *rows, const int64_t *rowptr,
                               const int64_t *cols, const double *vals)
{
    return guard([&] {
        NMF_LS_CHECK();
        l->impl->set_robin_local(nrows, rows, rowptr, cols, vals);
    });
}

int nmf_linsys_solve(nmf_linsys l, double tol, int maxit, nmf_solve_info *info)
{
    return guard([&] {
        NMF_LS_CHECK();
        NMF_REQUIRE(tol > 0.0 && maxit > 0, NMF_ERR_INVALID, "tol and maxit must be positive");
        SolveInfo r = l->impl->solve(tol, maxit);
        if (info) {
            info->iterations = r.iterations;
            info->outer_iterations = r.outer_iterations;
            info->converged = r.converged;
            info->final_error = r.final_error;
            info->rhs_magnitude = r.rhs_magnitude;
            info->solve_ms = r.solve_ms;
        }
    });
}

int nmf_linsys_get_solution(nmf_linsys l, double *x_host)
{
    return guard([&] {
        NMF_LS_CHECK();
        NMF_REQUIRE(x_host, NMF_ERR_INVALID, "null output");
        const double *x = l->impl->solution();
        NMF_HIP(hipDeviceSynchronize());
        NMF_HIP(hipMemcpy(x_host, x, (size_t)l->impl->nglobal() * sizeof(double), hipMemcpyDeviceToHost));
    });
}

int nmf_linsys_get_solution_local(nmf_linsys l, double *xloc_dev)
{
    return guard([&] {
        NMF_LS_CHECK();
        l->impl->solution_local(xloc_dev);
    });
}

int nmf_linsys_sc_diagonal(nmf_linsys l, double *d_host)
{
    return guard([&] {
        NMF_LS_CHECK();
        NMF_REQUIRE(d_host, NMF_ERR_INVALID, "null output");
        const size_t n = (size_t)std::max(l->impl->nglobal(), 1);
        DeviceBuffer d(n * sizeof(double));
        l->impl->sc_diagonal(d.as<double>());
        if (l->impl->nglobal())
            NMF_HIP(hipMemcpy(d_host, d.get(), (size_t)l->impl->nglobal() * sizeof(double), hipMemcpyDeviceToHost));
    });
}

int nmf_linsys_get_inv_diag(nmf_linsys l, double *d_host)
{
    return guard([&] {
        NMF_LS_CHECK();
        NMF_REQUIRE(d_host, NMF_ERR_INVALID, "null output");
        l->impl->refresh();
        NMF_HIP(hipDeviceSynchronize());
        NMF_HIP(hipMemcpy(d_host, l->impl->inv_diag(), (size_t)l->impl->nglobal() * sizeof(double),
                          hipMemcpyDeviceToHost));
    });
}

int nmf_linsys_apply_global(nmf_linsys l, const double *x_dev, double *y_dev)
{
    return guard([&] {
        NMF_LS_CHECK();
        NMF_REQUIRE(x_dev && y_dev, NMF_ERR_INVALID, "null vector");
        l->impl->apply_global(x_dev, y_dev);
    });
}
