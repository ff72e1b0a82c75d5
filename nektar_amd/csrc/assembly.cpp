// General C0 assembly and iterative Helmholtz solve (see assembly.hpp).
#include "assembly.hpp"

#include <algorithm>
#include <string>

#include "schur.hpp"
#include "static_cond.hpp"
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <numeric>

namespace nektar_amd {

namespace {
constexpr int kItersPerChunk = 8;  // CG iterations enqueued between state reads

bool verbose_cg()
{
    const char *v = std::getenv("NMF_VERBOSE");
    return v && *v && *v != '0';
}
}  // namespace

AssemblyMap::AssemblyMap(const int *map, const double *sign, int64_t nloc, int nglobal, int ndir)
    : nloc_(nloc), nglobal_(nglobal), ndir_(ndir), has_sign_(sign != nullptr)
{
    NMF_REQUIRE(nloc >= 0 && nglobal >= 0 && ndir >= 0 && ndir <= nglobal, NMF_ERR_INVALID,
                "assembly map: need 0 <= ndir <= nglobal and nloc >= 0");
    NMF_REQUIRE(nloc < (int64_t)1 << 31, NMF_ERR_UNSUPPORTED, "assembly map: more than 2^31 local entries");
    NMF_REQUIRE(nloc == 0 || map, NMF_ERR_INVALID, "assembly map: null map");
    // CSR of the transpose: the local entries of every global DOF in
    // increasing local index (a stable counting sort)
    std::vector<int> row(nglobal + 1, 0), col(nloc);
    for (int64_t i = 0; i < nloc; ++i) {
        NMF_REQUIRE(map[i] >= 0 && map[i] < nglobal, NMF_ERR_INVALID,
                    "assembly map: entry " + std::to_string(i) + " maps outside [0, nglobal)");
        if (sign)
            NMF_REQUIRE(sign[i] == 1.0 || sign[i] == -1.0, NMF_ERR_INVALID,
                        "assembly map: signs must be +1 or -1 (entry " + std::to_string(i) + ")");
        ++row[map[i] + 1];
    }
    std::partial_sum(row.begin(), row.end(), row.begin());
    // signed indices: the sign of an entry folded into its index (~i for -1)
    std::vector<int> fill(row.begin(), row.end() - 1), menc(std::max<int64_t>(nloc, 1));
    for (int64_t i = 0; i < nloc; ++i) {
        const bool neg = sign && sign[i] < 0.0;
        const int k = fill[map[i]]++;
        col[k] = neg ? ~(int)i : (int)i;
        menc[i] = neg ? ~map[i] : map[i];
    }
    map_.allocate(std::max<int64_t>(nloc, 1) * sizeof(int));
    row_.allocate(row.size() * sizeof(int));
    col_.allocate(std::max<int64_t>(nloc, 1) * sizeof(int));
    if (nloc) {
        NMF_HIP(hipMemcpy(map_.get(), menc.data(), nloc * sizeof(int), hipMemcpyHostToDevice));
        NMF_HIP(hipMemcpy(col_.get(), col.data(), nloc * sizeof(int), hipMemcpyHostToDevice));
    }
    NMF_HIP(hipMemcpy(row_.get(), row.data(), row.size() * sizeof(int), hipMemcpyHostToDevice));
}

void AssemblyMap::global_to_local(const double *glob, double *loc, hipStream_t s) const
{
    if (nloc_ == 0) return;
    NMF_HIP(solver::launch_gather_signed(s, (size_t)nloc_, map_.as<int>(), glob, loc));
}

void AssemblyMap::global_to_local_range(int64_t off, int64_t n, const double *glob, double *loc, hipStream_t s) const
{
    NMF_REQUIRE(off >= 0 && n >= 0 && off + n <= nloc_, NMF_ERR_INVALID, "gather range outside the map");
    if (n == 0) return;
    NMF_HIP(solver::launch_gather_signed(s, (size_t)n, map_.as<int>() + off, glob, loc));
}

void AssemblyMap::assemble(const double *loc, double *glob, hipStream_t s, bool zero_dir, const double *wdot,
                           double *partials, bool sq) const
{
    if (nglobal_ == 0) return;
    NMF_HIP(solver::launch_assemble_csr(s, nglobal_, ndir_, row_.as<int>(), col_.as<int>(), loc, glob, wdot, partials,
                                        zero_dir, sq));
}

void AssemblyMap::precon_local(bool init, const double *r, const double *inv_diag, double *wg, double *partials,
                               hipStream_t s) const
{
    NMF_HIP(solver::launch_cgloc_precon(s, init, nglobal_, ndir_, row_.as<int>(), col_.as<int>(), r, inv_diag, wg,
                                        partials));
}

void AssemblyMap::global_values(const double *loc, double *glob, hipStream_t s) const
{
    if (nglobal_ == 0) return;
    NMF_HIP(solver::launch_csr_first(s, nglobal_, row_.as<int>(), col_.as<int>(), loc, glob));
}

std::vector<int> AssemblyMap::host_map() const
{
    std::vector<int> h((size_t)nloc_);
    if (nloc_) NMF_HIP(hipMemcpy(h.data(), map_.get(), (size_t)nloc_ * sizeof(int), hipMemcpyDeviceToHost));
    for (int &m : h) m = m < 0 ? ~m : m;
    return h;
}

std::vector<double> AssemblyMap::host_sign() const
{
    std::vector<double> h;
    if (!has_sign_ || nloc_ == 0) return h;
    std::vector<int> m((size_t)nloc_);
    NMF_HIP(hipMemcpy(m.data(), map_.get(), (size_t)nloc_ * sizeof(int), hipMemcpyDeviceToHost));
    h.resize((size_t)nloc_);
    for (size_t i = 0; i < m.size(); ++i) h[i] = m[i] < 0 ? -1.0 : 1.0;
    return h;
}

IterativeSolver::IterativeSolver(const AssemblyMap &map, const std::vector<MFOperator *> &ops, hipStream_t stream,
                                 CGSpace space)
    : map_(map), ops_(ops), stream_(stream), space_(space)
{
    int64_t n = 0;
    for (MFOperator *o : ops_) {
        NMF_REQUIRE(o, NMF_ERR_INVALID, "null operator");
        NMF_REQUIRE(o->op == OpType::Helmholtz, NMF_ERR_INVALID,
                    "the iterative solve takes Helmholtz operators (symmetric positive definite)");
        NMF_REQUIRE(!o->fp32, NMF_ERR_INVALID, "the iterative solve takes FP64 operators");
        off_.push_back(n);
        n += (int64_t)o->nelmt * o->ncoeff;
    }
    NMF_REQUIRE(n == map_.nloc(), NMF_ERR_INVALID,
                "operators hold " + std::to_string(n) + " local coefficients, the map " + std::to_string(map_.nloc()));
    const int NG = map_.nglobal();
    nb_ = solver::reduction_blocks((size_t)std::max(NG, 1));
    // Krylov vectors and the right-hand side: global (NG) or local (n)
    const size_t nv = space_ != CGSpace::Global ? (size_t)std::max<int64_t>(n, 1) : (size_t)std::max(NG, 1);
    xloc_.allocate(std::max<int64_t>(n, 1) * sizeof(double));
    yloc_.allocate(std::max<int64_t>(n, 1) * sizeof(double));
    b_.allocate(nv * sizeof(double));
    invd_.allocate(std::max(NG, 1) * sizeof(double));
    xg_.allocate(std::max(NG, 1) * sizeof(double));
    for (auto &v : cg_) v.allocate(nv * sizeof(double));
    state_.allocate(sizeof(solver::CGState));
    pa_.allocate((size_t)nb_ * 3 * sizeof(double));
    pmu_.allocate((size_t)nb_ * sizeof(double));
    NMF_HIP(hipHostMalloc((void **)&host_state_, sizeof(solver::CGState), hipHostMallocDefault));
    NMF_HIP(solver::launch_zero<double>(stream_, nv, b_.as<double>()));
    NMF_HIP(solver::launch_zero<double>(stream_, nv, cg_[0].as<double>()));
    for (MFOperator *o : ops_) o->prepare();
    build_diagonal();
    if (space_ == CGSpace::StaticCond) setup_static_cond();
    for (MFOperator *o : ops_) versions_.push_back(o->version);
}

void IterativeSolver::refresh()
{
    bool changed = false;
    for (size_t k = 0; k < ops_.size(); ++k) changed |= ops_[k]->version != versions_[k];
    if (!changed) return;
    NMF_HIP(hipStreamSynchronize(stream_));
    for (MFOperator *o : ops_) o->prepare();
    build_diagonal();
    if (space_ == CGSpace::StaticCond) setup_static_cond();
    for (size_t k = 0; k < ops_.size(); ++k) versions_[k] = ops_[k]->version;
}

IterativeSolver::~IterativeSolver()
{
    (void)hipStreamSynchronize(stream_);
    if (host_state_) (void)hipHostFree(host_state_);
}

void IterativeSolver::apply_local(const double *xloc, double *yloc)
{
    for (size_t k = 0; k < ops_.size(); ++k)
        ops_[k]->apply(xloc + off_[k], yloc + off_[k], nullptr, nullptr, stream_);
    add_robin(xloc, false, yloc);
}

void IterativeSolver::add_robin(const double *x, bool global, double *yloc)
{
    if (nrobin_ == 0) return;
    NMF_HIP(solver::launch_robin(stream_, nrobin_, rrow_.as<int64_t>(), rptr_.as<int64_t>(), rcol_.as<int64_t>(),
                                 rval_.as<double>(), global ? map_.signed_map() : nullptr, x, yloc));
}

void IterativeSolver::set_robin_local(int64_t nrows, const int64_t *rows, const int64_t *rowptr,
                                      const int64_t *cols, const double *vals)
{
    NMF_REQUIRE(nrows >= 0, NMF_ERR_INVALID, "negative Robin row count");
    NMF_REQUIRE(nrows == 0 || space_ != CGSpace::StaticCond, NMF_ERR_UNSUPPORTED,
                "Robin conditions are not available with static condensation");
    if (nrows == 0) {
        nrobin_ = 0;
        build_diagonal();
        return;
    }
    NMF_REQUIRE(rows && rowptr && cols && vals, NMF_ERR_INVALID, "null Robin array");
    const int64_t n = map_.nloc(), nnz = rowptr[nrows];
    NMF_REQUIRE(rowptr[0] == 0 && nnz >= 0, NMF_ERR_INVALID, "Robin row pointers must start at 0");
    std::vector<char> seen((size_t)n, 0);
    for (int64_t k = 0; k < nrows; ++k) {
        NMF_REQUIRE(rows[k] >= 0 && rows[k] < n, NMF_ERR_INVALID, "Robin row outside the local vector");
        NMF_REQUIRE(!seen[(size_t)rows[k]], NMF_ERR_INVALID, "Robin rows must be distinct");
        seen[(size_t)rows[k]] = 1;
        NMF_REQUIRE(rowptr[k + 1] >= rowptr[k], NMF_ERR_INVALID, "Robin row pointers must not decrease");
        for (int64_t m = rowptr[k]; m < rowptr[k + 1]; ++m)
            NMF_REQUIRE(cols[m] >= 0 && cols[m] < n, NMF_ERR_INVALID, "Robin column outside the local vector");
    }
    NMF_HIP(hipStreamSynchronize(stream_));
    rrow_.allocate((size_t)nrows * sizeof(int64_t));
    rptr_.allocate((size_t)(nrows + 1) * sizeof(int64_t));
    rcol_.allocate((size_t)std::max<int64_t>(nnz, 1) * sizeof(int64_t));
    rval_.allocate((size_t)std::max<int64_t>(nnz, 1) * sizeof(double));
    NMF_HIP(hipMemcpy(rrow_.get(), rows, (size_t)nrows * sizeof(int64_t), hipMemcpyHostToDevice));
    NMF_HIP(hipMemcpy(rptr_.get(), rowptr, (size_t)(nrows + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    if (nnz) {
        NMF_HIP(hipMemcpy(rcol_.get(), cols, (size_t)nnz * sizeof(int64_t), hipMemcpyHostToDevice));
        NMF_HIP(hipMemcpy(rval_.get(), vals, (size_t)nnz * sizeof(double), hipMemcpyHostToDevice));
    }
    nrobin_ = nrows;
    build_diagonal();
}

void IterativeSolver::apply_from_global(const double *xg, double *yloc)
{
    for (size_t k = 0; k < ops_.size(); ++k) {
        MFOperator &o = *ops_[k];
        if (o.can_apply_mapped()) {
            o.apply_mapped(xg, map_.signed_map() + off_[k], yloc + off_[k], stream_);
        } else {
            const int64_t n = (int64_t)o.nelmt * o.ncoeff;
            map_.global_to_local_range(off_[k], n, xg, xloc_.as<double>() + off_[k], stream_);
            o.apply(xloc_.as<double>() + off_[k], yloc + off_[k], nullptr, nullptr, stream_);
        }
    }
    add_robin(xg, true, yloc);
}

// PreconditionerDiagonal.cpp:93-125: the diagonal of every element matrix,
// here by applying the operator to the unit vector of each local mode (all
// elements of a collection at once: ncoeff applies per collection), then
// assembled with sign^2 and inverted off the Dirichlet rows.
void IterativeSolver::build_diagonal()
{
    double *dloc = xloc_.as<double>();  // holds the local diagonal at the end
    DeviceBuffer unit(std::max<int64_t>(map_.nloc(), 1) * sizeof(double));
    for (size_t k = 0; k < ops_.size(); ++k) {
        MFOperator &o = *ops_[k];
        const size_t n = (size_t)o.nelmt * o.ncoeff;
        if (n == 0) continue;
        double *u = unit.as<double>() + off_[k], *y = yloc_.as<double>() + off_[k];
        for (int m = 0; m < o.ncoeff; ++m) {
            NMF_HIP(solver::launch_unit_mode(stream_, n, o.ncoeff, m, u));
            o.apply(u, y, nullptr, nullptr, stream_);
            NMF_HIP(solver::launch_take_mode(stream_, (size_t)o.nelmt, o.ncoeff, m, y, dloc + off_[k]));
        }
    }
    if (nrobin_)  // the Robin mass's diagonal entries
        NMF_HIP(solver::launch_robin_diag(stream_, nrobin_, rrow_.as<int64_t>(), rptr_.as<int64_t>(),
                                          rcol_.as<int64_t>(), rval_.as<double>(), dloc));
    map_.assemble(dloc, invd_.as<double>(), stream_, false, nullptr, nullptr, true);
    NMF_HIP(solver::launch_invert_diag_dir(stream_, map_.nglobal(), map_.ndir(), invd_.as<double>()));
    NMF_HIP(hipStreamSynchronize(stream_));
}

// The element matrices' columns come from operator applies to unit modes
// (every element of a collection at once); per element chunk: A_ii from the
// interior columns, its Cholesky factor, then the boundary columns in
// batches of schur::batch_columns(ni), each giving S_kk = A_kk - |L^-1 a_k|^2 in the
// element-local layout; assembled with the Dirichlet rows zeroed.
void IterativeSolver::sc_diagonal(double *dglob)
{
    NMF_REQUIRE(dglob || map_.nglobal() == 0, NMF_ERR_INVALID, "null output");
    refresh();
    // bytes of element blocks per chunk (NMF_SC_CHUNK_BYTES overrides the 8 GiB default)
    size_t kBudget = (size_t)8 << 30;
    if (const char *v = std::getenv("NMF_SC_CHUNK_BYTES"))
        if (std::strtoull(v, nullptr, 10) > 0) kBudget = std::strtoull(v, nullptr, 10);
    const std::vector<int> hmap = map_.host_map();
    std::vector<int> cnt(std::max(map_.nglobal(), 1), 0);
    for (int g : hmap) ++cnt[g];
    std::vector<int> stamp(std::max(map_.nglobal(), 1), 0);
    int stamp_id = 0;
    DeviceBuffer sloc(std::max<int64_t>(map_.nloc(), 1) * sizeof(double));
    NMF_HIP(solver::launch_zero<double>(stream_, (size_t)std::max<int64_t>(map_.nloc(), 1), sloc.as<double>()));
    for (size_t k = 0; k < ops_.size(); ++k) {
        MFOperator &o = *ops_[k];
        NMF_REQUIRE(o.equal_modes(), NMF_ERR_UNSUPPORTED,
                    "static condensation takes operators with equal modes in every direction");
        const std::vector<int> I = schur::interior_modes(o.shape, o.nm), Bd = schur::boundary_modes(o.shape, o.nm);
        const int E = o.nelmt, nc = o.ncoeff, ni = (int)I.size(), nb = (int)Bd.size();
        const int KB = schur::batch_columns(ni);
        NMF_REQUIRE(ni <= schur::kMaxInterior, NMF_ERR_UNSUPPORTED,
                    "static-condensation diagonal: more than " + std::to_string(schur::kMaxInterior) +
                        " interior modes per element");
        for (int e = 0; e < E; ++e) {
            const int *me = hmap.data() + off_[k] + (size_t)e * nc;
            for (int m : I)
                NMF_REQUIRE(cnt[me[m]] == 1, NMF_ERR_INVALID,
                            "static-condensation diagonal: interior mode " + std::to_string(m) + " of element " +
                                std::to_string(e) + " of collection " + std::to_string(k) +
                                " shares its global DOF");
            // distinct boundary DOFs: stamp each with the element's running number
            ++stamp_id;
            for (int j = 0; j < nb; ++j) {
                int &st = stamp[me[Bd[j]]];
                NMF_REQUIRE(st != stamp_id, NMF_ERR_UNSUPPORTED,
                            "static-condensation diagonal: two boundary modes of element " + std::to_string(e) +
                                " map to one global DOF (periodic single-element meshes)");
                st = stamp_id;
            }
        }
        if (E == 0 || nb == 0) continue;
        DeviceBuffer dI(std::max(ni, 1) * sizeof(int)), dB(nb * sizeof(int));
        if (ni) NMF_HIP(hipMemcpy(dI.get(), I.data(), ni * sizeof(int), hipMemcpyHostToDevice));
        NMF_HIP(hipMemcpy(dB.get(), Bd.data(), nb * sizeof(int), hipMemcpyHostToDevice));
        const size_t per = ((size_t)ni * ni + (size_t)ni * KB + KB) * sizeof(double);
        const int Ec = (int)std::max<size_t>(1, std::min<size_t>((size_t)E, kBudget / per));
        DeviceBuffer Aii(std::max<size_t>((size_t)Ec * ni * ni, 1) * sizeof(double));
        DeviceBuffer Ab(std::max<size_t>((size_t)Ec * ni * KB, 1) * sizeof(double));
        DeviceBuffer dkk((size_t)Ec * KB * sizeof(double)), info((size_t)Ec * sizeof(int));
        std::vector<int> hinfo(Ec);
        const size_t n = (size_t)E * nc;
        double *u = xloc_.as<double>() + off_[k], *y = yloc_.as<double>() + off_[k];
        auto column = [&](int m) {
            NMF_HIP(solver::launch_unit_mode(stream_, n, nc, m, u));
            o.apply(u, y, nullptr, nullptr, stream_);
        };
        for (int e0 = 0; e0 < E; e0 += Ec) {
            const int ne = std::min(Ec, E - e0);
            for (int c = 0; c < ni; ++c) {
                column(I[c]);
                NMF_HIP(schur::launch_take_rows(stream_, ne, e0, nc, ni, dI.as<int>(), 0, y, Aii.as<double>(),
                                                (size_t)ni * ni, c, ni));
            }
            if (ni) {
                NMF_HIP(schur::launch_cholesky(stream_, ne, ni, Aii.as<double>(), info.as<int>()));
                NMF_HIP(hipMemcpyAsync(hinfo.data(), info.get(), ne * sizeof(int), hipMemcpyDeviceToHost, stream_));
                NMF_HIP(hipStreamSynchronize(stream_));
                for (int e = 0; e < ne; ++e)
                    NMF_REQUIRE(hinfo[e] == 0, NMF_ERR_INVALID,
                                "static-condensation diagonal: the interior block of element " +
                                    std::to_string(e0 + e) + " of collection " + std::to_string(k) +
                                    " is not positive definite");
            }
            for (int b0 = 0; b0 < nb; b0 += KB) {
                const int kb = std::min(KB, nb - b0);
                for (int c = 0; c < kb; ++c) {
                    const int m = Bd[b0 + c];
                    column(m);
                    NMF_HIP(schur::launch_take_rows(stream_, ne, e0, nc, ni, dI.as<int>(), 0, y, Ab.as<double>(),
                                                    (size_t)ni * KB, c, ni));
                    NMF_HIP(schur::launch_take_rows(stream_, ne, e0, nc, 1, nullptr, m, y, dkk.as<double>(),
                                                    (size_t)KB, c, 1));
                }
                NMF_HIP(schur::launch_schur_diag(stream_, ne, ni, kb, KB, Aii.as<double>(), Ab.as<double>(),
                                                 dkk.as<double>(), dB.as<int>() + b0, nc, e0,
                                                 sloc.as<double>() + off_[k]));
            }
        }
        NMF_HIP(hipStreamSynchronize(stream_));  // the chunk buffers are freed here
    }
    map_.assemble(sloc.as<double>(), dglob, stream_, true, nullptr, nullptr, true);
    NMF_HIP(hipStreamSynchronize(stream_));
}

void IterativeSolver::set_rhs_local(const double *bloc)
{
    const size_t n = (size_t)map_.nloc();
    NMF_REQUIRE(bloc || n == 0, NMF_ERR_INVALID, "null right-hand side");
    bloc_.allocate(std::max<size_t>(n, 1) * sizeof(double));
    if (n) NMF_HIP(hipMemcpyAsync(bloc_.get(), bloc, n * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    have_rhs_ = true;
}

void IterativeSolver::set_dirichlet_local(const double *dloc)
{
    const size_t n = (size_t)map_.nloc();
    if (!dloc) {
        have_dir_ = false;
        return;
    }
    NMF_REQUIRE(map_.ndir() > 0, NMF_ERR_INVALID, "Dirichlet values given but the map has no Dirichlet DOFs");
    dloc_.allocate(std::max<size_t>(n, 1) * sizeof(double));
    dglob_.allocate(std::max(map_.nglobal(), 1) * sizeof(double));
    if (n) NMF_HIP(hipMemcpyAsync(dloc_.get(), dloc, n * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    have_dir_ = true;
}

// The right-hand side the CG sees: the caller's local one, less A dloc
// (element by element, GeneralMatrixOp) when there are Dirichlet values
// (GlobalLinSysIterativeFull.cpp:118-150); assembled for the global CG.
void IterativeSolver::load_rhs()
{
    const size_t n = (size_t)map_.nloc();
    const double *src = bloc_.as<double>();
    if (have_dir_ && n) {
        apply_local(dloc_.as<double>(), yloc_.as<double>());
        NMF_HIP(solver::launch_sub(stream_, n, bloc_.as<double>(), yloc_.as<double>(), yloc_.as<double>()));
        src = yloc_.as<double>();
    }
    if (space_ != CGSpace::Global) {
        if (n) NMF_HIP(hipMemcpyAsync(b_.get(), src, n * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    } else {
        map_.assemble(src, b_.as<double>(), stream_);
    }
}

const double *IterativeSolver::solution()
{
    const int NG = map_.nglobal();
    const double *x;
    if (space_ == CGSpace::Global) {
        x = cg_[0].as<double>();
    } else if (space_ == CGSpace::StaticCond) {
        x = sc_->xg.as<double>();
    } else {
        map_.global_values(cg_[0].as<double>(), xg_.as<double>(), stream_);
        x = xg_.as<double>();
    }
    if (!have_dir_ || NG == 0) return x;
    // the perturbation plus the Dirichlet values (one local copy per DOF)
    if (x != xg_.as<double>())
        NMF_HIP(hipMemcpyAsync(xg_.get(), x, (size_t)NG * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    map_.global_values(dloc_.as<double>(), dglob_.as<double>(), stream_);
    NMF_HIP(solver::launch_add<double>(stream_, (size_t)NG, xg_.as<double>(), dglob_.as<double>()));
    return xg_.as<double>();
}

void IterativeSolver::solution_local(double *xloc)
{
    const size_t n = (size_t)map_.nloc();
    if (n == 0) return;
    NMF_REQUIRE(xloc, NMF_ERR_INVALID, "null output");
    if (space_ == CGSpace::Global)
        map_.global_to_local(cg_[0].as<double>(), xloc, stream_);
    else
        NMF_HIP(hipMemcpyAsync(xloc, space_ == CGSpace::StaticCond ? sc_->xloc.get() : cg_[0].get(),
                               n * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    // GlobalLinSysIterativeFull.cpp:158-161: perturbation + Dirichlet values
    if (have_dir_) NMF_HIP(solver::launch_add<double>(stream_, n, xloc, dloc_.as<double>()));
}

void IterativeSolver::apply_global(const double *x, double *y)
{
    refresh();
    apply_from_global(x, yloc_.as<double>());
    map_.assemble(yloc_.as<double>(), y, stream_, true);
}

// NekLinSysIterCG::DoConjugateGradient (NekLinSysIterCG.cpp:118-372), the
// single-reduction recurrence of the structured solver on the general
// assembly: per iteration one update of the vectors, s = A w (gather,
// element operators, assembly with the s.w partial sums) and one scalar
// step on the device; the host reads the state every few iterations.
// CGSpace::Local runs NekLinSysIterCGLoc::DoConjugateGradient
// (NekLinSysIterCGLoc.cpp:114-308) instead: the same recurrence and scalar
// steps on element-local vectors, s = A w without assembly, and the
// residual assembled once per iteration by the preconditioner, whose
// assembled values also give r.w and the stopping norm.
SolveInfo IterativeSolver::solve(double tol, int maxit)
{
    NMF_REQUIRE(have_rhs_, NMF_ERR_STATE, "right-hand side not set");
    NMF_REQUIRE(maxit >= 0, NMF_ERR_INVALID, "maxit must be >= 0");
    refresh();
    load_rhs();
    if (space_ == CGSpace::StaticCond) return solve_static_cond(tol, maxit);
    tol = std::max(tol, 1.0e-16);  // NekLinSysIter.cpp:57
    const int NG = map_.nglobal(), ND = map_.ndir();
    const bool local = space_ == CGSpace::Local;
    const size_t NL = (size_t)map_.nloc();
    double *x = cg_[0].as<double>(), *r = cg_[1].as<double>(), *p = cg_[2].as<double>(), *q = cg_[3].as<double>(),
           *w = cg_[4].as<double>(), *s = cg_[5].as<double>();
    double *wg = xg_.as<double>();  // local CG: the assembled, scaled residual
    solver::CGState *st = state_.as<solver::CGState>();
    hipEvent_t e0, e1;
    NMF_HIP(hipEventCreate(&e0));
    NMF_HIP(hipEventCreate(&e1));
    SolveInfo info;
    try {
        solver::CGState init{};
        init.tol2 = tol * tol;
        init.maxit = maxit;
        *host_state_ = init;
        NMF_HIP(hipEventRecord(e0, stream_));
        NMF_HIP(hipMemcpyAsync(st, host_state_, sizeof(solver::CGState), hipMemcpyHostToDevice, stream_));
        if (local) {
            // NekLinSysIterCGLoc.cpp:145-196: r = b, x = 0, w = M^-1 r in
            // local form, s = A w element by element; eps, r.w, b.b from
            // the assembled residual, s.w as a local dot
            if (NL) {
                NMF_HIP(hipMemcpyAsync(r, b_.get(), NL * sizeof(double), hipMemcpyDeviceToDevice, stream_));
                for (double *v : {x, p, q}) NMF_HIP(solver::launch_zero<double>(stream_, NL, v));
            }
            map_.precon_local(true, r, invd_.as<double>(), wg, pa_.as<double>(), stream_);
            map_.global_to_local(wg, w, stream_);
            apply_local(w, s);
            NMF_HIP(solver::launch_dot(stream_, NL, nb_, s, w, pmu_.as<double>()));
        } else {
            NMF_HIP(solver::launch_cg_init_dirfirst(stream_, NG, ND, b_.as<double>(), invd_.as<double>(), x, r, p, q,
                                                    w, pa_.as<double>()));
            apply_from_global(w, yloc_.as<double>());
            map_.assemble(yloc_.as<double>(), s, stream_, true, w, pmu_.as<double>());
        }
        NMF_HIP(solver::launch_cg_scalar_init(stream_, st, pa_.as<double>(), pmu_.as<double>(), nb_));
        const int max_rounds = maxit / kItersPerChunk + 3;
        for (int round = 0;; ++round) {
            NMF_HIP(hipMemcpyAsync(host_state_, st, sizeof(solver::CGState), hipMemcpyDeviceToHost, stream_));
            NMF_HIP(hipStreamSynchronize(stream_));
            if (host_state_->done || round >= max_rounds) break;
            for (int it = 0; it < kItersPerChunk; ++it) {
                if (local) {
                    // NekLinSysIterCGLoc.cpp:210-252
                    NMF_HIP(solver::launch_cgloc_update(stream_, NL, st, p, q, x, r, w, s));
                    map_.precon_local(false, r, invd_.as<double>(), wg, pa_.as<double>(), stream_);
                    map_.global_to_local(wg, w, stream_);
                    apply_local(w, s);
                    NMF_HIP(solver::launch_dot(stream_, NL, nb_, s, w, pmu_.as<double>()));
                } else {
                    NMF_HIP(solver::launch_cg_update_n(stream_, (size_t)NG, st, p, q, x, r, w, s, invd_.as<double>(),
                                                       pa_.as<double>()));
                    apply_from_global(w, yloc_.as<double>());
                    map_.assemble(yloc_.as<double>(), s, stream_, true, w, pmu_.as<double>());
                }
                NMF_HIP(solver::launch_cg_scalar(stream_, st, pa_.as<double>(), pmu_.as<double>(), nb_));
            }
        }
        NMF_HIP(hipEventRecord(e1, stream_));
        NMF_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        NMF_HIP(hipEventElapsedTime(&ms, e0, e1));
        const solver::CGState &h = *host_state_;
        info.iterations = h.iter;
        info.outer_iterations = 1;
        info.converged = h.converged;
        info.final_error = h.final_error;
        info.rhs_magnitude = h.rhs_mag;
        info.solve_ms = ms;
        if (verbose_cg())
            std::fprintf(stderr, "CG iterations made = %d using tolerance of %.1e (error = %.6e, rhs_mag = %.6e)\n",
                         h.iter, tol, h.final_error, std::sqrt(h.rhs_mag));
    } catch (...) {
        (void)hipStreamSynchronize(stream_);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return info;
}

}  // namespace nektar_amd
