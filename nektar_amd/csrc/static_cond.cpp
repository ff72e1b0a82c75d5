// Iterative static condensation over the general assembly: the GPU
// counterpart of GlobalLinSysIterativeStaticCond with NekLinSysIterCG and the
// Diagonal preconditioner (GlobalLinSysStaticCond.cpp:98-230,
// GlobalLinSysIterativeStaticCond.cpp:378-492,
// PreconditionerDiagonal.cpp:132-154), single level.
//
// Setup, per collection: the interior blocks A_ii from the columns of the
// element matrices (operator applies to unit modes), their Cholesky
// factors, the explicit inverses A_ii^-1 (m_invD, kept) and the Schur
// diagonal for the preconditioner.  The reference applies the Schur
// complement S = A_bb - A_bi A_ii^-1 A_ib from dense blocks
// (m_denseBlocks); here S x = (A [x; 0])_b - (A [0; A_ii^-1 (A [x; 0])_i])_b
// with two matrix-free applies and one block product with A_ii^-1, which
// moves ni^2 instead of nb^2 doubles per element through HBM (hex P=6:
// 125^2 against 218^2).  Solve: condensation of the right-hand side, CG on
// the assembled boundary system, back-substitution of the interior.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "assembly.hpp"
#include "schur.hpp"
#include "static_cond.hpp"

namespace nektar_amd {

namespace {
constexpr int kItersPerChunk = 8;
constexpr double kRhsMagSmoothing = 0.9;  // m_rhs_mag_sm (GlobalLinSysIterative.cpp:71)
}  // namespace

const double *IterativeSolver::inv_diag() const
{
    return space_ == CGSpace::StaticCond ? sc_->invd_g.as<double>() : invd_.as<double>();
}

int IterativeSolver::nboundary() const { return sc_ ? sc_->bmap->nglobal() : 0; }

void IterativeSolver::setup_static_cond()
{
    sc_ = std::make_unique<StaticCondData>();
    StaticCondData &d = *sc_;
    const int NG = map_.nglobal(), ND = map_.ndir();
    const std::vector<int> hmap = map_.host_map();
    const std::vector<double> hsign = map_.host_sign();
    std::vector<int> cnt(std::max(NG, 1), 0);
    for (int g : hmap) ++cnt[g];

    // mode lists, offsets and the map checks
    int64_t nlb = 0, nli = 0;
    std::vector<std::vector<int>> bmodes(ops_.size());
    for (size_t k = 0; k < ops_.size(); ++k) {
        MFOperator &o = *ops_[k];
        StaticCondBlock b;
        NMF_REQUIRE(o.equal_modes(), NMF_ERR_UNSUPPORTED,
                    "static condensation takes operators with equal modes in every direction");
        const std::vector<int> I = schur::interior_modes(o.shape, o.nm);
        bmodes[k] = schur::boundary_modes(o.shape, o.nm);
        const std::vector<int> &Bd = bmodes[k];
        b.E = o.nelmt;
        b.nc = o.ncoeff;
        b.ni = (int)I.size();
        b.nb = (int)Bd.size();
        b.off = off_[k];
        b.boff = nlb;
        b.ioff = nli;
        NMF_REQUIRE(b.ni <= schur::kMaxInterior, NMF_ERR_UNSUPPORTED,
                    "static condensation: more than " + std::to_string(schur::kMaxInterior) +
                        " interior modes per element");
        for (int e = 0; e < b.E; ++e)
            for (int m : I) {
                const int g = hmap[b.off + (size_t)e * b.nc + m];
                NMF_REQUIRE(cnt[g] == 1 && g >= ND, NMF_ERR_INVALID,
                            "static condensation: interior mode " + std::to_string(m) + " of element " +
                                std::to_string(e) + " of collection " + std::to_string(k) +
                                " shares its global DOF or is a Dirichlet DOF");
            }
        b.dI.allocate(std::max(b.ni, 1) * sizeof(int));
        b.dB.allocate(std::max(b.nb, 1) * sizeof(int));
        if (b.ni) NMF_HIP(hipMemcpy(b.dI.get(), I.data(), b.ni * sizeof(int), hipMemcpyHostToDevice));
        if (b.nb) NMF_HIP(hipMemcpy(b.dB.get(), Bd.data(), b.nb * sizeof(int), hipMemcpyHostToDevice));
        nlb += (int64_t)b.E * b.nb;
        nli += (int64_t)b.E * b.ni;
        d.blk.push_back(std::move(b));
    }
    d.nlb = nlb;
    d.nli = nli;

    // the boundary system: Dirichlet DOFs keep their ids 0..ND-1, the other
    // boundary DOFs follow in increasing global id (the reference numbers
    // the global boundary DOFs first, AssemblyMapCG)
    {
        std::vector<char> isb(std::max(NG, 1), 0);
        std::vector<int> mapb(std::max<int64_t>(nlb, 1)), bid(std::max(NG, 1), -1);
        std::vector<double> signb(hsign.empty() ? 0 : std::max<int64_t>(nlb, 1));
        for (size_t k = 0; k < d.blk.size(); ++k) {
            const StaticCondBlock &b = d.blk[k];
            for (int e = 0; e < b.E; ++e)
                for (int j = 0; j < b.nb; ++j) isb[hmap[b.off + (size_t)e * b.nc + bmodes[k][j]]] = 1;
        }
        int nbg = ND;
        for (int g = 0; g < ND; ++g) bid[g] = g;
        for (int g = ND; g < NG; ++g)
            if (isb[g]) bid[g] = nbg++;
        std::vector<int> bglob(std::max(nbg, 1), 0);
        for (int g = 0; g < NG; ++g)
            if (bid[g] >= 0) bglob[bid[g]] = g;
        for (size_t k = 0; k < d.blk.size(); ++k) {
            const StaticCondBlock &b = d.blk[k];
            for (int e = 0; e < b.E; ++e)
                for (int j = 0; j < b.nb; ++j) {
                    const size_t i = b.off + (size_t)e * b.nc + bmodes[k][j];
                    mapb[b.boff + (size_t)e * b.nb + j] = bid[hmap[i]];
                    if (!hsign.empty()) signb[b.boff + (size_t)e * b.nb + j] = hsign[i];
                }
        }
        d.bmap = std::make_unique<AssemblyMap>(mapb.data(), hsign.empty() ? nullptr : signb.data(), nlb, nbg, ND);
        d.bglob.allocate(bglob.size() * sizeof(int));
        NMF_HIP(hipMemcpy(d.bglob.get(), bglob.data(), bglob.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    const int NB = d.bmap->nglobal();
    d.nred = solver::reduction_blocks((size_t)std::max(NB, 1));
    auto dbl = [](DeviceBuffer &b, size_t n) { b.allocate(std::max<size_t>(n, 1) * sizeof(double)); };
    dbl(d.fb, nlb);
    dbl(d.fi, nli);
    dbl(d.ti, nli);
    dbl(d.zi, nli);
    dbl(d.xb, nlb);
    dbl(d.sb, nlb);
    dbl(d.invd_b, NB);
    dbl(d.invd_g, NG);
    dbl(d.bvec, NB);
    dbl(d.red, NB);
    for (auto &v : d.cg) dbl(v, NB);
    dbl(d.pa, (size_t)d.nred * 3);
    dbl(d.pmu, d.nred);
    dbl(d.xloc, map_.nloc());
    dbl(d.xg, NG);
    NMF_HIP(solver::launch_zero<double>(stream_, (size_t)std::max<int64_t>(map_.nloc(), 1), d.xloc.as<double>()));
    NMF_HIP(solver::launch_zero<double>(stream_, (size_t)std::max(NG, 1), d.xg.as<double>()));

    // per collection: L L^T = A_ii, A_ii^-1, and S_kk = A_kk - |L^-1 a_k|^2
    DeviceBuffer sloc(std::max<int64_t>(map_.nloc(), 1) * sizeof(double));
    for (size_t k = 0; k < d.blk.size(); ++k) {
        StaticCondBlock &b = d.blk[k];
        MFOperator &o = *ops_[k];
        const int E = b.E, nc = b.nc, ni = b.ni, nb = b.nb, KB = schur::batch_columns(ni);
        dbl(b.invD, (size_t)E * ni * ni);
        if (E == 0) continue;
        const size_t n = (size_t)E * nc;
        double *u = xloc_.as<double>() + off_[k], *y = yloc_.as<double>() + off_[k];
        NMF_REQUIRE(o.equal_modes(), NMF_ERR_UNSUPPORTED,
                    "static condensation takes operators with equal modes in every direction");
        const std::vector<int> I = schur::interior_modes(o.shape, o.nm);
        const std::vector<int> &Bd = bmodes[k];
        DeviceBuffer L(std::max<size_t>((size_t)E * ni * ni, 1) * sizeof(double));
        DeviceBuffer Ab(std::max<size_t>((size_t)E * ni * KB, 1) * sizeof(double));
        DeviceBuffer dkk((size_t)E * KB * sizeof(double)), info((size_t)E * sizeof(int));
        for (int c = 0; c < ni; ++c) {
            NMF_HIP(solver::launch_unit_mode(stream_, n, nc, I[c], u));
            o.apply(u, y, nullptr, nullptr, stream_);
            NMF_HIP(schur::launch_take_rows(stream_, E, 0, nc, ni, b.dI.as<int>(), 0, y, L.as<double>(),
                                            (size_t)ni * ni, c, ni));
        }
        if (ni) {
            NMF_HIP(schur::launch_cholesky(stream_, E, ni, L.as<double>(), info.as<int>()));
            std::vector<int> hinfo(E);
            NMF_HIP(hipMemcpyAsync(hinfo.data(), info.get(), E * sizeof(int), hipMemcpyDeviceToHost, stream_));
            NMF_HIP(hipStreamSynchronize(stream_));
            for (int e = 0; e < E; ++e)
                NMF_REQUIRE(hinfo[e] == 0, NMF_ERR_INVALID,
                            "static condensation: the interior block of element " + std::to_string(e) +
                                " of collection " + std::to_string(k) + " is not positive definite");
            // A_ii^-1 = L^-T L^-1 I, KB columns per pass, in place
            NMF_HIP(schur::launch_identity_blocks(stream_, E, ni, b.invD.as<double>()));
            for (int c0 = 0; c0 < ni; c0 += KB)
                NMF_HIP(schur::launch_chol_solve(stream_, E, ni, std::min(KB, ni - c0), ni, L.as<double>(),
                                                 b.invD.as<double>() + (size_t)c0 * ni));
        }
        for (int b0 = 0; b0 < nb; b0 += KB) {
            const int kb = std::min(KB, nb - b0);
            for (int c = 0; c < kb; ++c) {
                NMF_HIP(solver::launch_unit_mode(stream_, n, nc, Bd[b0 + c], u));
                o.apply(u, y, nullptr, nullptr, stream_);
                NMF_HIP(schur::launch_take_rows(stream_, E, 0, nc, ni, b.dI.as<int>(), 0, y, Ab.as<double>(),
                                                (size_t)ni * KB, c, ni));
                NMF_HIP(schur::launch_take_rows(stream_, E, 0, nc, 1, nullptr, Bd[b0 + c], y, dkk.as<double>(),
                                                (size_t)KB, c, 1));
            }
            NMF_HIP(schur::launch_schur_diag(stream_, E, ni, kb, KB, L.as<double>(), Ab.as<double>(),
                                             dkk.as<double>(), b.dB.as<int>() + b0, nc, 0,
                                             sloc.as<double>() + b.off));
        }
        if (nb)
            NMF_HIP(schur::launch_take_rows(stream_, E, 0, nc, nb, b.dB.as<int>(), 0, sloc.as<double>() + b.off,
                                            d.sb.as<double>() + b.boff, (size_t)nb, 0, nb));
        NMF_HIP(hipStreamSynchronize(stream_));  // L, Ab, dkk and info are freed here
    }

    // the preconditioner: the Schur diagonal assembled over the boundary
    // system (PreconditionerDiagonal.cpp:132-154), inverted off the Dirichlet rows
    d.bmap->assemble(d.sb.as<double>(), d.invd_b.as<double>(), stream_, false, nullptr, nullptr, true);
    NMF_HIP(solver::launch_invert_diag_dir(stream_, NB, ND, d.invd_b.as<double>()));
    // in the caller's numbering: boundary rows carry it, interior rows 0
    NMF_HIP(solver::launch_zero<double>(stream_, (size_t)std::max(NG, 1), d.invd_g.as<double>()));
    NMF_HIP(solver::launch_scatter_index(stream_, NB, d.bglob.as<int>(), d.invd_b.as<double>(), d.invd_g.as<double>()));
    NMF_HIP(hipStreamSynchronize(stream_));
}

// GlobalLinSysStaticCond::v_Solve (GlobalLinSysStaticCond.cpp:98-230) with
// homogeneous Dirichlet data (V_bnd = 0 on entry), one level
SolveInfo IterativeSolver::solve_static_cond(double tol, int maxit)
{
    StaticCondData &d = *sc_;
    tol = std::max(tol, 1.0e-16);  // NekLinSysIter.cpp:57
    const int NB = d.bmap->nglobal(), ND = d.bmap->ndir();
    const double *bloc = b_.as<double>();
    double *fb = d.fb.as<double>(), *fi = d.fi.as<double>(), *ti = d.ti.as<double>(), *zi = d.zi.as<double>();
    double *xb = d.xb.as<double>(), *sb = d.sb.as<double>();
    double *x = d.cg[0].as<double>(), *r = d.cg[1].as<double>(), *p = d.cg[2].as<double>(),
           *q = d.cg[3].as<double>(), *w = d.cg[4].as<double>(), *s = d.cg[5].as<double>();
    solver::CGState *st = state_.as<solver::CGState>();
    // y = A [src_b ; src_i] for collection k (either part may be null), in yloc_
    auto apply_parts = [&](const StaticCondBlock &b, size_t k, const double *src_b, const double *src_i) {
        double *u = xloc_.as<double>() + b.off;
        NMF_HIP(solver::launch_zero<double>(stream_, (size_t)b.E * b.nc, u));
        if (src_b && b.nb)
            NMF_HIP(schur::launch_put_rows(stream_, b.E, 0, b.nc, b.nb, b.dB.as<int>(), src_b, (size_t)b.nb, 0, b.nb,
                                           1.0, u));
        if (src_i && b.ni)
            NMF_HIP(schur::launch_put_rows(stream_, b.E, 0, b.nc, b.ni, b.dI.as<int>(), src_i, (size_t)b.ni, 0, b.ni,
                                           1.0, u));
        ops_[k]->apply(u, yloc_.as<double>() + b.off, nullptr, nullptr, stream_);
        return yloc_.as<double>() + b.off;
    };
    auto take_b = [&](const StaticCondBlock &b, const double *y, double *dst, double alpha, bool acc) {
        NMF_HIP(schur::launch_take_rows(stream_, b.E, 0, b.nc, b.nb, b.dB.as<int>(), 0, y, dst + b.boff, (size_t)b.nb,
                                        0, b.nb, alpha, acc));
    };
    auto take_i = [&](const StaticCondBlock &b, const double *y, double *dst, double alpha, bool acc) {
        NMF_HIP(schur::launch_take_rows(stream_, b.E, 0, b.nc, b.ni, b.dI.as<int>(), 0, y, dst + b.ioff, (size_t)b.ni,
                                        0, b.ni, alpha, acc));
    };
    // S xb -> sb (local boundary layout), matrix-free
    auto schur_apply = [&]() {
        for (size_t k = 0; k < d.blk.size(); ++k) {
            const StaticCondBlock &b = d.blk[k];
            if (b.E == 0 || b.nb == 0) continue;
            const double *y = apply_parts(b, k, xb + b.boff, nullptr);
            take_b(b, y, sb, 1.0, false);
            if (b.ni == 0) continue;
            take_i(b, y, ti, 1.0, false);
            NMF_HIP(schur::launch_block_gemv(stream_, b.E, b.ni, b.invD.as<double>(), ti + b.ioff, zi + b.ioff));
            y = apply_parts(b, k, nullptr, zi + b.ioff);
            take_b(b, y, sb, -1.0, true);
        }
    };
    hipEvent_t e0, e1;
    NMF_HIP(hipEventCreate(&e0));
    NMF_HIP(hipEventCreate(&e1));
    SolveInfo info;
    try {
        NMF_HIP(hipEventRecord(e0, stream_));
        // F_bnd, F_int (LocalToLocalBnd / LocalToLocalInt)
        for (const StaticCondBlock &b : d.blk) {
            if (b.E == 0) continue;
            take_b(b, bloc + b.off, fb, 1.0, false);
            take_i(b, bloc + b.off, fi, 1.0, false);
        }
        // v_PreSolve (GlobalLinSysIterativeStaticCond.cpp:436-492): |AssembleBnd(F_bnd)|^2
        // over all boundary rows, smoothed over successive solves
        d.bmap->assemble(fb, d.red.as<double>(), stream_);
        NMF_HIP(solver::launch_sumsq(stream_, (size_t)NB, d.red.as<double>(), d.pa.as<double>()));
        NMF_HIP(solver::launch_sum_partials(stream_, d.pa.as<double>(), solver::reduction_blocks((size_t)NB),
                                            d.pmu.as<double>()));
        double mag = 0.0;
        NMF_HIP(hipMemcpyAsync(&mag, d.pmu.get(), sizeof(double), hipMemcpyDeviceToHost, stream_));
        NMF_HIP(hipStreamSynchronize(stream_));
        const double new_mag = mag > 1.0e-6 ? mag : 1.0;
        d.rhs_mag = d.rhs_mag < 0.0 ? new_mag : kRhsMagSmoothing * d.rhs_mag + (1.0 - kRhsMagSmoothing) * new_mag;
        // F_bnd -= A_bi A_ii^-1 F_int (m_BinvD): the boundary rows of A [0 ; A_ii^-1 F_int]
        for (size_t k = 0; k < d.blk.size(); ++k) {
            const StaticCondBlock &b = d.blk[k];
            if (b.E == 0 || b.ni == 0 || b.nb == 0) continue;
            NMF_HIP(schur::launch_block_gemv(stream_, b.E, b.ni, b.invD.as<double>(), fi + b.ioff, zi + b.ioff));
            take_b(b, apply_parts(b, k, nullptr, zi + b.ioff), fb, -1.0, true);
        }
        // the boundary system: b = AssembleBnd(F_bnd), NekLinSysIterCG on S
        // (GlobalLinSysIterativeStaticCond.cpp:403-430)
        d.bmap->assemble(fb, d.bvec.as<double>(), stream_);
        solver::CGState init{};
        init.tol2 = tol * tol;
        init.maxit = maxit;
        init.rhs_mag_override = d.rhs_mag;
        *host_state_ = init;
        NMF_HIP(hipMemcpyAsync(st, host_state_, sizeof(solver::CGState), hipMemcpyHostToDevice, stream_));
        auto matvec = [&]() {
            d.bmap->global_to_local(w, xb, stream_);
            schur_apply();
            d.bmap->assemble(sb, s, stream_, true, w, d.pmu.as<double>());
        };
        NMF_HIP(solver::launch_cg_init_dirfirst(stream_, NB, ND, d.bvec.as<double>(), d.invd_b.as<double>(), x, r, p,
                                                q, w, d.pa.as<double>()));
        matvec();
        NMF_HIP(solver::launch_cg_scalar_init(stream_, st, d.pa.as<double>(), d.pmu.as<double>(), d.nred));
        const int max_rounds = maxit / kItersPerChunk + 3;
        for (int round = 0;; ++round) {
            NMF_HIP(hipMemcpyAsync(host_state_, st, sizeof(solver::CGState), hipMemcpyDeviceToHost, stream_));
            NMF_HIP(hipStreamSynchronize(stream_));
            if (host_state_->done || round >= max_rounds) break;
            for (int it = 0; it < kItersPerChunk; ++it) {
                NMF_HIP(solver::launch_cg_update_n(stream_, (size_t)NB, st, p, q, x, r, w, s, d.invd_b.as<double>(),
                                                   d.pa.as<double>()));
                matvec();
                NMF_HIP(solver::launch_cg_scalar(stream_, st, d.pa.as<double>(), d.pmu.as<double>(), d.nred));
            }
        }
        // V_bnd = GlobalToLocalBnd(x); V_int = A_ii^-1 (F_int - A_ib V_bnd)
        d.bmap->global_to_local(x, xb, stream_);
        for (size_t k = 0; k < d.blk.size(); ++k) {
            const StaticCondBlock &b = d.blk[k];
            if (b.E == 0) continue;
            double *xl = d.xloc.as<double>() + b.off;
            if (b.ni) {
                take_i(b, apply_parts(b, k, xb + b.boff, nullptr), fi, -1.0, true);
                NMF_HIP(schur::launch_block_gemv(stream_, b.E, b.ni, b.invD.as<double>(), fi + b.ioff, zi + b.ioff));
                NMF_HIP(schur::launch_put_rows(stream_, b.E, 0, b.nc, b.ni, b.dI.as<int>(), zi + b.ioff, (size_t)b.ni,
                                               0, b.ni, 1.0, xl));
            }
            if (b.nb)
                NMF_HIP(schur::launch_put_rows(stream_, b.E, 0, b.nc, b.nb, b.dB.as<int>(), xb + b.boff, (size_t)b.nb,
                                               0, b.nb, 1.0, xl));
        }
        map_.global_values(d.xloc.as<double>(), d.xg.as<double>(), stream_);
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
