// Host-side check of the runtime-size operator stages
// (nektar_amd/csrc/generic_kernels.hpp): the same element code compiled as
// plain C++, one "work-item" running every stage in order, built with
// AddressSanitizer, so that an out-of-bounds index into the intermediate
// buffers, tables or index lists aborts the run.  Results are compared with
// the oracle's per-direction operators (oracle/liboracle.so).  Geometry is
// packed here in the layouts the kernels read (MFOperator::pack_geometry).
// Prints "OK <max rel err>" and exits 0 on success.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "generic_kernels.hpp"

extern "C" {
void or_bwdtrans_aniso(int shape, const int *nm, const int *nq, int nelmt, const double *in, double *out);
void or_iproduct_aniso(int shape, const int *nm, const int *nq, int nelmt, int deformed, const double *jac,
                       const double *in, double *out);
void or_physderiv_aniso(int shape, const int *nm, const int *nq, int nelmt, int deformed, const double *df,
                        const double *in, double *out0, double *out1, double *out2);
void or_iproduct_deriv_aniso(int shape, const int *nm, const int *nq, int nelmt, int deformed, const double *jac,
                             const double *df, const double *in0, const double *in1, const double *in2, double *out);
void or_helmholtz_aniso(int shape, const int *nm, const int *nq, int nelmt, int deformed, const double *jac,
                        const double *df, double lambda, const double *diff6, const double *in, double *out);
void or_adv_diff_react_aniso(int shape, const int *nm, const int *nq, int nelmt, int deformed, const double *jac,
                             const double *df, const double *vel, double lambda, const double *diff6,
                             const double *in, double *out);
}

using namespace nektar_amd;

namespace {

double relerr(const std::vector<double> &a, const std::vector<double> &b)
{
    double m = 0, n = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        m = std::fmax(m, std::fabs(a[i] - b[i]));
        n = std::fmax(n, std::fabs(b[i]));
    }
    return m / (n > 0 ? n : 1.0);
}

// one element at a time, every buffer its own heap block of exactly `buf`
// values (so that AddressSanitizer sees an overrun of any of them)
void run(GenArgs a)
{
    for (int e = 0; e < a.nelmt; ++e) {
        std::vector<double> X0(a.buf), X1(a.buf), X2(a.buf), X3(a.buf);
        if (a.shape == Shape::Hex) gen::element<double, gen::SH_HEX>(a, e, X0.data(), X1.data(), X2.data(), X3.data());
        else if (a.shape == Shape::Prism)
            gen::element<double, gen::SH_PRISM>(a, e, X0.data(), X1.data(), X2.data(), X3.data());
        else if (a.shape == Shape::Pyr)
            gen::element<double, gen::SH_PYR>(a, e, X0.data(), X1.data(), X2.data(), X3.data());
        else gen::element<double, gen::SH_TET>(a, e, X0.data(), X1.data(), X2.data(), X3.data());
    }
}

struct Case {
    Shape shape;
    int nm[3], nq[3], E;
};

}  // namespace

int main()
{
    const Case cases[] = {
        {Shape::Hex, {3, 4, 5}, {4, 5, 6}, 5},   {Shape::Hex, {4, 4, 4}, {6, 5, 7}, 4},
        {Shape::Hex, {6, 3, 4}, {9, 4, 6}, 3},   {Shape::Hex, {2, 5, 3}, {3, 5, 4}, 4},
        {Shape::Hex, {4, 4, 4}, {8, 8, 8}, 2},   {Shape::Tet, {3, 4, 5}, {5, 5, 5}, 5},
        {Shape::Tet, {2, 3, 4}, {4, 6, 5}, 4},   {Shape::Tet, {4, 4, 4}, {6, 4, 7}, 3},
        {Shape::Tet, {5, 5, 5}, {7, 6, 6}, 3},   {Shape::Tet, {2, 2, 2}, {3, 2, 2}, 4},
        {Shape::Tet, {6, 6, 9}, {4, 3, 3}, 2},   {Shape::Prism, {3, 3, 3}, {4, 4, 3}, 5},
        {Shape::Prism, {3, 4, 5}, {5, 4, 5}, 4}, {Shape::Prism, {4, 2, 4}, {5, 6, 5}, 3},
        {Shape::Prism, {2, 2, 2}, {3, 3, 2}, 4}, {Shape::Prism, {6, 6, 6}, {8, 8, 7}, 2},
        {Shape::Pyr, {2, 2, 2}, {3, 3, 2}, 4},   {Shape::Pyr, {3, 3, 3}, {4, 4, 3}, 5},
        {Shape::Pyr, {4, 4, 4}, {6, 5, 5}, 3},   {Shape::Pyr, {5, 5, 5}, {7, 7, 6}, 2},
        {Shape::Pyr, {6, 6, 6}, {8, 8, 7}, 2},
    };
    double worst = 0.0;
    std::mt19937_64 rng(11);
    std::uniform_real_distribution<double> U(-1.0, 1.0), Upos(0.5, 1.5);
    for (const Case &c : cases) {
        const GenLayout L = generic_layout(c.shape, c.nm, c.nq);
        // S: the oracle's shape (0 hex, 1 tet, 2 prism, 3 pyramid); all but
        // the hex pack df and J
        const int S = (int)c.shape, E = c.E, nqt = L.nqtot, nc = L.ncoeff;
        const Basis1D P0 = make_basis(default_basis(c.shape, 0), c.nm[0], c.nq[0]),
                      P1 = make_basis(default_basis(c.shape, 1), c.nm[1], c.nq[1]),
                      P2 = make_basis(default_basis(c.shape, 2), c.nm[2], c.nq[2]);
        for (int deformed = 0; deformed < 2; ++deformed) {
            const size_t gp = deformed ? (size_t)E * nqt : (size_t)E;
            std::vector<double> jac(gp), df(9 * gp), x((size_t)E * nc), f0((size_t)E * nqt), f1(f0.size()),
                f2(f0.size()), vel(3 * f0.size());
            for (auto &v : jac) v = Upos(rng);
            for (auto *a : {&df, &x, &f0, &f1, &f2, &vel})
                for (auto &v : *a) v = U(rng);
            // weight of point q (the Radau pre-scaling included)
            auto W = [&](int q) {
                const int i = q % c.nq[0], j = (q / c.nq[0]) % c.nq[1], k = q / (c.nq[0] * c.nq[1]);
                return P0.pts.w[i] * weight_prescale(P0.pts.type) * P1.pts.w[j] * weight_prescale(P1.pts.type) *
                       P2.pts.w[k] * weight_prescale(P2.pts.type);
            };
            // the packed layouts: regular [e*G + c], deformed [(e*G + c)*nqt + q]
            auto at = [&](int G, int e, int cc, int q) {
                return deformed ? ((size_t)e * G + cc) * nqt + q : (size_t)e * G + cc;
            };
            std::vector<double> gH(gp * (S ? 10 : 7)), gD(gp * 9), gI(gp * (S ? 10 : 9));
            for (int e = 0; e < E; ++e)
                for (int q = 0; q < (deformed ? nqt : 1); ++q) {
                    const size_t pt = deformed ? (size_t)e * nqt + q : (size_t)e;
                    const double *f = &df[pt * 9];
                    const double w = deformed ? W(q) : 1.0;
                    for (int d = 0; d < 9; ++d) gD[at(9, e, d, q)] = f[d];
                    if (S) {
                        for (int d = 0; d < 9; ++d) gH[at(10, e, d, q)] = gI[at(10, e, d, q)] = f[d];
                        gH[at(10, e, 9, q)] = gI[at(10, e, 9, q)] = jac[pt];
                    } else {
                        const int idx[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
                        for (int m = 0; m < 6; ++m) {
                            const int d = idx[m][0], ee = idx[m][1];
                            gH[at(7, e, m, q)] = w * jac[pt] * (f[d] * f[ee] + f[3 + d] * f[3 + ee] + f[6 + d] * f[6 + ee]);
                        }
                        gH[at(7, e, 6, q)] = w * jac[pt];
                        for (int d = 0; d < 9; ++d) gI[at(9, e, d, q)] = w * jac[pt] * f[d];
                    }
                }
            // advection coefficients c_d = sum_j V_j df[j*3+d], tets with the
            // collapsed factors folded in
            std::vector<double> adv((size_t)E * 3 * nqt);
            for (int e = 0; e < E; ++e)
                for (int q = 0; q < nqt; ++q) {
                    const int i = q % c.nq[0], j = (q / c.nq[0]) % c.nq[1], k = q / (c.nq[0] * c.nq[1]);
                    const double *f = &df[(deformed ? (size_t)e * nqt + q : (size_t)e) * 9];
                    const double *v = &vel[((size_t)e * nqt + q) * 3];
                    double a3[3];
                    for (int d = 0; d < 3; ++d) a3[d] = v[0] * f[d] + v[1] * f[3 + d] + v[2] * f[6 + d];
                    if (S == 2) {
                        const double h0 = 0.5 * (1 + P0.pts.z[i]), h1 = 2 / (1 - P2.pts.z[k]);
                        a3[0] = h1 * (a3[0] + h0 * a3[2]);
                    } else if (S == 3) {
                        const double h0 = 0.5 * (1 + P0.pts.z[i]), h1 = 0.5 * (1 + P1.pts.z[j]),
                                     h2 = 2 / (1 - P2.pts.z[k]);
                        a3[0] = h2 * (a3[0] + h0 * a3[2]);
                        a3[1] = h2 * (a3[1] + h1 * a3[2]);
                    } else if (S) {
                        const double h0 = 0.5 * (1 + P0.pts.z[i]), h1 = 0.5 * (1 + P1.pts.z[j]),
                                     h2 = 2 / (1 - P1.pts.z[j]), h3 = 2 / (1 - P2.pts.z[k]);
                        const double b0 = a3[0], b1 = a3[1], b2 = a3[2];
                        a3[0] = h2 * h3 * (b0 + h0 * (b1 + b2));
                        a3[1] = h3 * (b1 + h1 * b2);
                        a3[2] = b2;
                    }
                    for (int d = 0; d < 3; ++d) adv[((size_t)e * 3 + d) * nqt + q] = a3[d];
                }
            GenArgs a{};
            a.shape = c.shape;
            a.deformed = deformed;
            a.nelmt = E;
            for (int d = 0; d < 3; ++d) {
                a.nm[d] = c.nm[d];
                a.nq[d] = c.nq[d];
                a.oB[d] = L.oB[d];
                a.oD[d] = L.oD[d];
                a.oW[d] = L.oW[d];
            }
            for (int h = 0; h < 4; ++h) a.oH[h] = L.oH[h];
            a.nqtot = nqt;
            a.ncoeff = nc;
            a.npq = L.npq;
            a.buf = L.buf;
            std::vector<double> tab = L.tab;  // exact-size copies
            std::vector<int> itab = L.itab;
            a.tab = tab.data();
            a.itab = itab.empty() ? nullptr : itab.data();
            std::vector<double> o0((size_t)E * nqt), o1(o0.size()), o2(o0.size()), oc((size_t)E * nc);
            std::vector<double> r0(o0.size()), r1(o0.size()), r2(o0.size()), rc(oc.size());
            // BwdTrans
            a.op = OpType::BwdTrans;
            a.in0 = x.data();
            a.out0 = o0.data();
            run(a);
            or_bwdtrans_aniso(S, c.nm, c.nq, E, x.data(), r0.data());
            worst = std::fmax(worst, relerr(o0, r0));
            // IProductWRTBase
            a.op = OpType::IProduct;
            a.geo = jac.data();
            a.in0 = f0.data();
            a.out0 = oc.data();
            run(a);
            or_iproduct_aniso(S, c.nm, c.nq, E, deformed, jac.data(), f0.data(), rc.data());
            worst = std::fmax(worst, relerr(oc, rc));
            // PhysDeriv
            a.op = OpType::PhysDeriv;
            a.geo = gD.data();
            a.in0 = f0.data();
            a.out0 = o0.data();
            a.out1 = o1.data();
            a.out2 = o2.data();
            run(a);
            or_physderiv_aniso(S, c.nm, c.nq, E, deformed, df.data(), f0.data(), r0.data(), r1.data(), r2.data());
            worst = std::fmax(worst, std::fmax(relerr(o0, r0), std::fmax(relerr(o1, r1), relerr(o2, r2))));
            // IProductWRTDerivBase
            a.op = OpType::IProductDeriv;
            a.geo = gI.data();
            a.in0 = f0.data();
            a.in1 = f1.data();
            a.in2 = f2.data();
            a.out0 = oc.data();
            run(a);
            or_iproduct_deriv_aniso(S, c.nm, c.nq, E, deformed, jac.data(), df.data(), f0.data(), f1.data(),
                                    f2.data(), rc.data());
            worst = std::fmax(worst, relerr(oc, rc));
            // Helmholtz
            a.op = OpType::Helmholtz;
            a.geo = gH.data();
            a.lambda = 0.7;
            a.in0 = x.data();
            run(a);
            or_helmholtz_aniso(S, c.nm, c.nq, E, deformed, jac.data(), df.data(), 0.7, nullptr, x.data(), rc.data());
            worst = std::fmax(worst, relerr(oc, rc));
            // LinearAdvectionDiffusionReaction
            a.op = OpType::AdvDiffReact;
            a.adv = adv.data();
            a.lambda = 1.3;
            run(a);
            or_adv_diff_react_aniso(S, c.nm, c.nq, E, deformed, jac.data(), df.data(), vel.data(), 1.3, nullptr,
                                    x.data(), rc.data());
            worst = std::fmax(worst, relerr(oc, rc));
        }
    }
    std::printf("OK %.3e\n", worst);
    return worst <= 1e-12 ? 0 : 1;
}
