// C++ client of the MI355X library written against the reference-style
// interface (include/nektar_amd/MatrixFreeOps.hpp), checked against the
// oracle.  Exercises factory keys, SetJac/SetDF, UpdateFactors (lambda and
// a constant diffusion tensor), Helmholtz and PhysDeriv on host arrays and
// the error path; anisotropic BasisKeys; the Collections::Operator mirror
// (include/nektar_amd/Collections.hpp: OperatorKey factory, the 5-argument
// and single-direction forms); then the MultiRegions mirror (include/nektar_amd/
// GlobalLinSys.hpp): AssemblyMapCG::Assemble and the three solver choices
// against the oracle's CG on a 3^3 box.  Prints "OK <max rel err>" and
// exits 0 on success.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "nektar_amd/Collections.hpp"
#include "nektar_amd/GlobalLinSys.hpp"
#include "nektar_amd/MatrixFreeOps.hpp"

extern "C" {
typedef struct {
    int iterations;
    double final_error;
    double rhs_magnitude;
} or_cg_info;
int or_hex_nglobal(int n, int nm);
void or_hex_l2g(int n, int nm, int *map, unsigned char *dirichlet);
int or_cg_helmholtz_hex(int n, int nm, int nq, int deformed, const double *jac, const double *df, double lambda,
                        const double *b_glob, const double *inv_diag, double tol, int maxit, double *x_glob,
                        or_cg_info *info);
void or_helmholtz(int shape, int nm, int nq0, int nelmt, int deformed, const double *jac, const double *df,
                  double lambda, const double *diff6, const double *in, double *out);
void or_physderiv(int shape, int nm, int nq0, int nelmt, int deformed, const double *df, const double *in,
                  double *out0, double *out1, double *out2);
void or_iproduct_deriv(int shape, int nm, int nq0, int nelmt, int deformed, const double *jac, const double *df,
                       const double *in0, const double *in1, const double *in2, double *out);
void or_helmholtz_aniso(int shape, const int *nm, const int *nq, int nelmt, int deformed, const double *jac,
                        const double *df, double lambda, const double *diff6, const double *in, double *out);
}

using namespace Nektar;

static double relerr(const std::vector<double> &a, const std::vector<double> &b)
{
    double m = 0, n = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        m = std::fmax(m, std::fabs(a[i] - b[i]));
        n = std::fmax(n, std::fabs(b[i]));
    }
    return m / n;
}

int main()
{
    const int nm = 6, nq = 7, ne = 5;
    std::vector<LibUtilities::BasisKey> basis(3, {nm, nq});
    // parallelepiped elements: one J and df per element
    std::vector<double> jac(ne), df(9 * ne);
    std::mt19937_64 rng(3);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    for (int e = 0; e < ne; ++e) {
        double A[3][3];  // d x_k / d xi_i = A[k][i]
        for (int k = 0; k < 3; ++k)
            for (int i = 0; i < 3; ++i) A[k][i] = (k == i ? 0.5 : 0.0) + 0.1 * U(rng);
        const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1], c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2],
                     c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
        const double det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
        double inv[3][3];  // inv[i][k] = d xi_i / d x_k
        inv[0][0] = c00 / det;
        inv[1][0] = c01 / det;
        inv[2][0] = c02 / det;
        inv[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
        inv[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
        inv[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
        inv[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
        inv[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
        inv[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
        jac[e] = std::fabs(det);
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) df[e * 9 + j * 3 + k] = inv[k][j];
    }
    auto op = MatrixFree::GetOperatorFactory().CreateInstance(
        "Helmholtz" + MatrixFree::GetOpstring(LibUtilities::eHexahedron, false), basis, ne);
    auto helm = std::dynamic_pointer_cast<MatrixFree::Helmholtz>(op);
    if (!helm || !op->NeedsDF() || !op->NeedsJac()) return 2;
    op->SetJac(jac);
    op->SetDF(df);
    Collections::FactorMap factors = {{Collections::eFactorLambda, 1.5},
                                      {Collections::eFactorCoeffD00, 2.0},
                                      {Collections::eFactorCoeffD01, 0.3}};
    Collections::UpdateFactors(*helm, factors);
    std::vector<double> in(op->GetInputSize()), out, ref(op->GetOutputSize());
    for (auto &v : in) v = U(rng);
    (*op)(in, out);
    const double d6[6] = {2.0, 0.3, 1.0, 0.0, 0.0, 1.0};
    or_helmholtz(0, nm, nq, ne, 0, jac.data(), df.data(), 1.5, d6, in.data(), ref.data());
    double err = relerr(out, ref);

    auto pd = std::dynamic_pointer_cast<MatrixFree::PhysDeriv>(
        MatrixFree::GetOperatorFactory().CreateInstance("PhysDeriv_Hex_Regular", basis, ne));
    pd->SetDF(df);
    std::vector<double> u(pd->GetInputSize());
    for (auto &v : u) v = U(rng);
    std::vector<std::vector<double>> d;
    (*pd)(u, d);
    std::vector<double> r0(u.size()), r1(u.size()), r2(u.size());
    or_physderiv(0, nm, nq, ne, 0, df.data(), u.data(), r0.data(), r1.data(), r2.data());
    err = std::fmax(err, std::fmax(relerr(d[0], r0), std::fmax(relerr(d[1], r1), relerr(d[2], r2))));

    bool threw = false;
    try {
        MatrixFree::GetOperatorFactory().CreateInstance("Helmholtz_Quad_Regular", basis, ne);
    } catch (const NekError &e) {
        threw = e.code == NMF_ERR_UNSUPPORTED;
    }
    if (!threw) return 3;

    // one BasisKey per direction: a single key is refused, an anisotropic
    // expansion (modes 4/5/6, points 6/7/8) runs and matches the oracle's
    // per-direction operator
    threw = false;
    try {
        MatrixFree::GetOperatorFactory().CreateInstance("Helmholtz_Hex_Regular",
                                                        std::vector<LibUtilities::BasisKey>(1, {nm, nq}), ne);
    } catch (const NekError &e) {
        threw = e.code == NMF_ERR_INVALID;
    }
    if (!threw) return 4;
    {
        const std::vector<LibUtilities::BasisKey> ab = {{4, 6}, {5, 7}, {6, 8}};
        auto ah = std::dynamic_pointer_cast<MatrixFree::Helmholtz>(
            MatrixFree::GetOperatorFactory().CreateInstance("Helmholtz_Hex_Regular", ab, ne));
        if (!ah || ah->GetInputSize() != ne * 4 * 5 * 6) return 5;
        ah->SetJac(jac);
        ah->SetDF(df);
        ah->SetLambda(0.8);
        std::vector<double> ax(ah->GetInputSize()), ay;
        for (auto &v : ax) v = U(rng);
        (*ah)(ax, ay);
        const int am[3] = {4, 5, 6}, aq[3] = {6, 7, 8};
        std::vector<double> ar(ay.size());
        or_helmholtz_aniso(0, am, aq, ne, 0, jac.data(), df.data(), 0.8, nullptr, ax.data(), ar.data());
        err = std::fmax(err, relerr(ay, ar));
    }

    // Collections::Operator: OperatorKey factory, the 5-argument and the
    // single-direction forms, UpdateFactors, GetWspSize, unsupported keys
    {
        Collections::CollectionExpansion ce;
        ce.shape = LibUtilities::eHexahedron;
        ce.basis = basis;
        ce.nElmt = ne;
        ce.jac = jac;
        ce.df = df;
        auto &CF = Collections::GetOperatorFactory();
        auto ch = CF.CreateInstance(
            Collections::OperatorKey(LibUtilities::eHexahedron, Collections::eHelmholtz, Collections::eMatrixFree,
                                     false),
            ce, Collections::FactorMap{{Collections::eFactorLambda, 0.6}});
        if (ch->GetNumElmt() != (unsigned)ne || ch->GetWspSize() != 0 || ch->GetCoeffSize() != ch->GetInputSize())
            return 6;
        std::vector<double> cin(ch->GetInputSize()), c0, c1, c2, wsp, cref(ch->GetOutputSize());
        for (auto &v : cin) v = U(rng);
        (*ch)(cin, c0, c1, c2, wsp);
        or_helmholtz(0, nm, nq, ne, 0, jac.data(), df.data(), 0.6, nullptr, cin.data(), cref.data());
        err = std::fmax(err, relerr(c0, cref));
        auto cp = CF.CreateInstance(Collections::OperatorKey(LibUtilities::eHexahedron, Collections::ePhysDeriv,
                                                             Collections::eMatrixFree, false),
                                    ce);
        if (cp->GetWspSize() != 2 * cp->GetPhysSize()) return 7;
        std::vector<double> pu(cp->GetInputSize()), p0, p1, p2, pdir;
        for (auto &v : pu) v = U(rng);
        (*cp)(pu, p0, p1, p2, wsp);
        std::vector<double> q0(pu.size()), q1(pu.size()), q2(pu.size());
        or_physderiv(0, nm, nq, ne, 0, df.data(), pu.data(), q0.data(), q1.data(), q2.data());
        err = std::fmax(err, std::fmax(relerr(p0, q0), std::fmax(relerr(p1, q1), relerr(p2, q2))));
        (*cp)(1, pu, pdir, wsp);
        if (pdir != p1) return 8;  // the same derivative, bit for bit
        // IProductWRTDerivBase: inputs in entries 0-2, output in entry 3
        auto cd = CF.CreateInstance(Collections::OperatorKey(LibUtilities::eHexahedron,
                                                             Collections::eIProductWRTDerivBase,
                                                             Collections::eMatrixFree, false),
                                    ce);
        std::vector<double> e1 = q1, e2 = q2, e3;
        (*cd)(q0, e1, e2, e3, wsp);
        std::vector<double> dref(cd->GetOutputSize());
        or_iproduct_deriv(0, nm, nq, ne, 0, jac.data(), df.data(), q0.data(), q1.data(), q2.data(), dref.data());
        err = std::fmax(err, relerr(e3, dref));
        bool threw = false;
        try {
            CF.CreateInstance(Collections::OperatorKey(LibUtilities::eHexahedron, Collections::eHelmholtz,
                                                       Collections::eSumFac, false),
                              ce);
        } catch (const NekError &e) {
            threw = e.code == NMF_ERR_UNSUPPORTED;
        }
        if (!threw) return 9;
        threw = false;
        try {
            (*ch)(0, cin, c0, wsp);
        } catch (const NekError &) {
            threw = true;
        }
        if (!threw) return 10;
    }

    // MultiRegions mirror: the structured 3^3 box, P=4, its C0 map renumbered
    // Dirichlet-first, solved by the three solver choices against the oracle's CG
    {
        const int n = 3, m = 5, q = 6, E = n * n * n, nc = m * m * m;
        const int ng = or_hex_nglobal(n, m);
        std::vector<int> smap(E * nc);
        std::vector<unsigned char> dir(ng);
        or_hex_l2g(n, m, smap.data(), dir.data());
        std::vector<int> renum(ng);
        int nd = 0;
        for (int g = 0; g < ng; ++g)
            if (dir[g]) renum[g] = nd++;
        for (int g = 0, k = nd; g < ng; ++g)
            if (!dir[g]) renum[g] = k++;
        std::vector<int> map(E * nc);
        for (int i = 0; i < E * nc; ++i) map[i] = renum[smap[i]];
        const double h = 1.0 / n;
        std::vector<double> J(E, h * h * h / 8), D(9 * E, 0.0);
        for (int e = 0; e < E; ++e) D[e * 9 + 0] = D[e * 9 + 4] = D[e * 9 + 8] = 2.0 / h;
        std::vector<double> bloc(E * nc), bglob(ng, 0.0), xref(ng);
        for (auto &v : bloc) v = U(rng);
        for (int i = 0; i < E * nc; ++i) bglob[smap[i]] += bloc[i];
        or_cg_info oi;
        or_cg_helmholtz_hex(n, m, q, 0, J.data(), D.data(), 2.0, bglob.data(), nullptr, 1e-13, 5000, xref.data(), &oi);
        MultiRegions::AssemblyMapCG amap(map, {}, ng, nd);
        std::vector<double> bg;
        amap.Assemble(bloc, bg);
        double aerr = 0, an = 0;
        for (int g = 0; g < ng; ++g) {
            aerr = std::fmax(aerr, std::fabs(bg[renum[g]] - bglob[g]));
            an = std::fmax(an, std::fabs(bglob[g]));
        }
        if (aerr > 1e-14 * an) return 4;
        std::vector<std::shared_ptr<MatrixFree::Operator>> ops = {MatrixFree::GetOperatorFactory().CreateInstance(
            "Helmholtz_Hex_Regular", std::vector<LibUtilities::BasisKey>(3, {m, q}), E)};
        ops[0]->SetJac(J);
        ops[0]->SetDF(D);
        std::dynamic_pointer_cast<MatrixFree::Helmholtz>(ops[0])->SetLambda(2.0);
        const char *soln[3][2] = {{"IterativeFull", "ConjugateGradient"},
                                  {"IterativeFull", "ConjugateGradientLoc"},
                                  {"IterativeStaticCond", "ConjugateGradient"}};
        double serr = 0;
        for (auto &c : soln) {
            MultiRegions::GlobalLinSysIterative ls(amap, ops, c[0], c[1]);
            std::vector<double> xloc;
            ls.Solve(bloc, xloc, 1e-12, 5000);
            if (!ls.GetLastSolveInfo().converged) return 5;
            double e = 0, nrm = 0;
            for (int i = 0; i < E * nc; ++i) {
                e = std::fmax(e, std::fabs(xloc[i] - (dir[smap[i]] ? 0.0 : xref[smap[i]])));
                nrm = std::fmax(nrm, std::fabs(xref[smap[i]]));
            }
            serr = std::fmax(serr, e / nrm);
        }
        if (serr > 1e-8) {
            std::printf("FAIL solve %.3e\n", serr);
            return 6;
        }
        bool bad = false;
        try {
            MultiRegions::GlobalLinSysIterative ls(amap, ops, "DirectFull");
        } catch (const NekError &e) {
            bad = e.code == NMF_ERR_UNSUPPORTED;
        }
        if (!bad) return 7;
    }
    std::printf("%s %.3e\n", err <= 1e-12 ? "OK" : "FAIL", err);
    return err <= 1e-12 ? 0 : 1;
}
