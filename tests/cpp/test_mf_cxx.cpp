// C++ client of the MI355X library written against the reference-style
// interface (include/nektar_amd/MatrixFreeOps.hpp), checked against the
// oracle.  Exercises factory keys, SetJac/SetDF, UpdateFactors (lambda and
// a constant diffusion tensor), Helmholtz and PhysDeriv on host arrays and
// the error path.  Prints "OK <max rel err>" and exits 0 on success.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "nektar_amd/MatrixFreeOps.hpp"

extern "C" {
void or_helmholtz(int shape, int nm, int nq0, int nelmt, int deformed, const double *jac, const double *df,
                  double lambda, const double *diff6, const double *in, double *out);
void or_physderiv(int shape, int nm, int nq0, int nelmt, int deformed, const double *df, const double *in,
                  double *out0, double *out1, double *out2);
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
        MatrixFree::GetOperatorFactory().CreateInstance("Helmholtz_Prism_Regular", basis, ne);
    } catch (const NekError &e) {
        threw = e.code == NMF_ERR_UNSUPPORTED;
    }
    if (!threw) return 3;
    std::printf("%s %.3e\n", err <= 1e-12 ? "OK" : "FAIL", err);
    return err <= 1e-12 ? 0 : 1;
}
