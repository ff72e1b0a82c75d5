// Quadrature and modified bases on the host (FP64).
//
// Points: zeros of the relevant Jacobi polynomial are the eigenvalues of its
// symmetric tridiagonal Jacobi matrix (Golub-Welsch), isolated here by Sturm
// bisection and polished by Newton steps on the polynomial itself.  Weights
// use the closed Gauss-Lobatto / Gauss-Radau-Jacobi formulas (Karniadakis &
// Sherwin, App. B); the derivative matrix uses barycentric weights with the
// negative-sum diagonal.  The point sets are those of
// library/LibUtilities/Foundations/GaussPoints.cpp:117-160 (GLL = zwglj(0,0),
// Radau-M = zwgrjm(alpha,0)).  The bases follow the definitions of
// Basis.cpp:466-645 (Modified_A/B/C) with dbdata = D * bdata (:493-495).
#include "basis.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace nektar_amd {

double jacobi(int n, double a, double b, double x)
{
    if (n == 0) return 1.0;
    double p0 = 1.0, p1 = 0.5 * (a - b + (a + b + 2.0) * x);
    for (int k = 2; k <= n; ++k) {
        const double c = 2.0 * k + a + b;
        const double a1 = 2.0 * k * (k + a + b) * (c - 2.0);
        const double a2 = (c - 1.0) * (a * a - b * b);
        const double a3 = (c - 2.0) * (c - 1.0) * c;
        const double a4 = 2.0 * (k + a - 1.0) * (k + b - 1.0) * c;
        const double p2 = ((a2 + a3 * x) * p1 - a4 * p0) / a1;
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

double jacobi_deriv(int n, double a, double b, double x)
{
    if (n == 0) return 0.0;
    return 0.5 * (n + a + b + 1.0) * jacobi(n - 1, a + 1.0, b + 1.0, x);
}

namespace {

// Zeros of P_n^{a,b} in ascending order.
std::vector<double> jacobi_zeros(int n, double a, double b)
{
    std::vector<double> z(n);
    if (n <= 0) return z;
    // Jacobi matrix of the monic recurrence: diagonal al[k], off-diagonal
    // squared be[k] (k >= 1).
    std::vector<double> al(n), be(n, 0.0);
    for (int k = 0; k < n; ++k) {
        const double c = 2.0 * k + a + b;
        al[k] = (k == 0) ? (b - a) / (a + b + 2.0) : (b * b - a * a) / (c * (c + 2.0));
        if (k >= 1)
            be[k] = 4.0 * k * (k + a) * (k + b) * (k + a + b) / (c * c * (c + 1.0) * (c - 1.0));
    }
    // Sturm count: number of eigenvalues < x.
    auto count_below = [&](double x) {
        int cnt = 0;
        double d = al[0] - x;
        if (d < 0) ++cnt;
        for (int k = 1; k < n; ++k) {
            if (d == 0.0) d = 1e-300;
            d = (al[k] - x) - be[k] / d;
            if (d < 0) ++cnt;
        }
        return cnt;
    };
    for (int m = 0; m < n; ++m) {
        double lo = -1.0, hi = 1.0;
        for (int it = 0; it < 200 && hi - lo > 0.0; ++it) {
            const double mid = 0.5 * (lo + hi);
            if (mid == lo || mid == hi) break;
            if (count_below(mid) > m) hi = mid;
            else lo = mid;
        }
        double x = 0.5 * (lo + hi);
        for (int it = 0; it < 3; ++it) {
            const double pd = jacobi_deriv(n, a, b, x);
            if (pd == 0.0) break;
            const double dx = jacobi(n, a, b, x) / pd;
            if (!(std::fabs(dx) < 1e-6)) break;
            x -= dx;
        }
        z[m] = x;
    }
    return z;
}

void derivative_matrix(const std::vector<double> &z, std::vector<double> &D)
{
    const int n = (int)z.size();
    std::vector<double> lam(n, 1.0);
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k)
            if (k != j) lam[j] /= (z[j] - z[k]);
    D.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        double diag = 0.0;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            const double v = (lam[j] / lam[i]) / (z[i] - z[j]);  // l_j'(z_i)
            D[(size_t)j * n + i] = v;
            diag -= v;
        }
        D[(size_t)i * n + i] = diag;
    }
}

}  // namespace

Points1D make_points(PointsType t, int nq)
{
    if (nq < 2 || nq > 32) throw std::invalid_argument("make_points: nq out of range [2,32]");
    Points1D P;
    P.nq = nq;
    P.type = t;
    P.z.assign(nq, 0.0);
    P.w.assign(nq, 0.0);
    if (t == PointsType::GLL) {
        P.z[0] = -1.0;
        P.z[nq - 1] = 1.0;
        auto zi = jacobi_zeros(nq - 2, 1.0, 1.0);
        for (int i = 0; i < nq - 2; ++i) P.z[i + 1] = zi[i];
        for (int i = 0; i < nq; ++i) {
            const double L = jacobi(nq - 1, 0.0, 0.0, P.z[i]);
            P.w[i] = 2.0 / (nq * (nq - 1.0) * L * L);
        }
    } else {
        const double a = (t == PointsType::RadauM_A1B0) ? 1.0 : 2.0, b = 0.0;
        P.z[0] = -1.0;
        auto zi = jacobi_zeros(nq - 1, a, b + 1.0);
        for (int i = 0; i < nq - 1; ++i) P.z[i + 1] = zi[i];
        // w_i = C (1 - z_i) / P_{nq-1}^{a,b}(z_i)^2, w_0 scaled by (b+1).
        const double C = std::pow(2.0, a + b) * std::tgamma(a + nq) * std::tgamma(b + nq) /
                         (std::tgamma((double)nq) * std::tgamma(a + b + nq + 1.0)) / (b + nq);
        for (int i = 0; i < nq; ++i) {
            const double L = jacobi(nq - 1, a, b, P.z[i]);
            P.w[i] = C * (1.0 - P.z[i]) / (L * L);
        }
        P.w[0] *= (b + 1.0);
    }
    derivative_matrix(P.z, P.D);
    return P;
}

namespace {

// Modified_B entry for row p, column q at x (Basis.cpp:499-589 packing).
double modified_b(int p, int q, double x)
{
    const double m = 0.5 * (1.0 - x), pl = 0.5 * (1.0 + x);
    if (p == 0) {
        if (q == 0) return m;
        if (q == 1) return pl;
        return m * pl * jacobi(q - 2, 1.0, 1.0, x);
    }
    if (p == 1) {
        if (q == 0) return m;
        return m * pl * jacobi(q - 1, 1.0, 1.0, x);
    }
    const double mp = std::pow(m, p);
    if (q == 0) return mp;
    return mp * pl * jacobi(q - 1, 2.0 * p - 1.0, 1.0, x);
}

// ModifiedPyr_C entry of pyramid mode (p, q, r) at x (Basis.cpp:646-741):
// Modified_B row max(p, q) when p or q is a vertex index (<= 1), else
// ((1-x)/2)^(p+q-2), times (1+x)/2 P_{r-1}^{(2p+2q-3, 1)}(x) for r >= 1.
double modified_pyr_c(int p, int q, int r, double x)
{
    if (p <= 1 || q <= 1) return modified_b(std::max(p, q), r, x);
    const double mp = std::pow(0.5 * (1.0 - x), p + q - 2);
    if (r == 0) return mp;
    return mp * 0.5 * (1.0 + x) * jacobi(r - 1, 2.0 * p + 2.0 * q - 3.0, 1.0, x);
}

double modified_a(int p, double x)
{
    const double m = 0.5 * (1.0 - x), pl = 0.5 * (1.0 + x);
    if (p == 0) return m;
    if (p == 1) return pl;
    return m * pl * jacobi(p - 2, 1.0, 1.0, x);
}

}  // namespace

Basis1D make_basis(BasisType t, int nm, int nq)
{
    if (nm < 2 || nm > 16) throw std::invalid_argument("make_basis: nm out of range [2,16]");
    Basis1D B;
    B.type = t;
    B.nm = nm;
    const PointsType pt = t == BasisType::ModifiedA   ? PointsType::GLL
                          : t == BasisType::ModifiedB ? PointsType::RadauM_A1B0
                                                      : PointsType::RadauM_A2B0;
    B.pts = make_points(pt, nq);
    const auto &z = B.pts.z;
    if (t == BasisType::ModifiedA) {
        B.nmodes = nm;
        B.bdata.resize((size_t)nm * nq);
        for (int p = 0; p < nm; ++p)
            for (int i = 0; i < nq; ++i) B.bdata[(size_t)p * nq + i] = modified_a(p, z[i]);
    } else if (t == BasisType::ModifiedB) {
        B.nmodes = nm * (nm + 1) / 2;
        B.bdata.resize((size_t)B.nmodes * nq);
        size_t m = 0;
        for (int p = 0; p < nm; ++p)
            for (int q = 0; q < nm - p; ++q, ++m)
                for (int i = 0; i < nq; ++i) B.bdata[m * nq + i] = modified_b(p, q, z[i]);
    } else if (t == BasisType::ModifiedPyrC) {
        // pyramid modes (p, q, r), r < nm - max(p, q), r fastest
        B.nmodes = nm * (nm + 1) * (2 * nm + 1) / 6;
        B.bdata.resize((size_t)B.nmodes * nq);
        size_t m = 0;
        for (int p = 0; p < nm; ++p)
            for (int q = 0; q < nm; ++q)
                for (int r = 0; r < nm - std::max(p, q); ++r, ++m)
                    for (int i = 0; i < nq; ++i) B.bdata[m * nq + i] = modified_pyr_c(p, q, r, z[i]);
    } else {
        B.nmodes = nm * (nm + 1) * (nm + 2) / 6;
        B.bdata.resize((size_t)B.nmodes * nq);
        size_t m = 0;
        for (int p = 0; p < nm; ++p)
            for (int q = 0; q < nm - p; ++q)
                for (int r = 0; r < nm - p - q; ++r, ++m)
                    for (int i = 0; i < nq; ++i) B.bdata[m * nq + i] = modified_b(p + q, r, z[i]);
    }
    // dbdata = D * bdata
    B.dbdata.assign(B.bdata.size(), 0.0);
    for (int m = 0; m < B.nmodes; ++m)
        for (int i = 0; i < nq; ++i) {
            double s = 0.0;
            for (int k = 0; k < nq; ++k) s += B.pts.D[(size_t)k * nq + i] * B.bdata[(size_t)m * nq + k];
            B.dbdata[(size_t)m * nq + i] = s;
        }
    return B;
}

int num_coeffs(Shape s, int nm)
{
    if (s == Shape::Prism) return nm * nm * (nm + 1) / 2;
    if (s == Shape::Pyr) return nm * (nm + 1) * (2 * nm + 1) / 6;
    return s == Shape::Hex ? nm * nm * nm : nm * (nm + 1) * (nm + 2) / 6;
}

// the MODIFIED defaults: hex Q, Q, Q; tet Q, Q-1, Q-1; prism and pyramid
// Q, Q, Q-1
void default_points(Shape s, int nq0, int nq[3])
{
    nq[0] = nq0;
    nq[1] = s == Shape::Tet ? nq0 - 1 : nq0;
    nq[2] = s == Shape::Hex ? nq0 : nq0 - 1;
}

BasisType default_basis(Shape s, int dir)
{
    if (s == Shape::Hex) return BasisType::ModifiedA;
    if (s == Shape::Prism) return dir == 2 ? BasisType::ModifiedB : BasisType::ModifiedA;
    if (s == Shape::Pyr) return dir == 2 ? BasisType::ModifiedPyrC : BasisType::ModifiedA;
    return dir == 0 ? BasisType::ModifiedA : dir == 1 ? BasisType::ModifiedB : BasisType::ModifiedC;
}

PointsType default_points_type(Shape s, int dir)
{
    if (s == Shape::Hex) return PointsType::GLL;
    if (s == Shape::Prism) return dir == 2 ? PointsType::RadauM_A1B0 : PointsType::GLL;
    if (s == Shape::Pyr) return dir == 2 ? PointsType::RadauM_A2B0 : PointsType::GLL;  // MeshGraph.cpp:1754-1770
    return dir == 0 ? PointsType::GLL : dir == 1 ? PointsType::RadauM_A1B0 : PointsType::RadauM_A2B0;
}

double weight_prescale(PointsType t)
{
    return t == PointsType::RadauM_A1B0 ? 0.5 : t == PointsType::RadauM_A2B0 ? 0.25 : 1.0;
}

std::vector<double> interp_matrix(PointsType t, int nin, int nout)
{
    const Points1D pi = make_points(t, nin), po = make_points(t, nout);
    std::vector<double> I((size_t)nin * nout);
    for (int p = 0; p < nin; ++p)
        for (int i = 0; i < nout; ++i) {
            // an output point that is an input point interpolates exactly
            int hit = -1;
            for (int k = 0; k < nin; ++k)
                if (std::fabs(po.z[i] - pi.z[k]) < 1e-15) hit = k;
            double h = 1.0;
            if (hit >= 0) {
                h = hit == p ? 1.0 : 0.0;
            } else {
                for (int k = 0; k < nin; ++k)
                    if (k != p) h *= (po.z[i] - pi.z[k]) / (pi.z[p] - pi.z[k]);
            }
            I[(size_t)p * nout + i] = h;
        }
    return I;
}

void interp_points(Shape s, int nq0, double factor, int out[3])
{
    int nq[3];
    default_points(s, nq0, nq);
    for (int d = 0; d < 3; ++d)
        out[d] = (nq0 - nq[d] == 1) ? (int)(nq0 * factor - 1) : (int)(nq[d] * factor);
}

}  // namespace nektar_amd
