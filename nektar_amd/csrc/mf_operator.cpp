// MatrixFree operator object: key parsing, 1-D tables, geometry packing and
// launch.  Reference behaviour: MatrixFree::Helper set-up
// (Operator.hpp:529-641), Collections::Helmholtz_MatrixFree::UpdateFactors
// (Collections/Helmholtz.cpp:615-691), CoalescedGeomData layouts
// (Collections/CoalescedGeomData.cpp:90-342).
#include "mf_operator.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace nektar_amd {

void parse_operator_key(const std::string &key, OpType &op, Shape &shape, bool &deformed)
{
    auto p1 = key.find('_');
    auto p2 = key.find('_', p1 == std::string::npos ? 0 : p1 + 1);
    NMF_REQUIRE(p1 != std::string::npos && p2 != std::string::npos, NMF_ERR_INVALID,
                "operator key '" + key + "' is not <Op>_<Shape>_<Regular|Deformed>");
    const std::string o = key.substr(0, p1), s = key.substr(p1 + 1, p2 - p1 - 1), g = key.substr(p2 + 1);
    if (o == "BwdTrans") op = OpType::BwdTrans;
    else if (o == "IProduct") op = OpType::IProduct;
    else if (o == "PhysDeriv") op = OpType::PhysDeriv;
    else if (o == "Helmholtz") op = OpType::Helmholtz;
    else if (o == "IProductWRTDerivBase") op = OpType::IProductDeriv;
    else if (o == "LinearAdvectionDiffusionReaction") op = OpType::AdvDiffReact;
    else if (o == "PhysInterp1DScaled") op = OpType::PhysInterp;
    else
        throw Error(NMF_ERR_UNSUPPORTED, "operator '" + o +
                                             "' is not provided (BwdTrans, IProduct, IProductWRTDerivBase, PhysDeriv, "
                                             "Helmholtz, LinearAdvectionDiffusionReaction, PhysInterp1DScaled)");
    if (s == "Hex") shape = Shape::Hex;
    else if (s == "Tet") shape = Shape::Tet;
    else if (s == "Prism") shape = Shape::Prism;
    else if (s == "Pyr") shape = Shape::Pyr;
    else throw Error(NMF_ERR_UNSUPPORTED, "shape '" + s + "' is not provided (Hex, Tet, Prism, Pyr)");
    if (g == "Regular") deformed = false;
    else if (g == "Deformed") deformed = true;
    else throw Error(NMF_ERR_INVALID, "geometry tag '" + g + "' must be Regular or Deformed");
}

std::vector<double> MFOperator::build_tables(Shape shape, int nm, int nq0)
{
    std::vector<double> t;
    if (shape == Shape::Hex) {
        Basis1D b = make_basis(BasisType::ModifiedA, nm, nq0);
        const int nq = nq0;
        t.insert(t.end(), b.bdata.begin(), b.bdata.end());  // B[p*nq + i]
        t.insert(t.end(), b.pts.D.begin(), b.pts.D.end());  // D[l*nq + i]
        t.insert(t.end(), b.pts.w.begin(), b.pts.w.end());
        // transposed copies, so that every contraction reads one
        // contiguous row per output: BT[i*nm + p] = B[p*nq + i],
        // DT[i*nq + l] = D[l*nq + i]
        for (int i = 0; i < nq; ++i)
            for (int p = 0; p < nm; ++p) t.push_back(b.bdata[p * nq + i]);
        for (int i = 0; i < nq; ++i)
            for (int l = 0; l < nq; ++l) t.push_back(b.pts.D[l * nq + i]);
        return t;
    }
    Basis1D b0 = make_basis(BasisType::ModifiedA, nm, nq0);
    Basis1D b1 = make_basis(BasisType::ModifiedB, nm, nq0 - 1);
    Basis1D b2 = make_basis(BasisType::ModifiedC, nm, nq0 - 1);
    for (auto *b : {&b0, &b1, &b2}) t.insert(t.end(), b->bdata.begin(), b->bdata.end());
    for (auto *b : {&b0, &b1, &b2}) t.insert(t.end(), b->pts.D.begin(), b->pts.D.end());
    for (auto *b : {&b0, &b1, &b2}) {
        const double f = weight_prescale(b->pts.type);
        for (double w : b->pts.w) t.push_back(f * w);
    }
    for (double z : b0.pts.z) t.push_back(0.5 * (1.0 + z));
    for (double z : b1.pts.z) t.push_back(0.5 * (1.0 + z));
    for (double z : b1.pts.z) t.push_back(2.0 / (1.0 - z));
    for (double z : b2.pts.z) t.push_back(2.0 / (1.0 - z));
    // transposed copies, so that every contraction reads one contiguous row
    // per output: B0T[i*nm + p] = B0[p*nq0 + i], DdT[i*nqd + l] = Dd[l*nqd + i]
    for (int i = 0; i < nq0; ++i)
        for (int p = 0; p < nm; ++p) t.push_back(b0.bdata[p * nq0 + i]);
    for (auto *b : {&b0, &b1, &b2}) {
        const int n = (int)b->pts.w.size();
        for (int i = 0; i < n; ++i)
            for (int l = 0; l < n; ++l) t.push_back(b->pts.D[l * n + i]);
    }
    // zero-padded B1 and B2 rows that the kernels stage into LDS: from a
    // multiple of 4, rows of pitch rp = nq1 rounded up to 4, each table
    // followed by nm zero rows (tet::Dims::oB1P, RP, NB1R, NB2R)
    const int nq1 = nq0 - 1, rp = (nq1 + 3) / 4 * 4;
    t.resize((t.size() + 3) / 4 * 4, 0.0);
    auto padded_rows = [&](const Basis1D &b, int nmodes) {
        for (int m = 0; m < nmodes + nm; ++m)
            for (int k = 0; k < rp; ++k) t.push_back(m < nmodes && k < nq1 ? b.bdata[m * nq1 + k] : 0.0);
    };
    padded_rows(b1, nm * (nm + 1) / 2);
    padded_rows(b2, nm * (nm + 1) * (nm + 2) / 6);
    return t;
}

MFOperator::MFOperator(const std::string &key, int nummodes, int numpoints0, int n, int precision)
{
    Shape s;
    OpType o;
    bool d;
    parse_operator_key(key, o, s, d);
    const int m[3] = {nummodes, nummodes, nummodes};
    int q[3];
    default_points(s, numpoints0, q);
    init(key, m, q, n, precision);
}

MFOperator::MFOperator(const std::string &key, const int nummodes[3], const int numpoints[3], int n, int precision)
{
    NMF_REQUIRE(nummodes && numpoints, NMF_ERR_INVALID, "null modes/points array");
    init(key, nummodes, numpoints, n, precision);
}

void MFOperator::init(const std::string &key, const int nummodes[3], const int numpoints[3], int n, int precision)
{
    parse_operator_key(key, op, shape, deformed);
    NMF_REQUIRE(precision == NMF_FP64 || precision == NMF_FP32, NMF_ERR_INVALID, "precision must be NMF_FP64 or NMF_FP32");
    NMF_REQUIRE(n >= 0, NMF_ERR_INVALID, "negative element count");
    fp32 = precision == NMF_FP32;
    nelmt = n;
    for (int d = 0; d < 3; ++d) {
        nmd[d] = nummodes[d];
        nq[d] = numpoints[d];
    }
    nm = nmd[0];
    nqtot = nq[0] * nq[1] * nq[2];
    // a compile-time kernel exists for equal modes in every direction and
    // the MODIFIED default points of an instantiated (modes, points) pair
    int dq[3];
    default_points(shape, nq[0], dq);
    const bool iso = nmd[1] == nm && nmd[2] == nm && nq[1] == dq[1] && nq[2] == dq[2];
    // (prisms and pyramids always run the runtime-size kernels)
    const bool compiled = iso && (shape == Shape::Hex   ? hex_op_supported(nm, nq[0])
                                  : shape == Shape::Tet ? tet_supported(nm, nq[0])
                                                        : false);
    if (compiled) {
        ncoeff = num_coeffs(shape, nm);
        tab.upload(build_tables(shape, nm, nq[0]), fp32);
    } else {
        NMF_REQUIRE(op != OpType::PhysInterp, NMF_ERR_UNSUPPORTED,
                    "PhysInterp1DScaled takes equal modes and the default points of the shape only");
        generic = true;
        glay = generic_layout(shape, nmd, nq);  // checks the sizes
        ncoeff = glay.ncoeff;
        tab.upload(glay.tab, fp32);
        if (!glay.itab.empty()) {
            gitab.allocate(glay.itab.size() * sizeof(int));
            NMF_HIP(hipMemcpy(gitab.get(), glay.itab.data(), glay.itab.size() * sizeof(int), hipMemcpyHostToDevice));
        }
    }
    if (op == OpType::BwdTrans || op == OpType::PhysInterp) deformed = false;  // no geometry
    if (op == OpType::PhysInterp) set_scaling_factor(1.5);  // the reference's default (Operator.hpp:136-141)
    if (!generic) tile = tuned_tile(tune_key());
}

void MFOperator::prepare()
{
    ensure_geometry();
    if (generic) (void)generic_args(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
}

GenArgs MFOperator::generic_args(const void *in0, const void *in1, const void *in2, void *o0, void *o1, void *o2,
                                 hipStream_t s)
{
    GenArgs a{};
    a.op = op;
    a.shape = shape;
    a.fp32 = fp32;
    a.deformed = deformed;
    a.nelmt = nelmt;
    for (int d = 0; d < 3; ++d) {
        a.nm[d] = nmd[d];
        a.nq[d] = nq[d];
        a.oB[d] = glay.oB[d];
        a.oD[d] = glay.oD[d];
        a.oW[d] = glay.oW[d];
    }
    for (int c = 0; c < 4; ++c) a.oH[c] = glay.oH[c];
    a.nqtot = nqtot;
    a.ncoeff = ncoeff;
    a.npq = glay.npq;
    a.buf = glay.buf;
    a.tab = tab.get();
    a.itab = gitab.as<int>();
    a.geo = geo.get();
    a.adv = adv.get();
    a.lambda = lambda;
    a.in0 = in0;
    a.in1 = in1;
    a.in2 = in2;
    a.out0 = o0;
    a.out1 = o1;
    a.out2 = o2;
    a.stream = s;
    a.groups = std::min(nelmt, kGenGroups);
    if (lds_bytes(a) > kGenMaxLds && nelmt > 0) {
        // the element's buffers live in a global workspace, one slice per
        // workgroup of the walking grid (allocated once, kept)
        gws.allocate((size_t)a.groups * 4 * a.buf * value_bytes());
        a.ws = gws.get();
    }
    return a;
}

void MFOperator::set_scaling_factor(double f)
{
    NMF_REQUIRE(op == OpType::PhysInterp, NMF_ERR_INVALID, "SetScalingFactor: not a PhysInterp1DScaled operator");
    NMF_REQUIRE(std::isfinite(f) && f > 0.0, NMF_ERR_INVALID, "SetScalingFactor: factor must be positive");
    int no[3];
    interp_points(shape, nq[0], f, no);
    for (int d = 0; d < 3; ++d)
        NMF_REQUIRE(no[d] >= 1, NMF_ERR_INVALID, "SetScalingFactor: factor leaves no points in a direction");
    NMF_REQUIRE(interp_lds_bytes(nq, no, fp32) <= kInterpMaxLds, NMF_ERR_UNSUPPORTED,
                "SetScalingFactor: " + std::to_string(no[0]) + "x" + std::to_string(no[1]) + "x" +
                    std::to_string(no[2]) + " output points exceed the 64 KiB workgroup staging of the kernel");
    // transposed for the kernel: IT[i*nin + p] = I[p*nout + i], the weights
    // of one output contiguous
    std::vector<double> t;
    for (int d = 0; d < 3; ++d) {
        const auto I = interp_matrix(default_points_type(shape, d), nq[d], no[d]);
        for (int i = 0; i < no[d]; ++i)
            for (int p = 0; p < nq[d]; ++p) t.push_back(I[(size_t)p * no[d] + i]);
    }
    t.resize(t.size() + 16, 0.0);  // the kernel reads whole rows of up to 12 values
    itab.upload(t, fp32);
    factor = f;
    for (int d = 0; d < 3; ++d) nqo[d] = no[d];
    ++version;
}

size_t MFOperator::in_per_elmt() const
{
    const bool coeffs = op == OpType::BwdTrans || op == OpType::Helmholtz || op == OpType::AdvDiffReact;
    return coeffs ? (size_t)ncoeff : (size_t)nqtot;
}

size_t MFOperator::out_per_elmt() const
{
    if (op == OpType::PhysInterp) return (size_t)nqo[0] * nqo[1] * nqo[2];
    const bool coeffs = op == OpType::IProduct || op == OpType::Helmholtz || op == OpType::IProductDeriv ||
                        op == OpType::AdvDiffReact;
    return coeffs ? (size_t)ncoeff : (size_t)nqtot;
}

void MFOperator::set_jac(const double *j)
{
    NMF_REQUIRE(j || nelmt == 0, NMF_ERR_INVALID, "SetJac: null pointer");
    jac.assign(j, j + geom_points());
    have_jac = true;
    geo_dirty = true;
    ++version;
}

void MFOperator::set_df(const double *d)
{
    NMF_REQUIRE(d || nelmt == 0, NMF_ERR_INVALID, "SetDF: null pointer");
    df.assign(d, d + 9 * geom_points());
    have_df = true;
    geo_dirty = true;
    adv_dirty = true;
    ++version;
}

void MFOperator::set_advection(const double *v)
{
    NMF_REQUIRE(op == OpType::AdvDiffReact, NMF_ERR_INVALID,
                "advection velocities belong to LinearAdvectionDiffusionReaction operators");
    NMF_REQUIRE(v || nelmt == 0, NMF_ERR_INVALID, "SetAdvectionVelocities: null pointer");
    vel.assign(v, v + (size_t)nelmt * nqtot * 3);
    have_vel = true;
    adv_dirty = true;
    ++version;
}

void MFOperator::set_diffusion(const double *d6)
{
    if (!d6) {
        has_diff = false;
        const double id[6] = {1, 0, 1, 0, 0, 1};
        std::memcpy(diff, id, sizeof(diff));
    } else {
        has_diff = true;
        std::memcpy(diff, d6, sizeof(diff));
    }
    geo_dirty = true;
    ++version;
}

namespace {

// Lower Cholesky factor of the symmetric tensor {d00,d01,d11,d02,d12,d22}.
void cholesky3(const double *d6, double L[3][3])
{
    const double K[3][3] = {{d6[0], d6[1], d6[3]}, {d6[1], d6[2], d6[4]}, {d6[3], d6[4], d6[5]}};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) L[i][j] = 0.0;
    for (int j = 0; j < 3; ++j) {
        double s = K[j][j];
        for (int k = 0; k < j; ++k) s -= L[j][k] * L[j][k];
        NMF_REQUIRE(s > 0.0, NMF_ERR_INVALID, "SetConstVarDiffusion: tensor is not positive definite");
        L[j][j] = std::sqrt(s);
        for (int i = j + 1; i < 3; ++i) {
            double t = K[i][j];
            for (int k = 0; k < j; ++k) t -= L[i][k] * L[j][k];
            L[i][j] = t / L[j][j];
        }
    }
}

// df'[a*3+d] = sum_j L[j][a] df[j*3+d], so that df'^T df' = df^T K df.
void apply_diffusion(const double L[3][3], const double *f, double *out)
{
    for (int a = 0; a < 3; ++a)
        for (int d = 0; d < 3; ++d) {
            double s = 0.0;
            for (int j = 0; j < 3; ++j) s += L[j][a] * f[j * 3 + d];
            out[a * 3 + d] = s;
        }
}

// metric m = {00, 01, 02, 11, 12, 22}, m_de = sum_j f[j*3+d] f[j*3+e]
// (HelmholtzKernels.hpp:440-462)
void metric6(const double *f, double *m)
{
    const int idx[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    for (int c = 0; c < 6; ++c) {
        const int d = idx[c][0], e = idx[c][1];
        m[c] = f[d] * f[e] + f[3 + d] * f[3 + e] + f[6 + d] * f[6 + e];
    }
}

}  // namespace

std::vector<double> MFOperator::pack_geometry() const
{
    const size_t gp = geom_points();
    std::vector<double> g;
    if (op == OpType::IProduct) return jac;
    if (op == OpType::PhysDeriv) {
        if (!deformed) return df;
        g.resize(df.size());
        for (int e = 0; e < nelmt; ++e)
            for (int q = 0; q < nqtot; ++q)
                for (int d = 0; d < 9; ++d)
                    g[((size_t)e * 9 + d) * nqtot + q] = df[((size_t)e * nqtot + q) * 9 + d];
        return g;
    }
    if (op == OpType::IProductDeriv && shape == Shape::Hex) {
        // J df per element (weights applied in the kernel), or W J df per point
        g.resize(gp * 9);
        if (!deformed) {
            for (size_t e = 0; e < gp; ++e)
                for (int c = 0; c < 9; ++c) g[e * 9 + c] = jac[e] * df[e * 9 + c];
        } else {
            const Points1D P0 = make_points(PointsType::GLL, nq[0]), P1 = make_points(PointsType::GLL, nq[1]),
                           P2 = make_points(PointsType::GLL, nq[2]);
            for (int e = 0; e < nelmt; ++e)
                for (int k = 0, q = 0; k < nq[2]; ++k)
                    for (int j = 0; j < nq[1]; ++j)
                        for (int i = 0; i < nq[0]; ++i, ++q) {
                            const size_t pt = (size_t)e * nqtot + q;
                            const double wj = P0.w[i] * P1.w[j] * P2.w[k] * jac[pt];
                            for (int c = 0; c < 9; ++c) g[((size_t)e * 9 + c) * nqtot + q] = wj * df[pt * 9 + c];
                        }
        }
        return g;
    }
    // Helmholtz (and the tet IProductWRTDerivBase: df and J, as for Helmholtz)
    double L[3][3];
    if (has_diff) cholesky3(diff, L);
    auto fetch_df = [&](size_t pt, double *f) {
        if (has_diff) apply_diffusion(L, &df[pt * 9], f);
        else std::memcpy(f, &df[pt * 9], 9 * sizeof(double));
    };
    if (shape == Shape::Hex) {
        if (!deformed) {
            g.resize(gp * 7);
            for (size_t e = 0; e < gp; ++e) {
                double f[9], m[6];
                fetch_df(e, f);
                metric6(f, m);
                for (int c = 0; c < 6; ++c) g[e * 7 + c] = jac[e] * m[c];
                g[e * 7 + 6] = jac[e];
            }
        } else {
            const Points1D P0 = make_points(PointsType::GLL, nq[0]), P1 = make_points(PointsType::GLL, nq[1]),
                           P2 = make_points(PointsType::GLL, nq[2]);
            g.resize(gp * 7);
            for (int e = 0; e < nelmt; ++e)
                for (int k = 0, q = 0; k < nq[2]; ++k)
                    for (int j = 0; j < nq[1]; ++j)
                        for (int i = 0; i < nq[0]; ++i, ++q) {
                            const size_t pt = (size_t)e * nqtot + q;
                            double f[9], m[6];
                            fetch_df(pt, f);
                            metric6(f, m);
                            const double wj = P0.w[i] * P1.w[j] * P2.w[k] * jac[pt];
                            for (int c = 0; c < 6; ++c) g[((size_t)e * 7 + c) * nqtot + q] = wj * m[c];
                            g[((size_t)e * 7 + 6) * nqtot + q] = wj;
                        }
        }
        return g;
    }
    // tet: df (possibly diffusion-transformed) and J, per element or point
    g.resize(gp * 10);
    if (!deformed) {
        for (size_t e = 0; e < gp; ++e) {
            fetch_df(e, &g[e * 10]);
            g[e * 10 + 9] = jac[e];
        }
    } else {
        for (int e = 0; e < nelmt; ++e)
            for (int q = 0; q < nqtot; ++q) {
                const size_t pt = (size_t)e * nqtot + q;
                double f[9];
                fetch_df(pt, f);
                for (int c = 0; c < 9; ++c) g[((size_t)e * 10 + c) * nqtot + q] = f[c];
                g[((size_t)e * 10 + 9) * nqtot + q] = jac[pt];
            }
    }
    return g;
}

// Advection in the reference directions, c_d = sum_j V_j df[j*3+d]
// (LinearAdvectionDiffusionReactionKernels.hpp:193-277); for tets the
// collapsed-coordinate factors of the derivatives are folded in
// (:279-386): c_0 = h2 h3 (a_0 + h0 (a_1 + a_2)), c_1 = h3 (a_1 + h1 a_2),
// c_2 = a_2.  Uses the geometric df (the diffusion tensor acts on the
// Laplacian only).
std::vector<double> MFOperator::pack_advection() const
{
    std::vector<double> c((size_t)nelmt * 3 * nqtot);
    std::vector<double> z0, z1, z2;
    if (shape != Shape::Hex) {
        z0 = make_points(default_points_type(shape, 0), nq[0]).z;
        z1 = make_points(default_points_type(shape, 1), nq[1]).z;
        z2 = make_points(default_points_type(shape, 2), nq[2]).z;
    }
    for (int e = 0; e < nelmt; ++e)
        for (int k = 0, q = 0; k < nq[2]; ++k)
            for (int j = 0; j < nq[1]; ++j)
                for (int i = 0; i < nq[0]; ++i, ++q) {
                    const size_t pt = (size_t)e * nqtot + q;
                    const double *f = &df[(deformed ? pt : (size_t)e) * 9];
                    const double *v = &vel[pt * 3];
                    double a[3];
                    for (int d = 0; d < 3; ++d) a[d] = v[0] * f[d] + v[1] * f[3 + d] + v[2] * f[6 + d];
                    if (shape == Shape::Tet) {
                        const double h0 = 0.5 * (1.0 + z0[i]), h1 = 0.5 * (1.0 + z1[j]);
                        const double h2 = 2.0 / (1.0 - z1[j]), h3 = 2.0 / (1.0 - z2[k]);
                        const double a0 = a[0], a1 = a[1], a2 = a[2];
                        a[0] = h2 * h3 * (a0 + h0 * (a1 + a2));
                        a[1] = h3 * (a1 + h1 * a2);
                        a[2] = a2;
                    } else if (shape == Shape::Pyr) {
                        // AdvectionPyrKernel (LinearAdvectionDiffusionReactionKernels.hpp:
                        // 475-560): d/dxi0 = h2 d0, d/dxi1 = h2 d1, d/dxi2 =
                        // h0 h2 d0 + h1 h2 d1 + d2 with h0 = (1+z0)/2,
                        // h1 = (1+z1)/2, h2 = 2/(1-z2)
                        const double h0 = 0.5 * (1.0 + z0[i]), h1 = 0.5 * (1.0 + z1[j]), h2 = 2.0 / (1.0 - z2[k]);
                        a[0] = h2 * (a[0] + h0 * a[2]);
                        a[1] = h2 * (a[1] + h1 * a[2]);
                    } else if (shape == Shape::Prism) {
                        // AdvectionPrismKernel (LinearAdvectionDiffusionReactionKernels.hpp:
                        // 384-470): d/dxi0 = h1 d0, d/dxi2 = h0 h1 d0 + d2 with
                        // h0 = (1+z0)/2, h1 = 2/(1-z2)
                        const double h0 = 0.5 * (1.0 + z0[i]), h1 = 2.0 / (1.0 - z2[k]);
                        a[0] = h1 * (a[0] + h0 * a[2]);
                    }
                    for (int d = 0; d < 3; ++d) c[((size_t)e * 3 + d) * nqtot + q] = a[d];
                }
    return c;
}

void MFOperator::ensure_geometry()
{
    if (needs_jac()) NMF_REQUIRE(have_jac, NMF_ERR_STATE, "Jacobian not set (SetJac) before apply");
    if (needs_df()) NMF_REQUIRE(have_df, NMF_ERR_STATE, "derivative factors not set (SetDF) before apply");
    if (op == OpType::AdvDiffReact) {
        NMF_REQUIRE(have_vel, NMF_ERR_STATE, "advection velocities not set (SetAdvectionVelocities) before apply");
        if (adv_dirty) {
            adv.upload(pack_advection(), fp32);
            adv_dirty = false;
        }
    }
    if (!geo_dirty || op == OpType::BwdTrans || op == OpType::PhysInterp) return;
    geo.upload(pack_geometry(), fp32);
    geo_dirty = false;
}

void MFOperator::apply_iproduct_deriv(const void *in0, const void *in1, const void *in2, void *out, hipStream_t s)
{
    NMF_REQUIRE(op == OpType::IProductDeriv, NMF_ERR_INVALID, "not an IProductWRTDerivBase operator");
    NMF_REQUIRE(nelmt == 0 || (in0 && in1 && in2 && out), NMF_ERR_INVALID, "null input/output pointer");
    ensure_geometry();
    if (generic) {
        const hipError_t e = launch_generic(generic_args(in0, in1, in2, out, nullptr, nullptr, s));
        NMF_REQUIRE(e == hipSuccess, NMF_ERR_DEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
        return;
    }
    KernelArgs a{op, fp32, deformed, nelmt, tab.get(), geo.get(), lambda, in0, out, nullptr, nullptr, s, tile, in1, in2};
    hipError_t e = shape == Shape::Hex ? launch_hex(nm, nq[0], a) : launch_tet(nm, nq[0], a);
    NMF_REQUIRE(e == hipSuccess, NMF_ERR_DEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
}

void MFOperator::apply(const void *in, void *out0, void *out1, void *out2, hipStream_t s)
{
    NMF_REQUIRE(op != OpType::IProductDeriv, NMF_ERR_INVALID, "IProductWRTDerivBase has three inputs");
    NMF_REQUIRE(nelmt == 0 || (in && out0), NMF_ERR_INVALID, "null input/output pointer");
    if (op == OpType::PhysDeriv) NMF_REQUIRE(nelmt == 0 || (out1 && out2), NMF_ERR_INVALID, "PhysDeriv needs three outputs");
    ensure_geometry();
    if (op == OpType::PhysInterp) {
        const InterpArgs ia{fp32, nelmt, {nq[0], nq[1], nq[2]}, {nqo[0], nqo[1], nqo[2]}, itab.get(), in, out0, s};
        const hipError_t e = launch_phys_interp(ia);
        NMF_REQUIRE(e == hipSuccess, NMF_ERR_DEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
        return;
    }
    if (generic) {
        const hipError_t e = launch_generic(generic_args(in, nullptr, nullptr, out0, out1, out2, s));
        NMF_REQUIRE(e == hipSuccess, NMF_ERR_DEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
        return;
    }
    KernelArgs a{op, fp32, deformed, nelmt, tab.get(), geo.get(), lambda, in, out0, out1, out2, s, tile};
    a.adv = adv.get();
    hipError_t e = shape == Shape::Hex ? launch_hex(nm, nq[0], a) : launch_tet(nm, nq[0], a);
    NMF_REQUIRE(e == hipSuccess, NMF_ERR_DEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
}

void MFOperator::apply_physderiv_dir(int dir, const void *in, void *out, void *wsp, hipStream_t s)
{
    NMF_REQUIRE(op == OpType::PhysDeriv, NMF_ERR_INVALID, "not a PhysDeriv operator");
    NMF_REQUIRE(dir >= 0 && dir < 3, NMF_ERR_INVALID, "PhysDeriv direction must be 0, 1 or 2");
    NMF_REQUIRE(nelmt == 0 || wsp, NMF_ERR_INVALID, "null workspace");
    void *o[3];
    char *w = static_cast<char *>(wsp);
    const size_t half = (size_t)nelmt * nqtot * value_bytes();
    for (int d = 0, k = 0; d < 3; ++d) o[d] = d == dir ? out : w + half * (k++);
    apply(in, o[0], o[1], o[2], s);
}

void MFOperator::apply_mapped(const void *xg, const int *menc, void *out, hipStream_t s)
{
    NMF_REQUIRE(can_apply_mapped(), NMF_ERR_INVALID, "mapped apply: FP64 hex Helmholtz operators only");
    NMF_REQUIRE(nelmt == 0 || (xg && menc && out), NMF_ERR_INVALID, "null input/output pointer");
    ensure_geometry();
    KernelArgs a{op, fp32, deformed, nelmt, tab.get(), geo.get(), lambda, xg, out, nullptr, nullptr, s, tile};
    a.menc = menc;
    const hipError_t e = launch_hex(nm, nq[0], a);
    NMF_REQUIRE(e == hipSuccess, NMF_ERR_DEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
}

TuneKey MFOperator::tune_key() const
{
    return TuneKey{(int)shape, (int)op, nm, nq[0], fp32 ? 1 : 0, deformed ? 1 : 0, 0};
}

std::vector<int> MFOperator::tile_candidates() const
{
    std::vector<int> c;
    if (generic || op == OpType::PhysInterp) return c;
    int buf[8], n = 0;
    if (shape == Shape::Hex && op == OpType::Helmholtz) n = hex_tile_candidates(nm, nq[0], fp32, buf, 8);
    if (shape == Shape::Tet) n = tet_tile_candidates(nq[0], buf, 8);
    c.assign(buf, buf + n);
    return c;
}

void MFOperator::set_tile(int t)
{
    if (t != 0) {
        const auto c = tile_candidates();
        bool ok = false;
        for (int x : c) ok |= x == t;
        NMF_REQUIRE(ok, NMF_ERR_INVALID, "tile " + std::to_string(t) + " is not instantiated for this operator");
    }
    tile = t;
    ++version;
}

int MFOperator::autotune(hipStream_t s, std::vector<double> *times_us)
{
    const auto cand = tile_candidates();
    if (times_us) times_us->clear();
    if (cand.empty() || nelmt == 0) return 0;
    ensure_geometry();
    if (op == OpType::AdvDiffReact && !adv.get()) return 0;  // no advection velocities yet
    const size_t es = fp32 ? sizeof(float) : sizeof(double);
    // three inputs and three outputs of the larger of the coefficient and
    // point counts cover every operator's arguments
    const size_t vals = (size_t)nelmt * std::max(ncoeff, nqtot);
    DeviceBuffer in(3 * vals * es), out(3 * vals * es);
    // every byte 0x3F: all values ~4.8e-4 (FP64) / ~0.75 (FP32), normal and
    // non-zero, so the timed arithmetic draws the power of real data
    NMF_HIP(hipMemsetAsync(in.get(), 0x3F, 3 * vals * es, s));
    char *ib = static_cast<char *>(in.get()), *ob = static_cast<char *>(out.get());
    KernelArgs a{op,       fp32, deformed,          nelmt, tab.get(), geo.get(), lambda, ib, ob, ob + vals * es,
                 ob + 2 * vals * es, s, 0, ib + vals * es, ib + 2 * vals * es};
    a.adv = adv.get();
    const int best = time_tiles(
        s, cand,
        [&](int t) {
            a.tile = t;
            return shape == Shape::Hex ? launch_hex(nm, nq[0], a) : launch_tet(nm, nq[0], a);
        },
        10, times_us);
    tile = best;
    ++version;
    record_tile(tune_key(), best);
    return best;
}

}  // namespace nektar_amd
