// Host side of the runtime-size operators: tables and index lists.
#include "generic.hpp"

#include <algorithm>
#include <string>

#include "common.hpp"

namespace nektar_amd {

int generic_ncoeffs(Shape s, const int nm[3])
{
    if (s == Shape::Hex) return nm[0] * nm[1] * nm[2];
    int n = 0;
    if (s == Shape::Prism) {
        for (int p = 0; p < nm[0]; ++p) n += nm[1] * (nm[2] - p);
        return n;
    }
    if (s == Shape::Pyr) {
        for (int p = 0; p < nm[0]; ++p)
            for (int q = 0; q < nm[1]; ++q) n += nm[2] - std::max(p, q);
        return n;
    }
    for (int p = 0; p < nm[0]; ++p)
        for (int q = 0; q < nm[1] - p; ++q) n += nm[2] - p - q;
    return n;
}

void generic_check(Shape s, const int nm[3], const int nq[3])
{
    constexpr int kMaxM = 32, kMaxQ = 64;
    for (int d = 0; d < 3; ++d) {
        NMF_REQUIRE(nm[d] >= 2 && nm[d] <= kMaxM, NMF_ERR_UNSUPPORTED,
                    "modes per direction must be 2.." + std::to_string(kMaxM) + " (direction " + std::to_string(d) +
                        " has " + std::to_string(nm[d]) + ")");
        NMF_REQUIRE(nq[d] >= 2 && nq[d] <= kMaxQ, NMF_ERR_UNSUPPORTED,
                    "points per direction must be 2.." + std::to_string(kMaxQ) + " (direction " + std::to_string(d) +
                        " has " + std::to_string(nq[d]) + ")");
    }
    if (s == Shape::Tet)
        NMF_REQUIRE(nm[0] <= nm[1] && nm[1] <= nm[2], NMF_ERR_UNSUPPORTED,
                    "tet modes must satisfy nm0 <= nm1 <= nm2 (the collapsed Modified_A/B/C expansion)");
    if (s == Shape::Prism)
        NMF_REQUIRE(nm[0] <= nm[2], NMF_ERR_UNSUPPORTED,
                    "prism modes must satisfy nm0 <= nm2 (the collapsed Modified_A/A/B expansion)");
    if (s == Shape::Pyr)  // BwdTrans3DWorkspace's precondition (BwdTransKernels.hpp:537-544)
        NMF_REQUIRE(nm[0] == nm[1] && nm[1] == nm[2], NMF_ERR_UNSUPPORTED,
                    "pyramid modes must be equal in the three directions (ModifiedPyr_C is built for nm0 = nm1 = nm2)");
}

GenLayout generic_layout(Shape s, const int nm[3], const int nq[3])
{
    generic_check(s, nm, nq);
    GenLayout L;
    Basis1D b[3];
    for (int d = 0; d < 3; ++d) b[d] = make_basis(default_basis(s, d), nm[d], nq[d]);
    for (int d = 0; d < 3; ++d) {
        L.oB[d] = (int)L.tab.size();
        L.tab.insert(L.tab.end(), b[d].bdata.begin(), b[d].bdata.end());
    }
    for (int d = 0; d < 3; ++d) {
        L.oD[d] = (int)L.tab.size();
        L.tab.insert(L.tab.end(), b[d].pts.D.begin(), b[d].pts.D.end());
    }
    for (int d = 0; d < 3; ++d) {
        L.oW[d] = (int)L.tab.size();
        const double f = weight_prescale(b[d].pts.type);
        for (double w : b[d].pts.w) L.tab.push_back(f * w);
    }
    // tet half-space factors h0 = (1+z0)/2, h1 = (1+z1)/2, h2 = 2/(1-z1),
    // h3 = 2/(1-z2) (HelmholtzKernels.hpp:1368-1389); prisms use h0 and h3
    // (the reference's prism h0, h1, :1390-1403), pyramids h0, h1 and h3
    // (the reference's pyramid h0, h1, h2, :1405-1423); unused for hexes
    L.oH[0] = (int)L.tab.size();
    for (double z : b[0].pts.z) L.tab.push_back(0.5 * (1.0 + z));
    L.oH[1] = (int)L.tab.size();
    for (double z : b[1].pts.z) L.tab.push_back(0.5 * (1.0 + z));
    L.oH[2] = (int)L.tab.size();
    for (double z : b[1].pts.z) L.tab.push_back(s == Shape::Tet ? 2.0 / (1.0 - z) : 0.0);
    L.oH[3] = (int)L.tab.size();
    for (double z : b[2].pts.z) L.tab.push_back(s != Shape::Hex ? 2.0 / (1.0 - z) : 0.0);

    const int m0 = nm[0], m1 = nm[1], m2 = nm[2], n0 = nq[0], n1 = nq[1], n2 = nq[2];
    L.nqtot = n0 * n1 * n2;
    L.ncoeff = generic_ncoeffs(s, nm);
    size_t buf = (size_t)L.nqtot;
    if (s == Shape::Hex) {
        buf = std::max({buf, (size_t)m2 * m1 * n0, (size_t)m2 * n1 * n0});
    } else if (s == Shape::Prism) {
        // columns (p, q) in coefficient order (p-major): first coefficient,
        // first Modified_B mode of (p, r = 0), r count nm2 - p, p; per
        // coefficient its p, column and Modified_B mode (p, r)
        std::vector<int> cP, cPair, cM2, pC0, pM20, pNR, pP;
        int c = 0, col = 0, mode_pr = 0;
        for (int p = 0; p < m0; ++p) {
            for (int q = 0; q < m1; ++q, ++col) {
                pC0.push_back(c);
                pM20.push_back(mode_pr);
                pNR.push_back(m2 - p);
                pP.push_back(p);
                for (int r = 0; r < m2 - p; ++r, ++c) {
                    cP.push_back(p);
                    cPair.push_back(col);
                    cM2.push_back(mode_pr + r);
                }
            }
            mode_pr += m2 - p;
        }
        L.npq = col;
        for (auto *v : {&cP, &cPair, &cM2, &pC0, &pM20, &pNR, &pP}) L.itab.insert(L.itab.end(), v->begin(), v->end());
        buf = std::max({buf, (size_t)L.npq * n2, (size_t)m0 * n2 * n1});
    } else if (s == Shape::Pyr) {
        // columns (p, q), p-major, each with its first coefficient, first
        // ModifiedPyr_C mode (the coefficient index: the table runs in
        // coefficient order), r count nm2 - max(p, q) and p; per coefficient
        // its p, column and ModifiedPyr_C mode
        std::vector<int> cP, cPair, cM2, pC0, pM20, pNR, pP;
        int c = 0, col = 0;
        for (int p = 0; p < m0; ++p)
            for (int q = 0; q < m1; ++q, ++col) {
                const int nr = m2 - std::max(p, q);
                pC0.push_back(c);
                pM20.push_back(c);
                pNR.push_back(nr);
                pP.push_back(p);
                for (int r = 0; r < nr; ++r, ++c) {
                    cP.push_back(p);
                    cPair.push_back(col);
                    cM2.push_back(c);
                }
            }
        L.npq = col;
        for (auto *v : {&cP, &cPair, &cM2, &pC0, &pM20, &pNR, &pP}) L.itab.insert(L.itab.end(), v->begin(), v->end());
        buf = std::max({buf, (size_t)L.npq * n2, (size_t)m0 * n2 * n1, (size_t)n2});
    } else {
        // coefficient lists: p, (p,q) pair and Modified_C mode of each
        // coefficient; pair lists: first coefficient, first mode, r count, p
        std::vector<int> cP, cPair, cM2, pC0, pM20, pNR, pP;
        int c = 0, mode2 = 0, pair = 0;
        for (int p = 0; p < m0; ++p) {
            for (int q = 0; q < m1 - p; ++q, ++pair) {
                pC0.push_back(c);
                pM20.push_back(mode2);
                pNR.push_back(m2 - p - q);
                pP.push_back(p);
                for (int r = 0; r < m2 - p - q; ++r, ++c, ++mode2) {
                    cP.push_back(p);
                    cPair.push_back(pair);
                    cM2.push_back(mode2);
                }
            }
            // Modified_C modes of the q the coefficients skip (nm1 - p <= q < nm2 - p)
            for (int q = m1 - p; q < m2 - p; ++q) mode2 += m2 - p - q;
        }
        L.npq = pair;
        for (auto *v : {&cP, &cPair, &cM2, &pC0, &pM20, &pNR, &pP}) L.itab.insert(L.itab.end(), v->begin(), v->end());
        buf = std::max({buf, (size_t)L.npq * n2, (size_t)m0 * n2 * n1, (size_t)2 * n2});
    }
    NMF_REQUIRE(buf < (size_t(1) << 28), NMF_ERR_UNSUPPORTED, "element too large for the runtime-size kernels");
    L.buf = (int)buf;
    return L;
}

}  // namespace nektar_amd
