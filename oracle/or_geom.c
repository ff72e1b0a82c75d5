/*
 * or_geom.c -- geometric factors, structured meshes and C0 assembly.
 *
 * TEST INFRASTRUCTURE ONLY (parity oracle).
 *
 *   hex/tet factors : SpatialDomains/GeomFactors.cpp:150-477 (ComputeDeriv,
 *                     ComputeJac, ComputeDerivFactors, Adjoint :709-752)
 *                     for straight-sided (trilinear / affine) maps, whose
 *                     derivatives the reference evaluates exactly;
 *   regularity test : SpatialDomains/HexGeom.cpp:83-153 (iso-parameters);
 *   assembly        : MultiRegions/AssemblyMap/AssemblyMapCG.cpp:2804-2886
 *                     (GlobalToLocal = gather, Assemble = zero + scatter-add)
 *                     on this project's structured numbering (parity of
 *                     the reference's own RCM numbering is not claimed).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Nektar hex vertex order: v0(-,-,-) v1(+,-,-) v2(+,+,-) v3(-,+,-)
 *                          v4(-,-,+) v5(+,-,+) v6(+,+,+) v7(-,+,+) */
static const int hex_sign[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                   {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

int or_hex_is_regular(const double *V)
{
    /* HexGeom.cpp:101-145 */
    const double kNekZeroTol = 1.0e-12;
    for (int i = 0; i < 3; ++i) {
        double A = V[0 * 3 + i], B = V[1 * 3 + i], C = V[2 * 3 + i], D = V[3 * 3 + i];
        double E = V[4 * 3 + i], F = V[5 * 3 + i], G = V[6 * 3 + i], H = V[7 * 3 + i];
        double iso[8];
        iso[0] = 0.125 * (A + B + C + D + E + F + G + H);
        iso[1] = 0.125 * (-A + B + C - D - E + F + G - H);
        iso[2] = 0.125 * (-A - B + C + D - E - F + G + H);
        iso[3] = 0.125 * (-A - B - C - D + E + F + G + H);
        iso[4] = 0.125 * (A - B + C - D + E - F + G - H);
        iso[5] = 0.125 * (A + B - C - D - E - F + G + H);
        iso[6] = 0.125 * (A - B - C + D - E + F + G - H);
        iso[7] = 0.125 * (-A + B - C + D + E - F + G - H);
        double tmp = (fabs(iso[1]) + fabs(iso[2]) + fabs(iso[3])) * kNekZeroTol;
        for (int d = 4; d < 8; ++d)
            if (fabs(iso[d]) > tmp) return 0;
    }
    return 1;
}

/* d x_k / d xi_i of the trilinear map at xi: deriv[i][k] */
static void hex_deriv(const double *V, const double xi[3], double deriv[3][3])
{
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) deriv[i][k] = 0.0;
    for (int v = 0; v < 8; ++v) {
        const int *s = hex_sign[v];
        double f0 = 1 + s[0] * xi[0], f1 = 1 + s[1] * xi[1], f2 = 1 + s[2] * xi[2];
        double dn[3] = {s[0] * f1 * f2 / 8.0, f0 * s[1] * f2 / 8.0, f0 * f1 * s[2] / 8.0};
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) deriv[i][k] += dn[i] * V[v * 3 + k];
    }
}

/* GeomFactors.cpp:248-332 and :401-477 for one point: jac = sqrt(det g),
 * df[k*3+j] = sum_i deriv[i][k] g^{ij} */
static void factors_from_deriv(double deriv[3][3], double *jac, double *df)
{
    double g[9], adj[9];
    for (int i = 0, l = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j, ++l) {
            g[l] = 0.0;
            for (int k = 0; k < 3; ++k) g[l] = deriv[i][k] * deriv[j][k] + g[l];
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            int a = ((i + 1) % 3) * 3 + ((j + 1) % 3);
            int b = ((i + 1) % 3) * 3 + ((j + 2) % 3);
            int c = ((i + 2) % 3) * 3 + ((j + 1) % 3);
            int d = ((i + 2) % 3) * 3 + ((j + 2) % 3);
            adj[j * 3 + i] = g[a] * g[d] - g[b] * g[c];
        }
    double det = 0.0;
    for (int i = 0; i < 3; ++i) det = g[i] * adj[i * 3] + det;
    for (int i = 0; i < 9; ++i) adj[i] /= det;
    if (jac) *jac = sqrt(det);
    if (df)
        for (int k = 0, l = 0; k < 3; ++k)
            for (int j = 0; j < 3; ++j, ++l) {
                double s = 0.0;
                for (int i = 0; i < 3; ++i) s = deriv[i][k] * adj[3 * i + j] + s;
                df[l] = s;
            }
}

void or_hex_geom(const double *verts, int nelmt, int nq, int deformed,
                 double *jac, double *df)
{
    double z[64];
    or_points(OR_GLL, nq, z, NULL, NULL);
    const int nqt = nq * nq * nq;
    for (int e = 0; e < nelmt; ++e) {
        const double *V = verts + (size_t)e * 24;
        if (!deformed) {
            double xi[3] = {0.0, 0.0, 0.0}, dv[3][3];
            hex_deriv(V, xi, dv);
            factors_from_deriv(dv, jac ? jac + e : NULL, df ? df + (size_t)e * 9 : NULL);
            continue;
        }
        for (int k = 0, c = 0; k < nq; ++k)
            for (int j = 0; j < nq; ++j)
                for (int i = 0; i < nq; ++i, ++c) {
                    double xi[3] = {z[i], z[j], z[k]}, dv[3][3];
                    hex_deriv(V, xi, dv);
                    size_t q = (size_t)e * nqt + c;
                    factors_from_deriv(dv, jac ? jac + q : NULL, df ? df + q * 9 : NULL);
                }
    }
}

void or_hex_coords(const double *verts, int nelmt, int nq, double *xyz)
{
    double z[64];
    or_points(OR_GLL, nq, z, NULL, NULL);
    const int nqt = nq * nq * nq;
    for (int e = 0; e < nelmt; ++e) {
        const double *V = verts + (size_t)e * 24;
        for (int k = 0, c = 0; k < nq; ++k)
            for (int j = 0; j < nq; ++j)
                for (int i = 0; i < nq; ++i, ++c) {
                    double xi[3] = {z[i], z[j], z[k]};
                    double *o = xyz + ((size_t)e * nqt + c) * 3;
                    o[0] = o[1] = o[2] = 0.0;
                    for (int v = 0; v < 8; ++v) {
                        const int *s = hex_sign[v];
                        double N = (1 + s[0] * xi[0]) * (1 + s[1] * xi[1]) * (1 + s[2] * xi[2]) / 8.0;
                        for (int d = 0; d < 3; ++d) o[d] += N * V[v * 3 + d];
                    }
                }
    }
}

/* Affine tet: x = v0 + (v1-v0)(1+xi0)/2 + (v2-v0)(1+xi1)/2 + (v3-v0)(1+xi2)/2,
 * the straight-sided TetGeom map; derivatives w.r.t. the standard
 * (uncollapsed) coordinates as GeomFactors uses them. */
void or_tet_geom(const double *verts, int nelmt, double *jac, double *df)
{
    for (int e = 0; e < nelmt; ++e) {
        const double *V = verts + (size_t)e * 12;
        double dv[3][3];
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) dv[i][k] = 0.5 * (V[(i + 1) * 3 + k] - V[k]);
        factors_from_deriv(dv, jac ? jac + e : NULL, df ? df + (size_t)e * 9 : NULL);
    }
}

static uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static double uniform_pm1(uint64_t *s)
{
    return 2.0 * (double)(splitmix64(s) >> 11) * 0x1.0p-53 - 1.0;
}

void or_hex_mesh(int n, double amp, uint64_t seed, double *verts)
{
    const int nv = n + 1;
    const double h = 1.0 / n;
    double *X = (double *)malloc((size_t)nv * nv * nv * 3 * sizeof(double));
    uint64_t st = seed;
    for (int iz = 0; iz < nv; ++iz)
        for (int iy = 0; iy < nv; ++iy)
            for (int ix = 0; ix < nv; ++ix) {
                double *x = X + ((size_t)(iz * nv + iy) * nv + ix) * 3;
                x[0] = ix * h;
                x[1] = iy * h;
                x[2] = iz * h;
                int interior = ix > 0 && ix < n && iy > 0 && iy < n && iz > 0 && iz < n;
                if (interior && amp != 0.0)
                    for (int d = 0; d < 3; ++d) x[d] += amp * h * uniform_pm1(&st);
            }
    for (int ez = 0; ez < n; ++ez)
        for (int ey = 0; ey < n; ++ey)
            for (int ex = 0; ex < n; ++ex) {
                double *V = verts + ((size_t)(ez * n + ey) * n + ex) * 24;
                for (int v = 0; v < 8; ++v) {
                    int ix = ex + (hex_sign[v][0] > 0), iy = ey + (hex_sign[v][1] > 0),
                        iz = ez + (hex_sign[v][2] > 0);
                    memcpy(V + v * 3, X + ((size_t)(iz * nv + iy) * nv + ix) * 3, 3 * sizeof(double));
                }
            }
    free(X);
}

int or_hex_nglobal(int n, int nm)
{
    int N = n * (nm - 1) + 1;
    return N * N * N;
}

/* 1-D global node of local mode p in element e_d (vertex modes 0 and 1 at
 * the element's left/right end, bubbles in between). */
static int node1d(int e, int p, int nm)
{
    return e * (nm - 1) + (p == 0 ? 0 : p == 1 ? nm - 1 : p - 1);
}

void or_hex_l2g(int n, int nm, int *map, uint8_t *dirichlet)
{
    const int N = n * (nm - 1) + 1;
    const int nc = nm * nm * nm;
    for (int ez = 0; ez < n; ++ez)
        for (int ey = 0; ey < n; ++ey)
            for (int ex = 0; ex < n; ++ex) {
                int e = (ez * n + ey) * n + ex;
                for (int r = 0; r < nm; ++r)
                    for (int q = 0; q < nm; ++q)
                        for (int p = 0; p < nm; ++p) {
                            int gx = node1d(ex, p, nm), gy = node1d(ey, q, nm), gz = node1d(ez, r, nm);
                            map[(size_t)e * nc + (r * nm + q) * nm + p] = (gz * N + gy) * N + gx;
                        }
            }
    if (dirichlet)
        for (int gz = 0; gz < N; ++gz)
            for (int gy = 0; gy < N; ++gy)
                for (int gx = 0; gx < N; ++gx)
                    dirichlet[(gz * N + gy) * N + gx] =
                        gx == 0 || gx == N - 1 || gy == 0 || gy == N - 1 || gz == 0 || gz == N - 1;
}

void or_assemble(int nloc, const int *map, const double *loc, int nglob,
                 double *glob)
{
    memset(glob, 0, (size_t)nglob * sizeof(double));
    for (int i = 0; i < nloc; ++i) glob[map[i]] += loc[i];
}

void or_global_to_local(int nloc, const int *map, const double *glob,
                        double *loc)
{
    for (int i = 0; i < nloc; ++i) loc[i] = glob[map[i]];
}
