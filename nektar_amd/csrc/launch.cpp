// Runtime (shape, NM, NQ) -> template instantiation dispatch.
#include "launch.hpp"

namespace nektar_amd {

#define NMF_DECL(n)                                              \
    hipError_t launch_hex_nm##n(int nq, const KernelArgs &a);    \
    hipError_t launch_tet_nm##n(int nq, const KernelArgs &a);    \
    hipError_t launch_hex_solver_nm##n(int nq, const SolverArgs &a); \
    int hex_tiles_nm##n(int nq, bool fp32, int *out, int max);        \
    hipError_t launch_hex_grouped_nm##n(int nq, const GroupArgs &a);  \
    hipError_t launch_tet_grouped_nm##n(int nq, const GroupArgs &a);
NMF_DECL(2) NMF_DECL(3) NMF_DECL(4) NMF_DECL(5) NMF_DECL(6) NMF_DECL(7) NMF_DECL(8) NMF_DECL(9) NMF_DECL(10)
#undef NMF_DECL

// hex: Q = NUMMODES + 1 (the MODIFIED default) and + 2 everywhere; the Collections operators
// also take Q = NUMMODES for nm >= 3 (the batched, grouped and solver kernels do not)
bool hex_supported(int nm, int nq) { return nm >= kMinModes && nm <= kMaxModes && (nq == nm + 1 || nq == nm + 2); }
bool hex_op_supported(int nm, int nq) { return hex_supported(nm, nq) || (nq == nm && nm >= 3 && nm <= kMaxModes); }
bool tet_supported(int nm, int nq0) { return nm >= kMinModes && nm <= kMaxModes && (nq0 == nm + 1 || nq0 == nm + 2); }

#define NMF_CASE(shape, n) \
    case n: return launch_##shape##_nm##n(nq, a);

hipError_t launch_hex(int nm, int nq, const KernelArgs &a)
{
    switch (nm) {
        NMF_CASE(hex, 2) NMF_CASE(hex, 3) NMF_CASE(hex, 4) NMF_CASE(hex, 5) NMF_CASE(hex, 6)
        NMF_CASE(hex, 7) NMF_CASE(hex, 8) NMF_CASE(hex, 9) NMF_CASE(hex, 10)
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_tet(int nm, int nq, const KernelArgs &a)
{
    switch (nm) {
        NMF_CASE(tet, 2) NMF_CASE(tet, 3) NMF_CASE(tet, 4) NMF_CASE(tet, 5) NMF_CASE(tet, 6)
        NMF_CASE(tet, 7) NMF_CASE(tet, 8) NMF_CASE(tet, 9) NMF_CASE(tet, 10)
    default: return hipErrorInvalidValue;
    }
}

int hex_tile_candidates(int nm, int nq, bool fp32, int *out, int max)
{
    switch (nm) {
#define NMF_TILES(n) \
    case n: return hex_tiles_nm##n(nq, fp32, out, max);
        NMF_TILES(2) NMF_TILES(3) NMF_TILES(4) NMF_TILES(5) NMF_TILES(6) NMF_TILES(7) NMF_TILES(8) NMF_TILES(9)
        NMF_TILES(10)
#undef NMF_TILES
    default: return 0;
    }
}

int tet_tile_candidates(int nq0, int *out, int max)
{
    const int cand[2] = {tet_default_tpe(nq0), tet_default_tpe(nq0) == tet_packed_tpe(nq0) ? tet_wave_tpe(nq0)
                                                                                          : tet_packed_tpe(nq0)};
    int n = 0;
    for (int c = 0; c < 2 && n < max; ++c)
        if (c == 0 || cand[1] != cand[0]) out[n++] = cand[c];
    return n;
}

hipError_t launch_hex_grouped(int nm, int nq, const GroupArgs &a)
{
    switch (nm) {
        NMF_CASE(hex_grouped, 2) NMF_CASE(hex_grouped, 3) NMF_CASE(hex_grouped, 4) NMF_CASE(hex_grouped, 5)
        NMF_CASE(hex_grouped, 6) NMF_CASE(hex_grouped, 7) NMF_CASE(hex_grouped, 8) NMF_CASE(hex_grouped, 9)
        NMF_CASE(hex_grouped, 10)
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_tet_grouped(int nm, int nq, const GroupArgs &a)
{
    switch (nm) {
        NMF_CASE(tet_grouped, 2) NMF_CASE(tet_grouped, 3) NMF_CASE(tet_grouped, 4) NMF_CASE(tet_grouped, 5)
        NMF_CASE(tet_grouped, 6) NMF_CASE(tet_grouped, 7) NMF_CASE(tet_grouped, 8) NMF_CASE(tet_grouped, 9)
        NMF_CASE(tet_grouped, 10)
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_hex_solver(int nm, int nq, const SolverArgs &a)
{
    switch (nm) {
        NMF_CASE(hex_solver, 2) NMF_CASE(hex_solver, 3) NMF_CASE(hex_solver, 4) NMF_CASE(hex_solver, 5)
        NMF_CASE(hex_solver, 6) NMF_CASE(hex_solver, 7) NMF_CASE(hex_solver, 8) NMF_CASE(hex_solver, 9)
        NMF_CASE(hex_solver, 10)
    default: return hipErrorInvalidValue;
    }
}

}  // namespace nektar_amd
