// State of the iterative static condensation (IterativeSolver in
// CGSpace::StaticCond, static_cond.cpp).
#pragma once

#include <memory>
#include <vector>

#include "assembly.hpp"

namespace nektar_amd {

// one collection's condensed blocks: its interior / boundary mode lists,
// the inverses of the interior blocks A_ii^-1 (E x ni x ni, column-major;
// m_invD), and its offsets in the local (off), local-boundary (boff) and
// local-interior (ioff) vectors
struct StaticCondBlock {
    int E = 0, nc = 0, ni = 0, nb = 0;
    int64_t off = 0, boff = 0, ioff = 0;
    DeviceBuffer dI, dB, invD;
};

struct StaticCondData {
    std::vector<StaticCondBlock> blk;
    int64_t nlb = 0, nli = 0;           // local boundary / interior entries
    std::unique_ptr<AssemblyMap> bmap;  // local boundary entries -> boundary system (Dirichlet first)
    int nred = 1;                       // reduction blocks of the boundary system
    DeviceBuffer fb, fi, ti, zi, xb, sb;  // local boundary / interior vectors
    DeviceBuffer invd_b, invd_g;        // inverse Schur diagonal: boundary system / caller's numbering
    DeviceBuffer bglob;                 // boundary id -> caller's global id
    DeviceBuffer cg[6], bvec, red, pa, pmu;  // boundary-system CG vectors (x r p q w s), rhs, scratch
    DeviceBuffer xloc, xg;              // the solution: local layout / caller's global numbering
    double rhs_mag = -1.0;              // v_PreSolve's smoothed magnitude (< 0: unset)
};

}  // namespace nektar_amd
