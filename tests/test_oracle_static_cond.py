"""The condensed boundary system of tests/sc_system.py and the oracle's CG
on it (or_cg_dense_mag) pinned on the CPU: the condensed CG solution,
back-substituted, is the full system's solution (dense solve) on every
local entry; its Schur diagonal is the oracle's static-condensation diagonal
(or_sc_diagonal); the rhs magnitude is |AssembleBnd(F_bnd)|^2 of the
uncondensed forcing, which differs from that of the condensed right-hand
side."""
import numpy as np
import pytest

import _oracle as O
import sc_system as SC
from test_oracle_cg_loc import _global, _hex_case, _tet_case

CASES = [("tet", (2, 5, 0.5)), ("tet", (2, 6, 1.0)), ("hex", (2, 5, 6, 2.0, 0.0, 4)), ("hex", (2, 6, 7, 1.0, 0.12, 9))]


def _prep(kind, args):
    blocks, mp, sg, dirichlet, bl = _tet_case(*args) if kind == "tet" else _hex_case(*args)
    m, ndir, A, b, invd = _global(blocks, mp, sg, dirichlet, bl)
    shape = O.TET if kind == "tet" else O.HEX
    fl = np.tile(O.interior_modes(shape, args[1]), len(blocks))
    return blocks, m, sg, ndir, A, b, fl, bl


@pytest.mark.parametrize("kind,args", CASES)
def test_condensed_cg_solves_the_full_system(kind, args):
    blocks, m, sg, ndir, A, b, fl, bl = _prep(kind, args)
    ng = len(b)
    sys = SC.condensed(blocks, m, sg, fl, bl, ng, ndir)
    S, bc, rows = sys["S"], sys["b"], sys["rows"]
    d = np.diag(S).copy()
    invd = np.zeros(len(d))
    invd[ndir:] = 1.0 / d[ndir:]
    ref_d = O.sc_diagonal(blocks, fl, m, sg, ng, ndir)
    assert np.max(np.abs(d[ndir:] - ref_d[rows[ndir:]])) <= 1e-12 * np.max(np.abs(d))
    xb, info = O.cg_dense_mag(S, bc, ndir, invd, sys["rhs_mag"], tol=1e-12)
    assert info.final_error < 1e-12 and info.rhs_magnitude == sys["rhs_mag"]
    xl = SC.back_substitute(sys, xb, len(m))
    xd = np.zeros(ng)
    xd[ndir:] = np.linalg.solve(A[ndir:, ndir:], b[ndir:])
    assert np.max(np.abs(xl - sg * xd[m])) <= 1e-9 * np.max(np.abs(xd))
    # the magnitude of the uncondensed boundary forcing differs from that of
    # the condensed rhs (the GPU test compares the reported magnitude)
    assert abs(float(bc @ bc) - sys["rhs_mag"]) > 1e-3 * sys["rhs_mag"]
