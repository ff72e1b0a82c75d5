"""GPU iterative static condensation (solver "IterativeStaticCond":
GlobalLinSysStaticCond.cpp:98-230, GlobalLinSysIterativeStaticCond.cpp:
378-492, PreconditionerDiagonal.cpp:132-154) on signed maps, against the
condensed system built on the CPU (tests/sc_system.py, pinned against dense
solves and the oracle's static-condensation diagonal) and the oracle's CG on
it:

* the preconditioner's inverse diagonal = 1 / the assembled Schur diagonal
  on the free boundary rows, 0 elsewhere, to 1e-12;
* the first 9 boundary iterates (fewer on a system that converges first)
  equal the oracle's to 1e-10;
* the rhs magnitude is v_PreSolve's, and successive solves smooth it
  0.9 / 0.1 as the reference does;
* iteration counts at 1e-10 within max(2, 2 %) of the oracle's;
* at 1e-12 the local solution equals the dense full solve within the
  stopping tolerance times the condition number, and the full-system
  GPU CG's solution."""
import numpy as np
import pytest

import _oracle as O
import sc_system as SC
from test_gpu_cg_loc import _cond_bound, _systems

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def nmf():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import nektar_amd
    nektar_amd.lib()
    return nektar_amd


CASES = [("tet", (2, 5, 0.5)), ("tet", (2, 7, 2.0)), ("hex", (3, 5, 6, 2.5, 0.0, 7)), ("hex", (2, 7, 8, 1.0, 0.12, 9)),
         ("hex", (2, 3, 4, 1.0, 0.0, 3))]


def _oracle_cg(sys, ndir, tol, maxit):
    d = np.diag(sys["S"]).copy()
    invd = np.zeros(len(d))
    invd[ndir:] = 1.0 / d[ndir:]
    return O.cg_dense_mag(sys["S"], sys["b"], ndir, invd, sys["rhs_mag"], tol=tol, maxit=maxit), invd


@pytest.mark.parametrize("kind,args", CASES)
def test_static_cond_against_oracle(nmf, kind, args):
    blocks, m, sg, ndir, A, b, invd_full, bl, ops, amap = _systems(nmf, kind, args)
    shape = O.TET if kind == "tet" else O.HEX
    fl = np.tile(O.interior_modes(shape, args[1]), len(blocks))
    ng = len(b)
    sys = SC.condensed(blocks, m, sg, fl, bl, ng, ndir)
    rows = sys["rows"]
    ls = nmf.LinearSystem(amap, ops, solver="IterativeStaticCond")
    # preconditioner
    (_, _), invd_b = _oracle_cg(sys, ndir, 1e-10, 1)
    d = ls.inv_diag()
    ref = np.zeros(ng)
    ref[rows] = invd_b
    assert np.max(np.abs(d - ref)) <= 1e-12 * np.max(np.abs(ref))
    bt = torch.from_numpy(bl).cuda()
    # first iterates (the rhs magnitude of the first solve is v_PreSolve's)
    ls.set_rhs_local(bt)
    info = ls.solve(tol=1e-14, maxit=8)
    (xo, io), _ = _oracle_cg(sys, ndir, 1e-14, 8)
    assert info.iterations == io.iterations <= 9  # fewer when the boundary system converges first
    assert abs(info.rhs_magnitude - sys["rhs_mag"]) <= 1e-12 * sys["rhs_mag"]
    xg = ls.solution()
    assert np.max(np.abs(xg[rows] - xo)) <= 1e-10 * np.max(np.abs(xo))
    # a fresh handle per tolerance keeps the magnitude unsmoothed
    ls = nmf.LinearSystem(amap, ops, solver="IterativeStaticCond")
    ls.set_rhs_local(bt)
    info = ls.solve(tol=1e-10, maxit=5000)
    (xo, io), _ = _oracle_cg(sys, ndir, 1e-10, 5000)
    assert info.converged
    assert abs(info.iterations - io.iterations) <= max(2, io.iterations // 50), (info.iterations, io.iterations)
    # a second solve on the handle: magnitude 0.9 old + 0.1 new
    b2 = 0.5 * bl
    sys2 = SC.condensed(blocks, m, sg, fl, b2, ng, ndir)
    ls.set_rhs_local(torch.from_numpy(b2).cuda())
    info2 = ls.solve(tol=1e-12, maxit=5000)
    smoothed = 0.9 * sys["rhs_mag"] + 0.1 * sys2["rhs_mag"]
    assert abs(info2.rhs_magnitude - smoothed) <= 1e-12 * smoothed
    assert info2.converged and info2.final_error < 1e-12
    xd = np.zeros(ng)
    xd[ndir:] = np.linalg.solve(A[ndir:, ndir:], 0.5 * b[ndir:])
    scale, bound = np.max(np.abs(xd)), _cond_bound(A, ndir, 1e-12)
    xs = ls.solution()
    assert np.max(np.abs(xs - xd)) <= bound * scale
    assert np.all(xs[:ndir] == 0.0)
    yl = torch.zeros(len(m), dtype=torch.float64, device="cuda")
    ls.solution_local(yl)
    torch.cuda.synchronize()
    assert np.array_equal(yl.cpu().numpy(), sg * xs[m])
    glo = nmf.LinearSystem(amap, ops)
    glo.set_rhs_local(torch.from_numpy(b2).cuda())
    glo.solve(tol=1e-12, maxit=5000)
    assert np.max(np.abs(xs - glo.solution())) <= bound * scale
