"""GPU element-local CG (LinSysIterSolver "ConjugateGradientLoc",
NekLinSysIterCGLoc.cpp:114-308) on caller-supplied signed maps, against the
oracle's local CG (or_cg_loc, pinned on the CPU against the assembled CG and
dense solves), the GPU's assembled-vector CG on the same system and dense
solves:

* the first 9 iterates equal to the oracle's to 1e-10;
* iteration counts at a 1e-10 relative residual within the spread that
  rounding alone gives on the CPU (the oracle's local against its assembled
  recurrence, equal in exact arithmetic, widened by half; at least
  max(2, 1 %)), for the local and the assembled GPU recurrences alike;
* the same rhs magnitude (<AssembleLoc(b), b> over all rows) to 1e-12;
* the solution equal to the dense solve within the stopping tolerance times
  the condition number of the Jacobi-scaled system;
* the local solution continuous, bit for bit, and equal to the global one
  scattered back; maxit reached at the same count as the oracle;
* an unknown solver name is rejected."""
import numpy as np
import pytest

import _oracle as O
import geometry_cases as G
from test_gpu_general_assembly import _setup, _tet_setup
from test_oracle_cg_loc import _global, _hex_case, _tet_case

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def nmf():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import nektar_amd
    nektar_amd.lib()
    return nektar_amd


def _cond_bound(A, ndir, tol):
    Af = A[ndir:, ndir:]
    dh = 1.0 / np.sqrt(np.diag(Af))
    return 10 * tol * np.linalg.cond(dh[:, None] * Af * dh[None, :])


def _systems(nmf, kind, args):
    if kind == "tet":
        n, nm, lam = args
        blocks, mp, sg, dirichlet, bl = _tet_case(n, nm, lam)
        ops = _tet_setup(nmf, n, nm, lam, len(blocks) // 3)["ops"]
    else:
        n, nm, nq, lam, amp, seed = args
        blocks, mp, sg, dirichlet, bl = _hex_case(n, nm, nq, lam, amp, seed)
        ops = _setup(nmf, n, nm, nq, lam, amp, seed, max(1, n ** 3 // 3))["ops"]
    m, ndir, A, b, invd = _global(blocks, mp, sg, dirichlet, bl)
    amap = nmf.AssemblyMap(m, sg, len(b), ndir)
    return blocks, m, sg, ndir, A, b, invd, bl, ops, amap


CASES = [("tet", (2, 5, 0.5)), ("tet", (3, 4, 2.0)), ("hex", (3, 5, 6, 2.5, 0.0, 7)),
         ("hex", (3, 6, 7, 1.0, 0.12, 9))]


@pytest.mark.parametrize("kind,args", CASES)
def test_local_cg_against_oracle_and_global_cg(nmf, kind, args):
    blocks, m, sg, ndir, A, b, invd, bl, ops, amap = _systems(nmf, kind, args)
    ng = len(b)
    loc = nmf.LinearSystem(amap, ops, solver="ConjugateGradientLoc")
    glo = nmf.LinearSystem(amap, ops)
    assert np.array_equal(loc.inv_diag(), glo.inv_diag())
    assert np.max(np.abs(loc.inv_diag() - invd)) <= 1e-12 * np.max(np.abs(invd))
    bt = torch.from_numpy(bl).cuda()
    loc.set_rhs_local(bt)
    glo.set_rhs_local(bt)
    # the first iterates: the same recurrence as the oracle's, before
    # rounding differences grow (9 iterations: maxit 8)
    il = loc.solve(tol=1e-14, maxit=8)
    xo, io = O.cg_loc(blocks, m, sg, ng, ndir, bl, invd, tol=1e-14, maxit=8)
    assert il.iterations == io.iterations == 9
    first = np.zeros(ng)
    first[m] = sg * xo
    assert np.max(np.abs(loc.solution() - first)) <= 1e-10 * np.max(np.abs(first))
    # iteration counts at 1e-10: both GPU recurrences and the oracle's
    # differ only by rounding (the element operators sum in another order)
    il = loc.solve(tol=1e-10, maxit=5000)
    ig = glo.solve(tol=1e-10, maxit=5000)
    xo, io = O.cg_loc(blocks, m, sg, ng, ndir, bl, invd, tol=1e-10, maxit=5000)
    _, iog = O.cg_dense(A, b, ndir, invd, tol=1e-10, maxit=5000)
    assert il.converged and ig.converged
    # slack: the spread between the oracle's local and assembled recurrences
    # (equal in exact arithmetic) on the same inputs, widened by half
    slack = max(2, io.iterations // 100, (3 * abs(io.iterations - iog.iterations)) // 2)
    assert abs(il.iterations - io.iterations) <= slack, (il.iterations, io.iterations)
    assert abs(ig.iterations - io.iterations) <= slack, (ig.iterations, io.iterations)
    assert abs(il.rhs_magnitude - io.rhs_magnitude) <= 1e-12 * io.rhs_magnitude
    assert il.final_error < 1e-10
    # solutions at 1e-12 against the dense solve and the global CG
    tol = 1e-12
    il = loc.solve(tol=tol, maxit=5000)
    ig = glo.solve(tol=tol, maxit=5000)
    assert il.converged and ig.converged and il.final_error < tol
    xd = np.zeros(ng)
    xd[ndir:] = np.linalg.solve(A[ndir:, ndir:], b[ndir:])
    scale, bound = np.max(np.abs(xd)), _cond_bound(A, ndir, tol)
    xl, xg = loc.solution(), glo.solution()
    assert np.max(np.abs(xl - xd)) <= bound * scale
    assert np.max(np.abs(xl - xg)) <= bound * scale
    assert np.all(xl[:ndir] == 0.0)
    # the local solution: continuous bit for bit, the global values scattered back
    yl = torch.zeros(len(m), dtype=torch.float64, device="cuda")
    loc.solution_local(yl)
    torch.cuda.synchronize()
    assert np.array_equal(yl.cpu().numpy(), sg * xl[m])
    # the global solver's local solution is GlobalToLocal of its global one
    yg = torch.zeros(len(m), dtype=torch.float64, device="cuda")
    glo.solution_local(yg)
    torch.cuda.synchronize()
    assert np.array_equal(yg.cpu().numpy(), sg * xg[m])


def test_maxit_zero_rhs_and_bad_names(nmf):
    blocks, m, sg, ndir, A, b, invd, bl, ops, amap = _systems(nmf, "tet", (2, 4, 1.0))
    loc = nmf.LinearSystem(amap, ops, solver="ConjugateGradientLoc")
    loc.set_rhs_local(torch.from_numpy(bl).cuda())
    info = loc.solve(tol=1e-14, maxit=3)
    xo, io = O.cg_loc(blocks, m, sg, len(b), ndir, bl, invd, tol=1e-14, maxit=3)
    assert info.iterations == io.iterations == 4 and not info.converged
    first = np.zeros(len(b))
    first[m] = sg * xo
    assert np.max(np.abs(loc.solution() - first)) <= 1e-11 * np.max(np.abs(first))
    # a second solve on the same handle with a zero right-hand side
    loc.set_rhs_local(torch.zeros(len(m), dtype=torch.float64, device="cuda"))
    info = loc.solve(tol=1e-10, maxit=100)
    assert info.iterations == 0 and info.converged and info.rhs_magnitude == 1.0
    assert np.all(loc.solution() == 0.0)
    for bad in ("GMRES", "ConjugateGradientLocal", ""):
        with pytest.raises(nmf.NektarError):
            nmf.LinearSystem(amap, ops, solver=bad)
