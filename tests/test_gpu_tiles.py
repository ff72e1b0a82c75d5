"""Workgroup tiles of the hex Helmholtz kernels and the autotuner.

The vector-ALU tiles compute each element with the same arithmetic in the
same order, so their outputs must be bit-identical; so must the
matrix-core (MFMA) tiles among themselves (tile >= TILE_MFMA, from
NUMMODES 5 on), whose products sum in another order.  Each family's FP64
output is checked against the oracle (relative max error <= 1e-12) and
the FP32 output against it at 2e-5.  The solver's tuned gather-apply must
leave the CG iterates bit-identical within a family."""
import numpy as np
import pytest

import _oracle as O
import geometry_cases as G

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def nmf():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import nektar_amd
    nektar_amd.lib()
    return nektar_amd


def _op(nmf, nm, nq, n, deformed, fp32):
    _, jac, df = G.hex_geometry(n, nq, deformed)
    key = "Helmholtz" + nmf.GetOpstring("Hex", deformed)
    op = nmf.GetOperatorFactory().CreateInstance(key, nm, nq, n ** 3, nmf.FP32 if fp32 else nmf.FP64)
    op.SetJac(jac)
    op.SetDF(df)
    op.SetLambda(1.5)
    return op, jac, df


@pytest.mark.parametrize("nm,nq", [(3, 4), (4, 6), (5, 6), (7, 8), (8, 10), (10, 11)])
@pytest.mark.parametrize("deformed", [False, True], ids=["regular", "deformed"])
@pytest.mark.parametrize("fp32", [False, True], ids=["fp64", "fp32"])
def test_every_tile_bit_identical(nmf, nm, nq, deformed, fp32):
    n = 5  # 125 elements: a ragged last workgroup for every tile > 1
    op, jac, df = _op(nmf, nm, nq, n, deformed, fp32)
    cand = op.tile_candidates()
    assert len(cand) >= 2 and 1 in cand, cand
    mfma = [t for t in cand if t >= nmf.TILE_MFMA]
    assert bool(mfma) == (nm >= 5), cand
    dt = torch.float32 if fp32 else torch.float64
    x = O.splitmix_uniform(7, op.input_size())
    xin = torch.from_numpy(x).to(dt).cuda()
    outs = {}
    for t in cand:
        op.tile = t
        assert op.tile == t
        y = torch.full((op.output_size(),), float("nan"), dtype=dt, device="cuda")
        op(xin, y)
        torch.cuda.synchronize()
        outs[t] = y.cpu().numpy()
    ref = O.helmholtz(O.HEX, nm, nq, n ** 3, deformed, jac, df, 1.5, x)
    for family in ([t for t in cand if t < nmf.TILE_MFMA], mfma):
        if not family:
            continue
        base = outs[family[0]]
        assert np.all(np.isfinite(base))
        for t in family:
            assert np.array_equal(outs[t], base), f"tile {t} differs from tile {family[0]}"
        err = np.max(np.abs(base.astype(np.float64) - ref)) / np.max(np.abs(ref))
        assert err <= (2e-5 if fp32 else 1e-12), (family[0], err)


def test_bad_tile_rejected(nmf):
    op, _, _ = _op(nmf, 4, 5, 2, False, False)
    with pytest.raises(nmf.NektarError):
        op.tile = 3
    op.tile = 0
    assert op.tile == 0


def test_non_helmholtz_has_no_tiles(nmf):
    op = nmf.GetOperatorFactory().CreateInstance("BwdTrans_Hex_Regular", 4, 5, 8)
    assert op.tile_candidates() == []
    t, times = op.autotune()
    assert t == 0 and times == {}


def test_autotune_records_choice(nmf):
    nmf.clear_tuning()
    op, jac, df = _op(nmf, 7, 8, 8, True, False)
    best, times = op.autotune()
    assert best in op.tile_candidates()
    assert set(times) == set(op.tile_candidates())
    assert all(np.isfinite(v) and v > 0 for v in times.values())
    assert times[best] == min(times.values())
    assert op.tile == best
    # a later operator of the same kind starts from the recorded tile
    op2, _, _ = _op(nmf, 7, 8, 2, True, False)
    assert op2.tile == best
    # other kinds are unaffected
    op3, _, _ = _op(nmf, 7, 8, 2, False, False)
    assert op3.tile == 0
    nmf.clear_tuning()
    op4, _, _ = _op(nmf, 7, 8, 2, True, False)
    assert op4.tile == 0


def test_solver_autotune_keeps_iterates(nmf):
    nmf.clear_tuning()
    s = nmf.HelmholtzSolver(4, 5, 6, 1.0, perturb=0.1, seed=5)
    s.set_forcing_manufactured(1)
    a = s.solve(False, 1e-10, 2000)
    xa = s.solution()
    t64, t32 = s.autotune()
    assert t64 > 0 and t32 > 0
    b = s.solve(False, 1e-10, 2000)
    assert a.converged and b.converged
    if t64 < nmf.TILE_MFMA:  # the default tile is a vector-ALU tile
        assert a.iterations == b.iterations
        assert np.array_equal(xa, s.solution())
    else:  # matrix-core products round differently
        assert abs(a.iterations - b.iterations) <= 2
        assert np.max(np.abs(xa - s.solution())) <= 1e-8 * np.max(np.abs(xa))
    m = s.solve(True, 1e-10, 4000)
    assert m.converged
    # a second solver of the same kind picks the recorded tiles up
    s2 = nmf.HelmholtzSolver(3, 5, 6, 1.0, perturb=0.1, seed=6)
    s2.set_forcing_manufactured(1)
    assert s2.solve(False, 1e-10, 2000).converged
    nmf.clear_tuning()
