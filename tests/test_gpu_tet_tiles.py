"""Work-item tiles of the tet kernels: packed (NQ0*NQ1 work-items per
element, elements sharing waves) and wave-aligned (rounded up to a
multiple of 64, so that at Q <= 8 every element is one wave that
synchronises only itself).

Both tiles run the same arithmetic in the same order, so every operator's
output must be bit-identical between them; the FP64 output is checked
against the oracle (relative max error <= 1e-12) and FP32 at 2e-5 (2e-4
above 7 modes).  The element counts leave a ragged last workgroup for both
tiles."""
import numpy as np
import pytest

import _oracle as O
import geometry_cases as G

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

OPS = ["BwdTrans", "IProduct", "PhysDeriv", "Helmholtz", "IProductWRTDerivBase",
       "LinearAdvectionDiffusionReaction"]


@pytest.fixture(scope="module")
def nmf():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import nektar_amd
    nektar_amd.lib()
    return nektar_amd


def _run(nmf, opname, nm, nq0, nel, deformed, fp32, jac, df, vel, xs, tile):
    op = nmf.GetOperatorFactory().CreateInstance(f"{opname}{nmf.GetOpstring('Tet', deformed)}", nm, nq0, nel,
                                                 nmf.FP32 if fp32 else nmf.FP64)
    if opname not in ("BwdTrans", "PhysDeriv"):
        op.SetJac(jac)
    if opname not in ("BwdTrans", "IProduct"):
        op.SetDF(df)
    if opname in ("Helmholtz", "LinearAdvectionDiffusionReaction"):
        op.SetLambda(1.3)
    if opname == "LinearAdvectionDiffusionReaction":
        op.SetAdvectionVelocities(vel)
    cand = op.tile_candidates()
    assert len(cand) == 2, cand
    op.tile = cand[tile]
    dt = torch.float32 if fp32 else torch.float64
    xin = [torch.from_numpy(x).to(dt).cuda() for x in xs]
    nout = 3 if opname == "PhysDeriv" else 1
    outs = [torch.full((op.output_size(),), float("nan"), dtype=dt, device="cuda") for _ in range(nout)]
    if opname == "IProductWRTDerivBase":
        op(xin, outs[0])
    elif opname == "PhysDeriv":
        op(xin[0], *outs)
    else:
        op(xin[0], outs[0])
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in outs], cand


def _oracle(opname, nm, nq0, nel, deformed, jac, df, vel, xs):
    if opname == "BwdTrans":
        return [O.bwdtrans(O.TET, nm, nq0, nel, xs[0])]
    if opname == "IProduct":
        return [O.iproduct(O.TET, nm, nq0, nel, deformed, jac, xs[0])]
    if opname == "PhysDeriv":
        return list(O.physderiv(O.TET, nm, nq0, nel, deformed, df, xs[0]))
    if opname == "IProductWRTDerivBase":
        return [O.iproduct_deriv(O.TET, nm, nq0, nel, deformed, jac, df, *xs)]
    if opname == "LinearAdvectionDiffusionReaction":
        return [O.adv_diff_react(O.TET, nm, nq0, nel, deformed, jac, df, vel, 1.3, xs[0])]
    return [O.helmholtz(O.TET, nm, nq0, nel, deformed, jac, df, 1.3, xs[0])]


@pytest.mark.parametrize("opname", OPS)
@pytest.mark.parametrize("nm,nq0", [(2, 3), (4, 5), (5, 7), (7, 8), (7, 9), (8, 9), (8, 10), (9, 10), (10, 11),
                                     (10, 12)])
@pytest.mark.parametrize("deformed", [False, True], ids=["regular", "deformed"])
@pytest.mark.parametrize("fp32", [False, True], ids=["fp64", "fp32"])
def test_tiles_bit_identical_and_oracle(nmf, opname, nm, nq0, deformed, fp32):
    nel = 13
    _, jac, df = G.tet_geometry(nel, nq0, deformed)
    nqt = nq0 * (nq0 - 1) ** 2
    nin = nqt if opname in ("IProduct", "PhysDeriv", "IProductWRTDerivBase") else O.ncoeffs(O.TET, nm)
    xs = [G.coeffs(nel * nin, s) for s in ((31, 32, 33) if opname == "IProductWRTDerivBase" else (31,))]
    vel = G.coeffs(nel * nqt * 3, 34)
    got0, cand = _run(nmf, opname, nm, nq0, nel, deformed, fp32, jac, df, vel, xs, 0)
    got1, _ = _run(nmf, opname, nm, nq0, nel, deformed, fp32, jac, df, vel, xs, 1)
    for a, b in zip(got0, got1):
        assert np.all(np.isfinite(a))
        assert np.array_equal(a, b), f"tiles {cand} differ"
    ref = _oracle(opname, nm, nq0, nel, deformed, jac, df, vel, xs)
    scale = max(np.max(np.abs(r)) for r in ref)
    err = max(np.max(np.abs(a.astype(np.float64) - r)) for a, r in zip(got0, ref)) / scale
    # FP32: single precision tables and accumulation; above 7 modes the
    # collapsed-coordinate factors amplify the rounding of the tables
    assert err <= ((2e-5 if nm <= 7 else 2e-4) if fp32 else 1e-12), err


def test_autotune_picks_a_candidate(nmf):
    _, jac, df = G.tet_geometry(2000, 8, False)
    op = nmf.GetOperatorFactory().CreateInstance("Helmholtz_Tet_Regular", 7, 8, 2000)
    op.SetJac(jac)
    op.SetDF(df)
    op.SetLambda(1.0)
    t, times = op.autotune()
    assert t in op.tile_candidates() and set(times) == set(op.tile_candidates())
    assert op.tile == t
