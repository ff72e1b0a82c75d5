"""GPU parity of IProductWRTDerivBase (out = sum_j IProduct w.r.t. dphi/dx_j
of in_j) against the FP64 oracle: relative max error <= 1e-12 (FP64) and
<= 2e-5 (FP32); the composition with the GPU PhysDeriv equals the GPU
Helmholtz stiffness (lambda = 0) to 1e-12."""
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


def geometry(shape, nq0, deformed, nel_hint=27):
    if shape == O.HEX:
        n = round(nel_hint ** (1 / 3))
        _, jac, df = G.hex_geometry(n, nq0, deformed)
        return n ** 3, jac, df
    _, jac, df = G.tet_geometry(nel_hint, nq0, deformed)
    return nel_hint, jac, df


def make(nmf, shape, nm, nq0, nel, deformed, jac, df, fp32=False, op="IProductWRTDerivBase"):
    sname = "Hex" if shape == O.HEX else "Tet"
    o = nmf.GetOperatorFactory().CreateInstance(f"{op}{nmf.GetOpstring(sname, deformed)}", nm, nq0, nel,
                                                nmf.FP32 if fp32 else nmf.FP64)
    if op != "PhysDeriv":
        o.SetJac(jac)
    o.SetDF(df)
    if op == "Helmholtz":
        o.SetLambda(0.0)
    return o


@pytest.mark.parametrize("shape", [O.HEX, O.TET], ids=["hex", "tet"])
@pytest.mark.parametrize("deformed", [False, True], ids=["regular", "deformed"])
@pytest.mark.parametrize("nm", [2, 3, 5, 7, 10])
@pytest.mark.parametrize("fp32", [False, True], ids=["fp64", "fp32"])
def test_against_oracle(nmf, shape, deformed, nm, fp32):
    for nq0 in (nm + 1, nm + 2):
        nel, jac, df = geometry(shape, nq0, deformed)
        op = make(nmf, shape, nm, nq0, nel, deformed, jac, df, fp32)
        n = op.input_size()
        f = [G.coeffs(n, s) for s in (21, 22, 23)]
        dt = torch.float32 if fp32 else torch.float64
        fin = [torch.from_numpy(a).to(dt).cuda() for a in f]
        out = torch.full((op.output_size(),), float("nan"), dtype=dt, device="cuda")
        op(fin, out)
        torch.cuda.synchronize()
        ref = O.iproduct_deriv(shape, nm, nq0, nel, deformed, jac, df, *f)
        err = G.rel_err(out.double().cpu().numpy(), ref)
        assert err <= (2e-5 if fp32 else 1e-12), (nq0, err)


@pytest.mark.parametrize("shape", [O.HEX, O.TET], ids=["hex", "tet"])
def test_host_apply_and_composition(nmf, shape):
    nm, nq0, deformed = 6, 7, True
    nel, jac, df = geometry(shape, nq0, deformed, 8)
    ipd = make(nmf, shape, nm, nq0, nel, deformed, jac, df)
    pd = make(nmf, shape, nm, nq0, nel, deformed, jac, df, op="PhysDeriv")
    hm = make(nmf, shape, nm, nq0, nel, deformed, jac, df, op="Helmholtz")
    x = G.coeffs(nel * O.ncoeffs(shape, nm), 4)
    u = O.bwdtrans(shape, nm, nq0, nel, x)
    grad = pd.apply_host(u)
    y = ipd.apply_host(grad)
    assert G.rel_err(y, hm.apply_host(x)) <= 1e-12
    ref = O.iproduct_deriv(shape, nm, nq0, nel, deformed, jac, df, *grad)
    assert G.rel_err(y, ref) <= 1e-12


def test_errors(nmf):
    nel, jac, df = geometry(O.HEX, 5, False, 8)
    op = make(nmf, O.HEX, 4, 5, nel, False, jac, df)
    x = torch.zeros(op.input_size(), dtype=torch.float64, device="cuda")
    y = torch.zeros(op.output_size(), dtype=torch.float64, device="cuda")
    with pytest.raises(nmf.NektarError):
        op([x, x], y)
    rc = nmf.lib().nmf_operator_apply(op._h, x.data_ptr(), y.data_ptr(), None)
    assert rc == 1  # NMF_ERR_INVALID: three inputs needed
    fresh = nmf.GetOperatorFactory().CreateInstance("IProductWRTDerivBase_Tet_Regular", 4, 5, 3)
    with pytest.raises(nmf.NektarError):  # geometry not set
        fresh([x[:fresh.input_size()]] * 3, y[:fresh.output_size()])
    b = nmf.Batch()
    with pytest.raises(nmf.NektarError):
        b.add(op, x, y)
