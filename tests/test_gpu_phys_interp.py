"""GPU parity of PhysInterp1DScaled against the FP64 oracle: relative max
error <= 1e-14 (FP64, the same fused products in the same order) and
<= 1e-6 (FP32), hex and tet, Q = P+2 and P+3, scaling factors 1, 1.5
(default) and 2; every size of factors 1.5 (Q = 3..12) and 2 (Q = 3..11),
which run on the compile-time-size kernel (other factors run on the
run-time-size kernel); polynomial reproduction at the benchmark's size (32,768
elements, hex and tet, Q = 8); the factor setter changes the output size;
factors whose output would not fit the kernel's staging are rejected on the
host."""
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


def make(nmf, shape, nm, nq0, nel, fp32=False):
    sname = "Hex" if shape == O.HEX else "Tet"
    return nmf.GetOperatorFactory().CreateInstance(f"PhysInterp1DScaled{nmf.GetOpstring(sname, False)}", nm, nq0,
                                                   nel, nmf.FP32 if fp32 else nmf.FP64)


def apply(op, u, fp32):
    dt = torch.float32 if fp32 else torch.float64
    x = torch.from_numpy(u).to(dt).cuda()
    y = torch.full((op.output_size(),), float("nan"), dtype=dt, device="cuda")
    op(x, y)
    torch.cuda.synchronize()
    return y.double().cpu().numpy()


@pytest.mark.parametrize("shape", [O.HEX, O.TET], ids=["hex", "tet"])
@pytest.mark.parametrize("nm", [2, 4, 7, 10])
@pytest.mark.parametrize("fp32", [False, True], ids=["fp64", "fp32"])
def test_against_oracle(nmf, shape, nm, fp32):
    nel = 37
    for nq0 in (nm + 1, nm + 2):
        for factor in (1.5, 2.0, 1.0):
            op = make(nmf, shape, nm, nq0, nel, fp32)
            try:
                op.SetScalingFactor(factor)
            except nmf.NektarError:
                # only the largest orders at factor 2 exceed the staging
                assert factor == 2.0 and nq0 >= 10
                continue
            no = O.interp_nq(shape, nq0, factor)
            assert op.output_size() == nel * no[0] * no[1] * no[2]
            assert op.input_size() == nel * O.nqtot(shape, nq0)
            u = G.coeffs(op.input_size(), 90 + nm)
            y = apply(op, u, fp32)
            ref = O.phys_interp1d_scaled(shape, nq0, nel, factor, u)
            assert G.rel_err(y, ref) <= (1e-6 if fp32 else 1e-14), (nq0, factor)


@pytest.mark.parametrize("shape", [O.HEX, O.TET], ids=["hex", "tet"])
@pytest.mark.parametrize("fp32", [False, True], ids=["fp64", "fp32"])
@pytest.mark.parametrize("factor", [1.5, 2.0], ids=["f1.5", "f2"])
def test_every_compiled_size(nmf, shape, fp32, factor):
    nel = 29
    for nq0 in range(3, 13 if factor == 1.5 else 12):
        op = make(nmf, shape, max(nq0 - 2, 2), nq0, nel, fp32)
        op.SetScalingFactor(factor)
        no = O.interp_nq(shape, nq0, factor)
        assert op.output_size() == nel * no[0] * no[1] * no[2]
        u = G.coeffs(op.input_size(), 300 + nq0)
        y = apply(op, u, fp32)
        ref = O.phys_interp1d_scaled(shape, nq0, nel, factor, u)
        # the library and the oracle evaluate the interpolation tables from
        # their own point sets, which differ in the last bit; through three
        # 1-D sums of up to 12 terms that reaches 1.04e-14 at tet Q = 12
        assert G.rel_err(y, ref) <= (1e-6 if fp32 else 2e-14), nq0


@pytest.mark.parametrize("shape", [O.HEX, O.TET], ids=["hex", "tet"])
def test_full_size_reproduces_polynomials(nmf, shape):
    """32,768 elements at Q = 8 and the default factor (the sizes of the
    per-operator benchmark): a product of polynomials of degree n_d - 1 in
    each (collapsed) direction is reproduced exactly at the output points
    of the same families, with a different scale in every element."""
    nel, nq0 = 32768, 8
    ni = [nq0] * 3 if shape == O.HEX else [nq0, nq0 - 1, nq0 - 1]
    no = O.interp_nq(shape, nq0, 1.5)
    fam = [O.GLL] * 3 if shape == O.HEX else [O.GLL, O.RADAU_A1, O.RADAU_A2]
    rng = np.random.default_rng(5)
    coef = [rng.uniform(-1.0, 1.0, n) for n in ni]

    def line(d, n):
        return np.polynomial.polynomial.polyval(O.points(fam[d], n)[0], coef[d])

    fin = np.einsum("k,j,i->kji", line(2, ni[2]), line(1, ni[1]), line(0, ni[0])).ravel()
    fout = np.einsum("k,j,i->kji", line(2, no[2]), line(1, no[1]), line(0, no[0])).ravel()
    scale = 1.0 + np.arange(nel) / nel
    op = make(nmf, shape, nq0 - 2, nq0, nel)
    y = apply(op, (scale[:, None] * fin[None, :]).ravel(), False)
    assert G.rel_err(y, (scale[:, None] * fout[None, :]).ravel()) <= 1e-12


def test_default_factor_and_host_apply(nmf):
    nm, nq0, nel = 6, 8, 20
    op = make(nmf, O.TET, nm, nq0, nel)
    assert op.output_size() == nel * 12 * 11 * 11  # factor 1.5 by default
    u = G.coeffs(op.input_size(), 7)
    y = op.apply_host(u)
    assert G.rel_err(y, O.phys_interp1d_scaled(O.TET, nq0, nel, 1.5, u)) <= 1e-14


def test_rejects_what_does_not_fit(nmf):
    op = make(nmf, O.HEX, 10, 12, 4)
    with pytest.raises(nmf.NektarError):
        op.SetScalingFactor(3.0)
    with pytest.raises(nmf.NektarError):
        op.SetScalingFactor(0.0)
    # the operator keeps its previous, valid factor
    assert op.output_size() == 4 * 18 ** 3
