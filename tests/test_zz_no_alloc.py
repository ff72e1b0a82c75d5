"""nmf_operator_apply_physderiv_dir_noalloc: the single-direction PhysDeriv
with the two other directions in the library's static device scratch
instead of an allocated workspace.  Called through the C ABI and compared
bit for bit with nmf_operator_apply_physderiv_dir (the allocating form), on
two calls in a row (the second with fewer elements), in FP64 and FP32, plus
its size limit and its geometry-upload precondition."""
import ctypes

import numpy as np
import pytest

import _oracle as O
import geometry_cases as G

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SCRATCH_BYTES = 64 << 20
_P = ctypes.c_void_p


@pytest.fixture(scope="module")
def nmf():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import nektar_amd
    lib = nektar_amd.lib()
    fn = lib.nmf_operator_apply_physderiv_dir_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [_P, ctypes.c_int, _P, _P, _P]
    return nektar_amd


def _noalloc(nmf, op, d, x, y):
    return nmf.lib().nmf_operator_apply_physderiv_dir_noalloc(op._h, d, x.data_ptr(), y.data_ptr(), None)


def _alloc(nmf, op, d, x, y):
    return nmf.lib().nmf_operator_apply_physderiv_dir(op._h, d, x.data_ptr(), y.data_ptr(), None, None)


def _op(nmf, n, nm, nq, prec, seed):
    E = n ** 3
    verts = O.hex_mesh(n, 0.2, seed)
    jac, df = O.hex_geom(verts, E, nq, True)
    op = nmf.GetOperatorFactory().CreateInstance("PhysDeriv_Hex_Deformed", nm, nq, E, prec)
    op.SetDF(df)
    return op, E * nq ** 3


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_matches_allocating_form_twice(nmf, prec):
    dt = torch.float64 if prec == "fp64" else torch.float32
    p = nmf.FP64 if prec == "fp64" else nmf.FP32
    # first call: 6^3 elements at P=6, Q=8; second: 3^3 elements, P=4, Q=6
    for call, (n, nm, nq) in enumerate([(6, 7, 8), (3, 5, 6)]):
        op, size = _op(nmf, n, nm, nq, p, 10 + call)
        x = torch.from_numpy(G.coeffs(size, 20 + call)).to(dt).cuda()
        warm = [torch.empty(size, dtype=dt, device="cuda") for _ in range(3)]
        op(x, warm[0], warm[1], warm[2])  # uploads the geometry
        for d in range(3):
            ref = torch.full((size,), float("nan"), dtype=dt, device="cuda")
            got = torch.full((size,), float("nan"), dtype=dt, device="cuda")
            assert _alloc(nmf, op, d, x, ref) == 0, nmf.lib().nmf_last_error()
            assert _noalloc(nmf, op, d, x, got) == 0, nmf.lib().nmf_last_error()
            torch.cuda.synchronize()
            assert torch.equal(got, ref), (prec, call, d)
            assert torch.equal(got, warm[d]), (prec, call, d)
        print(f"{prec} call {call}: {n ** 3} elements P={nm - 1} Q={nq}: bitwise equal in 3 directions", flush=True)


def test_noalloc_rejects_oversize_and_pending_geometry(nmf):
    # 2 * nelmt * nqtot doubles above the 64 MiB scratch: refused before any launch
    E = SCRATCH_BYTES // (2 * 8 * 512) + 1
    big = nmf.GetOperatorFactory().CreateInstance("PhysDeriv_Hex_Regular", 7, 8, E, nmf.FP64)
    x = torch.zeros(1, dtype=torch.float64, device="cuda")
    rc = _noalloc(nmf, big, 0, x, x)
    assert rc == 2, rc  # NMF_ERR_UNSUPPORTED
    assert b"static scratch" in nmf.lib().nmf_last_error()
    # geometry set after the last apply: the upload would allocate, so refused
    op, size = _op(nmf, 2, 5, 6, nmf.FP64, 3)
    xs = torch.zeros(size, dtype=torch.float64, device="cuda")
    rc = _noalloc(nmf, op, 1, xs, torch.empty_like(xs))
    assert rc == 3, rc  # NMF_ERR_STATE
    bad = _noalloc(nmf, op, 3, xs, torch.empty_like(xs))
    assert bad == 1, bad  # NMF_ERR_INVALID: direction out of range
