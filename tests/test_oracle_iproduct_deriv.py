"""Pins of the oracle's IProductWRTDerivBase (out_p = sum_j int dphi_p/dx_j f_j,
IProductWRTDerivBase.h:592-825).

* hex: dense brute force from independently built Modified_A polynomials
  (tests/brute.py), dphi/dx_j = sum_d df[j*3+d] dphi/dxi_d;
* hex and tet: IProductWRTDerivBase(PhysDeriv(BwdTrans x)) equals the
  stiffness part of the Helmholtz operator (lambda = 0), an independent code
  path already pinned by brute force -- this fixes the df index convention
  and the collapsed-coordinate factors of the tet; it holds for any J and
  df arrays, so the perturbed per-point factors of the deformed cases
  exercise it too;
* linearity in each of the three inputs."""
import numpy as np
import pytest

import _oracle as O
import brute as B
import geometry_cases as G


def _elem(a, e, n):
    return a[e * n:(e + 1) * n]


@pytest.mark.parametrize("deformed", [False, True])
@pytest.mark.parametrize("nm,nq", [(2, 3), (4, 5), (5, 7), (7, 8)])
def test_hex_bruteforce(deformed, nm, nq):
    n = 2
    verts = O.hex_mesh(n, 0.2, 3) if deformed else G.affine_hex_mesh(n, 3)
    jac, df = O.hex_geom(verts, n ** 3, nq, deformed)
    nqt = nq ** 3
    f = [G.coeffs(n ** 3 * nqt, s) for s in (11, 12, 13)]
    out = O.iproduct_deriv(O.HEX, nm, nq, n ** 3, deformed, jac, df, *f)
    _, Dd, W = B.hex_dense(nm, nq)
    g = nqt if deformed else 1
    ref = []
    for e in range(n ** 3):
        je = np.broadcast_to(_elem(jac, e, g), (nqt,))
        dfe = np.broadcast_to(_elem(df, e, 9 * g).reshape(-1, 9), (nqt, 9))
        acc = np.zeros(nm ** 3)
        for d in range(3):
            t = sum(dfe[:, j * 3 + d] * _elem(f[j], e, nqt) for j in range(3))
            acc += Dd[d].T @ (W * je * t)
        ref.append(acc)
    assert G.rel_err(out, np.concatenate(ref)) < 1e-13


def _geometry(shape, nm, nq0, deformed, nel):
    if shape == O.HEX:
        n = round(nel ** (1 / 3))
        _, jac, df = G.hex_geometry(n, nq0, deformed)
        return n ** 3, jac, df
    _, jac, df = G.tet_geometry(nel, nq0, deformed)
    return nel, jac, df


@pytest.mark.parametrize("shape", [O.HEX, O.TET], ids=["hex", "tet"])
@pytest.mark.parametrize("deformed", [False, True], ids=["regular", "deformed"])
@pytest.mark.parametrize("nm", [2, 4, 6, 9])
def test_composition_with_physderiv_is_the_stiffness(shape, deformed, nm):
    nq0 = nm + 1
    nel, jac, df = _geometry(shape, nm, nq0, deformed, 8)
    x = G.coeffs(nel * O.ncoeffs(shape, nm), 5)
    u = O.bwdtrans(shape, nm, nq0, nel, x)
    grad = O.physderiv(shape, nm, nq0, nel, deformed, df, u)
    out = O.iproduct_deriv(shape, nm, nq0, nel, deformed, jac, df, *grad)
    stiff = O.helmholtz(shape, nm, nq0, nel, deformed, jac, df, 0.0, x)
    assert G.rel_err(out, stiff) < 1e-12


@pytest.mark.parametrize("shape", [O.HEX, O.TET], ids=["hex", "tet"])
def test_linear_in_each_input(shape):
    nm, nq0 = 5, 6
    nel, jac, df = _geometry(shape, nm, nq0, True, 8)
    n = nel * O.nqtot(shape, nq0)
    f = [G.coeffs(n, s) for s in (1, 2, 3)]
    zero = np.zeros(n)
    total = O.iproduct_deriv(shape, nm, nq0, nel, True, jac, df, *f)
    parts = [O.iproduct_deriv(shape, nm, nq0, nel, True, jac, df,
                              *[f[j] if j == i else zero for j in range(3)]) for i in range(3)]
    assert G.rel_err(parts[0] + parts[1] + parts[2], total) < 1e-13
    scaled = O.iproduct_deriv(shape, nm, nq0, nel, True, jac, df, *[2.5 * a for a in f])
    assert G.rel_err(scaled, 2.5 * total) < 1e-14
