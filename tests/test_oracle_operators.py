"""Pins of the oracle's operators.

* dense brute force (numpy/scipy, tests/brute.py) of BwdTrans, IProduct,
  PhysDeriv and Helmholtz for hex (regular, deformed, constant diffusion)
  and tet (affine);
* identities the mathematics fixes: symmetry and positive definiteness of
  the Helmholtz matrix, L*1 = 0 (vertex modes are a partition of unity),
  IProduct of 1 = the basis integrals, exact gradients of linear fields."""
import numpy as np
import pytest

import _oracle as O
import brute as B
import geometry_cases as G

TOL = 1e-13


def _elem(a, e, n):
    return a[e * n:(e + 1) * n]


@pytest.mark.parametrize("deformed", [False, True])
@pytest.mark.parametrize("nm,nq", [(2, 3), (3, 4), (4, 5), (5, 7), (6, 7)])
def test_hex_helmholtz_bruteforce(deformed, nm, nq):
    n = 2
    verts = O.hex_mesh(n, 0.2, 3) if deformed else G.affine_hex_mesh(n, 3)
    jac, df = O.hex_geom(verts, n ** 3, nq, deformed)
    x = G.coeffs(n ** 3 * nm ** 3, nm)
    y = O.helmholtz(O.HEX, nm, nq, n ** 3, deformed, jac, df, 1.3, x)
    g = nq ** 3 if deformed else 1
    yb = np.concatenate([B.hex_helmholtz_dense(nm, nq, _elem(jac, e, g), _elem(df, e, 9 * g), 1.3)
                         @ _elem(x, e, nm ** 3) for e in range(n ** 3)])
    assert G.rel_err(y, yb) < TOL


def test_hex_helmholtz_const_var_diffusion_bruteforce():
    nm, nq, n = 4, 5, 2
    d6 = np.array([2.0, 0.3, 1.5, -0.2, 0.1, 1.2])
    verts = O.hex_mesh(n, 0.2, 5)
    jac, df = O.hex_geom(verts, n ** 3, nq, True)
    x = G.coeffs(n ** 3 * nm ** 3, 9)
    y = O.helmholtz(O.HEX, nm, nq, n ** 3, True, jac, df, 0.4, x, d6)
    g = nq ** 3
    yb = np.concatenate([B.hex_helmholtz_dense(nm, nq, _elem(jac, e, g), _elem(df, e, 9 * g), 0.4, d6)
                         @ _elem(x, e, nm ** 3) for e in range(n ** 3)])
    assert G.rel_err(y, yb) < TOL


@pytest.mark.parametrize("nm,nq", [(3, 4), (5, 6), (7, 8)])
def test_hex_bwd_iprod_deriv_bruteforce(nm, nq):
    n = 2
    verts = O.hex_mesh(n, 0.2, 4)
    jac, df = O.hex_geom(verts, n ** 3, nq, True)
    Phi, Dd, W = B.hex_dense(nm, nq)
    nqt = nq ** 3
    x = G.coeffs(n ** 3 * nm ** 3, 1)
    u = O.bwdtrans(O.HEX, nm, nq, n ** 3, x)
    assert G.rel_err(u, np.concatenate([Phi @ _elem(x, e, nm ** 3) for e in range(n ** 3)])) < TOL
    f = G.coeffs(n ** 3 * nqt, 2)
    ip = O.iproduct(O.HEX, nm, nq, n ** 3, True, jac, f)
    ipb = np.concatenate([Phi.T @ (W * _elem(jac, e, nqt) * _elem(f, e, nqt)) for e in range(n ** 3)])
    assert G.rel_err(ip, ipb) < TOL
    # PhysDeriv of the polynomial field u (exact differentiation by collocation)
    d = O.physderiv(O.HEX, nm, nq, n ** 3, True, df, u)
    for j in range(3):
        db = []
        for e in range(n ** 3):
            dfe = _elem(df, e, 9 * nqt).reshape(nqt, 9)
            xe = _elem(x, e, nm ** 3)
            db.append(sum(dfe[:, j * 3 + k] * (Dd[k] @ xe) for k in range(3)))
        assert G.rel_err(d[j], np.concatenate(db)) < 1e-12


@pytest.mark.parametrize("nm", [2, 3, 4, 5, 6])
def test_tet_helmholtz_bruteforce(nm):
    ne = 3
    verts = G.random_tets(ne, 5)
    jac, df = O.tet_geom(verts, ne)
    nc = O.ncoeffs(O.TET, nm)
    x = G.coeffs(ne * nc, 2)
    y = O.helmholtz(O.TET, nm, nm + 1, ne, False, jac, df, 0.8, x)
    yb = np.concatenate([B.tet_dense_mass_stiff(nm, jac[e], df[e * 9:(e + 1) * 9], 0.8)[0] @ _elem(x, e, nc)
                         for e in range(ne)])
    assert G.rel_err(y, yb) < TOL


@pytest.mark.parametrize("nm", [2, 4, 6])
def test_tet_bwdtrans_bruteforce(nm):
    nq0 = nm + 1
    z0, _, _ = O.points(O.GLL, nq0)
    z1, _, _ = O.points(O.RADAU_A1, nq0 - 1)
    z2, _, _ = O.points(O.RADAU_A2, nq0 - 1)
    C, Bb, A = np.meshgrid(z2, z1, z0, indexing="ij")  # i fastest
    nc = O.ncoeffs(O.TET, nm)
    x = G.coeffs(nc, 3)
    u = O.bwdtrans(O.TET, nm, nq0, 1, x)
    ub = sum(x[m] * B.tet_basis_fn(nm, m)(A.ravel(), Bb.ravel(), C.ravel())[0] for m in range(nc))
    assert G.rel_err(u, ub) < TOL


@pytest.mark.parametrize("shape", [O.HEX, O.TET])
@pytest.mark.parametrize("nm", [3, 5, 8])
def test_helmholtz_symmetric_positive_definite(shape, nm):
    nq0 = nm + 1
    if shape == O.HEX:
        _, jac, df = G.ref_test_hex_geometry(1, nq0)
        deformed = True
    else:
        verts = G.random_tets(1, 8)
        jac, df = O.tet_geom(verts, 1)
        deformed = False
    nc = O.ncoeffs(shape, nm)
    A = np.stack([O.helmholtz(shape, nm, nq0, 1, deformed, jac, df, 0.9, np.eye(nc)[m]) for m in range(nc)], 1)
    assert np.max(np.abs(A - A.T)) < 1e-12 * np.max(np.abs(A))
    assert np.min(np.linalg.eigvalsh(0.5 * (A + A.T))) > 0
    # brute-force diagonal helper agrees with the matrix diagonal
    dg = O.helmholtz_diag(shape, nm, nq0, 1, deformed, jac, df, 0.9)
    assert np.array_equal(dg, np.diag(A))


@pytest.mark.parametrize("shape", [O.HEX, O.TET])
def test_laplacian_annihilates_constants(shape):
    """Vertex modes sum to 1 (partition of unity), so L*1 = 0 and the
    Helmholtz operator maps 1 to lambda * (integrals of the basis)."""
    nm, nq0 = 5, 6
    if shape == O.HEX:
        _, jac, df = G.hex_geometry(1, nq0, True)
        deformed = True
        ones = np.zeros(nm ** 3)
        for r in (0, 1):
            for q in (0, 1):
                for p in (0, 1):
                    ones[(r * nm + q) * nm + p] = 1.0
    else:
        verts = G.random_tets(1, 3)
        jac, df = O.tet_geom(verts, 1)
        deformed = False
        modes = B.tet_modes(nm)
        # tet vertices: (0,0,0), (1,0,0) [p=1,q=0,r=0], (0,1,0), top (0,0,1)
        ones = np.array([1.0 if (m in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]) else 0.0 for m in modes])
    u = O.bwdtrans(shape, nm, nq0, 1, ones)
    assert np.max(np.abs(u - 1.0)) < 1e-14
    y0 = O.helmholtz(shape, nm, nq0, 1, deformed, jac, df, 0.0, ones)
    assert np.max(np.abs(y0)) < 1e-13
    y1 = O.helmholtz(shape, nm, nq0, 1, deformed, jac, df, 2.0, ones)
    nqt = O.nqtot(shape, nq0)
    ip1 = O.iproduct(shape, nm, nq0, 1, deformed, jac, np.ones(nqt))
    assert G.rel_err(y1, 2.0 * ip1) < 1e-13


def test_physderiv_of_coordinates_is_identity():
    """d x_j / d x_k = delta_jk on a deformed (trilinear) hex."""
    nm, nq, n = 3, 4, 2
    verts = O.hex_mesh(n, 0.25, 6)
    jac, df = O.hex_geom(verts, n ** 3, nq, True)
    xyz = O.hex_coords(verts, n ** 3, nq)
    for c in range(3):
        d = O.physderiv(O.HEX, nm, nq, n ** 3, True, df, np.ascontiguousarray(xyz[:, :, c].ravel()))
        for k in range(3):
            assert np.max(np.abs(d[k] - (1.0 if k == c else 0.0))) < 1e-12
