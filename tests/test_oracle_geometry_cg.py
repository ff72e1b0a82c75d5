"""Pins of the oracle's geometric factors, assembly and CG.

Geometry: the inverse-metric factors must invert the finite-difference
Jacobian of the trilinear map; J must equal |det| of it; regular/deformed
classification follows the iso-parameter test.
Assembly/CG: the structured numbering gives a conforming C0 space (the
assembled operator is SPD on non-Dirichlet rows); the reference CG recurrence
agrees with a dense solve; the Helmholtz solve converges spectrally to a
manufactured solution."""
import numpy as np
import pytest

import _oracle as O
import geometry_cases as G


def _trilinear(V, xi):
    s = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)
    N = np.prod(1 + s * xi, axis=1) / 8.0
    return N @ V


def test_hex_factors_vs_finite_differences():
    n, nq = 2, 4
    verts = O.hex_mesh(n, 0.25, 9)
    jac, df = O.hex_geom(verts, n ** 3, nq, True)
    z, _, _ = O.points(O.GLL, nq)
    h = 1e-6
    for e in [0, 5]:
        V = verts[e * 24:(e + 1) * 24].reshape(8, 3)
        for q in [0, 13, nq ** 3 - 1]:
            i, j, k = q % nq, (q // nq) % nq, q // nq ** 2
            xi = np.array([z[i], z[j], z[k]])
            M = np.zeros((3, 3))  # M[k, i] = dx_k / dxi_i
            for d in range(3):
                dx = np.zeros(3)
                dx[d] = h
                M[:, d] = (_trilinear(V, xi + dx) - _trilinear(V, xi - dx)) / (2 * h)
            g = e * nq ** 3 + q
            assert abs(jac[g] - abs(np.linalg.det(M))) < 1e-8
            F = df[g * 9:(g + 1) * 9].reshape(3, 3)  # F[j, k] = dxi_k / dx_j
            assert np.max(np.abs(F.T @ M - np.eye(3))) < 1e-8


def test_regularity_classification():
    assert O.hex_is_regular(O.hex_mesh(1, 0.0, 1)[:24].reshape(8, 3))
    assert O.hex_is_regular(G.affine_hex_mesh(1, 4)[:24].reshape(8, 3))
    assert not O.hex_is_regular(G.REF_TEST_HEX)
    v = O.hex_mesh(3, 0.1, 2).reshape(27, 8, 3)
    assert not O.hex_is_regular(v[13])  # the centre element has perturbed vertices


def test_tet_factors():
    verts = G.random_tets(5, 1)
    jac, df = O.tet_geom(verts, 5)
    for e in range(5):
        V = verts[e * 12:(e + 1) * 12].reshape(4, 3)
        M = 0.5 * (V[1:] - V[0]).T  # dx_k / dxi_i
        assert abs(jac[e] - np.linalg.det(M)) < 1e-12
        assert np.max(np.abs(df[e * 9:(e + 1) * 9].reshape(3, 3).T @ M - np.eye(3))) < 1e-12


def test_mesh_perturbation_deterministic():
    a, b = O.hex_mesh(4, 0.1, 3), O.hex_mesh(4, 0.1, 3)
    assert np.array_equal(a, b)
    c = O.hex_mesh(4, 0.1, 4)
    assert not np.array_equal(a, c)
    v = a.reshape(-1, 3)
    assert v.min() == 0.0 and v.max() == 1.0


def _assembled_matrix(n, nm, nq, deformed, lam, amp=0.15):
    verts = O.hex_mesh(n, amp if deformed else 0.0, 3)
    jac, df = O.hex_geom(verts, n ** 3, nq, deformed)
    m, dirichlet = O.hex_l2g(n, nm)
    nglob = len(dirichlet)
    A = np.zeros((nglob, nglob))
    nc = nm ** 3
    for e in range(n ** 3):
        g = nq ** 3 if deformed else 1
        Ae = np.stack([O.helmholtz(O.HEX, nm, nq, 1, deformed, jac[e * g:(e + 1) * g],
                                   df[e * 9 * g:(e + 1) * 9 * g], lam, np.eye(nc)[c]) for c in range(nc)], 1)
        idx = m[e * nc:(e + 1) * nc]
        A[np.ix_(idx, idx)] += Ae
    return A, dirichlet, verts, jac, df


def test_assembled_operator_spd_and_conforming():
    A, dirichlet, *_ = _assembled_matrix(2, 4, 5, True, 1.0)
    nd = ~dirichlet
    Ai = A[np.ix_(nd, nd)]
    assert np.max(np.abs(Ai - Ai.T)) < 1e-12 * np.max(np.abs(Ai))
    assert np.min(np.linalg.eigvalsh(Ai)) > 0
    # conformity: the global constant reproduces lambda * integral of the
    # global basis (L annihilates the continuous constant)
    A0, *_ = _assembled_matrix(2, 4, 5, True, 0.0)
    nm, N = 4, 2 * 3 + 1
    g = np.arange(len(dirichlet))
    gx, gy, gz = g % N, (g // N) % N, g // (N * N)
    const = ((gx % (nm - 1) == 0) & (gy % (nm - 1) == 0) & (gz % (nm - 1) == 0)).astype(float)
    assert np.max(np.abs(A0 @ const)) < 1e-12


def test_cg_dense_matches_solve():
    rng = np.random.default_rng(0)
    n = 40
    Q = rng.standard_normal((n, n))
    A = Q @ Q.T + n * np.eye(n)
    b = rng.standard_normal(n)
    ndir = 3
    x, info = O.cg_dense(A, b, ndir=ndir, inv_diag=1.0 / np.diag(A), tol=1e-12)
    Ai = A[ndir:, ndir:]
    xs = np.linalg.solve(Ai, b[ndir:])
    assert np.max(np.abs(x[ndir:] - xs)) < 1e-9 * np.max(np.abs(xs))
    assert 1 <= info.iterations < 60
    assert info.final_error < 1e-12


def test_cg_helmholtz_matches_assembled_solve():
    n, nm, nq, lam = 2, 4, 5, 1.0
    A, dirichlet, verts, jac, df = _assembled_matrix(n, nm, nq, True, lam)
    rng = np.random.default_rng(1)
    b = rng.uniform(-1, 1, len(dirichlet))
    inv_diag = 1.0 / np.diag(A)
    x, info = O.cg_helmholtz_hex(n, nm, nq, True, jac, df, lam, b, inv_diag, 1e-12, 5000)
    nd = ~dirichlet
    xs = np.linalg.solve(A[np.ix_(nd, nd)], b[nd])
    assert np.max(np.abs(x[nd] - xs)) < 1e-9 * np.max(np.abs(xs))
    assert np.all(x[dirichlet] == 0.0)
    # reference stopping rule: ||r||^2 < tol^2 * ||b||^2 over all global rows
    assert info.rhs_magnitude == pytest.approx(float(b @ b))


def _manufactured_rhs(n, nm, nq, lam, k=1):
    """b = Assemble(IProduct(f)), f = (3 k^2 pi^2 + lam) prod sin(k pi x_i)."""
    verts = O.hex_mesh(n, 0.0, 1)
    jac, df = O.hex_geom(verts, n ** 3, nq, False)
    xyz = O.hex_coords(verts, n ** 3, nq)
    f = (3 * (k * np.pi) ** 2 + lam) * np.prod(np.sin(k * np.pi * xyz), axis=2).ravel()
    bl = O.iproduct(O.HEX, nm, nq, n ** 3, False, jac, f)
    m, dirichlet = O.hex_l2g(n, nm)
    b = np.zeros(len(dirichlet))
    np.add.at(b, m, bl)
    return b, jac, df, verts, m, dirichlet


@pytest.mark.parametrize("nm", [3, 5, 7])
def test_cg_manufactured_solution_converges(nm):
    n, nq, lam = 3, nm + 1, 1.0
    b, jac, df, verts, m, dirichlet = _manufactured_rhs(n, nm, nq, lam)
    # Jacobi diagonal from the brute-force local diagonals
    dl = O.helmholtz_diag(O.HEX, nm, nq, n ** 3, False, jac, df, lam)
    dg = np.zeros(len(dirichlet))
    np.add.at(dg, m, dl)
    x, info = O.cg_helmholtz_hex(n, nm, nq, False, jac, df, lam, b, 1.0 / dg, 1e-10, 5000)
    u = O.bwdtrans(O.HEX, nm, nq, n ** 3, x[m])
    xyz = O.hex_coords(verts, n ** 3, nq)
    exact = np.prod(np.sin(np.pi * xyz), axis=2).ravel()
    _, w, _ = O.points(O.GLL, nq)
    W = np.kron(w, np.kron(w, w))
    l2 = np.sqrt(np.sum(np.tile(W, n ** 3) * np.repeat(jac, nq ** 3) * (u - exact) ** 2))
    bound = {3: 5e-3, 5: 5e-5, 7: 2e-7}[nm]
    assert l2 < bound, (l2, info.iterations)
    assert info.final_error < 1e-10
