"""Pins of the unstructured tet mesh and its C0 map (tests/tet_mesh.py),
on the CPU with the oracle:

* the global dimension is that of the continuous degree-P space on the mesh,
  (P n + 1)^3 for the Kuhn-subdivided n^3 cube grid;
* fields assembled from random global coefficients are continuous: both
  elements of every shared face give the same values at common points;
* the Helmholtz system assembled from the oracle's tet element matrices is
  symmetric positive definite on the non-Dirichlet block, and the
  manufactured solution converges spectrally in P."""
import itertools

import numpy as np
import pytest

import _oracle as O
import brute as B
import tet_mesh as T


@pytest.mark.parametrize("nm", [2, 3, 4, 5])
def test_global_dimension(nm):
    n = 2
    _, tets, mp, sg, dirichlet = T.tet_c0_map(n, nm)
    assert mp.max() + 1 == ((nm - 1) * n + 1) ** 3
    assert len(dirichlet) == mp.max() + 1
    assert np.count_nonzero(~dirichlet) == ((nm - 1) * n - 1) ** 3
    assert set(np.unique(sg)) <= {-1.0, 1.0}


@pytest.mark.parametrize("nm", [3, 5])
def test_assembled_fields_are_continuous(nm):
    n = 2
    coords, tets, mp, sg, _ = T.tet_c0_map(n, nm)
    nc = len(B.tet_modes(nm))
    C = T.mode_polynomials(nm)
    x = np.random.default_rng(nm).uniform(-1, 1, mp.max() + 1)
    loc = (sg * x[mp]).reshape(len(tets), nc)
    faces = {}
    for e, tv in enumerate(tets):
        for c in itertools.combinations(range(4), 3):
            faces.setdefault(tuple(int(tv[i]) for i in c), []).append(e)
    W = np.array([[0.6, 0.3, 0.1], [0.2, 0.5, 0.3], [0.15, 0.25, 0.6], [1 / 3, 1 / 3, 1 / 3], [0.9, 0.05, 0.05]])
    shared = 0
    for key, els in faces.items():
        if len(els) != 2:
            continue
        pts = W @ coords[list(key)]
        v = [(T._eval_modes(C, nm - 1, coords[tets[e]], pts).T @ loc[e]) for e in els]
        assert np.max(np.abs(v[0] - v[1])) <= 1e-10 * max(1.0, np.max(np.abs(v[0])))
        shared += 1
    assert all(len(els) <= 2 for els in faces.values())
    # 4 faces per tet; each of the 6 faces of the box carries 2 n^2 boundary faces
    assert shared == (4 * len(tets) - 6 * 2 * n * n) // 2


def _dense_system(n, nm, lam):
    coords, tets, mp, sg, dirichlet = T.tet_c0_map(n, nm)
    nel, nc = len(tets), len(B.tet_modes(nm))
    jac, df = O.tet_geom(T.tet_vertices(coords, tets), nel)
    ng = mp.max() + 1
    A = np.zeros((ng, ng))
    M = mp.reshape(nel, nc)
    S = sg.reshape(nel, nc)
    for m in range(nc):
        u = np.zeros(nel * nc)
        u[m::nc] = 1.0
        col = O.helmholtz(O.TET, nm, nm + 1, nel, False, jac, df, lam, u).reshape(nel, nc)  # A_e[:, m]
        for e in range(nel):
            A[np.ix_(M[e], [M[e, m]])] += (S[e] * S[e, m] * col[e])[:, None]
    return coords, tets, mp, sg, dirichlet, jac, A


def _l2_error(coords, tets, nm, loc, k):
    C = T.mode_polynomials(nm)
    a, b, c, w = B.duffy_points(nm + 5)
    x2 = c
    x1 = (1 + b) * (1 - c) / 2 - 1
    x0 = (1 + a) * (-x1 - x2) / 2 - 1
    lam = (np.stack([x0, x1, x2], axis=1) + 1) / 2
    err = 0.0
    Mono = T._monomials(np.stack([x0, x1, x2], axis=1), nm - 1)
    for e, tv in enumerate(tets):
        V = coords[tv]
        Aff = np.stack([V[1] - V[0], V[2] - V[0], V[3] - V[0]], axis=0)
        X = V[0] + lam @ Aff
        uh = Mono @ C @ loc[e]
        ue = np.prod(np.sin(k * np.pi * X), axis=1)
        err += abs(np.linalg.det(Aff)) / 8.0 * np.sum(w * (uh - ue) ** 2)
    return np.sqrt(err)


def test_manufactured_solution_converges():
    n, lam, k = 2, 1.0, 1
    errs = []
    for nm in (2, 3, 4, 5):
        coords, tets, mp, sg, dirichlet, jac, A = _dense_system(n, nm, lam)
        nel, nc = len(tets), len(B.tet_modes(nm))
        free = ~dirichlet
        Af = A[np.ix_(free, free)]
        assert np.allclose(Af, Af.T, rtol=0, atol=1e-12 * np.abs(Af).max())
        assert np.linalg.eigvalsh(Af).min() > 0
        X = T.tet_quadrature_coords(coords, tets, nm + 1)
        f = (3 * (k * np.pi) ** 2 + lam) * np.prod(np.sin(k * np.pi * X), axis=2).ravel()
        bl = O.iproduct(O.TET, nm, nm + 1, nel, False, jac, f)
        b = np.zeros(len(dirichlet))
        np.add.at(b, mp, sg * bl)
        x = np.zeros(len(dirichlet))
        x[free] = np.linalg.solve(Af, b[free])
        loc = (sg * x[mp]).reshape(nel, nc)
        errs.append(_l2_error(coords, tets, nm, loc, k))
    assert all(e2 < 0.5 * e1 for e1, e2 in zip(errs, errs[1:])), errs
    assert errs[-1] < 2e-3, errs
