"""Seeded geometry and input generators shared by the parity tests.

Hex geometry comes from the oracle's restatement of GeomFactors for
trilinear maps; tet geometry from affine maps.  'Deformed' tet factors are
per-point perturbations of affine ones (any per-point field exercises the
operator's deformed code path; the operator does not care how it arose)."""
from __future__ import annotations

import numpy as np

import _oracle as O

# Vertex set of the reference's TestHexHelmholtz_MatrixFree_UniformP
# (library/UnitTests/Collections/TestHexCollection.cpp:4336-4351): the
# bi-unit cube with vertex 6 moved to (2, 3, 4), a deformed hex.
REF_TEST_HEX = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                         [-1, -1, 1], [1, -1, 1], [2, 3, 4], [-1, 1, 1]], dtype=np.float64)


def affine_hex_mesh(n, seed):
    """n^3 structured mesh mapped by a random (well-conditioned) affine map:
    parallelepipeds, i.e. 'regular' elements with a full metric tensor."""
    rng = np.random.default_rng(seed)
    A = np.eye(3) + 0.3 * rng.uniform(-1, 1, (3, 3))
    b = rng.uniform(-1, 1, 3)
    v = O.hex_mesh(n, 0.0, seed).reshape(-1, 8, 3)
    return (v @ A.T + b).reshape(-1)


def hex_geometry(nelmt_side, nq, deformed, seed=7, affine=True):
    n = nelmt_side
    if deformed:
        verts = O.hex_mesh(n, 0.2, seed)
    else:
        verts = affine_hex_mesh(n, seed) if affine else O.hex_mesh(n, 0.0, seed)
    jac, df = O.hex_geom(verts, n ** 3, nq, deformed)
    return verts, jac, df


def ref_test_hex_geometry(nelmt, nq):
    verts = np.tile(REF_TEST_HEX.reshape(-1), nelmt)
    jac, df = O.hex_geom(verts, nelmt, nq, True)
    return verts, jac, df


def random_tets(nelmt, seed, scale_range=(0.5, 2.0)):
    """Affine tets: a reference-shaped tet, randomly scaled, sheared and
    translated (positive orientation)."""
    rng = np.random.default_rng(seed)
    base = np.array([[0, 0, 0], [1, 0, 0], [0.2, 1, 0], [0.3, 0.25, 1.0]])
    out = np.empty((nelmt, 4, 3))
    for e in range(nelmt):
        s = rng.uniform(*scale_range)
        A = np.eye(3) + 0.25 * rng.uniform(-1, 1, (3, 3))
        if np.linalg.det(A) < 0:
            A[:, 0] *= -1
            A[:, 1] *= 1
        t = rng.uniform(-5, 5, 3)
        v = s * (base @ A.T) + t
        # keep orientation positive
        if np.linalg.det(np.stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]])) < 0:
            v[[1, 2]] = v[[2, 1]]
        out[e] = v
    return out.reshape(-1)


def tet_geometry(nelmt, nq0, deformed, seed=11):
    verts = random_tets(nelmt, seed)
    jac, df = O.tet_geom(verts, nelmt)
    if not deformed:
        return verts, jac, df
    nqt = O.nqtot(O.TET, nq0)
    rng = np.random.default_rng(seed + 1)
    jd = np.repeat(jac, nqt) * (1.0 + 0.1 * rng.uniform(-1, 1, nelmt * nqt))
    dd = np.repeat(df.reshape(nelmt, 1, 9), nqt, axis=1)
    dd = dd * (1.0 + 0.1 * rng.uniform(-1, 1, dd.shape))
    return verts, jd, dd.reshape(-1)


def coeffs(n, seed):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n)


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
