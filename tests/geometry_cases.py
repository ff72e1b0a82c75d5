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


# Nektar hex vertex order: corner signs along the three local axes
HEX_SIGN = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                     [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]])


def rotated_hex_mesh(n, nm, amp, seed):
    """The structured n^3 mesh of O.hex_mesh(n, amp, seed) with every element
    in a random local frame (axis permutation and reflections, 48 choices):
    local axis a runs along global axis perm[a], reversed where flip[a].
    Returns the element vertices in their local frames and, for every local
    coefficient, its global DOF in the structured numbering (O.hex_l2g) and
    its sign: a reflected axis swaps the two vertex modes and multiplies
    bubble mode p by (-1)^p (Modified_A: phi_0(-z) = phi_1(z),
    phi_p(-z) = (-1)^p phi_p(z) for p >= 2), the orientation signs an
    AssemblyMapCG holds for edges and faces met in opposite directions."""
    V = O.hex_mesh(n, amp, seed).reshape(-1, 8, 3)
    m, dirichlet = O.hex_l2g(n, nm)
    base = np.asarray(m).reshape(-1, nm, nm, nm)  # [e][r][q][p], global frame
    rng = np.random.default_rng(seed + 1000)
    corner = {tuple(sv): v for v, sv in enumerate(HEX_SIGN)}
    nel = V.shape[0]
    Vr = np.empty_like(V)
    mp = np.empty((nel, nm, nm, nm), dtype=np.int64)
    sg = np.empty((nel, nm, nm, nm))
    modes = np.arange(nm)
    for e in range(nel):
        perm, flip = rng.permutation(3), rng.integers(0, 2, 3)
        for v in range(8):
            g = np.empty(3, dtype=int)
            for a in range(3):
                g[perm[a]] = HEX_SIGN[v][a] * (-1 if flip[a] else 1)
            Vr[e, v] = V[e, corner[tuple(g)]]
        # per local axis: global-frame mode index and sign of every mode
        gm, gs = [], []
        for a in range(3):
            if flip[a]:
                gm.append(np.where(modes < 2, 1 - modes, modes))
                gs.append(np.where(modes < 2, 1.0, (-1.0) ** modes))
            else:
                gm.append(modes.copy())
                gs.append(np.ones(nm))
        P2, P1, P0 = np.meshgrid(modes, modes, modes, indexing="ij")  # local (p2, p1, p0)
        G = [None, None, None]
        G[perm[0]], G[perm[1]], G[perm[2]] = gm[0][P0], gm[1][P1], gm[2][P2]
        mp[e] = base[e][G[2], G[1], G[0]]
        sg[e] = gs[0][P0] * gs[1][P1] * gs[2][P2]
    return Vr.reshape(-1), mp.reshape(-1), sg.reshape(-1), np.asarray(dirichlet, dtype=bool)


def dirichlet_first(dirichlet):
    """Renumbering with the Dirichlet DOFs first (AssemblyMapCG's order):
    new[g] for every global DOF g, and the Dirichlet count."""
    order = np.concatenate([np.flatnonzero(dirichlet), np.flatnonzero(~dirichlet)])
    new = np.empty_like(order)
    new[order] = np.arange(len(order))
    return new, int(np.count_nonzero(dirichlet))
