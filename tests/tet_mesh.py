"""Unstructured tet meshes and their C0 assembly maps, for the general
assembly tests (test infrastructure).

Mesh: each cube of an n^3 grid on [0,1]^3 split into the 6 tets of the Kuhn
subdivision (monotone vertex paths from the cube's low corner to its high
corner), every tet's vertices in ascending global order.

Map: built from the basis functions themselves rather than from a table of
orientation conventions.  Every local mode (tests/brute.py, independently
built Modified_A/B/C tet modes with the singular-vertex terms, refitted as
polynomials of total degree <= P in the reference coordinates so that they
can be evaluated on the singular edges of the collapse) is evaluated on the
vertices, edges and faces of its element; it belongs to the
lowest-dimensional entity on which it does not vanish.  The modes of a
shared entity are matched between elements by their traces at sample
points given in physical space (the same points for every element): a mode
takes the global DOF of the mode with the same trace, with sign +1, or the
opposite trace, with sign -1.  A mode with no such partner raises, so the
builder cannot return a map that is not conforming mode by mode; the tests
then check continuity of assembled fields across every shared face and the
spectral convergence of a manufactured solve."""
import functools
import itertools

import numpy as np

import brute as B

# reference tet (Nektar): v0 (-1,-1,-1), v1 (1,-1,-1), v2 (-1,1,-1), v3 (-1,-1,1)
REF = np.array([[-1.0, -1, -1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]])


def kuhn_tets(n):
    """Vertex coordinates and the 6 n^3 tets (global vertex ids, ascending)."""
    nv = n + 1
    g = np.arange(nv)
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    coords = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1) / n
    vid = lambda i, j, k: (k * nv + j) * nv + i
    tets = []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                for perm in itertools.permutations(range(3)):
                    c = [i, j, k]
                    path = [vid(*c)]
                    for a in perm:
                        c[a] += 1
                        path.append(vid(*c))
                    tets.append(sorted(path))
    return coords, np.array(tets, dtype=np.int64)


def collapsed(xi):
    """Reference coordinates -> collapsed (a, b, c) (Nektar's tet collapse)."""
    x0, x1, x2 = xi[..., 0], xi[..., 1], xi[..., 2]
    c = x2
    b = 2 * (1 + x1) / (1 - x2) - 1
    a = 2 * (1 + x0) / (-x1 - x2) - 1
    return a, b, c


def ref_from_phys(verts, x):
    """Reference coordinates of physical points x in the affine tet verts."""
    A = np.stack([verts[1] - verts[0], verts[2] - verts[0], verts[3] - verts[0]], axis=1)  # x = v0 + A l
    lam = np.linalg.solve(A, (x - verts[0]).T).T  # barycentric l1, l2, l3
    return 2 * lam - 1  # xi_d = 2 l_d - 1


def _monomials(xi, P):
    """xi0^i xi1^j xi2^k, i + j + k <= P, at points xi [npts, 3]."""
    ex = [(i, j, k) for i in range(P + 1) for j in range(P + 1 - i) for k in range(P + 1 - i - j)]
    return np.stack([xi[:, 0] ** i * xi[:, 1] ** j * xi[:, 2] ** k for i, j, k in ex], axis=1)


@functools.lru_cache(maxsize=None)
def mode_polynomials(nm):
    """Every tet mode as a polynomial of total degree <= P = nm-1 in the
    reference coordinates xi (the modes span exactly that space): monomial
    coefficients fitted at interior Duffy points, where the collapsed
    coordinates are regular; the fit residual is checked.  Evaluating the
    fit reaches vertices and the singular edges of the collapse."""
    P = nm - 1
    fns = [B.tet_basis_fn(nm, m) for m in range(len(B.tet_modes(nm)))]
    a, b, c, _ = B.duffy_points(P + 4)
    x2 = c
    x1 = (1 + b) * (1 - c) / 2 - 1
    x0 = (1 + a) * (-x1 - x2) / 2 - 1
    xi = np.stack([x0, x1, x2], axis=1)
    M = _monomials(xi, P)
    vals = np.array([[f(ai, bi, ci)[0] for ai, bi, ci in zip(a, b, c)] for f in fns]).T  # [pt, mode]
    C, *_ = np.linalg.lstsq(M, vals, rcond=None)
    res = np.max(np.abs(M @ C - vals))
    assert res < 1e-11, f"tet modes are not reproduced by degree-{P} polynomials (residual {res})"
    return C  # [monomial, mode]


def _eval_modes(C, P, verts, pts):
    """Values of every mode at physical points of the affine tet verts."""
    xi = ref_from_phys(verts, pts)
    return (_monomials(xi, P) @ C).T  # [mode, point]


def _sample_points(kind, corners, rng_pts):
    """Physical sample points on an entity given by its corner coordinates
    (ordered by global vertex id), inside the entity and away from its
    boundary (barycentric weights in a fixed list)."""
    if kind == 0:
        return corners[:1].copy()
    return rng_pts[kind] @ corners


# fixed barycentric sample weights for edges (2 corners) and faces (3)
_W_EDGE = np.array([[0.83, 0.17], [0.61, 0.39], [0.47, 0.53], [0.29, 0.71], [0.12, 0.88], [0.71, 0.29],
                    [0.37, 0.63], [0.93, 0.07], [0.21, 0.79], [0.55, 0.45], [0.05, 0.95], [0.66, 0.34]])
_W_FACE = np.array([[0.6, 0.3, 0.1], [0.2, 0.5, 0.3], [0.15, 0.25, 0.6], [0.45, 0.45, 0.1], [0.1, 0.7, 0.2],
                    [0.33, 0.2, 0.47], [0.7, 0.1, 0.2], [0.25, 0.35, 0.4], [0.05, 0.15, 0.8], [0.5, 0.1, 0.4],
                    [0.4, 0.55, 0.05], [0.12, 0.43, 0.45], [0.08, 0.84, 0.08], [0.26, 0.64, 0.1],
                    [0.57, 0.32, 0.11], [0.19, 0.08, 0.73], [0.36, 0.41, 0.23], [0.81, 0.07, 0.12],
                    [0.03, 0.47, 0.5], [0.29, 0.14, 0.57], [0.48, 0.24, 0.28], [0.11, 0.27, 0.62],
                    [0.64, 0.22, 0.14], [0.22, 0.72, 0.06], [0.44, 0.06, 0.5], [0.09, 0.61, 0.3],
                    [0.52, 0.37, 0.11], [0.31, 0.31, 0.38], [0.74, 0.18, 0.08], [0.16, 0.16, 0.68]])
_W = {1: _W_EDGE, 2: _W_FACE}


def tet_c0_map(n, nm):
    """(coords, tets, map, sign, dirichlet): map and sign of every local
    coefficient (element-major, tet mode order) in a global numbering (not
    yet Dirichlet-first), dirichlet[g] for every global DOF."""
    coords, tets = kuhn_tets(n)
    modes = B.tet_modes(nm)
    nc = len(modes)
    C = mode_polynomials(nm)
    entities = {}  # key -> (list of trace vectors of its global DOFs, first global id)
    nglob = 0
    mp = np.zeros((len(tets), nc), dtype=np.int64)
    sg = np.zeros((len(tets), nc))
    boundary_dofs = []
    for e, tv in enumerate(tets):
        V = coords[tv]
        # entities of this tet: vertices, edges, faces (corner index tuples, ascending ids)
        ents = [(0, (i,)) for i in range(4)] + [(1, c) for c in itertools.combinations(range(4), 2)] + \
               [(2, c) for c in itertools.combinations(range(4), 3)]
        # classify every mode: values on each entity's samples
        vals = {}
        for kind, c in ents:
            pts = _sample_points(kind, V[list(c)], _W)
            vals[(kind, c)] = _eval_modes(C, nm - 1, V, pts)
        owner = [None] * nc
        for m in range(nc):
            for kind, c in ents:  # lowest dimension first
                if np.max(np.abs(vals[(kind, c)][m])) > 1e-8:
                    owner[m] = (kind, c)
                    break
        # interior modes get fresh global ids
        for m in range(nc):
            if owner[m] is None:
                mp[e, m], sg[e, m] = nglob, 1.0
                nglob += 1
        # boundary modes: match traces on the entity
        for kind, c in ents:
            ms = [m for m in range(nc) if owner[m] == (kind, c)]
            if not ms:
                continue
            key = (kind, tuple(int(tv[i]) for i in c))
            T = vals[(kind, c)][ms]
            if key not in entities:
                ids = list(range(nglob, nglob + len(ms)))
                nglob += len(ms)
                entities[key] = (T.copy(), ids)
                for m, gid in zip(ms, ids):
                    mp[e, m], sg[e, m] = gid, 1.0
                cc = coords[list(key[1])]
                face_of_box = any(np.all(cc[:, d] == 0) or np.all(cc[:, d] == 1) for d in range(3))
                if face_of_box:
                    boundary_dofs.extend(ids)
                continue
            Tref, ids = entities[key]
            for row, m in enumerate(ms):
                t = T[row]
                scale = np.max(np.abs(t))
                hit = None
                for j in range(len(ids)):
                    if np.max(np.abs(t - Tref[j])) <= 1e-8 * scale:
                        hit = (ids[j], 1.0)
                    elif np.max(np.abs(t + Tref[j])) <= 1e-8 * scale:
                        hit = (ids[j], -1.0)
                    if hit:
                        break
                if hit is None:
                    raise ValueError(f"mode {modes[m]} of tet {e} matches no mode of entity {key} up to sign")
                mp[e, m], sg[e, m] = hit
    dirichlet = np.zeros(nglob, dtype=bool)
    dirichlet[boundary_dofs] = True
    return coords, tets, mp.reshape(-1), sg.reshape(-1), dirichlet


def tet_quadrature_coords(coords, tets, nq0):
    """Physical coordinates of every tet's quadrature points (i fastest over
    the collapsed grid: GLL nq0, Gauss-Radau(1,0) and (2,0) nq0-1)."""
    import _oracle as O
    za = O.points(O.GLL, nq0)[0]
    zb = O.points(O.RADAU_A1, nq0 - 1)[0]
    zc = O.points(O.RADAU_A2, nq0 - 1)[0]
    C, Bb, A = np.meshgrid(zc, zb, za, indexing="ij")
    a, b, c = A.ravel(), Bb.ravel(), C.ravel()
    x2 = c
    x1 = (1 + b) * (1 - c) / 2 - 1
    x0 = (1 + a) * (-x1 - x2) / 2 - 1
    xi = np.stack([x0, x1, x2], axis=1)
    lam = (xi + 1) / 2
    out = []
    for tv in tets:
        V = coords[tv]
        out.append(V[0] + lam @ np.stack([V[1] - V[0], V[2] - V[0], V[3] - V[0]], axis=0))
    return np.array(out)  # [elmt, q, 3]


def tet_vertices(coords, tets):
    return np.ascontiguousarray(coords[tets].reshape(-1))
