"""The oracle's static-condensation diagonal (or_sc_diagonal,
Preconditioner.cpp:185-236 and PreconditionerDiagonal.cpp:132-154) pinned on
the CPU:

* the interior-mode flags are exactly the modes that vanish on every face of
  the element (hex: brute-force Modified_A values at z = +-1; tet: the
  trace classification of tests/tet_mesh.py);
* per element, the diagonal equals that of A_bb - A_bi A_ii^-1 A_ib formed
  with numpy's dense solve, and the Schur complement's inverse is the
  boundary block of A^-1 (block elimination);
* assembled, the condensed system S x_b = f_b - A_bi A_ii^-1 f_i reproduces
  the boundary values of the full solve, which fails if a shared mode is
  taken for an interior one; its diagonal is the oracle's on the free
  boundary rows, zero on Dirichlet and interior rows."""
import numpy as np
import pytest

import _oracle as O
import brute as B
import geometry_cases as G
import tet_mesh as TM
from test_oracle_cg_loc import _blocks, _global, _hex_case, _tet_case


@pytest.mark.parametrize("nm", [2, 3, 4, 5, 6, 7])
def test_interior_flags_are_the_modes_vanishing_on_every_face(nm):
    fh = O.interior_modes(O.HEX, nm)
    assert fh.sum() == max(nm - 2, 0) ** 3
    vals = np.array([[B.modified_a(p)(z) for z in (-1.0, 1.0)] for p in range(nm)])
    on_face = np.abs(vals).max(axis=1) > 1e-14  # mode p is non-zero at some end
    r, q, p = np.meshgrid(np.arange(nm), np.arange(nm), np.arange(nm), indexing="ij")
    assert np.array_equal(fh, ~(on_face[p] | on_face[q] | on_face[r]).ravel())
    ft = O.interior_modes(O.TET, nm)
    assert len(ft) == len(B.tet_modes(nm))
    assert ft.sum() == (nm - 2) * (nm - 3) * (nm - 4) // 6
    if nm >= 2:
        C = TM.mode_polynomials(nm)
        V = TM.REF
        vanish = np.ones(len(ft), dtype=bool)
        for face in ([0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]):
            pts = TM._W_FACE @ V[face]
            vanish &= np.max(np.abs(TM._eval_modes(C, nm - 1, V, pts)), axis=1) <= 1e-10
        assert np.array_equal(ft, vanish)


def _schur(A, fl):
    i, b = np.flatnonzero(fl), np.flatnonzero(~fl)
    return A[np.ix_(b, b)] - A[np.ix_(b, i)] @ np.linalg.solve(A[np.ix_(i, i)], A[np.ix_(i, b)]), i, b


@pytest.mark.parametrize("shape,nm,nq,deformed", [(O.HEX, 5, 6, False), (O.HEX, 6, 7, True), (O.TET, 6, 7, False),
                                                   (O.TET, 8, 9, False)])
def test_single_element_schur_complement(shape, nm, nq, deformed):
    rng = np.random.default_rng(nm)
    if shape == O.HEX:
        # Nektar hex vertex order: counter-clockwise bottom face, then top
        V = np.array([[x, y, z] for z in (0, 1) for (x, y) in ((0, 0), (1, 0), (1, 1), (0, 1))], dtype=float)
        V = V + (0.1 * rng.uniform(-1, 1, V.shape) if deformed else 0.0)
        jac, df = O.hex_geom(np.ascontiguousarray(V.ravel()), 1, nq, deformed)
    else:
        V = TM.REF * 0.5 + 0.1 * rng.uniform(-1, 1, (4, 3))
        jac, df = O.tet_geom(np.ascontiguousarray(V.ravel()), 1)
    A = _blocks(shape, nm, nq, 1, deformed, jac, df, 1.7)[0]
    fl = O.interior_modes(shape, nm)
    S, i, b = _schur(A, fl)
    d = O.sc_diagonal([A], fl, np.arange(len(A)), None, len(A), 0)
    assert np.max(np.abs(d[b] - np.diag(S))) <= 1e-12 * np.max(np.abs(np.diag(S)))
    assert np.all(d[i] == 0.0)
    Ainv = np.linalg.inv(A)
    assert np.max(np.abs(np.linalg.inv(S) - Ainv[np.ix_(b, b)])) <= 1e-9 * np.max(np.abs(Ainv))


CASES = [("tet", (2, 5, 0.5)), ("tet", (2, 6, 1.0)), ("hex", (2, 5, 6, 2.0, 0.0, 4)), ("hex", (2, 6, 7, 1.0, 0.12, 9))]


@pytest.mark.parametrize("kind,args", CASES)
def test_condensed_system_reproduces_the_full_solve(kind, args):
    blocks, mp, sg, dirichlet, bl = _tet_case(*args) if kind == "tet" else _hex_case(*args)
    m, ndir, A, b, invd = _global(blocks, mp, sg, dirichlet, bl)
    ng = len(b)
    shape, nm = (O.TET, args[1]) if kind == "tet" else (O.HEX, args[1])
    fl = np.tile(O.interior_modes(shape, nm), len(blocks))
    d = O.sc_diagonal(blocks, fl, m, sg, ng, ndir)
    # global interior DOFs: held by exactly one local entry, an interior one
    gi = np.zeros(ng, dtype=bool)
    gi[m[fl]] = True
    cnt = np.bincount(m, minlength=ng)
    assert np.all(cnt[gi] == 1)
    gb = ~gi
    # assembled Schur complement and condensed right-hand side, element by element
    S = np.zeros((ng, ng))
    fc = np.zeros(ng)
    off = 0
    for Ae in blocks:
        k = len(Ae)
        idx, s, f = m[off:off + k], sg[off:off + k], fl[off:off + k]
        Se, ii, bb = _schur(s[:, None] * Ae * s[None, :], f)
        S[np.ix_(idx[bb], idx[bb])] += Se
        be = np.zeros(k)
        np.add.at(be, np.arange(k), s * bl[off:off + k])
        fc[idx[bb]] += be[bb] - (s[:, None] * Ae * s[None, :])[np.ix_(bb, ii)] @ np.linalg.solve(
            (s[:, None] * Ae * s[None, :])[np.ix_(ii, ii)], be[ii])
        off += k
    free_b = np.flatnonzero(gb & (np.arange(ng) >= ndir))
    xb = np.linalg.solve(S[np.ix_(free_b, free_b)], fc[free_b])
    xd = np.zeros(ng)
    xd[ndir:] = np.linalg.solve(A[ndir:, ndir:], b[ndir:])
    assert np.max(np.abs(xb - xd[free_b])) <= 1e-10 * np.max(np.abs(xd))
    # the diagonal of the assembled Schur complement on the free boundary rows
    assert np.max(np.abs(d[free_b] - np.diag(S)[free_b])) <= 1e-12 * np.max(np.abs(np.diag(S)))
    assert np.all(d[:ndir] == 0.0) and np.all(d[gi] == 0.0)
    # and larger than zero, at most the full diagonal (S_kk = A_kk - |L^-1 a_k|^2)
    assert np.all(d[free_b] > 0.0) and np.all(d[free_b] <= np.diag(A)[free_b] * (1 + 1e-12))
