"""The oracle's element-local CG (or_cg_loc, NekLinSysIterCGLoc.cpp:114-308)
pinned on the CPU against the assembled-space CG (or_cg_dense, itself
pinned against dense solves) and against dense solves:

* the two recurrences are the same in exact arithmetic (every local dot
  equals the matching global one through <Assemble(a), b_glob> =
  <a, GlobalToLocal(b_glob)>), so they take the same iterations, give the
  same rhs magnitude and the same solution;
* the local solution is continuous: every copy of a global DOF holds the
  same value with its sign, exactly;
* with maxit reached, both stop at the same count;
* a zero right-hand side returns at once with x = 0.

Meshes: unstructured Kuhn tet meshes (tests/tet_mesh.py) and the hex mesh in
random element frames (tests/geometry_cases.py), both with signed maps."""
import numpy as np
import pytest

import _oracle as O
import brute as B
import geometry_cases as G
import tet_mesh as TM


def _blocks(shape, nm, nq, nel, deformed, jac, df, lam):
    nc = O.ncoeffs(shape, nm)
    cols = []
    for m in range(nc):
        u = np.zeros(nel * nc)
        u[m::nc] = 1.0
        cols.append(O.helmholtz(shape, nm, nq, nel, deformed, jac, df, lam, u).reshape(nel, nc))
    Ae = np.stack(cols, axis=2)  # [e, i, m]
    return [Ae[e] for e in range(nel)]


def _tet_case(n, nm, lam):
    coords, tets, mp, sg, dirichlet = TM.tet_c0_map(n, nm)
    nel = len(tets)
    jac, df = O.tet_geom(TM.tet_vertices(coords, tets), nel)
    X = TM.tet_quadrature_coords(coords, tets, nm + 1)
    f = (3 * np.pi ** 2 + lam) * np.prod(np.sin(np.pi * X), axis=2).ravel()
    bl = O.iproduct(O.TET, nm, nm + 1, nel, False, jac, f)
    return _blocks(O.TET, nm, nm + 1, nel, False, jac, df, lam), mp, sg, dirichlet, bl


def _hex_case(n, nm, nq, lam, amp, seed):
    Vr, mr, sr, dirichlet = G.rotated_hex_mesh(n, nm, amp, seed)
    nel, deformed = n ** 3, amp > 0
    jac, df = O.hex_geom(Vr, nel, nq, deformed)
    xyz = O.hex_coords(Vr, nel, nq)
    f = (3 * np.pi ** 2 + lam) * np.prod(np.sin(np.pi * xyz), axis=2).ravel()
    bl = O.iproduct(O.HEX, nm, nq, nel, deformed, jac, f)
    return _blocks(O.HEX, nm, nq, nel, deformed, jac, df, lam), np.asarray(mr, dtype=np.int64), sr, dirichlet, bl


def _global(blocks, mp, sg, dirichlet, bl):
    new, ndir = G.dirichlet_first(dirichlet)
    ng = len(dirichlet)
    m = new[mp]
    A = np.zeros((ng, ng))
    off = 0
    for Ae in blocks:
        k = len(Ae)
        idx, s = m[off:off + k], sg[off:off + k]
        A[np.ix_(idx, idx)] += s[:, None] * Ae * s[None, :]
        off += k
    b = np.zeros(ng)
    np.add.at(b, m, sg * bl)
    d = np.diag(A).copy()
    invd = np.zeros(ng)
    invd[ndir:] = 1.0 / d[ndir:]
    return m, ndir, A, b, invd


CASES = [("tet", (2, 3, 1.0)), ("tet", (2, 5, 0.5)), ("hex", (2, 4, 5, 2.0, 0.0, 4)), ("hex", (2, 5, 6, 1.0, 0.12, 9))]


@pytest.mark.parametrize("kind,args", CASES)
def test_local_cg_is_the_assembled_cg(kind, args):
    blocks, mp, sg, dirichlet, bl = _tet_case(*args) if kind == "tet" else _hex_case(*args)
    m, ndir, A, b, invd = _global(blocks, mp, sg, dirichlet, bl)
    ng = len(b)
    xl, il = O.cg_loc(blocks, m, sg, ng, ndir, bl, invd, tol=1e-12, maxit=5000)
    xg, ig = O.cg_dense(A, b, ndir, invd, tol=1e-12, maxit=5000)
    # the two forms sum their inner products in different orders (local
    # vectors against assembled ones): at 1e-12 that rounding moves the
    # stopping iteration by up to 2
    assert abs(il.iterations - ig.iterations) <= 2, (il.iterations, ig.iterations)
    assert abs(il.rhs_magnitude - ig.rhs_magnitude) <= 1e-13 * ig.rhs_magnitude
    assert il.final_error < 1e-12
    # continuity: every local copy of a DOF is sign * the same value, bit for bit
    first = np.full(ng, np.nan)
    seen = np.zeros(ng, dtype=bool)
    for i in range(len(m)):
        if not seen[m[i]]:
            first[m[i]], seen[m[i]] = sg[i] * xl[i], True
    assert np.array_equal(xl, sg * first[m])
    assert np.all(first[:ndir] == 0.0)
    # the same solution as the assembled CG and the dense solve
    xd = np.zeros(ng)
    xd[ndir:] = np.linalg.solve(A[ndir:, ndir:], b[ndir:])
    scale = np.max(np.abs(xd))
    # both stop at ||r|| < 1e-12 ||b||: their errors are bounded by that
    # times the condition number of the Jacobi-scaled free block
    Af = A[ndir:, ndir:]
    dh = 1.0 / np.sqrt(np.diag(Af))
    bound = 10 * 1e-12 * np.linalg.cond(dh[:, None] * Af * dh[None, :])
    assert np.max(np.abs(first - xg)) <= bound * scale
    assert np.max(np.abs(first - xd)) <= bound * scale


def test_maxit_and_zero_rhs():
    blocks, mp, sg, dirichlet, bl = _tet_case(2, 4, 1.0)
    m, ndir, A, b, invd = _global(blocks, mp, sg, dirichlet, bl)
    ng = len(b)
    xl, il = O.cg_loc(blocks, m, sg, ng, ndir, bl, invd, tol=1e-14, maxit=3)
    xg, ig = O.cg_dense(A, b, ndir, invd, tol=1e-14, maxit=3)
    assert il.iterations == ig.iterations == 4  # NekLinSysIterCGLoc.cpp:200: exits once iterations > maxit
    first = np.zeros(ng)
    first[m] = sg * xl
    assert np.max(np.abs(first - xg)) <= 1e-12 * np.max(np.abs(xg))
    x0, i0 = O.cg_loc(blocks, m, sg, ng, ndir, np.zeros_like(bl), invd)
    assert i0.iterations == 0 and np.all(x0 == 0.0) and i0.rhs_magnitude == 1.0


def test_unassembled_residual_norm_is_not_the_stopping_norm():
    # the stopping norm is <AssembleLoc(r), r> = ||Assemble(r)||^2 on the
    # free rows, not the local ||r||^2: on a signed map they differ, and
    # the oracle's rhs magnitude is the assembled one
    blocks, mp, sg, dirichlet, bl = _tet_case(2, 3, 1.0)
    m, ndir, A, b, invd = _global(blocks, mp, sg, dirichlet, bl)
    _, il = O.cg_loc(blocks, m, sg, len(b), ndir, bl, invd, tol=1e-10)
    assert abs(il.rhs_magnitude - float(b @ b)) <= 1e-13 * float(b @ b)
    assert abs(il.rhs_magnitude - float(bl @ bl)) > 1e-3 * float(b @ b)
