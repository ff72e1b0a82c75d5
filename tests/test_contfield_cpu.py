"""The host side of the Dirichlet Helmholtz solve (nektar_amd/contfield.py:
session reading, expression evaluation, the constrained boundary
projection and its imposition on the local coefficients), checked without a
GPU: the same steps with the oracle's operators and a dense solve reproduce
the reference's own regression metrics for
library/Demos/MultiRegions/Tests/Helmholtz3D_CG_Hex.tst (8 hexahedra,
NUMMODES 6, u = sin(pi x/2) sin(pi y/2) sin(pi z/2) on the boundary):
L2 1.11818e-07 and Linf 5.66374e-07, to the test's tolerance 1e-12."""
import os

import numpy as np
import pytest

import _oracle as O
import geometry_cases as G
from nektar_amd import contfield as cf
from nektar_amd import io as nio

HERE = os.path.dirname(os.path.abspath(__file__))
HEX = os.path.join(HERE, "golden", "helmholtz3d", "Helmholtz3D_Hex_P6.xml")


def test_expression_evaluator():
    v = cf.evaluate("-(Lambda + 3*PI*PI/4)*sin(PI/2*x)*sin(PI/2*y)^2", dict(Lambda=2.0, x=0.3, y=0.7))
    want = -(2.0 + 3 * np.pi ** 2 / 4) * np.sin(np.pi / 2 * 0.3) * np.sin(np.pi / 2 * 0.7) ** 2
    assert abs(v - want) < 1e-15
    for bad in ("__import__('os')", "x.real", "[1, 2]", "lambda q: q", "undefined_name + 1"):
        with pytest.raises(ValueError):
            cf.evaluate(bad, dict(x=1.0))


def test_session_reader():
    s = cf.read_session(HEX)
    assert s.params == {"Lambda": 1.0}
    assert s.regions == {0: [1, 2, 3, 4, 5, 6]}
    assert [bc.kind for bc in s.conditions[0]] == ["D"]
    assert set(s.functions) == {"Forcing", "ExactSolution"}


def test_constrained_projection_reproduces_polynomials():
    """Vertices interpolated, edges and face interior projected: a face
    function in the span of the modes is recovered exactly."""
    import nektar_amd as nmf

    nm, nq = 6, 7
    t = nmf.basis_tables(nmf.SHAPE_HEX, 0, nm, nq)
    B = t["bdata"].reshape(nm, nq)
    C = np.random.default_rng(1).uniform(-1, 1, (nm, nm))
    F = B.T @ C @ B  # F[i, j] = sum C[a, b] phi_a(z_i) phi_b(z_j)
    assert np.max(np.abs(cf._project_quad(B, t["w"], F) - C)) < 1e-12


def test_helmholtz3d_cg_hex_metrics_with_the_oracle():
    import nektar_amd as nmf

    s = cf.read_session(HEX)
    eids, lv, verts = nio.hex_elements(s.mesh)
    jac, df = nio.hex_geometry(verts)
    nm = s.mesh.expansions[0]["nummodes"]
    nq, E, nc = nm + 1, len(eids), nm ** 3
    dsets = [cf._face_vertices(s.mesh, f) for cid in s.regions[0] for _, ids in s.mesh.composites[cid] for f in ids]
    mp, sg, ng, nd = nio.hex_c0_map(lv, nm, dsets)
    t = nmf.basis_tables(nmf.SHAPE_HEX, 0, nm, nq)
    dloc = cf.hex_dirichlet_local(s, "u", lv, verts, nm, nq, mp, sg, nd, (t["bdata"].reshape(nm, nq), t["z"], t["w"]))
    assert np.all(dloc[mp >= nd] == 0.0)
    X = cf.hex_quadrature_points(verts, t["z"])
    names = dict(s.params, x=X[..., 0], y=X[..., 1], z=X[..., 2], t=0.0)
    f = cf.evaluate(s.functions["Forcing"]["u"], names).reshape(-1)
    exact = cf.evaluate(s.functions["ExactSolution"]["u"], names)
    lam = s.params["Lambda"]
    # ContField::v_HelmSolve's rhs and GlobalLinSysIterativeFull's lifting
    rl = -O.iproduct(O.HEX, nm, nq, E, False, jac, f) - O.helmholtz(O.HEX, nm, nq, E, False, jac, df, lam, dloc)
    A = np.zeros((ng, ng))
    for m in range(nc):
        x = np.zeros(E * nc)
        x[m::nc] = 1.0
        y = O.helmholtz(O.HEX, nm, nq, E, False, jac, df, lam, x)
        for e in range(E):
            i = e * nc + np.arange(nc)
            np.add.at(A, (mp[i], np.full(nc, mp[e * nc + m])), sg[i] * y[i] * sg[e * nc + m])
    b = np.zeros(ng)
    np.add.at(b, mp, sg * rl)
    xg = np.zeros(ng)
    xg[nd:] = np.linalg.solve(A[nd:, nd:], b[nd:])
    u = O.bwdtrans(O.HEX, nm, nq, E, sg * xg[mp] + dloc).reshape(E, -1)
    l2, linf = cf.errors(u, exact, jac, t["w"])
    assert abs(l2 - 1.11818e-07) <= 1e-12, l2
    assert abs(linf - 5.66374e-07) <= 1e-12, linf


TET = os.path.join(HERE, "golden", "helmholtz3d", "Helmholtz3D_Tet.xml")


def test_helmholtz3d_cg_tet_metrics_with_the_oracle():
    """library/Demos/MultiRegions/Tests/Helmholtz3D_CG_Tet.tst (48 tets,
    NUMMODES 5): L2 8.458e-05 (tolerance 1e-8), Linf 0.000862348 (1e-9).
    These values carry the reference's handling of triangular boundary
    faces (Radau edge values interpolated to GLL points, so the collapsed
    vertex takes an extrapolated value; later faces overwrite shared
    vertices and edges)."""
    s = cf.read_session(TET)
    P = cf.dirichlet_problem(s)
    nm, nq0, E, jac, df = P["nm"], P["nq0"], len(P["eids"]), P["jac"], P["df"]
    mp, sg, ng, nd, dloc, X = P["map"], P["sign"], P["nglobal"], P["ndir"], P["dloc"], P["X"]
    names = dict(s.params, x=X[..., 0], y=X[..., 1], z=X[..., 2], t=0.0)
    f = cf.evaluate(s.functions["Forcing"]["u"], names).reshape(-1)
    exact = cf.evaluate(s.functions["ExactSolution"]["u"], names)
    lam = s.params["Lambda"]
    nc = O.ncoeffs(O.TET, nm)
    rl = -O.iproduct(O.TET, nm, nq0, E, False, jac, f) - O.helmholtz(O.TET, nm, nq0, E, False, jac, df, lam, dloc)
    A = np.zeros((ng, ng))
    for m in range(nc):
        x = np.zeros(E * nc)
        x[m::nc] = 1.0
        y = O.helmholtz(O.TET, nm, nq0, E, False, jac, df, lam, x)
        for e in range(E):
            i = e * nc + np.arange(nc)
            np.add.at(A, (mp[i], np.full(nc, mp[e * nc + m])), sg[i] * y[i] * sg[e * nc + m])
    b = np.zeros(ng)
    np.add.at(b, mp, sg * rl)
    xg = np.zeros(ng)
    xg[nd:] = np.linalg.solve(A[nd:, nd:], b[nd:])
    u = O.bwdtrans(O.TET, nm, nq0, E, sg * xg[mp] + dloc)
    l2, linf = cf.tet_errors(u, exact, jac, *P["weights"])
    assert abs(l2 - 8.458e-05) <= 1e-8, l2
    assert abs(linf - 0.000862348) <= 1e-9, linf


def perturbed_hex_session(seed=3, amp=0.06):
    """The Helmholtz3D_CG_Hex session with every vertex coordinate that is
    not on the box boundary moved by U[-amp, amp]: trilinear elements whose
    boundary faces stay in the box planes but are no longer parallelograms."""
    s = cf.read_session(HEX)
    rng = np.random.default_rng(seed)
    for v in sorted(s.mesh.vertices):
        x = s.mesh.vertices[v].copy()
        inner = (x > 1e-12) & (x < 1 - 1e-12)
        x[inner] += rng.uniform(-amp, amp, int(inner.sum()))
        s.mesh.vertices[v] = x
    return s


def dense_solve(s, P):
    """The lifted C0 solve with the oracle's operators and a dense solve;
    returns the element-local solution and the L2 / Linf errors."""
    nm, nq, E, jac, df, X = P["nm"], P["nq0"], len(P["eids"]), P["jac"], P["df"], P["X"]
    mp, sg, ng, nd, dloc, deformed = P["map"], P["sign"], P["nglobal"], P["ndir"], P["dloc"], P["deformed"]
    nc = nm ** 3
    names = dict(s.params, x=X[..., 0], y=X[..., 1], z=X[..., 2], t=0.0)
    f = cf.evaluate(s.functions["Forcing"]["u"], names).reshape(-1)
    exact = cf.evaluate(s.functions["ExactSolution"]["u"], names)
    lam = s.params["Lambda"]
    # the Robin boundary mass as a dense local matrix (block diagonal by element)
    R = np.zeros((E * nc, E * nc))
    if P.get("robin") is not None:
        rows, rowptr, cols, vals = P["robin"]
        for k, r in enumerate(rows):
            R[r, cols[rowptr[k]:rowptr[k + 1]]] += vals[rowptr[k]:rowptr[k + 1]]
        assert np.allclose(R, R.T, rtol=0, atol=1e-15)
    rl = (-O.iproduct(O.HEX, nm, nq, E, deformed, jac, f) + P.get("bweak", 0.0)
          - O.helmholtz(O.HEX, nm, nq, E, deformed, jac, df, lam, dloc) - R @ dloc)
    A = np.zeros((ng, ng))
    for m in range(nc):
        x = np.zeros(E * nc)
        x[m::nc] = 1.0
        y = O.helmholtz(O.HEX, nm, nq, E, deformed, jac, df, lam, x) + R @ x
        for e in range(E):
            i = e * nc + np.arange(nc)
            np.add.at(A, (mp[i], np.full(nc, mp[e * nc + m])), sg[i] * y[i] * sg[e * nc + m])
    b = np.zeros(ng)
    np.add.at(b, mp, sg * rl)
    xg = np.zeros(ng)
    xg[nd:] = np.linalg.solve(A[nd:, nd:], b[nd:])
    xl = sg * xg[mp] + dloc
    u = O.bwdtrans(O.HEX, nm, nq, E, xl)
    return xl, cf.errors(u, exact, jac, P["weights"][0])


def test_trilinear_hexahedra_from_files():
    """Per-point factors of trilinear hexahedra equal the oracle's
    (GeomFactors' deformed case), the boundary projection on faces that are
    not parallelograms is orthogonal in the face's area-weighted inner
    product, and the lifted solve converges to the exact solution."""
    import nektar_amd as nmf

    s = perturbed_hex_session()
    P = cf.dirichlet_problem(s)
    assert P["deformed"]
    nm, nq, verts = P["nm"], P["nq0"], P["verts"]
    jo, dfo = O.hex_geom(np.ascontiguousarray(verts.reshape(-1)), len(verts), nq, True)
    assert G.rel_err(P["jac"], jo) < 1e-14 and G.rel_err(P["df"], dfo) < 1e-14
    # one boundary face: xi_2 = -1 of element 0 (z = 0 plane, a general quadrilateral)
    t = nmf.basis_tables(nmf.SHAPE_HEX, 0, nm, nq)
    B, z, w = t["bdata"].reshape(nm, nq), t["z"], t["w"]
    xi = np.zeros((nq, nq, 3))
    xi[:, :, 0], xi[:, :, 1], xi[:, :, 2] = z[:, None], z[None, :], -1.0
    X, dX = cf._hex_map(verts[0], xi.reshape(-1, 3))
    Jf = np.linalg.norm(np.cross(dX[:, :, 0], dX[:, :, 1]), axis=1).reshape(nq, nq)
    assert np.ptp(Jf) > 1e-3 * np.max(Jf)  # not a parallelogram
    F = np.sin(3 * X[:, 0]) * np.cos(2 * X[:, 1]).reshape(-1)
    F = F.reshape(nq, nq)
    C = cf._project_quad(B, w, F, Jf)
    g = B.T @ C @ B
    for a in range(2, nm):
        for b in range(2, nm):
            phi = np.outer(B[a], B[b])
            assert abs(np.sum(np.outer(w, w) * Jf * (F - g) * phi)) < 1e-14 * np.sum(np.outer(w, w) * Jf)
    # the solve
    _, (l2, linf) = dense_solve(s, P)
    assert l2 < 5e-7 and linf < 2e-6, (l2, linf)  # 1.3e-07 and 6.1e-07


ALLBCS = os.path.join(HERE, "golden", "helmholtz3d", "Helmholtz3D_Hex_AllBCs_P6.xml")


def test_helmholtz3d_cg_hex_allbcs_metrics_with_the_oracle():
    """library/Demos/MultiRegions/Tests/Helmholtz3D_CG_Hex_AllBCs.tst (27
    hexahedra, NUMMODES 4; Dirichlet, Neumann and Robin faces with
    PRIMCOEFF 2 and 3): the reference prints L2 0.000416575 and Linf
    0.000871589 (six significant digits, tolerance 1e-12 on the printed
    values).  The Neumann and Robin data enter the right-hand side as face
    integrals and the Robin faces add their alpha-weighted face mass to the
    operator; without either, the metrics differ."""
    s = cf.read_session(ALLBCS)
    assert sorted(bc.kind for bcs in s.conditions.values() for bc in bcs) == ["D", "D", "N", "N", "R", "R"]
    P = cf.dirichlet_problem(s)
    assert P["robin"] is not None and np.any(P["bweak"])
    _, (l2, linf) = dense_solve(s, P)
    assert f"{l2:.6g}" == "0.000416575", l2
    assert f"{linf:.6g}" == "0.000871589", linf
    # planted omissions: no Robin mass, or no Neumann/Robin right-hand side
    for drop in ("robin", "bweak"):
        Q = dict(P)
        Q[drop] = None if drop == "robin" else np.zeros_like(P["bweak"])
        _, (l2b, _) = dense_solve(s, Q)
        assert abs(l2b - l2) > 1e-6, drop
