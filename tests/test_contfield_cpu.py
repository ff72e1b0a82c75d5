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
