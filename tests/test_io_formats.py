"""Nektar++ file formats (nektar_amd/io.py) against the reference's own
FieldConvert test data (tests/golden/fieldconvert: the mesh and field of
utilities/FieldConvert/Tests/Hex_channel_C0helmsmoothing, 8 hexahedra on
the unit cube, NUMMODES 4, B64Z-LittleEndian compressed):

* the compressed mesh decodes to the unit cube's 27 vertices, 54 edges,
  36 quadrilateral faces and 8 hexahedra; local vertex orders follow
  HexGeom (right-handed elements), and the geometric factors equal the
  oracle's;
* the reference's four fields, read from the field file and evaluated in
  our element frames at the GLL points, agree exactly wherever elements
  meet -- which fails for a wrong vertex order, a wrong coefficient layout
  or a wrong field order (p varies in x, y and z);
* write_fld round-trips the reference's field file;
* the C0 map: (P n + 1)^3 global DOFs, random global coefficients give
  fields continuous across every shared point, and the boundary faces'
  DOFs are the Dirichlet ones;
* the tetrahedral Tet_channel_m3 mesh and fields: TetGeom frames (geometry
  equal to the oracle's, the tets fill the unit cube), and the reference's
  fields agree across every shared face in those frames (a swapped vertex
  pair breaks it)."""
import os
from collections import defaultdict

import numpy as np
import pytest

import _oracle as O
from nektar_amd import io as nio

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fieldconvert", "Hex_channel_C0helmsmoothing")


@pytest.fixture(scope="module")
def mesh():
    return nio.read_mesh_xml(G + ".xml")


def _frames(mesh):
    eids = sorted(mesh.elements)
    lv = np.array([nio.hex_local_vertices(mesh, e) for e in eids])
    verts = np.array([[mesh.vertices[v] for v in row] for row in lv])
    return eids, lv, verts


def _gll_coords(verts, nq):
    z = O.points(O.GLL, nq)[0]
    K, J, I = np.meshgrid(np.arange(nq), np.arange(nq), np.arange(nq), indexing="ij")
    xi = np.stack([z[I.ravel()], z[J.ravel()], z[K.ravel()]], axis=1)
    out = []
    for v in verts:
        A = np.stack([v[1] - v[0], v[3] - v[0], v[4] - v[0]], axis=1) / 2
        out.append(v[0] + (xi + 1) @ A.T)
    return np.array(out)


def _max_jump(X, u):
    d = defaultdict(list)
    for e in range(len(X)):
        for k in range(X.shape[1]):
            d[tuple(np.round(X[e, k], 9))].append(u[e, k])
    jumps = [max(v) - min(v) for v in d.values() if len(v) > 1]
    return len(jumps), max(jumps)


def test_mesh_reader(mesh):
    assert len(mesh.vertices) == 27 and len(mesh.edges) == 54 and len(mesh.faces) == 36
    assert all(k == "Q" for k, _ in mesh.faces.values())
    assert sorted(mesh.elements) == list(range(8)) and all(k == "H" for k, _ in mesh.elements.values())
    assert mesh.composites[0] == [("H", list(range(8)))] and mesh.domain == [0]
    assert [len(ids) for c in range(1, 7) for _, ids in mesh.composites[c]] == [4] * 6
    assert mesh.expansions[0]["nummodes"] == 4 and mesh.expansions[0]["fields"] == ["u", "v", "w", "p"]
    xyz = np.array(list(mesh.vertices.values()))
    assert np.allclose(sorted(set(xyz[:, 0])), [0, 0.5, 1])
    eids, lv, verts = _frames(mesh)
    jac, df = nio.hex_geometry(verts)
    jo, dfo = O.hex_geom(np.ascontiguousarray(verts.reshape(-1)), len(eids), 5, False)
    assert np.max(np.abs(jac - jo)) <= 1e-15 and np.max(np.abs(df - dfo)) <= 1e-13
    A = np.stack([verts[:, 1] - verts[:, 0], verts[:, 3] - verts[:, 0], verts[:, 4] - verts[:, 0]], axis=2)
    assert np.all(np.linalg.det(A) > 0)  # right-handed local frames
    assert np.allclose(jac, 0.25 ** 3)


def test_reference_fields_are_continuous_in_our_frames(mesh):
    eids, lv, verts = _frames(mesh)
    (blk,) = nio.read_fld(G + ".fld")
    assert blk.fields == ["u", "v", "w", "p"] and blk.shape == "Hexahedron" and blk.nummodes == [4, 4, 4]
    assert blk.ids == list(range(8)) and blk.data["u"].shape == (8, 64)
    nm, nq = 4, 5
    X = _gll_coords(verts, nq)
    pos = {e: i for i, e in enumerate(blk.ids)}
    for f in blk.fields:
        c = np.array([blk.data[f][pos[e]] for e in eids]).reshape(-1)
        u = O.bwdtrans(O.HEX, nm, nq, len(eids), c).reshape(len(eids), -1)
        n, jump = _max_jump(X, u)
        assert n == 217 and jump <= 1e-13 * np.max(np.abs(u)), (f, jump)
    # a wrong vertex order breaks it: swap the bottom and top faces of element 0
    bad = verts.copy()
    bad[0] = verts[0][[4, 5, 6, 7, 0, 1, 2, 3]]
    c = np.array([blk.data["p"][pos[e]] for e in eids]).reshape(-1)
    u = O.bwdtrans(O.HEX, nm, nq, len(eids), c).reshape(len(eids), -1)
    assert _max_jump(_gll_coords(bad, nq), u)[1] > 1e-3


def test_fld_round_trip(tmp_path):
    blocks = nio.read_fld(G + ".fld")
    out = tmp_path / "copy.fld"
    nio.write_fld(str(out), blocks, {"Time": 0})
    again = nio.read_fld(str(out))
    assert len(again) == 1
    a, b = blocks[0], again[0]
    assert (a.fields, a.shape, a.basis, a.nummodes, a.ids) == (b.fields, b.shape, b.basis, b.nummodes, b.ids)
    for f in a.fields:
        assert np.array_equal(a.data[f], b.data[f])


def test_c0_map(mesh):
    eids, lv, verts = _frames(mesh)
    nm, nq = 4, 5
    mp, sg, ng, nd = nio.hex_c0_map(lv, nm)
    assert ng == (2 * (nm - 1) + 1) ** 3 and nd == 0
    assert set(np.unique(sg)) <= {-1.0, 1.0}
    X = _gll_coords(verts, nq)
    rng = np.random.default_rng(5)
    for _ in range(3):
        x = rng.uniform(-1, 1, ng)
        u = O.bwdtrans(O.HEX, nm, nq, len(eids), sg * x[mp]).reshape(len(eids), -1)
        assert _max_jump(X, u)[1] <= 1e-12
    # a sign error in one element breaks it: flip the odd edge mode p = 3 on
    # element 0's edge xi1 = xi2 = +1 (inside the cube, shared with 3 elements)
    bad = sg.copy()
    bad[3 + nm * (1 + nm * 1)] *= -1
    u = O.bwdtrans(O.HEX, nm, nq, len(eids), bad * x[mp]).reshape(len(eids), -1)
    assert _max_jump(X, u)[1] > 1e-3
    # the six boundary composites as Dirichlet faces: every DOF on the cube's surface
    bfaces = [f for c in range(1, 7) for _, ids in mesh.composites[c] for f in ids]
    vsets = [{v for e in mesh.faces[f][1] for v in mesh.edges[e]} for f in bfaces]
    mp2, sg2, ng2, nd2 = nio.hex_c0_map(lv, nm, vsets)
    assert ng2 == ng and nd2 == ng - (2 * (nm - 1) - 1) ** 3
    assert np.all(mp2[np.isin(mp2, np.arange(nd2))] < nd2)


@pytest.mark.parametrize("nm,amp,seed", [(4, 0.0, 3), (5, 0.15, 11)])
def test_c0_map_in_random_frames(nm, amp, seed):
    # every element of a 3^3 mesh in a random axis permutation and reflection
    # (tests/geometry_cases.py): face modes then meet transposed and reversed
    import geometry_cases as GC
    n = 3
    Vr, mr, sr, dirichlet = GC.rotated_hex_mesh(n, nm, amp, seed)
    verts = Vr.reshape(-1, 8, 3)
    keys = {}
    lv = np.array([[keys.setdefault(tuple(np.round(v, 12)), len(keys)) for v in el] for el in verts])
    assert len(keys) == (n + 1) ** 3
    mp, sg, ng, nd = nio.hex_c0_map(lv, nm)
    assert ng == len(dirichlet)
    nq = nm + 1
    X = np.array([O.hex_coords(np.ascontiguousarray(v.reshape(-1)), 1, nq)[0] for v in verts])
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, ng)
    u = O.bwdtrans(O.HEX, nm, nq, len(verts), sg * x[mp]).reshape(len(verts), -1)
    assert _max_jump(X, u)[1] <= 1e-12
    # the same space as the trace-matched map of the generator: both maps are
    # bijections of the local entries onto the global DOFs with matching
    # signs up to one global sign per DOF
    pair = {}
    for a, b, s1, s2 in zip(mp, mr, sg, sr):
        pair.setdefault(a, set()).add((b, s1 * s2))
    assert all(len(v) == 1 for v in pair.values())


def test_plain_and_compressed_meshes_agree(mesh):
    # Hex_channel_m3.xml holds the same mesh as plain text (V, E, Q, H elements)
    plain = nio.read_mesh_xml(os.path.join(os.path.dirname(G), "Hex_channel_m3.xml"))
    assert plain.edges == mesh.edges and plain.faces == mesh.faces and plain.elements == mesh.elements
    assert all(np.array_equal(plain.vertices[k], mesh.vertices[k]) for k in mesh.vertices)
    ea, la, va = nio.hex_elements(plain)
    eb, lb, vb = nio.hex_elements(mesh)
    assert ea == eb and np.array_equal(la, lb) and np.array_equal(va, vb)


def test_tet_mesh_and_fields():
    # Tet_channel_m3 (utilities/FieldConvert/Tests): 48 tetrahedra on the unit
    # cube in plain text, NUMMODES 3, and its fields u, v, w, p
    import itertools

    import tet_mesh as TM
    d = os.path.dirname(G)
    tm = nio.read_mesh_xml(os.path.join(d, "Tet_channel_m3.xml"))
    assert len(tm.vertices) == 27 and len(tm.elements) == 48
    eids, lv, verts = nio.tet_elements(tm)
    jac, df = nio.tet_geometry(verts)
    jo, dfo = O.tet_geom(np.ascontiguousarray(verts.reshape(-1)), len(eids))
    assert np.max(np.abs(jac - jo)) <= 1e-15 and np.max(np.abs(df - dfo)) <= 1e-13
    assert abs(np.sum(jac) * 4 / 3 - 1.0) <= 1e-13  # the tets fill the unit cube
    A = np.stack([verts[:, 1] - verts[:, 0], verts[:, 2] - verts[:, 0], verts[:, 3] - verts[:, 0]], axis=2)
    assert np.all(np.linalg.det(A) > 0)
    (blk,) = nio.read_fld(os.path.join(d, "Tet_channel_m3.fld"))
    assert blk.shape == "Tetrahedron" and blk.nummodes == [3, 3, 3] and blk.data["p"].shape == (48, 10)
    # every field, evaluated in our frames on both sides of every shared face
    # (the tet modes as polynomials, tests/tet_mesh.py), agrees
    nm = 3
    C = TM.mode_polynomials(nm)
    pos = {e: i for i, e in enumerate(blk.ids)}
    faces = {}
    for k, e in enumerate(eids):
        for c in itertools.combinations(range(4), 3):
            faces.setdefault(frozenset(int(lv[k][i]) for i in c), []).append(k)
    shared = [(key, ks) for key, ks in faces.items() if len(ks) == 2]
    assert len(shared) == (4 * 48 - 6 * 2 * 2 * 2) // 2
    W = TM._W_FACE[:6]

    def worst_jump(vv):
        worst = 0.0
        for key, ks in shared:
            pts = W @ np.array([tm.vertices[v] for v in sorted(key)])
            vals = [TM._eval_modes(C, nm - 1, vv[k], pts).T for k in ks]
            for f in blk.fields:
                a = vals[0] @ blk.data[f][pos[eids[ks[0]]]]
                b = vals[1] @ blk.data[f][pos[eids[ks[1]]]]
                worst = max(worst, np.max(np.abs(a - b)))
        return worst

    assert worst_jump(verts) <= 1e-10
    # a wrong frame breaks it: vertices 1 and 2 of the first element swapped
    bad = verts.copy()
    bad[0] = verts[0][[0, 2, 1, 3]]
    assert worst_jump(bad) > 1e-2


@pytest.mark.parametrize("nm", [3, 4, 5])
def test_tet_c0_map(nm):
    # on the reference's tet mesh: (2 (nm-1) + 1)^3 DOFs, random global
    # coefficients continuous across every shared face, and a sign error on
    # one face mode breaks it; on the Kuhn meshes of tests/tet_mesh.py the
    # map corresponds one-to-one, signs included, with the trace-matched one
    import itertools

    import tet_mesh as TM
    tm = nio.read_mesh_xml(os.path.join(os.path.dirname(G), "Tet_channel_m3.xml"))
    eids, lv, verts = nio.tet_elements(tm)
    mp, sg, ng, nd = nio.tet_c0_map(lv, nm)
    assert ng == (2 * (nm - 1) + 1) ** 3 and nd == 0
    C = TM.mode_polynomials(nm)
    nc = C.shape[1]
    faces = {}
    for k in range(len(eids)):
        for c in itertools.combinations(range(4), 3):
            faces.setdefault(frozenset(int(lv[k][i]) for i in c), []).append(k)
    shared = [(key, ks) for key, ks in faces.items() if len(ks) == 2]

    def worst(s):
        loc = (s * x[mp]).reshape(len(eids), nc)
        w = 0.0
        for key, ks in shared:
            pts = TM._W_FACE[:8] @ np.array([tm.vertices[v] for v in sorted(key)])
            a, b = (TM._eval_modes(C, nm - 1, verts[k], pts).T @ loc[k] for k in ks)
            w = max(w, np.max(np.abs(a - b)))
        return w

    x = np.random.default_rng(nm).uniform(-1, 1, ng)
    assert worst(sg) <= 1e-12
    if nm >= 4:
        # flip, in one element only, the first interior mode of a shared face
        key, (k0, _) = shared[0]
        where = sorted(i for i in range(4) if int(lv[k0][i]) in key)
        first = {(0, 1, 2): (2, 1, 0), (0, 1, 3): (2, 0, 1), (1, 2, 3): (1, 1, 1), (0, 2, 3): (0, 2, 1)}[tuple(where)]
        order = [(p, q, r) for p in range(nm) for q in range(nm - p) for r in range(nm - p - q)]
        bad = sg.copy()
        bad[k0 * nc + order.index(first)] *= -1
        assert worst(bad) > 1e-3
    # Kuhn tets (vertices ascending): the same space as the trace-matched map
    coords, tets, mr, sr, dirichlet = TM.tet_c0_map(2, nm)
    m2, s2, ng2, _ = nio.tet_c0_map(tets, nm)
    assert ng2 == len(dirichlet)
    pair = {}
    for a, b, u, v in zip(m2, mr, s2, sr):
        pair.setdefault(a, set()).add((b, u * v))
    assert all(len(v) == 1 for v in pair.values())
