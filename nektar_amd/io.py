"""Nektar++ file formats for the solve path (host side).

* ``read_mesh_xml``: the session/mesh XML geometry (MeshGraphIOXml.cpp,
  MeshGraphIOXmlCompressed.cpp): VERTEX, EDGE, FACE (Q, T), ELEMENT (H, A),
  COMPOSITE, DOMAIN and EXPANSIONS, in plain text or in the
  ``B64Z-LittleEndian`` form (zlib-compressed, base64-encoded arrays of the
  int64/double records of SpatialDomains/MeshEntities.hpp:51-98).
* ``hex_local_vertices``: a hexahedron's local vertex order from its six
  faces, as HexGeom derives it (HexGeom.cpp:200-430); this order defines the
  element's reference map and so the meaning of its coefficients.
* ``hex_geometry``: Jacobian and derivative factors of straight-sided
  (parallelepiped) hexahedra, one set per element (GeomFactors.cpp:150-477
  for the regular case); ``hex_geometry_points``: the same at every
  quadrature point of trilinear hexahedra (the deformed case).
* ``tet_local_vertices`` / ``tet_geometry``: the same for tetrahedra
  (TetGeom.cpp:113-300).
* ``hex_c0_map``: the C0 local-to-global map with orientation signs of a
  conforming hex mesh (AssemblyMapCG's numbering role), from the elements'
  global vertex ids: vertex, edge, face and interior modes; edge and face
  modes are matched through a canonical frame per edge and face.
* ``read_fld`` / ``write_fld``: field files (FieldIOXml.cpp): ELEMENTS blocks
  with their FIELDS, SHAPE, BASIS, NUMMODESPERDIR and ID attributes and the
  coefficients field by field, element by element, in the ID order.
* ``tet_c0_map``: the same for tetrahedra, following the reference's edge
  and triangular-face trace maps (StdTetExp.cpp:1459-1660).
* ``c0_helmholtz_smoothing``: FieldConvert's C0Projection:helmsmoothing on
  hexahedral or tetrahedral meshes, from the files to the smoothed fields,
  on the GPU.
"""
from __future__ import annotations

import base64
import re
import xml.etree.ElementTree as ET
import zlib
from dataclasses import dataclass, field

import numpy as np

_B64Z = "B64Z-LittleEndian"


def _b64z_decode(text: str, dtype) -> np.ndarray:
    return np.frombuffer(zlib.decompress(base64.b64decode(text.strip())), dtype=dtype)


def _b64z_encode(a: np.ndarray) -> str:
    return base64.b64encode(zlib.compress(np.ascontiguousarray(a).tobytes())).decode("ascii")


def parse_ranges(s: str) -> list[int]:
    """'0-3,7' -> [0, 1, 2, 3, 7] (ParseUtils::GenerateSeqVector)."""
    out: list[int] = []
    for part in s.replace(" ", "").split(","):
        if not part:
            continue
        if "-" in part:
            a, b = part.split("-")
            out.extend(range(int(a), int(b) + 1))
        else:
            out.append(int(part))
    return out


def _ranges_string(ids) -> str:
    ids = list(ids)
    parts, i = [], 0
    while i < len(ids):
        j = i
        while j + 1 < len(ids) and ids[j + 1] == ids[j] + 1:
            j += 1
        parts.append(str(ids[i]) if i == j else f"{ids[i]}-{ids[j]}")
        i = j + 1
    return ",".join(parts)


@dataclass
class MeshGraph:
    vertices: dict = field(default_factory=dict)    # id -> xyz (3,)
    edges: dict = field(default_factory=dict)       # id -> (v0, v1)
    faces: dict = field(default_factory=dict)       # id -> ('Q' | 'T', [edge ids])
    elements: dict = field(default_factory=dict)    # id -> ('H' | 'A', [face ids])
    composites: dict = field(default_factory=dict)  # id -> [(letter, [ids])]
    domain: list = field(default_factory=list)      # composite ids
    expansions: list = field(default_factory=list)  # [{composites, nummodes, fields, type}]


def _composite_refs(text: str):
    """' H[0-7] F[1,2] ' -> [('H', [0..7]), ('F', [1, 2])]; 'C[0]' likewise."""
    return [(m.group(1), parse_ranges(m.group(2))) for m in re.finditer(r"([A-Z])\[([^\]]*)\]", text)]


def read_mesh_xml(path: str) -> MeshGraph:
    root = ET.parse(path).getroot()
    geo = root.find("GEOMETRY")
    if geo is None:
        raise ValueError(f"{path}: no GEOMETRY element")
    g = MeshGraph()

    def records(node, tag, width, dtype="<i8"):
        """(id, values...) rows of a plain or compressed entity list."""
        if node.get("COMPRESSED"):
            if node.get("COMPRESSED") != _B64Z or node.get("BITSIZE", "64") != "64":
                raise ValueError(f"unsupported compression {node.get('COMPRESSED')}/{node.get('BITSIZE')}")
            return _b64z_decode(node.text, dtype).reshape(-1, width)
        rows = []
        for child in node.findall(tag):
            rows.append([int(child.get("ID"))] + [float(x) for x in child.text.split()])
        return np.array(rows).reshape(-1, width)

    vnode = geo.find("VERTEX")
    if vnode.get("COMPRESSED"):
        v = _b64z_decode(vnode.text, np.dtype([("id", "<i8"), ("x", "<f8"), ("y", "<f8"), ("z", "<f8")]))
        for r in v:
            g.vertices[int(r["id"])] = np.array([r["x"], r["y"], r["z"]])
    else:
        for child in vnode.findall("V"):
            g.vertices[int(child.get("ID"))] = np.array([float(x) for x in child.text.split()])
    for r in records(geo.find("EDGE"), "E", 3):
        g.edges[int(r[0])] = (int(r[1]), int(r[2]))
    fnode = geo.find("FACE")
    for tag, n in (("Q", 4), ("T", 3)):
        for node in fnode.findall(tag) if fnode is not None else []:
            rows = records(node, tag, n + 1) if node.get("COMPRESSED") else [
                [int(node.get("ID"))] + [int(x) for x in node.text.split()]]
            for r in rows:
                g.faces[int(r[0])] = (tag, [int(x) for x in r[1:]])
    enode = geo.find("ELEMENT")
    for tag, n in (("H", 6), ("A", 4)):
        for node in enode.findall(tag):
            rows = records(node, tag, n + 1) if node.get("COMPRESSED") else [
                [int(node.get("ID"))] + [int(x) for x in node.text.split()]]
            for r in rows:
                g.elements[int(r[0])] = (tag, [int(x) for x in r[1:]])
    for tag in ("P", "R", "S"):  # prisms, pyramids, ... are not part of this path
        if enode.find(tag) is not None:
            raise NotImplementedError(f"element type {tag} is not supported")
    for c in geo.find("COMPOSITE").findall("C"):
        g.composites[int(c.get("ID"))] = _composite_refs(c.text)
    dom = geo.find("DOMAIN")
    if dom is not None:
        g.domain = [i for _, ids in _composite_refs(dom.text) for i in ids]
    exps = root.find("EXPANSIONS")
    for e in exps.findall("E") if exps is not None else []:
        g.expansions.append({"composites": [i for _, ids in _composite_refs(e.get("COMPOSITE")) for i in ids],
                             "nummodes": int(e.get("NUMMODES")), "type": e.get("TYPE", "MODIFIED"),
                             "fields": [f.strip() for f in e.get("FIELDS", "u").split(",")]})
    return g


def _shared_edge(mesh: MeshGraph, fa: int, fb: int) -> int:
    common = set(mesh.faces[fa][1]) & set(mesh.faces[fb][1])
    if len(common) != 1:
        raise ValueError(f"faces {fa} and {fb} share {len(common)} edges, expected 1")
    # the edge as face fa lists it (HexGeom takes m_faces[a]->GetEdge(i))
    return next(e for e in mesh.faces[fa][1] if e in common)


def hex_local_vertices(mesh: MeshGraph, eid: int) -> list[int]:
    """Global ids of the 8 local vertices of hexahedron eid (HexGeom.cpp:
    200-339 local edges from the faces, :341-430 vertices from the edges)."""
    kind, f = mesh.elements[eid]
    if kind != "H":
        raise ValueError(f"element {eid} is not a hexahedron")
    e = [None] * 12
    for k in range(1, 5):
        e[k - 1] = _shared_edge(mesh, f[0], f[k])
    e[4] = _shared_edge(mesh, f[1], f[4])
    for k in range(1, 4):
        e[k + 4] = _shared_edge(mesh, f[k], f[k + 1])
    for k in range(1, 5):
        e[k + 7] = _shared_edge(mesh, f[5], f[k])
    ev = [mesh.edges[x] for x in e]
    v = [None] * 8
    if ev[0][0] in ev[1]:
        v[0], v[1] = ev[0][1], ev[0][0]
    elif ev[0][1] in ev[1]:
        v[0], v[1] = ev[0][0], ev[0][1]
    else:
        raise ValueError(f"element {eid}: edges {e[0]} and {e[1]} share no vertex")
    for i in (1, 2):
        if ev[i][0] == v[i]:
            v[i + 1] = ev[i][1]
        elif ev[i][1] == v[i]:
            v[i + 1] = ev[i][0]
        else:
            raise ValueError(f"element {eid}: edge {e[i]} does not continue the bottom face")
    if ev[8][0] in ev[9]:
        v[4], v[5] = ev[8][1], ev[8][0]
    elif ev[8][1] in ev[9]:
        v[4], v[5] = ev[8][0], ev[8][1]
    else:
        raise ValueError(f"element {eid}: edges {e[8]} and {e[9]} share no vertex")
    for i in (9, 10):
        if ev[i][0] == v[i - 4]:
            v[i - 3] = ev[i][1]
        elif ev[i][1] == v[i - 4]:
            v[i - 3] = ev[i][0]
        else:
            raise ValueError(f"element {eid}: edge {e[i]} does not continue the top face")
    return v


def hex_elements(mesh: MeshGraph):
    """(element ids, local vertex ids [E, 8], local vertex coordinates
    [E, 8, 3]) of the mesh's hexahedra in increasing id."""
    eids = sorted(e for e, (k, _) in mesh.elements.items() if k == "H")
    lv = np.array([hex_local_vertices(mesh, e) for e in eids], dtype=np.int64).reshape(-1, 8)
    verts = np.array([[mesh.vertices[v] for v in row] for row in lv]).reshape(-1, 8, 3)
    return eids, lv, verts


def tet_local_vertices(mesh: MeshGraph, eid: int) -> list[int]:
    """Global ids of the 4 local vertices of tetrahedron eid (TetGeom.cpp:
    113-230 local edges from the faces, :232-300 vertices from the edges):
    edges 0-2 are the edges of face 0 that are the first edge of faces 1-3,
    edge 3 is shared by faces 1 and 3; v0, v1 from edge 0 (v1 on edge 1),
    v2 the far end of edge 1, v3 the far end of edge 3 from v0."""
    kind, f = mesh.elements[eid]
    if kind != "A":
        raise ValueError(f"element {eid} is not a tetrahedron")
    fe = [mesh.faces[x][1] for x in f]
    if any(mesh.faces[x][0] != "T" for x in f):
        raise ValueError(f"element {eid}: tetrahedron faces must be triangles")
    if fe[0][0] != fe[1][0]:
        raise ValueError(f"element {eid}: the first edges of faces 0 and 1 differ")
    e = [None] * 4
    for k in range(1, 4):
        hits = [x for x in fe[0] if x == fe[k][0]]
        if len(hits) != 1:
            raise ValueError(f"element {eid}: face {f[k]}'s first edge is not on face {f[0]}")
        e[k - 1] = hits[0]
    e[3] = _shared_edge(mesh, f[1], f[3])
    ev = [mesh.edges[x] for x in e]
    if ev[0][0] in ev[1]:
        v0, v1 = ev[0][1], ev[0][0]
    elif ev[0][1] in ev[1]:
        v0, v1 = ev[0][0], ev[0][1]
    else:
        raise ValueError(f"element {eid}: edges {e[0]} and {e[1]} share no vertex")
    if ev[1][0] == v1:
        v2 = ev[1][1]
    elif ev[1][1] == v1:
        v2 = ev[1][0]
    else:
        raise ValueError(f"element {eid}: edge {e[1]} does not continue edge {e[0]}")
    v3 = ev[3][1] if ev[3][0] == v0 else ev[3][0]
    return [v0, v1, v2, v3]


def tet_elements(mesh: MeshGraph):
    """(element ids, local vertex ids [E, 4], local vertex coordinates
    [E, 4, 3]) of the mesh's tetrahedra in increasing id."""
    eids = sorted(e for e, (k, _) in mesh.elements.items() if k == "A")
    lv = np.array([tet_local_vertices(mesh, e) for e in eids], dtype=np.int64).reshape(-1, 4)
    verts = np.array([[mesh.vertices[v] for v in row] for row in lv]).reshape(-1, 4, 3)
    return eids, lv, verts


def tet_geometry(verts: np.ndarray):
    """(jac, df) of straight-sided tetrahedra verts [E, 4, 3] (local vertex
    order; reference vertices (-1,-1,-1), (1,-1,-1), (-1,1,-1), (-1,-1,1)):
    x = v0 + sum_d (xi_d + 1)/2 (v_{d+1} - v0), J = |det(dx/dxi)|,
    df[j*3 + k] = dxi_k/dx_j, one set per element."""
    verts = np.asarray(verts, dtype=float)
    A = np.stack([verts[:, 1] - verts[:, 0], verts[:, 2] - verts[:, 0], verts[:, 3] - verts[:, 0]], axis=2) / 2
    inv = np.linalg.inv(A)
    df = np.transpose(inv, (0, 2, 1)).reshape(-1, 9)
    return np.abs(np.linalg.det(A)), np.ascontiguousarray(df.reshape(-1))


# reference coordinates of the hex's local vertices (StdHexExp)
HEX_REF = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                    [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)


def hex_geometry(verts: np.ndarray, tol: float = 1e-12):
    """(jac, df) of straight-sided parallelepiped hexahedra verts [E, 8, 3]
    (local vertex order): J = |det(dx/dxi)|, df[j*3 + k] = dxi_k/dx_j, one
    set per element.  Raises for elements that are not parallelepipeds (the
    deformed per-point factors are not built here)."""
    verts = np.asarray(verts, dtype=float)
    A = np.stack([verts[:, 1] - verts[:, 0], verts[:, 3] - verts[:, 0], verts[:, 4] - verts[:, 0]], axis=2) / 2
    # parallelepiped: every vertex is v0 + A (xi + 1)
    pred = verts[:, :1, :] + np.einsum("eki,vi->evk", A, HEX_REF + 1)
    scale = np.max(np.abs(verts)) + 1.0
    if np.max(np.abs(pred - verts)) > tol * scale:
        raise NotImplementedError("non-parallelepiped hexahedra need per-point factors (hex_geometry_points)")
    det = np.linalg.det(A)
    inv = np.linalg.inv(A)  # inv[e, k, j] = dxi_k / dx_j
    df = np.transpose(inv, (0, 2, 1)).reshape(-1, 9)  # df[j*3 + k]
    return np.abs(det), np.ascontiguousarray(df.reshape(-1))


def hex_is_parallelepiped(verts: np.ndarray, tol: float = 1e-12) -> np.ndarray:
    """Per element: the trilinear map has no xi-product terms, the regular
    geometry type of HexGeom::v_CalcGeomType (HexGeom.cpp:83-153)."""
    verts = np.asarray(verts, dtype=float)
    A = np.stack([verts[:, 1] - verts[:, 0], verts[:, 3] - verts[:, 0], verts[:, 4] - verts[:, 0]], axis=2) / 2
    pred = verts[:, :1, :] + np.einsum("eki,vi->evk", A, HEX_REF + 1)
    scale = np.max(np.abs(verts)) + 1.0 if verts.size else 1.0
    return np.max(np.abs(pred - verts), axis=(1, 2)) <= tol * scale


def hex_map_points(verts: np.ndarray, z: np.ndarray):
    """The trilinear map of hexahedra verts [E, 8, 3] at the tensor points
    (z_i, z_j, z_k), index (k Q + j) Q + i: coordinates x [E, Q^3, 3] and
    derivatives dx[E, Q^3, 3 (x_k), 3 (xi_d)]."""
    verts = np.asarray(verts, dtype=float)
    z = np.asarray(z, dtype=float)
    K, J, I = np.meshgrid(z, z, z, indexing="ij")
    xi = np.stack([I.reshape(-1), J.reshape(-1), K.reshape(-1)], axis=1)  # [Q^3, 3]
    s = HEX_REF  # [8, 3] vertex signs
    f = 1.0 + xi[:, None, :] * s[None, :, :]  # [Q^3, 8, 3]
    N = np.prod(f, axis=2) / 8.0  # shape functions [Q^3, 8]
    dN = np.empty(f.shape)
    for d in range(3):
        others = [c for c in range(3) if c != d]
        dN[..., d] = s[None, :, d] * f[..., others[0]] * f[..., others[1]] / 8.0
    x = np.einsum("qv,evk->eqk", N, verts)
    dx = np.einsum("qvd,evk->eqkd", dN, verts)
    return x, dx


def hex_geometry_points(verts: np.ndarray, z: np.ndarray):
    """(jac, df) at every quadrature point of trilinear hexahedra (the
    deformed factors of GeomFactors.cpp:150-477): jac[e*Q^3 + q] =
    |det(dx/dxi)|, df[(e*Q^3 + q)*9 + j*3 + k] = dxi_k/dx_j."""
    _, dx = hex_map_points(verts, z)
    E, nq3 = dx.shape[:2]
    A = dx.reshape(-1, 3, 3)  # A[k, d] = dx_k / dxi_d
    inv = np.linalg.inv(A)  # inv[d, k] = dxi_d / dx_k
    df = np.transpose(inv, (0, 2, 1)).reshape(-1, 9)  # df[k*3 + d]
    return np.abs(np.linalg.det(A)), np.ascontiguousarray(df.reshape(-1))


def hex_c0_map(local_vertices: np.ndarray, nm: int, dirichlet_vertex_sets=()):
    """C0 map of a conforming hex mesh from the elements' global vertex ids
    local_vertices [E, 8] (local order).  Returns (map, sign, nglobal, ndir)
    over the element-local coefficients (p fastest), numbered Dirichlet DOFs
    first, then vertices, edges, faces and interiors.

    Edge modes p >= 2 of a local edge running against its canonical
    direction (low to high global vertex id) take the sign (-1)^p
    (phi_p(-z) = (-1)^p phi_p(z)).  A face's canonical frame starts at its
    lowest-id corner, axis A towards the lower-id neighbour, axis B towards
    the other; a local face mode (p1, p2) maps to (p1, p2) or (p2, p1) on
    (A, B) and takes (-1)^p for each reversed axis.  dirichlet_vertex_sets:
    global vertex id sets of boundary faces whose DOFs are Dirichlet."""
    lv = np.asarray(local_vertices)
    E = len(lv)
    nb = nm - 2
    corner = {(a, b, c): i for i, (a, b, c) in enumerate(((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                                                          (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)))}
    dset = [frozenset(s) for s in dirichlet_vertex_sets]
    keys: dict = {}  # entity key -> (first id, kind)
    nxt = [0]

    def ident(key, count):
        if key not in keys:
            keys[key] = nxt[0]
            nxt[0] += count
        return keys[key]

    mp = np.zeros((E, nm, nm, nm), dtype=np.int64)  # [e, r, q, p]
    sg = np.ones((E, nm, nm, nm))
    ent = np.empty((E, nm, nm, nm), dtype=object)  # entity vertex sets, for Dirichlet marking
    for e in range(E):
        g = lv[e]
        for r in range(nm):
            for q in range(nm):
                for p in range(nm):
                    idx = (p, q, r)
                    bub = [d for d in range(3) if idx[d] >= 2]
                    if len(bub) == 0:
                        v = int(g[corner[idx]])
                        mp[e, r, q, p] = ident(("v", v), 1)
                        ent[e, r, q, p] = frozenset([v])
                    elif len(bub) == 1:
                        d = bub[0]
                        lo, hi = list(idx), list(idx)
                        lo[d], hi[d] = 0, 1
                        va, vb = int(g[corner[tuple(lo)]]), int(g[corner[tuple(hi)]])
                        base = ident(("e", min(va, vb), max(va, vb)), nb)
                        mp[e, r, q, p] = base + idx[d] - 2
                        sg[e, r, q, p] = 1.0 if va < vb else (-1.0) ** idx[d]
                        ent[e, r, q, p] = frozenset([va, vb])
                    elif len(bub) == 2:
                        d1, d2 = bub
                        fixed = [d for d in range(3) if d not in bub][0]
                        cg = {}
                        for s1 in (0, 1):
                            for s2 in (0, 1):
                                c = list(idx)
                                c[d1], c[d2] = s1, s2
                                cg[(s1, s2)] = int(g[corner[tuple(c)]])
                        c0 = min(cg, key=cg.get)
                        n1 = (1 - c0[0], c0[1])  # neighbour along d1
                        n2 = (c0[0], 1 - c0[1])  # neighbour along d2
                        d1_is_A = cg[n1] < cg[n2]
                        rev1, rev2 = c0[0] == 1, c0[1] == 1
                        p1, p2 = idx[d1], idx[d2]
                        a, b = (p1, p2) if d1_is_A else (p2, p1)
                        base = ident(("f",) + tuple(sorted(cg.values())), nb * nb)
                        mp[e, r, q, p] = base + (a - 2) * nb + (b - 2)
                        sg[e, r, q, p] = ((-1.0) ** p1 if rev1 else 1.0) * ((-1.0) ** p2 if rev2 else 1.0)
                        ent[e, r, q, p] = frozenset(cg.values())
                        del fixed
                    else:
                        mp[e, r, q, p] = ident(("i", e), nb ** 3) + ((r - 2) * nb + (q - 2)) * nb + (p - 2)
                        ent[e, r, q, p] = None
    nglob = nxt[0]
    dirichlet = np.zeros(nglob, dtype=bool)
    if dset:
        flat_m, flat_e = mp.reshape(-1), ent.reshape(-1)
        for m, s in zip(flat_m, flat_e):
            if s is not None and any(s <= f for f in dset):
                dirichlet[m] = True
    order = np.concatenate([np.flatnonzero(dirichlet), np.flatnonzero(~dirichlet)])
    new = np.empty(nglob, dtype=np.int64)
    new[order] = np.arange(nglob)
    return new[mp.reshape(-1)], sg.reshape(-1), nglob, int(dirichlet.sum())


@dataclass
class FieldBlock:
    fields: list
    shape: str
    basis: list
    nummodes: list
    ids: list
    data: dict  # field -> array [len(ids), ncoeff]


def read_fld(path: str) -> list[FieldBlock]:
    """ELEMENTS blocks of an XML field file (FieldIOXml.cpp): coefficients
    stored field by field, each field element by element in the ID order."""
    root = ET.parse(path).getroot()
    blocks = []
    for el in root.iter("ELEMENTS"):
        fields = [f.strip() for f in el.get("FIELDS").split(",")]
        nm = el.get("NUMMODESPERDIR")
        if not nm.startswith("UNIORDER:"):
            raise NotImplementedError("only UNIORDER expansions are read")
        nummodes = [int(x) for x in nm.split(":", 1)[1].split(",")]
        ids = parse_ranges(el.get("ID"))
        if el.get("COMPRESSED"):
            if el.get("COMPRESSED") != _B64Z or el.get("BITSIZE", "64") != "64":
                raise ValueError(f"unsupported compression {el.get('COMPRESSED')}")
            vals = _b64z_decode(el.text, "<f8")
        else:
            vals = np.array([float(x) for x in el.text.split()])
        total = len(vals) // len(fields)
        if total * len(fields) != len(vals) or total % len(ids):
            raise ValueError("ELEMENTS data size does not match its fields and elements")
        nc = total // len(ids)
        data = {f: vals[i * total:(i + 1) * total].reshape(len(ids), nc).copy() for i, f in enumerate(fields)}
        blocks.append(FieldBlock(fields, el.get("SHAPE"), el.get("BASIS").split(","), nummodes, ids, data))
    return blocks


def write_fld(path: str, blocks: list[FieldBlock], metadata: dict | None = None) -> None:
    """An XML field file with the blocks' coefficients, B64Z-LittleEndian
    compressed (FieldIOXml.cpp's default)."""
    root = ET.Element("NEKTAR")
    md = ET.SubElement(root, "Metadata")
    for k, v in (metadata or {}).items():
        ET.SubElement(md, k).text = str(v)
    for b in blocks:
        vals = np.concatenate([np.asarray(b.data[f], dtype="<f8").reshape(-1) for f in b.fields])
        el = ET.SubElement(root, "ELEMENTS", {
            "FIELDS": ",".join(b.fields), "SHAPE": b.shape, "BASIS": ",".join(b.basis),
            "NUMMODESPERDIR": "UNIORDER:" + ",".join(str(n) for n in b.nummodes),
            "ID": _ranges_string(b.ids), "COMPRESSED": _B64Z, "BITSIZE": "64"})
        el.text = _b64z_encode(vals)
    ET.ElementTree(root).write(path, encoding="utf-8", xml_declaration=True)


def _tet_index(nm):
    """(p, q, r) -> coefficient index of the tet order (p slowest, r fastest)."""
    idx, k = {}, 0
    for p in range(nm):
        for q in range(nm - p):
            for r in range(nm - p - q):
                idx[(p, q, r)] = k
                k += 1
    return idx


def _tet_trace_modes(nm):
    """The tet's local edges and faces with the modes that live on them
    (StdTetExp.cpp:1459-1660): edges (start, end, modes along the edge);
    faces (first-direction start, end, apex, [(mode, flips when the first
    direction is reversed)]); vertex modes in local vertex order."""
    edges = [(0, 1, [(i + 2, 0, 0) for i in range(nm - 2)]), (1, 2, [(1, i + 1, 0) for i in range(nm - 2)]),
             (0, 2, [(0, i + 2, 0) for i in range(nm - 2)]), (0, 3, [(0, 0, i + 2) for i in range(nm - 2)]),
             (1, 3, [(1, 0, i + 1) for i in range(nm - 2)]), (2, 3, [(0, 1, i + 1) for i in range(nm - 2)])]
    f0 = [((i, j, 0), i % 2 == 1) for i in range(2, nm) for j in range(1, nm - i)]
    f1 = [((i, 0, k), i % 2 == 1) for i in range(2, nm) for k in range(1, nm - i)]
    f2 = [((1, j, k), (j + 1) % 2 == 1) for j in range(1, nm - 1) for k in range(1, nm - 1 - j)]
    f3 = [((0, j, k), j % 2 == 1) for j in range(2, nm) for k in range(1, nm - j)]
    faces = [(0, 1, 2, f0), (0, 1, 3, f1), (1, 2, 3, f2), (0, 2, 3, f3)]
    vmodes = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]
    return edges, faces, vmodes


def tet_mode_local_entities(nm):
    """For every coefficient of the tet order, the set of local vertices of
    the entity its mode belongs to (vertex, edge or face), None for
    interior modes."""
    ix = _tet_index(nm)
    edges, faces, vmodes = _tet_trace_modes(nm)
    out = [None] * len(ix)
    for v, m in enumerate(vmodes):
        out[ix[m]] = frozenset([v])
    for a, b, modes in edges:
        for m in modes:
            out[ix[m]] = frozenset([a, b])
    for a, b, c, modes in faces:
        for m, _ in modes:
            out[ix[m]] = frozenset([a, b, c])
    return out


def tet_c0_map(local_vertices: np.ndarray, nm: int, dirichlet_vertex_sets=()):
    """C0 map of a conforming tet mesh from the elements' global vertex ids
    local_vertices [E, 4] (TetGeom order).  Returns (map, sign, nglobal,
    ndir) over the element-local coefficients (tet order), Dirichlet DOFs
    first, then vertices, edges, faces and interiors.

    Follows the reference's trace maps (StdTetExp.cpp:1459-1660): the
    interior modes of local edge k are listed along the edge and a backward
    edge (here: running from the higher to the lower global vertex id)
    flips every odd one; a triangular face keeps its apex (the collapsed
    direction) and may reverse its first direction, which flips the modes
    with an odd first index (for face 2 an even j); face modes correspond
    index by index between the two elements.  A face whose apex differs
    between its elements is rejected, as the reference does."""
    lv = np.asarray(local_vertices)
    E = len(lv)
    ix = _tet_index(nm)
    nc = len(ix)
    edges, faces, vmodes = _tet_trace_modes(nm)
    keys: dict = {}
    apex_of: dict = {}
    nxt = [0]

    def ident(key, count):
        if key not in keys:
            keys[key] = nxt[0]
            nxt[0] += count
        return keys[key]

    mp = np.full((E, nc), -1, dtype=np.int64)
    sg = np.ones((E, nc))
    ent = [[None] * nc for _ in range(E)]
    for e in range(E):
        g = [int(x) for x in lv[e]]
        for v, m in enumerate(vmodes):
            mp[e, ix[m]] = ident(("v", g[v]), 1)
            ent[e][ix[m]] = frozenset([g[v]])
        for a, b, modes in edges:
            base = ident(("e", min(g[a], g[b]), max(g[a], g[b])), nm - 2)
            back = g[a] > g[b]
            for i, m in enumerate(modes):
                mp[e, ix[m]] = base + i
                sg[e, ix[m]] = -1.0 if back and i % 2 == 1 else 1.0
                ent[e][ix[m]] = frozenset([g[a], g[b]])
        for a, b, c, modes in faces:
            key = ("f", frozenset([g[a], g[b], g[c]]))
            if key in apex_of and apex_of[key] != g[c]:
                raise ValueError(f"element {e}: a face is met with a different apex by its neighbour "
                                 "(tet vertex order not consistent)")
            apex_of[key] = g[c]
            base = ident(key, len(modes))
            back = g[a] > g[b]
            for i, (m, odd) in enumerate(modes):
                mp[e, ix[m]] = base + i
                sg[e, ix[m]] = -1.0 if back and odd else 1.0
                ent[e][ix[m]] = frozenset([g[a], g[b], g[c]])
        rest = [k for k in range(nc) if mp[e, k] < 0]
        base = ident(("i", e), len(rest))
        for i, k in enumerate(rest):
            mp[e, k] = base + i
    nglob = nxt[0]
    dirichlet = np.zeros(nglob, dtype=bool)
    dset = [frozenset(s) for s in dirichlet_vertex_sets]
    if dset:
        for e in range(E):
            for k in range(nc):
                s = ent[e][k]
                if s is not None and any(s <= f for f in dset):
                    dirichlet[mp[e, k]] = True
    order = np.concatenate([np.flatnonzero(dirichlet), np.flatnonzero(~dirichlet)])
    new = np.empty(nglob, dtype=np.int64)
    new[order] = np.arange(nglob)
    return new[mp.reshape(-1)], sg.reshape(-1), nglob, int(dirichlet.sum())


def c0_helmholtz_smoothing(session_path: str, fld_path: str, length: float, tol: float = 1e-10,
                           maxit: int = 5000, numpoints: int | None = None):
    """FieldConvert's ``-m C0Projection:helmsmoothing=L`` on a hexahedral or
    tetrahedral mesh (ProcessC0Projection.cpp:186-231): every field u of the
    field file is replaced by the C0 solution of (L + lam M) u_s = lam M u
    with lam = (2 pi / L)^2 and no Dirichlet DOFs (the module builds its C0
    expansion without boundary conditions), through the MI355X operators
    (BwdTrans, IProductWRTBase, Helmholtz) and the general CG.  numpoints:
    GLL points per direction (default NUMMODES + 1, the MODIFIED default;
    hexahedra also take NUMMODES).  Returns (FieldBlock of the smoothed
    coefficients, {field: SolveInfo})."""
    import torch

    from . import AssemblyMap, FP64, GetOperatorFactory, LinearSystem

    mesh = read_mesh_xml(session_path)
    kinds = {k for k, _ in mesh.elements.values()}
    geo = "Regular"
    if kinds == {"H"}:
        eids, lv, verts = hex_elements(mesh)
        tag, shape_name = "Hex", "Hexahedron"
        if np.all(hex_is_parallelepiped(verts)):
            jac, df = hex_geometry(verts)
        else:  # trilinear elements: factors at every quadrature point
            from . import SHAPE_HEX, basis_tables

            nq_ = numpoints or mesh.expansions[0]["nummodes"] + 1
            jac, df = hex_geometry_points(verts, basis_tables(SHAPE_HEX, 0, mesh.expansions[0]["nummodes"], nq_)["z"])
            geo = "Deformed"
    elif kinds == {"A"}:
        eids, lv, verts = tet_elements(mesh)
        tag, shape_name = "Tet", "Tetrahedron"
        jac, df = tet_geometry(verts)
    else:
        raise NotImplementedError("c0_helmholtz_smoothing takes meshes of hexahedra or of tetrahedra")
    nm = mesh.expansions[0]["nummodes"]
    nq, E = (numpoints or nm + 1), len(eids)
    mp, sg, ng, nd = (hex_c0_map if tag == "Hex" else tet_c0_map)(lv, nm)
    blocks = [b for b in read_fld(fld_path) if b.shape == shape_name]
    if len(blocks) != 1 or blocks[0].nummodes != [nm] * 3:
        raise NotImplementedError("one ELEMENTS block of the mesh's shape with the session's NUMMODES is expected")
    blk = blocks[0]
    pos = {e: i for i, e in enumerate(blk.ids)}
    lam = (2.0 * np.pi / length) ** 2
    F = GetOperatorFactory()
    bwd = F.CreateInstance(f"BwdTrans_{tag}_Regular", nm, nq, E, FP64)
    ipr = F.CreateInstance(f"IProduct_{tag}_{geo}", nm, nq, E, FP64)
    ipr.SetJac(jac)
    helm = F.CreateInstance(f"Helmholtz_{tag}_{geo}", nm, nq, E, FP64)
    helm.SetJac(jac)
    helm.SetDF(df)
    helm.SetLambda(lam)
    ls = LinearSystem(AssemblyMap(mp, sg, ng, nd), [helm])
    nc = helm.input_size() // max(E, 1)
    out, infos = {}, {}
    uq = torch.zeros(bwd.output_size(), dtype=torch.float64, device="cuda")
    bl = torch.zeros(E * nc, dtype=torch.float64, device="cuda")
    xs = torch.zeros(E * nc, dtype=torch.float64, device="cuda")
    for f in blk.fields:
        c = np.ascontiguousarray(np.array([blk.data[f][pos[e]] for e in eids]).reshape(-1))
        bwd(torch.from_numpy(c).cuda(), uq)
        ipr(uq, bl)
        ls.set_rhs_local(lam * bl)  # HelmSolve's rhs: -IProductWRTBase(-lam u)
        infos[f] = ls.solve(tol=tol, maxit=maxit)
        ls.solution_local(xs)
        torch.cuda.synchronize()
        out[f] = xs.cpu().numpy().reshape(E, nc).copy()
    return FieldBlock(list(blk.fields), blk.shape, list(blk.basis), list(blk.nummodes), list(eids), out), infos
