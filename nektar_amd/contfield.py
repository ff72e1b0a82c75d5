"""C0 Helmholtz solve of a Nektar++ session file with Dirichlet conditions
(host set-up, GPU operators and CG): what the reference's Helmholtz3D demo
does through ContField (library/Demos/MultiRegions/Helmholtz3D.cpp).

* ``Session``: PARAMETERS, BOUNDARYREGIONS, BOUNDARYCONDITIONS and FUNCTION
  blocks of the session XML (SessionReader), with a safe evaluator of the
  reference's expression syntax (LibUtilities/Interpreter: +, -, *, /, ^,
  the usual functions, PI, parameters, x, y, z, t).
* Dirichlet data: the boundary condition evaluated at the boundary
  expansion's quadrature points and projected with vertices interpolated,
  edges and face interiors L2-projected with what is already fixed held
  (ExpList::FwdTransBndConstrained, ExpList.cpp:2420; SegExp.cpp:81-180,
  QuadExp.cpp:67-160), then imposed on the element-local coefficients
  (ContField::v_ImposeDirichletConditions, ContField.cpp:557-608).
* ``helmholtz_solve``: ContField::v_HelmSolve (ContField.cpp:797-925) --
  rhs = -IProductWRTBase(f), the Dirichlet lifting of
  GlobalLinSysIterativeFull::v_Solve (GlobalLinSysIterativeFull.cpp:
  118-169) inside the GPU linear system, BwdTrans of the solution and the
  L2 / Linf errors against the exact solution at the quadrature points
  (ExpList::L2 / Linf, ExpList.cpp:3836-3863).
"""
from __future__ import annotations

import ast
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

import numpy as np

from . import io as nio

_FUNCS = {
    "sin": np.sin, "cos": np.cos, "tan": np.tan, "asin": np.arcsin, "acos": np.arccos, "atan": np.arctan,
    "atan2": np.arctan2, "sinh": np.sinh, "cosh": np.cosh, "tanh": np.tanh, "exp": np.exp, "log": np.log,
    "log10": np.log10, "sqrt": np.sqrt, "abs": np.abs, "fabs": np.abs, "pow": np.power, "min": np.minimum,
    "max": np.maximum, "floor": np.floor, "ceil": np.ceil, "sign": np.sign,
}
_CONSTS = {"PI": np.pi, "E": np.e}
_BINOPS = {ast.Add: np.add, ast.Sub: np.subtract, ast.Mult: np.multiply, ast.Div: np.divide, ast.Pow: np.power}


def evaluate(expr: str, names: dict):
    """Value of an expression of the session syntax; names maps variables
    and parameters to numbers or arrays.  Only arithmetic, the functions
    above and known names are accepted (no Python evaluation)."""
    tree = ast.parse(expr.replace("^", "**").strip(), mode="eval")

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)):
            return float(n.value)
        if isinstance(n, ast.BinOp) and type(n.op) in _BINOPS:
            return _BINOPS[type(n.op)](ev(n.left), ev(n.right))
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, (ast.USub, ast.UAdd)):
            v = ev(n.operand)
            return -v if isinstance(n.op, ast.USub) else v
        if isinstance(n, ast.Name):
            if n.id in names:
                return names[n.id]
            if n.id in _CONSTS:
                return _CONSTS[n.id]
            raise ValueError(f"unknown name '{n.id}' in expression '{expr}'")
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and n.func.id in _FUNCS and not n.keywords:
            return _FUNCS[n.func.id](*[ev(a) for a in n.args])
        raise ValueError(f"unsupported construct in expression '{expr}'")

    return ev(tree)


@dataclass
class BoundaryCondition:
    kind: str      # 'D' (Dirichlet), 'N' (Neumann), 'R' (Robin)
    var: str
    value: str
    primcoeff: str | None = None


@dataclass
class Session:
    mesh: nio.MeshGraph
    params: dict = field(default_factory=dict)
    variables: list = field(default_factory=list)
    regions: dict = field(default_factory=dict)      # region id -> composite ids
    conditions: dict = field(default_factory=dict)   # region id -> [BoundaryCondition]
    functions: dict = field(default_factory=dict)    # name -> {var: expression}

    def function(self, name: str, var: str, x, y, z, t: float = 0.0):
        expr = self.functions[name][var]
        return np.broadcast_to(evaluate(expr, dict(self.params, x=x, y=y, z=z, t=t)), np.shape(x)).astype(float)


def read_session(path: str) -> Session:
    """Mesh plus the CONDITIONS block of a session file (SessionReader:
    ReadParameters, ReadVariables, ReadBoundaryRegions, ReadBoundary-
    Conditions, ReadFunctions)."""
    s = Session(nio.read_mesh_xml(path))
    cond = ET.parse(path).getroot().find("CONDITIONS")
    if cond is None:
        return s
    par = cond.find("PARAMETERS")
    for p in par.findall("P") if par is not None else []:
        name, _, value = p.text.partition("=")
        s.params[name.strip()] = float(evaluate(value, dict(s.params)))
    var = cond.find("VARIABLES")
    s.variables = [v.text.strip() for v in var.findall("V")] if var is not None else []
    reg = cond.find("BOUNDARYREGIONS")
    for b in reg.findall("B") if reg is not None else []:
        s.regions[int(b.get("ID"))] = [i for _, ids in nio._composite_refs(b.text) for i in ids]
    bcs = cond.find("BOUNDARYCONDITIONS")
    for r in bcs.findall("REGION") if bcs is not None else []:
        s.conditions[int(r.get("REF"))] = [BoundaryCondition(c.tag, c.get("VAR"), c.get("VALUE"), c.get("PRIMCOEFF"))
                                           for c in r]
    for f in cond.findall("FUNCTION"):
        s.functions[f.get("NAME")] = {e.get("VAR"): e.get("VALUE") for e in f.findall("E")}
    return s


# ---------------------------------------------------------------------
# Constrained projection (FwdTransBndConstrained) in an element's frame

def _project_1d(B, w, f, v0, v1):
    """Modified_A coefficients of f sampled at the points (table B[m, i],
    weights w) with the vertex coefficients fixed to v0, v1 and the
    interior ones L2-projected (SegExp.cpp:150-180).  A constant Jacobian
    multiplies both sides and cancels."""
    nm = B.shape[0]
    c = np.zeros(nm)
    c[0], c[1] = v0, v1
    if nm > 2:
        M = (B * w) @ B.T
        rhs = (B * w) @ f - M @ c
        c[2:] = np.linalg.solve(M[2:, 2:], rhs[2:])
    return c


def _project_quad(B, w, F):
    """Modified_A x Modified_A coefficients C[a, b] of F[i, j] (points of
    direction a fastest along i) by the constrained projection of a
    quadrilateral (QuadExp.cpp:67-160): corners from the corner values,
    each edge by the 1-D projection of its values, the interior modes
    L2-projected with the boundary ones held."""
    nm = B.shape[0]
    C = np.zeros((nm, nm))
    C[0, 0], C[1, 0], C[0, 1], C[1, 1] = F[0, 0], F[-1, 0], F[0, -1], F[-1, -1]
    C[:, 0] = _project_1d(B, w, F[:, 0], F[0, 0], F[-1, 0])
    C[:, 1] = _project_1d(B, w, F[:, -1], F[0, -1], F[-1, -1])
    C[0, :] = _project_1d(B, w, F[0, :], F[0, 0], F[0, -1])
    C[1, :] = _project_1d(B, w, F[-1, :], F[-1, 0], F[-1, -1])
    if nm > 2:
        B2 = np.einsum("ai,bj->abij", B, B).reshape(nm * nm, -1)
        W2 = np.outer(w, w).reshape(-1)
        M = (B2 * W2) @ B2.T
        rhs = (B2 * W2) @ F.reshape(-1) - M @ C.reshape(-1)
        inner = np.array([a * nm + b for a in range(2, nm) for b in range(2, nm)])
        C.reshape(-1)[inner] = np.linalg.solve(M[np.ix_(inner, inner)], rhs[inner])
    return C


def _hex_face_frames(lv_e):
    """Local faces of a hexahedron as {frozenset(global vertex ids):
    (fixed direction d, side s)}: the face xi_d = s (s = -1 or +1)."""
    frames = {}
    for d in range(3):
        for s in (-1, 1):
            ids = [v for k, v in enumerate(lv_e) if nio.HEX_REF[k][d] == s]
            frames[frozenset(int(v) for v in ids)] = (d, s)
    return frames


def _face_vertices(mesh: nio.MeshGraph, fid: int) -> frozenset:
    return frozenset(v for e in mesh.faces[fid][1] for v in mesh.edges[e])


def hex_dirichlet_local(session: Session, var: str, lv, verts, nm: int, nq: int, mp, sg, ndir: int,
                        tables):
    """Element-local Dirichlet coefficients of a hexahedral mesh (zero off
    the Dirichlet DOFs): every Dirichlet face projected in the frame of the
    element that owns it, the values then taken by every local copy of the
    Dirichlet DOFs through the map (the reference's copy of boundary values
    into the elements that only touch the boundary,
    AssemblyMapCG.cpp:1955-2010)."""
    mesh = session.mesh
    B, z, w = tables
    E = len(lv)
    owner = {}
    for e in range(E):
        for key, frame in _hex_face_frames(lv[e]).items():
            owner.setdefault(key, []).append((e, frame))
    nc = nm ** 3
    dglob = np.full(int(np.max(mp)) + 1 if len(mp) else 0, np.nan)
    for rid, bcs in session.conditions.items():
        for bc in bcs:
            if bc.var != var or bc.kind != "D":
                continue
            for cid in session.regions[rid]:
                for letter, ids in mesh.composites[cid]:
                    if letter != "F":
                        raise NotImplementedError("Dirichlet regions of hexahedral meshes are made of faces")
                    for fid in ids:
                        own = owner.get(_face_vertices(mesh, fid), [])
                        if len(own) != 1:
                            raise ValueError(f"boundary face {fid} belongs to {len(own)} elements")
                        e, (d, s) = own[0]
                        a, b = [k for k in range(3) if k != d]
                        # face points (xi_a = z_i, xi_b = z_j, xi_d = s) in space
                        xi = np.zeros((nq, nq, 3))
                        xi[:, :, a] = z[:, None]
                        xi[:, :, b] = z[None, :]
                        xi[:, :, d] = s
                        A = np.stack([verts[e, 1] - verts[e, 0], verts[e, 3] - verts[e, 0],
                                      verts[e, 4] - verts[e, 0]], axis=1) / 2
                        X = verts[e, 0] + (xi + 1.0) @ A.T
                        F = evaluate(bc.value, dict(session.params, x=X[..., 0], y=X[..., 1], z=X[..., 2], t=0.0))
                        F = np.broadcast_to(F, X.shape[:2]).astype(float)
                        C = _project_quad(B, w, F)
                        md = 0 if s < 0 else 1
                        for pa in range(nm):
                            for pb in range(nm):
                                idx = [0, 0, 0]
                                idx[a], idx[b], idx[d] = pa, pb, md
                                li = e * nc + (idx[2] * nm + idx[1]) * nm + idx[0]
                                g = mp[li]
                                if g >= ndir:
                                    raise ValueError("a Dirichlet face mode maps to a non-Dirichlet DOF")
                                dglob[g] = sg[li] * C[pa, pb]
    if ndir and np.isnan(dglob[:ndir]).any():
        raise ValueError("some Dirichlet DOFs received no boundary value")
    dloc = np.where(mp < ndir, sg * np.nan_to_num(dglob)[mp], 0.0)
    return dloc


def hex_quadrature_points(verts, z):
    """Physical coordinates [E, Q^3, 3] of the GLL points z of
    parallelepiped hexahedra verts [E, 8, 3]: x = v0 + A (xi + 1), point
    (i, j, k) at index (k Q + j) Q + i."""
    xi = np.stack(np.meshgrid(z, z, z, indexing="ij"), axis=-1).transpose(2, 1, 0, 3).reshape(-1, 3)
    A = np.stack([verts[:, 1] - verts[:, 0], verts[:, 3] - verts[:, 0], verts[:, 4] - verts[:, 0]], axis=2) / 2
    return verts[:, None, 0, :] + np.einsum("ekd,qd->eqk", A, xi + 1.0)


def errors(u, exact, jac, w):
    """ExpList::L2 and Linf (ExpList.cpp:3836-3863) of u - exact at the
    quadrature points of hexahedra with one Jacobian each: sqrt(sum_e
    sum_q W_q J_e (u - exact)^2) and max |u - exact|."""
    wq = np.einsum("k,j,i->kji", w, w, w).reshape(-1)
    err = np.asarray(u).reshape(len(jac), -1) - np.asarray(exact).reshape(len(jac), -1)
    return float(np.sqrt(np.sum(wq[None, :] * np.asarray(jac)[:, None] * err * err))), float(np.max(np.abs(err)))


def helmholtz_solve(session_path: str, var: str = "u", tol: float = 1e-14, maxit: int = 20000,
                    solver: str = "ConjugateGradient"):
    """ContField::HelmSolve of the session's Forcing with its Dirichlet
    conditions and parameter Lambda (the Helmholtz3D demo): returns
    (L2 error, Linf error, SolveInfo, element-local solution) against the
    session's ExactSolution at the quadrature points.  Hexahedral meshes
    of parallelepipeds, NUMMODES per the expansion, Q = NUMMODES + 1
    GLL points (the MODIFIED default)."""
    import torch

    from . import AssemblyMap, FP64, GetOperatorFactory, LinearSystem, basis_tables, SHAPE_HEX

    s = read_session(session_path)
    kinds = {k for k, _ in s.mesh.elements.values()}
    if kinds != {"H"}:
        raise NotImplementedError("helmholtz_solve takes hexahedral meshes")
    for bcs in s.conditions.values():
        for bc in bcs:
            if bc.var == var and bc.kind != "D":
                raise NotImplementedError(f"boundary condition type {bc.kind} (only Dirichlet is supported)")
    eids, lv, verts = nio.hex_elements(s.mesh)
    jac, df = nio.hex_geometry(verts)
    nm = s.mesh.expansions[0]["nummodes"]
    nq, E = nm + 1, len(eids)
    dsets = []
    for rid, bcs in s.conditions.items():
        if any(bc.var == var and bc.kind == "D" for bc in bcs):
            for cid in s.regions[rid]:
                dsets += [_face_vertices(s.mesh, f) for letter, ids in s.mesh.composites[cid] for f in ids]
    mp, sg, ng, nd = nio.hex_c0_map(lv, nm, dsets)
    t = basis_tables(SHAPE_HEX, 0, nm, nq)
    B = t["bdata"].reshape(nm, nq)
    tables = (B, t["z"], t["w"])
    dloc = hex_dirichlet_local(s, var, lv, verts, nm, nq, mp, sg, nd, tables)

    lam = s.params.get("Lambda", 0.0)
    F = GetOperatorFactory()
    bwd = F.CreateInstance("BwdTrans_Hex_Regular", nm, nq, E, FP64)
    ipr = F.CreateInstance("IProduct_Hex_Regular", nm, nq, E, FP64)
    ipr.SetJac(jac)
    helm = F.CreateInstance("Helmholtz_Hex_Regular", nm, nq, E, FP64)
    helm.SetJac(jac)
    helm.SetDF(df)
    helm.SetLambda(lam)
    ls = LinearSystem(AssemblyMap(mp, sg, ng, nd), [helm], solver=solver)

    X = hex_quadrature_points(verts, t["z"])
    names = dict(s.params, x=X[..., 0], y=X[..., 1], z=X[..., 2], t=0.0)
    f = np.broadcast_to(evaluate(s.functions["Forcing"][var], names), X.shape[:2]).astype(float).reshape(-1)
    exact = np.broadcast_to(evaluate(s.functions["ExactSolution"][var], names), X.shape[:2]).astype(float)

    dev = "cuda"
    fq = torch.from_numpy(np.ascontiguousarray(f)).to(dev)
    bl = torch.zeros(ipr.output_size(), dtype=torch.float64, device=dev)
    ipr(fq, bl)
    ls.set_rhs_local(-bl)  # ContField.cpp:812-817: -IProductWRTBase(f)
    ls.set_dirichlet_local(torch.from_numpy(np.ascontiguousarray(dloc)).to(dev))
    info = ls.solve(tol=tol, maxit=maxit)
    xl = torch.zeros(E * nm ** 3, dtype=torch.float64, device=dev)
    ls.solution_local(xl)
    uq = torch.zeros(bwd.output_size(), dtype=torch.float64, device=dev)
    bwd(xl, uq)
    torch.cuda.synchronize()
    u = uq.cpu().numpy().reshape(E, -1)
    l2, linf = errors(u, exact, jac, t["w"])
    return l2, linf, info, xl.cpu().numpy()
