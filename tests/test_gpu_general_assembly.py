"""GPU general C0 assembly and iterative solve (AssemblyMapCG +
GlobalLinSysIterativeFull + NekLinSysIterCG with the Diagonal
preconditioner on a caller-supplied map with signs).

The mesh is the structured hex mesh of the structured solver with every
element in a random local frame (tests/geometry_cases.rotated_hex_mesh,
pinned on the CPU by the oracle), its map renumbered with the Dirichlet DOFs
first and split into two operator collections.  The assembled operator, the
inverse diagonal and the CG solution must equal the structured solver's on
the same mesh (an independent code path pinned against the oracle):
operator and diagonal to 1e-12, solution to 1e-9, iteration counts within
max(2, 1 %); gather/assemble are each other's transpose; an assembly is
deterministic."""
import numpy as np
import pytest

import _oracle as O
import geometry_cases as G

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def nmf():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import nektar_amd
    nektar_amd.lib()
    return nektar_amd


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _setup(nmf, n, nm, nq, lam, amp, seed, split):
    deformed = amp > 0
    Vr, mr, sr, dirichlet = G.rotated_hex_mesh(n, nm, amp, seed)
    new, ndir = G.dirichlet_first(dirichlet)
    nel, nc = n ** 3, nm ** 3
    jac, df = O.hex_geom(Vr, nel, nq, deformed)
    gp = O.nqtot(O.HEX, nq) if deformed else 1
    key = "Helmholtz" + nmf.GetOpstring("Hex", deformed)
    ops = []
    for a, b in ((0, split), (split, nel)):
        op = nmf.GetOperatorFactory().CreateInstance(key, nm, nq, b - a, nmf.FP64)
        op.SetJac(jac[a * gp:b * gp])
        op.SetDF(df[a * gp * 9:b * gp * 9])
        op.SetLambda(lam)
        ops.append(op)
    amap = nmf.AssemblyMap(new[mr], sr, len(dirichlet), ndir)
    ls = nmf.LinearSystem(amap, ops)
    return dict(Vr=Vr, mr=mr, sr=sr, dirichlet=dirichlet, new=new, ndir=ndir, jac=jac, ops=ops, amap=amap, ls=ls,
                deformed=deformed, nel=nel, nc=nc)


CASES = [(3, 4, 5, 1.0, 0.0, 2, 10), (4, 5, 6, 2.5, 0.15, 3, 29), (3, 7, 8, 0.5, 0.1, 5, 13)]


@pytest.mark.parametrize("n,nm,nq,lam,amp,seed,split", CASES)
def test_matches_structured_solver(nmf, n, nm, nq, lam, amp, seed, split):
    S = _setup(nmf, n, nm, nq, lam, amp, seed, split)
    s = nmf.HelmholtzSolver(n, nm, nq, lam, amp, seed)
    new, ng = S["new"], len(S["dirichlet"])
    assert s.nglobal == ng and s.ndirichlet == S["ndir"]
    # inverse diagonal (Dirichlet rows 0)
    d_new = S["ls"].inv_diag()
    assert _rel(d_new[new], s.inv_diag()) <= 1e-12
    # assembled operator
    x = np.random.default_rng(seed).uniform(-1, 1, ng)
    x[S["dirichlet"]] = 0.0
    xs = torch.from_numpy(x).cuda()
    ys = torch.zeros(ng, dtype=torch.float64, device="cuda")
    s.apply_global(xs, ys)
    xn = np.empty(ng)
    xn[new] = x
    yn = torch.zeros(ng, dtype=torch.float64, device="cuda")
    S["ls"].apply_global(torch.from_numpy(xn).cuda(), yn)
    torch.cuda.synchronize()
    assert _rel(yn.cpu().numpy()[new], ys.cpu().numpy()) <= 1e-12
    # manufactured solve: b = Assemble(IProductWRTBase(f)) in the rotated frames
    xyz = O.hex_coords(S["Vr"], S["nel"], nq)
    f = (3 * np.pi ** 2 + lam) * np.prod(np.sin(np.pi * xyz), axis=2).ravel()
    bloc = O.iproduct(O.HEX, nm, nq, S["nel"], S["deformed"], S["jac"], f)
    S["ls"].set_rhs_local(torch.from_numpy(bloc).cuda())
    info = S["ls"].solve(tol=1e-12, maxit=2000)
    s.set_forcing_manufactured(1)
    ref = s.solve(tol=1e-12, maxit=2000)
    assert info.converged and ref.converged
    # the element operators sum in different orders in the rotated frames:
    # long runs drift by a few iterations
    assert abs(info.iterations - ref.iterations) <= max(2, ref.iterations // 100), (info.iterations, ref.iterations)
    sol = S["ls"].solution()
    assert _rel(sol[new], s.solution()) <= 1e-9
    assert np.all(sol[:S["ndir"]] == 0.0)


def test_gather_is_the_transpose_of_assemble(nmf):
    n, nm = 3, 5
    Vr, mr, sr, dirichlet = G.rotated_hex_mesh(n, nm, 0.0, 8)
    new, ndir = G.dirichlet_first(dirichlet)
    ng, nloc = len(dirichlet), len(mr)
    amap = nmf.AssemblyMap(new[mr], sr, ng, ndir)
    rng = np.random.default_rng(2)
    g = torch.from_numpy(rng.uniform(-1, 1, ng)).cuda()
    l = torch.from_numpy(rng.uniform(-1, 1, nloc)).cuda()
    gl = torch.zeros(nloc, dtype=torch.float64, device="cuda")
    lg = torch.zeros(ng, dtype=torch.float64, device="cuda")
    amap.global_to_local(g, gl)
    amap.assemble(l, lg)
    torch.cuda.synchronize()
    # <G g, l> = <g, A l>
    a = float(torch.dot(gl, l))
    b = float(torch.dot(g, lg))
    assert abs(a - b) <= 1e-12 * max(abs(a), 1.0)
    # against numpy
    gn = g.cpu().numpy()
    assert np.array_equal(gl.cpu().numpy(), sr * gn[new[mr]])
    ref = np.zeros(ng)
    np.add.at(ref, new[mr], sr * l.cpu().numpy())
    assert _rel(lg.cpu().numpy(), ref) <= 1e-14
    # deterministic: the same bits on a second assembly
    lg2 = torch.zeros(ng, dtype=torch.float64, device="cuda")
    amap.assemble(l, lg2)
    torch.cuda.synchronize()
    assert torch.equal(lg, lg2)


def test_bad_maps_are_rejected(nmf):
    with pytest.raises(nmf.NektarError):
        nmf.AssemblyMap(np.array([0, 5]), None, 3, 0)        # outside [0, nglobal)
    with pytest.raises(nmf.NektarError):
        nmf.AssemblyMap(np.array([0, 1]), np.array([1.0, 0.5]), 3, 0)  # sign not +-1
    with pytest.raises(nmf.NektarError):
        nmf.AssemblyMap(np.array([0, 1]), None, 3, 4)        # ndir > nglobal


# ---- unstructured tet meshes (tests/tet_mesh.py, pinned on the CPU) -------
import brute as B  # noqa: E402
import tet_mesh as TM  # noqa: E402
from test_oracle_tet_assembly import _dense_system  # noqa: E402


def _tet_setup(nmf, n, nm, lam, split):
    coords, tets, mp, sg, dirichlet, jac, A = _dense_system(n, nm, lam)
    nel, nc = len(tets), len(B.tet_modes(nm))
    _, df = O.tet_geom(TM.tet_vertices(coords, tets), nel)
    ops = []
    for a, b in ((0, split), (split, nel)):
        op = nmf.GetOperatorFactory().CreateInstance("Helmholtz_Tet_Regular", nm, nm + 1, b - a, nmf.FP64)
        op.SetJac(jac[a:b])
        op.SetDF(df[9 * a:9 * b])
        op.SetLambda(lam)
        ops.append(op)
    return dict(coords=coords, tets=tets, mp=mp, sg=sg, dirichlet=dirichlet, jac=jac, A=A, ops=ops, nel=nel, nc=nc)


@pytest.mark.parametrize("n,nm,lam,split", [(2, 3, 1.0, 17), (2, 5, 0.5, 30), (3, 4, 2.0, 100)])
def test_tet_mesh_against_dense_oracle(nmf, n, nm, lam, split):
    S = _tet_setup(nmf, n, nm, lam, split)
    new, ndir = G.dirichlet_first(S["dirichlet"])
    ng = len(S["dirichlet"])
    amap = nmf.AssemblyMap(new[S["mp"]], S["sg"], ng, ndir)
    ls = nmf.LinearSystem(amap, S["ops"])
    free = ~S["dirichlet"]
    A = S["A"]
    # inverse diagonal on the free rows
    d = ls.inv_diag()[new]
    assert _rel(d[free], 1.0 / np.diag(A)[free]) <= 1e-12
    assert np.all(d[S["dirichlet"]] == 0.0)
    # assembled operator
    x = np.random.default_rng(n + nm).uniform(-1, 1, ng)
    x[S["dirichlet"]] = 0.0
    xn = np.empty(ng)
    xn[new] = x
    y = torch.zeros(ng, dtype=torch.float64, device="cuda")
    ls.apply_global(torch.from_numpy(xn).cuda(), y)
    torch.cuda.synchronize()
    ref = A @ x
    ref[S["dirichlet"]] = 0.0
    assert _rel(y.cpu().numpy()[new], ref) <= 1e-12
    # manufactured solve against the dense solve of the same system
    X = TM.tet_quadrature_coords(S["coords"], S["tets"], nm + 1)
    f = (3 * np.pi ** 2 + lam) * np.prod(np.sin(np.pi * X), axis=2).ravel()
    bl = O.iproduct(O.TET, nm, nm + 1, S["nel"], False, S["jac"], f)
    ls.set_rhs_local(torch.from_numpy(bl).cuda())
    info = ls.solve(tol=1e-12, maxit=5000)
    assert info.converged
    b = np.zeros(ng)
    np.add.at(b, S["mp"], S["sg"] * bl)
    xd = np.zeros(ng)
    xd[free] = np.linalg.solve(A[np.ix_(free, free)], b[free])
    assert _rel(ls.solution()[new], xd) <= 1e-9


def test_hex_and_tet_collections_in_one_system(nmf):
    # a hex region and a tet region in one system (not connected: a
    # conforming hex-tet interface needs prisms or pyramids); each part must
    # match its own independent solution
    lam = 1.0
    H = _setup(nmf, 2, 4, 5, lam, 0.0, 3, 4)
    Tt = _tet_setup(nmf, 2, 3, lam, 20)
    ngh, ngt = len(H["dirichlet"]), len(Tt["dirichlet"])
    mp = np.concatenate([H["mr"], Tt["mp"] + ngh])
    sg = np.concatenate([H["sr"], Tt["sg"]])
    dirichlet = np.concatenate([H["dirichlet"], Tt["dirichlet"]])
    new, ndir = G.dirichlet_first(dirichlet)
    amap = nmf.AssemblyMap(new[mp], sg, ngh + ngt, ndir)
    ops = H["ops"] + Tt["ops"]  # local vectors concatenated in this order
    ls = nmf.LinearSystem(amap, ops)
    xyz = O.hex_coords(H["Vr"], H["nel"], 5)
    fh = (3 * np.pi ** 2 + lam) * np.prod(np.sin(np.pi * xyz), axis=2).ravel()
    bh = O.iproduct(O.HEX, 4, 5, H["nel"], False, H["jac"], fh)
    X = TM.tet_quadrature_coords(Tt["coords"], Tt["tets"], 4)
    ft = (3 * np.pi ** 2 + lam) * np.prod(np.sin(np.pi * X), axis=2).ravel()
    bt = O.iproduct(O.TET, 3, 4, Tt["nel"], False, Tt["jac"], ft)
    ls.set_rhs_local(torch.from_numpy(np.concatenate([bh, bt])).cuda())
    info = ls.solve(tol=1e-12, maxit=5000)
    assert info.converged
    sol = ls.solution()[new]
    s = nmf.HelmholtzSolver(2, 4, 5, lam, 0.0, 3)
    s.set_forcing_manufactured(1)
    s.solve(tol=1e-12, maxit=2000)
    assert _rel(sol[:ngh], s.solution()) <= 1e-9
    free = ~Tt["dirichlet"]
    b = np.zeros(ngt)
    np.add.at(b, Tt["mp"], Tt["sg"] * bt)
    xd = np.zeros(ngt)
    xd[free] = np.linalg.solve(Tt["A"][np.ix_(free, free)], b[free])
    assert _rel(sol[ngh:], xd) <= 1e-9
