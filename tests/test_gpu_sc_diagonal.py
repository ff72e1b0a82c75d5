"""GPU static-condensation diagonal (nmf_linsys_sc_diagonal,
PreconditionerDiagonal.cpp:132-154 / Preconditioner.cpp:185-236) against the
oracle's (or_sc_diagonal, pinned on the CPU by dense Schur complements, block
elimination and the condensed solve) on signed maps: unstructured tet meshes,
hex meshes in random element frames (regular and deformed), hex and tet
collections in one system, interior blocks from 0 to 512 modes (the
Cholesky panels full and partial), element chunks smaller than a
collection; FP64, relative 1e-12.  Maps the condensation cannot take are
rejected."""
import numpy as np
import pytest

import _oracle as O
import geometry_cases as G
from test_gpu_general_assembly import _setup, _tet_setup
from test_oracle_cg_loc import _blocks, _hex_case, _tet_case

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


def _case(nmf, kind, args):
    if kind == "tet":
        n, nm, lam = args
        blocks, mp, sg, dirichlet, _ = _tet_case(n, nm, lam)
        ops = _tet_setup(nmf, n, nm, lam, len(blocks) // 3)["ops"]
        shape = O.TET
    else:
        n, nm, nq, lam, amp, seed = args
        blocks, mp, sg, dirichlet, _ = _hex_case(n, nm, nq, lam, amp, seed)
        ops = _setup(nmf, n, nm, nq, lam, amp, seed, max(1, n ** 3 // 3))["ops"]
        shape = O.HEX
    new, ndir = G.dirichlet_first(dirichlet)
    fl = np.tile(O.interior_modes(shape, nm), len(blocks))
    return blocks, new[mp], sg, len(dirichlet), ndir, fl, ops


CASES = [("tet", (2, 4, 1.0)), ("tet", (2, 7, 0.5)), ("tet", (2, 8, 2.0)), ("hex", (2, 3, 4, 1.0, 0.0, 3)),
         ("hex", (3, 5, 6, 2.5, 0.0, 7)), ("hex", (2, 7, 8, 1.0, 0.12, 9)), ("hex", (2, 8, 9, 0.7, 0.0, 5)),
         ("hex", (2, 10, 11, 1.0, 0.0, 6))]


@pytest.mark.parametrize("kind,args", CASES)
def test_sc_diagonal_against_oracle(nmf, kind, args):
    blocks, m, sg, ng, ndir, fl, ops = _case(nmf, kind, args)
    ref = O.sc_diagonal(blocks, fl, m, sg, ng, ndir)
    amap = nmf.AssemblyMap(m, sg, ng, ndir)
    ls = nmf.LinearSystem(amap, ops)
    d = ls.sc_diagonal()
    assert _rel(d, ref) <= 1e-12
    assert np.all(d[ref == 0.0] == 0.0)
    # the Jacobi diagonal bounds it from above on the boundary rows
    bnd = ref != 0.0
    assert np.all(d[bnd] <= (1.0 / ls.inv_diag()[bnd]) * (1 + 1e-12))


def test_chunks_and_mixed_collections(nmf, monkeypatch):
    blocks_h, mh, sh, ngh, _, flh, ops_h = _case(nmf, "hex", (2, 6, 7, 1.0, 0.0, 3))
    blocks_t, mt, st, ngt, _, flt, ops_t = _case(nmf, "tet", (2, 5, 1.0))
    # one system of both (not connected), Dirichlet rows first again
    _, _, _, dir_h = G.rotated_hex_mesh(2, 6, 0.0, 3)
    import tet_mesh as TM
    dir_t = TM.tet_c0_map(2, 5)[4]
    # undo the per-part renumbering: rebuild from the parts' own maps
    nh, _ = G.dirichlet_first(dir_h)
    nt, _ = G.dirichlet_first(dir_t)
    inv_h, inv_t = np.argsort(nh), np.argsort(nt)
    mp = np.concatenate([inv_h[mh], inv_t[mt] + ngh])
    dirichlet = np.concatenate([dir_h, dir_t])
    new, ndir = G.dirichlet_first(dirichlet)
    m = new[mp]
    sg = np.concatenate([sh, st])
    fl = np.concatenate([flh, flt])
    ref = O.sc_diagonal(blocks_h + blocks_t, fl, m, sg, ngh + ngt, ndir)
    amap = nmf.AssemblyMap(m, sg, ngh + ngt, ndir)
    ls = nmf.LinearSystem(amap, ops_h + ops_t)
    d_one = ls.sc_diagonal()
    assert _rel(d_one, ref) <= 1e-12
    # element chunks: an element's blocks hold ni^2 + kb ni + kb doubles,
    # kb = min(16, 8192 // ni) columns per pass (hex nm=6: ni = 64; tet
    # nm=5: ni = 1); chunks of 3 hexes, then of 3 tets (and single hexes):
    # the same values bit for bit
    for per in (64 * 64 + 64 * 16 + 16, 1 + 16 + 16):
        monkeypatch.setenv("NMF_SC_CHUNK_BYTES", str(3 * per * 8))
        assert np.array_equal(ls.sc_diagonal(), d_one)


def test_rejected_maps(nmf):
    blocks, m, sg, ng, ndir, fl, ops = _case(nmf, "hex", (2, 4, 5, 1.0, 0.0, 2))
    # an interior mode sharing the DOF of another local entry
    m2 = m.copy()
    i = int(np.flatnonzero(fl)[0])
    m2[i] = m2[i + 1]
    ls = nmf.LinearSystem(nmf.AssemblyMap(m2, sg, ng, ndir), ops)
    with pytest.raises(nmf.NektarError):
        ls.sc_diagonal()
    # two boundary modes of one element on one global DOF
    m3 = m.copy()
    b = np.flatnonzero(~fl[:64])
    m3[b[1]] = m3[b[0]]
    ls = nmf.LinearSystem(nmf.AssemblyMap(m3, sg, ng, ndir), ops)
    with pytest.raises(nmf.NektarError):
        ls.sc_diagonal()
