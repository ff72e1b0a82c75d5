"""Batches of operators of uneven kinds and sizes.

Each entry's output from a batch run (graph, grouped and per-item modes,
several replays) must be bit-identical to the same operator applied on its
own (same kernels, same arithmetic per element), and a few entries are
checked against the FP64 oracle (relative max error <= 1e-12)."""
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


def make_geometry(shape, nm, nq, nelmt, deformed, seed):
    if shape == O.HEX:
        n = int(np.ceil(nelmt ** (1 / 3)))
        _, jac, df = G.hex_geometry(n, nq, deformed, seed=seed)
        per = nq ** 3 if deformed else 1
        return jac[: nelmt * per].copy(), df[: nelmt * per * 9].copy()
    _, jac, df = G.tet_geometry(nelmt, nq, deformed, seed=seed)
    return jac, df


class Item:
    def __init__(self, nmf, opname, shape, nm, nq, nelmt, deformed, fp32, lam, seed):
        sname = "Hex" if shape == O.HEX else "Tet"
        self.shape, self.nm, self.nq, self.nelmt, self.deformed, self.lam = shape, nm, nq, nelmt, deformed, lam
        self.opname = opname
        self.op = nmf.GetOperatorFactory().CreateInstance(
            f"{opname}{nmf.GetOpstring(sname, deformed)}", nm, nq, nelmt, nmf.FP32 if fp32 else nmf.FP64)
        self.jac, self.df = make_geometry(shape, nm, nq, nelmt, deformed, seed)
        if opname in ("IProduct", "Helmholtz"):
            self.op.SetJac(self.jac)
        if opname in ("PhysDeriv", "Helmholtz"):
            self.op.SetDF(self.df)
        if opname == "Helmholtz":
            self.op.SetLambda(lam)
        dt = torch.float32 if fp32 else torch.float64
        self.x_host = G.coeffs(self.op.input_size(), seed)
        self.x = torch.from_numpy(self.x_host).to(dt).cuda()
        nout = 3 if opname == "PhysDeriv" else 1
        self.outs = [torch.full((self.op.output_size(),), float("nan"), dtype=dt, device="cuda")
                     for _ in range(nout)]

    def alone(self):
        outs = [torch.zeros_like(o) for o in self.outs]
        self.op(self.x, *outs)
        torch.cuda.synchronize()
        return [o.cpu().numpy() for o in outs]

    def got(self):
        return [o.cpu().numpy() for o in self.outs]

    def reset(self):
        for o in self.outs:
            o.fill_(float("nan"))


def uneven_items(nmf, fp32, count=14, seed=4):
    rng = np.random.default_rng(seed)
    items = []
    for i in range(count):
        shape = O.HEX if i % 2 == 0 else O.TET
        nm = int(rng.integers(4, 11))  # P in 3..9
        nq = nm + 1
        nelmt = int(np.exp(rng.uniform(np.log(3), np.log(300))))
        deformed = bool(i % 5 == 3)
        items.append(Item(nmf, "Helmholtz", shape, nm, nq, nelmt, deformed, fp32,
                          float(rng.uniform(0.1, 10)), seed * 100 + i))
    # a repeated kind so that grouping merges entries
    items.append(Item(nmf, "Helmholtz", O.HEX, 5, 6, 17, False, fp32, 2.0, 999))
    items.append(Item(nmf, "Helmholtz", O.HEX, 5, 6, 40, False, fp32, 3.0, 998))
    items.append(Item(nmf, "Helmholtz", O.TET, 5, 6, 33, False, fp32, 0.5, 997))
    items.append(Item(nmf, "Helmholtz", O.TET, 5, 6, 1, False, fp32, 0.7, 996))
    return items


@pytest.mark.parametrize("fp32", [False, True], ids=["fp64", "fp32"])
@pytest.mark.parametrize("mode", [0, 1, 2], ids=["graph", "grouped", "per_item"])
def test_batch_matches_individual(nmf, fp32, mode):
    items = uneven_items(nmf, fp32)
    ref = [it.alone() for it in items]
    b = nmf.Batch(mode=mode, streams=3)
    for it in items:
        b.add(it.op, it.x, it.outs[0])
    info = b.info()
    assert info["items"] == len(items)
    assert info["groups"] < len(items)  # repeated kinds were merged
    assert info["dofs"] == sum(it.op.input_size() for it in items)
    for rep in range(3):
        for it in items:
            it.reset()
        b.run()
        torch.cuda.synchronize()
        for k, it in enumerate(items):
            assert np.array_equal(it.got()[0], ref[k][0]), (rep, k, it.shape, it.nm, it.nelmt)


def test_batch_fp64_against_oracle(nmf):
    items = uneven_items(nmf, False, count=6, seed=9)
    b = nmf.Batch()
    for it in items:
        b.add(it.op, it.x, it.outs[0])
    b.run()
    torch.cuda.synchronize()
    for it in items:
        ref = O.helmholtz(it.shape, it.nm, it.nq, it.nelmt, it.deformed, it.jac, it.df, it.lam, it.x_host)
        assert G.rel_err(it.got()[0], ref) <= 1e-12


def test_batch_picks_up_changes(nmf):
    items = uneven_items(nmf, False, count=4, seed=5)
    b = nmf.Batch()
    for it in items:
        b.add(it.op, it.x, it.outs[0])
    b.run()
    items[1].op.SetLambda(7.25)
    items[2].op.SetJac(items[2].jac * 1.5)
    b.run()
    torch.cuda.synchronize()
    for it in items:
        assert np.array_equal(it.got()[0], it.alone()[0])


def test_batch_mixed_operator_types(nmf):
    items = [Item(nmf, "BwdTrans", O.HEX, 4, 5, 20, False, False, 1.0, 1),
             Item(nmf, "IProduct", O.TET, 5, 6, 11, True, False, 1.0, 2),
             Item(nmf, "PhysDeriv", O.HEX, 6, 7, 9, True, False, 1.0, 3),
             Item(nmf, "Helmholtz", O.TET, 3, 4, 30, False, False, 2.0, 4),
             Item(nmf, "Helmholtz", O.HEX, 3, 4, 30, False, False, 2.0, 5)]
    ref = [it.alone() for it in items]
    b = nmf.Batch(streams=2)
    for it in items:
        if it.opname == "PhysDeriv":
            b.add_physderiv(it.op, it.x, *it.outs)
        else:
            b.add(it.op, it.x, it.outs[0])
    for _ in range(2):
        for it in items:
            it.reset()
        b.run()
        torch.cuda.synchronize()
        for k, it in enumerate(items):
            for got, want in zip(it.got(), ref[k]):
                assert np.array_equal(got, want)


def test_batch_autotune_keeps_results(nmf):
    nmf.clear_tuning()
    items = uneven_items(nmf, False, count=6, seed=6)
    ref = [it.alone() for it in items]
    b = nmf.Batch()
    for it in items:
        b.add(it.op, it.x, it.outs[0])
    b.autotune()
    b.run()
    torch.cuda.synchronize()
    for k, it in enumerate(items):
        if it.op.tile < nmf.TILE_MFMA:
            # a vector-ALU tile computes the same arithmetic as the default
            assert np.array_equal(it.got()[0], ref[k][0])
        else:
            # a matrix-core tile rounds differently: equal to its own solo
            # apply bit for bit, to the default tile's within rounding
            assert np.array_equal(it.got()[0], it.alone()[0])
            assert np.max(np.abs(it.got()[0] - ref[k][0])) <= 1e-12 * np.max(np.abs(ref[k][0]))
        if it.shape == O.HEX:
            assert it.op.tile in it.op.tile_candidates()
    nmf.clear_tuning()


def test_batch_edge_cases(nmf):
    b = nmf.Batch()
    b.run()  # empty batch is a no-op
    torch.cuda.synchronize()
    assert b.info() == {"items": 0, "groups": 0, "dofs": 0}
    with pytest.raises(nmf.NektarError):
        b.set_mode(7)
    with pytest.raises(nmf.NektarError):
        b.set_streams(0)
    it = Item(nmf, "PhysDeriv", O.HEX, 4, 5, 8, False, False, 1.0, 1)
    with pytest.raises(nmf.NektarError):
        b.add(it.op, it.x, it.outs[0])
