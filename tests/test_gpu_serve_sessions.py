"""Serving HelmSolve requests given as session files
(nektar_amd.contfield.serve_sessions): the reference's own sessions
(Helmholtz3D_CG_Hex, _Tet, _Hex_AllBCs) and a trilinear-hex session,
arriving together and staggered, solved concurrently on streams of their
own.  Every result must equal the same request solved alone (the solve's
arithmetic does not depend on what runs beside it) and the reference's
recorded L2/Linf errors; a budget that fits one request at a time must
serialise the solves."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))
G = os.path.join(HERE, "golden", "helmholtz3d")
HEX, TET = os.path.join(G, "Helmholtz3D_Hex_P6.xml"), os.path.join(G, "Helmholtz3D_Tet.xml")
ALLBCS = os.path.join(G, "Helmholtz3D_Hex_AllBCs_P6.xml")
# the reference's printed errors and the test files' tolerances
# (Helmholtz3D_CG_Hex.tst, _Tet.tst, _Hex_AllBCs.tst; the last prints six digits)
REF = {HEX: (1.11818e-07, 5.66374e-07, 1e-12, 1e-12), TET: (8.458e-05, 0.000862348, 1e-8, 1e-9),
       ALLBCS: (0.000416575, 0.000871589, 5e-10, 5e-10)}


@pytest.fixture(scope="module")
def nmf():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import nektar_amd
    nektar_amd.lib()
    return nektar_amd


def _requests():
    import test_contfield_cpu as T

    tri = T.perturbed_hex_session()
    spec = [(HEX, "ConjugateGradient", 0.0), (TET, "ConjugateGradient", 0.0), (ALLBCS, "ConjugateGradient", 0.0),
            (HEX, "ConjugateGradientLoc", 0.01), (tri, "ConjugateGradient", 0.02), (TET, "IterativeStaticCond", 0.03)]
    return [dict(session=s, solver=sv, tol=1e-14, arrival_s=a) for s, sv, a in spec]


def test_sessions_served_concurrently_equal_alone(nmf):
    from nektar_amd import contfield as cf

    reqs = _requests()
    results, stats = cf.serve_sessions(reqs, hbm_budget_bytes=4 << 30, max_active=4)
    print(f"served {len(reqs)} sessions: makespan {stats['makespan_s']:.3f} s, peak active {stats['peak_active']}, "
          f"p50 {stats['p50_ms']:.1f} ms, p99 {stats['p99_ms']:.1f} ms", flush=True)
    assert stats["peak_active"] >= 2
    for r, got in zip(reqs, results):
        assert got["converged"], r
        l2, linf, info, _ = cf.helmholtz_solve(r["session"], tol=r["tol"], solver=r["solver"])
        assert (got["l2"], got["linf"], got["iterations"]) == (l2, linf, info.iterations), (r["solver"], got)
        if isinstance(r["session"], str):
            l2r, linfr, t2, tinf = REF[r["session"]]
            assert abs(got["l2"] - l2r) <= t2 and abs(got["linf"] - linfr) <= tinf, (r["session"], got)
        assert got["latency_ms"] >= got["queue_ms"] >= 0.0


def test_budget_serialises(nmf):
    from nektar_amd import contfield as cf

    reqs = [dict(session=HEX, solver="ConjugateGradient", tol=1e-14, arrival_s=0.0),
            dict(session=HEX, solver="ConjugateGradientLoc", tol=1e-14, arrival_s=0.0)]
    nb = cf.estimate_bytes(cf.dirichlet_problem(cf.read_session(HEX)))
    results, stats = cf.serve_sessions(reqs, hbm_budget_bytes=nb + nb // 2, max_active=4)
    assert stats["peak_active"] == 1 and stats["peak_bytes"] == nb
    assert results[1]["admit_s"] >= results[0]["finish_s"] or results[0]["admit_s"] >= results[1]["finish_s"]
    assert all(r["converged"] for r in results)
    with pytest.raises(ValueError):
        cf.serve_sessions(reqs[:1], hbm_budget_bytes=nb // 2)
