"""Serving loop: many Helmholtz solves admitted as they arrive.

Each request's result must be the same alone and in the mix: iterations,
final residual, L2 error against the manufactured solution and the norm of
the solution vector are compared for exact equality with a plain solve of
the same request on a fresh solver (the solve's arithmetic does not depend
on what runs beside it; a reused pooled solver gets a new lambda)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def nmf():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import nektar_amd
    nektar_amd.lib()
    return nektar_amd


def mixed_requests():
    # (n, nummodes, lambda, k, mixed); repeated shapes exercise pooled reuse
    spec = [(3, 4, 1.0, 1, 0), (4, 5, 2.5, 2, 1), (3, 4, 7.0, 3, 0), (5, 4, 0.3, 1, 0),
            (4, 6, 1.7, 2, 0), (4, 5, 9.1, 1, 0), (2, 7, 4.0, 1, 1), (3, 4, 0.1, 2, 1),
            (6, 4, 3.3, 3, 0), (4, 6, 5.5, 1, 1)]
    return [dict(n=n, nummodes=nm, numpoints=nm + 1, lam=lam, k=k, mixed=mx, tol=1e-10, maxit=5000,
                 arrival_s=0.0) for (n, nm, lam, k, mx) in spec]


def alone(nmf, r):
    s = nmf.HelmholtzSolver(r["n"], r["nummodes"], r["numpoints"], r["lam"])
    s.set_forcing_manufactured(r["k"])
    info = s.solve(mixed=bool(r["mixed"]), tol=r["tol"], maxit=r["maxit"])
    x = s.solution()
    return info, s.l2_error_manufactured(r["k"]), float(np.sqrt(np.sum(x * x)))


def check_same(nmf, reqs, results):
    for r, got in zip(reqs, results):
        info, l2, xnorm = alone(nmf, r)
        assert got["converged"] == 1 and info.converged, r
        assert got["iterations"] == info.iterations, (r, got["iterations"], info.iterations)
        assert got["outer_iterations"] == info.outer_iterations, r
        assert got["final_error"] == info.final_error, r
        assert got["l2_error"] == l2, (r, got["l2_error"], l2)
        # the server's norm is a device reduction, the alone one a host sum
        assert abs(got["solution_norm"] - xnorm) <= 1e-12 * xnorm, r


def test_same_alone_and_in_the_mix(nmf):
    reqs = mixed_requests()
    srv = nmf.Server(hbm_budget_bytes=8 << 30, streams=4, max_active=8)
    results, stats = srv.run(reqs, validate=True)
    assert stats["peak_active"] >= 4  # solves really overlapped
    for r, got in zip(reqs, results):
        assert got["unknowns"] == (r["n"] * (r["nummodes"] - 1) - 1) ** 3
        assert got["finish_s"] >= got["admit_s"] >= got["arrival_s"]
        # discretisation error (up to ~0.03 for k=3 on a 3^3 P=3 mesh); the
        # exact solution has L2 norm 0.35
        assert 0.0 < got["l2_error"] < 0.1
    check_same(nmf, reqs, results)


def test_estimate_bounds_device_bytes(nmf):
    """Admission reserves the estimate while a solver is still being built,
    so the estimate must bound what the solver then holds."""
    for r in mixed_requests():
        for mixed in (False, True):
            s = nmf.HelmholtzSolver(r["n"], r["nummodes"], r["numpoints"], r["lam"])
            s.set_forcing_manufactured(1)
            s.solve(mixed=mixed, tol=1e-6)
            sizes = [ctypes.c_int64() for _ in range(4)]
            nmf.check(nmf.lib().nmf_solver_sizes(s._h, *[ctypes.byref(v) for v in sizes]))
            held = sizes[3].value  # after the solve (the mixed solve adds its FP32 buffers)
            est = nmf.HelmholtzSolver.estimate_bytes(r["n"], r["nummodes"], r["numpoints"], False, mixed)
            assert held <= est, (r, mixed, held, est)


def test_budget_admission_and_eviction(nmf):
    reqs = mixed_requests()
    est = [nmf.HelmholtzSolver.estimate_bytes(r["n"], r["nummodes"], r["numpoints"], False, bool(r["mixed"]))
           for r in reqs]
    budget = int(max(est) * 1.5)  # room for one large request (plus change) at a time
    srv = nmf.Server(hbm_budget_bytes=budget, streams=2, max_active=8)
    results, stats = srv.run(reqs, validate=True)
    assert stats["peak_bytes"] <= budget
    assert stats["evictions"] >= 1
    assert stats["peak_active"] < len(reqs)
    check_same(nmf, reqs, results)


def test_timed_arrivals(nmf):
    reqs = mixed_requests()[:6]
    for i, r in enumerate(reqs):
        r["arrival_s"] = 0.02 * i
    srv = nmf.Server(hbm_budget_bytes=8 << 30, streams=4, max_active=4)
    results, stats = srv.run(reqs, validate=False)
    for r, got in zip(reqs, results):
        assert got["converged"] == 1
        assert got["admit_s"] >= r["arrival_s"]
        assert got["latency_ms"] == pytest.approx(1e3 * (got["finish_s"] - r["arrival_s"]))
        assert got["queue_ms"] >= 0.0 and got["device_ms"] > 0.0
    assert stats["makespan_s"] >= reqs[-1]["arrival_s"]


def test_one_at_a_time_matches(nmf):
    reqs = mixed_requests()[:4]
    srv = nmf.Server(hbm_budget_bytes=8 << 30, streams=1, max_active=1)
    results, stats = srv.run(reqs, validate=True)
    assert stats["peak_active"] == 1
    # requests 0 and 2 have the same mesh: the second reuses the pooled
    # solver with a new lambda
    assert stats["solvers_reused"] >= 1 and results[2]["reused_solver"] == 1
    check_same(nmf, reqs, results)


def test_server_errors(nmf):
    with pytest.raises(nmf.NektarError):
        nmf.Server(0, 1, 1)
    srv = nmf.Server(1 << 20, 1, 1)
    big = dict(n=8, nummodes=6, numpoints=7, lam=1.0, k=1, mixed=0, tol=1e-10, maxit=100, arrival_s=0.0)
    with pytest.raises(nmf.NektarError):
        srv.run([big])
    srv2 = nmf.Server(8 << 30, 1, 1)
    bad = dict(big, nummodes=12, numpoints=13)
    with pytest.raises(nmf.NektarError):
        srv2.run([bad])
    results, stats = srv2.run([])
    assert results == [] and stats["solvers_created"] == 0


def test_non_blocking_solves_interleaved(nmf):
    """Two solves driven chunk by chunk from one loop give the same results
    as blocking solves; while a solve runs, the solver refuses the calls
    that would share its work buffers."""
    a = nmf.HelmholtzSolver(4, 5, 6, 1.3)
    b = nmf.HelmholtzSolver(3, 6, 7, 0.7, perturb=0.1, seed=2)
    a.set_forcing_manufactured(1)
    b.set_forcing_manufactured(2)
    ra = a.solve(False, 1e-10, 5000)
    xa = a.solution()
    rb = b.solve(True, 1e-10, 5000)
    xb = b.solution()
    a.begin(False, 1e-10, 5000)
    b.begin(True, 1e-10, 5000)
    with pytest.raises(nmf.NektarError):
        a.finish()  # not finished yet
    x = torch.zeros(a.nglobal, dtype=torch.float64, device="cuda")
    for call in (lambda: a.solve(False, 1e-10, 10), lambda: a.set_forcing_manufactured(1),
                 lambda: a.l2_error_manufactured(1), lambda: a.apply_global(x, x), lambda: a.autotune(),
                 lambda: a.begin(False, 1e-10, 10)):
        with pytest.raises(nmf.NektarError) as err:
            call()
        assert err.value.code == 3  # NMF_ERR_STATE
    done_a = done_b = False
    steps = 0
    while not (done_a and done_b):
        if not done_a:
            done_a = a.advance(False)
        if not done_b:
            done_b = b.advance(False)
        steps += 1
        assert steps < 10_000_000
    ia, ib = a.finish(), b.finish()
    assert ia.iterations == ra.iterations and ia.final_error == ra.final_error
    assert ib.iterations == rb.iterations and ib.final_error == rb.final_error
    assert ib.outer_iterations == rb.outer_iterations
    assert np.array_equal(a.solution(), xa) and np.array_equal(b.solution(), xb)
    # usable again afterwards
    assert a.solve(False, 1e-10, 5000).iterations == ra.iterations
