#!/usr/bin/env python3
"""Benchmark of the matrix-free Helmholtz path on one MI355X.

Workload (one "step"): one FP64 Helmholtz operator apply y = (lambda M + L) x
over the 32x32x32 regular hex mesh of [0,1]^3, P=6 (7 modes), GLL Q=8,
lambda=1 (11,239,424 element-local DOFs), inputs resident in HBM.
value = element-local GDOF/s of that apply.

Also reported (same mesh): the FP32 apply used inside the mixed-precision
solve, and the C0 Helmholtz solve -lap(u)+u=f with u=sin(pi x)sin(pi y)sin(pi z)
to relative residual 1e-10 in FP64 CG and in mixed-precision
iterative-refinement CG (iterations, solve time, solve GDOF/s, L2 error).

roofline: achieved TFLOP/s of the dominant kernel (the fused Helmholtz
apply) = algorithmic flops per launch / average launch time from HIP events
on the launch stream; peak = MI355X FP64 78.6 TFLOP/s (vector FMAs; the
FP64 matrix peak is the same number).
cpu_baseline: the oracle (CPU FP64 restatement of the reference kernels,
vectorised across elements, OpenMP over element blocks) on a bounded sample:
medians on all cores and on one core, CPU model and SIMD width.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Helmholtz apply GDOF/s (FP64 vs mixed); served solves/s + p99 at 1e-10 res"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def hex_flops_per_elem(nm, nq):
    """Algorithmic flops of one sum-factorised hex Helmholtz apply: forward
    transform T, 3-direction collocation gradient G, pointwise metric+mass,
    transposes."""
    T = 2 * (nm ** 3 * nq + nm ** 2 * nq ** 2 + nm * nq ** 3)
    G = 6 * nq ** 4
    return 2 * T + 2 * G + 18 * nq ** 3


def hex_bytes_per_elem(nm, nq, deformed, sv):
    g = 7 * (nq ** 3 if deformed else 1)
    return 2 * nm ** 3 * sv + g * sv


def read_traffic(kernel_tag):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", "helmholtz_hex_p6_q8_32cubed_pmc.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(kernel_tag + " traffic_bytes_per_launch"):
            try:
                return float(line.split("=")[1])
            except (IndexError, ValueError):
                return None
    return None


def time_apply(op, x, y, steps, warmup, torch):
    """Average time of one apply over `steps` back-to-back launches on
    torch's current stream (events on that stream); the timed loop calls
    the C entry point with pre-resolved arguments so that host overhead
    does not open gaps between launches."""
    import nektar_amd as N

    fn = N.lib().nmf_operator_apply
    stream = torch.cuda.current_stream().cuda_stream
    args = (op._h, x.data_ptr(), y.data_ptr(), stream)
    for _ in range(warmup):
        N.check(fn(*args))
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    s.record()
    bad = 0
    for _ in range(steps):
        bad |= fn(*args)
    e.record()
    N.check(bad)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return s.elapsed_time(e) / steps, wall


def tet_flops_per_elem(nm, nq0):
    """Algorithmic flops of the fused tet Helmholtz apply (reference loop
    counts, BwdTransKernels.hpp:242-328, IProductKernels.hpp:369-454):
    F = 2 T + G + 30 nqt, T = 2(nq2 nc + nq1 nq2 nmb + nqt nm),
    G = 4 nqt (nq0 + nq1 + nq2)."""
    nq1 = nq2 = nq0 - 1
    nc = nm * (nm + 1) * (nm + 2) // 6
    nqt = nq0 * nq1 * nq2
    T = 2 * (nq2 * nc + nq1 * nq2 * nm * (nm + 1) // 2 + nqt * nm)
    G = 4 * nqt * (nq0 + nq1 + nq2)
    return 2 * T + G + 30 * nqt


def bench_ops(N, torch, steps=20, warmup=3):
    """Every operator of the Collections interface on 32,768 affine
    elements, P=6 (7 modes), Q=8, hex and tet, FP64 and FP32: average
    launch time over back-to-back launches, algorithmic HBM bytes (inputs,
    outputs, geometry, velocities) and, for the fused Helmholtz-type
    operators, algorithmic flops."""
    import numpy as np

    E, nm, nq = 32768, 7, 8
    rng = np.random.default_rng(6)
    s = rng.uniform(0.5, 2.0, E)[:, None, None]
    M = s * (np.eye(3)[None] + 0.25 * rng.uniform(-1, 1, (E, 3, 3))) * 0.05
    jac = np.abs(np.linalg.det(M))
    df = np.ascontiguousarray(np.transpose(np.linalg.inv(M), (0, 2, 1))).reshape(-1)
    out = {}
    for shape in ("Hex", "Tet"):
        nc = nm ** 3 if shape == "Hex" else nm * (nm + 1) * (nm + 2) // 6
        nqt = nq ** 3 if shape == "Hex" else nq * (nq - 1) ** 2
        fh = hex_flops_per_elem(nm, nq) if shape == "Hex" else tet_flops_per_elem(nm, nq)
        g = 7 if shape == "Hex" else 10
        vel = rng.uniform(-1, 1, E * nqt * 3)
        for prec, tag in ((N.FP64, "fp64"), (N.FP32, "fp32")):
            sv = 8 if prec == N.FP64 else 4
            for key in ("Helmholtz", "LinearAdvectionDiffusionReaction", "PhysInterp1DScaled", "BwdTrans",
                        "IProduct"):
                op = N.GetOperatorFactory().CreateInstance(f"{key}_{shape}_Regular", nm, nq, E, prec)
                if key not in ("BwdTrans", "PhysInterp1DScaled"):
                    op.SetJac(jac)
                if key in ("Helmholtz", "LinearAdvectionDiffusionReaction"):
                    op.SetDF(df)
                    op.SetLambda(1.0)
                if key == "LinearAdvectionDiffusionReaction":
                    op.SetAdvectionVelocities(vel)
                tile = op.autotune()[0] if op.tile_candidates() else 0
                dt = torch.float64 if prec == N.FP64 else torch.float32
                x = torch.rand(op.input_size(), dtype=dt, device="cuda") * 2 - 1
                y = torch.zeros(op.output_size(), dtype=dt, device="cuda")
                ms, _ = time_apply(op, x, y, steps, warmup, torch)
                us = ms * 1e3
                if key == "Helmholtz":
                    byts, flops = E * (2 * nc + g) * sv, E * fh
                elif key == "LinearAdvectionDiffusionReaction":
                    byts, flops = E * (2 * nc + g + 3 * nqt) * sv, E * (fh + 6 * nqt)
                elif key == "PhysInterp1DScaled":
                    byts, flops = (op.input_size() + op.output_size()) * sv, None
                elif key == "BwdTrans":
                    byts, flops = E * (nc + nqt) * sv, None
                else:
                    byts, flops = E * (nqt + nc + 1) * sv, None
                r = {"us": round(us, 2), "alg_gbs": round(byts / us / 1e3, 1),
                     "hbm_frac": round(byts / us / 1e3 / 8000.0, 3), "tile": tile}
                if flops:
                    r["tflops"] = round(flops / us / 1e6, 2)
                    r["fp_frac"] = round(flops / us / 1e6 / (PEAK_FP64_TFLOPS if sv == 8 else 157.3), 3)
                out[f"{key}_{shape}_{tag}"] = r
                del op, x, y
    return out


def bench_mfma(N, torch, steps=10, warmup=2):
    """Vector-ALU against matrix-core (MFMA) hex Helmholtz kernels at high
    order: P = 7, 8, 9 (NUMMODES 8, 9, 10), Q = P + 2, 13,824 elements
    (24^3), FP64 and FP32, regular geometry, and the deformed (per-point
    geometry) kernel at P = 7.  Every tile is timed by the autotuner; the
    fastest tile of each family is reported with its algorithmic TFLOP/s
    and fraction of the precision's peak, and for the deformed kernel (the
    apply of the curved P = 7 mesh, HBM-bound) its algorithmic GB/s and
    fraction of the 8 TB/s HBM peak."""
    import numpy as np

    E = 24 ** 3
    rng = np.random.default_rng(8)
    out = {}
    for nm, deformed in ((8, False), (9, False), (10, False), (8, True)):
        nq = nm + 1
        if deformed:
            nqt = nq ** 3
            jac = rng.uniform(0.5, 1.5, E * nqt) * 1e-4
            df = (np.tile(np.eye(3).reshape(-1), E * nqt) + 0.2 * rng.uniform(-1, 1, E * nqt * 9)) * 40.0
        else:
            M = (np.eye(3)[None] + 0.25 * rng.uniform(-1, 1, (E, 3, 3))) * 0.05
            jac = np.abs(np.linalg.det(M))
            df = np.ascontiguousarray(np.transpose(np.linalg.inv(M), (0, 2, 1))).reshape(-1)
        F = hex_flops_per_elem(nm, nq) * E
        for prec, tag in ((N.FP64, "fp64"), (N.FP32, "fp32")):
            key = "Helmholtz_Hex_" + ("Deformed" if deformed else "Regular")
            op = N.GetOperatorFactory().CreateInstance(key, nm, nq, E, prec)
            op.SetJac(jac)
            op.SetDF(df)
            op.SetLambda(1.0)
            best, times = op.autotune()
            peak = PEAK_FP64_TFLOPS if prec == N.FP64 else 157.3
            r = {"best_tile": best}
            for fam, sel in (("valu", lambda t: t < N.TILE_MFMA), ("mfma", lambda t: t >= N.TILE_MFMA)):
                cand = {t: v for t, v in times.items() if sel(t)}
                t = min(cand, key=cand.get)
                r[fam] = {"tile": t, "us": round(cand[t], 2), "tflops": round(F / cand[t] / 1e6, 2),
                          "fp_frac": round(F / cand[t] / 1e6 / peak, 3)}
                if deformed:
                    # per-point geometry makes this kernel HBM-bound: its
                    # algorithmic bytes (coefficients in and out, 7 geometry
                    # values per point) against the 8 TB/s peak
                    B = hex_bytes_per_elem(nm, nq, True, 8 if prec == N.FP64 else 4) * E
                    r[fam]["alg_gbs"] = round(B / cand[t] / 1e3, 1)
                    r[fam]["hbm_frac"] = round(B / cand[t] / 1e3 / PEAK_HBM_GBS, 3)
            r["mfma_over_valu"] = round(r["valu"]["us"] / r["mfma"]["us"], 3)
            out[f"P{nm - 1}_Q{nq}_{'deformed' if deformed else 'regular'}_{tag}"] = r
            del op
    return out


def bench_serve_sessions(repeats=3):
    """HelmSolve requests given as the reference's session files
    (Helmholtz3D_CG_Hex, _Tet, _Hex_AllBCs and a trilinear-hex session, each
    with the full and the element-local CG), all arriving at once: served
    with 4 solves in flight against one at a time.  Reading the sessions and
    building the problems on the host is part of every request's latency."""
    from nektar_amd import contfield as cf
    import test_contfield_cpu as T

    g = os.path.join(ROOT, "tests", "golden", "helmholtz3d")
    sessions = [cf.read_session(os.path.join(g, f)) for f in
                ("Helmholtz3D_Hex_P6.xml", "Helmholtz3D_Tet.xml", "Helmholtz3D_Hex_AllBCs_P6.xml")]
    sessions.append(T.perturbed_hex_session())
    reqs = [dict(session=s, solver=sv, tol=1e-14, arrival_s=0.0) for s in sessions
            for sv in ("ConjugateGradient", "ConjugateGradientLoc")] * repeats
    out = {"requests": len(reqs)}
    cf.serve_sessions(reqs[:4], max_active=4)  # warm (kernels, tables)
    for name, act in (("served", 4), ("one_at_a_time", 1)):
        res, st = cf.serve_sessions(reqs, max_active=act)
        out[name] = {"makespan_s": round(st["makespan_s"], 3), "p50_ms": round(st["p50_ms"], 1),
                     "p99_ms": round(st["p99_ms"], 1), "peak_active": st["peak_active"],
                     "all_converged": all(r["converged"] for r in res)}
    out["served_over_one_at_a_time"] = round(out["one_at_a_time"]["makespan_s"] / out["served"]["makespan_s"], 3)
    return out


def rotated_box_map(n, nm, seed=9):
    """The 32^3-type structured box with every element in a random local
    frame (axis permutation and reflections): map and sign of every local
    coefficient into the structured C0 numbering (N = n(nm-1)+1 nodes per
    direction, node of mode p in element e: e(nm-1) + {0, nm-1, p-1}), the
    Dirichlet mask, and the frames (perm, flip).  A reflected axis swaps the
    two vertex modes and gives bubble mode p the sign (-1)^p."""
    import numpy as np

    rng = np.random.default_rng(seed)
    E, N = n ** 3, n * (nm - 1) + 1
    perm = np.argsort(rng.random((E, 3)), axis=1)
    flip = rng.integers(0, 2, (E, 3))
    ez, ey, ex = np.unravel_index(np.arange(E), (n, n, n))
    ecoord = np.stack([ex, ey, ez], axis=1)
    p = np.arange(nm)
    off = np.where(p == 0, 0, np.where(p == 1, nm - 1, p - 1))
    stride = np.array([1, N, N * N])
    terms, signs = [], []
    for a in range(3):
        fa = flip[:, a][:, None]
        gm = np.where(fa == 1, np.where(p < 2, 1 - p, p), p)          # (E, nm) global-frame mode
        gs = np.where((fa == 1) & (p >= 2), (-1.0) ** p, 1.0)          # (E, nm)
        d = perm[:, a]
        terms.append((ecoord[np.arange(E), d][:, None] * (nm - 1) + off[gm]) * stride[d][:, None])
        signs.append(gs)
    gid = terms[2][:, :, None, None] + terms[1][:, None, :, None] + terms[0][:, None, None, :]
    sg = signs[2][:, :, None, None] * signs[1][:, None, :, None] * signs[0][:, None, None, :]
    g = np.arange(N ** 3)
    gx, gy, gz = g % N, (g // N) % N, g // (N * N)
    dirichlet = (gx == 0) | (gx == N - 1) | (gy == 0) | (gy == N - 1) | (gz == 0) | (gz == N - 1)
    return gid.reshape(-1), sg.reshape(-1), dirichlet, perm, flip


def bench_general(N, torch, n=32, nm=7, nq=8, lam=1.0):
    """The general assembly and CG (nmf_assembly / nmf_linsys) on the 32^3,
    P=6 box with every element in a random local frame (orientation signs
    on odd bubble modes), against the structured solver on the same box:
    solve time, iterations and the largest difference of the solutions, for
    the assembled-vector CG, the element-local CG and the iterative static
    condensation; and the static-condensation diagonal."""
    import numpy as np

    gid, sg, dirichlet, perm, flip = rotated_box_map(n, nm)
    order = np.concatenate([np.flatnonzero(dirichlet), np.flatnonzero(~dirichlet)])
    new = np.empty_like(order)
    new[order] = np.arange(len(order))
    E, h = n ** 3, 1.0 / n
    # geometry of an axis-aligned cube in a rotated frame: d x_{perm[a]} / d xi_a = s_a h/2
    s_ax = np.where(flip == 1, -1.0, 1.0)
    df = np.zeros((E, 9))
    for k in range(3):
        df[np.arange(E), perm[:, k] * 3 + k] = s_ax[:, k] * 2.0 / h
    jac = np.full(E, (h / 2) ** 3)
    op = N.GetOperatorFactory().CreateInstance("Helmholtz_Hex_Regular", nm, nq, E, N.FP64)
    op.SetJac(jac)
    op.SetDF(df.reshape(-1))
    op.SetLambda(lam)
    # forcing f = (3 pi^2 + lam) prod sin(pi x) at the local quadrature points
    z = N.basis_tables(0, 0, nm, nq)["z"]
    ez, ey, ex = np.unravel_index(np.arange(E), (n, n, n))
    centre = (np.stack([ex, ey, ez], axis=1) + 0.5) * h
    K, J, I = np.meshgrid(np.arange(nq), np.arange(nq), np.arange(nq), indexing="ij")
    loc_idx = [I.ravel(), J.ravel(), K.ravel()]
    X = np.repeat(centre[:, None, :], nq ** 3, axis=1)
    eye = np.eye(3)
    for a in range(3):
        # local axis a of element e runs along global axis perm[e, a]
        X += ((s_ax[:, a][:, None] * (h / 2)) * z[loc_idx[a]][None, :])[:, :, None] * eye[perm[:, a]][:, None, :]
    f = (3 * np.pi ** 2 + lam) * np.prod(np.sin(np.pi * X), axis=2).reshape(-1)
    ip = N.GetOperatorFactory().CreateInstance("IProduct_Hex_Regular", nm, nq, E, N.FP64)
    ip.SetJac(jac)
    fd = torch.from_numpy(f).cuda()
    bl = torch.zeros(ip.output_size(), dtype=torch.float64, device="cuda")
    ip(fd, bl)
    amap = N.AssemblyMap(new[gid], sg, len(dirichlet), int(dirichlet.sum()))
    ls = N.LinearSystem(amap, [op])
    ls.set_rhs_local(bl)
    ls.solve(tol=1e-10)  # warm
    info = ls.solve(tol=1e-10)
    s = N.HelmholtzSolver(n, nm, nq, lam)
    s.set_forcing_manufactured(1)
    s.solve(tol=1e-10)
    ref = s.solve(tol=1e-10)
    diff = float(np.max(np.abs(ls.solution()[new] - s.solution())) / np.max(np.abs(s.solution())))
    # static-condensation diagonal of the same system (host checks of the map included)
    t0 = time.perf_counter()
    scd = ls.sc_diagonal()
    sc_ms = (time.perf_counter() - t0) * 1e3
    del ls
    # the element-local CG (LinSysIterSolver ConjugateGradientLoc) on the same system
    ll = N.LinearSystem(amap, [op], solver="ConjugateGradientLoc")
    ll.set_rhs_local(bl)
    ll.solve(tol=1e-10)  # warm
    linfo = ll.solve(tol=1e-10)
    ldiff = float(np.max(np.abs(ll.solution()[new] - s.solution())) / np.max(np.abs(s.solution())))
    del ll
    # iterative static condensation (GlobalSysSoln IterativeStaticCond) of the same system:
    # setup (element Schur complements), then solves with the same right-hand side (the
    # smoothed rhs magnitude then stays the first solve's)
    t0 = time.perf_counter()
    sc = N.LinearSystem(amap, [op], solver="IterativeStaticCond")
    sc_setup_ms = (time.perf_counter() - t0) * 1e3
    sc.set_rhs_local(bl)
    sc.solve(tol=1e-10)  # warm
    sinfo = sc.solve(tol=1e-10)
    sdiff = float(np.max(np.abs(sc.solution()[new] - s.solution())) / np.max(np.abs(s.solution())))
    del sc
    nunk = int((~dirichlet).sum())
    return {"unknowns": nunk, "general_solve_ms": round(info.solve_ms, 3), "general_iterations": info.iterations,
            "general_converged": bool(info.converged),
            "general_solve_gdofs": round(nunk / (info.solve_ms * 1e-3) / 1e9, 4),
            "local_cg_solve_ms": round(linfo.solve_ms, 3), "local_cg_iterations": linfo.iterations,
            "local_cg_converged": bool(linfo.converged), "local_cg_rel_diff": ldiff,
            "sc_diagonal_ms": round(sc_ms, 1), "sc_diagonal_rows": int(np.count_nonzero(scd)),
            "static_cond_setup_ms": round(sc_setup_ms, 1), "static_cond_solve_ms": round(sinfo.solve_ms, 3),
            "static_cond_iterations": sinfo.iterations, "static_cond_converged": bool(sinfo.converged),
            "static_cond_rel_diff": sdiff,
            "structured_solve_ms": round(ref.solve_ms, 3), "structured_iterations": ref.iterations,
            "solution_rel_diff": diff, "negative_signs": int((sg < 0).sum())}


def bench_batch(N, torch, steps, warmup):
    """Batched uneven workload: Sum of element-local DOFs / time of one
    batch submission (graph of grouped launches), FP64 and FP32, against
    the entry-by-entry loop."""
    from nektar_amd.workloads import batch_workload

    work = batch_workload()
    res = {"items": len(work)}
    for prec, tag in ((N.FP64, "fp64"), (N.FP32, "fp32")):
        dt = torch.float64 if prec == N.FP64 else torch.float32
        ops, arrays = [], []
        for (shape, nm, nq, E, jac, df, lam) in work:
            op = N.GetOperatorFactory().CreateInstance(f"Helmholtz_{shape}_Regular", nm, nq, E, prec)
            op.SetJac(jac)
            op.SetDF(df)
            op.SetLambda(lam)
            g = torch.Generator(device="cuda").manual_seed(4)
            x = torch.rand(op.input_size(), dtype=dt, device="cuda", generator=g) * 2 - 1
            y = torch.zeros(op.output_size(), dtype=dt, device="cuda")
            ops.append(op)
            arrays.append((x, y))
        b = N.Batch(mode=N.Batch.GRAPH, streams=4)
        for op, (x, y) in zip(ops, arrays):
            b.add(op, x, y)
        b.autotune()
        info = b.info()
        r = {"kinds": info["groups"], "dofs": info["dofs"]}
        for mode, mtag in ((N.Batch.GRAPH, "graph"), (N.Batch.GROUPED, "grouped"), (N.Batch.PER_ITEM, "per_item")):
            b.set_mode(mode)
            for _ in range(warmup):
                b.run()
            torch.cuda.synchronize()
            s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s0.record()
            for _ in range(steps):
                b.run()
            s1.record()
            torch.cuda.synchronize()
            ms = s0.elapsed_time(s1) / steps
            r[f"{mtag}_us"] = round(ms * 1e3, 1)
            r[f"{mtag}_gdofs"] = round(info["dofs"] / (ms * 1e-3) / 1e9, 2)
        r["graph_over_per_item"] = round(r["per_item_us"] / r["graph_us"], 3)
        # each kind (shape, P, Q) alone as one grouped launch: its roofline
        # fraction, and how the sum of the kinds compares with the graph
        # (the graph runs kinds on 4 streams)
        kinds = {}
        for idx, (shape, nm, nq, E, _, _, _) in enumerate(work):
            kinds.setdefault((shape, nm, nq), []).append(idx)
        peak = PEAK_FP64_TFLOPS if prec == N.FP64 else 157.3
        per, total = {}, 0.0
        for (shape, nm, nq), idxs in sorted(kinds.items()):
            bk = N.Batch(mode=N.Batch.GROUPED, streams=1)
            for i in idxs:
                bk.add(ops[i], *arrays[i])
            for _ in range(warmup):
                bk.run()
            torch.cuda.synchronize()
            s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s0.record()
            for _ in range(steps):
                bk.run()
            s1.record()
            torch.cuda.synchronize()
            us = s0.elapsed_time(s1) / steps * 1e3
            fpe = hex_flops_per_elem(nm, nq) if shape == "Hex" else tet_flops_per_elem(nm, nq)
            fl = sum(work[i][3] for i in idxs) * fpe
            per[f"{shape}_P{nm - 1}_Q{nq}"] = {"items": len(idxs), "elements": sum(work[i][3] for i in idxs),
                                               "us": round(us, 1), "tflops": round(fl / us / 1e6, 2),
                                               "fp_frac": round(fl / us / 1e6 / peak, 3)}
            total += us
            del bk
        r["per_kind"] = per
        r["sum_of_kinds_us"] = round(total, 1)
        r["sum_of_kinds_over_graph"] = round(total / r["graph_us"], 3)
        res[tag] = r
        del b, ops, arrays
    res["fp32_over_fp64"] = round(res["fp64"]["graph_us"] / res["fp32"]["graph_us"], 3)
    return res


def bench_curved(N):
    """Curved 48^3 hex mesh (interior vertices +-10 % of the element width,
    seed 3), P=7, Q=9, per-point geometry: FP64 CG against the mixed
    FP32/FP64 reliable-update PCG, both to relative residual 1e-10."""
    import numpy as np

    s = N.HelmholtzSolver(48, 8, 9, 1.0, 0.1, 3)
    s.set_forcing_manufactured(1)
    tiles = s.autotune()
    nunk = s.nglobal - s.ndirichlet
    out = {"mesh": "48^3 curved hex, P=7, Q=9", "unknowns": nunk, "solver_tiles": list(tiles)}
    sol = {}
    for mixed, tag in ((False, "fp64"), (True, "mixed")):
        s.solve(mixed=mixed, tol=1e-10)  # warm (graph capture, FP32 buffers)
        info = s.solve(mixed=mixed, tol=1e-10)
        sol[tag] = s.solution()
        out[tag] = {"iterations": info.iterations, "outer": info.outer_iterations, "converged": bool(info.converged),
                    "final_rel_residual": info.final_error, "solve_ms": round(info.solve_ms, 2),
                    "solve_gdofs": round(nunk / (info.solve_ms * 1e-3) / 1e9, 4),
                    "l2_error": s.l2_error_manufactured(1)}
    out["mixed_speedup"] = round(out["fp64"]["solve_ms"] / out["mixed"]["solve_ms"], 3)
    out["mixed_vs_fp64_max_rel_diff"] = float(np.max(np.abs(sol["mixed"] - sol["fp64"])) / np.max(np.abs(sol["fp64"])))
    return out


def bench_serve(N, count, budget_gib, streams, max_active):
    """Serving stream: capacity of plain FP64 one-at-a-time solves, Poisson
    arrivals at 80 % of that capacity served one at a time and by the
    serving loop (FP64 and mixed), and the serving loop's own capacity."""
    import numpy as np
    from nektar_amd.workloads import poisson_arrivals, serve_requests

    reqs = serve_requests(count)
    budget = int(budget_gib * (1 << 30))
    # warm-up: load every order's kernels once
    warm = [dict(n=2, nummodes=nm, numpoints=nm + 1, lam=1.0, k=1, tol=1e-10, maxit=5000, mixed=m, arrival_s=0.0)
            for nm in range(4, 10) for m in (0, 1)]
    N.Server(budget, streams, max_active).run(warm)

    def run(arrivals, mixed, streams_, active_):
        rq = [dict(r, mixed=mixed, arrival_s=float(a)) for r, a in zip(reqs, arrivals)]
        res, st = N.Server(budget, streams_, active_).run(rq)
        lat = np.array([x["latency_ms"] for x in res])
        unk = float(sum(x["unknowns"] for x in res))
        span = st["makespan_s"]
        return {"solves_per_s": round(count / span, 3), "solve_gdofs": round(unk / span / 1e9, 4),
                "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
                "makespan_s": round(span, 3), "all_converged": all(x["converged"] for x in res),
                "max_final_error": max(x["final_error"] for x in res),
                "iterations": int(sum(x["iterations"] for x in res)),
                "peak_active": st["peak_active"], "peak_gib": round(st["peak_bytes"] / (1 << 30), 2),
                "solvers_reused": st["solvers_reused"], "evictions": st["evictions"]}

    zero = np.zeros(count)
    out = {"requests": count, "hbm_budget_gib": budget_gib, "streams": streams, "max_active": max_active,
           "unknowns_total": None}
    cap1 = run(zero, 0, 1, 1)
    rate = 0.8 * cap1["solves_per_s"]
    arr = poisson_arrivals(count, rate)
    out["poisson_rate_per_s"] = round(rate, 3)
    out["fp64_one_at_a_time_capacity"] = cap1
    out["fp64_one_at_a_time_poisson"] = run(arr, 0, 1, 1)
    out["served_fp64_poisson"] = run(arr, 0, streams, max_active)
    out["served_mixed_poisson"] = run(arr, 1, streams, max_active)
    out["served_fp64_capacity"] = run(zero, 0, streams, max_active)
    out["served_mixed_capacity"] = run(zero, 1, streams, max_active)
    out["unknowns_total"] = int(sum((r["n"] * (r["nummodes"] - 1) - 1) ** 3 for r in reqs))
    base = cap1["solve_gdofs"]
    out["served_fp64_over_one_at_a_time"] = round(out["served_fp64_capacity"]["solve_gdofs"] / base, 3)
    out["served_mixed_over_one_at_a_time"] = round(out["served_mixed_capacity"]["solve_gdofs"] / base, 3)
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n_side, nm, nq, target_s=10.0):
    """The oracle's FP64 Helmholtz apply on the host cores: medians of >= 10
    repeats on all cores and on one core, for a sample of 4,096 elements of
    the same mesh and order (1/8 of the mesh)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import _oracle as O

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cores = max(1, min(cores, 64))
    ne = 4096
    verts = O.hex_mesh(n_side, 0.0, 2)[: ne * 24]
    jac, df = O.hex_geom(verts, ne, nq, False)
    x = np.random.default_rng(2).uniform(-1, 1, ne * nm ** 3)

    def median_time(threads, budget_s):
        O.set_threads(threads)
        O.helmholtz(O.HEX, nm, nq, ne, False, jac, df, 1.0, x)  # warm-up
        times = []
        while len(times) < 10 or (sum(times) < budget_s and len(times) < 10000):
            t0 = time.perf_counter()
            O.helmholtz(O.HEX, nm, nq, ne, False, jac, df, 1.0, x)
            times.append(time.perf_counter() - t0)
        return float(np.median(times)), len(times), sum(times)

    t_all, n_all, s_all = median_time(cores, 0.7 * target_s)
    t_one, n_one, s_one = median_time(1, 0.3 * target_s)
    dofs = ne * nm ** 3
    return {"value": dofs / t_all / 1e9, "unit": "GDOF/s", "cores": cores, "kind": "port",
            "value_1core": dofs / t_one / 1e9, "statistic": "median",
            "cpu_model": _cpu_model(), "simd": "AVX2, 4 FP64 lanes (elements) per vector (-march=x86-64-v3)",
            "sample": f"{ne} elements of the 32^3 P=6 Q=8 mesh, FP64 Helmholtz apply: median of {n_all} reps "
                      f"({s_all:.1f} s) on {cores} cores (OpenMP over element blocks), median of {n_one} reps "
                      f"({s_one:.1f} s) on 1 core"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-solve", action="store_true", help="skip the CG solves")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched uneven workload")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving stream")
    ap.add_argument("--no-curved", action="store_true", help="skip the curved 48^3 P=7 solves")
    ap.add_argument("--no-general", action="store_true",
                    help="skip the general assembly solve (32^3 P=6 box in random element frames)")
    ap.add_argument("--no-ops", action="store_true",
                    help="skip the per-operator timings (hex and tet, 32,768 elements, P=6)")
    ap.add_argument("--no-mfma", action="store_true",
                    help="skip the vector against matrix-core (MFMA) Helmholtz comparison at P = 7..9")
    ap.add_argument("--no-serve-sessions", action="store_true",
                    help="skip serving the reference's session files")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain device copy of 512 MiB ten times (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read_per_launch": src.numel() * 8, "copy_bytes_written_per_launch": src.numel() * 8,
                          "launches": 10}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build runs on a single MI355X); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import nektar_amd as N

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    n_side, nm, nq, lam = 32, 7, 8, 1.0
    ne = n_side ** 3
    sys.path.insert(0, os.path.join(ROOT, "tests"))

    # geometry of the regular mesh: one J and df per element (axis-aligned cubes)
    import numpy as np
    h = 1.0 / n_side
    jac = np.full(ne, (h / 2) ** 3)
    df = np.tile(np.array([2 / h, 0, 0, 0, 2 / h, 0, 0, 0, 2 / h]), ne)

    results = {}
    for prec, tag in ((N.FP64, "fp64"), (N.FP32, "fp32")):
        op = N.GetOperatorFactory().CreateInstance("Helmholtz_Hex_Regular", nm, nq, ne, prec)
        op.SetJac(jac)
        op.SetDF(df)
        op.SetLambda(lam)
        tile, tile_us = op.autotune()  # elements per workgroup, measured on this mesh
        dt = torch.float64 if prec == N.FP64 else torch.float32
        g = torch.Generator(device="cuda").manual_seed(2)
        x = (torch.rand(op.input_size(), dtype=dt, device="cuda", generator=g) * 2 - 1).contiguous()
        y = torch.zeros(op.output_size(), dtype=dt, device="cuda")
        ms, wall = time_apply(op, x, y, args.steps, args.warmup, torch)
        # reported beside the headline, not in it: the GPU's clocks ramp up
        # over the first tens of milliseconds of sustained work after idle,
        # so the same K launches are timed again after ~30 ms of untimed ones
        settle = max(args.warmup, int(30.0 / max(ms, 1e-3)))
        ms_settled, _ = time_apply(op, x, y, args.steps, settle, torch)
        results[tag] = (ms, wall, tile, tile_us, ms_settled)
        del op, x, y

    ms64, wall64, tile64, tiles64, ms64s = results["fp64"]
    ms32, _, tile32, tiles32, ms32s = results["fp32"]
    ndof = ne * nm ** 3
    F = hex_flops_per_elem(nm, nq) * ne
    value = ndof / (ms64 * 1e-3) / 1e9
    achieved = F / (ms64 * 1e-3) / 1e12
    traffic = read_traffic(f"helmholtz_fp64_tile{tile64}")
    # the kernel is compute-bound (17.4 flop/B); the tile the autotuner chose
    # runs on FP64 vector FMAs or on the matrix cores (tiles >= TILE_MFMA);
    # MI355X's FP64 vector and FP64 matrix peaks are the same 78.6 TFLOP/s
    on_mfma = tile64 >= N.TILE_MFMA
    kname = (f"hexmf::helmholtz_kernel<double,7,8,false,{tile64 - N.TILE_MFMA}>" if on_mfma
             else f"hex::helmholtz_kernel<double,7,8,false,{tile64}>")
    roofline = {"bound": "mfma" if on_mfma else "valu_fp64", "achieved": round(achieved, 3),
                "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": traffic,
                "kernel": kname,
                "flops_per_launch": F,
                "alg_bytes_per_launch": hex_bytes_per_elem(nm, nq, False, 8) * ne}

    extra = {
        "apply_fp64_gdofs": round(value, 3),
        "apply_fp32_gdofs": round(ndof / (ms32 * 1e-3) / 1e9, 3),
        "apply_fp32_over_fp64": round(ms64 / ms32, 3),
        "apply_fp64_us": round(ms64 * 1e3, 2),
        "apply_fp32_us": round(ms32 * 1e3, 2),
        "apply_fp64_us_settled": round(ms64s * 1e3, 2),
        "apply_fp32_us_settled": round(ms32s * 1e3, 2),
        "apply_fp64_gdofs_settled": round(ndof / (ms64s * 1e-3) / 1e9, 3),
        "fp_frac_settled": round(F / (ms64s * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4),
        "tile_fp64": tile64, "tile_fp32": tile32,
        "autotune_us_fp64": {str(k): round(v, 2) for k, v in tiles64.items()},
        "autotune_us_fp32": {str(k): round(v, 2) for k, v in tiles32.items()},
    }
    if not args.no_solve:
        s = N.HelmholtzSolver(n_side, nm, nq, lam)
        s.set_forcing_manufactured(1)
        extra["solver_tiles"] = list(s.autotune())
        nunk = s.nglobal - s.ndirichlet
        for mixed, tag in ((False, "fp64"), (True, "mixed")):
            s.solve(mixed=mixed, tol=1e-10)  # warm (graph capture, FP32 buffers)
            info = s.solve(mixed=mixed, tol=1e-10)
            l2 = s.l2_error_manufactured(1)
            extra[f"solve_{tag}"] = {
                "iterations": info.iterations, "outer": info.outer_iterations,
                "converged": bool(info.converged), "final_rel_residual": info.final_error,
                "solve_ms": round(info.solve_ms, 3),
                "solve_gdofs": round(nunk / (info.solve_ms * 1e-3) / 1e9, 4),
                "solves_per_s": round(1e3 / info.solve_ms, 3), "l2_error": l2}
        extra["solve_mixed_speedup"] = round(extra["solve_fp64"]["solve_ms"] / extra["solve_mixed"]["solve_ms"], 3)
        extra["unknowns"] = nunk
        del s

    if not args.no_ops:
        extra["ops_p6_32768"] = bench_ops(N, torch)
    if not args.no_mfma:
        extra["mfma_p7_p9"] = bench_mfma(N, torch)
    if not args.no_general:
        extra["general_32_p6"] = bench_general(N, torch)
    if not args.no_curved:
        extra["curved_48_p7"] = bench_curved(N)
    if not args.no_batch:
        extra["batch"] = bench_batch(N, torch, steps=10, warmup=2)
    if not args.no_serve:
        extra["serve"] = bench_serve(N, args.serve_requests, budget_gib=64, streams=4, max_active=16)
    if not args.no_serve_sessions:
        extra["serve_sessions"] = bench_serve_sessions()

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "GDOF/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms64, 6),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "Helmholtz operator apply, 32^3 regular hex on [0,1]^3, P=6 (7 modes), GLL Q=8, "
                               "lambda=1, element-local coefficients U[-1,1)",
                   "elements": ne, "nummodes": nm, "numpoints": nq, "local_dofs": ndof},
        "roofline": roofline,
        "extra": extra,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(n_side, nm, nq)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
