#!/usr/bin/env python3
"""Solved DOF/s at 1e-10 relative residual on one MI355X: FP32-Krylov + FP64
iterative refinement (the value) beside plain FP64 Jacobi-PCG, on the
single large C3D8 workload (200x20x20 hexes, 265,923 DOF).

A step is one full linear-static solve with the mesh (coordinates,
connectivity, CSR pattern) already resident in HBM: device assembly of the
element matrices into the CSR values, Dirichlet constraints, the PCG solve
to 1e-10, stress recovery and reactions, with u and the von Mises fields
copied back to the host. The dominant kernel's roofline comes from the
solver's own SpMV kernel timed with HIP events on the solver's stream; the
CPU baseline is the oracle (the restated reference CPU path, reference
lookups and thread policy) on a bounded sample: the full assembly plus a
capped number of PCG iterations, scaled to the full iteration count.

Prints one JSON line.
"""
import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the FP64 SpMV on the 200x20x20 C3D8 block, from the
# FETCH_SIZE / WRITE_SIZE counters of rocprofv3 scaled by a 1 GiB device copy
# (profiles/cg64_spmv_c3d8_200x20x20.md); None where no counter run exists.
TRAFFIC_BYTES = {("cfg2", "fp64"): 207.0e6}
CONFIGS = {
    "cfg2": ("C3D8 200x20x20 block 20x2x2 m, jitter 0.1h, x=0 clamped, -1e6 Pa z-traction on x=20", "cfg2"),
    "cfg1": ("CPS4 400x40 cantilever 10x1 m, jitter 0.15h, -1e4 N tip load", "cfg1"),
    "cfg4": ("64 load cases on one C3D8 120x12x12 cantilever (61,347 DOF), 1-8 point loads each; "
             "one block solve vs 64 single solves", "cfg4"),
    "cfg5": ("served Poisson stream: 20% plane instances (DOF 2e3..2e5), 80% C3D8 blocks (DOF 1e4..3e5), "
             "200 GB admission budget", "cfg5"),
    "cfg3": ("512 uneven plane-stress instances, DOF log-uniform 2e3..2e5, CPS4 p=0.6 else CPS3, "
             "left edge clamped, tip load; one batched launch vs one-by-one FP64", "cfg3"),
}


def build_model(name, plane_size=None):
    from solvix_amd import decks
    if name == "cfg2":
        return decks.cfg2()
    return decks.cfg1(nx=plane_size[0], ny=plane_size[1]) if plane_size else decks.cfg1()


def cpu_baseline(model, full_iterations, cap):
    """Oracle in reference mode (linear-scan lookups, min(nproc, 8) threads,
    threads spawned per SpMV, serial dots) on: full assembly + `cap` PCG
    iterations; the PCG time is scaled to `full_iterations`."""
    import oracle
    t0 = time.perf_counter()
    r = oracle.solve(model, faithful=True, max_iter=cap, want_csr=False)
    wall = time.perf_counter() - t0
    t_asm = r.times_ms["assemble"] / 1e3
    t_pcg = r.times_ms["pcg"] / 1e3
    est = t_asm + t_pcg * full_iterations / max(1, r.iterations)
    return {
        "value": model.n_dofs() / est,
        "unit": "solved DOF/s",
        "cores": oracle.worker_threads(),
        "kind": "port",
        "sample": (f"oracle on {platform.node() or 'host'} ({os.cpu_count()} logical CPUs): full assembly "
                   f"{t_asm:.2f} s + {r.iterations} PCG iterations {t_pcg:.2f} s (sample wall {wall:.1f} s), "
                   f"PCG scaled to {full_iterations} iterations -> {est:.1f} s per solve"),
    }


def copy_calibration(mib):
    """Three device-to-device copies of `mib` MiB (torch's copy kernel): the
    reference point that scales hardware byte counters to bytes."""
    import torch
    n = mib << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(3):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / 3
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "avg_ms": ms,
                      "GBps_read_plus_write": 2 * n / ms / 1e6}), flush=True)


def bench_batch(args):
    """The batched workload: every instance's mesh (coordinates, connectivity,
    CSR pattern) resident; a step assembles every instance, applies its
    constraints, solves and recovers. Plain arm: FP64 PCG one instance after
    another. Batched arms: all instances in one Krylov engine (FP32-IR, and
    FP64 for reference), SpMV shape from the autotuner. Solved DOF counts
    the DOF of converged instances."""
    import numpy as np

    from solvix_amd import decks
    from solvix_amd.core import Batch, Context, Mesh, constrained_dofs, element_materials
    from solvix_amd.model import Precision

    models = decks.cfg3()
    ctx = Context(0)
    prep = []
    for m in models:
        E, nu = element_materials(m)
        ndof = m.ndof()
        pos = {int(nid): i for i, nid in enumerate(m.node_ids)}
        lr = np.array([pos[int(nid)] * ndof + int(d) - 1 for nid, d in zip(m.cload_node, m.cload_dof)], np.int64)
        prep.append((Mesh.from_model(m, ctx), E, nu, m.thickness, lr, m.cload_value, constrained_dofs(m)[0]))
    dofs = np.array([m.n_dofs() for m in models], np.int64)
    print(f"[bench] {len(models)} instances, {dofs.sum()} DOF, meshes resident", file=sys.stderr, flush=True)

    def assemble(i):
        mesh, E, nu, t, lr, lv, rows = prep[i]
        sy = mesh.assemble(E, nu, t, lr, lv)
        sy.apply_dirichlet(rows)
        return sy

    phase = {}

    def clock(key, t0):
        phase[key] = phase.get(key, 0.0) + time.perf_counter() - t0

    def sequential():
        us, its, ok = [], [], []
        for i in range(len(prep)):
            t0 = time.perf_counter()
            sy = assemble(i)
            clock("assemble", t0)
            t0 = time.perf_counter()
            try:
                u, it, _ = sy.solve(1e-10, 0, Precision.FP64)
                ok.append(True)
            except Exception:
                u, it = sy.solution(), sy.stats()["iterations"]
                ok.append(False)
            clock("solve", t0)
            t0 = time.perf_counter()
            sy.recover(prep[i][6])
            clock("recover", t0)
            us.append(u)
            its.append(it)
            sy.close()
        return us, np.array(its), np.array(ok)

    tuned = {}
    stats = []  # per-instance solve statistics of the last batched pass

    def batched(prec, mf=False):
        t0 = time.perf_counter()
        systems = [assemble(i) for i in range(len(prep))]
        clock("assemble", t0)
        t0 = time.perf_counter()
        b = Batch(ctx)
        for sy in systems:
            b.add(sy)
        t = tuned[(prec, mf)]  # the autotuned shape of this (2D-only) batch, node-block size 2
        ctx.set_tuning(unroll=t["unroll"][2], rows_per_tile=t["rows_per_tile"][2])
        ctx.set_operator("matrix_free" if mf else "assembled")
        try:
            res = b.solve(prec, raise_on_failure=False)
        finally:
            ctx.set_tuning()
            ctx.set_operator("assembled")
        clock("solve", t0)
        t0 = time.perf_counter()
        for i, sy in enumerate(systems):
            sy.recover(prep[i][6])
        clock("recover", t0)
        us = [sy.solution() for sy in systems]
        its = np.array([r[0] for r in res])
        ok = np.array([r[2] for r in res])
        stats.clear()
        stats.extend(sy.stats() for sy in systems)
        return b, systems, us, its, ok

    # Arms: label -> (precision, matrix-free operator)
    ARMS = {"fp64_batched": (Precision.FP64, False), "fp64_batched_mf": (Precision.FP64, True),
            "fp32_ir_batched": (Precision.FP32_IR, False)}
    labels = ("fp64_sequential", "fp64_batched", "fp64_batched_mf", "fp32_ir_batched")
    if args.batch_arms != "all":  # profiling runs: one batched engine alone
        labels = (args.batch_arms,)
    # autotune once per batched arm (untimed; the tuned shape depends on the batch only)
    for label in labels:
        if label not in ARMS:
            continue
        prec, mf = ARMS[label]
        systems = [assemble(i) for i in range(len(prep))]
        b = Batch(ctx)
        for sy in systems:
            b.add(sy)
        ctx.set_operator("matrix_free" if mf else "assembled")
        try:
            tuned[(prec, mf)] = b.autotune(prec, 3)
        finally:
            ctx.set_operator("assembled")
        print(f"[bench] autotune {label}: {tuned[(prec, mf)]}", file=sys.stderr, flush=True)
        b.close()
        for sy in systems:
            sy.close()

    out = {}
    for label in labels:
        for k in range(args.warmup + args.steps):
            phase.clear()
            t0 = time.perf_counter()
            if label == "fp64_sequential":
                us, its, ok = sequential()
                b = systems = None
            else:
                b, systems, us, its, ok = batched(*ARMS[label])
            dt = time.perf_counter() - t0
            timed = k >= args.warmup
            split = " ".join(f"{key} {v:.2f} s" for key, v in phase.items())
            print(f"[bench] {label}: pass {k + 1}/{args.warmup + args.steps} {dt:.2f} s ({split}), converged "
                  f"{int(ok.sum())}/{ok.size}, iterations max {int(its.max())} mean {its.mean():.0f}",
                  file=sys.stderr, flush=True)
            if timed:
                rec = out.setdefault(label, {"dt": 0.0, "solve_s": 0.0})
                rec["dt"] += dt
                rec["solve_s"] += phase.get("solve", 0.0)
                rec.update(us=us, its=its, ok=ok)
                if b is not None:
                    ms, by = b.time_spmv(20)
                    rec.update(spmv_ms=ms, spmv_bytes=by, stats=list(stats))
            if b is not None:
                b.close()
                for sy in systems:
                    sy.close()

    def dps(rec):
        return float(dofs[rec["ok"]].sum()) * args.steps / rec["dt"]

    if "fp64_sequential" not in out:
        lab = labels[0]
        rec = out[lab]
        print(json.dumps({"arm": lab, "value": round(dps(rec), 1), "unit": "solved DOF/s",
                          "ms_per_step": round(rec["dt"] / args.steps * 1e3, 3),
                          "solve_s": round(rec["solve_s"] / args.steps, 3), "converged": int(rec["ok"].sum()),
                          "iterations_mean": round(float(rec["its"].mean()), 1),
                          "spmv_us": round(rec.get("spmv_ms", 0.0) * 1e3, 2),
                          "spmv_bytes": rec.get("spmv_bytes")}), flush=True)
        return

    seq = out["fp64_sequential"]
    mix = out["fp32_ir_batched"]
    b64 = out["fp64_batched"]
    bmf = out["fp64_batched_mf"]

    def err(rec):
        return float(max(np.abs(a - c).max() / max(np.abs(c).max(), 1e-300) for a, c in zip(rec["us"], seq["us"])))

    # Algorithmic bytes of a whole batched solve: every instance's iterations
    # times the bytes of one iteration (SpMV with its direction step, update,
    # the per-chunk / per-period direction step), FP32-IR refreshes and FP64
    # fallback iterations counted as such; over the measured solve time this is
    # the bandwidth the whole solve sustains, tails and launches included.
    def solve_bytes(rec, prec, mf=False):
        total = 0.0
        for i, st in enumerate(rec["stats"]):
            mesh = prep[i][0]
            bs = models[i].ndof()
            n, nnz = float(mesh.n), float(mesh.nnz)
            nblk, ch = nnz / (bs * bs), math.ceil(mesh.n / bs / 64)
            it64 = 8 * nnz + 4 * nblk + 8 * ch + 32 * n + 64 * n + 24 * n / 32
            if mf:  # matrix-free product: nodes and connectivity instead of the matrix
                nen = 4 if int(models[i].kinds[0]) == 1 else 3
                it64 = 81 * n / bs + 4 * nen * models[i].n_elem + 64 * n + 24 * n / 32
            if prec == Precision.FP64:
                total += st["iterations"] * it64
                continue
            it32 = 4 * nnz + 4 * nblk + 8 * ch + 16 * n + 32 * n + 12 * n / 8
            refresh = 24 * n + (8 * nnz + 4 * nblk + 8 * ch + 28 * n) + 12 * n
            fp64_its = st["fp64_iterations"]
            total += (st["iterations"] - fp64_its) * it32 + st["refreshes"] * refresh + fp64_its * it64
        return total

    def whole(rec, prec, mf=False):
        by = solve_bytes(rec, prec, mf)
        t = rec["solve_s"] / args.steps
        return {"algorithmic_bytes": by, "solve_s": round(t, 3), "GBps": round(by / t / 1e9, 1),
                "frac": round(by / t / 1e9 / HBM_PEAK_GBS, 4)}

    def roof(rec, name):
        ach = rec["spmv_bytes"] / (rec["spmv_ms"] * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": name,
                "bytes_per_launch": rec["spmv_bytes"], "avg_launch_us": round(rec["spmv_ms"] * 1e3, 2)}

    def arm(rec, kernel, prec, mf=False):
        return {"value": round(dps(rec), 1), "ms_per_step": round(rec["dt"] / args.steps * 1e3, 3),
                "converged": int(rec["ok"].sum()), "iterations_max": int(rec["its"].max()),
                "iterations_mean": round(float(rec["its"].mean()), 1),
                "iterations_dof_weighted": round(float((rec["its"] * dofs).sum() / dofs.sum()), 1),
                "max_rel_err_u_vs_fp64": err(rec),
                "roofline": roof(rec, kernel), "whole_solve": whole(rec, prec, mf),
                "speedup_vs_fp64_sequential": round(dps(rec) / dps(seq), 3)}

    arms = {"fp64_batched_mf": arm(bmf, "cg64_mf_spmv_kernel<4|3> (whole batch)", Precision.FP64, True),
            "fp64_batched": arm(b64, "cg64_spmv_kernel<2,...> (whole batch)", Precision.FP64),
            "fp32_ir_batched": arm(mix, "ir32_spmv_kernel<2,...> (whole batch)", Precision.FP32_IR)}
    dtypes = {"fp64_batched_mf": "f64", "fp64_batched": "f64", "fp32_ir_batched": "f32 Krylov + f64 refinement"}
    # the value is the fastest batched arm that solves every instance within the error bar
    good = [k for k, a in arms.items() if a["converged"] == len(models) and a["max_rel_err_u_vs_fp64"] <= 1e-8]
    best = max(good or list(arms), key=lambda k: arms[k]["value"])
    a = arms[best]
    line = {
        "metric": "solved DOF/s at 1e-10 rel. residual: mixed-prec batched vs plain FP64 PCG",
        "value": a["value"],
        "unit": "solved DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": a["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtypes[best],
        "data": "synthetic (seeded SplitMix64 deck generator)",
        "config": {"workload": CONFIGS["cfg3"][0], "instances": len(models), "n_dof_total": int(dofs.sum()),
                   "tolerance": 1e-10},
        "arm": best,
        "converged": a["converged"],
        "iterations_max": a["iterations_max"],
        "max_rel_err_u_vs_fp64": a["max_rel_err_u_vs_fp64"],
        "roofline": a["roofline"],
        "whole_solve": a["whole_solve"],
        "speedup_vs_fp64": a["speedup_vs_fp64_sequential"],
        "autotune": {k: tuned[ARMS[k]] for k in ARMS},
        "fp64_sequential": {"value": round(dps(seq), 1), "ms_per_step": round(seq["dt"] / args.steps * 1e3, 3),
                            "converged": int(seq["ok"].sum()), "iterations_max": int(seq["its"].max()),
                            "iterations_mean": round(float(seq["its"].mean()), 1)},
        "arms": arms,
    }
    print(json.dumps(line), flush=True)


def bench_multi(args):
    """64 load cases on one assembled, constrained C3D8 system. A step solves
    all of them to 1e-10: either one FP64 PCG per case (set the load vector,
    solve) or one block PCG over all cases (the matrix read once per
    iteration for all of them)."""
    import numpy as np

    from solvix_amd import decks
    from solvix_amd.core import Context, assemble, constrained_dofs
    from solvix_amd.model import Precision

    model, cases = decks.cfg4()
    ctx = Context(0)
    mesh, sy = assemble(model, ctx)
    sy.apply_dirichlet(constrained_dofs(model)[0])
    n = mesh.n
    pos = {int(nid): i for i, nid in enumerate(model.node_ids)}
    F = np.zeros((len(cases), n))
    for k, (nodes, dofs, vals) in enumerate(cases):
        for nid, d, v in zip(nodes, dofs, vals):
            F[k, pos[int(nid)] * 3 + int(d) - 1] += v
    out = {}
    arms = [a for a in ("single", "block") if args.multi_arm in ("both", a)]
    for arm in arms:
        for k in range(args.warmup + args.steps):
            t0 = time.perf_counter()
            if arm == "single":
                U = np.zeros_like(F)
                its = np.zeros(len(cases), np.int64)
                for j in range(len(cases)):
                    sy.set_rhs(F[j])
                    U[j], its[j], _ = sy.solve(1e-10, 0, Precision.FP64)
                ok = np.ones(len(cases), bool)
            else:
                U, its, res, ok, dev_ms = sy.solve_multi(F, 1e-10, 0, Precision.FP64, raise_on_failure=False)
            dt = time.perf_counter() - t0
            print(f"[bench] {arm}: pass {k + 1}/{args.warmup + args.steps} {dt:.3f} s, converged "
                  f"{int(ok.sum())}/{ok.size}, iterations max {int(its.max())} mean {its.mean():.0f}",
                  file=sys.stderr, flush=True)
            if k >= args.warmup:
                rec = out.setdefault(arm, {"dt": 0.0})
                rec["dt"] += dt
                rec.update(U=U, its=its, ok=ok)
    if len(arms) < 2:
        return
    a, b = out["block"], out["single"]
    dof = float(n) * len(cases) * args.steps
    err = float(max(np.abs(a["U"][j] - b["U"][j]).max() / np.abs(b["U"][j]).max() for j in range(len(cases))))
    line = {
        "metric": "solved DOF/s at 1e-10 rel. residual: mixed-prec batched vs plain FP64 PCG",
        "value": round(dof / a["dt"], 1),
        "unit": "solved DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(a["dt"] / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 deck generator)",
        "config": {"workload": CONFIGS["cfg4"][0], "n_dof": n, "nnz": int(mesh.nnz), "load_cases": len(cases),
                   "tolerance": 1e-10},
        "iterations_max": int(a["its"].max()),
        "converged": int(a["ok"].sum()),
        "max_rel_err_u_vs_single": err,
        "single_fp64": {"value": round(dof / b["dt"], 1), "ms_per_step": round(b["dt"] / args.steps * 1e3, 3),
                        "iterations_max": int(b["its"].max())},
        "speedup_vs_fp64": round(b["dt"] / a["dt"], 3),
    }
    print(json.dumps(line), flush=True)


def bench_serve(args):
    """Open-loop served stream: requests are submitted at their Poisson arrival
    times from a client thread while the server solves; latency is submit to
    result. Requests arriving in the first `lead` seconds are served but not
    measured. Solved DOF/s counts the measured requests' DOF over the span
    from the first measured submission to the last measured completion."""
    import threading

    import numpy as np

    from solvix_amd import decks
    from solvix_amd.core import Context, Server, solve_with_method
    from solvix_amd.model import Method, Precision

    stream = list(decks.cfg5_stream(duration_s=args.lead + args.duration, rate=args.rate))
    print(f"[bench] {len(stream)} requests generated ({args.rate}/s over {args.lead + args.duration} s)",
          file=sys.stderr, flush=True)
    ctx = Context(0)
    prec = Precision.FP64 if args.serve_precision == "fp64" else Precision.FP32_IR
    ctx.set_operator(args.serve_operator)  # the server keeps the operator in force at its creation
    srv = Server(ctx, prec, hbm_budget=200 << 30, max_active=1024, use_graph=args.serve_graph == "on")
    ctx.set_operator("assembled")  # the FP64 references below multiply the assembled K
    subs = [None] * len(stream)
    t0 = time.perf_counter()

    def client():
        for i, req in enumerate(stream):
            delay = t0 + req.arrival_s - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            subs[i] = (time.perf_counter(), srv.submit(req.model))

    th = threading.Thread(target=client)
    th.start()
    done = []
    last_print = time.perf_counter()
    for i, req in enumerate(stream):
        while subs[i] is None:
            time.sleep(0.001)
        t_sub, ticket = subs[i]
        sol, info = srv.wait(ticket)
        done.append((req.arrival_s, t_sub - t0, time.perf_counter() - t0, info, req.model.n_dofs(), sol))
        if time.perf_counter() - last_print > 5.0:
            st = srv.stats()
            print(f"[bench] {i + 1}/{len(stream)} done at {time.perf_counter() - t0:.1f} s, active "
                  f"{st['peak_active']} peak, queued {st['queued']}", file=sys.stderr, flush=True)
            last_print = time.perf_counter()
    th.join()
    st = srv.stats()
    srv.close()
    meas = [d for d in done if d[0] >= args.lead]
    lat = np.array([d[3]["latency_ms"] for d in meas])
    span = max(d[2] for d in meas) - min(d[1] for d in meas)
    dof = float(sum(d[4] for d in meas))
    # per-request error vs FP64: the smallest measured requests solved alone in FP64
    sample = sorted(meas, key=lambda d: d[4])[:8]
    errs = []
    for d in sample:
        req = stream[[r.arrival_s for r in stream].index(d[0])]
        ref = solve_with_method(req.model, Method.AUTO, Precision.FP64, ctx)
        e = np.abs(d[5].displacements - ref.displacements).max() / np.abs(ref.displacements).max()
        errs.append(float(e))
    # plain FP64 PCG, one request after another, on a sample of the measured
    # requests (every k-th, about 60): the same mix solved without the server
    k = max(1, len(meas) // 60)
    plain = meas[::k]
    t1 = time.perf_counter()
    for d in plain:
        req = stream[[r.arrival_s for r in stream].index(d[0])]
        solve_with_method(req.model, Method.AUTO, Precision.FP64, ctx)
    plain_dps = float(sum(d[4] for d in plain)) / (time.perf_counter() - t1)
    print(f"[bench] plain FP64 one by one on {len(plain)} sampled requests: {plain_dps / 1e6:.2f} M DOF/s",
          file=sys.stderr, flush=True)
    line = {
        "metric": "solved DOF/s at 1e-10 rel. residual: mixed-prec batched vs plain FP64 PCG",
        "value": round(dof / span, 1),
        "unit": "solved DOF/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": round(span * 1e3, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == Precision.FP64 else "f32 Krylov + f64 refinement",
        "data": "synthetic (seeded SplitMix64 deck generator)",
        "config": {"workload": CONFIGS["cfg5"][0], "precision": args.serve_precision,
                   "operator": args.serve_operator if prec == Precision.FP64 else "assembled",
                   "rate_per_s": args.rate, "measured_s": args.duration, "graphs": args.serve_graph,
                   "lead_s": args.lead, "requests": len(stream), "measured_requests": len(meas),
                   "offered_dof_per_s": round(sum(r.model.n_dofs() for r in stream) / (args.lead + args.duration), 1)},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "max": float(lat.max())},
        "server": st,
        "operator": args.serve_operator if prec == Precision.FP64 else "assembled",
        "max_rel_err_u_vs_fp64_sample": max(errs) if errs else None,
        "fp64_sequential_sample": {"value": round(plain_dps, 1), "requests": len(plain)},
        "speedup_vs_fp64": round(dof / span / plain_dps, 3),
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="cfg2", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-cap", type=int, default=150, help="PCG iterations in the CPU sample")
    ap.add_argument("--arms", default="both", choices=["both", "fp64", "fp32_ir"],
                    help="run one precision only (profiling); the JSON line needs both")
    ap.add_argument("--spmv-reps", type=int, default=100, help="launches of the timed SpMV kernel")
    ap.add_argument("--rotation", type=int, default=2,
                    help="single C3D8 solve: distinct matrices the SpMV rotation leg cycles through (0: skip)")
    ap.add_argument("--batch-arms", default="all", choices=["all", "fp64_batched", "fp64_batched_mf", "fp32_ir_batched"],
                    help="batched workload: every arm, or one batched engine alone (profiling)")
    ap.add_argument("--operator", default="assembled", choices=["assembled", "matrix_free"],
                    help="single solves: operator of the FP64 arm (matrix_free: plane meshes)")
    ap.add_argument("--plane-size", default=None, metavar="NXxNY",
                    help="cfg1 only: the cantilever on an NX x NY mesh (profiling at other sizes)")
    ap.add_argument("--multi-arm", default="both", choices=["both", "single", "block"],
                    help="multi-RHS workload: run one arm only (profiling)")
    ap.add_argument("--rate", type=float, default=200.0, help="served stream: arrivals per second")
    ap.add_argument("--serve-operator", default="matrix_free", choices=["assembled", "matrix_free"],
                    help="served stream: operator of the server's FP64 iterations")
    ap.add_argument("--serve-precision", default="fp64", choices=["fp32_ir", "fp64"],
                    help="served stream: precision of the server's solves")
    ap.add_argument("--serve-graph", default="off", choices=["on", "off"],
                    help="served stream: run the server's chunks as captured hipGraphs (on) or kernel by kernel")
    ap.add_argument("--duration", type=float, default=60.0, help="served stream: measured seconds")
    ap.add_argument("--lead", type=float, default=10.0, help="served stream: unmeasured seconds before")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {args.gpus} requested): the solver path is a single-GPU "
              f"build; multi-GPU sharding is design-only.", file=sys.stderr)
        sys.exit(2)

    if args.config == "cfg3":
        bench_batch(args)
        return
    if args.config == "cfg5":
        bench_serve(args)
        return
    if args.config == "cfg4":
        bench_multi(args)
        return

    import numpy as np

    from solvix_amd.core import Context, Mesh, constrained_dofs, element_materials
    from solvix_amd.model import Precision

    size = tuple(int(v) for v in args.plane_size.split("x")) if args.plane_size else None
    model = build_model(args.config, size)
    n_dofs = model.n_dofs()
    ctx = Context(0)
    ctx.set_operator(args.operator)
    mesh = Mesh.from_model(model, ctx)
    E, nu = element_materials(model)
    ndof = model.ndof()
    pos = {int(n): i for i, n in enumerate(model.node_ids)}
    load_rows = np.array([pos[int(n)] * ndof + int(d) - 1 for n, d in zip(model.cload_node, model.cload_dof)],
                         np.int64)
    cons_rows, _, _ = constrained_dofs(model)

    def step(prec):
        s = mesh.assemble(E, nu, model.thickness, load_rows, model.cload_value)
        s.apply_dirichlet(cons_rows)
        u, its, res = s.solve(1e-10, 0, prec)
        vm, vn, react, cond = s.recover(cons_rows)
        return s, u, its, res

    arms = {}
    for prec, label in ((Precision.FP64, "fp64"), (Precision.FP32_IR, "fp32_ir")):
        if args.arms not in ("both", label):
            continue
        for _ in range(args.warmup):
            step(prec)
        print(f"[bench] {label}: warmup done", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        for k in range(args.steps):
            s, u, its, res = step(prec)
            print(f"[bench] {label}: step {k + 1}/{args.steps} done", file=sys.stderr, flush=True)
        dt = time.perf_counter() - t0
        ms_kernel, bytes_kernel = s.time_spmv(prec, args.spmv_reps)
        arms[label] = dict(sys=s, u=u, its=its, res=res, dt=dt, dofs_per_s=n_dofs * args.steps / dt,
                           spmv_ms=ms_kernel, spmv_bytes=bytes_kernel)
        print(f"[bench] {label}: {args.steps} solves in {dt:.3f} s, {its} iterations, residual {res:.2e}, "
              f"SpMV {ms_kernel * 1e3:.1f} us ({bytes_kernel / ms_kernel / 1e6:.0f} GB/s)", file=sys.stderr, flush=True)

    if args.arms != "both":
        return
    a, b = arms["fp32_ir"], arms["fp64"]
    err = float(np.abs(a["u"] - b["u"]).max() / np.abs(b["u"]).max())

    def roof(arm, name, label):
        ach = arm["spmv_bytes"] / (arm["spmv_ms"] * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None if (args.plane_size or (label == "fp64" and mf64)) else TRAFFIC_BYTES.get((args.config, label)),
                "kernel": name,
                "bytes_per_launch": arm["spmv_bytes"], "avg_launch_us": round(arm["spmv_ms"] * 1e3, 2)}

    spmv32 = "ir32_spmv_kernel<%d,...>" % ndof
    mf64 = args.operator == "matrix_free" and ndof == 2
    arm_out = {
        "fp64": {"value": round(b["dofs_per_s"], 1), "ms_per_step": round(b["dt"] / args.steps * 1e3, 3),
                 "dtype": "f64", "iterations": int(b["its"]), "residual": b["res"], "operator": args.operator,
                 "max_rel_err_u_vs_fp64": 0.0,
                 "roofline": roof(b, "cg64_mf_spmv_kernel<4>" if mf64 else "cg64_spmv_kernel<%d,...>" % ndof, "fp64")},
        "fp32_ir": {"value": round(a["dofs_per_s"], 1), "ms_per_step": round(a["dt"] / args.steps * 1e3, 3),
                    "dtype": "f32 Krylov + f64 refinement", "iterations": int(a["its"]), "residual": a["res"],
                    "max_rel_err_u_vs_fp64": err, "roofline": roof(a, spmv32, "fp32_ir")},
    }
    # the value is the faster arm (both solve to 1e-10; the FP32-IR arm within 1e-8 of the FP64 one)
    best = "fp32_ir" if a["dofs_per_s"] > b["dofs_per_s"] and err <= 1e-8 else "fp64"
    top = arm_out[best]
    out = {
        "metric": "solved DOF/s at 1e-10 rel. residual: mixed-prec batched vs plain FP64 PCG",
        "value": top["value"],
        "unit": "solved DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": top["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": top["dtype"],
        "data": "synthetic (seeded SplitMix64 deck generator)",
        "config": {"workload": CONFIGS[args.config][0], "n_dof": n_dofs, "nnz": int(mesh.nnz),
                   "tolerance": 1e-10},
        "arm": best,
        "iterations": top["iterations"],
        "residual": top["residual"],
        "roofline": top["roofline"],
        "fp64": arm_out["fp64"],
        "fp32_ir": arm_out["fp32_ir"],
        "speedup_fp32_ir_vs_fp64": round(a["dofs_per_s"] / b["dofs_per_s"], 3),
        "max_rel_err_u_fp32_ir_vs_fp64": err,
    }
    if args.config == "cfg2" and not args.plane_size and args.rotation > 1:
        out["spmv_rotation"] = spmv_rotation(args, mesh, arms, E, nu, model, load_rows, cons_rows)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(model, int(b["its"]), args.cpu_cap)
    print(json.dumps(out), flush=True)


def spmv_rotation(args, mesh, arms, E, nu, model, load_rows, cons_rows):
    """SpMV bandwidth beyond the Infinity Cache: the solver's SpMV launched in
    turn over `args.rotation` distinct 200x20x20 C3D8 matrices (the bench
    deck plus decks of other jitter seeds, each on its own mesh), so each
    launch streams a matrix that the previous launches evicted."""
    from solvix_amd import decks
    from solvix_amd.core import Mesh, System
    from solvix_amd.model import Precision

    extra = []
    for k in range(1, args.rotation):
        m = decks.cfg2(seed=decks.SEED_BASE + 2 + 0x100 * k)
        mk = Mesh.from_model(m, mesh.ctx)
        sk = mk.assemble(E, nu, m.thickness, load_rows, m.cload_value)
        sk.apply_dirichlet(cons_rows)
        extra.append((mk, sk))
    res = {"systems": args.rotation, "reps": args.spmv_reps}
    for prec, label in ((Precision.FP64, "fp64"), (Precision.FP32_IR, "fp32_ir")):
        for _, sk in extra:
            sk.solve(1e-10, 0, prec)
        systems = [arms[label]["sys"]] + [sk for _, sk in extra]
        ms, by = System.time_spmv_rotation(systems, prec, args.spmv_reps)
        ach = by / (ms * 1e-3) / 1e9
        res[label] = {"avg_launch_us": round(ms * 1e3, 2), "bytes_per_launch": by, "achieved": round(ach, 1),
                      "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                      "matrix_bytes_all": round(by * len(systems))}
        print(f"[bench] SpMV rotation {label}: {len(systems)} matrices, {ms * 1e3:.1f} us per launch, "
              f"{ach:.0f} GB/s", file=sys.stderr, flush=True)
    for mk, sk in extra:
        sk.close()
        mk.close()
    return res


if __name__ == "__main__":
    main()
