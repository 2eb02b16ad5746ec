#!/usr/bin/env python3
"""Solved DOF/s at 1e-10 relative residual on one MI355X: the mixed-precision
batched solver (FP32 Krylov iterations, FP64 refinement, residual and final
test) against plain FP64 Jacobi-PCG solving the same instances one by one.

Default workload: 512 uneven plane-stress instances (DOF log-uniform over
2e3..2e5, CPS4 with p = 0.6 else CPS3, left edge clamped, tip load; 22.4 M DOF
in all). A step is one pass over the batch with every instance's mesh
(coordinates, connectivity, CSR pattern) resident in HBM: device assembly of
every K and f, Dirichlet constraints, one batched Krylov engine solving all
instances to 1e-10 (autotuned SpMV launch shape), stress recovery and
reactions, u and the von Mises fields copied back. `value` is the FP32-IR
batched arm with the matrix-free operator in force (FP32 inner iterations on
the assembled FP32 K, FP64 refreshes on the assembled K, the FP64 phase on the
operator); the comparator is plain FP64 PCG one instance after another over
the same 512 instances, timed once in the same run; FP32-IR with the assembled
FP64 phase and the FP64 batched arms (the assembled K, and the matrix-free
operator) are reported beside them.

The roofline is the dominant kernel's: of the value arm's two products (the
batch SpMV of the FP32 inner iteration, the matrix-free product of its FP64
phase), the one with the larger share of the solve, each timed with HIP
events on the solver's stream.
The CPU baseline is the oracle (the restated reference CPU path: linear-scan
lookups, min(nproc, 8) threads spawned per SpMV, serial dots) on a seeded
sample of the instances, which also gives every arm's error against the
oracle. A second leg runs the single 265,923-DOF C3D8 solve (FP64 and FP32-IR)
with one full oracle solve beside it, and the SpMV over distinct matrices
that together exceed the Infinity Cache.

Other workloads (--config): cfg2 (the single C3D8 solve alone), cfg1 (the
CPS4 cantilever), cfg4 (64 load cases, block solve against single solves),
cfg5 (served Poisson stream). Prints one JSON line.
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
METRIC = "solved DOF/s at 1e-10 rel. residual: mixed-prec batched vs plain FP64 PCG"
# HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE counters of rocprofv3,
# each scaled by what the counter reads for a 1 GiB device copy (profiles/);
# None where no counter run exists.
TRAFFIC_BYTES = {("cfg2", "fp64"): 184.9e6, ("cfg2", "fp32_ir"): 96.6e6,
                 ("cfg3", "fp32_ir_batched"): 2797.0e6, ("cfg3", "fp32_ir_batched_mf"): 2797.0e6,  # one FP32 SpMV
                 ("cfg3", "fp64_batched"): 5020.4e6, ("cfg3", "fp64_batched_mf"): 1281.8e6}
CONFIGS = {
    "cfg2": "C3D8 200x20x20 block 20x2x2 m (265,923 DOF), jitter 0.1h, x=0 clamped, -1e6 Pa z-traction on x=20",
    "cfg1": "CPS4 400x40 cantilever 10x1 m (32,882 DOF), jitter 0.15h, -1e4 N tip load",
    "cfg4": ("64 load cases on one C3D8 120x12x12 cantilever (61,347 DOF), 1-8 point loads each; "
             "one block solve vs 64 single solves"),
    "cfg5": ("served Poisson stream: 20% plane instances (DOF 2e3..2e5), 80% C3D8 blocks (DOF 1e4..3e5), "
             "200 GB admission budget"),
    "cfg3": ("512 uneven plane-stress instances, DOF log-uniform 2e3..2e5, CPS4 p=0.6 else CPS3, "
             "left edge clamped, tip load; one batched launch vs one-by-one FP64"),
}


def host_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "host": platform.node() or "host"}


def rel_err(a, b):
    import numpy as np
    a, b = np.asarray(a, float), np.asarray(b, float)
    d = np.abs(b).max() if b.size else 0.0
    return float(np.abs(a - b).max() / d) if d > 0 else float(np.abs(a - b).max(initial=0.0))


def oracle_full(model):
    """One full oracle solve in reference mode (the reference's linear-scan
    lookups and thread policy), timed; also the error reference."""
    import oracle
    t0 = time.perf_counter()
    r = oracle.solve(model, faithful=True, want_csr=False)
    wall = time.perf_counter() - t0
    return r, wall


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


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------- batched
def cpu_sample_indices(n_inst, dof, budget_dof, seed):
    """A seeded shuffle of the instances, taken in that order while they fit
    in `budget_dof` DOF (the oracle's reference-mode lookups grow with the
    square of an instance's size, so the sample is bounded by DOF; an
    instance that would overflow the budget is skipped)."""
    from solvix_amd import decks
    order = list(range(n_inst))
    rng = decks.SplitMix64(seed)
    for i in range(n_inst - 1, 0, -1):  # Fisher-Yates
        j = int(rng.next_u64(1)[0] % (i + 1))
        order[i], order[j] = order[j], order[i]
    out, tot = [], 0
    for i in order:
        if tot + int(dof[i]) > budget_dof:
            continue
        out.append(i)
        tot += int(dof[i])
    return sorted(out)


def bench_batch(args):
    import numpy as np

    from solvix_amd import decks
    from solvix_amd.core import Batch, Context, Mesh, constrained_dofs, element_materials
    from solvix_amd.model import Precision

    models = decks.cfg3()
    t_setup = time.perf_counter()
    ctx = Context(0)
    ctx.set_tuning(ru_delta=args.ru_delta, use_graph=not args.no_graph)
    if args.batch_streams:
        ctx.set_batch_streams(args.batch_streams)
    prep = []
    for m in models:
        E, nu = element_materials(m)
        ndof = m.ndof()
        pos = {int(nid): i for i, nid in enumerate(m.node_ids)}
        lr = np.array([pos[int(nid)] * ndof + int(d) - 1 for nid, d in zip(m.cload_node, m.cload_dof)], np.int64)
        prep.append((Mesh.from_model(m, ctx), E, nu, m.thickness, lr, m.cload_value, constrained_dofs(m)[0]))
    dofs = np.array([m.n_dofs() for m in models], np.int64)
    log(f"{len(models)} instances, {dofs.sum()} DOF, meshes resident ({time.perf_counter() - t_setup:.1f} s)")
    sample = cpu_sample_indices(len(models), dofs, args.cpu_sample_dof, decks.SEED_BASE + 3 ^ 0xC0FFEE)

    def assemble(i):
        mesh, E, nu, t, lr, lv, rows = prep[i]
        sy = mesh.assemble(E, nu, t, lr, lv)
        sy.apply_dirichlet(rows)
        return sy

    phase = {}

    def clock(key, t0):
        phase[key] = phase.get(key, 0.0) + time.perf_counter() - t0

    def sequential():
        us, its, ok, vms, stats = [], [], [], {}, []
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
            vm = sy.recover(prep[i][6])[0]
            clock("recover", t0)
            if i in sample:
                vms[i] = vm
            stats.append(sy.stats())
            us.append(u)
            its.append(it)
            sy.close()
        return us, np.array(its), np.array(ok), vms, stats

    tuned = {}

    def batched(prec, mf=False):
        t0 = time.perf_counter()
        systems = [assemble(i) for i in range(len(prep))]
        clock("assemble", t0)
        t0 = time.perf_counter()
        b = Batch(ctx)
        for sy in systems:
            b.add(sy)
        t = tuned[(prec, mf)]  # the autotuned shape of this (2D-only) batch, node-block size 2
        ctx.set_tuning(unroll=t["unroll"][2], rows_per_tile=t["rows_per_tile"][2], use_graph=not args.no_graph)
        ctx.set_tile_order("xcd" if t["tile_order"][2] else "linear")
        ctx.set_operator("matrix_free" if mf else "assembled")
        try:
            res = b.solve(prec, raise_on_failure=False)
        finally:
            ctx.set_tuning(use_graph=not args.no_graph)
            ctx.set_tile_order("xcd")
            ctx.set_operator("assembled")
        clock("solve", t0)
        t0 = time.perf_counter()
        vms = {}
        for i, sy in enumerate(systems):
            vm = sy.recover(prep[i][6])[0]
            if i in sample:
                vms[i] = vm
        clock("recover", t0)
        us = [sy.solution() for sy in systems]
        its = np.array([r[0] for r in res])
        ok = np.array([r[2] for r in res])
        stats = [sy.stats() for sy in systems]
        return b, systems, us, its, ok, vms, stats

    # FP32-IR with the matrix-free operator in force: FP32 inner iterations on
    # the assembled FP32 K, the FP64 phase (finish, fallback) on the operator
    ARMS = {"fp32_ir_batched_mf": (Precision.FP32_IR, True), "fp32_ir_batched": (Precision.FP32_IR, False),
            "fp64_batched": (Precision.FP64, False), "fp64_batched_mf": (Precision.FP64, True)}
    labels = ["fp32_ir_batched_mf", "fp32_ir_batched", "fp64_sequential", "fp64_batched", "fp64_batched_mf"]
    if args.batch_arms != "all":  # one arm alone
        labels = [args.batch_arms]
    for label in labels:  # autotune once per batched arm (untimed; the shape depends on the batch only)
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
        log(f"autotune {label}: {tuned[(prec, mf)]}")
        b.close()
        for sy in systems:
            sy.close()

    out = {}
    for label in labels:
        # the value arm runs warmup + steps passes; the others one timed pass
        # (plus one warmup pass for the batched ones)
        value_arm = label == labels[0]
        passes = (args.warmup, args.steps) if value_arm else ((0, 1) if label == "fp64_sequential" else (1, 1))
        for k in range(sum(passes)):
            phase.clear()
            t0 = time.perf_counter()
            if label == "fp64_sequential":
                us, its, ok, vms, stats = sequential()
                b = systems = None
            else:
                b, systems, us, its, ok, vms, stats = batched(*ARMS[label])
            dt = time.perf_counter() - t0
            split = " ".join(f"{key} {v:.2f} s" for key, v in phase.items())
            log(f"{label}: pass {k + 1}/{sum(passes)} {dt:.2f} s ({split}), converged {int(ok.sum())}/{ok.size}, "
                f"iterations max {int(its.max())} mean {its.mean():.0f}")
            if k >= passes[0]:
                rec = out.setdefault(label, {"dt": 0.0, "solve_s": 0.0, "passes": 0})
                rec["dt"] += dt
                rec["passes"] += 1
                rec["solve_s"] += phase.get("solve", 0.0)
                rec.update(us=us, its=its, ok=ok, vms=vms, stats=stats)
                if b is not None and k == sum(passes) - 1:
                    ms, by = b.time_spmv(args.spmv_reps)
                    rec.update(spmv_ms=ms, spmv_bytes=by)
                    if ARMS[label] == (Precision.FP32_IR, True):  # the product of its FP64 phase too
                        ctx.set_operator("matrix_free")
                        try:
                            b.prepare(Precision.FP64)
                            ms64, by64 = b.time_spmv(args.spmv_reps)
                        finally:
                            ctx.set_operator("assembled")
                        rec.update(spmv64_ms=ms64, spmv64_bytes=by64)
            if b is not None:
                b.close()
                for sy in systems:
                    sy.close()

    # CPU baseline and error reference: the oracle on the seeded sample
    import oracle
    cpu = {"dof": 0, "s": 0.0, "assemble": 0.0, "pcg": 0.0, "recover": 0.0}
    refs = {}
    if not args.no_cpu_baseline:
        for i in sample:
            r, wall = oracle_full(models[i])
            refs[i] = r
            cpu["dof"] += int(dofs[i]) if r.status == 0 else 0
            cpu["s"] += wall
            for key in ("assemble", "pcg", "recover"):
                cpu[key] += r.times_ms[key] / 1e3
        log(f"oracle on {len(sample)} sampled instances ({cpu['dof']} DOF): {cpu['s']:.1f} s")

    def dps(rec):
        return float(dofs[rec["ok"]].sum()) * rec["passes"] / rec["dt"]

    def oracle_err(rec):
        if not refs:
            return None, None
        eu = max(rel_err(rec["us"][i], refs[i].u) for i in refs)
        ev = max(rel_err(rec["vms"][i], refs[i].von_mises) for i in refs)
        return eu, ev

    def stop_summary(rec):
        st = rec["stats"]
        return {"max_residual": max(s["residual"] for s in st),
                "max_true_residual": max(s["true_residual"] for s in st),
                "max_true_residual_over_floor": max(s["true_residual"] / max(s["true_floor"], 1e-300) for s in st),
                "stopped_on_true_residual": sum(s["converged_by"] == 2 for s in st),
                "stopped_on_recursive_residual": sum(s["converged_by"] == 1 for s in st),
                "fp64_finish_from_floor": sum(s["fallback_reason"] == 2 for s in st),
                "fp64_finish_from_stall": sum(s["fallback_reason"] == 1 for s in st),
                "fp64_iterations": int(sum(s["fp64_iterations"] for s in st))}

    # Algorithmic bytes of a whole batched solve: every instance's iterations
    # times the bytes of one iteration (SpMV with its direction step, update,
    # the per-chunk / per-period direction step), FP32-IR refreshes and FP64
    # iterations counted as such; over the measured solve time this is the
    # bandwidth the whole solve sustains, tails and launches included.
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
        t = rec["solve_s"] / rec["passes"]
        return {"algorithmic_bytes": by, "solve_s": round(t, 3), "GBps": round(by / t / 1e9, 1),
                "frac": round(by / t / 1e9 / HBM_PEAK_GBS, 4)}

    def roof(rec, name, key, which="spmv"):
        ach = rec[which + "_bytes"] / (rec[which + "_ms"] * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": TRAFFIC_BYTES.get(key), "kernel": name,
                "bytes_per_launch": rec[which + "_bytes"], "avg_launch_us": round(rec[which + "_ms"] * 1e3, 2)}

    seq = out.get("fp64_sequential")

    def arm(label, kernel, prec, mf=False):
        rec = out[label]
        eu, ev = oracle_err(rec)
        a = {"value": round(dps(rec), 1), "ms_per_step": round(rec["dt"] / rec["passes"] * 1e3, 3),
             "passes_timed": rec["passes"], "converged": int(rec["ok"].sum()),
             "iterations_max": int(rec["its"].max()), "iterations_mean": round(float(rec["its"].mean()), 1),
             "iterations_dof_weighted": round(float((rec["its"] * dofs).sum() / dofs.sum()), 1),
             "max_rel_err_u_vs_oracle_sample": eu, "max_rel_err_vm_vs_oracle_sample": ev,
             "stopping": stop_summary(rec)}
        if seq is not None and label != "fp64_sequential":
            a["max_rel_err_u_vs_fp64_sequential"] = float(
                max(rel_err(x, c) for x, c in zip(rec["us"], seq["us"])))
            a["speedup_vs_fp64_sequential"] = round(dps(rec) / dps(seq), 3)
        if kernel:
            a["roofline"] = roof(rec, kernel, ("cfg3", label))
            a["whole_solve"] = whole(rec, prec, mf)
        if "spmv64_ms" in rec:
            # FP32 phase and FP64 phase: each product's share of the solve,
            # from the DOF-weighted iterations of each phase times its launch
            # time per DOF; the roofline is the dominant one's
            it64 = np.array([st["fp64_iterations"] for st in rec["stats"]], float)
            it32 = rec["its"] - it64
            share32 = float((it32 * dofs).sum()) * rec["spmv_ms"]
            share64 = float((it64 * dofs).sum()) * rec["spmv64_ms"]
            r64 = roof(rec, "cg64_mf_spmv_kernel<4|3> (whole batch)", ("cfg3", "fp64_batched_mf"), "spmv64")
            a["product_time_share"] = {"fp32_phase": round(share32 / (share32 + share64), 3),
                                       "fp64_phase": round(share64 / (share32 + share64), 3)}
            if share64 > share32:
                a["roofline"], a["roofline_fp32_phase"] = r64, a["roofline"]
            else:
                a["roofline_fp64_phase"] = r64
        return a

    kernels = {"fp32_ir_batched": "ir32_spmv_kernel<2,...> (whole batch)",
               "fp32_ir_batched_mf": "ir32_spmv_kernel<2,...> (whole batch)",
               "fp64_batched": "cg64_spmv_kernel<2,...> (whole batch)",
               "fp64_batched_mf": "cg64_mf_spmv_kernel<4|3> (whole batch)"}
    arms = {}
    for label in labels:
        if label == "fp64_sequential":
            arms[label] = arm(label, None, Precision.FP64)
            arms[label]["timed"] = "once, the same 512 instances one after another"
        else:
            arms[label] = arm(label, kernels[label], *ARMS[label])
    head = arms[labels[0]]
    line = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "solved DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 Krylov + f64 refinement" if labels[0].startswith("fp32_ir") else "f64",
        "data": "synthetic (seeded SplitMix64 deck generator)",
        "config": {"workload": CONFIGS["cfg3"], "instances": len(models), "n_dof_total": int(dofs.sum()),
                   "tolerance": 1e-10},
        "arm": labels[0],
        "converged": head["converged"],
        "max_rel_err_u_vs_oracle_sample": head["max_rel_err_u_vs_oracle_sample"],
        "max_rel_err_vm_vs_oracle_sample": head["max_rel_err_vm_vs_oracle_sample"],
        "roofline": head.get("roofline"),
        "whole_solve": head.get("whole_solve"),
        "speedup_vs_fp64": head.get("speedup_vs_fp64_sequential"),
        "comparator": "fp64_sequential" if seq is not None else None,
        "autotune": {k: tuned[ARMS[k]] for k in ARMS if ARMS[k] in tuned},
        "arms": arms,
    }
    if refs:
        line["cpu_baseline"] = {
            "value": round(cpu["dof"] / cpu["s"], 1), "unit": "solved DOF/s", "cores": oracle.worker_threads(),
            "kind": "port",
            "sample": (f"oracle in reference mode on {len(sample)} seeded instances of the 512 ({cpu['dof']} DOF: "
                       f"{', '.join(str(i) for i in sample)}), full solves: assembly {cpu['assemble']:.1f} s, "
                       f"PCG {cpu['pcg']:.1f} s, recovery {cpu['recover']:.2f} s, wall {cpu['s']:.1f} s"),
            **host_info()}
    ctx.close()
    if args.second_leg and args.batch_arms == "all":
        log("second leg: the single 265,923-DOF C3D8 solve")
        line["second_leg"] = bench_single(args, "cfg2", emit=False)
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------- single
def build_model(name, size=None):
    """cfg2 / cfg1, or the same block / cantilever at another mesh size
    (`size` = (nx, ny, nz) or (nx, ny), the element size kept)."""
    from solvix_amd import decks
    if name == "cfg2":
        return decks.cfg2(nx=size[0], ny=size[1], nz=size[2]) if size else decks.cfg2()
    return decks.cfg1(nx=size[0], ny=size[1]) if size else decks.cfg1()


def custom_size(args, config):
    spec = args.solid_size if config == "cfg2" else args.plane_size
    return tuple(int(v) for v in spec.split("x")) if spec else None


def bench_single(args, config, emit=True):
    """One large solve: FP64 and FP32-IR arms (warmup + steps full solves
    each) beside one full oracle solve (the CPU baseline and the error
    reference)."""
    import numpy as np

    from solvix_amd.core import Context, Mesh, constrained_dofs, element_materials
    from solvix_amd.model import Precision

    size = custom_size(args, config)
    model = build_model(config, size)
    n_dofs = model.n_dofs()
    ctx = Context(0)
    ctx.set_operator(args.operator)
    if args.no_graph:
        ctx.set_tuning(use_graph=False)
    mesh = Mesh.from_model(model, ctx)
    E, nu = element_materials(model)
    ndof = model.ndof()
    pos = {int(n): i for i, n in enumerate(model.node_ids)}
    load_rows = np.array([pos[int(n)] * ndof + int(d) - 1 for n, d in zip(model.cload_node, model.cload_dof)],
                         np.int64)
    cons_rows, _, _ = constrained_dofs(model)
    steps, warmup = (args.steps, args.warmup) if emit else (args.leg_steps, 1)

    def step(prec):
        s = mesh.assemble(E, nu, model.thickness, load_rows, model.cload_value)
        s.apply_dirichlet(cons_rows)
        u, its, res = s.solve(1e-10, 0, prec)
        vm, vn, react, cond = s.recover(cons_rows)
        return s, u, its, res, vm

    arms = {}
    for prec, label in ((Precision.FP64, "fp64"), (Precision.FP32_IR, "fp32_ir")):
        if args.arms not in ("both", label):
            continue
        for _ in range(warmup):
            step(prec)
        t0 = time.perf_counter()
        for k in range(steps):
            s, u, its, res, vm = step(prec)
        dt = time.perf_counter() - t0
        ms_kernel, bytes_kernel = s.time_spmv(prec, args.spmv_reps)
        arms[label] = dict(sys=s, u=u, vm=vm, its=its, res=res, dt=dt, dofs_per_s=n_dofs * steps / dt,
                           spmv_ms=ms_kernel, spmv_bytes=bytes_kernel, stats=s.stats())
        log(f"{config} {label}: {steps} solves in {dt:.3f} s, {its} iterations, residual {res:.2e}, "
            f"SpMV {ms_kernel * 1e3:.1f} us ({bytes_kernel / ms_kernel / 1e6:.0f} GB/s)")
    if args.arms != "both":
        return None
    ref = wall = None
    if not args.no_cpu_baseline:
        ref, wall = oracle_full(model)
        log(f"{config} oracle: {wall:.1f} s ({ref.iterations} iterations)")
    mf64 = args.operator == "matrix_free" and ndof == 2

    def roof(a, name, label):
        ach = a["spmv_bytes"] / (a["spmv_ms"] * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": None if (size or (label == "fp64" and mf64)) else TRAFFIC_BYTES.get((config, label)),
                "kernel": name, "bytes_per_launch": a["spmv_bytes"], "avg_launch_us": round(a["spmv_ms"] * 1e3, 2)}

    def arm(a, label, kernel, dtype):
        st = a["stats"]
        o = {"value": round(a["dofs_per_s"], 1), "ms_per_step": round(a["dt"] / steps * 1e3, 3), "dtype": dtype,
             "iterations": int(a["its"]), "fp64_iterations": int(st["fp64_iterations"]), "residual": a["res"],
             "stopped_on": {1: "recursive FP64 residual", 2: "true FP64 residual"}.get(st["converged_by"], "-"),
             "fp64_finish": {0: None, 1: "stalled above the floor", 2: "at the FP64 evaluation floor"}[st["fallback_reason"]],
             "true_residual": st["true_residual"], "true_floor": st["true_floor"],
             "roofline": roof(a, kernel, label)}
        if ref is not None:
            o["max_rel_err_u_vs_oracle"] = rel_err(a["u"], ref.u)
            o["max_rel_err_vm_vs_oracle"] = rel_err(a["vm"], ref.von_mises)
        return o

    a64, a32 = arms["fp64"], arms["fp32_ir"]
    arm_out = {"fp64": arm(a64, "fp64", "cg64_mf_spmv_kernel<4>" if mf64 else f"cg64_spmv_kernel<{ndof},...>", "f64"),
               "fp32_ir": arm(a32, "fp32_ir", f"ir32_spmv_kernel<{ndof},...>", "f32 Krylov + f64 refinement")}
    out = {
        "metric": METRIC,
        "value": arm_out["fp32_ir"]["value"],
        "unit": "solved DOF/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": arm_out["fp32_ir"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 Krylov + f64 refinement",
        "data": "synthetic (seeded SplitMix64 deck generator)",
        "config": {"workload": CONFIGS[config] if not size else
                   (f"C3D8 block {args.solid_size}" if config == "cfg2" else f"CPS4 cantilever {args.plane_size}"),
                   "n_dof": n_dofs, "nnz": int(mesh.nnz), "tolerance": 1e-10},
        "arm": "fp32_ir",
        "roofline": arm_out["fp32_ir"]["roofline"],
        "comparator": "fp64",
        "speedup_fp32_ir_vs_fp64": round(a32["dofs_per_s"] / a64["dofs_per_s"], 3),
        "max_rel_err_u_fp32_ir_vs_fp64": rel_err(a32["u"], a64["u"]),
        "fp64": arm_out["fp64"],
        "fp32_ir": arm_out["fp32_ir"],
    }
    if ref is not None:
        out["cpu_baseline"] = {
            "value": round(n_dofs / wall, 1), "unit": "solved DOF/s", "cores": __import__("oracle").worker_threads(),
            "kind": "port",
            "sample": (f"one full oracle solve in reference mode ({ref.iterations} PCG iterations): assembly "
                       f"{ref.times_ms['assemble'] / 1e3:.2f} s, PCG {ref.times_ms['pcg'] / 1e3:.2f} s, recovery "
                       f"{ref.times_ms['recover'] / 1e3:.2f} s, wall {wall:.1f} s"),
            **host_info()}
    if config == "cfg2" and not size and args.rotation > 1:
        out["spmv_rotation"] = spmv_rotation(args, mesh, arms, E, nu, model, load_rows, cons_rows)
    for a in arms.values():
        a["sys"].close()
    mesh.close()
    ctx.close()
    if emit:
        print(json.dumps(out), flush=True)
    return out


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
    ctx = mesh.ctx
    # both load hints in both precisions; the headline per precision is the
    # faster hint on these launches: nontemporal matrix loads for FP64 (the
    # matrix no longer evicts the gathered vector from L2), ordinary loads for
    # FP32 (its 94 MB launches keep the vector in L2 either way, and the
    # nontemporal path costs more than it saves); the other one beside it
    runs = (("fp64", Precision.FP64, "streaming"), ("fp64", Precision.FP64, "cached"),
            ("fp32_ir", Precision.FP32_IR, "cached"), ("fp32_ir", Precision.FP32_IR, "streaming"))
    for label, prec, hint in runs:
        ctx.set_load_hint(hint)
        systems = [arms[label]["sys"]] + [sk for _, sk in extra]
        for sk in systems:
            sk.prepare_solver(prec)
        ms, by = System.time_spmv_rotation(systems, prec, args.spmv_reps)
        ach = by / (ms * 1e-3) / 1e9
        rec = {"avg_launch_us": round(ms * 1e3, 2), "bytes_per_launch": by, "achieved": round(ach, 1),
               "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
               "matrix_bytes_all": round(by * len(systems))}
        if label in res:
            res[label]["cached_loads" if hint == "cached" else "nontemporal_loads"] = {
                k: rec[k] for k in ("avg_launch_us", "achieved", "frac")}
        else:
            res[label] = dict(rec, load_hint=hint)
        log(f"SpMV rotation {label}{' ' + hint if hint else ''}: {len(systems)} matrices, {ms * 1e3:.1f} us per "
            f"launch, {ach:.0f} GB/s")
    ctx.set_load_hint("auto")
    for mk, sk in extra:
        sk.close()
        mk.close()
    return res


# ---------------------------------------------------------------- multi-RHS
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
    if args.multi_group:
        ctx.set_multi_group(args.multi_group)
    if args.no_graph:
        ctx.set_tuning(use_graph=False)
    mesh, sy = assemble(model, ctx)
    sy.apply_dirichlet(constrained_dofs(model)[0])
    n = mesh.n
    pos = {int(nid): i for i, nid in enumerate(model.node_ids)}
    F = np.zeros((len(cases), n))
    for k, (nodes, dofs, vals) in enumerate(cases):
        for nid, d, v in zip(nodes, dofs, vals):
            F[k, pos[int(nid)] * 3 + int(d) - 1] += v
    out = {}
    if args.multi_arm == "spmv":  # the block SpMV alone (profiler runs: no solve, no graphs)
        us, by = sy.multi_spmv_timing(len(cases), args.spmv_reps)
        print(json.dumps({"block_spmv_us": round(us, 2), "bytes_per_launch": by,
                          "GBps": round(by / us / 1e3, 1)}), flush=True)
        return
    arms = [a for a in ("single", "block", "block_fp32_ir") if args.multi_arm in ("both", a)]
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
                prec = Precision.FP64 if arm == "block" else Precision.FP32_IR
                U, its, res, ok, dev_ms = sy.solve_multi(F, 1e-10, 0, prec, raise_on_failure=False)
            dt = time.perf_counter() - t0
            log(f"{arm}: pass {k + 1}/{args.warmup + args.steps} {dt:.3f} s, converged {int(ok.sum())}/{ok.size}, "
                f"iterations max {int(its.max())} mean {its.mean():.0f}")
            if k >= args.warmup:
                rec = out.setdefault(arm, {"dt": 0.0})
                rec["dt"] += dt
                rec.update(U=U, its=its, ok=ok)
    # the block FP64 solve's SpMV (direction step fused in), timed alone in
    # the solve's launch shape
    spmv = None
    if "block" in arms:
        us, by = sy.multi_spmv_timing(len(cases), args.spmv_reps)
        grouped = args.multi_group >= 0
        spmv = {"bound": "hbm", "achieved": round(by / us / 1e3, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(by / us / 1e3 / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": f"blk_spmv_group_kernel<3,{args.multi_group or 4}>" if grouped else "blk_spmv_node_kernel<3>",
                "bytes_per_launch": by, "avg_launch_us": round(us, 2)}
        log(f"block SpMV: {us:.2f} us per launch, {by / 1e6:.1f} MB algorithmic, {by / us / 1e3:.1f} GB/s")
    if len(arms) < 3:
        if spmv:
            print(json.dumps({"block_spmv": spmv, "block_ms": round(out["block"]["dt"] / args.steps * 1e3, 3)}),
                  flush=True)
        return
    mixed, blk, b = out["block_fp32_ir"], out["block"], out["single"]
    dof = float(n) * len(cases) * args.steps
    # the oracle (reference mode, full solve per case) on a fixed sample of the
    # load cases: the per-case error reference and the CPU baseline
    ref, cpu_s = {}, 0.0
    if not args.no_cpu_baseline:
        import dataclasses
        for j in range(0, len(cases), max(1, len(cases) // 16)):
            nodes, dofs, vals = cases[j]
            r, wall = oracle_full(dataclasses.replace(model, cload_node=nodes, cload_dof=dofs, cload_value=vals))
            ref[j] = r.u
            cpu_s += wall
        log(f"oracle on load cases {sorted(ref)}: {cpu_s:.1f} s")

    def err(a):
        return float(max(rel_err(a["U"][j], b["U"][j]) for j in range(len(cases))))

    def arm(a):
        o = {"value": round(dof / a["dt"], 1), "ms_per_step": round(a["dt"] / args.steps * 1e3, 3),
             "iterations_max": int(a["its"].max()), "iterations_mean": round(float(a["its"].mean()), 1),
             "converged": int(a["ok"].sum()), "max_rel_err_u_vs_single_fp64": err(a),
             "speedup_vs_single_fp64": round(b["dt"] / a["dt"], 3)}
        if ref:
            o["max_rel_err_u_vs_oracle_sample"] = float(max(rel_err(a["U"][j], u) for j, u in ref.items()))
        return o
    line = {
        "metric": METRIC, "value": round(dof / mixed["dt"], 1), "unit": "solved DOF/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(mixed["dt"] / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 Krylov + f64 refinement",
        "data": "synthetic (seeded SplitMix64 deck generator)",
        "config": {"workload": CONFIGS["cfg4"], "n_dof": n, "nnz": int(mesh.nnz), "load_cases": len(cases),
                   "tolerance": 1e-10},
        "arm": "block_fp32_ir", "comparator": "single_fp64",
        "speedup_vs_fp64": round(b["dt"] / mixed["dt"], 3),
        "block_fp32_ir": arm(mixed), "block_fp64": dict(arm(blk), roofline=spmv),
        "single_fp64": {"value": round(dof / b["dt"], 1), "ms_per_step": round(b["dt"] / args.steps * 1e3, 3),
                        "iterations_max": int(b["its"].max()), "iterations_mean": round(float(b["its"].mean()), 1)},
    }
    if ref:
        line["single_fp64"]["max_rel_err_u_vs_oracle_sample"] = float(
            max(rel_err(b["U"][j], u) for j, u in ref.items()))
        line["cpu_baseline"] = {
            "value": round(n * len(ref) / cpu_s, 1), "unit": "solved DOF/s",
            "cores": __import__("oracle").worker_threads(), "kind": "port",
            "sample": f"full oracle solves in reference mode of load cases {sorted(ref)}, {cpu_s:.1f} s",
            **host_info()}
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------- served
def bench_serve(args):
    """Open-loop served stream: requests are submitted at their Poisson arrival
    times from a client thread while the server solves; latency is submit to
    result. Requests arriving in the first `lead` seconds are served but not
    measured. Solved DOF/s counts the measured requests' DOF over the span
    from the first measured submission to the last measured completion. The
    samples (oracle errors, plain FP64 one by one) are chosen before the run;
    every other request's deck is dropped once the server has copied it."""
    import threading

    import numpy as np

    from solvix_amd import decks
    from solvix_amd.core import Context, Server, solve_with_method
    from solvix_amd.model import Method, Precision

    t_gen = time.perf_counter()
    stream = list(decks.cfg5_stream(duration_s=args.lead + args.duration, rate=args.rate))
    dofs = [r.model.n_dofs() for r in stream]
    offered = sum(dofs) / (args.lead + args.duration)
    measured = [i for i, r in enumerate(stream) if r.arrival_s >= args.lead]
    # per-request error against the FP64 oracle: a seeded sample of the
    # measured requests bounded by DOF; plain FP64 one by one: every k-th
    rng = decks.SplitMix64(decks.SEED_BASE + 5 ^ 0xC0FFEE)
    order = sorted(measured, key=lambda k: int(rng.next_u64(1)[0]))
    oracle_ids, tot = [], 0
    for i in order:
        if tot + dofs[i] <= args.serve_oracle_dof:
            oracle_ids.append(i)
            tot += dofs[i]
    plain_ids = measured[::max(1, len(measured) // 60)]
    keep = set(oracle_ids) | set(plain_ids)
    log(f"{len(stream)} requests generated ({args.rate}/s over {args.lead + args.duration} s, offered "
        f"{offered / 1e6:.2f} M DOF/s) in {time.perf_counter() - t_gen:.1f} s")
    ctx = Context(0)
    prec = Precision.FP64 if args.serve_precision == "fp64" else Precision.FP32_IR
    ctx.set_operator(args.serve_operator)  # the server keeps the operator in force at its creation
    srv = Server(ctx, prec, hbm_budget=200 * 10 ** 9, max_active=1024, use_graph=args.serve_graph == "on")
    ctx.set_operator("assembled")  # the FP64 references below multiply the assembled K
    subs = [None] * len(stream)
    models = {i: stream[i].model for i in keep}
    t0 = time.perf_counter()

    def client():
        for i, req in enumerate(stream):
            delay = t0 + req.arrival_s - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            subs[i] = (time.perf_counter(), srv.submit(req.model))
            req.model = None  # the server holds its own copy

    th = threading.Thread(target=client)
    th.start()
    done = {}
    sols = {}
    last_print = time.perf_counter()
    for i in range(len(stream)):
        while subs[i] is None:
            time.sleep(0.001)
        t_sub, ticket = subs[i]
        sol, info = srv.wait(ticket)
        done[i] = (t_sub - t0, time.perf_counter() - t0, info["latency_ms"])
        if i in oracle_ids:
            sols[i] = sol
        if time.perf_counter() - last_print > 5.0:
            st = srv.stats()
            log(f"{i + 1}/{len(stream)} done at {time.perf_counter() - t0:.1f} s, peak active {st['peak_active']}, "
                f"queued {st['queued']}, resident {st['resident_bytes'] / 1e9:.1f} GB")
            last_print = time.perf_counter()
    th.join()
    st = srv.stats()
    srv.close()
    lat = np.array([done[i][2] for i in measured])
    span = max(done[i][1] for i in measured) - min(done[i][0] for i in measured)
    dof = float(sum(dofs[i] for i in measured))
    import oracle
    errs = []
    for i in oracle_ids:
        r = oracle.solve(models[i], want_csr=False)
        errs.append((rel_err(sols[i].displacements, r.displacements), rel_err(sols[i].von_mises, r.von_mises)))
    t1 = time.perf_counter()
    for i in plain_ids:
        solve_with_method(models[i], Method.AUTO, Precision.FP64, ctx)
    plain_dps = float(sum(dofs[i] for i in plain_ids)) / (time.perf_counter() - t1)
    log(f"plain FP64 one by one on {len(plain_ids)} sampled requests: {plain_dps / 1e6:.2f} M DOF/s")
    line = {
        "metric": METRIC, "value": round(dof / span, 1), "unit": "solved DOF/s", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": round(span * 1e3, 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if prec == Precision.FP64 else "f32 Krylov + f64 refinement",
        "data": "synthetic (seeded SplitMix64 deck generator)",
        "config": {"workload": CONFIGS["cfg5"], "precision": args.serve_precision,
                   "operator": args.serve_operator,  # FP32-IR: the operator of its FP64 phase
                   "rate_per_s": args.rate, "measured_s": args.duration, "graphs": args.serve_graph,
                   "lead_s": args.lead, "requests": len(stream), "measured_requests": len(measured),
                   "hbm_budget_bytes": 200 * 10 ** 9, "offered_dof_per_s": round(offered, 1)},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "max": float(lat.max())},
        "server": st,
        "max_rel_err_u_vs_oracle_sample": max(e[0] for e in errs) if errs else None,
        "max_rel_err_vm_vs_oracle_sample": max(e[1] for e in errs) if errs else None,
        "oracle_sample": {"requests": len(errs), "dof": tot},
        "fp64_sequential_sample": {"value": round(plain_dps, 1), "requests": len(plain_ids)},
        "speedup_vs_fp64": round(dof / span / plain_dps, 3),
    }
    print(json.dumps(line), flush=True)


def profile_spmv(args):
    """The roofline kernel alone, for profiler runs: the solver set-up of the
    workload without a solve (svx_batch_prepare / svx_system_prepare_solver:
    the same tile tables and launch shape as the solve), then the SpMV
    launched args.spmv_reps times on the solver's stream and timed with HIP
    events, as the bench's roofline does after its solves. No graph is
    captured in this mode."""
    import numpy as np

    from solvix_amd import decks
    from solvix_amd.core import Batch, Context, assemble, constrained_dofs
    from solvix_amd.model import Precision

    prec = Precision.FP64 if args.profile_spmv == "fp64" else Precision.FP32_IR
    ctx = Context(0)
    ctx.set_operator(args.operator)
    ctx.set_load_hint(args.load_hint)
    keep = []
    if args.config == "cfg3":
        b = Batch(ctx)
        for m in decks.cfg3():
            mesh, sy = assemble(m, ctx)
            sy.apply_dirichlet(constrained_dofs(m)[0])
            keep.append((mesh, sy))
            b.add(sy)
        if args.profile_shape:  # the autotuned shape of a bench run: no tuning launches here
            u, r, *x = (int(v) for v in args.profile_shape.split(","))
            tuned = {"unroll": [0, 0, u, 0], "rows_per_tile": [0, 0, r, 0], "tile_order": [1, 1, x[0] if x else 1, 1],
                     "spmv_ms": None}
        else:
            tuned = b.autotune(prec, 3)
        ctx.set_tuning(unroll=tuned["unroll"][2], rows_per_tile=tuned["rows_per_tile"][2])
        ctx.set_tile_order("xcd" if tuned["tile_order"][2] else "linear")
        b.prepare(prec)
        ms, by = b.time_spmv(args.spmv_reps)
        if args.profile_orders:  # the other tile order in the same process, after the first
            ctx.set_tile_order("linear" if tuned["tile_order"][2] else "xcd")
            b.prepare(prec)
            ms2, _ = b.time_spmv(args.spmv_reps)
            extra_orders = {"other_tile_order_avg_launch_us": round(ms2 * 1e3, 2)}
        else:
            extra_orders = {}
        kernel = ("cg64_mf_spmv_kernel<4|3>" if args.operator == "matrix_free" else
                  ("cg64_spmv_kernel<2,%d>" if prec == Precision.FP64 else "ir32_spmv_kernel<2,%d>") % tuned["unroll"][2])
        extra = {"instances": len(keep), "autotune": tuned, "load_hint": args.load_hint, **extra_orders}
    else:
        model = build_model(args.config, custom_size(args, args.config))
        mesh, sy = assemble(model, ctx)
        sy.apply_dirichlet(constrained_dofs(model)[0])
        keep.append((mesh, sy))
        sy.prepare_solver(prec)
        ms, by = sy.time_spmv(prec, args.spmv_reps)
        extra = {}
        if args.profile_orders:  # the linear tile order in the same process
            ctx.set_tile_order("linear")
            sy.prepare_solver(prec)
            extra["other_tile_order_avg_launch_us"] = round(sy.time_spmv(prec, args.spmv_reps)[0] * 1e3, 2)
        kernel = ("cg64_spmv_kernel<%d,...>" if prec == Precision.FP64 else "ir32_spmv_kernel<%d,...>") % model.ndof()
        extra["n_dof"] = model.n_dofs()
        extra["load_hint"] = args.load_hint
        if custom_size(args, args.config):
            extra["mesh_size"] = args.solid_size if args.config == "cfg2" else args.plane_size
    ach = by / (ms * 1e-3) / 1e9
    print(json.dumps({"profile_spmv": args.profile_spmv, "config": args.config, "operator": args.operator,
                      "kernel": kernel, "launches": args.spmv_reps + 1, "avg_launch_us": round(ms * 1e3, 2),
                      "bytes_per_launch": by, "achieved_GBps": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4),
                      **extra}), flush=True)
    for mesh, sy in keep:
        sy.close()
        mesh.close()
    ctx.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="cfg3", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-dof", type=int, default=150000,
                    help="DOF in the seeded instance sample the oracle solves (CPU baseline and error reference)")
    ap.add_argument("--no-second-leg", dest="second_leg", action="store_false",
                    help="batched workload: skip the single C3D8 solve leg")
    ap.add_argument("--leg-steps", type=int, default=3, help="second leg: timed solves per arm")
    ap.add_argument("--arms", default="both", choices=["both", "fp64", "fp32_ir"],
                    help="single solves: run one precision only (profiling); the JSON line needs both")
    ap.add_argument("--spmv-reps", type=int, default=100, help="launches of the timed SpMV kernel")
    ap.add_argument("--rotation", type=int, default=3,
                    help="single C3D8 solve: distinct matrices the SpMV rotation leg cycles through (0: skip)")
    ap.add_argument("--batch-arms", default="all",
                    choices=["all", "fp32_ir_batched_mf", "fp32_ir_batched", "fp64_batched", "fp64_batched_mf",
                             "fp64_sequential"],
                    help="batched workload: every arm, or one alone")
    ap.add_argument("--operator", default="assembled", choices=["assembled", "matrix_free"],
                    help="single solves: operator of the FP64 arm (matrix_free: plane meshes)")
    ap.add_argument("--solid-size", default=None, metavar="NXxNYxNZ",
                    help="cfg2 (and its SpMV profile) on a C3D8 block of this mesh size, same element size")
    ap.add_argument("--plane-size", default=None, metavar="NXxNY",
                    help="cfg1 only: the cantilever on an NX x NY mesh (profiling at other sizes)")
    ap.add_argument("--batch-streams", type=int, default=0,
                    help="batched workload: engines run side by side on their own streams (0: library default)")
    ap.add_argument("--no-graph", action="store_true",
                    help="solves enqueue each chunk kernel by kernel instead of as a captured graph "
                         "(profiler runs)")
    ap.add_argument("--multi-group", type=int, default=0, choices=[-1, 0, 4, 8],
                    help="multi-RHS workload: nodes per group of the block SpMV (0: default, -1: no groups)")
    ap.add_argument("--multi-arm", default="both", choices=["both", "single", "block", "block_fp32_ir", "spmv"],
                    help="multi-RHS workload: run one arm only (profiling)")
    ap.add_argument("--rate", type=float, default=200.0, help="served stream: arrivals per second")
    ap.add_argument("--serve-oracle-dof", type=int, default=1000000,
                    help="served stream: DOF budget of the seeded request sample checked against the oracle")
    ap.add_argument("--serve-operator", default="matrix_free", choices=["assembled", "matrix_free"],
                    help="served stream: operator of the server's FP64 iterations")
    ap.add_argument("--serve-precision", default="fp64", choices=["fp32_ir", "fp64"],
                    help="served stream: precision of the server's solves")
    ap.add_argument("--serve-graph", default="off", choices=["on", "off"],
                    help="served stream: run the server's chunks as captured hipGraphs (on) or kernel by kernel")
    ap.add_argument("--duration", type=float, default=60.0, help="served stream: measured seconds")
    ap.add_argument("--lead", type=float, default=10.0, help="served stream: unmeasured seconds before")
    ap.add_argument("--profile-spmv", default=None, choices=["fp64", "fp32_ir"],
                    help="cfg3 / cfg2: time the solver's SpMV of this precision alone after the solver set-up, "
                         "without solving (profiler runs)")
    ap.add_argument("--profile-shape", default=None, metavar="UNROLL,ROWS[,XCD]",
                    help="--profile-spmv on cfg3: the plane launch shape (blocks in flight, rows per tile, tile "
                         "order 1 = XCD / 0 = linear) instead of autotuning")
    ap.add_argument("--ru-delta", type=float, default=0.0,
                    help="FP32-IR: refresh once the recursive residual falls below this fraction of its maximum "
                         "since the last refresh (0: the library default, 0.1)")
    ap.add_argument("--load-hint", default="auto", choices=["auto", "cached", "streaming"],
                    help="SpMV profile: load hint of the FP64 ELL product (svx_set_load_hint)")
    ap.add_argument("--profile-orders", action="store_true",
                    help="--profile-spmv: also time the other tile order in the same process")
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
    if args.profile_spmv:
        profile_spmv(args)
    elif args.config == "cfg3":
        bench_batch(args)
    elif args.config == "cfg5":
        bench_serve(args)
    elif args.config == "cfg4":
        bench_multi(args)
    else:
        bench_single(args, args.config)


if __name__ == "__main__":
    main()
