"""Airfoil benchmark: steps/s at 720k cells, tuned + graphed vs plain HIP.

One step is one airfoil time step (save_soln, then twice adt_calc, res_calc,
bres_calc, update with the rms reduction) over the whole 1200x600 mesh
(720,000 cells, 1,438,200 edges), FP64, with all data resident in HBM.

  plain    the reference's execution model on this backend: LDS-staged,
           two-level-coloured indirect loops, device sync and gbl read-back
           after every loop;
  reference design, best
           the same execution model with the fastest of the reference's own
           schemes (staged or flat-coloured, any block size) per indirect loop;
  tuned    per-loop plan and block size picked by timing every variant on the
           same mesh (exhaustive sweep), run eagerly (sync where a gbl result
           is needed);
  graphed  the tuned plans with each time step replayed from one HIP graph,
           rms/maxerr/errloc read back from a device history at the end.

Two input orderings are measured (both named by the workload): the headline
uses the RCM-renumbered mesh; the randomly numbered mesh is reported beside
it, as given, after the runtime's op_renumber, and with the device alone
working in that numbering (OP_HIP_RENUMBER; the application keeps its own).  Parity: every run's rms at
every one of the W + K iterations (warm-up, eager and graph-replayed alike)
and its final q against the oracle's OpenMP run of the same mesh, relative
error at most 1e-10; otherwise the line carries value null and the exit
status is 1.
"""
import argparse
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "op2-common_amd"))
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "airfoil steps/s, 720k cells: tuned/graphed vs plain HIP at rms rel err ≤1e-10"
PARITY_TOL = 1e-10  # rms (every iteration) and final q, relative, vs the oracle
PLANS = ("atomic", "color", "staged", "gather", "owner")


def profiled_traffic(loop_plan, workload):
    """HBM bytes per launch of a loop's kernel, from the committed rocprofv3
    counter summary for that (loop/plan/block, mesh/ordering), if there is one:
    a line 'traffic: <loop/plan/block> <mesh/ordering> <bytes>' in a text file
    under profiles/.  The summary records the FETCH_SIZE/WRITE_SIZE runs and
    the copy-kernel scale factors the bytes were derived with."""
    prof = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")
    if not os.path.isdir(prof):
        return None
    for name in sorted(os.listdir(prof)):
        if not name.endswith((".txt", ".md")):
            continue
        with open(os.path.join(prof, name)) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 4 and parts[0] == "traffic:" and parts[1] == loop_plan \
                        and parts[2] == workload:
                    return float(parts[3])
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


STREAM_VARIANTS = ("copy, one 16-byte vector per thread", "copy, four per thread",
                   "copy, four per thread, non-temporal", "copy, eight per thread",
                   "copy, persistent grid striding four per iteration",
                   "read only, four per thread", "write only, four per thread")


def stream_ceiling(op2amd, size=1 << 30, reps=20):
    """The achievable HBM streaming rate on this device: copy (and read-only,
    write-only) kernels of several shapes over 1 GiB buffers, the fastest of
    `reps` calls each (op2h_stream_probe).  Returns (best copy GB/s, all)."""
    import ctypes as C
    lib = op2amd._lib.lib
    lib.op2h_stream_probe.restype = C.c_int
    lib.op2h_stream_probe.argtypes = [C.c_size_t, C.c_int, C.c_int, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]
    res = {}
    for v, name in enumerate(STREAM_VARIANTS):
        g, us = C.c_double(), C.c_double()
        if lib.op2h_stream_probe(size, v, reps, C.byref(g), C.byref(us)) == 0:
            res[name] = round(g.value, 1)
    best = max(v for k, v in res.items() if k.startswith("copy"))
    return best, res


def with_fractions(per_loop, ceiling):
    for v in per_loop.values():
        v["frac_of_peak"] = round(v["achieved_GBs"] / HBM_PEAK_GBS, 4)
        v["frac_of_stream_ceiling"] = round(v["achieved_GBs"] / ceiling, 4)
    return per_loop


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.abs(b)))


def sweep(op2amd, mesh, iters, renumber):
    """Exhaustive sweep: per indirect loop, every plan x block size, timed
    per call with HIP events; then the block size of the direct loops.
    Returns the fastest plans, and the fastest among the reference's own
    schemes (LDS-staged two-level colouring, flat colouring) per indirect
    loop under "_reference_design"."""
    best, ref_best = {}, {}
    for loop in ("res_calc", "bres_calc"):
        times = {}
        for plan in PLANS:
            for bs in (128, 256, 512):
                kw = {"plan_res": "atomic", "plan_bres": "atomic"}
                kw["plan_res" if loop == "res_calc" else "plan_bres"] = plan
                kw["block_res" if loop == "res_calc" else "block_bres"] = bs
                r = op2amd.airfoil.run(mesh, iters, mode="tuned", warmup=2,
                                       kernel_timing=True, renumber=renumber, **kw)
                times[(plan, bs)] = r["loop_kernel_ms"][loop] / max(r["loop_calls"][loop], 1)
        best[loop] = min(times, key=times.get)
        ref_best[loop] = min((k for k in times if k[0] in ("staged", "color")), key=times.get)
        log(f"  sweep {loop}: " + ", ".join(
            f"{p}/{b}={t * 1e3:.1f}us" for (p, b), t in sorted(times.items(), key=lambda kv: kv[1])))
    direct = {}
    for bs in (128, 256, 512, 1024):
        r = op2amd.airfoil.run(mesh, iters, mode="tuned", warmup=2, kernel_timing=True,
                               renumber=renumber, plan_res=best["res_calc"][0],
                               block_res=best["res_calc"][1], block_direct=bs, block_adt=bs)
        direct[bs] = sum(r["loop_kernel_ms"][lp] / max(r["loop_calls"][lp], 1)
                         for lp in ("save_soln", "adt_calc", "update"))
    bd = min(direct, key=direct.get)
    return dict(plan_res=best["res_calc"][0], block_res=best["res_calc"][1],
                plan_bres=best["bres_calc"][0], block_bres=best["bres_calc"][1],
                block_direct=bd, block_adt=bd,
                _reference_design=dict(plan_res=ref_best["res_calc"][0],
                                       block_res=ref_best["res_calc"][1],
                                       plan_bres=ref_best["bres_calc"][0],
                                       block_bres=ref_best["bres_calc"][1]))


def parse_plans(text):
    # "res=color:256,bres=atomic:128,direct=1024"
    kw = {}
    for part in text.split(","):
        k, v = part.split("=")
        if k in ("res", "bres"):
            p, b = v.split(":")
            kw[f"plan_{k}"] = p
            kw[f"block_{k}"] = int(b)
        elif k == "direct":
            kw["block_direct"] = kw["block_adt"] = int(v)
    return kw


def measure(op2amd, mesh, K, W, args, renumber=False, plans=None, plain_too=True):
    """Step rates of the plain backend, the best of the reference's own
    schemes with the reference's per-loop syncs, the tuned plans run eagerly,
    and the tuned plans replayed from graphs, with and without loop-chain
    fusion; all over the same W + K iterations, whose trajectories are
    returned for the parity check."""
    n = K + W
    out, runs = {}, {}
    if plain_too:
        plain = op2amd.airfoil.run(mesh, n, mode="plain", warmup=W, renumber=renumber)
        out["plain_steps_per_s"] = K / plain["seconds"]
        runs["plain"] = plain
        log(f"  plain: {out['plain_steps_per_s']:.1f} steps/s")
    kw = dict(plans or sweep(op2amd, mesh, args.sweep_iters, renumber))
    ref_design = kw.pop("_reference_design", None)
    out["tuned_plans"] = kw
    if plain_too and ref_design:
        # the reference's schemes at their best block sizes, synced per loop
        rd = op2amd.airfoil.run(mesh, n, mode="plain", warmup=W, renumber=renumber,
                                **ref_design)
        out["reference_design_best_plans"] = ref_design
        out["reference_design_best_steps_per_s"] = K / rd["seconds"]
        runs["reference_design_best"] = rd
        log(f"  best staged/colour plans {ref_design}: "
            f"{out['reference_design_best_steps_per_s']:.1f} steps/s")
    eager = op2amd.airfoil.run(mesh, n, mode="tuned", warmup=W, renumber=renumber, **kw)
    out["tuned_eager_steps_per_s"] = K / eager["seconds"]
    runs["tuned_eager"] = eager
    log(f"  tuned plans, eager: {out['tuned_eager_steps_per_s']:.1f} steps/s")
    graphed = op2amd.airfoil.run(mesh, n, mode="graphed", warmup=W, renumber=renumber,
                                 graph_steps=args.graph_steps, **kw)
    runs["graphed"] = graphed
    out["graphed_steps_per_s"] = K / graphed["seconds"]
    out["steps_per_graph_launch"] = graphed["graph_steps"]
    # host-to-device work outside the timed region: declarations (copies and
    # uploads of every dat and map), plan builds, graph capture
    out["graphed_setup_seconds"] = graphed.get("setup_seconds")
    log(f"  graphed {kw}: {out['graphed_steps_per_s']:.1f} steps/s")
    # the same graphs without loop-chain fusion: what the graphs alone give
    api = op2amd.api
    api.op2h_set_fusion(0)
    try:
        unfused = op2amd.airfoil.run(mesh, n, mode="graphed", warmup=W, renumber=renumber,
                                     graph_steps=args.graph_steps, **kw)
    finally:
        api.op2h_set_fusion(1)
    runs["graphed_unfused"] = unfused
    out["graphed_unfused_steps_per_s"] = K / unfused["seconds"]
    out["fusion_gain"] = out["graphed_steps_per_s"] / out["graphed_unfused_steps_per_s"]
    log(f"  graphed without loop-chain fusion: {out['graphed_unfused_steps_per_s']:.1f} steps/s")
    out["graph_gain_vs_tuned_eager"] = out["graphed_steps_per_s"] / out["tuned_eager_steps_per_s"]
    if plain_too:
        out["speedup_vs_plain"] = out["graphed_steps_per_s"] / out["plain_steps_per_s"]
        if "reference_design_best_steps_per_s" in out:
            out["speedup_vs_reference_design_best"] = (
                out["graphed_steps_per_s"] / out["reference_design_best_steps_per_s"])
        out["rms_rel_err_graphed_vs_plain"] = rel(graphed["rms"], plain["rms"])
        out["q_rel_err_graphed_vs_plain"] = float(
            np.max(np.abs(graphed["q"] - plain["q"])) / np.max(np.abs(plain["q"])))
    timed = op2amd.airfoil.run(mesh, min(K, 100) + 2, mode="tuned", warmup=2,
                               kernel_timing=True, renumber=renumber, **kw)
    per_loop = {}
    for lp in op2amd.airfoil.LOOPS:
        us = 1e3 * timed["loop_kernel_ms"][lp] / max(timed["loop_calls"][lp], 1)
        b = timed["loop_bytes"][lp]
        per_loop[lp] = {"avg_us": round(us, 3), "bytes_per_call": b,
                        "achieved_GBs": round(b / (us * 1e-6) / 1e9, 1) if us > 0 else 0.0}
    out["per_loop"] = per_loop
    out["step_bytes"] = sum(timed["loop_bytes"][lp] * (1 if lp == "save_soln" else 2)
                            for lp in op2amd.airfoil.LOOPS)
    out["_runs"] = runs
    return out


def parity(runs, orc):
    """Per run: the largest relative difference of rms from the oracle's over
    every iteration (warm-up, eager and graph-replayed ones alike), and of
    the final q, relative to max |q|."""
    res = {}
    for name, r in runs.items():
        q_err = float(np.max(np.abs(r["q"] - orc["q"])) / np.max(np.abs(orc["q"])))
        rms_err = rel(r["rms"], orc["rms"])
        # a NaN anywhere is a failed check, never a pass
        res[name] = {"rms_rel_err_max": rms_err if np.isfinite(rms_err) else float("inf"),
                     "q_rel_err": q_err if np.isfinite(q_err) else float("inf")}
    return res


def workload_large(op2amd, args):
    """4800x2400 (11.5M cells): FP64 vs FP32 mixed mode, both tuned+graphed.
    Error reference: the FP64 run with the owner partition plan, whose q,
    maxerr and errloc equal the sequential executor's bit for bit
    (tests/test_gpu_airfoil.py)."""
    I, J = (int(v) for v in args.large_mesh.split("x"))
    mesh = op2amd.mesh.generate(I, J, 0, args.ordering)
    n, W = args.large_iters, 20
    out = {}
    # block sizes: owner patches of 256 or 512 cells, direct loops 256..1024,
    # each timed per call with launch-attached events
    def per_call(kw, loop):
        t = op2amd.airfoil.run(mesh, 8, mode="tuned", warmup=2, kernel_timing=True, **kw)
        return t["loop_kernel_ms"][loop] / max(t["loop_calls"][loop], 1)
    base = dict(plan_res="owner", plan_bres="atomic", block_bres=128)
    res_t = {b: per_call(dict(base, block_res=b), "res_calc") for b in (256, 512)}
    block_res = min(res_t, key=res_t.get)
    adt_t = {b: per_call(dict(base, block_res=block_res, block_adt=b), "adt_calc")
             for b in (256, 512, 1024)}
    block_adt = min(adt_t, key=adt_t.get)
    log(f"  owner patch size: {res_t}; adt_calc block: {adt_t}")
    out["block_sweep_ms"] = {"res_calc": res_t, "adt_calc": adt_t}
    common = dict(plan_res="owner", block_res=block_res, plan_bres="atomic", block_bres=128,
                  block_direct=1024, block_adt=block_adt)
    common["graph_steps"] = args.graph_steps
    ref = op2amd.airfoil.run(mesh, n, mode="graphed", warmup=W, **common)
    out["fp64_steps_per_s"] = (n - W) / ref["seconds"]
    mixed = op2amd.airfoil.run(mesh, n, mode="graphed", warmup=W, mixed=True, **common)
    out["mixed_steps_per_s"] = (n - W) / mixed["seconds"]
    out["mixed_rms_rel_err_last"] = float(abs(mixed["rms"][-1] / ref["rms"][-1] - 1))
    out["mixed_rms_rel_err_max"] = rel(mixed["rms"], ref["rms"])
    out["mixed_q_rel_err"] = float(np.max(np.abs(mixed["q"] - ref["q"])) / np.max(np.abs(ref["q"])))
    for label, kw in (("fp64", {}), ("mixed", {"mixed": True})):
        t = op2amd.airfoil.run(mesh, 12, mode="tuned", warmup=2, kernel_timing=True, **common, **kw)
        out[f"{label}_per_loop"] = {
            lp: {"avg_us": round(1e3 * t["loop_kernel_ms"][lp] / max(t["loop_calls"][lp], 1), 2),
                 "bytes_per_call": t["loop_bytes"][lp],
                 "achieved_GBs": round(t["loop_bytes"][lp] / (1e-3 * t["loop_kernel_ms"][lp] /
                                                              max(t["loop_calls"][lp], 1)) / 1e9, 1)}
            for lp in op2amd.airfoil.LOOPS}
    ceiling, ceiling_all = stream_ceiling(op2amd)
    out["stream_ceiling_GBs"], out["stream_probe_GBs"] = ceiling, ceiling_all
    for label in ("fp64", "mixed"):
        with_fractions(out[f"{label}_per_loop"], ceiling)
    dom = out["fp64_per_loop"]["res_calc"]
    line = {
        "metric": "airfoil steps/s, 11.5M cells, FP64 vs FP32-flux mixed mode",
        "value": round(out["fp64_steps_per_s"], 2), "unit": "steps/s", "n_gpus": 1,
        "steps": n - W, "warmup": W, "ms_per_step": round(1e3 / out["fp64_steps_per_s"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"airfoil {I}x{J} ({mesh.ncell} cells), {args.ordering} "
                               f"ordering, {n} iterations", **out},
        "roofline": {"bound": "hbm",
                     "kernel": f"res_calc ({common['plan_res']}, block {common['block_res']})",
                     "achieved": dom["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dom["achieved_GBs"] / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(
                         f"res_calc/{common['plan_res']}/{common['block_res']}",
                         f"{args.large_mesh}/{args.ordering}")},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def workload_ensemble(op2amd, args):
    """256 uneven meshes (40x20 .. 640x320, per-member Mach and alpha): each
    loop one persistent launch over all members' work items, the step
    graph-captured, workgroup sizes picked by the ensemble's autotuner over
    the 10 warm-up steps, vs one launch per loop per member.  Per-member step
    times: inside the batched step from device-clock stamps of every work
    item (a separate, shorter run), in the looped run from HIP events."""
    sp = op2amd.ensemble.spec(nmesh=args.ensemble_size)
    n, W = args.ensemble_iters, 10
    batched = op2amd.ensemble.run(sp, n, warmup=W, batched=True, graphed=True, autotune=True)
    looped = op2amd.ensemble.run(sp, min(n, 30), warmup=W, batched=False)
    timing = op2amd.ensemble.run(sp, 12, warmup=2, batched=False, member_timing=True)
    btiming = op2amd.ensemble.run(sp, 20, warmup=W + 1, batched=True, graphed=True,
                                  autotune=True, member_timing=True)
    nc = batched["total_cells"]
    b_ups = nc * (n - W) / batched["seconds"]
    l_ups = nc * (min(n, 30) - W) / looped["seconds"]
    per = timing["member_step_ms"]
    lat, busy = btiming["member_step_ms"], btiming["member_busy_ms"]
    same = bool(np.array_equal(batched["rms"][:min(n, 30)], looped["rms"][:min(n, 30)]))
    pct = lambda v, p: round(float(np.percentile(v, p)), 4)
    line = {
        "metric": "ensemble cell-updates/s (256 uneven airfoil meshes, FP64)",
        "value": round(b_ups, 1), "unit": "cell-steps/s", "n_gpus": 1, "steps": n - W,
        "warmup": W, "ms_per_step": round(1e3 * batched["seconds"] / (n - W), 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{sp.nmesh} members, {nc} cells, {batched['total_edges']} edges",
                   "batched_graphed_cell_steps_per_s": b_ups,
                   "looped_cell_steps_per_s": l_ups,
                   "speedup_batched_vs_looped": b_ups / l_ups,
                   "batched_step_ms": 1e3 * batched["seconds"] / (n - W),
                   "looped_step_ms": 1e3 * looped["seconds"] / (min(n, 30) - W),
                   "batched_member_step_ms_p50": pct(lat, 50),
                   "batched_member_step_ms_p99": pct(lat, 99),
                   "batched_member_busy_ms_p50": pct(busy, 50),
                   "batched_member_busy_ms_p99": pct(busy, 99),
                   "looped_member_step_ms_p50": pct(per, 50),
                   "looped_member_step_ms_p99": pct(per, 99),
                   "looped_member_step_ms_sum": float(per.sum()),
                   "rms_identical_batched_vs_looped": same,
                   "batched_workgroup_sizes": batched["blocks"],
                   "batched_persistent_grids": batched["grids"],
                   "dead_res": batched["dead_res"], "fused_bres": batched["fused_bres"]},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def workload_translated(op2amd, args):
    """The airfoil program as op2amd.translator emits it for this backend
    (tests/apps/airfoil/airfoil_tr.cpp, built by __graft_entry__.build()):
    per-loop kernel times and GB/s of the generated kernels, from the
    program's op_timing_output with the OP_HIP_KERNEL_TIMING token, next to the
    hand-written kernels' on the same mesh with the same plans.  The program
    reads rms back after every update (it prints it each iteration), so its
    step rate is an eager one."""
    import subprocess
    import tempfile
    exe = ROOT / "tests" / "apps" / "_build" / "airfoil_tr" / "airfoil_tr"
    if not exe.exists():
        raise SystemExit(f"{exe} is missing: run __graft_entry__.build() first")
    I, J = (int(v) for v in args.mesh.split("x"))
    mesh = op2amd.mesh.generate(I, J, 0, args.ordering)
    iters = args.translated_iters
    plans = {"res_calc": ("owner", 128), "bres_calc": ("atomic", 128)}
    tokens = ["OP_HIP_KERNEL_TIMING"] + [f"OP_HIP_PLAN={lp}:{p}:{b}" for lp, (p, b) in plans.items()]
    with tempfile.TemporaryDirectory() as d:
        grid = os.path.join(d, "new_grid.dat")
        mesh.write(grid)
        log(f"translated airfoil: {I}x{J}, {iters} iterations, kernel timing on")
        r = subprocess.run([str(exe), grid, str(iters)] + tokens, capture_output=True, text=True,
                           timeout=600)
        if r.returncode != 0:
            raise SystemExit(f"airfoil_tr failed ({r.returncode}): {r.stdout[-2000:]}{r.stderr[-2000:]}")
        # the step rate without the per-loop timing events (which wait for
        # each kernel), as an application would run it; the program's clock
        # also covers the plans built at each loop's first call, so the run is
        # long enough to amortise them
        iters2 = max(iters, 2000)
        log(f"translated airfoil: {I}x{J}, {iters2} iterations, kernel timing off")
        r2 = subprocess.run([str(exe), grid, str(iters2)] + tokens[1:], capture_output=True,
                            text=True, timeout=600)
        if r2.returncode != 0:
            raise SystemExit(f"airfoil_tr failed ({r2.returncode}): {r2.stdout[-2000:]}{r2.stderr[-2000:]}")
        # and as launches of a captured 10-step graph after one eager step
        iters3 = iters2 + 1
        log(f"translated airfoil: {I}x{J}, {iters3} iterations, graphs of 10 steps")
        r3 = subprocess.run([str(exe), grid, str(iters3), "graph=10"] + tokens[1:],
                            capture_output=True, text=True, timeout=600)
        if r3.returncode != 0:
            raise SystemExit(f"airfoil_tr failed ({r3.returncode}): {r3.stdout[-2000:]}{r3.stderr[-2000:]}")
    secs = secs_untimed = secs_graph = None
    for line in r2.stdout.splitlines():
        if line.startswith("time "):
            secs_untimed = float(line.split()[1])
    for line in r3.stdout.splitlines():
        if line.startswith("time "):
            secs_graph = float(line.split()[1])
    gen = {}
    for line in r.stdout.splitlines():
        if line.startswith("time "):
            secs = float(line.split()[1])
        f = [t.strip() for t in line.split(";")]
        if len(f) == 6 and f[0].isdigit():
            count, kms, gbs, name = int(f[0]), float(f[3]), float(f[4]), f[5]
            gen[name] = {"calls": count, "avg_us": round(1e3 * kms / max(count, 1), 3),
                         "achieved_GBs": round(gbs, 1)}
    hand = op2amd.airfoil.run(mesh, iters + 2, mode="tuned", warmup=2, kernel_timing=True,
                              plan_res="owner", block_res=128, plan_bres="atomic",
                              block_bres=128)
    hw = {}
    for lp in op2amd.airfoil.LOOPS:
        us = 1e3 * hand["loop_kernel_ms"][lp] / max(hand["loop_calls"][lp], 1)
        hw[lp] = {"avg_us": round(us, 3),
                  "achieved_GBs": round(hand["loop_bytes"][lp] / (us * 1e-6) / 1e9, 1)}
    res_gen = gen.get("res_calc", {}).get("achieved_GBs", 0.0)
    line = {
        "metric": "translated airfoil: res_calc kernel GB/s of the generated code (720k cells)",
        "value": res_gen, "unit": "GB/s", "n_gpus": 1, "steps": iters, "warmup": 0,
        "ms_per_step": round(1e3 * secs_untimed / iters2, 4) if secs_untimed else None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"airfoil {I}x{J} ({mesh.ncell} cells), {args.ordering} "
                               f"ordering, {iters} iterations, eager (rms read every update)",
                   "plans": {lp: f"{p}:{b}" for lp, (p, b) in plans.items()},
                   "translated_per_loop": gen, "hand_written_per_loop": hw,
                   "translated_steps_per_s": round(iters2 / secs_untimed, 1) if secs_untimed else None,
                   "translated_steps_per_s_iterations": iters2,
                   "translated_graphed_steps_per_s": round(iters3 / secs_graph, 1) if secs_graph else None,
                   "translated_steps_per_s_kernel_timing": round(iters / secs, 1) if secs else None},
        "roofline": {"bound": "hbm", "kernel": "res_calc (generated, owner, block 128)",
                     "achieved": res_gen, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(res_gen / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic("res_calc/generated-owner/128",
                                                 f"{I}x{J}/{args.ordering}")},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def workload_copy(op2amd, args):
    """Counter calibration: save_soln (qold = q, a plain copy) on a set of
    16.8M elements, 512 MiB read and 512 MiB written per call, far beyond the
    L2 and the Infinity Cache.  Under rocprofv3 --pmc its FETCH_SIZE and
    WRITE_SIZE per dispatch, divided by 512 MiB, give the counters' scale."""
    api = op2amd.api
    n = 1 << 24
    api.init()
    cells = api.op_decl_set(n, b"cells")
    q = api.decl_dat(cells, 4, "double", np.ones((n, 4)), "q")
    qold = api.decl_dat(cells, 4, "double", np.zeros((n, 4)), "qold")
    a_q = api.arg_dat(q, -1, None, 4, "double", 0)
    a_qold = api.arg_dat(qold, -1, None, 4, "double", 1)
    loop = api.loops["op_par_loop_airfoil_1_save_soln"]
    for _ in range(args.warmup):
        loop(b"save_soln", cells, a_q, a_qold)
    api.op2h_device_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loop(b"save_soln", cells, a_q, a_qold)
    api.op2h_device_sync()
    dt = (time.perf_counter() - t0) / args.steps
    bytes_per_call = 2.0 * n * 32
    out = api.fetch(qold, (n, 4), np.float64)
    line = {"metric": "copy kernel bandwidth (counter calibration)",
            "value": round(bytes_per_call / dt / 1e9, 1), "unit": "GB/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": f"save_soln copy of {n} x 4 doubles (512 MiB in, 512 MiB out)",
                       "read_bytes_per_call": n * 32, "write_bytes_per_call": n * 32,
                       "copy_correct": bool(np.all(out == 1.0))},
            "roofline": None, "cpu_baseline": None}
    print(json.dumps(line), flush=True)


def workload_costmodel(op2amd, args):
    """Plan selection over synthetic op_par_loop descriptors: exhaustive sweep
    of each loop's variants on the GPU, an ensemble of MLPs (2 x 64, FP32)
    trained on 80 % of the loops with the library's HIP kernels (epoch count
    chosen on a validation part of the training loops), regret on the held-out
    20 % against the sweep's best, the runtime's static default and the best
    single static rule fitted on the training loops."""
    cm = op2amd.costmodel
    descs = cm.random_descs(args.cm_loops, args.cm_seed)
    t0 = time.time()
    recs = cm.sweep(descs, reps=5, log=log)
    t_sweep = time.time() - t0
    if args.cm_dump:
        with open(args.cm_dump, "w") as f:
            json.dump(recs, f)
    perm = np.random.default_rng(args.cm_seed).permutation(len(recs))
    ntest = max(1, len(recs) // 5)
    test_ids, train_ids = list(perm[:ntest]), list(perm[ntest:])
    t0 = time.time()
    model = cm.CostModel(seed=args.cm_seed).fit_validated(recs, train_ids,
                                                           max_epochs=args.cm_epochs)
    t_train = time.time() - t0
    log(f"  validation curve (epochs, mean regret): {model.validation_curve}")
    t0 = time.time()
    reg_model = cm.regret(recs, test_ids, model.choose)
    t_pick = (time.time() - t0) / len(test_ids)
    reg_default = cm.regret(recs, test_ids, cm.default_choice)
    reg_fixed = cm.regret(recs, test_ids, cm.best_fixed_choice(recs, train_ids))
    reg_train = cm.regret(recs, train_ids, model.choose)
    reg_top2 = cm.regret(recs, test_ids, lambda r: model.choose_timed(r, 2))
    # the model inside the runtime: each airfoil loop's plan picked at its
    # first call (op2h_set_plan_model), against the headline's exhaustive sweep
    mesh = op2amd.mesh.generate(1200, 600, 0, "rcm")
    n, W = 220, 20
    model.install(top=1)
    try:
        picked = op2amd.airfoil.run(mesh, n, mode="graphed", warmup=W,
                                    graph_steps=args.graph_steps)
    finally:
        cm.CostModel.uninstall()
    kw = {k: v for k, v in sweep(op2amd, mesh, args.sweep_iters, False).items()
          if not k.startswith("_")}  # the plans alone
    swept = op2amd.airfoil.run(mesh, n, mode="graphed", warmup=W, graph_steps=args.graph_steps,
                               **kw)
    model_rate, swept_rate = (n - W) / picked["seconds"], (n - W) / swept["seconds"]
    log(f"  airfoil 1200x600: model plans {picked['chosen']} {model_rate:.0f} steps/s, "
        f"swept {kw} {swept_rate:.0f} steps/s")
    line = {
        "metric": "cost-model plan selection: mean time regret vs exhaustive sweep, held-out loops",
        "value": round(float(reg_model.mean()), 4), "unit": "fraction", "n_gpus": 1,
        "steps": len(recs), "warmup": 0, "ms_per_step": None, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (model), f32/f64 (loops)",
        "data": "synthetic",
        "config": {"workload": f"{len(recs)} synthetic op_par_loop descriptors, 80/20 split",
                   "train_loops": len(train_ids), "test_loops": len(test_ids),
                   "ensemble": model.members, "epochs": model.epochs,
                   "weight_decay": model.wd,
                   "regret_model_mean": float(reg_model.mean()),
                   "regret_model_median": float(np.median(reg_model)),
                   "regret_model_p90": float(np.percentile(reg_model, 90)),
                   "loops_within_5pct_model": float(np.mean(reg_model <= 0.05)),
                   "regret_model_train_mean": float(reg_train.mean()),
                   "regret_model_top2_timed_mean": float(reg_top2.mean()),
                   "regret_default_plan_mean": float(reg_default.mean()),
                   "regret_best_fixed_rule_mean": float(reg_fixed.mean()),
                   "loops_within_5pct_best_fixed_rule": float(np.mean(reg_fixed <= 0.05)),
                   "sweep_seconds": t_sweep, "train_seconds": t_train,
                   "model_pick_seconds_per_loop": t_pick,
                   "airfoil_1200x600_model_plans": {lp: list(v) for lp, v in
                                                    picked["chosen"].items()},
                   "airfoil_1200x600_model_steps_per_s": model_rate,
                   "airfoil_1200x600_swept_plans": kw,
                   "airfoil_1200x600_swept_steps_per_s": swept_rate,
                   "airfoil_1200x600_model_regret": swept_rate / model_rate - 1.0},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="airfoil",
                    choices=["airfoil", "airfoil-large", "ensemble", "costmodel", "copy",
                             "translated"],
                    help="airfoil: the headline 720k line (default); the others "
                         "measure the 11.5M-cell FP64/mixed config and the ensemble")
    ap.add_argument("--large-mesh", default="4800x2400")
    ap.add_argument("--large-iters", type=int, default=2000)
    ap.add_argument("--ensemble-size", type=int, default=256)
    ap.add_argument("--ensemble-iters", type=int, default=200)
    ap.add_argument("--translated-iters", type=int, default=100)
    ap.add_argument("--cm-loops", type=int, default=200)
    ap.add_argument("--cm-epochs", type=int, default=4000)
    ap.add_argument("--cm-dump", default="", help="write the sweep records (JSON) here")
    ap.add_argument("--cm-seed", type=int, default=0)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--mesh", default="1200x600")
    ap.add_argument("--ordering", default="rcm", choices=["random", "rcm"],
                    help="input ordering of the headline measurement")
    ap.add_argument("--renumber", default="none", choices=["none", "transparent"],
                    help="headline measurement with the device in its own RCM numbering "
                         "(OP_HIP_RENUMBER=pecell; the application keeps the mesh's)")
    ap.add_argument("--plans", default=None,
                    help="skip the sweep, e.g. res=color:256,bres=atomic:128,direct=1024")
    ap.add_argument("--only-headline", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sweep-iters", type=int, default=6)
    ap.add_argument("--graph-steps", type=int, default=10,
                    help="time steps captured per graph (lowered to a divisor of --steps)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the library runs one mesh on one GPU; it has no multi-GPU sharding)")
        sys.exit(2)

    import op2amd
    if args.workload == "airfoil-large":
        return workload_large(op2amd, args)
    if args.workload == "ensemble":
        return workload_ensemble(op2amd, args)
    if args.workload == "costmodel":
        return workload_costmodel(op2amd, args)
    if args.workload == "copy":
        return workload_copy(op2amd, args)
    if args.workload == "translated":
        return workload_translated(op2amd, args)
    from oracle import oracle
    I, J = (int(v) for v in args.mesh.split("x"))
    K, W = args.steps, args.warmup
    plans = parse_plans(args.plans) if args.plans else None

    meshes = {}
    for ordering in ("rcm", "random"):
        if args.only_headline and ordering != args.ordering:
            continue
        t0 = time.time()
        meshes[ordering] = op2amd.mesh.generate(I, J, 0, ordering)
        log(f"mesh {I}x{J} {ordering}: {meshes[ordering].ncell} cells "
            f"({time.time() - t0:.1f}s)")
    mesh = meshes[args.ordering]

    ceiling, ceiling_all = stream_ceiling(op2amd)
    log(f"streaming ceiling: {ceiling:.1f} GB/s ({ceiling_all})")
    log(f"headline: {args.ordering} ordering")
    head = measure(op2amd, mesh, K, W, args, plans=plans,
                   renumber="transparent" if args.renumber == "transparent" else False)
    with_fractions(head["per_loop"], ceiling)
    value = head["graphed_steps_per_s"]
    head_runs = head.pop("_runs")

    others, other_runs = {}, {}
    if not args.only_headline:
        other = "random" if args.ordering == "rcm" else "rcm"
        log(f"{other} ordering, as given")
        # the headline's plans (a full sweep per mesh would triple the run)
        plans_other = dict(head["tuned_plans"])
        if "reference_design_best_plans" in head:
            plans_other["_reference_design"] = head["reference_design_best_plans"]
        others[f"{other}_ordering"] = measure(op2amd, meshes[other], K, W, args,
                                              plans=plans_other)
        other_runs[f"{other}_ordering"] = others[f"{other}_ordering"].pop("_runs")
        if other == "random":
            log("random ordering after op_renumber(pecell)")
            r = measure(op2amd, meshes["random"], K, W, args, renumber=True,
                        plans=plans_other)
            other_runs["random_ordering_op_renumber"] = r.pop("_runs")
            others["random_ordering_op_renumber"] = r
            log("random ordering, renumbered on the device only (OP_HIP_RENUMBER=pecell)")
            r = measure(op2amd, meshes["random"], K, W, args, renumber="transparent",
                        plans=plans_other)
            other_runs["random_ordering_transparent_renumber"] = r.pop("_runs")
            r["vs_rcm_ordering"] = r["graphed_steps_per_s"] / head["graphed_steps_per_s"]
            others["random_ordering_transparent_renumber"] = r

    # parity of every run over all W + K iterations against the oracle's
    # OpenMP executor on the same mesh; its run doubles as the CPU baseline's
    # sample when that is long enough
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(os.cpu_count() or 1, 16)
    n_all = K + W
    iters = n_all
    if not args.no_cpu:
        def cpu_call(n):
            t = time.time()
            oracle.airfoil_run(mesh, n, 1, "openmp", nthreads=nthreads)
            return time.time() - t
        # the per-step cost from two run lengths, so the plan builds and
        # allocations of each call cancel; then a ~20 s sample
        t_short = cpu_call(2)
        one = max((cpu_call(12) - t_short) / 10, 1e-4)
        iters = max(n_all, int(max(10, min(20000, 20.0 / one))))
    log(f"oracle: OpenMP executor, {iters} iterations on {nthreads} threads")
    cpu_run = oracle.airfoil_run(mesh, iters, 1, "openmp", nthreads=nthreads)
    orc = {"rms": cpu_run["rms"][:n_all]}
    if iters == n_all:
        orc["q"] = cpu_run["q"]
    else:
        orc["q"] = oracle.airfoil_run(mesh, n_all, 1, "openmp", nthreads=nthreads)["q"]
    checks = {args.ordering: parity(head_runs, orc)}
    for name, runs in other_runs.items():
        if name == "random_ordering_op_renumber" or name.startswith("random"):
            m_o = meshes["random"]
        else:
            m_o = meshes["rcm"]
        o2 = oracle.airfoil_run(m_o, n_all, 1, "openmp", nthreads=nthreads)
        checks[name] = parity(runs, {"rms": o2["rms"], "q": o2["q"]})
    worst = max(v for c in checks.values() for r in c.values() for v in r.values())
    parity_ok = bool(worst <= PARITY_TOL)
    log(f"parity vs the oracle over {n_all} iterations: worst {worst:.2e} "
        f"({'ok' if parity_ok else 'FAILED'})")

    cpu = None
    if not args.no_cpu:
        secs = float(sum(cpu_run["loop_seconds"]))  # the time-step loops only
        # per-loop CPU time and GB/s on the GPU line's byte model
        calls = {"save_soln": 1, "adt_calc": 2, "res_calc": 2, "bres_calc": 2, "update": 2}
        cpu_loops = {}
        for i, lp in enumerate(op2amd.airfoil.LOOPS):
            us = 1e6 * cpu_run["loop_seconds"][i] / (iters * calls[lp])
            b = head["per_loop"][lp]["bytes_per_call"]
            cpu_loops[lp] = {"avg_us": round(us, 2), "achieved_GBs": round(b / (us * 1e-6) / 1e9, 1)}
        others["cpu_openmp_per_loop"] = cpu_loops
        cpu = {"value": round(iters / secs, 3), "unit": "steps/s", "cores": nthreads,
               "kind": "port",
               "sample": f"{iters} airfoil steps on the same {I}x{J} {args.ordering} mesh "
                         f"({secs:.1f} s in the five loops, plan builds excluded): the "
                         "oracle's OpenMP executor (block colours in sequence, blocks in "
                         "parallel), g++ -O3 -fopenmp"}
        log(f"cpu baseline: {cpu['value']:.2f} steps/s on {nthreads} threads")
        # single-thread sequential executor beside it (a ~5 s sample)
        t = time.time()
        oracle.airfoil_run(mesh, 1, 1, "seq")
        one = max(time.time() - t, 1e-3)
        iters_seq = int(max(1, min(100, 5.0 / one)))
        t = time.time()
        oracle.airfoil_run(mesh, iters_seq, 1, "seq")
        others["cpu_seq_1thread_steps_per_s"] = round(iters_seq / (time.time() - t), 3)
        log(f"cpu seq, 1 thread: {others['cpu_seq_1thread_steps_per_s']:.2f} steps/s")

    dom = head["per_loop"]["res_calc"]
    kw = head["tuned_plans"]
    ordering = args.ordering + ("+device-rcm" if args.renumber == "transparent" else "")
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(1e3 / value, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"airfoil {I}x{J} ({mesh.ncell} cells, {mesh.nedge} edges), "
                        f"{args.ordering} ordering"
                        f"{', renumbered on the device only' if args.renumber == 'transparent' else ''}"
                        ", FP64; one step = save_soln + 2 x (adt_calc, res_calc, bres_calc, update)",
            "mesh": args.mesh, "ordering": ordering,
            "stream_ceiling_GBs": ceiling, "stream_probe_GBs": ceiling_all,
            "parity_vs_oracle": {
                "oracle": f"OpenMP executor on {nthreads} threads, same mesh and ordering, "
                          f"all {n_all} iterations (warm-up, eager and graph-replayed)",
                "tolerance": PARITY_TOL, "worst": worst, "ok": parity_ok, "runs": checks},
            **{k: v for k, v in head.items()},
            **others,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"res_calc ({kw['plan_res']}, block {kw['block_res']})",
            "achieved": dom["achieved_GBs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(dom["achieved_GBs"] / HBM_PEAK_GBS, 4),
            "traffic": profiled_traffic(f"res_calc/{kw['plan_res']}/{kw['block_res']}",
                                        f"{args.mesh}/{ordering}"),
        },
        "cpu_baseline": cpu,
    }
    if not parity_ok:
        # a rate whose results differ from the oracle's is not reported
        out["value"] = None
        out["ms_per_step"] = None
    print(json.dumps(out), flush=True)
    if not parity_ok:
        sys.exit(1)


if __name__ == "__main__":
    main()
