"""Airfoil benchmark: steps/s at 720k cells, tuned + graphed vs plain HIP.

One step is one airfoil time step (save_soln, then twice adt_calc, res_calc,
bres_calc, update with the rms reduction) over the whole 1200x600 mesh
(720,000 cells, 1,438,200 edges), FP64, with all data resident in HBM.

  plain    the reference's execution model on this backend: LDS-staged,
           two-level-coloured indirect loops, device sync and gbl read-back
           after every loop;
  tuned    per-loop plan and block size picked by timing every variant on the
           same mesh (exhaustive sweep);
  graphed  the tuned plans with each time step replayed from one HIP graph,
           rms/maxerr/errloc read back from a device history at the end.

Two input orderings are measured (both named by the workload): the headline
uses the RCM-renumbered mesh; the randomly numbered mesh is reported beside
it, as given and after the runtime's op_renumber.  Parity: the graphed rms
trajectory against the plain one at every iteration, and against the
oracle's OpenMP run for the first iterations (max relative error).
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
PLANS = ("atomic", "color", "staged", "gather")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.abs(b)))


def sweep(op2amd, mesh, iters, renumber):
    """Exhaustive sweep: per indirect loop, every plan x block size, timed
    per call with HIP events; then the block size of the direct loops."""
    best = {}
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
                block_direct=bd, block_adt=bd)


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
    n = K + W
    out = {}
    if plain_too:
        plain = op2amd.airfoil.run(mesh, n, mode="plain", warmup=W, renumber=renumber)
        out["plain_steps_per_s"] = K / plain["seconds"]
        log(f"  plain: {out['plain_steps_per_s']:.1f} steps/s")
    kw = plans or sweep(op2amd, mesh, args.sweep_iters, renumber)
    out["tuned_plans"] = kw
    graphed = op2amd.airfoil.run(mesh, n, mode="graphed", warmup=W, renumber=renumber, **kw)
    out["graphed_steps_per_s"] = K / graphed["seconds"]
    log(f"  graphed {kw}: {out['graphed_steps_per_s']:.1f} steps/s")
    if plain_too:
        out["speedup_vs_plain"] = out["graphed_steps_per_s"] / out["plain_steps_per_s"]
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
    out["_graphed"] = graphed
    return out


def workload_large(op2amd, args):
    """4800x2400 (11.5M cells): FP64 vs FP32 mixed mode, both tuned+graphed.
    Error reference: the FP64 owner-compute run, whose q, maxerr and errloc
    equal the sequential executor's bit for bit (tests/test_gpu_airfoil.py)."""
    I, J = (int(v) for v in args.large_mesh.split("x"))
    mesh = op2amd.mesh.generate(I, J, 0, args.ordering)
    n, W = args.large_iters, 20
    out = {}
    common = dict(plan_res="gather", block_res=512, plan_bres="atomic", block_bres=128,
                  block_direct=1024, block_adt=1024)
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
    dom = out["fp64_per_loop"]["res_calc"]
    line = {
        "metric": "airfoil steps/s, 11.5M cells, FP64 vs FP32-flux mixed mode",
        "value": round(out["fp64_steps_per_s"], 2), "unit": "steps/s", "n_gpus": 1,
        "steps": n - W, "warmup": W, "ms_per_step": round(1e3 / out["fp64_steps_per_s"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"airfoil {I}x{J} ({mesh.ncell} cells), {args.ordering} "
                               f"ordering, {n} iterations", **out},
        "roofline": {"bound": "hbm", "kernel": "res_calc (gather, block 512)",
                     "achieved": dom["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dom["achieved_GBs"] / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def workload_ensemble(op2amd, args):
    """256 uneven meshes (40x20 .. 640x320, per-member Mach and alpha):
    one launch per loop over all members, graph-captured, vs one launch per
    loop per member."""
    sp = op2amd.ensemble.spec(nmesh=args.ensemble_size)
    n, W = args.ensemble_iters, 5
    batched = op2amd.ensemble.run(sp, n, warmup=W, batched=True, graphed=True)
    looped = op2amd.ensemble.run(sp, min(n, 30), warmup=W, batched=False)
    timing = op2amd.ensemble.run(sp, 12, warmup=2, batched=False, member_timing=True)
    nc = batched["total_cells"]
    b_ups = nc * (n - W) / batched["seconds"]
    l_ups = nc * (min(n, 30) - W) / looped["seconds"]
    per = timing["member_step_ms"]
    same = bool(np.array_equal(batched["rms"][:min(n, 30)], looped["rms"][:min(n, 30)]))
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
                   "looped_member_step_ms_p50": float(np.percentile(per, 50)),
                   "looped_member_step_ms_p99": float(np.percentile(per, 99)),
                   "looped_member_step_ms_sum": float(per.sum()),
                   "rms_identical_batched_vs_looped": same},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def workload_costmodel(op2amd, args):
    """Plan selection over synthetic op_par_loop descriptors: exhaustive sweep
    of each loop's variants on the GPU, an MLP (2 x 64, FP32) trained on 80 %
    of the loops with the library's HIP kernels, regret on the held-out 20 %."""
    cm = op2amd.costmodel
    rng = np.random.default_rng(args.cm_seed)
    recs = []
    t_sweep = time.time()
    for i in range(args.cm_loops):
        n = float(10 ** rng.uniform(4, 7))
        arity = int(rng.integers(1, 5))
        dim = int(rng.choice([1, 2, 4, 8]))
        access = str(rng.choice(["read", "inc", "rw"]))
        dbl = int(rng.integers(0, 2))
        ordering = str(rng.choice(["random", "rcm", "hilbert"]))
        k = int(rng.integers(4, 65))
        d = cm.desc_from(n, arity, dim, access, dbl, ordering, k, 100 + i)
        var = cm.variants(access)
        ms, shape = cm.time_variants(d, var, reps=3)
        recs.append(dict(arity=arity, dim=dim, access=access, is_double=dbl, ordering=ordering,
                         k=k, variants=var, ms=ms, **shape))
        best = min(t for t in ms if t > 0)
        log(f"  loop {i}: n={shape['set_size']} a={arity} d={dim} {access} "
            f"{'f64' if dbl else 'f32'} {ordering} k={k} best {best * 1e3:.1f}us")
    t_sweep = time.time() - t_sweep
    perm = rng.permutation(len(recs))
    ntest = max(1, len(recs) // 5)
    test_ids, train_ids = perm[:ntest], perm[ntest:]

    def rows(ids):
        X, y, owner = [], [], []
        for i in ids:
            r = recs[i]
            for (plan, bs), t in zip(r["variants"], r["ms"]):
                if t > 0:
                    X.append(cm.features(r, plan, bs))
                    y.append(np.log(t))
                    owner.append(i)
        return np.array(X, np.float32), np.array(y, np.float32), owner
    Xtr, ytr, _ = rows(train_ids)
    mu, sd = Xtr.mean(axis=0), Xtr.std(axis=0) + 1e-6
    ymu, ysd = ytr.mean(), ytr.std() + 1e-6
    mlp = cm.MLP(Xtr.shape[1], 64, seed=args.cm_seed)
    t_train = time.time()
    loss = mlp.train((Xtr - mu) / sd, (ytr - ymu) / ysd, epochs=args.cm_epochs, lr=3e-3)
    t_train = time.time() - t_train

    def regret(ids, pick):
        out = []
        for i in ids:
            r = recs[i]
            valid = [(v, t) for v, t in zip(r["variants"], r["ms"]) if t > 0]
            best = min(t for _, t in valid)
            out.append(pick(r, valid) / best - 1.0)
        return np.array(out)

    def model_pick(r, valid):
        X = np.array([cm.features(r, p, b) for (p, b), _ in valid], np.float32)
        pred = mlp.forward((X - mu) / sd)
        return valid[int(np.argmin(pred))][1]

    def default_pick(r, valid):
        want = (cm.DIRECT, 256) if r["access"] == "read" else (op2amd.api.PLAN_STAGED, 256)
        return dict(valid)[want]
    t_inf = time.time()
    reg_model = regret(test_ids, model_pick)
    t_inf = (time.time() - t_inf) / len(test_ids)
    reg_default = regret(test_ids, default_pick)
    reg_model_train = regret(train_ids, model_pick)
    line = {
        "metric": "cost-model plan selection: mean time regret vs exhaustive sweep, held-out loops",
        "value": round(float(reg_model.mean()), 4), "unit": "fraction", "n_gpus": 1,
        "steps": len(recs), "warmup": 0, "ms_per_step": None, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (model), f32/f64 (loops)",
        "data": "synthetic",
        "config": {"workload": f"{len(recs)} synthetic op_par_loop descriptors, 80/20 split",
                   "train_loops": int(len(train_ids)), "test_loops": int(len(test_ids)),
                   "train_rows": int(Xtr.shape[0]), "final_train_loss": loss,
                   "regret_model_mean": float(reg_model.mean()),
                   "regret_model_p90": float(np.percentile(reg_model, 90)),
                   "regret_model_within_5pct": float(np.mean(reg_model <= 0.05)),
                   "regret_model_train_mean": float(reg_model_train.mean()),
                   "regret_default_plan_mean": float(reg_default.mean()),
                   "regret_default_plan_p90": float(np.percentile(reg_default, 90)),
                   "sweep_seconds": t_sweep, "train_seconds": t_train,
                   "model_pick_seconds_per_loop": t_inf},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="airfoil",
                    choices=["airfoil", "airfoil-large", "ensemble", "costmodel"],
                    help="airfoil: the headline 720k line (default); the others "
                         "measure the 11.5M-cell FP64/mixed config and the ensemble")
    ap.add_argument("--large-mesh", default="4800x2400")
    ap.add_argument("--large-iters", type=int, default=2000)
    ap.add_argument("--ensemble-size", type=int, default=256)
    ap.add_argument("--ensemble-iters", type=int, default=200)
    ap.add_argument("--cm-loops", type=int, default=200)
    ap.add_argument("--cm-epochs", type=int, default=3000)
    ap.add_argument("--cm-seed", type=int, default=0)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--mesh", default="1200x600")
    ap.add_argument("--ordering", default="rcm", choices=["random", "rcm"],
                    help="input ordering of the headline measurement")
    ap.add_argument("--plans", default=None,
                    help="skip the sweep, e.g. res=color:256,bres=atomic:128,direct=1024")
    ap.add_argument("--only-headline", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sweep-iters", type=int, default=6)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the single-mesh path has no multi-GPU sharding in this build)")
        sys.exit(2)

    import op2amd
    if args.workload == "airfoil-large":
        return workload_large(op2amd, args)
    if args.workload == "ensemble":
        return workload_ensemble(op2amd, args)
    if args.workload == "costmodel":
        return workload_costmodel(op2amd, args)
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

    log(f"headline: {args.ordering} ordering")
    head = measure(op2amd, mesh, K, W, args, plans=plans)
    value = head["graphed_steps_per_s"]
    graphed = head.pop("_graphed")

    others = {}
    if not args.only_headline:
        other = "random" if args.ordering == "rcm" else "rcm"
        log(f"{other} ordering, as given")
        others[f"{other}_ordering"] = measure(op2amd, meshes[other], K, W, args)
        others[f"{other}_ordering"].pop("_graphed")
        if other == "random":
            log("random ordering after op_renumber(pecell)")
            r = measure(op2amd, meshes["random"], K, W, args, renumber=True)
            r.pop("_graphed")
            others["random_ordering_op_renumber"] = r

    # parity of the headline against the oracle's OpenMP executor
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(os.cpu_count() or 1, 16)
    ncheck = 5
    orc = oracle.airfoil_run(mesh, ncheck, 1, "openmp", nthreads=nthreads)
    err_vs_oracle = rel(graphed["rms"][:ncheck], orc["rms"])

    cpu = None
    if not args.no_cpu:
        t = time.time()
        oracle.airfoil_run(mesh, 1, 1, "openmp", nthreads=nthreads)
        one = max(time.time() - t, 1e-3)
        iters = int(max(2, min(200, 15.0 / one)))
        t = time.time()
        oracle.airfoil_run(mesh, iters, 1, "openmp", nthreads=nthreads)
        secs = time.time() - t
        cpu = {"value": round(iters / secs, 3), "unit": "steps/s", "cores": nthreads,
               "kind": "port",
               "sample": f"{iters} airfoil steps on the same {I}x{J} {args.ordering} mesh "
                         f"({secs:.1f} s): the oracle's OpenMP executor (block colours "
                         "in sequence, blocks in parallel), g++ -O3 -fopenmp"}
        log(f"cpu baseline: {cpu['value']:.2f} steps/s on {nthreads} threads")

    dom = head["per_loop"]["res_calc"]
    kw = head["tuned_plans"]
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
                        f"{args.ordering} ordering, FP64; one step = save_soln + 2 x "
                        "(adt_calc, res_calc, bres_calc, update)",
            "mesh": args.mesh, "ordering": args.ordering,
            "rms_rel_err_vs_oracle_first_steps": err_vs_oracle,
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
            "traffic": None,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
