"""The benchmark configurations at their full sizes, on the GPU, against the
oracle:

B  1200x600 (720k cells), random and RCM numberings, 1000 iterations: the
   benchmark's graphed plans and the plain staged backend against the
   oracle's OpenMP executor (rms and the first update's raw rms at every
   iteration, final q, 1e-10 relative); the owner + gather plans bit for bit
   against the sequential executor (q, maxerr and errloc at every iteration).
C  4800x2400 (11.5M cells): one call of every loop against the sequential
   oracle (bit for bit with owner and gather increments), then FP64 against
   the FP32-flux mixed mode over 2000 iterations, with the measured error
   bounds.
D  the 256-member ensemble spec (40x20 .. 640x320, per-member Mach and
   alpha), 10 steps: batched (one launch per loop, graph-captured) equals one
   launch per member bit for bit; the smallest, the largest and six seeded
   members equal the oracle's sequential run of that member.
E  200 synthetic loop descriptors, set sizes log-uniform in 1e4..1e7: for
   each increment (OP_INC) descriptor, every plan's output agrees with the
   gather plan's (1e-12 FP64, 2e-5 FP32), and the owner plan's bit for bit.
   OP_READ descriptors have one plan (direct); an OP_RW through a map
   (w = 0.5 w + v) depends on the order the elements reach a target, which
   the reference leaves to the colouring, so its plans are not compared.

The sequential executor's results at 720k come from the oracle's "seqpar"
evaluation (bit for bit the sequential executor's, pinned in
tests/test_oracle.py)."""
import ctypes as C
import os

import numpy as np
import pytest

import op2amd
import op2amd.api as api
from conftest import pointwise_rel_err, rel_err
from oracle import oracle
from test_gpu_loops import App, inputs

pytestmark = pytest.mark.gpu
TOL = 1e-10
NT = min(int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1), 16)
HEADLINE = dict(plan_res="owner", block_res=128, plan_bres="atomic", block_bres=128,
                block_direct=512, block_adt=512)


@pytest.mark.parametrize("ordering", ["rcm", "random"])
def test_config_b_720k_1000_iterations(ordering):
    mesh = op2amd.mesh.generate(1200, 600, 0, ordering)
    niter = 1000
    print(f"config B {ordering}: oracle runs on {NT} threads", flush=True)
    omp = oracle.airfoil_run(mesh, niter, 1, "openmp", nthreads=NT)
    seq = oracle.airfoil_run(mesh, niter, 1, "seqpar", nthreads=NT)
    runs = {
        "graphed": op2amd.airfoil.run(mesh, niter, mode="graphed", warmup=20, graph_steps=10,
                                      **HEADLINE),
        "plain": op2amd.airfoil.run(mesh, niter, mode="plain"),
    }
    for name, r in runs.items():
        assert pointwise_rel_err(r["rms"], omp["rms"]) <= TOL, name
        assert pointwise_rel_err(r["rms_k0"], omp["rms_k0"]) <= TOL, name
        assert pointwise_rel_err(r["maxerr"], omp["maxerr"]) <= TOL, name
        assert rel_err(r["q"], omp["q"]) <= TOL, name
    exact = op2amd.airfoil.run(mesh, niter, mode="graphed", warmup=20, graph_steps=10,
                               plan_res="owner", block_res=128, plan_bres="gather")
    assert exact["graph_steps"] == 10
    assert np.array_equal(exact["q"], seq["q"])
    assert np.array_equal(exact["maxerr"], seq["maxerr"])
    assert np.array_equal(exact["errloc"], seq["errloc"])
    # rms: the same squares summed by a tree instead of in element order
    # (measured 2.7e-13 over 2.88M terms)
    assert pointwise_rel_err(exact["rms"], seq["rms"]) <= 2e-12
    assert pointwise_rel_err(exact["rms_k0"], seq["rms_k0"]) <= 2e-12


@pytest.fixture(scope="module")
def mesh_c():
    return op2amd.mesh.generate(4800, 2400, 0, "rcm")


@pytest.mark.parametrize("plan,block", [("owner", 256), ("gather", 256)])
def test_config_c_11m_loops_bit_exact(mesh_c, plan, block):
    mesh = mesh_c
    q, res0 = inputs(mesh, 6)
    app = App(mesh, q, res0)
    try:
        api.op2h_set_loop_plan(b"res_calc", api.PLAN_NAMES[plan], block)
        api.op2h_set_loop_plan(b"bres_calc", api.PLAN_NAMES["gather"], 128)
        app.adt_calc()
        adt = app.fetch(app.adt, (mesh.ncell,))
        app.res_calc()
        res1 = app.fetch(app.res, (mesh.ncell, 4))
        app.bres_calc()
        res2 = app.fetch(app.res, (mesh.ncell, 4))
        rms, maxerr, errloc = C.c_double(0.0), C.c_double(0.0), C.c_int(0)
        app.update(rms, maxerr, errloc)
        qn = app.fetch(app.q, (mesh.ncell, 4))
        rn = app.fetch(app.res, (mesh.ncell, 4))
    finally:
        app.close()
    adt_o = oracle.adt_calc(mesh, q)
    assert np.array_equal(adt, adt_o)
    res1_o = oracle.res_calc(mesh, q, adt_o, res0)
    assert np.array_equal(res1, res1_o)
    res2_o = oracle.bres_calc(mesh, q, adt_o, res1_o)
    assert np.array_equal(res2, res2_o)
    qn_o, rn_o, rms_o, mx_o, loc_o = oracle.update(q, res2_o, adt_o, 0.0, 0.0, 0)
    assert np.array_equal(qn, qn_o) and np.array_equal(rn, rn_o)
    # the same 46M squares summed by a tree (measured 2.8e-13 relative)
    assert abs(rms.value - rms_o) <= 2e-12 * rms_o
    assert maxerr.value == mx_o and errloc.value == loc_o


def test_config_c_11m_mixed_precision_2000_iterations(mesh_c):
    """FP32 fluxes, adt and res with FP64 q, update and reductions, against
    the FP64 run of the same plans (whose per-call results equal the
    sequential oracle's bit for bit, above) over 2000 iterations.  Measured on
    MI355X: rms 1.7e-6 (pointwise max over the iterations), 1.6e-6 at the
    last, q 1.9e-6 (relative to max |q|); the bounds leave a 5x margin."""
    common = dict(plan_res="owner", block_res=256, plan_bres="atomic", block_bres=128,
                  block_direct=1024, block_adt=1024, graph_steps=10)
    fp64 = op2amd.airfoil.run(mesh_c, 2000, mode="graphed", warmup=20, **common)
    mixed = op2amd.airfoil.run(mesh_c, 2000, mode="graphed", warmup=20, mixed=True, **common)
    e_rms = pointwise_rel_err(mixed["rms"], fp64["rms"])
    e_last = abs(mixed["rms"][-1] / fp64["rms"][-1] - 1)
    e_q = rel_err(mixed["q"], fp64["q"])
    print(f"config C mixed vs FP64: rms max {e_rms:.3e}, rms last {e_last:.3e}, q {e_q:.3e}")
    assert np.all(np.isfinite(mixed["rms"])) and np.all(np.isfinite(mixed["q"]))
    assert e_rms <= 1e-5
    assert e_last <= 1e-5
    assert e_q <= 1e-5


def test_config_d_256_member_ensemble():
    sp = op2amd.ensemble.spec()
    niter = 10
    batched = op2amd.ensemble.run(sp, niter, batched=True, graphed=True, want_q=True)
    looped = op2amd.ensemble.run(sp, niter, batched=False, want_q=True)
    members = batched["members"]
    assert len(members) == 256
    for key in ("rms", "maxerr", "errloc", "q"):
        assert np.array_equal(looped[key], batched[key]), key
    sizes = np.array([m["ncell"] for m in members])
    offs = np.concatenate([[0], np.cumsum(sizes)])
    rng = np.random.default_rng(11)
    pick = {int(np.argmin(sizes)), int(np.argmax(sizes))}
    pick |= set(int(i) for i in rng.choice(256, 6, replace=False))
    for m in sorted(pick):
        mb = members[m]
        mesh = op2amd.mesh.generate(mb["I"], mb["J"], sp.seed + 1 + m, "rcm")
        ref = oracle.airfoil_run_qinf(mesh, niter, mb["qinf"])
        assert np.array_equal(batched["q"][offs[m]:offs[m + 1]], ref["q"]), m
        assert np.array_equal(batched["maxerr"][:, m], ref["maxerr"]), m
        assert np.array_equal(batched["errloc"][:, m], ref["errloc"]), m
        assert np.max(np.abs(batched["rms"][:, m] / ref["rms"] - 1)) <= 1e-12, m


@pytest.mark.timeout(900)
def test_config_e_200_descriptors_plans_agree():
    """Every plan of every increment descriptor at block sizes 128 and 256
    against the gather plan (each target's increments in element order).
    Meshes are generated on host threads ahead of the GPU checks."""
    from concurrent.futures import ThreadPoolExecutor
    cm = op2amd.costmodel
    descs = cm.random_descs(200, seed=0, n_lo=1e4, n_hi=1e7)
    ds = [cm.desc_from(p["n"], p["arity"], p["dim"], p["access"], p["is_double"],
                       p["ordering"], p["k"], p["seed"]) for p in descs]
    sizes = [(d.I * d.J * (2 if d.arity == 2 else 1)) for d in ds]
    assert len(ds) == 200 and min(sizes) < 2e4 and max(sizes) > 5e6
    inc = [(p, d) for p, d in zip(descs, ds) if p["access"] == "inc"]
    bad = []
    checked = 0
    ahead = 3
    with ThreadPoolExecutor(max_workers=ahead) as pool:
        futs = [pool.submit(cm.desc_mesh, d) for _, d in inc[:ahead]]
        for i, (p, d) in enumerate(inc):
            mesh = futs[i].result()
            futs[i] = None
            if i + ahead < len(inc):
                futs.append(pool.submit(cm.desc_mesh, inc[i + ahead][1]))
            var = [(v, b) for v, b in cm.variants(p["access"]) if b in (128, 256)]
            out = cm.check_plans(d, var, mesh=mesh)
            tol = 1e-12 if p["is_double"] else 2e-5
            for (plan, block), (dev, mis) in zip(var, out):
                if dev < 0:
                    assert plan in (cm.PLAN_STAGED, cm.PLAN_OWNER), (i, plan, block)
                    continue
                checked += 1
                if not dev <= tol or (plan == cm.PLAN_OWNER and mis != 0):
                    bad.append((i, p, plan, block, dev, mis))
            del mesh
    print(f"config E: {len(inc)} increment descriptors, {checked} plan outputs checked, "
          f"{len(bad)} outside tolerance")
    assert len(inc) >= 40 and checked >= 8 * len(inc) - 20
    assert not bad, bad[:5]
