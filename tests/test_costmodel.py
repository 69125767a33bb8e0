"""The MLP cost model's kernels against the float64 numpy oracle, and the
synthetic loops behind its training data."""
import numpy as np
import pytest

import op2amd
from oracle import mlp_ref

NIN, H = 19, 64


def batch(n=300, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, NIN)).astype(np.float32)
    t = (x[:, 0] - 0.5 * x[:, 3] + np.sin(x[:, 5])).astype(np.float32)
    return x, t


def test_oracle_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = rng.standard_normal(NIN * H + H + H * H + H + H + 1) * 0.2
    x, t = batch(40)
    _, g = mlp_ref.loss_grad(p, x, t, NIN, H)
    for i in rng.choice(len(p), 60, replace=False):
        e = np.zeros_like(p)
        e[i] = 1e-6
        fd = (mlp_ref.loss_grad(p + e, x, t, NIN, H)[0] -
              mlp_ref.loss_grad(p - e, x, t, NIN, H)[0]) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(g[i]), i


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 37, 300, 1100])
def test_hip_mlp_forward_and_gradients_match_oracle(n):
    """Forward pass and gradients on the matrix cores against the float64
    oracle: one sample, sizes off the 16-row tiles and the 4-deep steps, and
    several 256-sample splits of the weight gradients."""
    m = op2amd.costmodel.MLP(NIN, H, seed=7)
    p = m.params()
    x, t = batch(n)
    y = m.forward(x)
    y_ref, _ = mlp_ref.forward(p, x, NIN, H)
    assert np.max(np.abs(y - y_ref)) <= 1e-4 * max(1.0, np.max(np.abs(y_ref)))
    loss, g = m.loss_grad(x, t)
    loss_ref, g_ref = mlp_ref.loss_grad(p, x, t, NIN, H)
    assert abs(loss - loss_ref) <= 1e-5 * loss_ref
    assert np.max(np.abs(g - g_ref)) <= 1e-4 * np.max(np.abs(g_ref))
    _, g2 = m.loss_grad(x, t)
    assert np.array_equal(g, g2)  # split partials summed in a fixed order


@pytest.mark.gpu
def test_hip_mlp_training_fits():
    m = op2amd.costmodel.MLP(NIN, H, seed=3)
    x, t = batch(2000)
    l0, _ = m.loss_grad(x, t)
    l1 = m.train(x, t, epochs=400, lr=3e-3)
    assert l1 < 0.05 * l0


@pytest.mark.gpu
@pytest.mark.parametrize("is_double,tol", [(1, 1e-12), (0, 2e-5)])
@pytest.mark.parametrize("arity", [2, 4])
def test_synthetic_inc_loop_agrees_across_plans(is_double, tol, arity):
    cm = op2amd.costmodel
    d = cm.desc_from(3000, arity, 4, "inc", is_double, "rcm", 12, 5)
    size = 0
    outs = {}
    for plan, block in [(cm.PLAN_ATOMIC, 128), (cm.PLAN_COLOR, 256), (cm.PLAN_STAGED, 128),
                        (cm.PLAN_GATHER, 64), (cm.PLAN_OWNER, 128)]:
        if not size:
            _, shape = cm.time_variants(d, [(plan, block)], reps=1)
            size = shape["target_size"] * 4
        outs[plan] = cm.run_once(d, plan, block, size)
    ref = outs[cm.PLAN_GATHER]
    for plan, o in outs.items():
        assert np.max(np.abs(o - ref)) <= tol * np.max(np.abs(ref)), plan
    # owner and gather both add each target's increments in element order
    assert np.array_equal(outs[cm.PLAN_OWNER], ref)


@pytest.mark.gpu
def test_synthetic_variants_time_and_refuse_invalid():
    cm = op2amd.costmodel
    d = cm.desc_from(5000, 2, 2, "rw", 0, "random", 8, 1)
    ms, shape = cm.time_variants(d, [(cm.PLAN_COLOR, 128), (cm.PLAN_ATOMIC, 128),
                                     (cm.PLAN_STAGED, 256), (cm.PLAN_GATHER, 128)], reps=2)
    assert ms[0] > 0 and ms[2] > 0 and ms[3] > 0
    assert ms[1] == -1.0  # atomics cannot implement OP_RW
    assert 4000 <= shape["set_size"] <= 6000


@pytest.mark.gpu
def test_model_picked_plans_beat_the_static_default():
    """Small version of the plan-selection experiment: the cost model trained
    on 24 swept loops picks faster variants on 8 held-out loops than the
    runtime's static default (the reference's staged plan, block 256)."""
    cm = op2amd.costmodel
    recs = cm.sweep(cm.random_descs(32, seed=3, n_lo=2e4, n_hi=3e5), reps=3)
    ids = np.random.default_rng(0).permutation(len(recs))
    test, train = list(ids[:8]), list(ids[8:])
    model = cm.CostModel(members=3, seed=1).fit(recs, train, epochs=1500)
    r_model = cm.regret(recs, test, model.choose)
    r_default = cm.regret(recs, test, cm.default_choice)
    assert np.all(r_model >= 0) and np.all(r_default >= 0)
    assert r_model.mean() < r_default.mean()
    # on its own training loops the model is near the sweep's best
    assert cm.regret(recs, train, model.choose).mean() < 0.1


@pytest.mark.gpu
def test_runtime_picks_each_loop_plan_with_the_model():
    """An installed model picks each AUTO loop's plan at the loop's first
    call, on the device: the runtime describes res_calc as this module does
    (sizes, arity, dim, precision, access, arithmetic, the map's mean
    bandwidth), its pick is the model's best prediction over the variants
    res_calc implements, and the run keeps the reference trajectory."""
    from conftest import golden, pointwise_rel_err
    cm = op2amd.costmodel
    recs = cm.sweep(cm.random_descs(12, 3, n_lo=1e4, n_hi=6e4), reps=1)
    model = cm.CostModel(members=2, seed=1).fit(recs, list(range(len(recs))), epochs=300)
    g = golden("airfoil_120x60_random_a1")
    mesh = op2amd.mesh.generate(int(g["I"]), int(g["J"]), 0, str(g["ordering"]))
    model.install(top=1)
    try:
        r = op2amd.airfoil.run(mesh, 200, mode="tuned")
        d = cm.model_loop("res_calc")
        dd = cm.model_loop("save_soln")
    finally:
        cm.CostModel.uninstall()
    e = mesh.ecell.astype(np.int64)
    assert (d["set_size"], d["target_size"], d["arity"], d["dim"], d["is_double"], d["access"],
            d["k"]) == (mesh.nedge, mesh.ncell, 2, 4, 1, "inc", 60)
    assert abs(d["mean_bandwidth"] - np.abs(e[:, 0] - e[:, 1]).mean()) <= 1e-9 * d["mean_bandwidth"]
    assert (dd["access"], dd["target_size"], dd["mean_bandwidth"]) == ("read", mesh.ncell, 0.0)
    cands = [(p, b) for p in (cm.PLAN_ATOMIC, cm.PLAN_COLOR, cm.PLAN_STAGED, cm.PLAN_GATHER,
                              cm.PLAN_OWNER)
             for b in cm.BLOCKS
             if not (p == cm.PLAN_STAGED and not 128 <= b <= 512) and not (p == cm.PLAN_OWNER and b > 256)]
    pred = model.predict([cm.features(d, p, b) for p, b in cands])
    chosen = tuple(r["chosen"]["res_calc"])
    assert chosen in cands
    assert pred[cands.index(chosen)] <= pred.min() + 1e-4
    assert r["chosen"]["save_soln"][1] in cm.BLOCKS
    assert pointwise_rel_err(r["rms"], g["rms"][:200]) <= 1e-10
