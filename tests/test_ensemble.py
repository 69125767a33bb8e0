"""Ensemble member definitions (CPU) and the batched ensemble on the GPU."""
import math

import numpy as np
import pytest

import op2amd
from oracle import oracle

SMALL = dict(nmesh=5, min_IJ=(8, 4), max_IJ=(40, 20), seed=3, ordering="rcm")


def test_members_deterministic_and_in_range():
    sp = op2amd.ensemble.spec()
    a = op2amd.ensemble.describe(sp)
    b = op2amd.ensemble.describe(op2amd.ensemble.spec())
    assert a == b and len(a) == 256
    for m in a:
        assert 40 <= m["I"] <= 640 and 20 <= m["J"] <= 320 and abs(m["I"] - 2 * m["J"]) <= 1
        assert 0.3 <= m["mach"] <= 0.6 and 0.0 <= m["alpha_deg"] <= 5.0
        u = math.hypot(m["qinf"][1], m["qinf"][2])
        assert abs(u - math.sqrt(1.4 * 1.0) * m["mach"]) < 1e-6
        assert abs(math.degrees(math.atan2(m["qinf"][2], m["qinf"][1])) - m["alpha_deg"]) < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("ordering", ["rcm", "random"])
def test_batched_equals_looped_and_each_member_equals_sequential_run(ordering):
    """Batched (graphed and eager) and looped runs give the same bits, and each
    member equals the oracle's sequential run of that member.  With random
    member numberings the owner patches span more than 16-bit offsets and
    read 16-byte instance records; with RCM they fit the 8-byte form."""
    # random: two 400x200 members (80,000 cells, spans beyond 65535)
    cfg = SMALL if ordering == "rcm" else dict(nmesh=2, min_IJ=(400, 200), max_IJ=(400, 200),
                                               seed=3, ordering="random")
    sp = op2amd.ensemble.spec(**cfg)
    niter = 60
    batched = op2amd.ensemble.run(sp, niter, batched=True, graphed=True, want_q=True)
    eager = op2amd.ensemble.run(sp, niter, batched=True, graphed=False, want_q=True)
    looped = op2amd.ensemble.run(sp, niter, batched=False, want_q=True)
    for r in (eager, looped):
        for key in ("rms", "maxerr", "errloc", "q"):
            assert np.array_equal(r[key], batched[key]), key
    off = 0
    for m, mb in enumerate(batched["members"]):
        mesh = op2amd.mesh.generate(mb["I"], mb["J"], cfg["seed"] + 1 + m, ordering)
        ref = oracle.airfoil_run_qinf(mesh, niter, mb["qinf"])
        q = batched["q"][off:off + mesh.ncell]
        off += mesh.ncell
        assert np.array_equal(q, ref["q"]), m
        assert np.array_equal(batched["maxerr"][:, m], ref["maxerr"]), m
        assert np.array_equal(batched["errloc"][:, m], ref["errloc"]), m
        # rms: the same terms summed in another order (per-tile partials)
        assert np.max(np.abs(batched["rms"][:, m] / ref["rms"] - 1)) <= 1e-12, m


@pytest.mark.gpu
def test_autotuned_block_sizes_change_no_bits():
    """The batched run with its workgroup sizes picked by timing the first
    ten steps (one candidate per step, each twice) gives the bits of the default launch
    shapes; the picks are candidates."""
    sp = op2amd.ensemble.spec(**SMALL)
    plain = op2amd.ensemble.run(sp, 30, batched=True, graphed=True, want_q=True)
    tuned = op2amd.ensemble.run(sp, 30, batched=True, graphed=True, want_q=True,
                                autotune=True)
    for key in ("rms", "maxerr", "errloc", "q"):
        assert np.array_equal(tuned[key], plain[key]), key
    # bres_calc runs inside res_calc's launch: no workgroup size of its own
    assert plain["blocks"] == {"update_adt": 256, "res_calc": 256, "bres_calc": 0,
                               "update_save_adt": 256}
    assert plain["fused_bres"] and tuned["fused_bres"]
    assert set(tuned["blocks"].values()) <= {0, 64, 128, 256, 512, 1024}
    assert tuned["blocks"]["res_calc"] == 256
    assert plain["dead_res"] and tuned["dead_res"]


@pytest.mark.gpu
@pytest.mark.parametrize("max_workgroups", [1, 3, 7])
def test_persistent_workgroups_walking_many_items_give_the_same_bits(max_workgroups):
    """Grids capped far below the work items (each workgroup walks many tiles,
    patches and boundary runs of several members) give the bits of the
    uncapped batched run, of the looped run and of the oracle."""
    cfg = dict(nmesh=4, min_IJ=(40, 20), max_IJ=(120, 60), seed=5, ordering="rcm")
    sp = op2amd.ensemble.spec(**cfg)
    niter = 12
    ref = op2amd.ensemble.run(sp, niter, batched=False, want_q=True)
    capped = op2amd.ensemble.run(sp, niter, batched=True, graphed=True, want_q=True,
                                 max_workgroups=max_workgroups)
    assert max(capped["grids"].values()) <= max_workgroups
    for key in ("rms", "maxerr", "errloc", "q"):
        assert np.array_equal(capped[key], ref[key]), key
    m = 1
    mb = capped["members"][m]
    off = capped["members"][0]["ncell"]
    mesh = op2amd.mesh.generate(mb["I"], mb["J"], cfg["seed"] + 1 + m, "rcm")
    want = oracle.airfoil_run_qinf(mesh, niter, mb["qinf"])
    assert np.array_equal(capped["q"][off:off + mesh.ncell], want["q"])


@pytest.mark.gpu
@pytest.mark.parametrize("max_workgroups", [0, 5])
def test_bres_calc_fused_into_res_calc_gives_the_same_bits(max_workgroups):
    """bres_calc fused into res_calc's owner launch (each boundary cell adds
    its boundary edges after its interior ones) and bres_calc in launches of
    its own, one wave per run of boundary cells, give the same bits, with the
    tuned workgroup sizes and with grids capped so that workgroups walk many
    items."""
    sp = op2amd.ensemble.spec(nmesh=5, min_IJ=(40, 20), max_IJ=(160, 80), seed=9)
    kw = dict(batched=True, graphed=True, want_q=True, max_workgroups=max_workgroups,
              autotune=max_workgroups == 0)
    fused = op2amd.ensemble.run(sp, 20, **kw)
    apart = op2amd.ensemble.run(sp, 20, separate_bres=True, **kw)
    assert fused["fused_bres"] and not apart["fused_bres"]
    assert fused["blocks"]["bres_calc"] == 0
    assert apart["blocks"]["bres_calc"] in (64, 128, 256, 512, 1024)
    for key in ("rms", "maxerr", "errloc", "q"):
        assert np.array_equal(fused[key], apart[key]), key


@pytest.mark.gpu
@pytest.mark.parametrize("patch_cells", [64, 128, 1024])
def test_res_calc_patch_size_changes_no_bits(patch_cells):
    """Owner patches of 64, 128 or 1024 cells (one workgroup each) give the
    bits of the default 256-cell patches: each cell sums its edges' fluxes in
    edge order whatever patch holds it."""
    sp = op2amd.ensemble.spec(nmesh=4, min_IJ=(40, 20), max_IJ=(160, 80), seed=7)
    ref = op2amd.ensemble.run(sp, 10, batched=True, graphed=True, want_q=True)
    run = op2amd.ensemble.run(sp, 10, batched=True, graphed=True, want_q=True,
                              patch_cells=patch_cells)
    assert run["blocks"]["res_calc"] == patch_cells and ref["blocks"]["res_calc"] == 256
    for key in ("rms", "maxerr", "errloc", "q"):
        assert np.array_equal(run[key], ref[key]), key


@pytest.mark.gpu
def test_member_step_times_inside_the_batched_step():
    """Device-clock stamps of every work item give each member's latency
    inside the graphed batched step (from the step's first item to the
    member's last) and its busy time (its items' spans summed over the
    launches, which run one after another, so busy <= latency).  Stamping
    changes no bits, and the mean step span is within the wall time per step."""
    sp = op2amd.ensemble.spec(nmesh=12, min_IJ=(40, 20), max_IJ=(320, 160), seed=2)
    niter, W = 10, 2
    plain = op2amd.ensemble.run(sp, niter, warmup=W, batched=True, graphed=True, want_q=True)
    timed = op2amd.ensemble.run(sp, niter, warmup=W, batched=True, graphed=True, want_q=True,
                                member_timing=True)
    for key in ("rms", "maxerr", "errloc", "q"):
        assert np.array_equal(timed[key], plain[key]), key
    lat, busy = timed["member_step_ms"], timed["member_busy_ms"]
    assert np.all(lat > 0) and np.all(busy > 0)
    assert np.all(busy <= lat * (1 + 1e-9))
    wall_ms = 1e3 * timed["seconds"] / (niter - W)
    assert lat.max() <= wall_ms * 1.01, (lat.max(), wall_ms)
