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
    """The batched run with its block sizes picked by timing the first five
    steps (one candidate per step) gives the bits of the default launch
    shapes; the picks are candidates."""
    sp = op2amd.ensemble.spec(**SMALL)
    plain = op2amd.ensemble.run(sp, 30, batched=True, graphed=True, want_q=True)
    tuned = op2amd.ensemble.run(sp, 30, batched=True, graphed=True, want_q=True,
                                autotune=True)
    for key in ("rms", "maxerr", "errloc", "q"):
        assert np.array_equal(tuned[key], plain[key]), key
    assert plain["blocks"] == {"save_soln": 256, "adt_calc": 1024, "res_calc": 128,
                               "bres_calc": 128}
    assert set(tuned["blocks"].values()) <= {64, 128, 256, 512, 1024}
    assert tuned["blocks"]["res_calc"] == 128
