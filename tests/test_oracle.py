"""Pins the oracle (oracle/op2_oracle.cpp) against the reference's own code.

Golden vectors come from the reference's sequential executor and airfoil
kernels compiled from their sources (tests/golden/make_golden.py).  What they
leave open is pinned by closed forms and invariants of the discretisation.
"""
import math

import numpy as np
import pytest

import op2amd
from conftest import golden, rel_err
from oracle import oracle


def mesh_of(g):
    return op2amd.mesh.generate(int(g["I"]), int(g["J"]), 0, str(g["ordering"]))


@pytest.mark.parametrize("name", ["airfoil_120x60_random_a1", "airfoil_120x60_rcm_a1",
                                  "airfoil_24x12_random_a1", "airfoil_24x12_structured_a0"])
def test_seq_oracle_bit_exact_with_reference(name):
    g = golden(name)
    m = mesh_of(g)
    r = oracle.airfoil_run(m, int(g["niter"]), int(g["alpha_mode"]), "seq")
    assert np.array_equal(r["rms_k0"], g["rms_k0_raw"])
    assert np.array_equal(r["rms"], g["rms"])
    assert np.array_equal(r["maxerr"], g["maxerr"])
    assert np.array_equal(r["errloc"], g["errloc"])
    assert np.array_equal(r["q"], g["q"])


def test_seq_oracle_matches_unmodified_reference_app():
    g = golden("airfoil_24x12_random_app_a0")
    r = oracle.airfoil_run(mesh_of(g), 1000, 0, "seq")
    assert np.array_equal(r["q"], g["q"])


def test_single_loops_bit_exact_first_iteration():
    g = golden("airfoil_24x12_random_a1")
    m = mesh_of(g)
    c = oracle.consts(1)
    q = np.tile(np.array(c.qinf[:]), (m.ncell, 1))
    adt = oracle.adt_calc(m, q)
    assert np.array_equal(adt, g["first_adt"])
    res = oracle.res_calc(m, q, adt, np.zeros((m.ncell, 4)))
    assert np.array_equal(res, g["first_res_after_res"])
    res = oracle.bres_calc(m, q, adt, res)
    assert np.array_equal(res, g["first_res_after_bres"])
    qn, res2, rms, mx, loc = oracle.update(q, res, adt)
    assert np.array_equal(qn, g["first_q_after_update"])
    assert np.all(res2 == 0.0)


def test_openmp_oracle_close_to_seq_and_thread_count_independent():
    g = golden("airfoil_120x60_random_a1")
    m = mesh_of(g)
    runs = [oracle.airfoil_run(m, 200, 1, "openmp", nthreads=t) for t in (1, 3, 8)]
    for r in runs[1:]:
        # colour-ordered increments do not depend on the thread count
        assert np.array_equal(r["q"], runs[0]["q"])
    seq = oracle.airfoil_run(m, 200, 1, "seq")
    assert rel_err(runs[0]["q"], seq["q"]) < 1e-12
    assert np.max(np.abs(runs[0]["rms"] / seq["rms"] - 1)) < 1e-10


@pytest.mark.parametrize("fixture", ["plan_60x30_random_p0", "plan_60x30_rcm_p64"])
@pytest.mark.parametrize("loop", ["adt_calc", "res_calc", "bres_calc"])
def test_plan_restatement_equals_reference_plan(fixture, loop):
    g = golden(fixture)
    m = op2amd.mesh.generate(int(g["I"]), int(g["J"]), 0, str(g["ordering"]))
    n, args = oracle.airfoil_plan_args(m, loop)
    p = oracle.plan_build(n, args, int(g["part_size"]))
    assert p["nblocks"] == int(g[f"{loop}_nblocks"])
    assert p["ncolors"] == int(g[f"{loop}_ncolors"])
    for key in ("offset", "nelems", "nthrcol", "thrcol", "ncolblk", "blkmap",
                "nindirect", "ind_sizes", "ind_offs"):
        assert np.array_equal(p[key], g[f"{loop}_{key}"]), key
    for i in range(p["ninds"]):
        assert np.array_equal(p["ind_maps"][i], g[f"{loop}_ind_map{i}"])
    for k, lm in enumerate(p["loc_maps"]):
        if lm is not None:
            assert np.array_equal(lm.astype(np.int32), g[f"{loop}_loc_map{k}"])


def test_res_calc_conserves():
    """Each edge adds +f to one cell and -f to the other (res_calc.h:20-32):
    the sum over cells vanishes up to rounding."""
    m = op2amd.mesh.generate(60, 30, 3, "random")
    rng = np.random.default_rng(5)
    c = oracle.consts(1)
    q = np.tile(np.array(c.qinf[:]), (m.ncell, 1)) * (1 + 0.05 * rng.standard_normal((m.ncell, 4)))
    adt = oracle.adt_calc(m, q)
    res = oracle.res_calc(m, q, adt, np.zeros((m.ncell, 4)))
    total = res.sum(axis=0)
    scale = np.abs(res).sum(axis=0)
    assert np.all(np.abs(total) <= 1e-13 * scale)


def test_free_stream_is_preserved_with_far_field_everywhere():
    m = op2amd.mesh.generate(40, 20, 1, "random")
    m.bound[:] = 2
    c = oracle.consts(1)
    q = np.tile(np.array(c.qinf[:]), (m.ncell, 1))
    adt = oracle.adt_calc(m, q)
    res = oracle.bres_calc(m, q, adt, oracle.res_calc(m, q, adt, np.zeros((m.ncell, 4))))
    assert np.max(np.abs(res)) < 1e-14


def test_adt_closed_form_on_square_cell():
    """adt = (sum over faces |u dy - v dx| + c |face|) / cfl; on an axis-aligned
    square of side h: (2h(|u|+|v|) + 4ch) / cfl."""
    h = 0.125
    x = np.array([[0, 0], [h, 0], [h, h], [0, h]], dtype=np.float64)
    m = op2amd.mesh.Mesh(4, 1, 0, 4, x=x, cell=np.array([[0, 1, 2, 3]]),
                         edge=np.zeros((0, 2)), ecell=np.zeros((0, 2)),
                         bedge=np.array([[1, 0], [2, 1], [3, 2], [0, 3]]),
                         becell=np.zeros(4), bound=np.full(4, 2))
    c = oracle.consts(1)
    q = np.array([c.qinf[:]])
    adt = oracle.adt_calc(m, q)
    u, v = q[0, 1] / q[0, 0], q[0, 2] / q[0, 0]
    cs = math.sqrt(c.gam * c.gm1 * (q[0, 3] / q[0, 0] - 0.5 * (u * u + v * v)))
    expect = (2 * h * (abs(u) + abs(v)) + 4 * cs * h) / c.cfl
    assert abs(adt[0] - expect) <= 4e-16 * expect


def test_update_rms_matches_compensated_sum():
    rng = np.random.default_rng(2)
    n = 5000
    qold = 1 + rng.random((n, 4))
    res = rng.standard_normal((n, 4)) * 1e-3
    adt = 1 + rng.random(n)
    _, _, rms, mx, loc = oracle.update(qold, res, adt)
    dels = ((1.0 / adt)[:, None] * res) ** 2
    assert abs(rms - math.fsum(dels.ravel())) <= 1e-14 * rms
    assert mx == dels.max()
    assert loc == int(np.argmax(dels.max(axis=1) == mx))


def test_colouring_checker_rejects_a_planted_conflict():
    g = golden("plan_60x30_random_p0")
    m = op2amd.mesh.generate(int(g["I"]), int(g["J"]), 0, str(g["ordering"]))
    thrcol = g["res_calc_thrcol"].copy()
    blk = np.repeat(np.arange(len(g["res_calc_nelems"])), g["res_calc_nelems"])
    assert oracle.colouring_conflicts(m.ecell, thrcol, blk) == 0
    # plant: give an edge the colour of another edge of its block sharing a cell
    e0 = 0
    c0 = set(m.ecell[e0])
    e1 = next(e for e in range(1, int(g["res_calc_nelems"][0]))
              if c0 & set(m.ecell[e]))
    thrcol[e1] = thrcol[e0]
    assert oracle.colouring_conflicts(m.ecell, thrcol, blk) > 0


@pytest.mark.parametrize("I,J,ordering,niter,alpha", [(120, 60, "random", 200, 1),
                                                      (120, 60, "rcm", 200, 1),
                                                      (24, 12, "random", 300, 1),
                                                      (24, 12, "structured", 50, 0)])
@pytest.mark.parametrize("nthreads", [1, 3, 8])
def test_seqpar_is_the_sequential_executor_bit_for_bit(I, J, ordering, niter, alpha, nthreads):
    """The parallel evaluation of the sequential executor ("seqpar": each
    edge's increments evaluated once, each cell adding its own in edge order,
    reductions summed serially) gives the sequential executor's bits: every
    per-iteration output and the final q."""
    m = op2amd.mesh.generate(I, J, 0, ordering)
    a = oracle.airfoil_run(m, niter, alpha, "seq")
    b = oracle.airfoil_run(m, niter, alpha, "seqpar", nthreads=nthreads)
    for k in ("rms_k0", "rms", "maxerr", "errloc", "q"):
        assert np.array_equal(a[k], b[k]), k


def test_seqpar_matches_unmodified_reference_app():
    g = golden("airfoil_24x12_random_app_a0")
    r = oracle.airfoil_run(mesh_of(g), 1000, 0, "seqpar", nthreads=4)
    assert np.array_equal(r["q"], g["q"])
