"""GPU tests of batched stepping, schedules and checkpoints (through the C ABI).

* A batch of problems with different grids, moment orders, materials and
  filter strengths, stepped in lockstep by the persistent kernel (eagerly and
  as a replayed hipGraph), is bit-identical to stepping each handle alone.
* Every problem of an FP32 batch is within 1e-5 relative L2 per moment field
  of its FP64 oracle solve (fields below 1e-8 of the largest are measured
  against the largest).
* Changing the tile shape or LDS budget (smp_set_schedule) does not change a
  single bit of the result; checkpoints restore the state exactly; the
  autotuner leaves the solution unchanged.
"""
import numpy as np
import pytest

from oracle import starmap_oracle as so
from starmap_amd import autotune
from starmap_amd import init as I
from starmap_amd import problems as P
from starmap_amd.batch import BatchSolver, count_steps
from starmap_amd.solver import HipSolver
from tests.fp32_bounds import check_fp32_bound

pytestmark = pytest.mark.gpu


def small_sweep(n=5, side_div=8, orders=(3, 5, 7, 9), tfinal=0.25):
    # the sweep's seeded sides (divided by side_div), materials and filter
    # strengths, with moment orders cycling through `orders`
    out = []
    for i, d in enumerate(P.batch_draws(n_instances=n)):
        prob = P.batch_instance(max(4, d["side"] // side_div), orders[i % len(orders)], d["f_eff"],
                                d["sigma_a"], d["sigma_s0"], tfinal=tfinal)
        out.append(I.starmap_init(prob))
    return out


def states(s):
    return [s.get_state(j) for j in range(s.n_sys)]


def single_runs(pars, precision):
    out = []
    for p in pars:
        s = HipSolver(p, precision=precision)
        try:
            dt = s.prepare()
            s.step(dt, count_steps(dt, p["tfinal"]))
            out.append(states(s))
        finally:
            s.close()
    return out


@pytest.mark.parametrize("precision,cw", [("fp32", 4), ("fp64", 4), ("fp32", 8)])
@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_bitwise_equal_single_handles(precision, cw, use_graph):
    pars = small_sweep()
    print("batch", precision, cw, "graph" if use_graph else "eager",
          [(p["n"][1], p["n_mom"]) for p in pars], flush=True)
    ref = single_runs(pars, precision)
    b = BatchSolver(pars, precision=precision, cell_waves=cw)
    try:
        for rep in range(2):          # the second run replays the kept graph
            dts, ns = b.prepare()
            b.step(ns, use_graph=use_graph)
            for i, s in enumerate(b.solvers):
                for j, (x, y) in enumerate(zip(states(s), ref[i])):
                    assert np.array_equal(x, y), (rep, i, j)
        tot, ker, launches, by, mcu, grid = b.timing()
        # one launch per substep position, two where some problems take
        # their last (one-application) odd substep while others go on
        mixed = sum(1 for k in set(int(v) for v in ns) if k < int(max(ns)))
        assert launches == 2 * int(max(ns)) + 1 + mixed and grid >= 256 and tot > 0
        assert mcu == sum(s.n_sys * int(np.prod(s.n)) * int(n) for s, n in zip(b.solvers, ns))
    finally:
        b.close()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_batch_of_2d_and_3d_problems_bitwise_equal_single_handles(precision):
    """2D and 3D problems in one batch: the launch takes the records of the
    problem with the most terms (10 for 3D P7, against 8 for 2D P_N), so the
    2D problems run the wider kernel, whose extra terms add exact zeros; each
    problem equals its single-handle run bit for bit, as does a 3D problem
    with a non-cubic grid and extrapolation boundaries."""
    probs = [P.lattice_2d(n=40, n_mom=9, tfinal=0.2),
             P.gauss_periodic_3d(n=14, n_mom=7, tfinal=0.2, std=0.2),
             dict(P.gauss_periodic_3d(n=11, n_mom=3, tfinal=0.2, std=0.2), n=[11, 9, 7], bc=[1, 1, 1]),
             P.batch_instance(48, 15, 5.0, np.full((4, 4), 0.5), np.full((4, 4), 2.0), tfinal=0.2)]
    pars = [I.starmap_init(p) for p in probs]
    print("batch 2d+3d", precision, [tuple(p["n"]) for p in pars], flush=True)
    ref = single_runs(pars, precision)
    b = BatchSolver(pars, precision=precision)
    try:
        for use_graph in (False, True):
            dts, ns = b.prepare()            # from the initial state again
            b.step(ns, use_graph=use_graph)
            for i, s in enumerate(b.solvers):
                for j, (x, y) in enumerate(zip(states(s), ref[i])):
                    assert np.array_equal(x, y), (use_graph, i, j)
    finally:
        b.close()


@pytest.mark.parametrize("n,orders", [(4, (7, 3, 5)), (2, (31, 39))])
def test_fp32_batch_within_1e5_of_fp64_oracle(n, orders):
    pars = small_sweep(n=n, orders=orders, tfinal=0.2)
    b = BatchSolver(pars, precision="fp32")
    try:
        b.run()
        for i, (p, s) in enumerate(zip(pars, b.solvers)):
            print("oracle", i, p["n"][1], p["n_mom"], flush=True)
            # every moment field, whatever the problem's own mom_output
            ref, rpar = so.starmap_solver(dict(p, mom_output="all"), verbose=False, return_state=True)
            assert rpar["_oracle"]["step_count"] == b.n_steps[i]
            assert len(ref) == s.n_sys
            got = [s.get_state(j).reshape(np.shape(r["U"]), order="F") for j, r in enumerate(ref)]
            e = check_fp32_bound(got, [r["U"] for r in ref])
            print("  fp32 errors", e, flush=True)
    finally:
        b.close()


def test_schedule_change_keeps_every_bit():
    prob = P.lattice_2d(n=60, n_mom=9, tfinal=0.3)
    par = I.starmap_init(prob)
    res = []
    # 31x7: threads by slot column, a row of threads without a cell; 15x17:
    # 16x17 slot elements above 256, threads numbered along the tile
    for tile, bpc, cw in [(None, 0, 0), ((127, 2), 1, 8), ((31, 8), 2, 4), ((63, 4), 1, 8),
                          ((31, 7), 1, 8), ((15, 17), 2, 4)]:
        s = HipSolver(par, precision="fp64")
        try:
            dt = s.prepare()
            s.step(dt, 2)
            if tile is not None:
                s.set_schedule(tile, bpc, cw)
                assert s.get_schedule() == (tuple(tile) + (1,), bpc, cw)
            s.step(dt, 5)
            res.append(states(s))
        finally:
            s.close()
    for other in res[1:]:
        for x, y in zip(res[0], other):
            assert np.array_equal(x, y)


def test_checkpoint_round_trip_and_autotune_keep_state():
    par = I.starmap_init(P.gauss_periodic_3d(n=16, n_mom=3, tfinal=0.3, std=0.2))
    s = HipSolver(par, precision="fp32")
    r = HipSolver(par, precision="fp32")
    try:
        dt = s.prepare()
        r.prepare()
        s.step(dt, 2)
        s.checkpoint_save(3)
        a = states(s)
        s.step(dt, 3)
        s.checkpoint_load(3)
        assert all(np.array_equal(x, y) for x, y in zip(a, states(s)))
        best, times = autotune.tune(s, steps=1, use_cache=False)
        assert times and s.get_schedule() == best
        assert all(np.array_equal(x, y) for x, y in zip(a, states(s)))
        s.step(dt, 3)
        r.step(dt, 5)
        assert all(np.array_equal(x, y) for x, y in zip(states(r), states(s)))
        s.checkpoint_free(-1)
    finally:
        s.close()
        r.close()


def test_batch_autotune():
    pars = small_sweep(n=3)
    b = BatchSolver(pars, precision="fp32")
    ref = single_runs(pars, "fp32")
    try:
        dts, ns = b.prepare()
        best, times = autotune.tune(b, steps=1, use_cache=False)
        assert times
        for x in b.solvers:
            tile, bpc, cw = x.get_schedule()
            assert (bpc, cw) == best[1:]
            if best[0] == autotune.FIT:      # each problem on its own tile
                assert tile in [tuple(t) for t in autotune.fit_tiles(x)]
            else:
                assert tile == best[0]
        b.step(ns)
        for i, s in enumerate(b.solvers):
            assert all(np.array_equal(x, y) for x, y in zip(states(s), ref[i]))
    finally:
        b.close()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_batch_mixed_tiles_bitwise(precision):
    """Each handle on its own tile shape (the autotuner's "fit" schedule):
    one batched launch then mixes tile shapes and slot sizes, and every
    problem still equals its single-handle run bit for bit."""
    pars = small_sweep(n=6, side_div=2, tfinal=0.05)
    ref = single_runs(pars, precision)
    b = BatchSolver(pars, precision=precision)
    try:
        b.prepare()
        autotune.apply(b, autotune.FIT, 2, 4)
        tiles = {tuple(s.get_schedule()[0]) for s in b.solvers}
        print("batch fit", precision, [(p["n"][1], p["n_mom"]) for p in pars], sorted(tiles), flush=True)
        assert len(tiles) > 1
        dts, ns = b.prepare()
        b.step(ns, use_graph=True)
        for i, s in enumerate(b.solvers):
            for j, (x, y) in enumerate(zip(states(s), ref[i])):
                assert np.array_equal(x, y), (i, j)
    finally:
        b.close()


def _spatial_filter(par, dt, k, x, y):
    # a filter field (varies in space): the kernel reads it from global memory
    return np.exp(-dt * 8.0 * (1 + 0.1 * x) * (k / par["n_mom"]) ** 4)


@pytest.mark.parametrize("case", ["lattice2d", "lattice2d_filter_field", "gauss3d"])
def test_pair_tiles_keep_every_bit(case):
    """FP32 tiles of two cells per thread (up to 512 cells) give the same
    bits as the one-cell tiles: every cell's arithmetic is unchanged."""
    if case == "lattice2d":
        prob, tiles = P.lattice_2d(n=70, n_mom=9, tfinal=0.3), [(63, 8), (127, 4), (31, 16), (15, 32)]
    elif case == "lattice2d_filter_field":
        prob = dict(P.lattice_2d(n=50, n_mom=7, tfinal=0.3), filterfunction=_spatial_filter)
        tiles = [(63, 8), (31, 16)]
    else:
        prob = P.gauss_periodic_3d(n=18, n_mom=5, tfinal=0.3, std=0.2)
        tiles = [(31, 4, 4), (15, 8, 4)]
    par = I.starmap_init(prob)
    res = []
    for tile, bpc, cw in [(None, 0, 4)] + [(t, b, c) for t in tiles for b, c in ((1, 8), (2, 4))]:
        print("pair", case, tile, bpc, cw, flush=True)
        s = HipSolver(par, precision="fp32", cell_waves=4)
        try:
            dt = s.prepare()
            s.step(dt, 2)
            if tile is not None:
                s.set_schedule(tile, bpc, cw)
            s.step(dt, 5)
            res.append(states(s))
        finally:
            s.close()
    for other in res[1:]:
        for x, y in zip(res[0], other):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("use_graph,bpc,cw", [(False, 2, 4), (True, 2, 4), (True, 1, 8)])
def test_batch_with_pair_tiles_bitwise_equal_single_handles(use_graph, bpc, cw):
    pars = small_sweep()
    ref = single_runs(pars, "fp32")
    b = BatchSolver(pars, precision="fp32", cell_waves=cw)
    try:
        autotune.apply(b, autotune.FIT_PAIR, bpc, cw)
        assert all(int(np.prod(s.get_schedule()[0])) > 256 or max(s.n[1:]) < 64
                   for s in b.solvers)
        dts, ns = b.prepare()
        b.step(ns, use_graph=use_graph)
        for i, s in enumerate(b.solvers):
            for j, (x, y) in enumerate(zip(states(s), ref[i])):
                assert np.array_equal(x, y), (i, j)
    finally:
        b.close()
