"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle on the
same inputs.

Tolerances (L2 error per moment field at the final step):
  FP64 state: 1e-12 of max(the field's norm, 1e-3 of the largest field's
  norm), see rel_errors
  FP32 state: the bound of tests/fp32_bounds.py (1e-5 relative over all
  moments and per field down to 1e-2 of the largest field; every field
  within 1e-6 of the largest field's norm)
"""
import numpy as np
import pytest

from oracle import starmap_oracle as so
from starmap_amd import init as I
from starmap_amd import problems as P
from starmap_amd.solver import HipSolver
from tests.fp32_bounds import check_fp32_bound
from tests.mms import POINTS as MMS_POINTS
from tests.mms import mms_errors, mms_problem

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 1e-5}


def rel_errors(out, ref):
    """L2 error per moment field over max(its norm, 1e-3 of the largest
    field's norm).  Rounding differences (the device's fused multiply-adds,
    the coefficient -M/h rounded once instead of dividing by h) are white
    noise at about eps64 of the largest field, which the transport spreads
    over every moment; a field far below the largest sees it magnified by
    the ratio of the norms (MMS P3 32^2: 2e-16 of the largest field on a
    field 7e-5 below it, 3e-12 of its own norm)."""
    norms = np.array([np.linalg.norm(r["U"]) for r in ref])
    big = norms.max()
    errs, own = [], []
    for o, r, nr in zip(out, ref, norms):
        assert np.shape(o["U"]) == np.shape(r["U"])
        d = np.linalg.norm(np.asarray(o["U"]) - np.asarray(r["U"]))
        errs.append(d / max(nr, 1e-3 * big))
        own.append(d / nr if nr > 0 else (0.0 if d == 0 else np.inf))
    print("  worst error over own norm %.2e (field %d), over the measure %.2e"
          % (max(own), int(np.argmax(own)), max(errs)), flush=True)
    return np.array(errs)


def run_both(prob, precision="fp64"):
    """The oracle and the HIP path on the same problem, every moment field in
    the outputs (whatever the problem's own mom_output)."""
    prob = dict(prob, mom_output="all")
    ref, rpar = so.starmap_solver(so.starmap_init(prob), verbose=False, return_state=True)
    s = HipSolver(I.starmap_init(prob), precision=precision)
    try:
        out = s.run(verbose=False)
        assert s.step_count == rpar["_oracle"]["step_count"]
        assert s.dt == rpar["_oracle"]["dt"]
        for o, r in zip(out, ref):
            for k in ("x", "y", "z"):
                if k in r:
                    assert np.array_equal(o[k], r[k])
    finally:
        s.close()
    return out, ref


CASES = {
    # 2D P5 100x100 periodic Gaussian, t_final 1.0 (67 steps): full size
    "gauss2d_p5_100": lambda: P.gauss_periodic_2d(),
    # 2D lattice, extrapolation BC, exponential filter at substep position
    "lattice2d_p7_70": lambda: P.lattice_2d(n=70, n_mom=7, tfinal=0.4),
    "lattice2d_p39_48": lambda: P.lattice_2d(n=48, n_mom=39, tfinal=0.3),
    "lattice2d_p8_unfiltered": lambda: P.lattice_2d(n=50, n_mom=8, tfinal=0.3, f_position=""),
    "lattice2d_full_filter": lambda: P.lattice_2d(n=40, n_mom=5, tfinal=0.3, f_position="full"),
    # 3D periodic with anisotropic scattering
    "gauss3d_p7_20": lambda: P.gauss_periodic_3d(n=20, n_mom=7, tfinal=0.2, std=0.15),
    "gauss3d_p3_odd": lambda: dict(P.gauss_periodic_3d(n=13, n_mom=3, tfinal=0.3, std=0.2),
                                   n=[13, 11, 9]),
    # SP_N with a time-dependent source
    "boxes_sp5": lambda: P.boxes_2d(n=40, n_mom=5, closure="SP", tfinal=0.3),
    "boxes_p3": lambda: P.boxes_2d(n=36, n_mom=3, closure="P", tfinal=0.2),
    # parabolic time step (starmap_solver.m:171-204, :253-265)
    "ap_p1_parabolic": lambda: P.asymptotic_preserving_2d(n=40, n_plot=3),
    "ap_p3_parabolic_relative": lambda: P.asymptotic_preserving_2d(n=30, n_mom=3, timestep=2,
                                                                   sigma_a=0.5, n_plot=3),
    "ap_p1_parabolic_timedep": lambda: P.asymptotic_preserving_2d(n=30, n_plot=3,
                                                                  time_dependent=True),
    # manufactured solution of starmap_2d_create_mms.m: time-dependent
    # sigma_a, sigma_s0 and sigma_sm (hyperbolic time-dependent branch,
    # materials averaged over each step, last step ending on tfinal), source in
    # every moment
    "mms_p3_32": lambda: mms_problem(32, 3, "P")[0],
    "mms_sp5_24": lambda: dict(mms_problem(24, 5, "SP")[0], tfinal=0.5),
    # starmap_2d_ex_controlrod.m: sigma_a(x,y,t) under the hyperbolic step,
    # outputs during the rod's motion
    "controlrod_sp3_40": lambda: P.controlrod_2d(n=40, tfinal=0.7, n_plot=8),
    # a filter with a time argument, re-evaluated every step
    # (starmap_solver.m:284), at substep and full position
    "lattice2d_moving_filter_substep": lambda: dict(
        P.lattice_2d(n=40, n_mom=5, tfinal=0.4), filterfunction=P.exp_filter_moving(4, 10.0)),
    "lattice2d_moving_filter_full": lambda: dict(
        P.lattice_2d(n=36, n_mom=4, tfinal=0.4, f_position="full"),
        filterfunction=P.exp_filter_moving(2, 5.0)),
    # a filter that varies in space (starmap_solver.m:284-305 with x, y): the
    # kernel reads the filter fields from global memory (its variant with
    # global-memory coefficients)
    "lattice2d_filter_field": lambda: dict(P.lattice_2d(n=50, n_mom=7, tfinal=0.3),
                                           filterfunction=spatial_filter),
    # the 1D examples at full size (starmap_1d_ex_gauss.m, P19 1000 cells;
    # starmap_1d_ex_planesource.m, SP7 1000 cells): 1D problems lifted to a
    # 1 x 1 x n grid, one output time each
    "gauss1d_p19_1000": lambda: P.gauss_1d(),
    "planesource1d_sp7_1000": lambda: P.planesource_1d(),
    # SP_N in 3D (the 3D SP_N closure's coupling)
    "gauss3d_sp5_12": lambda: P.gauss_periodic_3d(n=12, n_mom=5, closure="SP", tfinal=0.6, std=0.2),
    # periodic in x, extrapolation in y, on a domain the Gaussian's tails
    # reach; off centre, or the periodic neighbour of an edge cell would
    # equal the cell itself and both conditions would give the same result
    "gauss2d_mixed_bc": lambda: dict(P.gauss_periodic_2d(n=40, n_mom=5, tfinal=0.6, half=0.5),
                                     bc=[0, 1], ic=_offset_gaussian),
}


def _offset_gaussian(x, y):
    return np.exp(-((x - 0.2) ** 2 + (y + 0.1) ** 2) / (2 * 0.2 ** 2))


def spatial_filter(par, dt, k, x, y):
    return np.exp(-dt * 8.0 * (1 + 0.1 * x) * (k / par["n_mom"]) ** 4)


@pytest.mark.parametrize("case", sorted(CASES))
def test_fp64_matches_oracle(case):
    print("case", case, flush=True)
    out, ref = run_both(CASES[case]())
    errs = rel_errors(out, ref)
    assert errs.max() <= TOL["fp64"], (case, errs.max(), int(errs.argmax()))


def test_3d_extrapolation_filter():
    prob = P.gauss_periodic_3d(n=10, n_mom=3, tfinal=0.3, std=0.2)
    prob.update(bc=[1, 1, 1], filterfunction=P.exp_filter(2, 5.0), f_position="substep")
    out, ref = run_both(prob)
    assert rel_errors(out, ref).max() <= TOL["fp64"]


@pytest.mark.parametrize("case", ["gauss2d_p5_100", "lattice2d_p7_70", "lattice2d_p39_48",
                                  "gauss3d_p7_20", "lattice2d_filter_field",
                                  "gauss1d_p19_1000", "planesource1d_sp7_1000"])
def test_fp32_within_1e5(case):
    """The FP32 state against the FP64 oracle, within the bound of
    tests/fp32_bounds.py."""
    print("case", case, flush=True)
    out, ref = run_both(CASES[case](), precision="fp32")
    e = check_fp32_bound([o["U"] for o in out], [r["U"] for r in ref])
    print("  fp32 errors", e, flush=True)


def test_mms_second_order_on_gpu():
    """The manufactured solution of starmap_2d_create_mms.m converges at
    second order on the GPU as on the oracle (grids 2.^(4:7), P3, tfinal 1):
    every moment order's L1, L2 and max error falls by 4 per halving of h."""
    errs = []
    for n in MMS_POINTS:
        print("mms", n, flush=True)
        prob, solution = mms_problem(n, 3, "P")
        par = I.starmap_init(prob)
        s = HipSolver(par, precision="fp64")
        try:
            sol = s.run(verbose=False)
            assert abs(s.t - prob["tfinal"]) <= 1e-14
        finally:
            s.close()
        errs.append(mms_errors(sol, solution, prob["tfinal"], np.atleast_1d(par["mom_order"]), n))
    p = np.array([[np.log2(errs[i][k] / errs[i + 1][k]) for i in range(len(errs) - 1)]
                  for k in range(3)])
    assert np.all(np.abs(p[:, -1, :] - 2.0) <= 0.04), p[:, -1, :]
    assert np.all(p >= 1.9), p


def test_output_callback_interpolation():
    """Callbacks receive the same interpolated fields as in the oracle,
    several output times per step included."""
    got = {"gpu": {}, "ref": {}}

    def make(tag):
        def output(par, x, y, U, step):
            got[tag][step] = [u.copy() for u in U]
        return output

    base = P.gauss_periodic_2d(n=40, n_mom=3, std=0.1, tfinal=0.2, mom_output=[1, 2, 5])
    tp = [0.0, 0.031, 0.0311, 0.1, 0.2]
    pr = dict(base, t_plot=tp, output=make("ref"))
    so.starmap_solver(so.starmap_init(pr), verbose=False)
    pg = dict(base, t_plot=tp, output=make("gpu"))
    s = HipSolver(I.starmap_init(pg))
    s.run(verbose=False)
    s.close()
    assert sorted(got["gpu"]) == sorted(got["ref"]) == [1, 2, 3, 4, 5]
    for k in got["ref"]:
        for a, b in zip(got["gpu"][k], got["ref"][k]):
            assert np.linalg.norm(a - b) <= 1e-12 * max(np.linalg.norm(b), 1e-300) + 1e-300


def test_observables_reductions():
    prob = P.lattice_2d(n=30, n_mom=3, tfinal=0.2)
    s = HipSolver(I.starmap_init(prob))
    out = s.run(verbose=False)
    sums, sq, bad = s.observables()
    s.close()
    assert not bad
    h = 7.0 / 30
    n = np.array([1, 30, 30])
    for j, o in enumerate(out):
        U = o["U"]
        assert abs(sums[j] - U.sum()) <= 1e-12 * max(np.abs(U).sum(), 1e-300)
        # staggered trapezoid weights of starmap_2d_ex_l2norm.m:93-99
        st = [U.shape[0] == 30, U.shape[1] == 30]
        w1 = np.ones(U.shape[0])
        w1[[0, -1]] = st[0] / 2
        w2 = np.ones(U.shape[1])
        w2[[0, -1]] = st[1] / 2
        want = h * h * w1 @ (U ** 2) @ w2
        assert abs(sq[j] - want) <= 1e-12 * max(want, 1e-300)


def test_errors_are_reported_not_crashes():
    from starmap_amd import _lib
    from starmap_amd._lib import SmpError
    s = HipSolver(I.starmap_init(P.gauss_periodic_2d(n=8, n_mom=1)))
    bad = np.zeros(5)
    with pytest.raises(SmpError, match="count does not match"):
        _lib.check(_lib.lib().smp_set_state(s.handle, 0, bad.ctypes.data, bad.size))
    with pytest.raises(SmpError, match="sigma"):
        s.step(0.01, 1)           # materials never set
    with pytest.raises(SmpError, match="dt must be positive"):
        s.step(-1.0, 1)
    s.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fused_steps_bitwise_equal_single_steps(precision):
    """n steps in one call (substep 3 fused with the next substep 1) give the
    same bits as n calls of one step (three launches each)."""
    prob = P.lattice_2d(n=45, n_mom=9, tfinal=1.0)
    outs = []
    for split in (False, True):
        s = HipSolver(I.starmap_init(prob), precision=precision)
        dt = s.prepare()
        if split:
            for _ in range(7):
                s.step(dt, 1)
        else:
            s.step(dt, 7)
        outs.append([s.get_state(j) for j in range(s.n_sys)])
        s.close()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("f_position", ["substep", "full"])
def test_step_timing(f_position):
    """smp_last_step_timing after one call: the substep launch count of the
    call (2n+1 for n fused steps, 3n with the 'full' filter), their time
    within the call's elapsed time (equal to it when only substep kernels
    run; with the 'full' filter the scale kernels run between them), and
    their algorithmic bytes (each source component read once, each target
    read and written once per launch)."""
    prob = P.lattice_2d(n=45, n_mom=9, tfinal=1.0, f_position=f_position)
    s = HipSolver(I.starmap_init(prob), precision="fp32")
    try:
        dt = s.prepare()
        n = 5
        s.step(dt, n)
        tot, sub, launches, by = s.timing()
        n_even = int(np.sum((s.gt - 1).sum(axis=1) % 2 == 0))
        n_odd = s.n_sys - n_even
        cells = int(np.prod(s.n))
        if f_position == "full":
            assert launches == 3 * n
            assert 0 < sub < tot
            kinds = [0, 1, 0] * n
        else:
            assert launches == 2 * n + 1
            assert 0 < sub == tot
            kinds = [0] + [1, 0] * n
        want = sum(4 * cells * (n_even + 2 * n_odd if k == 0 else n_odd + 2 * n_even) for k in kinds)
        assert by == want
    finally:
        s.close()


def test_summary_print_matches_oracle(capsys):
    """The overview lines of starmap_solver.m:477-482 are the oracle's, and
    the CPU-time lines have the reference's format."""
    import re
    prob = P.lattice_2d(n=30, n_mom=5, tfinal=0.2)
    so.starmap_solver(so.starmap_init(prob), verbose=True)
    ref = capsys.readouterr().out.splitlines()
    s = HipSolver(I.starmap_init(prob))
    try:
        s.run(verbose=True)
    finally:
        s.close()
    got = capsys.readouterr().out.splitlines()
    assert got[:3] == ref[:3] and got[0] == "Lattice 2D with FP5"
    for ln, lab in zip(got[3:6], ("advection:", "absorb/scatter:", "plotting:")):
        assert re.fullmatch(r" %s *\d+\.\d\ds *\d+%%" % re.escape(lab), ln), ln
        assert len(ln) == len(ref[3 + ("advection:", "absorb/scatter:", "plotting:").index(lab)])
    pct = [int(ln.rsplit("s", 1)[1].rstrip("%")) for ln in got[3:6]]
    assert abs(sum(pct) - 100) <= 2


def test_callback_returning_par_changes_the_run():
    """A callback that returns par (here halving tfinal at the first output)
    changes the rest of the run as in the oracle (starmap_solver.m:441-469);
    callbacks see t_plot with inf appended."""
    def make(seen):
        def output(par, x, y, U, step):
            seen.append((step, list(par["t_plot"])))
            if step == 1:
                return dict(par, tfinal=par["tfinal"] / 2)
            return None
        return output

    base = P.lattice_2d(n=30, n_mom=3, tfinal=0.6, t_plot=[0.1, 0.6])
    seen_r, seen_g = [], []
    ref, rpar = so.starmap_solver(so.starmap_init(dict(base, output=make(seen_r))),
                                  verbose=False, return_state=True)
    s = HipSolver(I.starmap_init(dict(base, output=make(seen_g))))
    try:
        out = s.run(verbose=False)
        assert s.step_count == rpar["_oracle"]["step_count"]
    finally:
        s.close()
    assert seen_g == seen_r and [k for k, _ in seen_g] == [1] and seen_g[0][1][-1] == np.inf
    assert rel_errors(out, ref).max() <= TOL["fp64"]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_unstable_time_step_is_reported(precision):
    """A time step 8x beyond the stability limit makes the state overflow
    (the reference runs on with Inf/NaN): the library's scan in smp_step
    returns SMP_E_UNSTABLE, raised as SmpUnstable."""
    from starmap_amd._lib import SmpUnstable
    prob = P.gauss_periodic_2d(n=16, n_mom=3, std=0.2, tfinal=300.0)
    prob["cn"] = prob["cn"] * 8
    s = HipSolver(I.starmap_init(prob), precision=precision)
    try:
        with pytest.raises(SmpUnstable, match="non-finite"):
            s.run(verbose=False)
    finally:
        s.close()


def test_stable_state_passes_the_finite_scan():
    import ctypes
    from starmap_amd import _lib
    s = HipSolver(I.starmap_init(P.gauss_periodic_2d(n=16, n_mom=3, std=0.2, tfinal=0.5)))
    try:
        s.run(verbose=False)
        bad = ctypes.c_int32(7)
        _lib.check(_lib.lib().smp_check_finite(s.handle, ctypes.byref(bad)))
        assert bad.value == 0
    finally:
        s.close()


def test_out_of_memory_is_reported():
    """A state beyond the GPU's memory (P39 on 12000^2 cells: 946 GB in
    FP64) fails in smp_create with SMP_E_OOM, not a crash."""
    from starmap_amd import _lib
    prob = P.lattice_2d(n=12000, n_mom=39, f_position="")
    with pytest.raises(_lib.SmpError) as e:
        HipSolver(I.starmap_init(prob), precision="fp64")
    assert e.value.code == _lib.SMP_E_OOM, str(e.value)
    # the failed allocation leaves no error behind for the next solve
    out, ref = run_both(P.lattice_2d(n=20, n_mom=3, tfinal=0.2))
    assert rel_errors(out, ref).max() <= TOL["fp64"]
