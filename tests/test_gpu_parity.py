"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle on the
same inputs.

Tolerances (relative L2 per moment field at the final step, measured against
the largest field norm when a field is zero in the oracle):
  FP64 state: 1e-12
  FP32 state: 1e-5
"""
import numpy as np
import pytest

from oracle import starmap_oracle as so
from starmap_amd import init as I
from starmap_amd import problems as P
from starmap_amd.solver import HipSolver

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 1e-5}


def rel_errors(out, ref):
    norms = np.array([np.linalg.norm(r["U"]) for r in ref])
    big = norms.max()
    errs = []
    for o, r, nr in zip(out, ref, norms):
        assert np.shape(o["U"]) == np.shape(r["U"])
        d = np.linalg.norm(np.asarray(o["U"]) - np.asarray(r["U"]))
        errs.append(d / nr if nr > 1e-8 * big else d / big)
    return np.array(errs)


def run_both(prob, precision="fp64"):
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
}


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


@pytest.mark.parametrize("case", ["gauss2d_p5_100", "lattice2d_p7_70", "gauss3d_p7_20"])
def test_fp32_within_1e5(case):
    print("case", case, flush=True)
    out, ref = run_both(CASES[case](), precision="fp32")
    errs = rel_errors(out, ref)
    assert errs.max() <= TOL["fp32"], (case, errs.max())


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
