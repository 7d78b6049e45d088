"""Full-size checks of the HIP path where the oracle is too slow: exact
linearity, and the exact mass decay of the zeroth moment on a periodic grid.

For fixed materials and filters the scheme is linear in the state and the
source, and scaling by 2 commutes with rounding (the values here are far from
overflow and from the subnormal range).  So a run from 2*U0 with source 2*Q
equals twice the run from U0 with Q, bit for bit, after any number of steps,
at every size.  Both runs start from the same seeded random state in every
component, so every component carries values.

This catches reads of anything that is not state (coefficient fields, stale
or uninitialised memory, the sink), races between groups, tiles or streams,
and non-linear mistakes, at the full sizes: the 2D P39 lattice on 1000x1000
(820 components, 4016 tiles per launch) and the 3D P7 Gaussian on 200^3.  It
does not catch a wrong linear stencil, which the oracle parity cases check at
smaller sizes (P39 among them).

Checked through the device moment sums of every component (sums of squares
and of squared neighbour differences exactly 4x, in FP64) and a sample of
whole fields (exactly 2x).
"""
import inspect

import numpy as np
import pytest

from oracle import starmap_oracle as so
from starmap_amd import init as I
from starmap_amd import problems as P
from starmap_amd.observables import moment_terms
from starmap_amd.solver import HipSolver

pytestmark = pytest.mark.gpu


def twice(f):
    """2*f with f's number of arguments (which tells the set-up whether the
    source depends on time)."""
    n = len(inspect.signature(f).parameters)
    g = lambda *x: 2.0 * np.asarray(f(*x), dtype=np.float64)
    return {2: lambda a, b: g(a, b), 3: lambda a, b, c: g(a, b, c),
            4: lambda a, b, c, d: g(a, b, c, d)}[n]


def doubled(prob):
    q = dict(prob)
    if callable(prob.get("source")):
        q["source"] = twice(prob["source"])
    return q


def check_linear(prob, precision, steps, seed=11):
    a = HipSolver(I.starmap_init(prob), precision=precision)
    b = HipSolver(I.starmap_init(doubled(prob)), precision=precision)
    try:
        dt = a.prepare()
        assert b.prepare() == dt
        rng = np.random.default_rng(seed)
        for j in range(a.n_sys):
            u = rng.random(a.shape(j)) - 0.5
            if precision == "fp32":
                u = u.astype(np.float32).astype(np.float64)
            a.set_state(j, u)
            b.set_state(j, 2.0 * u)
        a.step(dt, steps)
        b.step(dt, steps)
        sa, da = moment_terms(a)
        sb, db = moment_terms(b)
        print("linearity", precision, a.n_sys, "components, sum of squares of component 0:", sa[0],
              flush=True)
        assert np.all(sa > 0)
        assert np.array_equal(sb, 4.0 * sa)
        assert np.array_equal(db, 4.0 * da)
        for j in sorted({0, 1, 2, a.n_sys // 2, a.n_sys - 2, a.n_sys - 1}):
            assert np.array_equal(b.get_state(j), 2.0 * a.get_state(j)), j
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_lattice_p39_1000_linear(precision):
    """2D P39 lattice, 1000x1000, extrapolation boundaries, substep filter,
    source in the centre block: three steps (seven launches)."""
    check_linear(P.lattice_2d(n=1000, n_mom=39, tfinal=3.2), precision, 3)


def test_gauss3d_p7_200_linear():
    """3D P7, 200^3 periodic, anisotropic scattering: two steps in FP64."""
    check_linear(P.gauss_periodic_3d(n=200, n_mom=7, tfinal=0.75), "fp64", 2)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-6)])
def test_gauss3d_p7_200_mass_decay(precision, tol):
    """3D P7, 200^3 periodic, sigma_a = 0.1, no source: the flux of the
    zeroth moment sums to zero over the periodic grid, so its total decays by
    exactly exp(-sigma_a*dt) per step (two half-step applications, each
    multiplying it by 1 - dt/2*sigma_a*expm1div(-sigma_a*dt/2) =
    exp(-sigma_a*dt/2)), to the rounding of the state and of the sums."""
    s = HipSolver(I.starmap_init(P.gauss_periodic_3d(n=200, n_mom=7, tfinal=0.75)), precision=precision)
    try:
        dt = s.prepare()
        m0 = s.observables()[0][0]
        s.step(dt, 4)
        m4 = s.observables()[0][0]
        expect = m0 * np.exp(-0.1 * dt * 4)
        print("mass", precision, m0, m4, expect, (m4 - expect) / expect, flush=True)
        assert m0 > 0
        assert abs(m4 - expect) <= tol * expect
    finally:
        s.close()


def varied_ic(x, y, z, k):
    """A different field in every component k (1-based), with wavelengths from
    about 8 to 160 cells of the full-size grids: deterministic in (point, k),
    so the oracle and the HIP path start from the same state."""
    f = (k * 0.6180339887) % 1.0
    g = (k * 0.4142135624) % 1.0
    return 0.5 * np.cos((5.0 + 95.0 * f) * x + (3.0 + 60.0 * g) * y + (7.0 + 40.0 * f * g) * z
                        + 6.283 * g) / (1.0 + 0.01 * k)


def varied_ic_2d(x, y, k):
    return varied_ic(x, y, 0.0, k)


@pytest.mark.parametrize("workload,precision", [("c2", "fp64"), ("c3", "fp64"), ("c2", "fp32")])
def test_full_size_matches_oracle(workload, precision):
    """C2 (2D P39 lattice 1000^2: substep filter, source, absorber blocks)
    and C3 (3D P7 200^3 periodic, anisotropic scattering) at their full sizes
    against the oracle, every moment field at 1e-12 relative L2 (the measure
    of the parity tests) in FP64, within the FP32 bound (tests/fp32_bounds.py)
    for the FP32 state: a varied state in every component, two steps (final
    time 1.5 dt, so the output interpolates between the steps' states,
    starmap_solver.m:425-428)."""
    from tests.fp32_bounds import check_fp32_bound
    from tests.test_gpu_parity import rel_errors
    if workload == "c2":
        prob = P.lattice_2d(n=1000, n_mom=39, tfinal=3.2)
    else:
        prob = P.gauss_periodic_3d(n=200, n_mom=7, tfinal=0.75)
    prob["ic"] = varied_ic_2d if workload == "c2" else varied_ic
    par = I.starmap_init(prob)
    s = HipSolver(par, device=-1, precision=precision)   # host only: the first step's dt
    try:
        dt = min(s.dt_hyp, par["tfinal"] / par["cn"]) * par["cn"]
    finally:
        s.close()
    prob.update(tfinal=1.5 * dt, t_plot=[1.5 * dt], mom_output="all")
    print("oracle", workload, flush=True)
    ref, rpar = so.starmap_solver(so.starmap_init(prob), verbose=False, return_state=True)
    print("hip", workload, precision, flush=True)
    s = HipSolver(I.starmap_init(prob), precision=precision)
    try:
        out = s.run(verbose=False)
        assert s.step_count == rpar["_oracle"]["step_count"] == 2
        assert s.dt == rpar["_oracle"]["dt"]
    finally:
        s.close()
    assert len(out) == len(ref) == (820 if workload == "c2" else 64)
    if precision == "fp32":
        e = check_fp32_bound([o["U"] for o in out], [r["U"] for r in ref])
        print("full size", workload, precision, e, flush=True)
        return
    errs = rel_errors(out, ref)
    print("full size", workload, "worst", errs.max(), flush=True)
    assert errs.max() <= 1e-12, (workload, errs.max(), int(errs.argmax()))
