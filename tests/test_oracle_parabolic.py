"""CPU checks of the parabolic time-step selection (starmap_solver.m:171-204,
:253-265, growth_factor :506-514): the oracle's restatement against closed
forms, and the product's host module (starmap_amd/timestep.py) against the
oracle.

Closed forms that pin growth_factor:
* one component, no coupling: the growth factor of the step is
  exp(-sigma dt) (four half-steps of exact absorption), 1 in the relative
  mode (timestep 2);
* two components with symbol -i m/h [[0,1],[1,0]] on the two parities and
  no decay: the step matrix has determinant 1 and trace 2 - 4 s^2,
  s = dt m/h, so its largest eigenvalue modulus is 1 for s <= 1 and
  |1 - 2 s^2| + sqrt((1 - 2 s^2)^2 - 1) beyond.
A dropped half-step, a wrong parity mask or a wrong decay factor breaks
these.  The search then returns the stable limit h/m (with a small decay)
to the bisection's 2%.
"""
import math

import numpy as np
import pytest

from oracle import starmap_oracle as so
from starmap_amd import problems as P
from starmap_amd import timestep as T


@pytest.mark.parametrize("d,dt", [(0.7, 0.3), (1e-5, 0.1), (3.0, 2.0), (0.0, 1.0)])
def test_growth_factor_single_component(d, dt):
    M = np.zeros((1, 1), dtype=complex)
    assert so.growth_factor(M, [0], [], [d], dt, 1) == pytest.approx(math.exp(-d * dt), rel=1e-14)
    assert so.growth_factor(M, [0], [], [d], dt, 2) == pytest.approx(1.0, rel=1e-14)


@pytest.mark.parametrize("s", [0.3, 0.9, 1.0, 1.05, 1.5, 3.0])
def test_growth_factor_staggered_pair(s):
    m, h = 0.6, 0.1
    M = -1j * np.array([[0, m], [m, 0]]) / h
    la = so.growth_factor(M, [0], [1], [0.0, 0.0], s * h / m, 1)
    tr = 2 - 4 * s * s
    ref = 1.0 if abs(tr) <= 2 else abs(tr) / 2 + math.sqrt(tr * tr / 4 - 1)
    assert la == pytest.approx(ref, rel=1e-12)


@pytest.mark.parametrize("dt0", [0.01, 0.5])
def test_search_finds_stability_limit(dt0):
    m, h = 0.6, 0.1
    M = -1j * np.array([[0, m], [m, 0]]) / h
    dt, la0 = so._stable_dt(M, [0], [1], [1e-3, 1e-3], dt0, 1e9, 1)
    assert 0.98 <= dt / (h / m) <= 1.0
    assert so.growth_factor(M, [0], [1], [1e-3, 1e-3], dt, 1) < la0


def test_product_module_matches_oracle():
    rng = np.random.default_rng(4)
    for n in (3, 6, 10):
        A = rng.normal(size=(n, n))
        M = -1j * (A + A.T) * 3.0
        even = list(range(0, n, 2))
        odd = list(range(1, n, 2))
        decay = rng.random(n) * 5
        for dt in (0.01, 0.05, 0.2):
            for ts in (1, 2):
                a = T.growth_factor(M, even, odd, decay, dt, ts)
                b = so.growth_factor(M, even, odd, decay, dt, ts)
                assert a == pytest.approx(b, rel=1e-12)
        for ts in (1, 2):
            assert T.select_dt(M, even, odd, decay, 0.01, 10.0, ts) == so._stable_dt(M, even, odd, decay, 0.01, 10.0, ts)


def test_expm1div_matches_oracle():
    # (exp(x)-1)/x cancels about log10(1/|x|) digits just above the series
    # threshold: one ulp of exp there is 1e-16/|x| relative
    x = np.concatenate([np.linspace(-3, 3, 101), np.linspace(-3e-4, 3e-4, 61)])
    tol = 4e-16 * np.maximum(1.0, 1.0 / np.maximum(np.abs(x), 1e-300))
    assert np.all(np.abs(T.expm1div(x) - so.expm1div(x)) <= tol * np.abs(so.expm1div(x)))


def test_parabolic_steps_fewer_and_stable():
    hyp = so.starmap_solver(so.starmap_init(P.asymptotic_preserving_2d(n=30, n_plot=3, timestep=0)),
                            verbose=False, return_state=True)[1]["_oracle"]
    par = so.starmap_init(P.asymptotic_preserving_2d(n=30, n_plot=3, timestep=1))
    _, p2 = so.starmap_solver(par, verbose=False, return_state=True)
    o = p2["_oracle"]
    assert o["step_count"] * 10 < hyp["step_count"]
    assert o["dt"] > 10 * o["dt_hyp"]
