"""GPU tests of the observables entry points (through the C ABI).

* smp_moment_norms: per component sum of squares and of squared neighbour
  differences along each output dimension, against numpy on the state read
  back in the reference layout (periodic faces, whose stored order is
  rotated, extrapolation, FP32, 3D): relative 1e-12.
* smp_probe: every component at a point against scipy's linear
  RegularGridInterpolator on the reference grid vectors (interp2 / interp3
  semantics, NaN outside): 1e-12.
* The moment-order norms and the angular reconstruction built on them.
"""
import math

import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator

from starmap_amd import init as I
from starmap_amd import observables as O
from starmap_amd import problems as P
from starmap_amd.solver import HipSolver

pytestmark = pytest.mark.gpu


def stepped(prob, precision="fp64", steps=4):
    s = HipSolver(I.starmap_init(prob), precision=precision)
    dt = s.prepare()
    s.step(dt, steps)
    return s


def active(s):
    return [a for a in range(3) if s.n[a] > 1]


CASES = [
    ("periodic 2D P3", lambda: P.gauss_periodic_2d(n=12, n_mom=3), "fp64"),
    ("lattice 2D P5", lambda: P.lattice_2d(n=14, n_mom=5, tfinal=0.1), "fp64"),
    ("lattice 2D P5 fp32", lambda: P.lattice_2d(n=14, n_mom=5, tfinal=0.1), "fp32"),
    ("periodic 3D P3", lambda: P.gauss_periodic_3d(n=6, n_mom=3, tfinal=0.1), "fp64"),
]


@pytest.mark.parametrize("name,make,precision", CASES, ids=[c[0] for c in CASES])
def test_moment_terms_match_numpy(name, make, precision):
    s = stepped(make(), precision)
    try:
        sumsq, diffsq = O.moment_terms(s)
        act = active(s)
        for j in range(s.n_sys):
            U = s.get_state(j)
            assert abs(sumsq[j] - (U ** 2).sum()) <= 1e-12 * max((U ** 2).sum(), 1e-300), (name, j)
            for e, a in enumerate(act):
                ref = (np.diff(U, axis=a) ** 2).sum()
                assert abs(diffsq[j, e] - ref) <= 1e-12 * max(ref, 1e-300) + 1e-300, (name, j, e)
            for e in range(len(act), 3):
                assert diffsq[j, e] == 0.0
    finally:
        s.close()


@pytest.mark.parametrize("name,make,precision", CASES, ids=[c[0] for c in CASES])
def test_probe_matches_linear_interpolation(name, make, precision):
    s = stepped(make(), precision)
    try:
        act = active(s)
        ax = np.asarray(s.par["ax"], dtype=np.float64)
        rng = np.random.default_rng(7)
        pts = [ax[[2 * a for a in act]] + 0.37 * s.h[act],           # inside, off grid
               np.asarray([s.coords[a][0][2] for a in act]),          # a cell centre
               np.asarray([s.coords[a][1][-1] for a in act]),         # the last face
               ax[[2 * a + 1 for a in act]] + 0.1]                    # outside
        pts += [ax[[2 * a for a in act]] + rng.random(len(act)) * (ax[[2 * a + 1 for a in act]] -
                                                                    ax[[2 * a for a in act]])
                for _ in range(4)]
        for p in pts:
            got = O.probe(s, p)
            for j in range(s.n_sys):
                U = s.get_state(j)
                grids = [s.coords[a][s.gt[j, a] - 1] for a in act]
                V = U.reshape([U.shape[a] for a in act])
                f = RegularGridInterpolator(grids, V, method="linear", bounds_error=False,
                                            fill_value=np.nan)
                ref = float(f(p[None, :])[0])
                if np.isnan(ref):
                    assert np.isnan(got[j]), (name, j, p, got[j])
                else:
                    assert abs(got[j] - ref) <= 1e-12 * max(np.abs(V).max(), 1e-300), (name, j, p, got[j], ref)
    finally:
        s.close()


def test_moment_norms_and_reconstruction():
    prob = P.gauss_periodic_2d(n=16, n_mom=5)
    s = stepped(prob, steps=6)
    try:
        M, D = O.moment_norms(s)
        sumsq, diffsq = O.moment_terms(s)
        order = np.asarray(s.par["mom_order"]).ravel()
        n1 = prob["n"][0]
        for l in range(len(M)):
            assert M[l] == pytest.approx(math.sqrt(sumsq[order == l].sum()) / n1, rel=1e-14)
            assert D[l] == pytest.approx(math.sqrt(diffsq[order == l].sum()) / n1, rel=1e-14)
        point = (0.1, 0.2)
        theta, phi, U_sol = O.angular_reconstruction(s, point, n_recon=20)
        assert U_sol.shape == (21, 21)
        u = O.probe(s, point)
        assert np.array_equal(U_sol, O.reconstruct(u, 5, theta, phi))
        # the angular mean of the reconstruction is the zeroth moment / sqrt(4 pi)
        x, w = np.polynomial.legendre.leggauss(8)
        nphi = 14
        g = O.reconstruct(u, 5, np.arccos(x), np.arange(nphi) * 2 * math.pi / nphi)
        mean = (g * w[:, None]).sum() * 2 * math.pi / nphi / (4 * math.pi)
        assert mean == pytest.approx(u[0] / math.sqrt(4 * math.pi), rel=1e-12)
    finally:
        s.close()
