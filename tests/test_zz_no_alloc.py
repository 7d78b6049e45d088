"""smp_probe_noalloc -- the allocation-free variant of smp_probe -- against
smp_probe, through the C ABI.

Two calls in a row on the library's shared fixed array: first a P9 problem
(55 components), then a P3 problem (10 components, a smaller input), each
at interior points, a boundary point and a point outside (NaN); every value
must equal smp_probe's bit for bit (same table, same kernel).  The rejection
of a problem with more components than the array holds is checked on the CPU
in test_abi.py.
"""
import numpy as np
import pytest

from starmap_amd import _lib
from starmap_amd import init as I
from starmap_amd import problems as P
from starmap_amd.solver import HipSolver

pytestmark = pytest.mark.gpu


def _call(name, s, point):
    ax = np.asarray(s.par["ax"], dtype=np.float64)
    lo = np.ascontiguousarray(ax[[0, 2, 4]])
    sp = np.zeros(3)
    p = np.asarray(point, dtype=np.float64)
    sp[3 - p.size:] = p
    out = np.full(s.n_sys, 7.0)
    _lib.check(getattr(_lib.lib(), name)(s.handle, lo.ctypes.data, sp.ctypes.data, out.ctypes.data))
    return out


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_probe_noalloc_equals_probe(precision):
    # per point: all values finite ("in"), all NaN ("out"), or mixed (None:
    # on the domain edge some staggered grids do not reach)
    cases = [(P.lattice_2d(n=40, n_mom=9, tfinal=0.3),
              [((3.3, 3.6), "in"), ((0.0, 7.0), None), ((5.0, 1.2), "in"), ((8.0, 1.0), "out")]),
             (P.gauss_periodic_2d(n=24, n_mom=3, std=0.2, tfinal=0.2),
              [((0.1, -0.2), "in"), ((-1.5, 1.5), None), ((1.2, 0.7), "in")])]
    for prob, points in cases:
        s = HipSolver(I.starmap_init(prob), precision=precision)
        try:
            s.run(verbose=False)
            print("noalloc", precision, s.n_sys, flush=True)
            for pt, where in points:
                a = _call("smp_probe", s, pt)
                b = _call("smp_probe_noalloc", s, pt)
                assert np.array_equal(a, b, equal_nan=True), (pt, a, b)
                if where == "in":
                    assert np.all(np.isfinite(b)), (pt, b)
                elif where == "out":
                    assert np.all(np.isnan(b)), (pt, b)
        finally:
            s.close()
