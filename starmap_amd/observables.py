"""Observables the reference's example scripts compute from the solution,
on the device state of a solver handle (starmap_amd.solver.HipSolver).

* Moment-order norms and their decay rates
  (starmap_2d_ex_lattice_filter_convergence.m:63-82,
  starmap_2d_ex_gauss_filter_convergence.m): per moment order l,
  ``Moments(l) = sqrt(sum over components of order l of sum(U.^2)) / n(1)``
  and ``DiffXMoments(l)`` the same over the squared differences of
  neighbours along each output dimension.  The per-component sums come from
  ``smp_moment_norms`` (device reduction in FP64).
* Angular reconstruction at a point
  (starmap_2d_ex_linesource_filter_reconstruction.m:21-45, :112-121): every
  component interpolated at the point (``smp_probe``, linear on its own
  staggered grid), the real 2D P_N moments turned into complex spherical
  harmonic coefficients (the matrix S of :21-33) and summed against
  Y_l^m(theta, phi) (``sph`` of :135-141) on the reconstruction grid.

The staggered-trapezoid L2 norm of starmap_2d_ex_l2norm.m is
``HipSolver.observables``.
"""
import math

import numpy as np

from . import _lib


def moment_terms(solver):
    """(sumsq[n_sys], diffsq[n_sys, 3]): per component the sum of U_j^2 over
    its grid points and, per output dimension, the sum of squared
    differences of neighbours (MATLAB ``diff(U, [], d)``), in FP64."""
    n = solver.n_sys
    sumsq = np.zeros(n)
    diffsq = np.zeros(3 * n)
    _lib.check(_lib.lib().smp_moment_norms(solver.handle, sumsq.ctypes.data, diffsq.ctypes.data))
    return sumsq, diffsq.reshape(n, 3)


def moment_norms(solver, n_cells=None):
    """(Moments, DiffXMoments), indexed by moment order 0..N: the estimates
    of starmap_2d_ex_lattice_filter_convergence.m:63-76.  Components are
    grouped by the closure's moment order (par.mom_order, which equals the
    script's ``ceil(sqrt(2*(1:n_sys)+1/4)-3/2)`` for 2D P_N); n_cells is
    the script's prob.n(1), by default the cells of the first physical
    axis."""
    sumsq, diffsq = moment_terms(solver)
    order = np.asarray(solver.par["mom_order"]).ravel().astype(int)
    if n_cells is None:
        n_cells = int(np.asarray(solver.par["prob"]["n"]).ravel()[0])
    n_ord = int(order.max()) + 1
    m = np.zeros(n_ord)
    d = np.zeros(n_ord)
    for j in range(solver.n_sys):
        m[order[j]] += sumsq[j]
        d[order[j]] += diffsq[j].sum()
    return np.sqrt(m) / n_cells, np.sqrt(d) / n_cells


def decay_rates(norms, orders):
    """Decay rates of starmap_2d_ex_lattice_filter_convergence.m:79-82:
    for consecutive entries N1 < N2 of ``orders`` (moment approximation
    orders, MATLAB values), row 0 is the rate of the norms of moment order
    N-1 and row 1 of order N: ``-diff(log(M([orders; orders+1])), [], 2)
    ./ diff(log([orders-1; orders]), [], 2)``.  ``norms`` is indexed by
    moment order (norms[l] = the script's Moments(l+1))."""
    o = np.asarray(orders, dtype=np.int64)
    M = np.asarray(norms, dtype=np.float64)
    a = np.log(np.vstack([M[o - 1], M[o]]))
    b = np.log(np.vstack([o - 1, o]).astype(np.float64))
    return -np.diff(a, axis=1) / np.diff(b, axis=1)


def probe(solver, point):
    """Every component linearly interpolated at a physical point (x, y) or
    (x, y, z) on its own staggered grid; NaN outside the grid (MATLAB
    interp2/interp3 defaults)."""
    par = solver.par
    ax = np.asarray(par["ax"], dtype=np.float64)
    lo = np.ascontiguousarray(ax[[0, 2, 4]])
    p = np.asarray(point, dtype=np.float64).ravel()
    # solver axes: physical 2D (x, y) is solver (y, z), 3D is (x, y, z)
    sp = np.zeros(3)
    sp[3 - p.size:] = p
    out = np.zeros(solver.n_sys)
    _lib.check(_lib.lib().smp_probe(solver.handle, lo.ctypes.data, sp.ctypes.data, out.ctypes.data))
    return out


def _order_vec(n_sys):
    j = np.arange(1, n_sys + 1)
    order = np.ceil(np.sqrt(2 * j + 0.25) - 1.5).astype(int)
    vec = 2 * (j - 1) - order * (order + 2)
    return order, vec


def real_to_complex(n_mom):
    """The matrix S of starmap_2d_ex_linesource_filter_reconstruction.m:21-33
    for 2D P_N (n_sys = (N+1)(N+2)/2): real moments u = S c for complex
    spherical-harmonic coefficients c (order l = m-1 of the script's loop
    variable m; within an order the pairs (r-1, r) of real moments take
    the coefficients of +/-|vec|)."""
    n_sys = (n_mom + 1) * (n_mom + 2) // 2
    S = np.zeros((n_sys, n_sys), dtype=np.complex128)
    for m in range(2, n_mom + 2):
        for i in range(1, m // 2 + 1):
            r = m * (m - 1) // 2 + 2 * i          # 1-based row of the pair's second moment
            c_lo = m * (m - 1) // 2 + i           # 1-based columns
            c_hi = m * (m + 1) // 2 + 1 - i
            sgn = (-1.0) ** (m + 1)
            S[r - 2, c_lo - 1] = 1
            S[r - 1, c_lo - 1] = -1j
            S[r - 2, c_hi - 1] = sgn
            S[r - 1, c_hi - 1] = sgn * 1j
    S /= math.sqrt(2.0)
    for m in range(1, n_mom + 2, 2):
        S[m * (m + 1) // 2 - 1, (m * m + 1) // 2 - 1] = 1
    return S


def spherical_harmonic(l, m, theta, phi):
    """Y_l^m on the grid theta (rows) x phi (columns) as the script's ``sph``:
    sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) (-1)^m P_l^m(-cos theta) e^{i m phi}
    with the Condon-Shortley phase in P_l^m (as MATLAB's legendre), and
    Y_l^{-m} = (-1)^m conj(Y_l^m)."""
    from scipy.special import lpmv
    theta = np.asarray(theta, dtype=np.float64).ravel()
    phi = np.asarray(phi, dtype=np.float64).ravel()
    if m < 0:
        return (-1.0) ** m * np.conj(spherical_harmonic(l, -m, theta, phi))
    norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - m) / math.factorial(l + m))
    p = lpmv(m, l, -np.cos(theta))
    return (norm * (-1.0) ** m * p)[:, None] * np.exp(1j * m * phi)[None, :]


def reconstruct(u_point, n_mom, theta, phi, return_complex=False):
    """Angular function of the real 2D P_N moments u_point on theta x phi:
    ``real(sum_j Y_j * (S \\ u)_j)`` (:119-120 of the reconstruction
    script)."""
    n_sys = (n_mom + 1) * (n_mom + 2) // 2
    u = np.asarray(u_point, dtype=np.float64).ravel()
    if u.size != n_sys:
        raise ValueError("u_point must have (N+1)(N+2)/2 = %d entries" % n_sys)
    c = np.linalg.solve(real_to_complex(n_mom), u.astype(np.complex128))
    order, vec = _order_vec(n_sys)
    f = np.zeros((np.size(theta), np.size(phi)), dtype=np.complex128)
    for j in range(n_sys):
        f += spherical_harmonic(int(order[j]), int(vec[j]), theta, phi) * c[j]
    return f if return_complex else f.real


def angular_reconstruction(solver, point, n_recon=100):
    """(theta, phi, U_sol): the angular reconstruction of a 2D P_N solution
    at a point, on the script's grid theta = 0:pi/n:pi, phi = 0:2pi/n:2pi."""
    par = solver.par
    if par.get("closure", "P") != "P" or int(par.get("dim", 2)) != 2:
        raise ValueError("the angular reconstruction is defined for 2D P_N")
    n_mom = int(par["n_mom"])
    delta = math.pi / n_recon
    theta = np.arange(n_recon + 1) * delta
    phi = np.arange(n_recon + 1) * 2 * delta
    return theta, phi, reconstruct(probe(solver, point), n_mom, theta, phi)
