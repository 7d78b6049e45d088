"""Parabolic time-step selection of the reference (host side, dense LAPACK
eigenvalues as in MATLAB): the largest eigenvalue of the von Neumann growth
factor of one full step, a doubling/halving bracket and a geometric
bisection for the largest stable dt.

* ``growth_factor``  starmap_solver.m:506-514
* ``select_dt``      starmap_solver.m:174-204 (the bracket and bisection)
* ``decay_rates``    the smallest sigma per component, :176-180 / :256-260

``par.timestep == 1`` bounds the eigenvalue by 1 (+1e-8); ``== 2`` measures
it relative to the decay of the zeroth moment (:513).
"""
import numpy as np


def expm1div(x):
    """(exp(x)-1)/x with the series 1 + x/2 + x^2/6 for |x| <= 2e-4
    (starmap_solver.m:531-535)."""
    x = np.asarray(x, dtype=np.float64)
    big = np.abs(x) > 2e-4
    out = 1 + x * .5 + x * x / 6
    xb = np.where(big, x, 1.0)
    return np.where(big, (np.exp(xb) - 1) / xb, out)


def symbol(Mx, My, Mz, h):
    """M = -1i*(Mx/h(1) + My/h(2) + Mz/h(3)) (starmap_solver.m:108)."""
    return -1j * (np.asarray(Mx, dtype=np.float64) / h[0] + np.asarray(My, dtype=np.float64) / h[1] +
                  np.asarray(Mz, dtype=np.float64) / h[2])


def growth_factor(M, c_even, c_odd, decay, dt, timestep):
    """Largest eigenvalue modulus of the growth-factor matrix of one step
    [odd, even, even, odd] with decay rates ``decay`` (one per component)."""
    decay = np.asarray(decay, dtype=np.float64)
    n = decay.size
    G = (dt * expm1div(-decay * dt / 2))[:, None] * (M + np.diag(-decay / 2))
    eye = np.eye(n)
    Ge = G.copy()
    Ge[list(c_odd), :] = 0
    Ge = Ge + eye
    Go = G.copy()
    Go[list(c_even), :] = 0
    Go = Go + eye
    la = float(np.max(np.abs(np.linalg.eigvals(Go @ Ge @ Ge @ Go))))
    if timestep == 2:
        la = la * np.exp(decay[0] * dt)
    return la


def select_dt(M, c_even, c_odd, decay, dt, tleft, timestep):
    """The bracket search and bisection of starmap_solver.m:181-203 from
    the initial guess dt.  Returns (dt_bracket[0], la0): the largest
    admissible step found (before the CFL factor) and the eigenvalue bound
    used (1 + 1e-8, lowered to 1 once an eigenvalue below 1 is seen)."""
    gf = lambda d: growth_factor(M, c_even, c_odd, decay, d, timestep)
    la = gf(dt)
    dtb = [dt, dt]
    lab = [la, la]
    la0 = 1 + 1e-8
    while (lab[0] < la0) == (lab[1] < la0) and (dtb[0] <= tleft or lab[0] > la0):
        stable = lab[0] < la0
        if stable:       # dt(1) stable: [dt(2), 2 dt(2)]
            dtb = [dtb[1], 2 * dtb[1]]
            lab = [lab[1], lab[0]]
            lab[1] = gf(dtb[1])
        else:            # dt(1) unstable: [dt(1)/2, dt(1)]
            dtb = [dtb[0] / 2, dtb[0]]
            lab = [lab[1], lab[0]]
            lab[0] = gf(dtb[0])
        if min(lab) < 1:
            la0 = 1
    while dtb[0] / dtb[1] < 0.98 and dtb[0] <= tleft:
        dt_mid = np.sqrt(dtb[0] * dtb[1])
        la_mid = gf(dt_mid)
        if la_mid >= la0:
            dtb[1], lab[1] = dt_mid, la_mid
        else:
            dtb[0], lab[0] = dt_mid, la_mid
    return float(dtb[0]), la0


def decay_rates(sA111, sT, gkey, mom_order):
    """decay(1) = min(sA{1,1,1}), decay(j) = min(sT{g_j, order_j}) for
    j >= 2 (0-based: component 0 and j >= 1)."""
    n = len(gkey)
    d = np.empty(n)
    d[0] = float(np.min(sA111))
    for j in range(1, n):
        d[j] = float(np.min(sT[gkey[j] + (int(mom_order[j]),)]))
    return d
