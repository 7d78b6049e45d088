"""Problem definitions for the benchmark and test workloads.

These play the role of the reference's example scripts
(``starmap_2d_ex_*.m``, ``starmap_3d_ex_*.m``): each builder returns a
``prob`` dict for ``starmap_init`` with problem functions that follow the
reference's argument conventions (README.txt:100-158), e.g. 2D
``sigma_a(x, y)``, ``source(x, y)``, ``filterfunction(par, dt, k)``.

Every builder takes the grid size so that tests can run the same problem
at reduced sizes.
"""
import numpy as np


def legendre_max_root(n):
    """Largest root of the Legendre polynomial P_n: the largest eigenvalue
    magnitude of the 1D P_{n-1} flux matrix, and of Mz for P_{n-1} in 2D/3D."""
    c = np.zeros(n + 1)
    c[n] = 1.0
    return float(np.max(np.polynomial.legendre.legroots(c)))


def gaussian_2d(std):
    """Isotropic Gaussian of unit mass (starmap_2d_ex_l2norm.m:56-59 form:
    1/(4 pi tau) exp(-r^2/(4 tau)) with tau = std^2/2)."""
    def ic(x, y):
        return np.exp(-(x ** 2 + y ** 2) / (2 * std ** 2)) / (2 * np.pi * std ** 2)
    return ic


def gaussian_3d(std):
    def ic(x, y, z):
        return np.exp(-(x ** 2 + y ** 2 + z ** 2) / (2 * std ** 2)) / \
            (2 * np.pi * std ** 2) ** 1.5
    return ic


def exp_filter(p, f_eff):
    """Exponential FP_N filter, starmap_2d_ex_lattice_filter_convergence.m:143-147:
    f = exp(log(eps)*(k/(N+1))^p)^(dt*f_b),
    f_b = -f_eff/log(exp(log(eps)*(N/(N+1))^p))."""
    eps = np.finfo(np.float64).eps

    def filterfunction(par, dt, k):
        N = par["n_mom"]
        f_b = -f_eff / np.log(np.exp(np.log(eps) * (N / (N + 1)) ** p))
        return np.exp(np.log(eps) * (k / (N + 1)) ** p) ** (dt * f_b)
    return filterfunction


def gauss_periodic_2d(n=100, n_mom=5, closure="P", tfinal=1.0, std=0.03, half=1.5,
                      sigma_a=0.0, sigma_s0=1.0, mom_output="all"):
    """2D P_N, periodic [-half,half]^2, unit-mass Gaussian zeroth moment,
    homogeneous sigma_a/sigma_s0, dt = 0.5*dx (cn = lambda_max).
    Defaults: 2D P_5, 100x100, std 0.03, t_final 1.0."""
    lam = legendre_max_root(n_mom + 1)
    return {
        "name": "Gaussian periodic 2D",
        "closure": closure,
        "n_mom": n_mom,
        "ic": gaussian_2d(std),
        "sigma_a": (lambda x, y: sigma_a),
        "sigma_s0": (lambda x, y: sigma_s0),
        "ax": [-half, half, -half, half],
        "n": [n, n],
        "bc": [0, 0],
        "cn": lam,
        "t_plot": [tfinal],
        "tfinal": tfinal,
        "mom_output": mom_output,
        "output": None,
    }


def lattice_blocks(n_absorbers=12, seed=2024):
    """Seeded choice of absorber blocks in the 7x7 checkerboard of unit blocks
    on [0,7]^2: n_absorbers blocks drawn without replacement from the 48
    non-centre blocks with numpy's default_rng(seed).  Returns a 7x7 0/1
    array indexed [bx-1, by-1]."""
    cand = [(bx, by) for bx in range(1, 8) for by in range(1, 8) if (bx, by) != (4, 4)]
    rng = np.random.default_rng(seed)
    pick = rng.choice(len(cand), size=n_absorbers, replace=False)
    g = np.zeros((7, 7))
    for i in pick:
        bx, by = cand[i]
        g[bx - 1, by - 1] = 1.0
    return g


def _block_lookup(table, x, y):
    # ceil() block convention of starmap_2d_ex_lattice.m:50-53; coordinate 0
    # (the left/bottom boundary faces) is clamped to block 1
    cx = np.clip(np.ceil(x), 1, 7).astype(np.int64)
    cy = np.clip(np.ceil(y), 1, 7).astype(np.int64)
    return table[cx - 1, cy - 1]


def lattice_2d(n=1000, n_mom=39, closure="P", tfinal=3.2, n_absorbers=12, seed=2024,
               filter_p=4, f_eff=10.0, f_position="substep", mom_output=1,
               t_plot=None):
    """2D P_N lattice on [0,7]^2 with extrapolation BC (bc=[1 1], as
    starmap_2d_ex_lattice.m:38): absorber blocks sigma_a=10/sigma_s0=0,
    other blocks sigma_a=0/sigma_s0=1, unit isotropic source in the centre
    block (3,4)^2, zero initial state, exponential filter order filter_p and
    strength f_eff at 'substep' position, dt = 0.5*dx."""
    lam = legendre_max_root(n_mom + 1)
    g = lattice_blocks(n_absorbers, seed)

    def sigma_a(x, y):
        return _block_lookup(g, x, y) * 10.0

    def sigma_s0(x, y):
        return 1.0 - _block_lookup(g, x, y)

    def source(x, y):
        return ((3 < x) & (x < 4) & (3 < y) & (y < 4)).astype(np.float64)

    prob = {
        "name": "Lattice 2D",
        "closure": closure,
        "n_mom": n_mom,
        "sigma_a": sigma_a,
        "sigma_s0": sigma_s0,
        "source": source,
        "ax": [0.0, 7.0, 0.0, 7.0],
        "n": [n, n],
        "bc": [1, 1],
        "cn": lam,
        "t_plot": [tfinal] if t_plot is None else t_plot,
        "tfinal": tfinal,
        "mom_output": mom_output,
        "output": None,
    }
    if f_position:
        prob["filterfunction"] = exp_filter(filter_p, f_eff)
        prob["f_position"] = f_position
    return prob


def gauss_periodic_3d(n=200, n_mom=7, closure="P", tfinal=0.75, std=0.05,
                      sigma_a=0.1, sigma_s0=1.0, aniso=0.5, mom_output=1):
    """3D P_N, periodic [-1,1]^3, Gaussian zeroth moment of std 0.05,
    sigma_a=0.1, sigma_s0=1, anisotropic sigma_sm = aniso^m (m >= 1),
    dt = 0.5*dx (cn = 1.5*lambda_max, since dt_hyp = h/(3 lambda))."""
    lam = legendre_max_root(n_mom + 1)
    prob = {
        "name": "Gaussian periodic 3D",
        "closure": closure,
        "n_mom": n_mom,
        "ic": gaussian_3d(std),
        "sigma_a": (lambda x, y, z: sigma_a),
        "sigma_s0": (lambda x, y, z: sigma_s0),
        "ax": [-1.0, 1.0, -1.0, 1.0, -1.0, 1.0],
        "n": [n, n, n],
        "bc": [0, 0, 0],
        "cn": 1.5 * lam,
        "t_plot": [tfinal],
        "tfinal": tfinal,
        "mom_output": mom_output,
        "output": None,
    }
    if aniso is not None:
        prob["sigma_sm"] = (lambda x, y, z, m: aniso ** m)
    return prob


def boxes_2d(n=50, n_mom=9, closure="P", tfinal=0.2):
    """starmap_2d_ex_boxes.m:38-80: homogeneous sigma_a=0.9, sigma_s0=0.1,
    piecewise-constant time-dependent source, bc=[1 1]; the P_N and SP_N
    solutions coincide for this setup (starmap_2d_ex_boxes.m:6-16)."""
    def source(x, y, t):
        box = (((1.75 < x) & (x <= 2.25) & (1.75 < y) & (y <= 2.25)) |
               ((2.75 < x) & (x <= 3.25) & (1.5 < y) & (y <= 2.50)) |
               ((1.75 < x) & (x <= 2.25) & (2.75 < y) & (y <= 3.25)) |
               ((3.50 < x) & (x <= 4.25) & (3.5 < y) & (y <= 3.75)))
        return (2 + np.sin(4 * np.pi * t) * np.exp(-t / 3)) * box
    return {
        "name": "Boxes Test",
        "closure": closure,
        "n_mom": n_mom,
        "sigma_a": (lambda x, y: 0.9),
        "sigma_s0": (lambda x, y: 0.1),
        "source": source,
        "ax": [0.0, 5.0, 0.0, 5.0],
        "n": [n, n],
        "bc": [1, 1],
        "t_plot": [tfinal],
        "tfinal": tfinal,
        "output": None,
    }


def gauss_1d(n=1000, n_mom=19, closure="P", tfinal=0.4, n_plot=21):
    """starmap_1d_ex_gauss.m:32-65: 1D P19 on [0,1] with 1000 cells,
    periodic, zeroth moment exp(-500 (x-1/2)^2), sigma_a = sigma_s0 = 1.5
    (constant handles returning scalars), outputs at 21 times up to 0.4."""
    return {
        "name": "1D Gaussian",
        "closure": closure,
        "n_mom": n_mom,
        "n": [n],
        "sigma_a": lambda x: 1.5,
        "sigma_s0": lambda x: 1.5,
        "ic": lambda x, k: np.exp(-500 * (x - 0.5) ** 2) * (k == 1),
        "ax": [0.0, 1.0],
        "bc": [0],
        "t_plot": list(np.linspace(0, tfinal, n_plot)),
        "output": None,
    }


def planesource_1d(n=1000, n_mom=7, closure="SP", tfinal=0.4, n_plot=21):
    """starmap_1d_ex_planesource.m:32-67: 1D SP7 on [-0.6,0.6] with 1000
    cells, periodic, the plane source smoothed to a Gaussian of pseudo-time
    3.2e-3, pure scattering (sigma_a = 0, sigma_s0 = 1)."""
    t = 3.2e-3
    return {
        "name": "Plane Source Test",
        "closure": closure,
        "n_mom": n_mom,
        "n": [n],
        "sigma_a": lambda x: 0.0,
        "sigma_s0": lambda x: 1.0,
        "ic": lambda x, k: 1 / (4 * np.pi * t) * np.exp(-(x ** 2) / (4 * t)) * (k == 1),
        "ax": [-0.6, 0.6],
        "bc": [0],
        "t_plot": list(np.linspace(0, tfinal, n_plot)),
        "output": None,
    }


def controlrod_2d(n=251, n_mom=3, closure="SP", tfinal=0.9, n_plot=46):
    """starmap_2d_ex_controlrod.m:35-81: a control rod moves in and out of a
    multiplying medium.  Effective absorption rod(x,y)*s(t), with s rising
    linearly to 10 over [0,0.2], held to 0.4 and falling to 0 at 0.6;
    effective scattering sigma_s + nu*sigma_f = 1 + 0.9*2; initial zeroth
    moment 1e6; periodic [-1,1]^2.  sigma_a takes t, so the solver runs its
    time-dependent branch (materials averaged over each step, the last step
    ending on tfinal)."""
    def rod(x, y):
        return (((-0.5 <= x) & (x <= 0.2) & (-0.5 <= y) & (y <= 0)) |
                ((-0.3 <= x) & (x <= 0) & (0 < y) & (y <= 0.6)) |
                ((0 < x) & (x <= 0.5) & (0 < y) & (y < 0.3))).astype(np.float64)

    def time_dependence(t):
        T, dT, smax = 0.6, 0.2, 10.0
        return (smax / dT * t * ((0 < t) & (t <= dT)) + smax * ((dT < t) & (t <= T - dT)) +
                smax / dT * (T - t) * ((T - dT < t) & (t <= T)))

    return {
        "name": "Control Rod Test",
        "closure": closure,
        "n_mom": n_mom,
        "sigma_a": lambda x, y, t: rod(x, y) * time_dependence(t),
        "sigma_s0": lambda x, y: (1.0 + 0.9 * 2.0) * np.ones(np.shape(x)),
        "ic": lambda x, y: 1e6 * np.ones(np.shape(x)),
        "ax": [-1.0, 1.0, -1.0, 1.0],
        "n": [n, n],
        "bc": [0, 0],
        "t_plot": list(np.linspace(0, tfinal, n_plot)),
        "tfinal": tfinal,
        "output": None,
    }


def exp_filter_moving(p, f_eff, speed=2.0):
    """A filter with the argument list (par, dt, k, x, y, t), which the
    solver re-evaluates every step (starmap_solver.m:284): the exponential
    filter of order p whose strength varies in space and time,
    exp(-dt * f_eff * (1 + 0.5 sin(speed t) + 0.1 x) * (k/N)^p)."""
    def filterfunction(par, dt, k, x, y, t):
        N = par["n_mom"]
        return np.exp(-dt * f_eff * (1 + 0.5 * np.sin(speed * t) + 0.1 * x) * (k / N) ** p)
    return filterfunction


BATCH_SIDES = tuple(range(64, 513, 32))     # 64, 96, ..., 512
BATCH_ORDERS = tuple(range(3, 40, 2))       # odd N = 3, 5, ..., 39


def batch_draws(n_instances=64, seed=2025, sides=BATCH_SIDES, orders=BATCH_ORDERS):
    """Seeded parameters of the batched sweep: per instance a grid side, an
    odd moment order N, a filter strength f_eff ~ U[0,20] and, on a 4x4 block
    partition of the domain, sigma_a and sigma_s0 log-uniform on [0.01,10].
    One numpy default_rng(seed) draws, in this order: sides (choice), orders
    (choice), f_eff (uniform), then per instance 16 sigma_a and 16 sigma_s0
    exponents (uniform on [-2,1])."""
    rng = np.random.default_rng(seed)
    side = rng.choice(np.asarray(sides), size=n_instances)
    order = rng.choice(np.asarray(orders), size=n_instances)
    f_eff = rng.uniform(0.0, 20.0, size=n_instances)
    blocks = []
    for _ in range(n_instances):
        sa = 10.0 ** rng.uniform(-2.0, 1.0, size=(4, 4))
        ss = 10.0 ** rng.uniform(-2.0, 1.0, size=(4, 4))
        blocks.append((sa, ss))
    return [dict(side=int(side[i]), n_mom=int(order[i]), f_eff=float(f_eff[i]),
                 sigma_a=blocks[i][0], sigma_s0=blocks[i][1]) for i in range(n_instances)]


def _block4(table, x, y):
    # 4x4 blocks of width 0.5 on [-1,1]^2 with the ceil() convention of
    # starmap_2d_ex_lattice.m:50-53: a face on a block boundary belongs to
    # the block on its left/lower side; the boundary at -1 to block 1
    bx = np.clip(np.ceil((np.asarray(x) + 1.0) / 0.5), 1, 4).astype(np.int64)
    by = np.clip(np.ceil((np.asarray(y) + 1.0) / 0.5), 1, 4).astype(np.int64)
    return table[bx - 1, by - 1]


def batch_instance(side, n_mom, f_eff, sigma_a, sigma_s0, tfinal=1.0, std=0.05,
                   filter_p=4, closure="P"):
    """One instance of the batched sweep: 2D P_N on the periodic [-1,1]^2
    with side x side cells, unit-mass Gaussian zeroth moment (std 0.05),
    piecewise-constant sigma_a / sigma_s0 on 4x4 blocks, exponential filter
    of order filter_p and strength f_eff at 'substep' position,
    dt = 0.5*dx (cn = lambda_max)."""
    lam = legendre_max_root(n_mom + 1)
    sa, ss = np.asarray(sigma_a), np.asarray(sigma_s0)
    return {
        "name": "Batched instance",
        "closure": closure,
        "n_mom": n_mom,
        "ic": gaussian_2d(std),
        "sigma_a": (lambda x, y: _block4(sa, x, y)),
        "sigma_s0": (lambda x, y: _block4(ss, x, y)),
        "ax": [-1.0, 1.0, -1.0, 1.0],
        "n": [side, side],
        "bc": [0, 0],
        "cn": lam,
        "t_plot": [],
        "tfinal": tfinal,
        "mom_output": 1,
        "output": None,
        "filterfunction": exp_filter(filter_p, f_eff),
        "f_position": "substep",
    }


def batch_sweep(n_instances=64, seed=2025, tfinal=1.0, side_div=1, max_order=None):
    """The batched sweep: 64 seeded instances (batch_draws).  side_div
    divides every side (reduced sizes for tests); max_order caps N."""
    out = []
    for d in batch_draws(n_instances, seed):
        n_mom = d["n_mom"] if max_order is None else min(d["n_mom"], max_order)
        out.append(batch_instance(max(4, d["side"] // side_div), n_mom, d["f_eff"],
                                  d["sigma_a"], d["sigma_s0"], tfinal=tfinal))
    return out


def asymptotic_preserving_2d(n=150, n_mom=1, ep=1e-3, timestep=1, sigma_a=0.0, n_plot=21,
                             time_dependent=False, half=0.6):
    """starmap_2d_ex_asymptotic_preserving.m: 2D P_N on [-half,half]^2 with
    extrapolation BC, isotropic scattering 1/ep (times 1 + 0.5 sin(t) if
    time_dependent), smoothed Dirac initial zeroth moment (pseudo-time
    3.2e-4), output times linspace(0, 0.005, n_plot)/ep; timestep 0
    (hyperbolic), 1 or 2 (parabolic)."""
    def ic(x, y, k):
        t0 = 3.2e-4
        return 1 / (4 * np.pi * t0) * np.exp(-(x ** 2 + y ** 2) / (4 * t0)) * (k == 1)

    prob = {
        "name": "Asymptotic Preserving Test",
        "closure": "P",
        "n_mom": n_mom,
        "sigma_a": (lambda x, y: sigma_a + 0 * x),
        "ic": ic,
        "ax": [-half, half, -half, half],
        "n": [n, n],
        "bc": [1, 1],
        "t_plot": list(np.linspace(0, 0.005, n_plot) / ep),
        "timestep": timestep,
        "output": None,
    }
    if time_dependent:
        prob["sigma_s0"] = (lambda x, y, t: (1 + 0.5 * np.sin(t)) / ep + 0 * x)
    else:
        prob["sigma_s0"] = (lambda x, y: 1 / ep + 0 * x)
    return prob


def gaussian_random_field(n, corr=0.2, std=0.5, seed=2026, length=2.0):
    """Periodic Gaussian random field on the n x n cell grid of a square of
    side `length`: white noise from numpy default_rng(seed) filtered in
    Fourier space with a Gaussian kernel so that the covariance is
    exp(-r^2 / (2 corr^2)), then shifted and scaled to mean 0 and std `std`."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((n, n))
    k = 2 * np.pi * np.fft.fftfreq(n, d=length / n)
    kx, ky = np.meshgrid(k, k, indexing="ij")
    # kernel of std corr/sqrt(2): the smoothed noise has covariance std corr
    s = corr / np.sqrt(2.0)
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * np.exp(-(kx ** 2 + ky ** 2) * s ** 2 / 2)))
    f = f - f.mean()
    return f * (std / f.std())


def cell_owner(n, half=1.0):
    """Index (ix*n + iy) of the cell that owns a point (x, y) of [-half,half]^2
    with the ceil() convention of starmap_2d_ex_lattice.m:50-53: a face on a
    cell boundary belongs to the cell on its left/lower side, the boundary
    at -half to the first cell."""
    h = 2.0 * half / n

    def owner(x, y):
        ix = np.clip(np.ceil((np.asarray(x) + half) / h), 1, n).astype(np.int64) - 1
        iy = np.clip(np.ceil((np.asarray(y) + half) / h), 1, n).astype(np.int64) - 1
        return ix * n + iy
    return owner


def adjoint_grf(n=512, n_mom=15, steps=2000, corr=0.2, grf_std=0.5, seed=2026,
                sigma_a=None, sigma_s0=None, src_std=0.05, closure="P"):
    """2D P_N on [-1,1]^2 with extrapolation BC, per-cell sigma_a =
    0.5*exp(Gaussian random field, correlation length corr, std grf_std)
    and sigma_s0 = 1 (or the given n x n arrays), unit-mass Gaussian source
    of std src_std at (-0.5, 0), zero initial state, no filter,
    dt = 0.5*dx, `steps` steps (t_final = steps*dt, exact in binary for
    n a power of two).  Materials on the staggered grids take the value of
    the owning cell (cell_owner)."""
    lam = legendre_max_root(n_mom + 1)
    if sigma_a is None:
        sigma_a = 0.5 * np.exp(gaussian_random_field(n, corr, grf_std, seed))
    if sigma_s0 is None:
        sigma_s0 = np.ones((n, n))
    sa = np.asarray(sigma_a, dtype=np.float64).ravel()
    ss = np.asarray(sigma_s0, dtype=np.float64).ravel()
    owner = cell_owner(n)
    dt = 0.5 * 2.0 / n

    def source(x, y):
        return np.exp(-((x + 0.5) ** 2 + y ** 2) / (2 * src_std ** 2)) / (2 * np.pi * src_std ** 2)

    return {
        "name": "Adjoint GRF 2D",
        "closure": closure,
        "n_mom": n_mom,
        "sigma_a": (lambda x, y: sa[owner(x, y)]),
        "sigma_s0": (lambda x, y: ss[owner(x, y)]),
        "source": source,
        "ax": [-1.0, 1.0, -1.0, 1.0],
        "n": [n, n],
        "bc": [1, 1],
        "cn": lam,
        "t_plot": [],
        "tfinal": steps * dt,
        "mom_output": 1,
        "output": None,
        "_owner": owner,
    }
