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
