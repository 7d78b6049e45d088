"""CPU checks of the host side of the observables (starmap_amd/observables.py).

The reference's scripts cannot run here (no MATLAB), so the reconstruction
is pinned by what the mathematics fixes: S is unitary, the sum over the
complex coefficients of real moments is real, Parseval's identity holds
under a quadrature exact for the products of harmonics, and the zeroth
moment alone gives the constant 1/sqrt(4 pi).  A wrong sign or conjugate in
S or in Y_l^{-m} leaves an imaginary part; a wrong factor breaks Parseval.
The orientation convention (Condon-Shortley phase, P_l^m at -cos(theta))
follows the script's formulas and stays unpinned against its plots.
"""
import math

import numpy as np
import pytest

from starmap_amd import init as I
from starmap_amd import observables as O
from starmap_amd import problems as P


@pytest.mark.parametrize("n_mom", [1, 2, 5, 9])
def test_real_to_complex_is_unitary(n_mom):
    S = O.real_to_complex(n_mom)
    assert np.abs(S.conj().T @ S - np.eye(S.shape[0])).max() < 1e-14


@pytest.mark.parametrize("n_mom", [1, 3, 6])
def test_reconstruction_real_and_parseval(n_mom):
    rng = np.random.default_rng(n_mom)
    n = (n_mom + 1) * (n_mom + 2) // 2
    u = rng.normal(size=n)
    x, w = np.polynomial.legendre.leggauss(n_mom + 3)
    theta = np.arccos(x)
    nphi = 2 * n_mom + 3
    phi = np.arange(nphi) * 2 * math.pi / nphi
    f = O.reconstruct(u, n_mom, theta, phi, return_complex=True)
    assert np.abs(f.imag).max() <= 1e-13 * np.abs(f.real).max()
    integral = (f.real ** 2 * w[:, None]).sum() * 2 * math.pi / nphi
    assert abs(integral - (u ** 2).sum()) <= 1e-12 * (u ** 2).sum()
    e0 = np.zeros(n)
    e0[0] = 1.0
    g = O.reconstruct(e0, n_mom, theta, phi)
    assert np.allclose(g, 1 / math.sqrt(4 * math.pi), rtol=0, atol=1e-15)


def test_spherical_harmonic_orthonormal():
    x, w = np.polynomial.legendre.leggauss(12)
    theta = np.arccos(x)
    phi = np.arange(24) * 2 * math.pi / 24
    ys = [(l, m, O.spherical_harmonic(l, m, theta, phi)) for l in range(5) for m in range(-l, l + 1)]
    for l1, m1, a in ys:
        for l2, m2, b in ys:
            ip = (a * np.conj(b) * w[:, None]).sum() * 2 * math.pi / 24
            assert abs(ip - (1.0 if (l1, m1) == (l2, m2) else 0.0)) < 1e-13, (l1, m1, l2, m2, ip)


def test_decay_rates_of_power_law():
    # norms ~ (l+1)^-p: every rate equals p when the orders differ by the
    # script's construction (orders-1 and orders for N1 -> N2)
    p = 2.5
    l = np.arange(40)
    M = (l + 0.0) ** -p
    M[0] = 1.0
    rates = O.decay_rates(M, [3, 5, 9, 17])
    assert rates.shape == (2, 3)
    assert np.allclose(rates, p, rtol=1e-13)


@pytest.mark.parametrize("n_mom", [3, 7, 13])
def test_order_grouping_matches_script_formula(n_mom):
    par = I.starmap_init(P.gauss_periodic_2d(n=8, n_mom=n_mom))
    n_sys = (n_mom + 1) * (n_mom + 2) // 2
    j = np.arange(1, n_sys + 1)
    script = np.ceil(np.sqrt(2 * j + 0.25) - 1.5).astype(int)
    assert np.array_equal(np.asarray(par["mom_order"]).ravel(), script)
