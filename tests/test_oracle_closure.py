"""Pins of the oracle's closure restatement (oracle/closure.py) against closed
forms that do not restate its formulas.

* P_N and SP_N flux matrices: the largest eigenvalue magnitude is the
  largest root of the Legendre polynomial P_{N+1}.
* P_N matrices represent multiplication by the direction cosines in an
  orthonormal basis: they are symmetric, rotation invariant (every direction
  n gives the same spectrum for n.M), and in 3D Mx^2+My^2+Mz^2 = I on all
  moments of order < N (|Omega| = 1 is exact below the truncation order).
* 1D tridiagonal entries (k+1)/(2k+1) and (k+1)/(2k+3) of the Legendre
  recurrence.
* sizes, nonzero counts and grid-class sizes at the benchmark orders.
"""
import numpy as np
import pytest

from oracle.closure import closure_pn, closure_spn
from oracle.starmap_oracle import grid_classes


def legendre_max_root(n):
    c = np.zeros(n + 1)
    c[n] = 1.0
    return np.max(np.polynomial.legendre.legroots(c))


@pytest.mark.parametrize("N,dim", [(1, 2), (2, 2), (5, 2), (8, 2), (15, 2), (39, 2),
                                   (1, 3), (2, 3), (4, 3), (7, 3)])
def test_pn_lambda_max_is_legendre_root(N, dim):
    _, Mx, My, Mz = closure_pn(N, dim)
    lam = np.max(np.abs(np.linalg.eigvalsh(Mz)))
    assert abs(lam - legendre_max_root(N + 1)) < 2e-14


@pytest.mark.parametrize("N,dim", [(3, 2), (9, 2), (4, 2), (3, 3), (7, 3)])
def test_spn_lambda_max_is_legendre_root(N, dim):
    _, Mx, My, Mz = closure_spn(N, dim)
    lam = np.max(np.abs(np.linalg.eigvals(Mz)))
    assert abs(lam - legendre_max_root(N + 1)) < 2e-14


@pytest.mark.parametrize("N", [3, 6, 11])
def test_pn_2d_symmetric_and_rotation_invariant(N):
    _, Mx, My, Mz = closure_pn(N, 2)
    assert np.count_nonzero(Mx) == 0
    assert np.array_equal(My, My.T) and np.array_equal(Mz, Mz.T)
    ref = np.sort(np.linalg.eigvalsh(Mz))
    rng = np.random.default_rng(N)
    for th in rng.uniform(0, 2 * np.pi, 5):
        ev = np.sort(np.linalg.eigvalsh(np.cos(th) * My + np.sin(th) * Mz))
        assert np.max(np.abs(ev - ref)) < 1e-13


@pytest.mark.parametrize("N", [2, 4, 7])
def test_pn_3d_rotation_invariance_and_unit_sphere(N):
    _, Mx, My, Mz = closure_pn(N, 3)
    for M in (Mx, My, Mz):
        assert np.array_equal(M, M.T)
    ref = np.sort(np.linalg.eigvalsh(Mz))
    rng = np.random.default_rng(N)
    for _ in range(5):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        ev = np.sort(np.linalg.eigvalsh(n[0] * Mx + n[1] * My + n[2] * Mz))
        assert np.max(np.abs(ev - ref)) < 1e-13
    lam = legendre_max_root(N + 1)
    rho = np.max(np.abs(np.linalg.eigvalsh(Mx + My + Mz)))
    assert abs(rho - np.sqrt(3) * lam) < 1e-13
    S = Mx @ Mx + My @ My + Mz @ Mz
    low = N ** 2                       # components of order < N
    assert np.max(np.abs(S[:low, :low] - np.eye(low))) < 1e-14


@pytest.mark.parametrize("clo", [closure_pn, closure_spn])
def test_1d_tridiagonal(clo):
    N = 6
    mo, Mx, My, Mz = clo(N, 1)
    assert not Mx.any() and not My.any()
    for k in range(N):
        assert Mz[k, k + 1] == (k + 1) / (2 * k + 1)
        assert Mz[k + 1, k] == (k + 1) / (2 * k + 3)
    assert np.count_nonzero(Mz) == 2 * N
    assert mo[0] == N + 1


# sizes at the benchmark orders: n_sys, nnz(Mx,My,Mz), class sizes
@pytest.mark.parametrize("N,dim,n_sys,nnz,classes", [
    (5, 2, 21, (0, 50, 50), {"111": 6, "121": 6, "112": 6, "122": 3}),
    (39, 2, 820, (0, 3042, 3042), {"111": 210, "121": 210, "112": 210, "122": 190}),
    (7, 3, 64, (170, 170, 98), {"111": 10, "121": 10, "112": 10, "211": 10,
                                "122": 6, "212": 6, "221": 6, "222": 6}),
    (15, 2, 136, (0, 450, 450), {"111": 36, "121": 36, "112": 36, "122": 28}),
])
def test_sizes_and_classes(N, dim, n_sys, nnz, classes):
    mo, Mx, My, Mz = closure_pn(N, dim)
    assert Mz.shape == (n_sys, n_sys)
    assert tuple(np.count_nonzero(M) for M in (Mx, My, Mz)) == nnz
    gt, cls = grid_classes(Mx, My, Mz)
    got = {k: len(v) for k, v in cls.items() if v}
    assert got == classes
    # bipartite: every coupling joins components whose grid type differs
    # exactly along the coupling direction
    for d, M in enumerate((Mx, My, Mz)):
        j, k = np.nonzero(M)
        flip = np.zeros(3, dtype=bool)
        flip[d] = True
        assert np.all((gt[j] != gt[k]) == flip)


def test_grid_classes_match_bfs_two_colouring():
    """The fixed-point closure of the reference equals a breadth-first
    two-colouring of the coupling graph per axis."""
    for N, dim, clo in [(7, 3, closure_pn), (6, 2, closure_pn), (5, 3, closure_spn),
                        (9, 2, closure_spn)]:
        _, Mx, My, Mz = clo(N, dim)
        n = Mz.shape[0]
        gt_ref = np.zeros((n, 3), dtype=int)
        gt_ref[0] = 1
        seen = {0}
        queue = [0]
        while queue:
            j = queue.pop(0)
            for d, M in enumerate((Mx, My, Mz)):
                for k in np.nonzero(M[j])[0]:
                    g = gt_ref[j].copy()
                    g[d] = 3 - g[d]
                    if k in seen:
                        assert np.array_equal(gt_ref[k], g)
                    else:
                        gt_ref[k] = g
                        seen.add(k)
                        queue.append(k)
        gt, _ = grid_classes(Mx, My, Mz)
        assert np.array_equal(gt, gt_ref)


def test_moment_order_counts():
    mo, *_ = closure_pn(39, 2)
    assert np.array_equal(np.bincount(mo), np.arange(1, 41))
    mo, *_ = closure_pn(7, 3)
    assert np.array_equal(np.bincount(mo), 2 * np.arange(8) + 1)
