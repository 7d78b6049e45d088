"""P_N and SP_N moment-system matrices for the host side of the solver.

The flux matrices are uploaded once per problem as coupling tables.  The
entries follow the reference closures (starmap_closure_pn.m:28-98,
starmap_closure_spn.m:31-88): real-spherical-harmonic recurrence
coefficients D, F, B with the sqrt(2) rescaling of the m=0 harmonics, and the
Olbrant et al. SP_N blocks.  Indices here are 0-based.

Returns ``(mom_order, Mx, My, Mz)`` as dense float64 arrays in solver axes:
in 2D physical x is solver y and physical y is solver z (Mx = 0); in 1D the
matrix is Mz.
"""
import numpy as np

SQ2 = np.sqrt(2.0)


# divided by (2l+1) and then by (2l-1), as the reference evaluates them
def _coef_d(l, m):
    return np.sqrt((l - m) * (l - m - 1) / (2 * l + 1) / (2 * l - 1))


def _coef_f(l, m):
    return np.sqrt((l + m) * (l + m - 1) / (2 * l + 1) / (2 * l - 1))


def _coef_b(l, m):
    return np.sqrt((l - m) * (l + m) / (2 * l + 1) / (2 * l - 1))


def _legendre_1d(n_mom):
    n = n_mom + 1
    k = np.arange(n - 1)
    Mz = np.zeros((n, n))
    Mz[k, k + 1] = (k + 1) / (2 * k + 1)
    Mz[k + 1, k] = (k + 1) / (2 * k + 3)
    z = np.zeros((n, n))
    return np.array([n_mom + 1]), z, z.copy(), Mz


def _pn_2d(N):
    n = (N + 1) * (N + 2) // 2
    A = np.zeros((n, n))      # physical x
    B = np.zeros((n, n))      # physical y
    for m in range(1, N + 1):
        first = m * (m - 1) // 2          # first row of degree m-1
        for i in range(1, m + 1):
            r = first + i - 1
            s = 1 if i % 2 == 0 else -1   # (-1)^i
            v = _coef_d(m, -m + 2 * ((i + 1) // 2 - 1))
            A[r, r + m] = v
            B[r, r + m - s] = -s * v
        for i in range(1, m):
            r = first + i - 1
            s = 1 if i % 2 == 0 else -1
            v = _coef_f(m, -m + 2 + 2 * ((i + 1) // 2 - 1))
            A[r, r + m + 2] = -v
            B[r - s, r + m + 2] = s * v
    rows = [m * (m + 1) // 2 - 1 for m in range(1, N + 1, 2)]
    cols = [(m + 1) * (m + 2) // 2 - 1 for m in range(2, N + 1, 2)]
    for M in (A, B):
        M[rows, :] *= SQ2
        M[:, cols] *= SQ2
    A = (A + A.T) / 2
    B = (B + B.T) / 2
    j = np.arange(1, n + 1)
    order = np.ceil(np.sqrt(2 * j + 0.25) - 1.5).astype(np.int64)
    return order, np.zeros((n, n)), A, B


def _pn_3d(N):
    n = (N + 1) ** 2
    X = np.zeros((n, n))
    Y = np.zeros((n, n))
    Z = np.zeros((n, n))
    for m in range(1, N + 1):
        first = (m - 1) ** 2
        for i in range(1, 2 * m):
            r = first + i - 1
            s = 1 if i % 2 == 0 else -1
            c = (i + 1) // 2
            v = _coef_f(m, m + 1 - c)
            X[r, r + 2 * m - 1] = v
            Y[r, r + 2 * m - 1 - s] = -s * v
            Z[r, r + 2 * m + 1] = _coef_b(m, m - c)
        cnt = 2 * m - 3
        for i in range(1, cnt + 1):
            r = first + i - 1
            c = (i + 1) // 2
            v = _coef_d(m, m - 1 - c)
            X[r, r + 2 * m + 3] = -v
            ri, ii = r, i
            if m > 1 and i == cnt:
                ri, ii = r + 1, i + 1
            s = 1 if ii % 2 == 0 else -1
            Y[ri, ri + 2 * m + 3 - s] = -s * v
    rows = [m * m - 1 for m in range(1, N + 1)]
    cols = [(m + 1) ** 2 - 1 for m in range(2, N + 1)]
    for M in (X, Y):
        M[rows, :] *= SQ2
        M[:, cols] *= SQ2
    X = (X + X.T) / 2
    Y = (Y + Y.T) / 2
    Z = Z + Z.T
    order = np.ceil(np.sqrt(np.arange(1, n + 1)) - 1).astype(np.int64)
    return order, X, Y, Z


def closure_pn(n_mom, dim):
    if dim == 1:
        return _legendre_1d(n_mom)
    if dim == 2:
        return _pn_2d(n_mom)
    if dim == 3:
        return _pn_3d(n_mom)
    raise ValueError("dim must be 1, 2 or 3")


def closure_spn(n_mom, dim):
    if dim == 1:
        return _legendre_1d(n_mom)
    if dim not in (2, 3):
        raise ValueError("dim must be 1, 2 or 3")
    ns = (n_mom + 2) // 2                 # ceil((N+1)/2)
    i = np.arange(ns, dtype=np.float64)
    k = 2 * i * (2 * i - 1) / ((4 * i + 1) * (4 * i - 1))
    keep = ((i < ns - 1) | (n_mom % 2 == 1)).astype(np.float64)
    l = 4 * i ** 2 / ((4 * i + 1) * (4 * i - 1)) + (2 * i + 1) ** 2 / ((4 * i + 1) * (4 * i + 3)) * keep
    u = 2 * (2 * i + 1) * (i + 1) / ((4 * i + 1) * (4 * i + 3))
    blk = np.diag(l) + np.diag(k[1:], -1) + np.diag(u[:-1], 1)
    even = np.arange(0, n_mom + 1, 2)
    per = dim + 1                          # components per even moment: 1 + dim
    n = ns * per
    X = np.zeros((n, n))
    rows = ns + np.arange(ns) * dim        # first flux component of each block
    X[np.ix_(rows, np.arange(ns))] = blk
    X[np.ix_(np.arange(ns), rows)] = np.eye(ns)
    q = np.arange(ns * dim)
    head = np.arange(ns)
    if dim == 2:
        # swap the two flux components of every block
        perm_y = np.concatenate([head, ns + q + np.where(q % 2 == 0, 1, -1)])
        Y = X[np.ix_(perm_y, perm_y)]
        order = np.concatenate([even, np.repeat(even + 1, 2)]).astype(np.int64)
        return order, np.zeros((n, n)), X, Y
    # dim == 3: y swaps the first two flux components of every block,
    # z swaps the first and the third
    perm_y = np.concatenate([head, ns + q + np.array([1, -1, 0])[q % 3]])
    perm_z = np.concatenate([head, ns + q + np.array([2, 0, -2])[q % 3]])
    Y = X[np.ix_(perm_y, perm_y)]
    Z = X[np.ix_(perm_z, perm_z)]
    order = np.concatenate([even, np.repeat(even + 1, 3)]).astype(np.int64)
    return order, X, Y, Z
