"""Oracle: P_N and SP_N moment-system matrices (CPU, numpy, FP64).

TEST INFRASTRUCTURE ONLY.  This module is part of the oracle: it may be
imported by ``tests/``, by ``__graft_entry__.smoke()`` and by the CPU-baseline
leg of ``bench.py``, and by nothing else.  The product path has its own
closure builder.

A statement-by-statement restatement of the reference closures, kept in
MATLAB's 1-based index arithmetic (converted to 0-based only at the final
array store) so that each line can be checked against the source:

* ``closure_pn``  follows starmap_closure_pn.m:28-98
* ``closure_spn`` follows starmap_closure_spn.m:31-88

Both return ``(moment_order, Mx, My, Mz)`` as dense float64 arrays in the
solver's axis convention (2D: physical x -> solver y, physical y -> solver z,
Mx = 0; 1D: everything in Mz).

Pinned by tests/test_oracle_closure.py against closed forms (symmetry,
Legendre roots, rotation invariance, explicit entries).
"""
import numpy as np


def _D(l, m):
    # starmap_closure_pn.m:91-92
    return np.sqrt((l - m) * (l - m - 1) / (2 * l + 1) / (2 * l - 1))


def _F(l, m):
    # starmap_closure_pn.m:94-95
    return np.sqrt((l + m) * (l + m - 1) / (2 * l + 1) / (2 * l - 1))


def _B(l, m):
    # starmap_closure_pn.m:97-98
    return np.sqrt((l - m) * (l + m) / (2 * l + 1) / (2 * l - 1))


def _ceil_half(i):
    return np.ceil(i / 2.0)


def _assign(M, rows1, cols1, vals):
    """MATLAB ``M(sub2ind(size(M),rows,cols)) = vals`` with 1-based rows/cols;
    later duplicates overwrite earlier ones, as in MATLAB."""
    rows1 = np.atleast_1d(np.asarray(rows1, dtype=np.int64))
    cols1 = np.atleast_1d(np.asarray(cols1, dtype=np.int64))
    vals = np.broadcast_to(np.asarray(vals, dtype=np.float64), rows1.shape)
    for r, c, v in zip(rows1, cols1, vals):
        M[r - 1, c - 1] = v


def _tridiag_1d(n_mom):
    # starmap_closure_pn.m:29-42 (identical in starmap_closure_spn.m:32-45)
    n_sys = n_mom + 1
    sub = np.zeros(n_sys)
    sup = np.zeros(n_sys)
    for k in range(0, n_sys - 1):
        sub[k] = (k + 1) / (2 * k + 3)       # subdiag(k+1)
        sup[k + 1] = (k + 1) / (2 * k + 1)   # superdiag(k+2)
    Mz = np.zeros((n_sys, n_sys))
    # spdiags with m>=n: sub-diagonal takes the upper part of its column,
    # super-diagonal the lower part.
    for i in range(n_sys - 1):
        Mz[i + 1, i] = sub[i]
        Mz[i, i + 1] = sup[i + 1]
    Mx = np.zeros((n_sys, n_sys))
    My = np.zeros((n_sys, n_sys))
    return np.array([n_mom + 1]), Mx, My, Mz


def closure_pn(n_mom, dim):
    """starmap_closure_pn.m:28-89."""
    if dim == 1:
        return _tridiag_1d(n_mom)
    if dim == 2:
        n_sys = (n_mom + 1) * (n_mom + 2) // 2
        Mx = np.zeros((n_sys, n_sys))
        My = np.zeros((n_sys, n_sys))
        for m in range(1, n_mom + 1):
            i = np.arange(1, m + 1)
            p = m * (m - 1) // 2 + i
            v = _D(m, -m + 2 * (_ceil_half(i) - 1))
            _assign(Mx, p, p + m, v)
            sgn = (-1.0) ** i
            _assign(My, p, p + m - sgn.astype(np.int64), -sgn * v)
            i = np.arange(1, m)
            p = m * (m - 1) // 2 + i
            v = _F(m, -m + 2 + 2 * (_ceil_half(i) - 1))
            _assign(Mx, p, p + m + 2, -v)
            sgn = (-1.0) ** i
            _assign(My, p - sgn.astype(np.int64), p + m + 2, sgn * v)
        m = np.arange(1, n_mom + 1, 2)
        i = m * (m + 1) // 2
        Mx[i - 1, :] *= np.sqrt(2.0)
        My[i - 1, :] *= np.sqrt(2.0)
        m = np.arange(2, n_mom + 1, 2)
        i = (m + 1) * (m + 2) // 2
        Mx[:, i - 1] *= np.sqrt(2.0)
        My[:, i - 1] *= np.sqrt(2.0)
        Mx = (Mx + Mx.T) / 2
        My = (My + My.T) / 2
        Mz = My
        My = Mx
        Mx = np.zeros((n_sys, n_sys))
        j = np.arange(1, n_sys + 1)
        moment_order = np.ceil(np.sqrt(2 * j + 0.25) - 1.5).astype(np.int64)
        return moment_order, Mx, My, Mz
    if dim == 3:
        n_sys = (n_mom + 1) ** 2
        Mx = np.zeros((n_sys, n_sys))
        My = np.zeros((n_sys, n_sys))
        Mz = np.zeros((n_sys, n_sys))
        for m in range(1, n_mom + 1):
            i = np.arange(1, 2 * m)
            p = (m - 1) ** 2 + i
            v = _F(m, m + 1 - _ceil_half(i))
            _assign(Mx, p, p + 2 * m - 1, v)
            sgn = (-1.0) ** i
            _assign(My, p, p + 2 * m - 1 - sgn.astype(np.int64), -sgn * v)
            v = _B(m, m - _ceil_half(i))
            _assign(Mz, p, p + 2 * m + 1, v)
            i = np.arange(1, 2 * m - 2)
            p = (m - 1) ** 2 + i
            v = _D(m, m - 1 - _ceil_half(i))
            _assign(Mx, p, p + 2 * m + 3, -v)
            if m > 1:
                p = p.copy()
                i = i.copy()
                p[-1] += 1
                i[-1] += 1
            sgn = (-1.0) ** i
            _assign(My, p, p + 2 * m + 3 - sgn.astype(np.int64), -sgn * v)
        m = np.arange(1, n_mom + 1)
        i = m ** 2
        Mx[i - 1, :] *= np.sqrt(2.0)
        My[i - 1, :] *= np.sqrt(2.0)
        m = np.arange(2, n_mom + 1)
        i = (m + 1) ** 2
        Mx[:, i - 1] *= np.sqrt(2.0)
        My[:, i - 1] *= np.sqrt(2.0)
        Mx = (Mx + Mx.T) / 2
        My = (My + My.T) / 2
        Mz = Mz + Mz.T
        j = np.arange(1, n_sys + 1)
        moment_order = np.ceil(np.sqrt(j) - 1).astype(np.int64)
        return moment_order, Mx, My, Mz
    raise ValueError("dim must be 1, 2 or 3")


def _diag(v, k, n):
    """MATLAB diag(v,k) embedded in an n x n matrix."""
    M = np.zeros((n, n))
    for r, val in enumerate(v):
        if k >= 0:
            M[r, r + k] = val
        else:
            M[r - k, r] = val
    return M


def closure_spn(n_mom, dim):
    """starmap_closure_spn.m:31-88."""
    if dim == 1:
        return _tridiag_1d(n_mom)
    nstar = int(np.ceil((n_mom + 1) / 2))
    i = np.arange(0, nstar, dtype=np.float64)
    k = 2 * i * (2 * i - 1) / ((4 * i + 1) * (4 * i - 1))
    odd_top = (i < nstar - 1) | (np.floor(n_mom / 2) * 2 < n_mom)
    l = 4 * i ** 2 / ((4 * i + 1) * (4 * i - 1)) + \
        (2 * i + 1) ** 2 / ((4 * i + 1) * (4 * i + 3)) * odd_top
    m = 2 * (2 * i + 1) * (i + 1) / ((4 * i + 1) * (4 * i + 3))
    block = _diag(k[1:], -1, nstar) + _diag(l, 0, nstar) + _diag(m[:-1], 1, nstar)
    r = np.arange(1, nstar + 1)
    if dim == 2:
        Mx = np.zeros((nstar * 3, nstar * 3))
        rows = nstar + r * 2 - 1
        Mx[np.ix_(rows - 1, r - 1)] = block
        Mx[np.ix_(r - 1, rows - 1)] = np.eye(nstar)
        ii = np.arange(1, nstar * 2 + 1)
        perm = np.concatenate([np.arange(1, nstar + 1),
                               nstar + ii + (np.ceil(ii / 2) * 2 != ii) * 2 - 1])
        My = Mx[np.ix_(perm - 1, perm - 1)]
        Mz = My
        My = Mx
        Mx = np.zeros(Mz.shape)
        ev = np.arange(0, n_mom + 1, 2)
        moment_order = np.concatenate([ev, np.repeat(ev + 1, 2)]).astype(np.int64)
        return moment_order, Mx, My, Mz
    if dim == 3:
        Mx = np.zeros((nstar * 4, nstar * 4))
        rows = nstar + r * 3 - 2
        Mx[np.ix_(rows - 1, r - 1)] = block
        Mx[np.ix_(r - 1, rows - 1)] = np.eye(nstar)
        ii = np.arange(1, nstar * 3 + 1)
        perm = np.concatenate([np.arange(1, nstar + 1), nstar + ii + np.mod(ii + 1, 3) - 1])
        My = Mx[np.ix_(perm - 1, perm - 1)]
        perm = np.concatenate([np.arange(1, nstar + 1), nstar + ii + 2 * (1 - np.mod(ii + 2, 3))])
        Mz = Mx[np.ix_(perm - 1, perm - 1)]
        ev = np.arange(0, n_mom + 1, 2)
        moment_order = np.concatenate([ev, np.repeat(ev + 1, 3)]).astype(np.int64)
        return moment_order, Mx, My, Mz
    raise ValueError("dim must be 1, 2 or 3")
