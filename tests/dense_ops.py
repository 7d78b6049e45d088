"""Independent linear-algebra models of one StaRMAP time step, used to pin the
oracle stepper on small problems.

* ``dense_step_operator`` assembles the one-step map on a tiny grid as an
  explicit matrix from Kronecker products of 1D difference matrices and the
  coupling matrices (no index-extension vectors, no sumcell loop).
* ``fourier_step_matrix`` builds the n_sys x n_sys amplification matrix of a
  single Fourier mode on a periodic grid, from the phase-centred difference
  symbol 2i sin(xi h / 2) / h.

Both follow the scheme of starmap_solver.m:307-421 as mathematics: odd
half-step, two even half-steps with the same flux, odd half-step, each
half-step u <- u + dt/2 * (W + Q - s u) * E with E = expm1div(-s dt/2),
substep filter slots k=1 (first odd, first even), k=2 (second even, last odd).
"""
import numpy as np


def two_colour(Mx, My, Mz):
    """Grid type per component and axis by breadth-first search: crossing a
    coupling along axis d flips the grid type along d."""
    n = Mz.shape[0]
    gt = np.zeros((n, 3), dtype=int)
    gt[0] = 1
    seen = {0}
    queue = [0]
    while queue:
        j = queue.pop(0)
        for d, M in enumerate((Mx, My, Mz)):
            for k in np.nonzero(M[j])[0]:
                if k not in seen:
                    g = gt[j].copy()
                    g[d] = 3 - g[d]
                    gt[k] = g
                    seen.add(k)
                    queue.append(k)
    return gt


def _diff_1d(n, bc, src_g, h):
    """1D difference matrix from a source grid to the opposite grid.
    Cell centres: n points.  Faces: n points (periodic, face i at x_lo+(i+1)h)
    or n+1 points (extrapolation, face i at x_lo+i h)."""
    if src_g == 1:     # centres -> faces
        if bc == 0:
            A = np.zeros((n, n))
            for f in range(n):
                A[f, (f + 1) % n] += 1
                A[f, f] -= 1
        else:
            A = np.zeros((n + 1, n))
            for f in range(1, n):
                A[f, f] += 1
                A[f, f - 1] -= 1
    else:              # faces -> centres
        if bc == 0:
            A = np.zeros((n, n))
            for c in range(n):
                A[c, c] += 1
                A[c, (c - 1) % n] -= 1
        else:
            A = np.zeros((n, n + 1))
            for c in range(n):
                A[c, c + 1] += 1
                A[c, c] -= 1
    return A / h


def extents(gt, n, bc):
    return np.array([[n[d] if gt[j, d] == 1 else n[d] + bc[d] for d in range(3)]
                     for j in range(gt.shape[0])])


def dense_step_operator(Mx, My, Mz, n, bc, h, dt, s, E, F1=None, F2=None):
    """Return (A, offsets) such that the state vector (components
    concatenated, each raveled column-major) after one step is A @ u0 when
    there is no source.  s, E: per-component scalars.  F1, F2: per-component
    substep filter scalars (None: no filter)."""
    gt = two_colour(Mx, My, Mz)
    ext = extents(gt, n, bc)
    size = ext.prod(axis=1)
    off = np.concatenate([[0], np.cumsum(size)])
    N = off[-1]
    nsys = Mz.shape[0]
    even = [j for j in range(nsys) if (gt[j] - 1).sum() % 2 == 0]
    odd = [j for j in range(nsys) if (gt[j] - 1).sum() % 2 == 1]

    def deriv(j, k, d):
        mats = []
        for a in range(3):
            if a == d:
                mats.append(_diff_1d(n[a], bc[a], gt[k, a], h[a]))
            else:
                mats.append(np.eye(ext[k, a]))
        # column-major ravel: index = i1 + e1*(i2 + e2*i3)
        return np.kron(mats[2], np.kron(mats[1], mats[0]))

    def flux_rows(targets):
        W = np.zeros((N, N))
        for j in targets:
            for d, M in enumerate((Mx, My, Mz)):
                for k in np.nonzero(M[j])[0]:
                    W[off[j]:off[j + 1], off[k]:off[k + 1]] -= M[j, k] * deriv(j, k, d)
        return W

    def half(targets, W, slot):
        A = np.eye(N)
        for j in targets:
            sl = slice(off[j], off[j + 1])
            blk = np.eye(N)[sl] + dt / 2 * E[j] * (W[sl] - s[j] * np.eye(N)[sl])
            if F1 is not None and j != 0:
                blk = blk * (F1[j] if slot == 1 else F2[j])
            A[sl] = blk
        return A

    Wo = flux_rows(odd)
    We = flux_rows(even)
    A1 = half(odd, Wo, 1)
    Ae1 = half(even, We, 1)
    Ae2 = half(even, We, 2)
    A3 = half(odd, Wo, 2)
    return A3 @ Ae2 @ Ae1 @ A1, off, gt


def fourier_step_matrix(Mx, My, Mz, xi, h, dt, s, E, F1=None, F2=None):
    """Amplification matrix of one step for the mode exp(i xi.x), where each
    component is sampled at its own staggered positions."""
    gt = two_colour(Mx, My, Mz)
    nsys = Mz.shape[0]
    sym = [2j * np.sin(xi[d] * h[d] / 2) / h[d] for d in range(3)]
    L = -(sym[0] * Mx + sym[1] * My + sym[2] * Mz)
    odd = np.array([(gt[j] - 1).sum() % 2 == 1 for j in range(nsys)])

    def half(rows, slot):
        A = np.eye(nsys, dtype=complex)
        for j in np.nonzero(rows)[0]:
            r = np.zeros(nsys, dtype=complex)
            r[j] = 1.0
            r = r + dt / 2 * E[j] * (L[j] - s[j] * np.eye(nsys)[j])
            if F1 is not None and j != 0:
                r = r * (F1[j] if slot == 1 else F2[j])
            A[j] = r
        return A

    return half(odd, 2) @ half(~odd, 2) @ half(~odd, 1) @ half(odd, 1), gt


# ---------------------------------------------------------------------------
# Affine one-step model with per-point coefficients
# ---------------------------------------------------------------------------
def flux_matrix(Mx, My, Mz, n, bc, h):
    """The flux operator W = -(sum_d M_d (x) D_d) as one dense matrix over
    the concatenated state (each component raveled column-major on its own
    staggered grid), plus the component offsets and grid types."""
    gt = two_colour(Mx, My, Mz)
    ext = extents(gt, n, bc)
    size = ext.prod(axis=1)
    off = np.concatenate([[0], np.cumsum(size)])
    N = off[-1]
    W = np.zeros((N, N))
    for j in range(Mz.shape[0]):
        for d, M in enumerate((Mx, My, Mz)):
            for k in np.nonzero(M[j])[0]:
                mats = []
                for a in range(3):
                    if a == d:
                        mats.append(_diff_1d(n[a], bc[a], gt[k, a], h[a]))
                    else:
                        mats.append(np.eye(ext[k, a]))
                D = np.kron(mats[2], np.kron(mats[1], mats[0]))
                W[off[j]:off[j + 1], off[k]:off[k + 1]] -= M[j, k] * D
    return W, off, gt


def staggered_points(gt_j, n, bc, lo, h):
    """Coordinates (X, Y, Z), raveled column-major, of the points of a
    component with grid types gt_j: cell centres lo+(i+1/2)h; faces lo+(i+1)h
    (periodic, n of them) or lo+ih (extrapolation, n+1 of them)."""
    axes = []
    for a in range(3):
        if gt_j[a] == 1:
            axes.append(lo[a] + (np.arange(n[a]) + 0.5) * h[a])
        elif bc[a] == 0:
            axes.append(lo[a] + (np.arange(n[a]) + 1.0) * h[a])
        else:
            axes.append(lo[a] + np.arange(n[a] + 1) * h[a])
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    return X.ravel(order="F"), Y.ravel(order="F"), Z.ravel(order="F")


def affine_step(W, off, odd, even, dt, s, E, Q=None, F1=None, F2=None, Ffull=None):
    """One time step as the affine map u -> A u + b.

    W: flux matrix (flux_matrix); s, E: per component, the decay rate and
    expm1div(-s dt/2) at each of its points (arrays); Q: per component the
    source at each point, or None; F1, F2: substep filter factors per
    component and point (slot 1: first odd and first even half-step; slot 2:
    second even and last odd), or None; Ffull: filter factors applied to
    every component but the zeroth after the step, or None.

    Each half-step on its targets j is u_j <- f_j (u_j + dt/2 E_j (W_j u + Q_j
    - s_j u_j)), the exact solution of u' = W u + Q - s u over dt/2 with W u
    and Q frozen, times the filter factor f_j (1 for j = 0)."""
    N = W.shape[0]
    I = np.eye(N)

    def half(targets, slot):
        A = np.eye(N)
        b = np.zeros(N)
        for j in targets:
            sl = slice(off[j], off[j + 1])
            e = np.broadcast_to(np.asarray(E[j], dtype=float), (sl.stop - sl.start,))
            sig = np.broadcast_to(np.asarray(s[j], dtype=float), e.shape)
            rows = I[sl] + (dt / 2) * e[:, None] * (W[sl] - sig[:, None] * I[sl])
            bj = (dt / 2) * e * (np.asarray(Q[j], dtype=float) if Q is not None and Q[j] is not None
                                 else 0.0)
            if F1 is not None and j != 0:
                f = np.broadcast_to(np.asarray(F1[j] if slot == 1 else F2[j], dtype=float), e.shape)
                rows = f[:, None] * rows
                bj = f * bj
            A[sl] = rows
            b[sl] = bj
        return A, b

    def then(second, first):
        (A2, b2), (A1, b1) = second, first
        return A2 @ A1, A2 @ b1 + b2

    step = half(odd, 1)
    step = then(half(even, 1), step)
    step = then(half(even, 2), step)
    step = then(half(odd, 2), step)
    if Ffull is not None:
        d = np.ones(N)
        for j in range(1, len(off) - 1):
            d[off[j]:off[j + 1]] = Ffull[j]
        step = (d[:, None] * step[0], d * step[1])
    return step
