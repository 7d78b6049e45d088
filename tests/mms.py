"""The manufactured-solution problem of starmap_2d_create_mms.m, derived here
with sympy as the reference derives it with the Symbolic Math Toolbox.

Test helper: the reference's generator (starmap_2d_create_mms.m:32-76,
:232-251) prescribes
  * u_1 = exp(-t) sin(2 pi x)^2, every other moment zero;
  * sigma_a = t cos(2 pi y), sigma_s0 = 1, sigma_sm(m) = 0.9^m;
  * the source q = du/dt + Mx du/dx + My du/dy + (sigma_a + sigma_s0) u
    - sigma_sm(m) .* u, with Mx, My the closure's matrices of the physical
    x and y directions ([mom_order,~,Mx,My] = closure(n_mom,2), :56-59);
and solves on [0,1]^2 with extrapolation boundaries up to tfinal = 1 on the
grids 2.^(4:7) (:33-48).  Every material function takes t, so the solver
runs its time-dependent branch (material averaging, last step shortened).
"""
import numpy as np
import sympy as sp

from oracle.closure import closure_pn, closure_spn

POINTS = (16, 32, 64, 128)


def _derive(n_mom, closure):
    x, y, t = sp.symbols("x y t", real=True)
    fun = {"P": closure_pn, "SP": closure_spn}[closure]
    mom_order, _, Mx, My = fun(n_mom, 2)
    mom_order = np.atleast_1d(mom_order).astype(int)
    n_sys = Mx.shape[0]
    u = sp.zeros(n_sys, 1)
    u[0] = sp.exp(-t) * sp.sin(2 * sp.pi * x) ** 2
    absorb = sp.cos(2 * sp.pi * y) * t
    g = sp.Rational(9, 10)
    scatter0 = g ** 0
    scatterm = sp.Matrix([g ** int(m) for m in mom_order])
    Mxs = sp.Matrix(np.asarray(Mx, dtype=float).tolist())
    Mys = sp.Matrix(np.asarray(My, dtype=float).tolist())
    q = u.diff(t) + Mxs * u.diff(x) + Mys * u.diff(y) + (absorb + scatter0) * u \
        - scatterm.multiply_elementwise(u)
    q = q.applyfunc(sp.simplify)
    return (x, y, t), u, q, mom_order


def _vectorize(expr, syms):
    f = sp.lambdify(syms, expr, "numpy")

    def g(*a):
        return np.asarray(f(*a), dtype=float) + np.zeros(np.shape(a[0]))
    return g


def mms_problem(n, n_mom=3, closure="P", tfinal=1.0):
    """``prob`` of the generated starmap_2d_ex_mms_auto.m on an n x n grid,
    plus the exact solution ``solution(x, y, k, t)`` (k 1-based)."""
    syms, u, q, _ = _derive(n_mom, closure)
    n_sys = u.shape[0]
    us = [_vectorize(u[k], syms) for k in range(n_sys)]
    qs = [_vectorize(q[k], syms) for k in range(n_sys)]

    def initial(x, y, k):
        return us[k - 1](x, y, 0.0)

    def source(x, y, t, k):
        return qs[k - 1](x, y, t)

    def solution(x, y, k, t):
        return us[k - 1](x, y, t)

    prob = {
        "name": "Manufactured Solution Test",
        "closure": closure,
        "n_mom": n_mom,
        "tfinal": tfinal,
        "sigma_a": lambda x, y, t: np.cos(2 * np.pi * y) * t,
        "sigma_s0": lambda x, y, t: np.ones(np.shape(x)),
        "sigma_sm": lambda x, y, m, t: 0.9 ** m * np.ones(np.shape(x)),
        "source": source,
        "ic": initial,
        "ax": [0.0, 1.0, 0.0, 1.0],
        "n": [n, n],
        "bc": [1, 1],
        "t_plot": [],
        "output": None,
    }
    return prob, solution


def mms_errors(sol, solution, tfinal, mom_order, n):
    """Scaled L1, L2 and max errors per moment order, as the generated
    script accumulates them (starmap_2d_create_mms.m:140-149)."""
    nord = int(np.max(mom_order)) + 1
    e1, e2, einf = np.zeros(nord), np.zeros(nord), np.zeros(nord)
    for j, s in enumerate(sol):
        X, Y = np.meshgrid(s["x"], s["y"], indexing="ij")
        D = np.abs(s["U"] - solution(X, Y, j + 1, tfinal))
        i = int(mom_order[j])
        e1[i] += D.sum() / n ** 2
        e2[i] += (D ** 2).sum() / n ** 2
        einf[i] = max(einf[i], D.max())
    return e1, np.sqrt(e2), einf
