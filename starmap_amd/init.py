"""``par = starmap_init(prob)`` -- the reference's problem-setup API
(starmap_init.m:1) for the host side of the HIP solver.

``prob`` is a dict with the reference's field names (README.txt:82-158):
name, closure ('P'/'SP'), n_mom, ic, sigma_a, sigma_s0, sigma_sm, source,
source_ind, ax, n, bc, t_plot, tfinal, cn, timestep, mom_output,
filterfunction, f_position, output, plus any problem-specific fields.
Problem functions are Python callables taking the reference's arguments; the
number of positional parameters plays the role of MATLAB's ``nargin``.

As in the reference, 1D and 2D problems are lifted to the solver's 3D data
structures (starmap_init.m:75-205): a 2D problem on (x, y) is solved on
solver axes (dummy, y, z) with n = [1 nx ny].
"""
import inspect

import numpy as np

from . import closure as _closure


def nargin(f):
    """MATLAB nargin of a callable: the number of positional parameters, or
    -(fixed+1) for a callable with *args (MATLAB's varargin)."""
    params = list(inspect.signature(f).parameters.values())
    fixed = [p for p in params if p.kind in (p.POSITIONAL_ONLY, p.POSITIONAL_OR_KEYWORD)]
    if any(p.kind == p.VAR_POSITIONAL for p in params):
        return -(len(fixed) + 1)
    return len(fixed)


def zero(*args):
    """Zero function: zeros of the size of the first argument."""
    return np.zeros(np.shape(args[0]))


def capargs(fct, *args):
    """Call fct with as many leading arguments as it takes (at least one)."""
    return fct(*args[:max(nargin(fct), 1)])


# argument lifting per dimension: nargin of the user function -> wrapper with
# the solver's (x, y, z[, ...]) argument list
def _lift(dim, name, f):
    na = nargin(f)
    if dim == 1:
        table = {
            "ic": {1: lambda: (lambda _a, _b, z: f(z)),
                   2: lambda: (lambda _a, _b, z, k: f(z, k))},
            "sigma_a": {1: lambda: (lambda _a, _b, z: f(z)),
                        2: lambda: (lambda _a, _b, z, t: f(z, t))},
            "sigma_s0": {1: lambda: (lambda _a, _b, z: f(z)),
                         2: lambda: (lambda _a, _b, z, t: f(z, t))},
            # the reference's 2-argument wrapper hands (m, x) to f(x, m)
            # (starmap_init.m:109); kept as is
            "sigma_sm": {2: lambda: (lambda _a, _b, m, z: f(z, m)),
                         3: lambda: (lambda _a, _b, z, m, t: f(z, m, t))},
            "source": {1: lambda: (lambda _a, _b, z: f(z)),
                       2: lambda: (lambda _a, _b, z, t: f(z, t)),
                       3: lambda: (lambda _a, _b, z, t, k: f(z, t, k))},
            "filterfunction": {3: lambda: f,
                               4: lambda: (lambda p, dt, k, _a, _b, z: f(p, dt, k, z)),
                               5: lambda: (lambda p, dt, k, _a, _b, z, t: f(p, dt, k, z, t))},
        }
    else:
        table = {
            "ic": {2: lambda: (lambda _a, x, y: f(x, y)),
                   3: lambda: (lambda _a, x, y, k: f(x, y, k))},
            "sigma_a": {2: lambda: (lambda _a, x, y: f(x, y)),
                        3: lambda: (lambda _a, x, y, t: f(x, y, t))},
            "sigma_s0": {2: lambda: (lambda _a, x, y: f(x, y)),
                         3: lambda: (lambda _a, x, y, t: f(x, y, t))},
            "sigma_sm": {3: lambda: (lambda _a, x, y, m: f(x, y, m)),
                         4: lambda: (lambda _a, x, y, m, t: f(x, y, m, t))},
            "source": {2: lambda: (lambda _a, x, y: f(x, y)),
                       3: lambda: (lambda _a, x, y, t: f(x, y, t)),
                       4: lambda: (lambda _a, x, y, t, k: f(x, y, t, k))},
            "filterfunction": {3: lambda: f,
                               5: lambda: (lambda p, dt, k, _a, x, y: f(p, dt, k, x, y)),
                               6: lambda: (lambda p, dt, k, _a, x, y, t: f(p, dt, k, x, y, t))},
        }
    if na not in table[name]:
        raise ValueError("%s has incompatible number of arguments." % name)
    return table[name][na]()


def starmap_init(prob):
    """Configure the solver data structures from a problem description
    (starmap_init.m:30-231)."""
    par = dict(prob)
    par["prob"] = prob
    if "ax" in prob:
        dim = len(prob["ax"]) // 2
    elif "n" in prob:
        dim = len(prob["n"])
    elif "bc" in prob:
        dim = len(prob["bc"])
    else:
        raise ValueError("At least one of ax, n, or bc should be defined")
    par["dim"] = dim
    if "ax" in prob and "n" in prob and len(prob["ax"]) // 2 != len(prob["n"]):
        raise ValueError("Incompatible problem variables ax and n.")
    if "ax" in prob and "bc" in prob and len(prob["ax"]) // 2 != len(prob["bc"]):
        raise ValueError("Incompatible problem variables ax and bc.")
    if "ax" not in prob and "n" in prob and "bc" in prob and len(prob["n"]) != len(prob["bc"]):
        raise ValueError("Incompatible problem variables n and bc.")
    builders = {"P": _closure.closure_pn, "SP": _closure.closure_spn}
    if par.get("closure") not in builders:
        raise ValueError("closure must be 'P' or 'SP'")
    par["mom_order"], par["Mx"], par["My"], par["Mz"] = builders[par["closure"]](par["n_mom"], dim)
    n_sys = par["Mz"].shape[0]
    if dim in (1, 2):
        if "ax" in prob:
            ax = list(prob["ax"])
            par["ax"] = ax * 3 if dim == 1 else ax[0:2] + ax
        if "n" in prob:
            par["n"] = [1, 1] + list(prob["n"]) if dim == 1 else [1] + list(prob["n"])
        if "bc" in prob:
            bc = list(prob["bc"])
            par["bc"] = bc * 3 if dim == 1 else [bc[0]] + bc
        for name in ("ic", "sigma_a", "sigma_s0", "sigma_sm", "source", "filterfunction"):
            if name in prob and prob[name] is not None:
                par[name] = _lift(dim, name, prob[name])
    defaults = {"name": "Output", "ic": zero, "sigma_a": zero, "sigma_s0": zero,
                "source": zero, "ax": [0, 1, 0, 1, 0, 1], "n": [100, 100, 100],
                "timestep": 0, "cn": 0.99, "bc": [0, 0, 0],
                "t_plot": [k / 10 for k in range(11)], "mom_output": 1}
    for k, v in defaults.items():
        par.setdefault(k, v)
    if "source_ind" not in par:
        na = nargin(par["source"])
        # zeroth moment for an isotropic source, all moments for a source
        # that takes the component index (starmap_init.m:218-221)
        cnt = int(na >= 0) + int(na > 4) * (n_sys - 1)
        par["source_ind"] = list(range(1, cnt + 1))
    if "tfinal" not in par:
        par["tfinal"] = par["t_plot"][-1]
    par["closure_mod"] = "F" if "filterfunction" in par else ""
    return par
