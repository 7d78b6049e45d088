"""GPU tests of the discrete adjoint (through the C ABI).

* dJ/dU(0): J is affine in the initial state for fixed materials, so the
  adjoint state at every probed point -- interior, edges and corners, both
  parities, extrapolation and periodic boundaries -- equals a central
  difference of the FP64 forward to rounding (1e-9 relative).  This pins
  the transposed flux stencil and its boundary rules.
* dJ/dsigma_a, dJ/dsigma_s0 per cell against FP64 central differences of
  the forward (relative step 1e-6 of the cell value), 1e-5 relative to the
  largest probed gradient; with and without the substep filter, with
  checkpoints every few steps.
* The same gradients against central differences of the oracle (the CPU
  restatement of the reference step) on a small grid.
* FP32 state with FP64 accumulation against the FP64 adjoint (1e-4).
* Dot product: <lambda(0), dU0> = J(dU0) - J(0) for a random dU0 in every
  component (2D P15 128^2, 100 steps in 7 checkpoint segments, 1e-11
  relative to the sum of |lambda*dU0|).
* The torch.autograd.Function returns the adjoint gradients.
"""
import numpy as np
import pytest

from oracle import starmap_oracle as so
from starmap_amd import init as I
from starmap_amd import problems as P
from starmap_amd.adjoint import Adjoint, objective_value, strip_objective
from starmap_amd.batch import count_steps
from starmap_amd.solver import HipSolver

pytestmark = pytest.mark.gpu


def small(n=12, n_mom=3, steps=9, bc=(1, 1), filt=False, seed=5, xmin=0.5):
    rng = np.random.default_rng(seed)
    sa = 0.5 * np.exp(rng.normal(0, 0.5, (n, n)))
    ss = 1.0 + 0.3 * rng.random((n, n))
    prob = P.adjoint_grf(n=n, n_mom=n_mom, steps=steps, sigma_a=sa, sigma_s0=ss, src_std=0.2)
    prob["bc"] = list(bc)
    if filt:
        prob["filterfunction"] = P.exp_filter(4, 5.0)
        prob["f_position"] = "substep"
    return prob, sa, ss


def forward_J(prob, mask, precision="fp64", state=None):
    s = Adjoint(I.starmap_init(prob), precision=precision)
    try:
        dt = s.prepare()
        if state:
            for j, arr in state.items():
                s.set_state(j, arr)
        from starmap_amd.batch import count_steps
        s.step(dt, count_steps(dt, s.par["tfinal"]))
        return objective_value(s, mask)
    finally:
        s.close()


def adjoint_of(prob, precision="fp64", every=4):
    s = Adjoint(I.starmap_init(prob), precision=precision)
    mask = strip_objective(s, xmin=0.2)
    s.run({0: mask}, every=every)
    return s, mask


@pytest.mark.parametrize("bc", [(1, 1), (0, 0)])
def test_lambda0_equals_finite_difference_in_initial_state(bc):
    prob, _, _ = small(bc=bc, filt=True)
    s, mask = adjoint_of(prob)
    try:
        rng = np.random.default_rng(1)
        for j in (0, 1, 2, 5):
            lam = s.lam(j)
            shp = lam.shape
            pts = [(0, 0, 0), (0, shp[1] - 1, shp[2] - 1), (0, 0, shp[2] - 1), (0, shp[1] - 1, 0)]
            pts += [(0, int(rng.integers(shp[1])), int(rng.integers(shp[2]))) for _ in range(3)]
            for p in pts:
                e = np.zeros(shp)
                e[p] = 1.0
                Jp = forward_J(prob, mask, state={j: e})
                Jm = forward_J(prob, mask, state={j: -e})
                fd = (Jp - Jm) / 2
                print("lambda", bc, j, p, lam[p], fd, flush=True)
                assert abs(lam[p] - fd) <= 1e-9 * max(abs(fd), np.abs(lam).max()), (j, p, lam[p], fd)
    finally:
        s.close()


def fd_cells(prob, sa, ss, cells, mask, which, rel=1e-6, oracle=False):
    out = []
    for c in cells:
        vals = []
        for sgn in (1, -1):
            a, b = sa.copy(), ss.copy()
            arr = a if which == "a" else b
            eps = rel * arr.flat[c]
            arr.flat[c] += sgn * eps
            p2 = P.adjoint_grf(n=sa.shape[0], n_mom=prob["n_mom"], steps=None or 0, sigma_a=a,
                               sigma_s0=b, src_std=0.2)
            p2.update({k: prob[k] for k in ("bc", "tfinal") if k in prob})
            for k in ("filterfunction", "f_position"):
                if k in prob:
                    p2[k] = prob[k]
            if oracle:
                ref, _ = so.starmap_solver(so.starmap_init(p2), verbose=False, return_state=True)
                vals.append(float(np.sum(np.reshape(ref[0]["U"], mask.shape, order="F") * mask)))
            else:
                vals.append(forward_J(p2, mask))
        out.append((vals[0] - vals[1]) / (2 * rel * (sa if which == "a" else ss).flat[c]))
    return np.array(out)


@pytest.mark.parametrize("filt", [False, True])
def test_material_gradients_match_finite_differences(filt):
    prob, sa, ss = small(filt=filt)
    s, mask = adjoint_of(prob, every=2)
    try:
        n = sa.shape[0]
        owner = dict(prob, sigma_a=(lambda x, y: prob["_owner"](x, y).astype(np.float64)))
        ga, gs = s.cell_gradients(owner, n * n)
    finally:
        s.close()
    cells = [0, n - 1, n * n - 1, 5 * n + 7, 7 * n + 3, 9 * n + 10]
    for which, g in (("a", ga), ("s", gs)):
        fd = fd_cells(prob, sa, ss, cells, mask, which)
        print("dJ/dsigma_" + which, g[cells], fd, flush=True)
        scale = np.abs(fd).max()
        assert np.all(np.abs(g[cells] - fd) <= 1e-5 * scale), (which, g[cells], fd)


def test_material_gradient_matches_oracle_differences():
    prob, sa, ss = small(n=10, steps=6)
    s, mask = adjoint_of(prob, every=3)
    try:
        owner = dict(prob, sigma_a=(lambda x, y: prob["_owner"](x, y).astype(np.float64)))
        ga, gs = s.cell_gradients(owner, 100)
    finally:
        s.close()
    cells = [44, 57, 91]
    for which, g in (("a", ga), ("s", gs)):
        fd = fd_cells(prob, sa, ss, cells, mask, which, oracle=True)
        print("oracle dJ/dsigma_" + which, g[cells], fd, flush=True)
        assert np.all(np.abs(g[cells] - fd) <= 1e-5 * np.abs(fd).max())


def test_fp32_adjoint_close_to_fp64():
    prob, sa, ss = small(n=24, n_mom=5, steps=20)
    res = {}
    for prec in ("fp64", "fp32"):
        s, mask = adjoint_of(prob, precision=prec, every=5)
        try:
            owner = dict(prob, sigma_a=(lambda x, y: prob["_owner"](x, y).astype(np.float64)))
            res[prec] = s.cell_gradients(owner, 24 * 24)
        finally:
            s.close()
    for a64, a32 in zip(res["fp64"], res["fp32"]):
        assert np.linalg.norm(a32 - a64) <= 1e-4 * np.linalg.norm(a64)


def test_torch_autograd_function():
    torch = pytest.importorskip("torch")
    from starmap_amd.adjoint import StarmapObjective
    n = 12
    prob0, sa, ss = small(n=n)

    def make_prob(a, b):
        p = P.adjoint_grf(n=n, n_mom=3, steps=9, sigma_a=a, sigma_s0=b, src_std=0.2)
        return p

    ta = torch.tensor(sa, requires_grad=True)
    ts = torch.tensor(ss, requires_grad=True)
    J = StarmapObjective.apply(ta, ts, make_prob, "fp64", 4)
    J.backward()
    s = Adjoint(I.starmap_init(make_prob(sa, ss)), precision="fp64")
    try:
        mask = strip_objective(s, xmin=0.8)
        s.run({0: mask}, every=4)
        assert abs(objective_value(s, mask) - float(J)) <= 1e-12 * abs(float(J))
        owner = dict(prob0, sigma_a=(lambda x, y: prob0["_owner"](x, y).astype(np.float64)))
        ga, gs = s.cell_gradients(owner, n * n)
    finally:
        s.close()
    assert np.allclose(ta.grad.numpy().ravel(), ga, rtol=1e-12, atol=0)
    assert np.allclose(ts.grad.numpy().ravel(), gs, rtol=1e-12, atol=0)


def small_3d(n=(6, 5, 4), n_mom=3, steps=5, bc=(1, 0, 1)):
    """3D P_N with mixed boundaries, spatially varying sigma_a, sigma_s0 and
    source (fields on every staggered grid)."""
    lam = P.legendre_max_root(n_mom + 1)
    h = 2.0 / max(n)
    return {
        "name": "adjoint 3D", "closure": "P", "n_mom": n_mom,
        "sigma_a": (lambda x, y, z: 0.4 + 0.2 * np.sin(3 * x + 1) * np.cos(2 * y) + 0.1 * z),
        "sigma_s0": (lambda x, y, z: 1.0 + 0.3 * np.cos(x * y + z)),
        "source": (lambda x, y, z: np.exp(-((x + 0.3) ** 2 + y ** 2 + (z - 0.2) ** 2) / 0.2)),
        "ax": [-1.0, 1.0, -1.0, 1.0, -1.0, 1.0], "n": list(n), "bc": list(bc),
        "cn": 1.5 * lam, "t_plot": [], "tfinal": steps * 0.5 * h, "mom_output": 1, "output": None,
    }


def test_lambda0_equals_finite_difference_3d():
    """The transposed stencil on three axes, periodic and extrapolation
    boundaries mixed: lambda(0) against central differences of the FP64
    forward at corners and random points of components of both parities."""
    prob = small_3d()
    s = Adjoint(I.starmap_init(prob), precision="fp64")
    try:
        rng = np.random.default_rng(3)
        s.prepare()
        mask = rng.random(s.shape(0))
        s.run({0: mask}, every=2)
        for j in (0, 1, 2, 3, 7):
            lam = s.lam(j)
            shp = lam.shape
            corners = [tuple(c * (d - 1) for c, d in zip(cs, shp))
                       for cs in ((0, 0, 0), (1, 1, 1), (1, 0, 1), (0, 1, 0))]
            pts = corners + [tuple(int(rng.integers(d)) for d in shp) for _ in range(3)]
            for p in pts:
                e = np.zeros(shp)
                e[p] = 1.0
                Jp = forward_J(prob, mask, state={j: e})
                Jm = forward_J(prob, mask, state={j: -e})
                fd = (Jp - Jm) / 2
                print("lambda3d", j, p, lam[p], fd, flush=True)
                assert abs(lam[p] - fd) <= 1e-9 * max(abs(fd), np.abs(lam).max()), (j, p, lam[p], fd)
    finally:
        s.close()


def test_dot_product_every_component():
    """<lambda(0), dU0> against J(dU0) - J(0) for a random dU0 in every
    component and point: J = <w, U_1(T)> is affine in the initial state, so
    the two agree to rounding.  One check of the whole transposed step (every
    coupling, both parities, the boundaries, the checkpoint segments) on
    2D P15, 128^2, 100 steps in 7 segments."""
    prob = P.adjoint_grf(n=128, n_mom=15, steps=100)
    rng = np.random.default_rng(7)
    s = Adjoint(I.starmap_init(prob), precision="fp64")
    try:
        s.prepare()
        w = rng.random(s.shape(0))
        s.run({0: w}, every=16)
        lam = [s.lam(j) for j in range(s.n_sys)]
    finally:
        s.close()
    d = {j: rng.standard_normal(np.shape(lam[j])) for j in range(len(lam))}
    dJ = forward_J(prob, w, state=d) - forward_J(prob, w)
    ip = sum(float(np.sum(lam[j] * d[j])) for j in d)
    scale = sum(float(np.sum(np.abs(lam[j] * d[j]))) for j in d)
    print("dot product", dJ, ip, scale, flush=True)
    assert abs(dJ - ip) <= 1e-11 * scale


def test_torch_loss_of_final_moment_drives_the_adjoint():
    """StarmapFinalMoment: U1(T) as a torch tensor; a nonlinear torch loss of
    it, L = sum(w * U1^2) + sum(U1), differentiated by autograd (backward
    seeds the adjoint with dL/dU1(T) = 2 w U1 + 1) against Richardson-
    extrapolated FP64 central differences of L through the forward at seeded
    cells."""
    torch = pytest.importorskip("torch")
    from starmap_amd.adjoint import StarmapFinalMoment
    n = 12
    prob0, sa, ss = small(n=n)

    def make_prob(a, b):
        return P.adjoint_grf(n=n, n_mom=3, steps=9, sigma_a=a, sigma_s0=b, src_std=0.2)

    w = np.random.default_rng(5).uniform(0.5, 1.5, size=(n, n))

    def loss_np(a, b):
        s = HipSolver(I.starmap_init(make_prob(a, b)), precision="fp64")
        try:
            dt = s.prepare()
            s.step(dt, count_steps(dt, s.par["tfinal"]))
            U1 = s.get_state(0).reshape(n, n, order="F")
        finally:
            s.close()
        return float(np.sum(w * U1 ** 2) + np.sum(U1))

    ta = torch.tensor(sa, requires_grad=True)
    ts = torch.tensor(ss, requires_grad=True)
    U1 = StarmapFinalMoment.apply(ta, ts, make_prob, "fp64", 4)
    L = (torch.tensor(w) * U1 ** 2).sum() + U1.sum()
    assert abs(float(L) - loss_np(sa, ss)) <= 1e-12 * abs(float(L))
    L.backward()
    for c in np.random.default_rng(9).choice(n * n, size=4, replace=False):
        for grad, which in ((ta.grad, "a"), (ts.grad, "s")):
            base = sa if which == "a" else ss

            def central(eps):
                vals = []
                for sgn in (1, -1):
                    a, b = sa.copy(), ss.copy()
                    (a if which == "a" else b).flat[c] += sgn * eps
                    vals.append(loss_np(a, b))
                return (vals[0] - vals[1]) / (2 * eps)
            # Richardson extrapolation of two central differences: the
            # O(eps^2) truncation error of the quadratic loss cancels; what
            # is left is the rounding of L over the 2*eps difference, about
            # 3e-11 absolute here (the adjoint is exact: cells with
            # gradients of 1e-3 agree to 1e-9)
            e = 1e-4 * base.flat[c]
            fd = (4 * central(e) - central(2 * e)) / 3
            assert abs(float(grad.numpy().flat[c]) - fd) <= 1e-7 * abs(fd) + 1e-10, (which, c)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_schedules_under_memory_budgets_give_identical_results(precision):
    """The same adjoint under four schedules -- a checkpoint per segment (no
    budget), everything recorded in the forward pass, a few segments recorded
    and the rest recomputed, and binomial checkpointing with re-advanced
    segments -- gives the same lambda(0), gradients and final state bit for
    bit: every schedule recomputes each launch from the same inputs."""
    prob = P.adjoint_grf(n=32, n_mom=5, steps=37, src_std=0.2)
    every = 4                              # 10 segments, the last of 1 step
    res = {}
    s = Adjoint(I.starmap_init(prob), precision=precision)
    try:
        mask = strip_objective(s, xmin=0.2)
        s.run({0: mask}, every=every)
        st0 = s.schedule_stats()
        state_bytes = st0["peak_bytes"] // st0["peak_buffers"]
        assert st0["recomputed_segments"] == 10 and st0["recorded_in_forward"] == 0
        # buffers of a segment's records: 5 per record kind (states, and the
        # W+Q the recording launches store)
        rb = (st0["peak_buffers"] - 10) // 2
        assert rb in (5, 10)
        budgets = {"default": None, "all": 10 * rb + 1, "some": 6 * rb + 1 + 4, "binomial": 2 * rb + 4}
        for name, nbuf in budgets.items():
            s.set_memory(0 if nbuf is None else nbuf * state_bytes)
            s.run({0: mask}, every=every)
            st = s.schedule_stats()
            print(name, st, flush=True)
            if nbuf is not None:
                assert st["peak_bytes"] <= nbuf * state_bytes
            res[name] = ([s.lam(j) for j in range(s.n_sys)],
                         [g for pair in s.material_gradients().values() for g in pair],
                         [s.get_state(j) for j in range(s.n_sys)], st)
    finally:
        s.close()
    assert res["all"][3]["recomputed_segments"] == 0
    assert 0 < res["some"][3]["recorded_in_forward"] < 10
    assert res["binomial"][3]["extra_segment_advances"] > 0
    ref = res["default"]
    for name in ("all", "some", "binomial"):
        for a, b in zip(ref[0] + ref[1] + ref[2], res[name][0] + res[name][1] + res[name][2]):
            assert np.array_equal(a, b), name


def _adjoint_results(prob, precision, every, kernel, monkeypatch, seed_fn=None, record_w="1"):
    monkeypatch.setenv("SMP_ADJ_KERNEL", kernel)
    monkeypatch.setenv("SMP_ADJ_RECORD_W", record_w)
    s = Adjoint(I.starmap_init(prob), precision=precision)
    try:
        s.prepare()
        w = (seed_fn or (lambda sh: np.random.default_rng(11).random(sh)))(s.shape(0))
        s.run({0: w}, every=every)
        lam = [s.lam(j) for j in range(s.n_sys)]
        grads = [g for pair in s.material_gradients().values() for g in pair]
    finally:
        s.close()
    return lam, grads


@pytest.mark.parametrize("record_w", ["0", "1"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("case", ["2d_extrapolation_filter", "2d_periodic", "3d_mixed"])
def test_lds_adjoint_update_equals_gather_kernel(case, precision, record_w, monkeypatch):
    """The LDS-staged adjoint update (state and mu tiles in LDS, the
    transposed flux as forward-form differences of mu extended by zero,
    boundary faces masked per coupling axis) against the gather kernels,
    which read every sample from global memory with the boundary masks of
    the transposed stencil: lambda(0) of every component bit for bit, the
    gradients on every grid type to the rounding of their sum over targets
    (the LDS kernel adds two partial sums per grid type and point).  Covers
    extrapolation and periodic axes, the substep filter, sources and
    spatially varying materials, 2D and 3D.  SMP_ADJ_KERNEL=lds makes the
    library fail rather than fall back to the gather kernels.  With
    record_w = 1 both read W+Q from the forward launches' records (the LDS
    kernel then stages no state tiles)."""
    if case == "3d_mixed" and precision == "fp64":
        pytest.skip("3D FP64: two groups' source tiles (480-element FP64 slots) do not fit in the "
                    "LDS; the library runs the gather kernels for it")
    if case == "2d_extrapolation_filter":
        prob, _, _ = small(n=20, n_mom=5, steps=7, filt=True)
    elif case == "2d_periodic":
        prob, _, _ = small(n=18, n_mom=7, steps=6, bc=(0, 0))
    else:
        prob = small_3d(n=(7, 6, 5), n_mom=3, steps=4)
    ref = _adjoint_results(prob, precision, 3, "gather", monkeypatch, record_w=record_w)
    new = _adjoint_results(prob, precision, 3, "lds", monkeypatch, record_w=record_w)
    for a, b in zip(ref[0], new[0]):
        assert np.array_equal(a, b), (case, precision, float(np.abs(a - b).max()))
    for a, b in zip(ref[1], new[1]):
        scale = max(float(np.abs(a).max()), 1e-300)
        assert float(np.abs(a - b).max()) <= 1e-12 * scale, (case, precision)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_recorded_flux_gives_the_same_gradients(precision, monkeypatch):
    """The forward launches that record a segment also store W+Q of their
    targets, and the update reads it instead of gathering the flux again:
    lambda(0) is unchanged bit for bit (it does not depend on W), and the
    gradients agree to the rounding of W (the forward's flux in the state
    precision against the update's FP64 sum of the same differences)."""
    prob, _, _ = small(n=20, n_mom=5, steps=7, filt=True)
    a = _adjoint_results(prob, precision, 3, "gather", monkeypatch, record_w="0")
    b = _adjoint_results(prob, precision, 3, "gather", monkeypatch, record_w="1")
    for x, y in zip(a[0], b[0]):
        assert np.array_equal(x, y)
    tol = 1e-12 if precision == "fp64" else 1e-5
    for x, y in zip(a[1], b[1]):
        assert float(np.abs(x - y).max()) <= tol * float(np.abs(x).max())
