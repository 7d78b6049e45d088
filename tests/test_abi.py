"""The C ABI library loads and exports every entry point declared in
include/starmap_hip.h; host-side logic that needs no GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from starmap_amd import _lib, closure as C, init as I
from oracle import closure as O
from oracle import starmap_oracle as so

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "starmap_hip.h")).read()
    return sorted(set(re.findall(r"\b(smp_[a-z_0-9]+)\s*\(", src)))


def test_header_and_binding_agree():
    assert sorted(_lib.EXPORTS) == _declared()


def test_library_exports_every_symbol():
    L = _lib.lib()
    for name in _declared():
        assert hasattr(L, name), name
    assert b"gfx950" in L.smp_version()


def test_create_without_gpu_reports_an_error():
    """Without a visible device smp_create fails with a status, never a
    crash; invalid descriptors are rejected before any device call."""
    L = _lib.lib()
    h = ctypes.c_void_p()
    d = _lib.Desc()
    d.n_sys = 0
    rc = L.smp_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == _lib.SMP_E_INVALID
    assert L.smp_last_error()
    assert L.smp_create(None, ctypes.byref(h)) == _lib.SMP_E_INVALID
    assert L.smp_step(None, 0.1, 1) == _lib.SMP_E_INVALID


@pytest.mark.parametrize("clo", ["closure_pn", "closure_spn"])
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_product_closure_matches_oracle_bitwise(clo, dim):
    for N in range(1, 16):
        a = getattr(O, clo)(N, dim)
        b = getattr(C, clo)(N, dim)
        for x, y in zip(a, b):
            assert np.array_equal(np.asarray(x), np.asarray(y))


def test_init_lifting_matches_oracle():
    from starmap_amd import problems as P
    for prob in (P.gauss_periodic_2d(n=10), P.lattice_2d(n=14, n_mom=5),
                 P.gauss_periodic_3d(n=6, n_mom=3), P.boxes_2d(n=10, n_mom=3),
                 P.gauss_1d(n=20), P.planesource_1d(n=20)):
        a = I.starmap_init(prob)
        b = so.starmap_init(prob)
        for k in ("dim", "ax", "n", "bc", "source_ind", "tfinal", "cn", "closure_mod"):
            assert list(np.atleast_1d(a[k])) == list(np.atleast_1d(b[k])), k
        for k in ("ic", "sigma_a", "sigma_s0", "source"):
            assert I.nargin(a[k]) == so.nargin(b[k]), k


def test_init_errors_match_reference_messages():
    with pytest.raises(ValueError, match="Incompatible problem variables ax and n"):
        I.starmap_init({"closure": "P", "n_mom": 1, "ax": [0, 1, 0, 1], "n": [4]})
    with pytest.raises(ValueError, match="At least one of ax, n, or bc"):
        I.starmap_init({"closure": "P", "n_mom": 1})
    with pytest.raises(ValueError, match="sigma_a has incompatible number of arguments"):
        I.starmap_init({"closure": "P", "n_mom": 1, "ax": [0, 1, 0, 1],
                        "sigma_a": lambda x: 0.0})


@pytest.mark.parametrize("clo,N,dim", [("P", 5, 2), ("P", 39, 2), ("P", 7, 3), ("SP", 9, 2),
                                       ("SP", 7, 3), ("P", 4, 1)])
def test_host_only_handle_grid_types_and_schedule(clo, N, dim, monkeypatch):
    """A handle created with device = -1 validates the coupling, computes the
    grid types (equal to the reference's class closure, restated by the
    oracle) and builds the substep schedules without a GPU."""
    from starmap_amd.solver import HipSolver
    ax = [0.0, 1.0] * dim
    prob = {"closure": clo, "n_mom": N, "ax": ax, "n": [12] * dim, "bc": [1] * dim}
    par = I.starmap_init(prob)
    # one workgroup per CU: the whole LDS holds the live set, every source is
    # staged once per tile
    monkeypatch.setenv("SMP_BLOCKS_PER_CU", "1")
    s = HipSolver(par, device=-1)
    gt_ref, _ = so.grid_classes(par["Mx"], par["My"], par["Mz"])
    assert np.array_equal(s.gt, gt_ref)
    for kind in (0, 1):
        groups, slots, loads, sources = s.plan_stats(kind)
        assert groups >= 1 and slots >= 1
        assert loads == sources
    s.close()
    # two workgroups per CU: re-staging stays bounded
    monkeypatch.setenv("SMP_BLOCKS_PER_CU", "2")
    s = HipSolver(par, device=-1)
    for kind in (0, 1):
        groups, slots, loads, sources = s.plan_stats(kind)
        assert sources <= loads <= 2 * sources
    from starmap_amd._lib import SmpError
    with pytest.raises(SmpError, match="without a device"):
        s.step(0.1, 1)
    s.close()


def test_invalid_coupling_is_rejected():
    from starmap_amd.solver import HipSolver
    from starmap_amd._lib import SmpError
    par = I.starmap_init({"closure": "P", "n_mom": 3, "ax": [0, 1, 0, 1], "n": [8, 8]})
    bad = dict(par)
    My = par["My"].copy()
    My[0, 0] = 1.0             # a self-coupling breaks the two-colouring
    bad["My"] = My
    with pytest.raises(SmpError, match="not bipartite"):
        HipSolver(bad, device=-1)


def test_batch_and_schedule_host_checks():
    """Batched stepping rejects host-only handles and bad groups before any
    device call; re-making a schedule works on a host-only handle, and a
    tile the coupling or LDS cannot take leaves the previous one in force."""
    from starmap_amd.solver import HipSolver
    from starmap_amd._lib import SmpError
    L = _lib.lib()
    par = I.starmap_init({"closure": "P", "n_mom": 5, "ax": [0, 1, 0, 1], "n": [20, 20],
                          "bc": [0, 0]})
    s = HipSolver(par, device=-1)
    try:
        arr = (ctypes.c_void_p * 1)(s.handle.value)
        b = ctypes.c_void_p()
        assert L.smp_batch_create(arr, 1, ctypes.byref(b)) == _lib.SMP_E_STATE
        assert L.smp_batch_create(arr, 0, ctypes.byref(b)) == _lib.SMP_E_INVALID
        assert L.smp_batch_create(None, 1, ctypes.byref(b)) == _lib.SMP_E_INVALID
        assert s.get_schedule() == ((63, 4, 1), 1, 8)
        before = [s.plan_stats(k) for k in (0, 1)]
        s.set_schedule((127, 2), 2, 8)
        assert s.get_schedule() == ((127, 2, 1), 2, 8)
        with pytest.raises(SmpError, match="power of two"):
            s.set_schedule((30, 4), 1)
        with pytest.raises(SmpError, match="more than 256 cells"):
            s.set_schedule((63, 8), 1)
        assert s.get_schedule() == ((127, 2, 1), 2, 8)
        with pytest.raises(SmpError, match="cell_waves"):
            s.set_schedule(None, 1, 6)
        s.set_schedule(None, 1)
        assert [s.plan_stats(k) for k in (0, 1)] == before
        with pytest.raises(SmpError, match="without a device"):
            s.checkpoint_save(0)
    finally:
        s.close()


def test_batch_step_counts_follow_the_reference_loop():
    from starmap_amd.batch import count_steps
    # t = 0; while t < tfinal: t = t + dt
    for dt, tf in [(1 / 96, 1.0), (1 / 64, 1.0), (0.1, 0.3), (0.0035, 3.2)]:
        t, n = 0.0, 0
        while t < tf:
            t += dt
            n += 1
        assert count_steps(dt, tf) == n
    assert count_steps(1 / 64, 1.0) == 64


def test_batch_sweep_draws_are_seeded():
    from starmap_amd import problems as P
    a, b = P.batch_draws(64), P.batch_draws(64)
    assert [x["side"] for x in a] == [x["side"] for x in b]
    assert {x["side"] for x in a} <= set(P.BATCH_SIDES)
    assert {x["n_mom"] for x in a} <= set(P.BATCH_ORDERS)
    assert all(0 <= x["f_eff"] <= 20 for x in a)
    assert all(0.01 <= x["sigma_a"].min() and x["sigma_a"].max() <= 10 for x in a)


def test_adjoint_entry_points_need_a_device():
    from starmap_amd.adjoint import Adjoint
    from starmap_amd import problems as P
    from starmap_amd._lib import SmpError
    s = Adjoint(I.starmap_init(P.adjoint_grf(n=8, n_mom=3, steps=4)), device=-1)
    try:
        L = _lib.lib()
        assert L.smp_adjoint(s.handle, 0.1, 4, 2) == _lib.SMP_E_STATE
        assert L.smp_adjoint(s.handle, -0.1, 4, 2) == _lib.SMP_E_INVALID
        assert L.smp_adjoint(s.handle, 0.1, 4, 0) == _lib.SMP_E_INVALID
        with pytest.raises(SmpError, match="without a device"):
            s.reset()
    finally:
        s.close()


def test_probe_noalloc_rejects_more_components_than_its_array():
    """smp_probe_noalloc's fixed array holds SMP_PROBE_NOALLOC_MAX_COMPS = 4096
    components: a 2D P90 problem (4186 components) is refused with
    SMP_E_INVALID before any device call."""
    from starmap_amd import problems as P
    from starmap_amd.solver import HipSolver
    s = HipSolver(I.starmap_init(P.gauss_periodic_2d(n=4, n_mom=90)), device=-1)
    try:
        assert s.n_sys == 4186
        lo, pt, out = np.zeros(3), np.zeros(3), np.zeros(s.n_sys)
        rc = _lib.lib().smp_probe_noalloc(s.handle, lo.ctypes.data, pt.ctypes.data, out.ctypes.data)
        assert rc == _lib.SMP_E_INVALID
        assert b"at most 4096 components" in _lib.lib().smp_last_error()
    finally:
        s.close()


def test_matlab_gateway_matches_header_and_wrapper():
    """The MEX gateway (not compiled here: no MATLAB in this image) calls only
    entry points the header declares, and the M-side wrapper sends only
    commands the gateway implements."""
    mdir = os.path.join(ROOT, "starmap_amd", "matlab")
    c_src = open(os.path.join(mdir, "starmap_hip_mex.c")).read()
    m_src = open(os.path.join(mdir, "starmap_solver_hip.m")).read()
    called = set(re.findall(r"\b(smp_[a-z_0-9]+)\s*\(", c_src))
    assert called and called <= set(_declared()), called - set(_declared())
    commands = set(re.findall(r'strcmp\(cmd, "([a-z_0-9]+)"\)', c_src))
    sent = set(re.findall(r"starmap_hip_mex\('([a-z_0-9]+)'", m_src))
    assert sent and sent <= commands, sent - commands


def test_pair_tiles_need_fp32():
    """Tiles of more than 256 cells (two cells per thread) are accepted for
    the FP32 state and refused for FP64."""
    from starmap_amd import problems as P
    from starmap_amd.solver import HipSolver
    par = I.starmap_init(P.lattice_2d(n=40, n_mom=5))
    s = HipSolver(par, device=-1, precision="fp32", cell_waves=4)
    try:
        s.set_schedule((63, 8), 2, 4)
        assert s.get_schedule() == ((63, 8, 1), 2, 4)
        s.set_schedule((127, 4), 1, 8)
        assert s.get_schedule() == ((127, 4, 1), 1, 8)
        with pytest.raises(_lib.SmpError, match="at most 256 cells"):
            s.set_schedule((63, 9), 2, 4)
    finally:
        s.close()
    s = HipSolver(par, device=-1, precision="fp64", cell_waves=4)
    try:
        with pytest.raises(_lib.SmpError, match="FP32"):
            s.set_schedule((63, 8), 2, 4)
    finally:
        s.close()
