"""The adjoint's checkpoint schedule (smp_adjoint_schedule, host only).

Every schedule is replayed symbolically: each buffer holds "the state at the
start of segment s" (or a record range "the records of segment s"), and
every action checks that what it reads is what it needs.  This catches a
reversal that starts from the wrong state, a checkpoint overwritten before
its last use, a segment reversed out of order or twice, a record range
reused while its segment is still to be reversed, and a schedule that
needs more buffers than its budget.
"""
import pytest

from starmap_amd import _lib
from starmap_amd.adjoint import adjoint_schedule

WORK, TAIL_START, PREV, NONE = -2, -3, -4, -1


def replay(n_steps, every, buffers, rec_factor=1):
    st, fwd, rev = adjoint_schedule(n_steps, every, buffers, rec_factor)
    k = min(every, n_steps)
    U = -(-n_steps // k)
    assert st["segments"] == U and st["steps_per_segment"] == k
    T = st["recorded_in_forward"]
    H = U - T
    buf = {}        # buffer id -> segment whose start state it holds
    rng = {}        # record range -> segment whose records it holds
    state = 0       # the handle's state: start of segment `state`
    fwd_at = 0      # next segment the forward pass computes
    used_slots, used_ranges, work, tail_start = set(), set(), False, False

    def start_ok(seg, src, r):
        if src == PREV:
            return rng.get(r - 1) == seg - 1
        return buf.get(src) == seg

    for kind, s0, s1, src, dst, r in fwd:
        if kind == "checkpoint":
            assert s0 == state == fwd_at
            buf[dst] = state
            if dst >= 0:
                used_slots.add(dst)
            elif dst == TAIL_START:
                tail_start = True
        elif kind == "run":
            assert s0 == state == fwd_at and s1 > s0
            state = fwd_at = s1
        elif kind == "record":
            assert s0 == fwd_at and start_ok(s0, src, r)
            rng[r] = s0
            used_ranges.add(r)
            fwd_at = s0 + 1
        elif kind == "final":
            assert rng.get(r) == U - 1 == s0
            state = U
        else:
            raise AssertionError("unexpected forward action " + kind)
    assert fwd_at == U and state == U, "the forward pass ends at the final time"

    nxt = U - 1
    recomputed, advanced = 0, 0
    for kind, s0, s1, src, dst, r in rev:
        if kind == "advance":
            assert buf.get(src) == s0 and s1 > s0
            buf[dst] = s1
            advanced += s1 - s0
            if dst >= 0:
                used_slots.add(dst)
            else:
                assert dst == WORK
                work = True
        elif kind == "recompute":
            assert buf.get(src) == s0
            # the range must not hold a segment still to be reversed
            assert rng.get(r) is None or rng[r] > nxt
            rng[r] = s0
            used_ranges.add(r)
            recomputed += 1
        elif kind == "reverse":
            assert s0 == nxt, "segments are reversed from the last to the first"
            assert rng.get(r) == s0, "the segment's records are in its range"
            assert start_ok(s0, src, r), "the segment's start state is at hand"
            nxt -= 1
        else:
            raise AssertionError("unexpected reverse action " + kind)
    assert nxt == -1, "every segment is reversed"
    assert recomputed == st["recomputed_segments"] == H
    assert advanced == st["extra_segment_advances"]
    peak = ((max(used_ranges) + 1 if used_ranges else 0) * rec_factor * (k + 1) + len(used_slots) + work
            + tail_start)
    assert peak <= st["peak_buffers"]
    if buffers >= 0:
        assert st["peak_buffers"] <= buffers
    return st


CASES = [(1, 1), (1, 50), (7, 2), (50, 50), (123, 10), (2000, 50), (333, 7), (40, 1)]


@pytest.mark.parametrize("n_steps,every", CASES)
def test_default_schedule_is_one_checkpoint_per_segment(n_steps, every):
    st = replay(n_steps, every, -1)
    k = min(every, n_steps)
    U = -(-n_steps // k)
    assert st["recorded_in_forward"] == 0 and st["checkpoint_slots"] == U
    assert st["recomputed_segments"] == U and st["extra_segment_advances"] == 0
    assert st["peak_buffers"] == U + 2 * (k + 1)


@pytest.mark.parametrize("rec_factor", [1, 2])
@pytest.mark.parametrize("n_steps,every", CASES)
def test_budget_schedules_replay_and_trade_memory_for_recomputation(n_steps, every, rec_factor):
    """rec_factor 2: a step's records hold the launches' W+Q beside their
    states, so a recorded segment takes twice the buffers."""
    k = min(every, n_steps)
    U = -(-n_steps // k)
    full = U * rec_factor * (k + 1) + 1
    rb = rec_factor * (k + 1)
    budgets = sorted({full, full + 5, full - 1, full // 2, full // 4, 2 * rb + 4,
                      2 * rb + 3, 3 * rb + 2, U + 2 * rb + 1})
    cost = []
    for b in budgets:
        if b < 2 * rec_factor * (k + 1) + 3:
            continue
        st = replay(n_steps, every, b, rec_factor)
        cost.append((b, st["recomputed_segments"] + st["extra_segment_advances"]))
    # more memory never costs more forward work
    for (b0, c0), (b1, c1) in zip(cost, cost[1:]):
        assert c1 <= c0, (b0, c0, b1, c1)
    # everything recorded in the forward pass when it fits
    st = replay(n_steps, every, full, rec_factor)
    assert st["recorded_in_forward"] == U and st["recomputed_segments"] == 0


def test_small_budget_uses_binomial_checkpointing():
    # 40 segments, room for two record ranges and a few checkpoints only
    st = replay(2000, 50, 2 * 51 + 6)
    assert st["recorded_in_forward"] == 0
    assert st["checkpoint_slots"] < st["segments"]
    assert st["extra_segment_advances"] > 0
    # with a single checkpoint every segment is re-advanced from the start
    st1 = replay(20, 1, 2 * 2 + 2)
    assert st1["checkpoint_slots"] == 1
    assert st1["extra_segment_advances"] == 20 * 19 // 2


def test_binomial_advances_are_optimal_on_small_cases():
    """Against an exhaustive search of the split points: the fewest extra
    segment advances to reverse l segments that the forward pass crosses
    (advancing through them on its way is free), with f slots beside the
    one holding their start."""
    import functools

    @functools.lru_cache(None)
    def N(l, f):
        if l == 1:
            return 0
        best = l * (l - 1) // 2   # every segment from the start
        if f > 0:
            for m in range(1, l):
                best = min(best, m + N(l - m, f - 1) + N(m, f))
        return best

    @functools.lru_cache(None)
    def S(l, f):
        if l == 1:
            return 0
        best = l * (l - 1) // 2
        if f > 0:
            for m in range(1, l):
                best = min(best, S(l - m, f - 1) + N(m, f))
        return best

    for U in (5, 9, 14):
        for c in (1, 2, 3, 4):
            # one-step segments; a budget with room for the two record
            # ranges (2 buffers each), c slots and the scratch state, too
            # small for any recorded segment (3 ranges would be needed)
            b = 2 * 2 + c + 1
            st = replay(U, 1, b)
            if st["recorded_in_forward"] or st["checkpoint_slots"] != c:
                continue
            assert st["extra_segment_advances"] == S(U, c - 1), (U, c)


def test_budget_below_one_segment_is_an_error():
    with pytest.raises(_lib.SmpError) as e:
        adjoint_schedule(2000, 50, 10)
    assert e.value.code == _lib.SMP_E_OOM
    with pytest.raises(_lib.SmpError):
        adjoint_schedule(0, 50, -1)
