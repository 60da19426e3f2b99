"""The binomial (revolve) checkpointing schedule the adjoint runs, on the host (no device).

Pins: replaying the actions on a model of the run (the step whose state the context holds, the
states the slots hold) shows that every adjoint step finds its own input state, that they come in
reverse order from the first turn on, and that no more slots are used than granted; every step
runs forward at most r times, r the smallest with C(s + r, r) >= n; the number of forward steps
equals the minimum over every recursive split schedule (a dynamic program over all splits, small n
and s) and Griewank's closed form r n - C(s + r, r - 1) for larger n.  s counts the slots with the
one holding the initial state; the adjoint of a step records its own forward step, which the count
leaves out.
"""
import math
from functools import lru_cache

import pytest


@pytest.fixture(scope="module")
def sched(g):
    return g.revolve_schedule


def replay(actions, n, s):
    """(forward steps, forward runs per step, slots used); asserts the schedule's invariants."""
    cur, slots, adj, runs, fwd = 0, {}, [], [0] * n, 0
    for op, step, slot in actions:
        if op == "snapshot":
            assert step == cur
            slots[slot] = cur
        elif op == "restore":
            assert slots.get(slot) == step
            cur = step
        elif op == "advance":
            assert step > cur
            for t in range(cur, step):
                runs[t] += 1
            fwd += step - cur
            cur = step
        elif op == "first_turn":
            assert step == cur == n - 1 and not adj
            adj.append(step)
        else:
            assert op == "adjoint" and step == cur
            adj.append(step)
    assert adj == list(range(n - 1, -1, -1))
    assert 0 in slots and max(slots) < min(s, n)
    return fwd, runs, len(slots)


@lru_cache(maxsize=None)
def best(n, s):
    """The fewest forward steps that reverse n steps from a held state with s slots (its own
    included), over every split; with one slot each step is recomputed from the held state."""
    if n == 1:
        return 0
    if s == 1:
        return n * (n - 1) // 2
    return min(m + best(n - m, s - 1) + best(m, s) for m in range(1, n))


def repetitions(n, s):
    r = 1
    while math.comb(s + r, r) < n:
        r += 1
    return r


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 11, 13, 21, 34, 55, 60])
@pytest.mark.parametrize("s", [1, 2, 3, 4, 6])
def test_the_schedule_is_valid_and_minimal(sched, n, s):
    fwd, runs, _ = replay(sched(n, s), n, s)
    if n > 1:
        assert max(runs) <= repetitions(n, min(s, n))
    assert fwd == best(n, min(s, n))


@pytest.mark.parametrize("n,s", [(200, 31), (200, 11), (200, 5), (200, 2), (1000, 9), (5000, 20)])
def test_the_forward_count_is_griewanks_minimum(sched, n, s):
    fwd, runs, used = replay(sched(n, s), n, s)
    r = repetitions(n, s)
    assert fwd == r * n - math.comb(s + r, r - 1)
    assert max(runs) <= r and used <= s


def test_bad_arguments_are_refused(sched):
    for n, s in [(0, 3), (5, 0), (-1, 2)]:
        with pytest.raises(ValueError):
            sched(n, s)
