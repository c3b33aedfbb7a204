"""The adjoint sweep's binomial checkpoint schedule (emu_amd.adjoint,
segment="binomial"), on the CPU: every step is reversed once, last first;
no more than the given number of step states is held; and the forward steps
it spends are the fewest any schedule of this kind can spend, found by brute
force over every split, and equal to Griewank's closed form for revolve."""
import functools
import math
import weakref

import pytest

from emu_amd.adjoint import binomial_reverse, binomial_split


class State:
    def __init__(self, n):
        self.n = n


@functools.lru_cache(maxsize=None)
def fewest_steps(t, s):
    """Least forward steps to reverse t steps from a held state with s free
    slots, over every choice of where each snapshot goes."""
    if t <= 1:
        return 0
    if s == 0:
        return t * (t - 1) // 2
    return min(m + fewest_steps(t - m, s - 1) + fewest_steps(m, s) for m in range(1, t))


def griewank(t, s):
    """Revolve's forward-step count for t steps with S = s + 1 states (the
    start counted): r t - C(S + r, r - 1), r the least repetition count with
    C(S + r, S) >= t (Griewank, ACM TOMS 26, 2000)."""
    if t <= 1:
        return 0
    S = s + 1
    r = 1
    while math.comb(S + r, S) < t:
        r += 1
    return r * t - math.comb(S + r, r - 1)


def run(t, s):
    live = weakref.WeakSet()
    steps = [0]
    order = []
    peak = [0]

    def advance(state, n):
        assert n >= state.n
        steps[0] += n - state.n
        x = State(n)
        live.add(x)
        return x

    def reverse(state, n):
        assert state.n == n
        order.append(n)
        peak[0] = max(peak[0], len(live))

    start = State(0)
    binomial_reverse(start, 0, t, s, advance, reverse)
    return steps[0], order, peak[0]


@pytest.mark.parametrize("s", [0, 1, 2, 3, 5])
def test_binomial_schedule_reverses_every_step_within_its_slots(s):
    for t in range(0, 70):
        steps, order, peak = run(t, s)
        assert order == list(range(t - 1, -1, -1))
        # s snapshots, plus the state handed to the step being reversed
        assert peak <= s + 1


@pytest.mark.parametrize("s", [1, 2, 3, 4])
def test_binomial_schedule_is_the_fewest_forward_steps(s):
    for t in range(1, 90):
        steps, _, _ = run(t, s)
        assert steps == fewest_steps(t, s), (t, s)
        assert steps == griewank(t, s), (t, s)


def test_binomial_split_bounds():
    with pytest.raises(ValueError):
        binomial_split(1, 2)
    with pytest.raises(ValueError):
        binomial_split(5, 0)
    for s in range(1, 5):
        for t in range(2, 200):
            m = binomial_split(t, s)
            assert 1 <= m < t


def test_recompute_without_slots():
    steps, order, _ = run(10, 0)
    assert steps == 45 and order == list(range(9, -1, -1))
