"""The in-place permutation plan of svx_permute (host code, no GPU): the steps compose to the asked
permutation, each step moves only bits of one 12-bit sweep tile (the low 6 bits plus at most 6
others, so every workgroup rewrites the amplitudes it read), and the sweep counts are the ones the
header states.  The reference exchanges bank ranges in place (kernels/sv/sv_swap.cpp:10-39,
host/sv/runtime.py:235-259); here any permutation is a short sequence of such sweeps."""
import numpy as np
import pytest

from u55c_statevector_30q_amd import libsvx


def _compose(steps, n):
    cur = np.arange(n)   # cur[b]: original bit held at position b
    for s in steps:
        cur = cur[s]     # position b after the step holds what position s[b] held before it
    return cur


def _check_plan(perm):
    n = len(perm)
    steps = libsvx.permute_plan(perm, in_place=True)
    assert np.array_equal(_compose(steps, n), np.asarray(perm)), (perm, steps)
    for s in steps:
        moved = [b for b in range(n) if s[b] != b]
        assert sorted(s) == list(range(n))
        assert sum(1 for b in moved if b >= 6) <= min(12, n) - min(6, n), (perm, s)
        assert set(int(s[b]) for b in moved) == set(moved)   # the step maps its tile bits onto themselves
    return len(steps)


@pytest.mark.parametrize("n", [1, 3, 6, 7, 12, 13, 20, 30, 34, 40])
def test_bit_reversal_sweeps(n):
    perm = list(range(n))[::-1]
    steps = _check_plan(perm)
    pairs_high = max(0, n // 2 - 6)       # pairs (i, n-1-i) with both bits >= 6
    expect = 0 if n == 1 else (1 if n <= 12 else 1 + -(-pairs_high // 3))
    assert steps == expect, (n, steps)


def test_random_permutations_compose():
    rng = np.random.default_rng(11)
    worst = 0
    for _ in range(300):
        n = int(rng.integers(1, 41))
        perm = [int(x) for x in rng.permutation(n)]
        worst = max(worst, _check_plan(perm))
    assert worst <= 40


def test_identity_and_tile_local_permutations_are_at_most_one_sweep():
    assert len(libsvx.permute_plan(list(range(30)))) == 0
    perm = list(range(30))
    perm[0], perm[29] = 29, 0
    perm[3], perm[17] = 17, 3
    assert _check_plan(perm) == 1
    assert len(libsvx.permute_plan(perm, in_place=False)) == 1


def test_bad_permutations_raise():
    with pytest.raises(ValueError):
        libsvx.permute_plan([0, 0, 1])
    with pytest.raises(ValueError):
        libsvx.permute_plan([0, 3, 1])
