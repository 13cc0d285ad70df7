"""CPU model of the learned seed index's bounded search (lsi_range,
lsi_fence_step, lsi_last_below and lsi_keys_part in
haystackfm_amd/csrc/hsfm_lsi_dev.h): the lower bound of a key lies in [lo, hi
+ 1] = [pred - err, pred + err + 1]; fence levels of every 2^(5+4v)-th key
narrow it, each read as one round of at most 9 16-byte pairs, and the last
round reads slots [lo - 1, hi + 1] as at most 18 pairs, giving the lower bound
and both neighbour keys.  The model reads exactly the entries the kernel
reads, so a level count that leaves more than 17 entries for a round loses
keys here; the check is bisection on the same sorted array.  The GPU tests
assert that the library builds the level count this rule gives
(hsfm_lsi_info.fence_levels)."""
import bisect

import numpy as np
import pytest

MAX_LEVELS = 7
U64 = (1 << 64) - 1


def shift(v):
    return 5 + 4 * v


def levels_for(emax):
    if 2 * emax + 1 <= 32:
        return 0
    nl = 1
    while nl < MAX_LEVELS and ((2 * emax + 1) >> shift(nl - 1)) > 16:
        nl += 1
    return nl


class VirtualKeys:
    """Strictly increasing keys k(i) = 3 i + ((i * 2654435761) >> 7) % 3 for
    i < n, computed on demand (n up to 2^32 slots); entries past the end read
    as the largest key value, as the padded device buffers do."""

    def __init__(self, n):
        self.n = n

    def __getitem__(self, i):
        if i >= self.n:
            return U64
        return 3 * i + ((i * 2654435761) >> 7) % 3

    def fence(self, sh):
        return lambda j: self[j << sh] if (j << sh) < self.n else U64


def last_below(f, a, z, j, key):
    if a <= z:
        b = a & ~1
        c = 0
        for u in range(9):
            e = b + 2 * u
            x, y = (f(e), f(e + 1)) if e <= z else (U64, U64)
            c += (a <= e <= z and x < key) + (a <= e + 1 <= z and y < key)
        if c:
            j = a + c - 1
    return j


def lookup(keys, nlev, pred, err, key):
    """(lower bound, key before it or None, key at it or None) as the kernel
    computes them (leaf start 0)."""
    n = keys.n
    lo = min(max(pred - err, 0), n)
    hi = min(pred + err + 1, n - 1)
    hi = max(hi, lo - 1)
    if lo <= hi:
        for lv in range(nlev - 1, -1, -1):
            sh = shift(lv)
            j = last_below(keys.fence(sh), (lo >> sh) + 1, hi >> sh, lo >> sh, key)
            lo, hi = max(lo, j << sh), min(hi, (j + 1) << sh)
    first, last = max(lo - 1, 0), min(hi + 1, n - 1)
    below, pv, sv = 0, None, None
    b = first & ~1
    for part in range(2):
        bb = b + 18 * part
        if part == 1 and (sv is not None or bb > last):
            break
        for u in range(9):
            a0 = bb + 2 * u
            x, y = (keys[a0], keys[a0 + 1]) if a0 <= last else (U64, U64)
            for idx, v in ((a0, x), (a0 + 1, y)):
                if first <= idx <= last:
                    if v < key:
                        below += 1
                        pv = v
                    elif sv is None:
                        sv = v
    return first + below, pv, sv


def _cases(rng, n, emax, count):
    for _ in range(count):
        slot = int(rng.integers(1, n))
        err = int(rng.integers(0, emax + 1))
        pred = slot + int(rng.integers(-err - 1, err + 1))  # the lower bound lies in [pred - err, pred + err + 1]
        yield slot, pred, err


@pytest.mark.parametrize("emax", [0, 7, 255, 4095, 69_632, 340_899, 1 << 20, 1 << 24, (1 << 31) - 1])
def test_levels_rule_finds_every_lower_bound(emax):
    rng = np.random.default_rng(emax)
    keys = VirtualKeys((1 << 32) - 1)
    nlev = levels_for(emax)
    for slot, pred, err in _cases(rng, keys.n, emax, 300):
        for key in (keys[slot], keys[slot] - 1 if keys[slot] - 1 > keys[slot - 1] else keys[slot]):
            lb, pv, sv = lookup(keys, nlev, pred, err, key)
            assert (lb, pv, sv) == (slot, keys[slot - 1], keys[slot]), (emax, nlev, slot, pred, err)


def test_one_level_fewer_loses_keys():
    """The rule is needed: with one level fewer than it gives, windows of the
    largest error lose lower bounds (as a fixed two levels did beyond
    max_err = 69,631, once the coarsest round could not cover its window)."""
    rng = np.random.default_rng(5)
    keys = VirtualKeys(1 << 32)
    for emax in (200_000, 1 << 20, 1 << 26):
        nlev = levels_for(emax)
        assert nlev > 2
        misses = 0
        for slot, pred, err in _cases(rng, keys.n, emax, 200):
            lb, _, _ = lookup(keys, nlev - 1, pred, err, keys[slot])
            misses += lb != slot
        assert misses > 0


def test_small_arrays_match_bisection():
    rng = np.random.default_rng(9)
    for n in (1, 2, 3, 33, 34, 35, 36, 100, 5000):
        arr = np.unique(rng.integers(0, 10 * n + 10, size=n * 2, dtype=np.uint64))[:n]
        n = len(arr)
        keys = VirtualKeys(n)
        keys.__class__ = type("Arr", (VirtualKeys,), {"__getitem__": lambda s, i: int(arr[i]) if i < n else U64})
        vals = arr.tolist()
        for q in range(0, int(arr[-1]) + 3):
            lb = bisect.bisect_left(vals, q)
            err = int(rng.integers(0, n + 1))
            pred = lb + int(rng.integers(-err - 1, err + 1))
            got = lookup(keys, levels_for(err), pred, err, q)
            assert got == (lb, vals[lb - 1] if lb else None, vals[lb] if lb < n else None), (n, q, pred, err)


def test_levels_rule_values():
    assert levels_for(0) == 0
    assert levels_for(15) == 0            # the last round alone covers [lo, hi]
    assert levels_for(16) == 1
    assert levels_for(271) == 1           # (2 emax + 1) >> 5 <= 16
    assert levels_for(272) == 2
    assert levels_for(4351) == 2          # (2 emax + 1) >> 9 <= 16
    assert levels_for(4352) == 3
    assert levels_for(69_631) == 3
    assert levels_for(69_632) == 4
    assert levels_for(340_899) == 4
    assert levels_for((1 << 32) - 1) == 7


def leaf_range(pred, err, s0, n):
    """lsi_range: [lo, hi] with the lower bound in [lo, hi + 1]."""
    lo = min(max(pred - err, s0), n)
    hi = max(min(pred + err + 1, n - 1), lo - 1)
    return lo, hi


def test_leaf_range_holds_every_lower_bound():
    """Brute force over small leaves: for every key of a leaf's key space (not
    only its entries), the lower bound lies in [lo, hi + 1] when the leaf's
    error covers its entries and its prediction at the leaf's largest key past
    the leaf's end (k_lsi_err after k_lsi_err_top); with the entries' error
    alone, keys above the last entry (and the keys of empty leaves) are missed
    whenever the model predicts past the end."""
    rng = np.random.default_rng(5)
    space = 64  # the keys of one leaf: 0 .. 63, u = key / 64
    misses_entries_only = 0
    for _ in range(400):
        s0 = int(rng.integers(0, 50))
        ent = np.sort(rng.choice(space, size=int(rng.integers(0, 8)), replace=False))
        after = int(rng.integers(0, 50))
        n = s0 + len(ent) + after
        a, b = float(rng.uniform(0, 40)), float(rng.uniform(-10, 10))

        def pred(w):  # non-decreasing in the key, as the stage-1 ramps and a >= 0 make it
            return s0 + int(np.floor(a * (w / space) + b))

        err_entries = max((abs(pred(int(x)) - (s0 + i)) for i, x in enumerate(ent)), default=0)
        end = s0 + len(ent)
        err = max(err_entries, pred(space - 1) - end)
        for w in range(space):
            lb = s0 + int(np.searchsorted(ent, w))
            lo, hi = leaf_range(pred(w), err, s0, n)
            assert lo <= lb <= hi + 1, (s0, ent.tolist(), a, b, w)
            lo, hi = leaf_range(pred(w), err_entries, s0, n)
            misses_entries_only += not (lo <= lb <= hi + 1)
    assert misses_entries_only > 0
