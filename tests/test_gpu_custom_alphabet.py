"""GPU SMEM/MEM with a caller-supplied alphabet (the compat_mask of the index
view, alphabet_fns.compatible_fn in the reference, src/alphabet.rs:203-212),
checked against brute force over every text position.

The alphabet here matches each base only to itself and N only to N: no base
set describes it (N has no bases yet matches something), so the device keeps
byte comparison for single-row extensions and restarts windows holding N
unseeded.  The checker states the SMEM definition directly: for each read
position i the longest compatible match end R(i); (i, R(i)) is reported when
R(i) - i >= min_len and R(i - 1) < R(i); its occurrences are every text
position matching q[i..R(i)), one record per distinct matched text string."""
import numpy as np
import pytest

import haystackfm_amd as hf
from oracle import oracle as O
from tests.helpers import random_text

pytestmark = pytest.mark.gpu

N = 5


def _strict_compatible(code: int):
    return [code] if 1 <= code <= 5 else []


STRICT = hf.AlphabetFns(_strict_compatible, (1, 2, 3, 4), 200)


def _compat_table():
    t = np.zeros((256, 16), bool)
    for b in range(256):
        for s in _strict_compatible(b):
            t[b, s] = True
    return t


def _brute_smems(text, bounds, q, min_len):
    """[(start, end, [(count, [(ref, off), ...]) per distinct string sorted by string])]"""
    tab = _compat_table()
    n, m = len(text), len(q)
    ends = []
    for i in range(m):
        alive = np.ones(n, bool)
        e = i
        for k in range(i, m):
            nxt = np.zeros(n, bool)
            p = np.nonzero(alive)[0]
            p = p[p + (k - i) < n]
            ok = tab[q[k], text[p + (k - i)]]
            nxt[p[ok]] = True
            if not nxt.any():
                break
            alive = nxt
            e = k + 1
        ends.append((e, np.nonzero(alive)[0] if e > i else np.zeros(0, np.int64)))
    out = []
    for i in range(m):
        e, occ = ends[i]
        if e - i < min_len or (i > 0 and ends[i - 1][0] >= e):
            continue
        groups = {}
        for p in occ.tolist():
            key = bytes(text[p:p + e - i])
            ref = int(np.searchsorted(bounds, p, side="right"))
            start = int(bounds[ref - 1]) if ref else 0
            groups.setdefault(key, []).append((ref, p - start))
        recs = [(len(v), sorted(v)) for _, v in sorted(groups.items())]
        out.append((i, e, recs))
    return out


@pytest.mark.parametrize("seed", [0, 1])
def test_strict_alphabet_smems_match_brute_force(seed):
    rng = np.random.default_rng(900 + seed)
    seqs = []
    for _ in range(3):
        s = random_text(rng, int(rng.integers(800, 1500)))
        s[rng.random(len(s)) < 0.03] = N
        seqs.append(O.decode(s))
    o = O.OracleBidir.build(seqs, sample_rate=8, threads=4)
    data = o.to_bytes()
    (fl,) = np.frombuffer(data[:4], "<u4")
    fwd = hf.FmIndex(data[4:4 + int(fl)], alphabet_fns=STRICT)
    rev = hf.FmIndex(data[4 + int(fl):], alphabet_fns=STRICT)
    g = hf.BidirFmIndex(fwd, rev)
    text, bounds = O.concatenate(seqs)
    text = np.asarray(text, np.uint8)
    bounds = np.asarray(bounds, np.uint32)
    reads = []
    for _ in range(30):
        a = int(rng.integers(0, len(text) - 70))
        q = text[a:a + 60].copy()
        q[q == 0] = 1
        sub = rng.random(60) < 0.04
        q[sub] = rng.integers(1, 6, int(sub.sum()))
        reads.append(q)
    min_len = 14
    got = g.find_smems_gpu(reads, min_len, bounds, 1 << 20)
    for q, hits in zip(reads, got):
        want = _brute_smems(text, bounds, q, min_len)
        have = {}
        for h in hits:
            assert not h.truncated
            have.setdefault((h.query_start, h.query_end), []).append((h.match_count, sorted(h.positions)))
        assert sorted(have) == [(i, e) for i, e, _ in want]
        for i, e, recs in want:
            assert sorted(have[(i, e)]) == sorted(recs), (i, e)
