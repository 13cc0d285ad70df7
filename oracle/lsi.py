"""CPU oracle of the learned seed index (float64 numpy).

TEST INFRASTRUCTURE ONLY: tests/, the smoke test and bench.py's cpu_baseline
leg use this module as the checker; the product package never imports it.

The learned seed index answers, for an exact-ACGT k-mer W (k = 21 by
default), the bidirectional SA interval that plain backward search gives:
extend_left over W's codes from the full interval (bidir.rs:112-130, the
exact single-code extension used by LookupTable::build, lookup.rs:37-91).
Keys pack W two bits per base, FIRST base MOST significant (A=0 .. T=3), so
the sorted key order is the suffix-array order of the k-mers.

What the oracle restates, each pinned by a test:
  kmer_table       every distinct exact-ACGT k-mer of the text with its
                   (fwd_lo, fwd_hi, rev_lo); pinned by the naive suffix array
                   (fwd_lo = #suffixes below W, fwd_hi - fwd_lo = #occurrences)
  mlp_forward /    stage-1 model f(x) = w2 . clamp(W1 x + b1, 0, 1) + b2 (a sum
  mlp_grad         of ramps: non-decreasing when W1, w2 >= 0, which training
                   keeps) and the gradient of mean((f - y)^2); pinned by
                   central finite differences
  fit_leaf         least-squares line of one leaf; pinned by numpy.linalg.lstsq
  lookup           key -> interval or None, by binary search of the table
"""
from __future__ import annotations

import numpy as np

from . import oracle as O

ACGT = (1, 2, 3, 4)  # codes of A, C, G, T


def key_of(codes) -> int:
    """Two bits per base, first base most significant; None if not ACGT."""
    key = 0
    for c in codes:
        c = int(c)
        if c < 1 or c > 4:
            return None
        key = (key << 2) | (c - 1)
    return key


def codes_of(key: int, k: int) -> np.ndarray:
    return np.array([((key >> (2 * (k - 1 - t))) & 3) + 1 for t in range(k)], np.uint8)


def exact_search(bidir, codes):
    """Plain exact backward search of codes on a BidirFmIndex (bidir.rs:112-130)."""
    n = bidir.fwd.text_len
    iv = (0, n, 0, n)
    for c in reversed([int(x) for x in codes]):
        iv = O.extend_left(bidir.fwd, iv, c)
        if iv is None or iv[1] <= iv[0]:
            return None
    return iv[0], iv[1], iv[2]


def kmer_table(text: np.ndarray, bidir, k: int):
    """Sorted distinct exact-ACGT k-mers of text with their intervals."""
    keys = set()
    t = np.asarray(text)
    for p in range(len(t) - k + 1):
        key = key_of(t[p:p + k])
        if key is not None:
            keys.add(key)
    keys = np.array(sorted(keys), np.uint64)
    iv = np.zeros((len(keys), 3), np.uint32)
    for i, key in enumerate(keys):
        iv[i] = exact_search(bidir, codes_of(int(key), k))
    return keys, iv


def features(keys: np.ndarray, k: int) -> np.ndarray:
    """x = key / 4^k in [0, 1)."""
    return keys.astype(np.float64) / float(4 ** k)


def mlp_forward(params: dict, x: np.ndarray):
    h_pre = np.outer(x, params["w1"]) + params["b1"]
    h = np.clip(h_pre, 0.0, 1.0)
    return h @ params["w2"] + params["b2"], h_pre, h


def mlp_loss(params: dict, x: np.ndarray, y: np.ndarray) -> float:
    f, _, _ = mlp_forward(params, x)
    return float(np.mean((f - y) ** 2))


def mlp_grad(params: dict, x: np.ndarray, y: np.ndarray) -> dict:
    """Gradient of mean((f - y)^2) with respect to every parameter."""
    f, h_pre, h = mlp_forward(params, x)
    df = 2.0 * (f - y) / len(x)
    dh = np.outer(df, params["w2"]) * ((h_pre > 0) & (h_pre < 1))
    return {"w1": dh.T @ x, "b1": dh.sum(axis=0), "w2": h.T @ df, "b2": np.array(df.sum())}


def fit_leaf(dx: np.ndarray, y: np.ndarray) -> tuple[float, float]:
    """Least-squares y ~ a * dx + b (a = 0 for fewer than two distinct dx)."""
    n = len(dx)
    if n == 0:
        return 0.0, 0.0
    sx, sy = dx.sum(), y.sum()
    sxx, sxy = (dx * dx).sum(), (dx * y).sum()
    den = n * sxx - sx * sx
    if n < 2 or den <= 0:
        return 0.0, sy / n
    a = (n * sxy - sx * sy) / den
    return a, (sy - a * sx) / n


def lookup(keys: np.ndarray, iv: np.ndarray, key: int):
    i = int(np.searchsorted(keys, np.uint64(key)))
    if i < len(keys) and int(keys[i]) == key:
        return tuple(int(v) for v in iv[i])
    return None


# ---------------------------------------------------------------------------
# Restart lookup model (the learned index's entries, hsfm_lsi_dev.h): a CPU
# statement of what the device builds and how a lookup answers, checked
# against brute force over the text by tests/test_lsi_restart_model.py.
#
# Entry keys hold a string's codes from the top of a 64-bit word down, LAST
# code first (A=0 .. T=3), so sorted keys put strings that share a suffix
# next to each other; the low bits are metadata.
#   full entries:  every distinct text window t of k codes with no cutting
#                  symbol, once per base string y compatible with t; payload =
#                  t (its interval on the device); metadata 0 for an all-ACGT t,
#                  VARIANT otherwise
#   short entries: windows of r < k codes after a contig start (position 0 or
#                  one past a cutting symbol), once per base string; metadata
#                  SHORT with the neighbour bounds G and H
# A cutting symbol is one no base is compatible with (the sentinel; under
# ExactDna every ambiguity code).
# ---------------------------------------------------------------------------
VARIANT = 1
SHORT = 1 << 13
G_SHIFT, H_SHIFT = 8, 3


def base_sets(tag: int = 0) -> list[int]:
    """qb[code]: bit b set when base b+1 (A, C, G, T) is compatible with text
    code `code` (alphabet.rs compat table)."""
    out = []
    for c in range(16):
        m = 0
        for b in range(4):
            if c in O.compatible(b + 1, tag):
                m |= 1 << b
        out.append(m)
    return out


def restart_key(digits, r: int) -> int:
    """Key of a base string of r codes (digits 0..3, first code first)."""
    key = 0
    for i, d in enumerate(digits):
        key |= int(d) << (64 - 2 * r + 2 * i)
    return key


def _expansions(codes, qb):
    out = [[]]
    for c in codes:
        m = qb[int(c)]
        out = [p + [b] for p in out for b in range(4) if (m >> b) & 1]
    return out


def restart_entries(text, k: int, tag: int = 0):
    """Sorted entries [(key, reach, payload)], payload = the text window (tuple
    of codes) for full entries, None for short ones; G and H of short entries
    are in their keys."""
    t = np.asarray(text)
    n = len(t)
    qb = base_sets(tag)
    live = [qb[int(c)] != 0 for c in t]
    exact = [int(c) in (1, 2, 3, 4) for c in t]
    ent = {}
    for s in range(n - k + 1):
        if all(live[s:s + k]):
            w = tuple(int(c) for c in t[s:s + k])
            meta = 0 if all(exact[s:s + k]) else VARIANT
            for y in _expansions(w, qb):
                ent[(restart_key(y, k) | meta, w)] = (k, w)
    starts = [0] + [p + 1 for p in range(n) if not live[p]]
    for s in starts:
        for r in range(1, k):
            if s + r > n or not live[s + r - 1]:
                break
            for x in _expansions(t[s:s + r], qb):
                key = restart_key(x, r) | SHORT | (r << G_SHIFT) | (r << H_SHIFT)
                ent[(key, None)] = (r, None)
    rows = sorted(((key, reach, w) for (key, w), (reach, _) in ent.items()), key=lambda e: (e[0], e[2] or ()))
    smask = ~0 << (64 - 2 * k) & ((1 << 64) - 1)
    keys = [e[0] for e in rows]
    reach = [e[1] for e in rows]

    def lcs(a, b):
        x = (a ^ b) & smask
        return k if x == 0 else min(k, (64 - x.bit_length()) // 2)

    G = [0] * len(rows)
    H = [0] * len(rows)
    for i in range(len(rows)):  # G_i = max(r_i, min(G_{i-1}, lcs(x_{i-1}, x_i)))
        G[i] = reach[i] if i == 0 else max(reach[i], min(G[i - 1], lcs(keys[i - 1], keys[i])))
    for i in range(len(rows) - 1, -1, -1):
        H[i] = reach[i] if i == len(rows) - 1 else max(reach[i], min(H[i + 1], lcs(keys[i], keys[i + 1])))
    out = []
    for i, (key, r, w) in enumerate(rows):
        if key & SHORT:
            key = (key & ~((31 << G_SHIFT) | (31 << H_SHIFT))) | (G[i] << G_SHIFT) | (H[i] << H_SHIFT)
        out.append((key, r, w))
    return out


def restart_lookup(entries, k: int, codes):
    """(set of text windows compatible with the base string `codes`, longest
    suffix length) as the device lookup answers: the run of full entries at
    the lower bound, else the neighbour formula over G / H."""
    smask = ~0 << (64 - 2 * k) & ((1 << 64) - 1)
    w = restart_key([int(c) - 1 for c in codes], k)
    keys = [e[0] for e in entries]
    lb = int(np.searchsorted(np.array(keys, dtype=np.uint64), np.uint64(w), side="left")) if keys else 0
    run = []
    j = lb
    while j < len(entries) and (keys[j] & smask) == w and not (keys[j] & SHORT):
        run.append(entries[j][2])
        j += 1
    if run:
        return set(run), k

    def lcs(a, b):
        x = (a ^ b) & smask
        return k if x == 0 else min(k, (64 - x.bit_length()) // 2)

    lam = 0
    if lb > 0:
        e = keys[lb - 1]
        lam = max(lam, min(lcs(w, e), (e >> G_SHIFT) & 31 if e & SHORT else k))
    if lb < len(keys):
        e = keys[lb]
        lam = max(lam, min(lcs(w, e), (e >> H_SHIFT) & 31 if e & SHORT else k))
    return set(), lam


def restart_brute(text, k: int, codes, tag: int = 0):
    """Brute force over the text: the distinct windows of k codes compatible
    with `codes`, and the longest suffix of `codes` with a compatible
    occurrence (k if the set is not empty)."""
    t = [int(c) for c in np.asarray(text)]
    qb = base_sets(tag)
    q = [int(c) for c in codes]

    def ok(qc, tc):
        return (qb[tc] >> (qc - 1)) & 1

    sets = {tuple(t[s:s + k]) for s in range(len(t) - k + 1) if all(ok(q[i], t[s + i]) for i in range(k))}
    if sets:
        return sets, k
    best = 0
    for p in range(1, len(t) + 1):
        l = 0
        while l < k and p - l - 1 >= 0 and ok(q[k - 1 - l], t[p - l - 1]):
            l += 1
        best = max(best, l)
    return set(), best
