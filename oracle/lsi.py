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
  mlp_forward /    stage-1 model f(x) = w2 . relu(W1 x + b1) + b2 and the
  mlp_grad         gradient of mean((f - y)^2); pinned by central finite
                   differences
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
    h = np.maximum(h_pre, 0.0)
    return h @ params["w2"] + params["b2"], h_pre, h


def mlp_loss(params: dict, x: np.ndarray, y: np.ndarray) -> float:
    f, _, _ = mlp_forward(params, x)
    return float(np.mean((f - y) ** 2))


def mlp_grad(params: dict, x: np.ndarray, y: np.ndarray) -> dict:
    """Gradient of mean((f - y)^2) with respect to every parameter."""
    f, h_pre, h = mlp_forward(params, x)
    df = 2.0 * (f - y) / len(x)
    dh = np.outer(df, params["w2"]) * (h_pre > 0)
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
