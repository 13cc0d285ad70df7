"""Pins of the learned-seed-index oracle (CPU only)."""
import numpy as np
import pytest

from oracle import lsi as L
from oracle import oracle as O


@pytest.fixture(scope="module")
def small():
    text, bounds = O.gen_reference(3, 4000, 3, gc=0.45, repeat_frac=0.3, n_units=4, unit_min=50, unit_max=200,
                                   iupac_frac=0.01)
    b = O.OracleBidir.build(text=text, bounds=bounds, sample_rate=4)
    return text, b


def test_kmer_table_matches_naive_suffix_array(small):
    text, b = small
    k = 7
    keys, iv = L.kmer_table(text, b, k)
    sa = O.suffix_array(text, naive=True)
    prefixes = [L.key_of(text[p:p + k]) if p + k <= len(text) else None for p in sa]
    for key, (lo, hi, _) in zip(keys[::7], iv[::7]):
        rows = [i for i, x in enumerate(prefixes) if x == int(key)]
        assert rows == list(range(lo, hi))
    # key order is suffix-array order
    assert np.all(np.diff(iv[:, 0].astype(np.int64)) > 0)


def test_kmer_table_rev_lo_matches_right_extension(small):
    text, b = small
    k = 9
    keys, iv = L.kmer_table(text, b, k)
    n = b.fwd.text_len
    for key, row in zip(keys[::11], iv[::11]):
        cur = (0, n, 0, n)
        for c in L.codes_of(int(key), k):
            cur = O.extend_right(b.rev, cur, int(c))
        assert (cur[0], cur[1], cur[2]) == tuple(int(v) for v in row)


def test_absent_kmers_have_no_interval(small):
    text, b = small
    k = 10
    keys, iv = L.kmer_table(text, b, k)
    present = set(int(x) for x in keys)
    rng = np.random.default_rng(4)
    for key in rng.integers(0, 4 ** k, 200):
        hit = L.exact_search(b, L.codes_of(int(key), k))
        assert (hit is not None) == (int(key) in present)
        assert L.lookup(keys, iv, int(key)) == hit


def test_mlp_gradient_matches_finite_differences():
    rng = np.random.default_rng(7)
    H = 8
    params = {"w1": rng.normal(size=H), "b1": rng.normal(size=H) * 0.1, "w2": rng.normal(size=H) * 0.3,
              "b2": np.array(0.05)}
    x = rng.random(64)
    y = x ** 1.3
    g = L.mlp_grad(params, x, y)
    eps = 1e-6
    for name in params:
        flat = np.atleast_1d(params[name]).astype(np.float64)
        for i in range(len(flat)):
            up = {kk: np.array(v, dtype=np.float64, copy=True) for kk, v in params.items()}
            dn = {kk: np.array(v, dtype=np.float64, copy=True) for kk, v in params.items()}
            np.atleast_1d(up[name])[i] += eps
            np.atleast_1d(dn[name])[i] -= eps
            num = (L.mlp_loss(up, x, y) - L.mlp_loss(dn, x, y)) / (2 * eps)
            assert abs(num - np.atleast_1d(g[name])[i]) < 1e-6 * max(1.0, abs(num))


def test_fit_leaf_matches_lstsq():
    rng = np.random.default_rng(9)
    dx = rng.random(300) * 1e6
    y = 3.5e-4 * dx + 17 + rng.normal(size=300)
    a, b = L.fit_leaf(dx, y)
    (a2, b2), *_ = np.linalg.lstsq(np.stack([dx, np.ones_like(dx)], 1), y, rcond=None)
    assert abs(a - a2) < 1e-9 and abs(b - b2) < 1e-6
    assert L.fit_leaf(np.array([5.0]), np.array([2.0])) == (0.0, 2.0)
