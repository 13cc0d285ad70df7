"""GPU checks of the learned seed index: exactness against plain exact search
and the oracle, training-kernel gradients against the float64 oracle, and what
the trained stage 1 gains over the untrained model."""
import numpy as np
import pytest

import haystackfm_amd as hf
from oracle import lsi as L
from oracle import oracle as O
from tests.test_lsi_fences import levels_for

pytestmark = pytest.mark.gpu


def _bidir(seed, total, **kw):
    text, bounds = O.gen_reference(seed, total, 3, gc=0.41, **kw)
    o = O.OracleBidir.build(text=text, bounds=bounds, sample_rate=16, threads=8)
    return text, o, hf.BidirFmIndex.from_bytes(o.to_bytes())


def _query_keys(rng, keys, k, n_absent=2000):
    absent = rng.integers(0, 4 ** k, n_absent, dtype=np.uint64)
    mutated = keys[rng.integers(0, len(keys), 2000)] ^ (np.uint64(1) << (2 * rng.integers(0, k, 2000)).astype(np.uint64))
    return np.concatenate([keys, absent, mutated, keys[:1] + np.uint64(1)])


@pytest.mark.parametrize("k", [9, 21])
def test_lsi_matches_oracle_table(k):
    """Every present k-mer of a small text and random absent ones: the learned
    index, plain exact search on the device and the oracle agree."""
    rng = np.random.default_rng(k)
    text, o, g = _bidir(60 + k, 12_000, repeat_frac=0.3, n_units=5, unit_min=100, unit_max=600, iupac_frac=0.01)
    keys, iv = L.kmer_table(text, o, k)
    lsi = hf.LearnedSeedIndex(g, k=k, leaves=64, hidden=8, train_steps=100, sample=4096)
    info = lsi.info()
    assert info["num_keys"] == len(keys)
    q = _query_keys(rng, keys, k)
    got = lsi.seed(q)
    plain = hf.kmer_search(g, q, k)
    assert np.array_equal(got, plain)
    for key, row in zip(q[::5], got[::5]):
        exp = L.lookup(keys, iv, int(key))
        assert (exp is None and row[3] == 0) or (exp is not None and tuple(row) == exp + (1,))


def test_lsi_exact_on_larger_index():
    """200 kbp with repeats and N: 21-mers of reads (1% substitutions) and of
    random sequence, learned index == plain exact search."""
    rng = np.random.default_rng(3)
    text, o, g = _bidir(77, 200_000, repeat_frac=0.45, n_units=40, unit_min=300, unit_max=3000,
                        iupac_frac=0.002, iupac_codes=[5])
    lsi = hf.LearnedSeedIndex(g, k=21, leaves=4096, hidden=16, train_steps=300)
    info = lsi.info()
    assert 0 < info["num_keys"] <= len(text)
    flat, off = O.gen_reads(5, text, o.fwd.seq_boundaries, 3000, 150, sub=0.01, both_strands=True)
    q = hf.kmer_keys(flat, 21)
    q = q[q != np.uint64(0xFFFFFFFFFFFFFFFF)]
    q = np.concatenate([q, rng.integers(0, 4 ** 21, 5000, dtype=np.uint64)])
    got = lsi.seed(q)
    assert np.array_equal(got, hf.kmer_search(g, q, 21))
    assert got[:, 3].sum() > len(q) // 3


def test_lsi_exact_with_wide_windows():
    """Untrained stage 1 and one leaf: error bounds of thousands of keys, so
    the lookup adds a third fence level; still equal to plain exact search."""
    leaves = 1
    rng = np.random.default_rng(leaves)
    text, o, g = _bidir(78, 200_000, repeat_frac=0.3, n_units=20, unit_min=300, unit_max=3000)
    lsi = hf.LearnedSeedIndex(g, k=21, leaves=leaves, hidden=8, train_steps=0)
    info = lsi.info()
    assert (2 * info["max_err"]) >> 9 > 16  # wider than 17 entries of the 512-key fence
    assert info["fence_levels"] == levels_for(info["max_err"]) >= 3
    flat, _ = O.gen_reads(6, text, o.fwd.seq_boundaries, 2000, 150, sub=0.01, both_strands=True)
    q = hf.kmer_keys(flat, 21)
    q = q[q != np.uint64(0xFFFFFFFFFFFFFFFF)]
    q = np.concatenate([q, _query_keys(rng, q, 21), np.array([0, 4 ** 21 - 1, q.min(), q.max()], np.uint64)])
    got = lsi.seed(q)
    assert np.array_equal(got, hf.kmer_search(g, q, 21))
    assert got[:, 3].sum() > len(q) // 3


@pytest.mark.parametrize("path", ["global", "lds"])
def test_lsi_both_lookup_kernels(path):
    """The lookup with its leaves in global memory and in LDS: equal to plain
    exact search, with two fence levels and with more."""
    text, o, g = _bidir(79, 150_000, repeat_frac=0.3, n_units=20, unit_min=300, unit_max=3000)
    flat, _ = O.gen_reads(7, text, o.fwd.seq_boundaries, 1500, 150, sub=0.01, both_strands=True)
    q = hf.kmer_keys(flat, 21)
    q = q[q != np.uint64(0xFFFFFFFFFFFFFFFF)]
    q = np.concatenate([q, _query_keys(np.random.default_rng(4), q, 21)])
    plain = hf.kmer_search(g, q, 21)
    for leaves, steps in ((4096, 300), (1, 0)):  # narrow windows; wide ones with a third level
        lsi = hf.LearnedSeedIndex(g, k=21, leaves=leaves, hidden=8, train_steps=steps, lookup=path)
        assert np.array_equal(lsi.seed(q), plain)


@pytest.mark.parametrize("path", ["global", "lds"])
def test_lsi_exact_beyond_three_fence_levels(path):
    """A 3 Mbp text with GC 0.30, one leaf and the untrained stage 1: the
    largest error bound is above 69,631 keys, where a fixed set of fence
    levels (a top level plus the 512- and 32-key ones) lost present keys; the
    build adds levels until the widest window fits, and every key of the
    reads, every mutated key and random keys match plain exact search."""
    rng = np.random.default_rng(12)
    text, bounds = O.gen_reference(80, 3_000_000, 3, gc=0.3)
    o = O.OracleBidir.build(text=text, bounds=bounds, sample_rate=32, threads=8)
    g = hf.BidirFmIndex.from_bytes(o.to_bytes())
    lsi = hf.LearnedSeedIndex(g, k=21, leaves=1, hidden=8, train_steps=0, lookup=path)
    info = lsi.info()
    assert info["max_err"] >= 69_632
    assert info["fence_levels"] == levels_for(info["max_err"]) >= 4
    flat, _ = O.gen_reads(9, text, bounds, 3000, 150, sub=0.01, both_strands=True)
    q = hf.kmer_keys(flat, 21)
    q = q[q != np.uint64(0xFFFFFFFFFFFFFFFF)]
    q = np.concatenate([q, _query_keys(rng, q, 21), np.array([0, 4 ** 21 - 1, q.min(), q.max()], np.uint64)])
    got = lsi.seed(q)
    assert np.array_equal(got, hf.kmer_search(g, q, 21))
    assert got[:, 3].sum() > len(q) // 3


def test_lsi_gradient_matches_oracle():
    rng = np.random.default_rng(11)
    H = 16
    params = {"w1": rng.normal(size=H), "b1": rng.normal(size=H) * 0.3, "w2": rng.normal(size=H) * 0.2,
              "b2": np.array(0.1)}
    x = rng.random(5000)
    y = np.sqrt(x) + 0.01 * rng.normal(size=5000)
    flat = np.concatenate([params["w1"], params["b1"], params["w2"], [params["b2"]]])
    loss, g = hf.lsi_gradient(flat.astype(np.float32), H, x.astype(np.float32), y.astype(np.float32))
    p32 = {kk: np.asarray(v, np.float32).astype(np.float64) for kk, v in params.items()}
    x32, y32 = x.astype(np.float32).astype(np.float64), y.astype(np.float32).astype(np.float64)
    ref = L.mlp_grad(p32, x32, y32)
    ref_flat = np.concatenate([ref["w1"], ref["b1"], ref["w2"], [float(ref["b2"])]])
    assert abs(loss - L.mlp_loss(p32, x32, y32)) <= 1e-4 * max(1.0, abs(loss))
    scale = np.abs(ref_flat).max()
    assert np.abs(g - ref_flat).max() <= 2e-4 * scale


def test_trained_stage1_beats_untrained():
    """What training gains.  Leaves are k-mer prefixes; inside a leaf the
    keys follow the shape of the global key CDF, which the trained stage 1
    learns.  On a 2 Mbp text with 16 leaves (about 120k keys each) the
    trained model's error windows are far narrower than the untrained f(x) =
    x (one straight line per leaf), the bounded search needs no more fence
    levels and loads fewer bytes per key, and both stay exact."""
    text, o, g = _bidir(91, 2_000_000, repeat_frac=0.3, n_units=40, unit_min=300, unit_max=3000)
    base = hf.LearnedSeedIndex(g, k=21, leaves=16, hidden=16, train_steps=0)
    trained = hf.LearnedSeedIndex(g, k=21, leaves=16, hidden=16, train_steps=400)
    bi, ti = base.info(), trained.info()
    assert ti["num_keys"] == bi["num_keys"]
    assert ti["mse_final"] < 0.5 * bi["mse_final"]
    assert ti["mean_window"] < 0.6 * bi["mean_window"]
    assert ti["max_err"] < bi["max_err"]
    assert ti["fence_levels"] <= bi["fence_levels"]
    flat, _ = O.gen_reads(8, text, o.fwd.seq_boundaries, 3000, 150, sub=0.01)
    q = hf.kmer_keys(flat, 21)
    q = q[q != np.uint64(0xFFFFFFFFFFFFFFFF)][:50000]
    plain = hf.kmer_search(g, q, 21)
    sb, st = {}, {}
    assert np.array_equal(base.seed(q, sb), plain)
    assert np.array_equal(trained.seed(q, st), plain)
    assert st["probes"] < sb["probes"]


def test_learned_index_gains_over_the_methods_without_it():
    """What the trained model gains: training lowers the stage-1 loss (the
    in-leaf CDF, about 70 keys per leaf here, so a coarse target) below the
    untrained f(x) = x, and a lookup needs far fewer dependent key probes
    (bounded search in the predicted window) than a binary search of the whole
    key array loads, or than the k exact left extensions of plain search."""
    rng = np.random.default_rng(2)
    text, o, g = _bidir(91, 300_000, repeat_frac=0.45, n_units=40, unit_min=300, unit_max=3000)
    base = hf.LearnedSeedIndex(g, k=21, leaves=4096, hidden=16, train_steps=0)
    trained = hf.LearnedSeedIndex(g, k=21, leaves=4096, hidden=16, train_steps=400)
    bi, ti = base.info(), trained.info()
    assert ti["num_keys"] == bi["num_keys"]
    assert ti["mse_final"] < bi["mse_final"]
    flat, _ = O.gen_reads(8, text, o.fwd.seq_boundaries, 2000, 150, sub=0.01)
    q = hf.kmer_keys(flat, 21)
    q = q[q != np.uint64(0xFFFFFFFFFFFFFFFF)][:20000]
    st = {}
    got = trained.seed(q, st)
    assert np.array_equal(got, hf.kmer_search(g, q, 21))
    loads = st["probes"] / len(q)  # 64-byte units loaded per lookup (fence and key rounds)
    assert loads < 0.5 * np.log2(ti["num_keys"])  # vs a binary search of all keys
    assert loads < 21 / 2                          # vs 21 dependent exact extensions
