"""Learned seeds inside the SMEM/MEM restarts (hsfm_mem_seed_mode): every
seeding mode -- the packed tables, plain backward search from the full
interval, learned 21-mer seeds with either fallback -- gives the oracle's
MemHit lists, on IUPAC texts (where a window's set also holds the intervals of
ambiguity-code windows compatible with it, stored as entries of their own) and
ExactDna (where ambiguity codes cut windows, leaving short entries), with
min_len below, at and above k.  At min_len = k a restart of an all-base read
never falls back to plain search: the lookup gives the set or L(j)."""
import numpy as np
import pytest

import haystackfm_amd as hf
from oracle import oracle as O
from tests.helpers import random_text
from tests.test_gpu_mem import check as check_mems

pytestmark = pytest.mark.gpu

MODES = [hf.HSFM_SEED_TABLES, hf.HSFM_SEED_PLAIN, hf.HSFM_SEED_LEARNED, hf.HSFM_SEED_LEARNED_TABLES]


def _raw(g, reads, min_len, bounds, mode):
    return g.find_mems_batch(reads, min_len, mode, bounds, 64)


@pytest.mark.parametrize("tag", [0, 1])
def test_every_seed_mode_matches_the_oracle(tag):
    text, bounds = O.gen_reference(70 + tag, 300_000, 3, gc=0.41, repeat_frac=0.35, n_units=30, unit_min=300,
                                   unit_max=3000, iupac_frac=0.004)
    o = O.OracleBidir.build(text=text, bounds=bounds, sample_rate=16, tag=tag, threads=8)
    g = hf.BidirFmIndex.from_bytes(o.to_bytes())
    lsi = hf.LearnedSeedIndex(g, k=21, leaves=1024, hidden=16, train_steps=200)
    info = lsi.info()
    assert info["mem_seeding"] == 1
    if tag == 0:
        assert info["variant_entries"] > 0  # ambiguity codes in the text expand into base strings
    else:
        assert info["short_entries"] > 0  # under ExactDna they cut windows short
    flat, off = O.gen_reads(71, text, bounds, 1500, 150, sub=0.01, indel=0.001, both_strands=True)
    reads = (flat, off)
    ref = {}
    for min_len in (19, 21, 30):
        for mode in MODES:
            g.set_seed_mode(mode, lsi if mode >= hf.HSFM_SEED_LEARNED else None)
            got = g.find_mems_batch(reads, min_len, hf.HSFM_MODE_SMEM, bounds, 64)
            if min_len not in ref:
                ref[min_len] = got
            else:
                for a, b in zip(got, ref[min_len]):
                    assert np.array_equal(a, b), (min_len, mode)
        # the learned mode against the oracle on a subset
        g.set_seed_mode(hf.HSFM_SEED_LEARNED, lsi)
        sub = [flat[off[i]:off[i + 1]] for i in range(0, 1500, 5)]
        check_mems(o, g, sub, min_len, cap=64)
    g.set_seed_mode(hf.HSFM_SEED_TABLES)


def test_learned_seeding_is_refused_when_sets_cannot_be_complete():
    """A run of 40 N: its windows have 4^21 expansions, more than the build
    stores per window, so the index cannot seed MEM restarts exactly and the
    learned modes are refused (the tables and plain search still work)."""
    rng = np.random.default_rng(3)
    seq = random_text(rng, 30_000)
    seq[10_000:10_040] = 5
    o = O.OracleBidir.build([O.decode(seq)], sample_rate=8)
    g = hf.BidirFmIndex.from_bytes(o.to_bytes())
    lsi = hf.LearnedSeedIndex(g, k=21, leaves=64, hidden=8, train_steps=50)
    assert lsi.info()["mem_seeding"] == 0
    with pytest.raises(hf.GpuError, match="cannot seed MEM restarts"):
        g.set_seed_mode(hf.HSFM_SEED_LEARNED, lsi)
    g.set_seed_mode(hf.HSFM_SEED_PLAIN)
    q = [seq[9_900:10_100], seq[20_000:20_150]]
    check_mems(o, g, q, 21)


def test_freeing_the_learned_index_restores_table_seeding():
    text, bounds = O.gen_reference(72, 60_000, 2, gc=0.41)
    o = O.OracleBidir.build(text=text, bounds=bounds, sample_rate=8)
    g = hf.BidirFmIndex.from_bytes(o.to_bytes())
    lsi = hf.LearnedSeedIndex(g, k=21, leaves=64, hidden=8, train_steps=50)
    flat, off = O.gen_reads(73, text, bounds, 200, 150, sub=0.01)
    g.set_seed_mode(hf.HSFM_SEED_LEARNED, lsi)
    a = g.find_mems_batch((flat, off), 21, hf.HSFM_MODE_MEM, bounds, 16)
    g._seed_lsi = None
    del lsi
    b = g.find_mems_batch((flat, off), 21, hf.HSFM_MODE_MEM, bounds, 16)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("tag", [0, 1])
def test_learned_restarts_never_fall_back_at_min_len_k(tag):
    """Reads of bases only, min_len = k: every restart window is all bases and
    j >= k, so the learned lookup answers every restart (its set, or L(j) from
    the longest occurring suffix) and none runs the unseeded search; the MEM
    lists equal plain search and the oracle.  Twelve contigs, ambiguity codes
    in the text (expanded under IupacDna, cutting windows under ExactDna)."""
    text, bounds = O.gen_reference(90 + tag, 200_000, 12, gc=0.41, repeat_frac=0.4, n_units=25, unit_min=200,
                                   unit_max=2000, iupac_frac=0.01)
    o = O.OracleBidir.build(text=text, bounds=bounds, sample_rate=16, tag=tag, threads=8)
    g = hf.BidirFmIndex.from_bytes(o.to_bytes())
    lsi = hf.LearnedSeedIndex(g, k=21, leaves=256, hidden=16, train_steps=200)
    info = lsi.info()
    assert info["mem_seeding"] == 1 and info["entries"] > info["num_keys"]
    flat, off = O.gen_reads(91, text, bounds, 2000, 150, sub=0.02, both_strands=True, n_random=200)
    rng = np.random.default_rng(tag)
    amb = (flat < 1) | (flat > 4)  # reads copy the text's ambiguity codes: make them bases
    flat[amb] = rng.integers(1, 5, int(amb.sum())).astype(np.uint8)
    reads = (flat, off)
    g.set_seed_mode(hf.HSFM_SEED_PLAIN)
    plain = g.find_mems_batch(reads, 21, hf.HSFM_MODE_SMEM, bounds, 64)
    g.set_seed_mode(hf.HSFM_SEED_LEARNED, lsi)
    got = g.find_mems_batch(reads, 21, hf.HSFM_MODE_SMEM, bounds, 64)
    st = g.last_stats()
    g.set_seed_mode(hf.HSFM_SEED_TABLES)
    for a, b in zip(got, plain):
        assert np.array_equal(a, b)
    assert st["diag"][3] == 0  # unseeded restarts
    sub = [flat[off[i]:off[i + 1]] for i in range(0, 2200, 11)]
    g.set_seed_mode(hf.HSFM_SEED_LEARNED, lsi)
    check_mems(o, g, sub, 21, cap=64)
    g.set_seed_mode(hf.HSFM_SEED_TABLES)


@pytest.mark.parametrize("leaves", [1 << 12, 1 << 16, 1 << 18])
def test_learned_restarts_exact_with_many_leaves(leaves):
    """Many small leaves (down to a key or two per leaf): a window above every
    key of its leaf has its lower bound one past the leaf, which the search
    range must still hold however far the leaf's model predicts past its end;
    the MEM lists equal plain search and the oracle."""
    text, bounds = O.gen_reference(95, 200_000, 6, gc=0.41, repeat_frac=0.4, n_units=25, unit_min=200,
                                   unit_max=2000, iupac_frac=0.005)
    o = O.OracleBidir.build(text=text, bounds=bounds, sample_rate=16, threads=8)
    g = hf.BidirFmIndex.from_bytes(o.to_bytes())
    lsi = hf.LearnedSeedIndex(g, k=21, leaves=leaves, hidden=16, train_steps=200)
    assert lsi.info()["mem_seeding"] == 1
    flat, off = O.gen_reads(96, text, bounds, 2000, 150, sub=0.02, both_strands=True, n_random=300)
    reads = (flat, off)
    g.set_seed_mode(hf.HSFM_SEED_PLAIN)
    plain = g.find_mems_batch(reads, 21, hf.HSFM_MODE_SMEM, bounds, 64)
    try:
        g.set_seed_mode(hf.HSFM_SEED_LEARNED, lsi)
        got = g.find_mems_batch(reads, 21, hf.HSFM_MODE_SMEM, bounds, 64)
        for a, b in zip(got, plain):
            assert np.array_equal(a, b)
        sub = [flat[off[i]:off[i + 1]] for i in range(0, 2300, 23)]
        check_mems(o, g, sub, 21, cap=64)
    finally:
        g.set_seed_mode(hf.HSFM_SEED_TABLES)
