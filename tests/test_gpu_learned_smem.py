"""Learned seeds inside the SMEM/MEM restarts (hsfm_mem_seed_mode): every
seeding mode -- the packed tables, plain backward search from the full
interval, learned 21-mer seeds with either fallback -- gives the oracle's
MemHit lists, on IUPAC texts (where k-mers compatible with an ambiguity-code
window are flagged and fall back) and ExactDna, with min_len below, at and
above k."""
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
        assert info["flagged"] > 0  # N and other codes in the text make some 21-mers ambiguous
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


def test_learned_seeding_is_refused_when_flags_cannot_be_complete():
    """A run of 40 N: its windows have 4^21 expansions, more than the flag
    pass enumerates, so the index cannot seed MEM restarts exactly and the
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
