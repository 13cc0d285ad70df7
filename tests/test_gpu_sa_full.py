"""GPU checks of the device-built full suffix array that resolve reads: every
row's value equals the LF walk to the samples (FmIndex::resolve_sa,
src/fm_index/query.rs:282-291), so locate and MEM positions are unchanged."""
import numpy as np
import pytest

import haystackfm_amd as hf
from oracle import oracle as O
from tests.helpers import random_text

pytestmark = pytest.mark.gpu

SYMBOLS = ["A", "C", "G", "T", "N", "R", "Y", "S", "W", "K", "M", "B", "D", "H", "V"]


@pytest.mark.parametrize("rate", [1, 5, 32])
@pytest.mark.parametrize("tag", [0, 1])
def test_every_row_resolves_like_the_walk(monkeypatch, rate, tag):
    """Single-symbol queries cover every non-sentinel row, including the rows
    after the last sample of each sequence (reached from its sentinel row)."""
    rng = np.random.default_rng(100 * rate + tag)
    lens = [1, 2, 3, 31, 33, 64, 700] + [int(x) for x in rng.integers(1, 2500, 9)]
    seqs = [O.decode(random_text(rng, L, iupac_frac=0.02 if tag == 0 else 0.0)) for L in lens]
    o = O.OracleIndex.build(seqs, sample_rate=rate, tag=tag)
    g = hf.FmIndex.from_bytes(o.to_bytes())
    qs = SYMBOLS[:4] if tag else SYMBOLS
    offs, hits = g.locate_batch(qs)
    assert g.last_stats()["sa_full_rows"] == o.text_len
    eo, eh = o.locate_batch(qs, threads=4)
    assert np.array_equal(offs, eo) and np.array_equal(hits, eh)
    monkeypatch.setenv("HSFM_SA_FULL", "0")
    offs2, hits2 = g.locate_batch(qs)
    assert g.last_stats()["sa_full_rows"] == 0
    assert np.array_equal(offs, offs2) and np.array_equal(hits, hits2)


def test_mem_positions_with_and_without_full_sa(monkeypatch):
    rng = np.random.default_rng(5)
    text, bounds = O.gen_reference(6, 120_000, 5, gc=0.41, repeat_frac=0.3, n_units=10, unit_min=300,
                                   unit_max=1500, iupac_frac=0.005)
    o = O.OracleBidir.build(text=text, bounds=bounds, sample_rate=16, threads=8)
    g = hf.BidirFmIndex.from_bytes(o.to_bytes())
    flat, off = O.gen_reads(7, text, bounds, 300, 150, sub=0.02, both_strands=True)
    a = g.find_mems_batch((flat, off), 18, hf.HSFM_MODE_SMEM, bounds, 64)
    assert g.last_stats()["sa_full_rows"] == o.fwd.text_len
    monkeypatch.setenv("HSFM_SA_FULL", "0")
    b = g.find_mems_batch((flat, off), 18, hf.HSFM_MODE_SMEM, bounds, 64)
    assert g.last_stats()["sa_full_rows"] == 0
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
