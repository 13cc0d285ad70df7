"""The product's host index builder must emit the same bytes as the oracle's
restatement of build_cpu + to_bytes (pins both constructions)."""
import numpy as np
import pytest

from workloads import index_build as IB
from oracle import oracle as O
from tests.helpers import random_text


@pytest.mark.parametrize("enc", [0, 1])
def test_index_bytes_identical(enc):
    rng = np.random.default_rng(50 + enc)
    for trial in range(6):
        seqs = [O.decode(random_text(rng, int(rng.integers(1, 900)), iupac_frac=0.03 * trial))
                for _ in range(int(rng.integers(1, 5)))]
        text, bounds = O.concatenate(seqs)
        rate = int(rng.choice([1, 3, 32]))
        depth = int(rng.choice([0, 2, 5]))
        tag = int(rng.integers(0, 2))
        a = IB.build_index_bytes(text, bounds, sample_rate=rate, lookup_depth=depth, encoding=enc, tag=tag, threads=3)
        b = O.OracleIndex.build(text=text, bounds=bounds, sample_rate=rate, lookup_depth=depth, encoding=enc,
                                tag=tag).to_bytes()
        assert a == b


def test_bidir_bytes_identical():
    text, bounds = O.gen_reference(3, 50_000, 3, gc=0.41, repeat_frac=0.2, n_units=5, unit_min=300,
                                   unit_max=900, iupac_frac=0.005)
    a = IB.build_bidir_bytes(text, bounds, headers=["a", "b", "c"], sample_rate=32, lookup_depth=6, threads=4)
    b = O.OracleBidir.build(text=text, bounds=bounds, headers=["a", "b", "c"], sample_rate=32,
                            lookup_depth=6, threads=4).to_bytes()
    assert a == b


def test_synth_shapes():
    text, bounds = IB.synth_reference(1, 100_000, 8, gc=0.41, repeat_frac=0.2, n_units=20, unit_min=300,
                                      unit_max=2000, div_min=0.05, div_max=0.15, iupac_frac=0.005)
    assert len(bounds) == 8 and int(bounds[-1]) == len(text) == 100_008
    assert (text[bounds - 1] == 0).all() and int((text == 0).sum()) == 8
    body = text[text != 0]
    gc = float(np.isin(body, [2, 3]).mean())
    assert 0.38 < gc < 0.44
    amb = float((body >= 5).mean())
    assert 0.003 < amb < 0.007
    flat, off = IB.synth_reads(2, text, bounds, 1000, 150, sub=0.01, both_strands=True)
    assert len(off) == 1001 and int(off[-1]) == 150_000
    flat, off = IB.synth_reads(3, text, bounds, 500, 150, sigma=1.0, len_min=32, len_max=10_000, sub=0.015,
                               shuffle=True)
    lens = np.diff(off)
    assert lens.min() >= 32 and lens.max() <= 10_000 and 100 < np.median(lens) < 220
