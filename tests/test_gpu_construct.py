"""GPU index construction (hsfm_build_index_device): the byte stream of
FmIndex::build_cpu + to_bytes and BidirFmIndex::build_cpu + to_bytes, equal
byte for byte to the specification-derived streams, to the oracle's builder
and to the host builder, on texts with several contigs, ambiguity codes, long
repeats (many doubling rounds), both Occ encodings, both alphabets, sample
rates and lookup depths."""
import numpy as np
import pytest

import haystackfm_amd as hf
from workloads import index_build as IB
from oracle import oracle as O
from tests.helpers import random_text
from tests.test_index_bytes_spec import CASES, SpecIndex, _codes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", range(len(CASES)))
def test_matches_the_spec_streams(case):
    seqs, headers, rate, depth, enc, tag = CASES[case]
    spec = SpecIndex([_codes(s) for s in seqs], headers, rate, depth, enc, tag)
    text = np.array(spec.text, np.uint8)
    got = hf.build_index_bytes_gpu(text, spec.bounds, headers=headers, sample_rate=rate, lookup_depth=depth,
                                   encoding=enc, tag=tag)
    assert got == spec.to_bytes()
    rev = SpecIndex([list(reversed(spec.text[:-1]))], None, rate, 0, enc, tag, concat=False)
    f, r = spec.to_bytes(), rev.to_bytes()
    got = hf.build_index_bytes_gpu(text, spec.bounds, headers=headers, sample_rate=rate, lookup_depth=depth,
                                   encoding=enc, tag=tag, bidir=True)
    assert got == len(f).to_bytes(4, "little") + f + r


@pytest.mark.parametrize("seed", range(6))
def test_matches_the_oracle_builder(seed):
    rng = np.random.default_rng(seed)
    nseq = int(rng.integers(1, 6))
    seqs = [O.decode(random_text(rng, int(rng.integers(0, 5000)), iupac_frac=[0, 0.01, 0.2][seed % 3]))
            for _ in range(nseq)]
    text, bounds = O.concatenate(seqs)
    rate = [1, 2, 5, 32, 64, 7][seed]
    depth = [0, 1, 4, 8, 3, 6][seed]
    enc, tag = seed % 2, (seed // 2) % 2
    o = O.OracleBidir.build(text=text, bounds=bounds, sample_rate=rate, lookup_depth=depth, encoding=enc, tag=tag,
                            threads=4)
    got = hf.build_index_bytes_gpu(text, bounds, sample_rate=rate, lookup_depth=depth, encoding=enc, tag=tag,
                                   bidir=True)
    assert got == o.to_bytes()


def test_long_repeats_and_runs():
    """Runs of one code and tandem copies: suffixes share prefixes thousands
    of codes long, so prefix doubling runs many rounds."""
    rng = np.random.default_rng(9)
    unit = O.decode(random_text(rng, 700))
    seqs = ["A" * 20000, unit * 30, "AC" * 5000 + O.decode(random_text(rng, 3000)), "N" * 3000 + "ACGT"]
    text, bounds = O.concatenate(seqs)
    o = O.OracleBidir.build(text=text, bounds=bounds, sample_rate=16, lookup_depth=5, threads=4)
    assert hf.build_index_bytes_gpu(text, bounds, sample_rate=16, lookup_depth=5, bidir=True) == o.to_bytes()


def test_matches_the_host_builder_on_a_repeat_rich_reference():
    text, bounds = IB.synth_reference(0xBEEF, 3_000_000, 5, gc=0.41, repeat_frac=0.3, n_units=60, unit_min=300,
                                      unit_max=6000, div_min=0.01, div_max=0.1, iupac_frac=0.005)
    want = IB.build_bidir_bytes(text, bounds, sample_rate=32, lookup_depth=10, threads=8)
    got = hf.build_index_bytes_gpu(text, bounds, sample_rate=32, lookup_depth=10, bidir=True)
    assert got == want
    # and the device-built index answers queries as the uploaded host-built one does
    g = hf.BidirFmIndex.from_bytes(got)
    flat, off = IB.synth_reads(3, text, bounds, 2000, 150, sub=0.01, both_strands=True)
    a = g.fwd.locate_batch((flat, off))
    assert int(a[0][-1]) > 100


def test_rejects_malformed_input():
    text = np.array([1, 2, 3, 0], np.uint8)
    with pytest.raises(hf.GpuError, match="last boundary"):
        hf.build_index_bytes_gpu(text, [3])
    with pytest.raises(hf.GpuError, match="sentinel"):
        hf.build_index_bytes_gpu(np.array([1, 2, 3, 4], np.uint8), [4])
    with pytest.raises(hf.GpuError, match="code"):
        hf.build_index_bytes_gpu(np.array([1, 20, 3, 0], np.uint8), [4])
