"""GPU parity on the index variants the reference supports beside the default
(Bitplane Occ, IupacDna): the OneHot Occ encoding (src/occ/mod.rs:93-97) for
locate and SMEM/MEM, ExactDna on SMEM/MEM (src/alphabet.rs:282-305), and an
inconsistent index whose LF walks outrun the sample rate, which the call
reports as HSFM_EINDEX (the guard of src/gpu/locate.rs / locate_resolve.wgsl;
src/suffix_array/mod.rs:75-81 names the invariant) while every kernel exits
normally."""
import numpy as np
import pytest

import haystackfm_amd as hf
from oracle import oracle as O
from tests.helpers import random_text
from tests.test_gpu_locate import _check_batch
from tests.test_gpu_mem import check as check_mems
from tests.test_index_bytes_spec import SpecIndex

pytestmark = pytest.mark.gpu


def _queries(rng, text, count, maxlen, iupac=0.1):
    qs = []
    for _ in range(count):
        m = int(rng.integers(1, maxlen))
        if rng.random() < 0.7 and len(text) > m + 1:
            st = int(rng.integers(0, len(text) - m))
            q = text[st:st + m].copy()
            q[q == 0] = 1
        else:
            q = random_text(rng, m, alphabet="ACGTN", iupac_frac=iupac)
        qs.append(q)
    return qs


@pytest.mark.parametrize("tag", [0, 1])
def test_onehot_locate(tag):
    rng = np.random.default_rng(40 + tag)
    seqs = [O.decode(random_text(rng, int(rng.integers(500, 4000)), iupac_frac=0.03)) for _ in range(3)]
    o = O.OracleIndex.build(seqs, sample_rate=8, lookup_depth=4, encoding=1, tag=tag)
    g = hf.FmIndex.from_bytes(o.to_bytes())
    text, _ = O.concatenate(seqs)
    _check_batch(o, g, _queries(rng, text, 800, 16))


@pytest.mark.parametrize("encoding,tag", [(1, 0), (0, 1), (1, 1)])
def test_onehot_and_exact_alphabet_mems(encoding, tag):
    """SMEM and MEM lists, intervals and positions against the oracle on
    IUPAC-bearing texts: OneHot records, and the ExactDna alphabet (only
    equal codes match, so N in a read matches N in the text and nothing
    else)."""
    rng = np.random.default_rng(50 + 2 * encoding + tag)
    seqs = [O.decode(random_text(rng, int(rng.integers(3000, 9000)), iupac_frac=0.02)) for _ in range(3)]
    o = O.OracleBidir.build(seqs, sample_rate=8, encoding=encoding, tag=tag, threads=4)
    g = hf.BidirFmIndex.from_bytes(o.to_bytes())
    text, bounds = O.concatenate(seqs)
    flat, off = O.gen_reads(60 + tag, text, bounds, 300, 150, sub=0.02, both_strands=True)
    reads = [flat[off[i]:off[i + 1]] for i in range(len(off) - 1)]
    reads += _queries(rng, text, 100, 60, iupac=0.2)
    check_mems(o, g, reads, 12)
    check_mems(o, g, reads, 25, cap=3)


def _corrupt_index_bytes():
    """A 2-contig text, sample rate 8, with the samples of text positions 8
    and 96 dropped (their marker bits cleared, their values left out of
    sa_vals, the rank records consistent): walks from positions 9..15 and
    97..103 need more than 8 LF steps."""
    rng = np.random.default_rng(7)
    a = [int(c) for c in random_text(rng, 120)]
    b = [int(c) for c in random_text(rng, 90)]
    sp = SpecIndex([a, b], rate=8, drop_samples=(8, 96))
    return sp, sp.to_bytes()


def test_inconsistent_index_reports_eindex():
    sp, data = _corrupt_index_bytes()
    g = hf.FmIndex.from_bytes(data)
    text = np.array(sp.text, np.uint8)
    # a pattern occurring only at position 10 (inside the unsampled run)
    q = text[10:30].copy()
    with pytest.raises(hf.GpuError, match="exceeded the sample rate"):
        g.locate_batch([q])
    assert g.last_stats()["guard_hits"] > 0
    # patterns that resolve within the sample rate still work on the same index
    q_ok = text[40:60].copy()
    offs, hits = g.locate_batch([q_ok])
    assert int(offs[-1]) == 1 and tuple(hits[0]) == (0, 40)
    # the library stays usable: a valid index gives the oracle's answers
    o = O.OracleIndex.build(text=text, bounds=np.array(sp.bounds, np.uint32), sample_rate=8, naive_sa=True)
    good = hf.FmIndex.from_bytes(o.to_bytes())
    _check_batch(o, good, [q, q_ok])


def test_inconsistent_bidir_index_reports_eindex_on_mems():
    sp, fwd = _corrupt_index_bytes()
    rev = SpecIndex([list(reversed(sp.text[:-1]))], rate=8, concat=False).to_bytes()
    data = len(fwd).to_bytes(4, "little") + fwd + rev
    g = hf.BidirFmIndex.from_bytes(data)
    text = np.array(sp.text, np.uint8)
    q = text[12:42].copy()  # its one position, 12, walks 12 LF steps to the sample at 0
    with pytest.raises(hf.GpuError, match="exceeded the sample rate"):
        g.find_smems_gpu([q], 20, sp.bounds, 16)
    # no positions asked: no walk, no error
    got = g.find_smems_gpu([q], 20, [], 16)
    assert [(h.query_start, h.query_end, h.match_count) for h in got[0]] == [(0, 30, 1)]
