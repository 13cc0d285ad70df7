"""GPU parity: batched count / locate through the C ABI vs the oracle."""
import numpy as np
import pytest

import haystackfm_amd as hf
from oracle import oracle as O
from tests.helpers import random_text

pytestmark = pytest.mark.gpu


def _pair(seqs=None, **kw):
    o = O.OracleIndex.build(seqs, **kw)
    return o, hf.FmIndex.from_bytes(o.to_bytes())


def _check_batch(o, g, queries):
    offs, hits = g.locate_batch(queries)
    if isinstance(queries, tuple):
        eo, eh = o.locate_batch(qflat=queries[0], qoff=queries[1], threads=4)
    else:
        eo, eh = o.locate_batch(queries, threads=4)
    assert np.array_equal(offs, eo)
    # identical order too: ascending SA rows, as FmIndex::locate returns them
    assert np.array_equal(hits, eh)
    assert np.array_equal(g.count_gpu(queries), np.diff(eo).astype(np.uint32))
    return offs, hits


def test_known_answers(golden):
    for text, pat, exp in golden["count"]["cases"]:
        o, g = _pair([text], sample_rate=1)
        assert int(g.count_gpu([pat])[0]) == exp
        _check_batch(o, g, [pat])
    for text, pat, exp in golden["locate"]["cases"]:
        o, g = _pair([text], sample_rate=1)
        assert sorted(p for _, p in g.locate_gpu([pat])[0]) == exp


def test_iupac_parity_cases(golden):
    for seqs, q in golden["iupac_locate"]["cases"]:
        o, g = _pair(seqs, sample_rate=1)
        text, bounds = O.concatenate(seqs)
        want = sorted(O.ref_map(p, bounds) for p in O.naive_locate(text, q))
        got = g.locate_gpu([q])[0]
        assert sorted((g.seq_headers.index(h), p) for h, p in got) == want


def test_multi_sequence_rates(golden):
    gm = golden["multi_sequence_rates"]
    for rate in gm["rates"]:
        o, g = _pair(gm["seqs"], sample_rate=rate)
        _check_batch(o, g, gm["patterns"])


def test_locate_many_with_branching(golden):
    gm = golden["locate_many"]
    o, g = _pair(gm["seqs"], sample_rate=3)
    _check_batch(o, g, gm["queries"] + [""])


@pytest.mark.parametrize("tag", [0, 1])
@pytest.mark.parametrize("lookup", [0, 4])
def test_random_iupac(tag, lookup):
    rng = np.random.default_rng(31 + tag + lookup)
    seqs = [O.decode(random_text(rng, int(rng.integers(200, 3000)), iupac_frac=0.03)) for _ in range(3)]
    o, g = _pair(seqs, sample_rate=8, lookup_depth=lookup, tag=tag)
    qs = []
    text, _ = O.concatenate(seqs)
    for _ in range(600):
        m = int(rng.integers(0, 14))
        if rng.random() < 0.6 and len(text) > m + 1:
            st = int(rng.integers(0, len(text) - m))
            q = text[st:st + m].copy()
            q[q == 0] = 1
        else:
            q = random_text(rng, m, alphabet="ACGTN", iupac_frac=0.15)
        qs.append(q)
    _check_batch(o, g, qs)


def test_wave_fallback_path_exact():
    """Queries of ambiguity codes over an IUPAC-rich text overflow the register
    fast path; the wave path must still return every hit."""
    rng = np.random.default_rng(77)
    seqs = [O.decode(random_text(rng, 20000, iupac_frac=0.2))]
    o, g = _pair(seqs, sample_rate=16)
    qs = [random_text(rng, int(rng.integers(1, 9)), alphabet="NRYSWKMBDHV") for _ in range(300)]
    qs += [np.full(6, 5, np.uint8), np.full(2, 5, np.uint8)]
    _check_batch(o, g, qs)
    st = g.last_stats()
    assert st["fallback_reads"] > 0
    assert st["guard_hits"] == 0


def test_empty_batch_and_invalid_bytes():
    o, g = _pair(["ACGTACGTAC"], sample_rate=2)
    offs, hits = g.locate_batch([])
    assert len(offs) == 1 and offs[0] == 0 and len(hits) == 0
    # bytes outside the alphabet match nothing; $ never matches
    _check_batch(o, g, [np.array([0], np.uint8), np.array([200, 1], np.uint8), np.array([16], np.uint8)])


def test_config1_shape_small():
    """The 1 Mbp configuration's shape at 200 kbp: ACGT + 0.1% N, rate 32,
    32-mers at random offsets with 1% substitutions."""
    text, bounds = O.gen_reference(1, 200_000, 1, gc=0.5, iupac_frac=0.001, iupac_codes=[5])
    o = O.OracleIndex.build(text=text, bounds=bounds, sample_rate=32, threads=4)
    g = hf.FmIndex.from_bytes(o.to_bytes())
    flat, off = O.gen_reads(2, text, bounds, 4000, 32, sub=0.01)
    _check_batch(o, g, (flat, off))


def test_lookup12_with_iupac_reference():
    """The 100 Mbp configuration's shape at 1 Mbp: 16 symbols, lookup depth 12,
    reads from both strands."""
    text, bounds = O.gen_reference(5, 1_000_000, 8, gc=0.41, repeat_frac=0.2, n_units=20,
                                   unit_min=300, unit_max=3000, iupac_frac=0.005)
    o = O.OracleIndex.build(text=text, bounds=bounds, sample_rate=32, lookup_depth=12, threads=8)
    g = hf.FmIndex.from_bytes(o.to_bytes())
    flat, off = O.gen_reads(6, text, bounds, 20000, 150, sub=0.01, both_strands=True)
    offs, hits = _check_batch(o, g, (flat, off))
    assert int(offs[-1]) > 0
    assert g.last_stats()["guard_hits"] == 0


def test_second_and_workgroup_levels():
    """Reads the register kernel gives up on (sets past 4 intervals) run the
    16-lanes-per-read pass and, past 16 intervals, the workgroup kernel;
    their hits equal the oracle's, including reads whose final set is
    published through the fallback interval list."""
    text, bounds = O.gen_reference(13, 300_000, 3, gc=0.41, repeat_frac=0.3, n_units=10, unit_min=300,
                                   unit_max=1000, iupac_frac=0.02)
    o = O.OracleIndex.build(text=text, bounds=bounds, sample_rate=16, lookup_depth=8, threads=8)
    g = hf.FmIndex.from_bytes(o.to_bytes())
    flat, off = O.gen_reads(14, text, bounds, 5000, 40, sub=0.01, both_strands=True)
    _check_batch(o, g, (flat, off))
    assert g.last_stats()["fallback_reads"] > 0


def test_second_register_pass_with_seed_records():
    """A text above 2^20 positions with 2% ambiguity codes and lookup depth 12:
    reads leave the first register pass at ambiguity codes near their end and
    the second pass seeds the unseedable ones from the packed 12-mer records
    (several intervals per record included); every hit equals the oracle's."""
    rng = np.random.default_rng(15)
    text, bounds = O.gen_reference(16, 1_300_000, 4, gc=0.41, repeat_frac=0.3, n_units=20, unit_min=300,
                                   unit_max=1500, iupac_frac=0.02)
    o = O.OracleIndex.build(text=text, bounds=bounds, sample_rate=32, lookup_depth=12, threads=8)
    g = hf.FmIndex.from_bytes(o.to_bytes())
    flat, off = O.gen_reads(17, text, bounds, 20000, 60, sub=0.005, both_strands=True)
    _check_batch(o, g, (flat, off))
    assert g.last_stats()["fallback_reads"] > 0


def test_locate_positions_match_the_text():
    """locate_positions(_many) (query.rs:46-208): absolute offsets into the
    concatenated text, the rows of locate_batch in the same order, each a
    compatible occurrence (brute force over the text)."""
    rng = np.random.default_rng(41)
    seqs = [O.decode(random_text(rng, int(n), iupac_frac=0.02)) for n in (3000, 5000, 800)]
    o = O.OracleIndex.build(seqs, sample_rate=4)
    g = hf.FmIndex.from_bytes(o.to_bytes())
    text, bounds = O.concatenate(seqs)
    qs = [text[int(s):int(s) + int(m)].copy() for s, m in zip(rng.integers(0, len(text) - 30, 300),
                                                             rng.integers(1, 25, 300))]
    qs = [np.where(q == 0, 1, q).astype(np.uint8) for q in qs] + [np.zeros(0, np.uint8)]
    offs, pos = g.locate_positions_batch(qs)
    lo, hits = g.locate_batch(qs)
    assert np.array_equal(offs, lo)
    starts = np.concatenate([[0], np.asarray(bounds, np.int64)[:-1]])
    assert np.array_equal(pos.astype(np.int64), starts[hits[:, 0].astype(np.int64)] + hits[:, 1])
    for i, q in enumerate(qs[:120]):
        assert sorted(pos[int(offs[i]):int(offs[i + 1])].tolist()) == sorted(O.naive_locate(text, q)), i
    assert [p.tolist() for p in g.locate_positions_many(qs[:5])] == [
        pos[int(offs[i]):int(offs[i + 1])].tolist() for i in range(5)]
