"""GPU parity: find_smems_gpu / find_mems_gpu through the C ABI vs the oracle."""
import numpy as np
import pytest

import haystackfm_amd as hf
from oracle import oracle as O
from tests.helpers import random_text, smem_random_cases

pytestmark = pytest.mark.gpu


def _pair(seqs=None, **kw):
    o = O.OracleBidir.build(seqs, **kw)
    return o, hf.BidirFmIndex.from_bytes(o.to_bytes())


def expected_hits(o, mems, ref_bounds, cap):
    """MemHit list the device must return for the oracle's Mems: one record
    per SA interval, intervals of a MEM sorted by start, positions = the first
    min(size, cap) rows of each interval (bidir_index.rs:307-359, 452-468)."""
    out = []
    for m in mems:
        for lo, hi, _, _ in sorted(m.ivs.tolist()):
            size = hi - lo
            k = min(size, cap) if len(ref_bounds) else 0
            if k:
                pos = o.fwd.resolve_rows(np.arange(lo, lo + k, dtype=np.uint32))
                plist = [O.ref_map(int(p), ref_bounds) for p in pos]
            else:
                plist = []
            out.append((m.query_start, m.query_end, size, size > cap, plist))
    return out


def _as_tuples(hits):
    return [(h.query_start, h.query_end, h.match_count, h.truncated, h.positions) for h in hits]


def check(o, g, queries, min_len, ref_bounds=None, cap=1024, modes=(0, 1)):
    rb = o.fwd.seq_boundaries if ref_bounds is None else np.asarray(ref_bounds, np.uint32)
    for mode in modes:
        exp = o.find_mems_batch(queries, min_len=min_len, mode=mode, threads=4)
        fn = g.find_smems_gpu if mode == 0 else g.find_mems_gpu
        got = fn(queries, min_len, rb, cap)
        assert len(got) == len(exp)
        for qi, (gh, em) in enumerate(zip(got, exp)):
            assert _as_tuples(gh) == expected_hits(o, em, rb, cap), (mode, qi)
        st = g.last_stats()
        assert st["guard_hits"] == 0


def test_known_answers(golden):
    for case in golden["smem"]["exact"]:
        o, g = _pair(case["ref"], sample_rate=1)
        check(o, g, [case["query"]], case["min_len"])


def test_brute_force_cases(golden):
    for ref, q in golden["smem"]["brute_force"]:
        o, g = _pair([ref], sample_rate=1)
        check(o, g, [q], 1)
        text, _ = O.concatenate([ref])
        got = g.find_smems_gpu([q], 1, o.fwd.seq_boundaries)[0]
        spans = sorted({(h.query_start, h.query_end) for h in got})
        assert spans == [(s, e) for s, e, _ in O.brute_mems(text, q, 1)]


def test_regressions(golden):
    gs = golden["smem"]
    q = gs["regression_query"]
    for reg in gs["regressions"]:
        seqs = [q[a:e] for a, e, _ in reg["refs"]]
        o, g = _pair(seqs, headers=[h for _, _, h in reg["refs"]], sample_rate=1)
        check(o, g, [q], reg["min_len"])
        got = g.find_smems_gpu([q], reg["min_len"], o.fwd.seq_boundaries)[0]
        assert {r for h in got for r, _ in h.positions} == {0, 1}


def test_randomized_lcg_cases():
    cases = smem_random_cases()
    for seqs, names, full in cases[:40]:
        o, g = _pair(seqs, headers=names, sample_rate=1)
        check(o, g, [full], 15)


@pytest.mark.parametrize("seed", range(3))
def test_random_iupac_batches(seed):
    rng = np.random.default_rng(300 + seed)
    seqs = [O.decode(random_text(rng, int(rng.integers(500, 4000)), iupac_frac=0.02)) for _ in range(3)]
    o, g = _pair(seqs, sample_rate=int(rng.choice([1, 8, 32])))
    text, _ = O.concatenate(seqs)
    qs = []
    for _ in range(120):
        m = int(rng.integers(0, 60))
        if rng.random() < 0.7 and len(text) > m + 1:
            st = int(rng.integers(0, len(text) - m))
            q = text[st:st + m].copy()
            q[q == 0] = 1
            flip = rng.random(m) < 0.05
            q[flip] = random_text(rng, int(flip.sum()), alphabet="ACGTN")
        else:
            q = random_text(rng, m, alphabet="ACGT")
        qs.append(q)
    check(o, g, qs, int(rng.integers(1, 12)))


def test_truncation_and_no_boundaries():
    o, g = _pair(["ACGT" * 300, "ACGTTACG" * 40], sample_rate=4)
    qs = ["ACGTACGT", "ACG", "TTACGACGT"]
    check(o, g, qs, 3, cap=7)
    # empty ref_boundaries: counts and truncation flags only, no positions
    got = g.find_mems_gpu(qs, 3, [], 7)
    exp = o.find_mems_batch(qs, min_len=3, mode=1)
    for gh, em in zip(got, exp):
        assert _as_tuples(gh) == expected_hits(o, em, [], 7)
        assert all(h.positions == [] for h in gh)


def test_user_ref_boundaries_differ_from_index():
    """ref_map uses the caller's boundaries, not the index's (ref_map.wgsl:31-59)."""
    o, g = _pair(["ACGTTGCAACGGT", "TTGACCAGTAGG"], sample_rate=2)
    rb = [5, 14, 30]
    check(o, g, ["ACGTTG", "CAGTA"], 2, ref_bounds=rb)


def test_empty_inputs():
    o, g = _pair(["ACGTACGT"], sample_rate=1)
    assert g.find_smems_gpu([], 3, o.fwd.seq_boundaries) == []
    got = g.find_smems_gpu(["", "AC", "ACGT"], 0, o.fwd.seq_boundaries)
    assert got == [[], [], []]
    check(o, g, ["", np.array([0, 0], np.uint8), np.array([250, 1, 2], np.uint8)], 1)


def test_iupac_heavy_sets_use_capacity_retries():
    """An IUPAC-rich text and N-rich queries make interval sets large; every
    read must still come back exactly."""
    rng = np.random.default_rng(5)
    seqs = [O.decode(random_text(rng, 30000, iupac_frac=0.25))]
    o, g = _pair(seqs, sample_rate=16)
    qs = [random_text(rng, int(rng.integers(5, 25)), alphabet="ACGTN", iupac_frac=0.3) for _ in range(60)]
    check(o, g, qs, 2, modes=(0,))


def _mem_reads(rng, text, nreads, length, sub, amb):
    qs = []
    for _ in range(nreads):
        st = int(rng.integers(0, len(text) - length))
        q = text[st:st + length].copy()
        q[q == 0] = 1
        flip = rng.random(length) < sub
        q[flip] = random_text(rng, int(flip.sum()), alphabet="ACGT")
        ambm = rng.random(length) < amb
        q[ambm] = random_text(rng, int(ambm.sum()), alphabet="NRYSWKMBDHV")
        if rng.random() < 0.5:
            q = q[::-1].copy()  # unrelated orientation: mostly short matches
        qs.append(q)
    return qs


@pytest.mark.parametrize("seed", range(2))
def test_thread_path_seed_table(seed):
    """min_len >= 12 runs the persistent kernel seeded from the packed 12-mer
    records, including ambiguity codes in the seed windows."""
    rng = np.random.default_rng(700 + seed)
    text, bounds = O.gen_reference(40 + seed, 300_000, 4, gc=0.41, repeat_frac=0.2, n_units=30, unit_min=300,
                                   unit_max=2000, iupac_frac=0.005)
    o = O.OracleBidir.build(text=text, bounds=bounds, sample_rate=32, lookup_depth=10, threads=8)
    g = hf.BidirFmIndex.from_bytes(o.to_bytes())
    qs = _mem_reads(rng, text, 400, 150, 0.02, 0.01)
    qs += [random_text(rng, 150, alphabet="ACGT") for _ in range(40)]
    check(o, g, qs, 12 + 3 * seed)


@pytest.mark.parametrize("other", ["wave"])
def test_persistent_and_other_paths_agree(monkeypatch, other):
    """The persistent kernel (with its deferred items on the workgroup
    kernel) and the workgroup kernel for every read give identical records."""
    rng = np.random.default_rng(11)
    text, bounds = O.gen_reference(9, 200_000, 3, gc=0.41, repeat_frac=0.2, n_units=20, unit_min=300,
                                   unit_max=1500, iupac_frac=0.01)
    o = O.OracleBidir.build(text=text, bounds=bounds, sample_rate=16, threads=8)
    g = hf.BidirFmIndex.from_bytes(o.to_bytes())
    qs = _mem_reads(rng, text, 500, 150, 0.02, 0.02)
    a = g.find_mems_batch(qs, 18, 0, bounds, 64)
    monkeypatch.setenv("HSFM_MEM_PATH", other)
    b = g.find_mems_batch(qs, 18, 0, bounds, 64)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("iupac", [0.005, 0.04])
def test_packed_seed_tables_12_14_15(monkeypatch, iupac):
    """Restarts seeded from the packed 12-, 14- and 15-mer records (one
    interval inline, several through the interval list) give the oracle's
    MEMs; an IUPAC-rich text makes many 14- and 15-mers own several compatible
    intervals.  min_len 15 runs the 15-mer setting on 14-mers (seeds are
    shorter than min_len)."""
    rng = np.random.default_rng(57)
    text, bounds = O.gen_reference(58, 300_000, 4, gc=0.41, repeat_frac=0.2, n_units=30, unit_min=300,
                                   unit_max=2000, iupac_frac=iupac)
    o = O.OracleBidir.build(text=text, bounds=bounds, sample_rate=32, threads=8)
    g = hf.BidirFmIndex.from_bytes(o.to_bytes())
    qs = _mem_reads(rng, text, 300, 150, 0.02, 0.01)
    qs += [random_text(rng, 150, alphabet="ACGT") for _ in range(20)]
    for F in ("12", "14", "15"):
        monkeypatch.setenv("HSFM_SEED_F", F)
        for min_len in (15, 19):
            check(o, g, qs, min_len, modes=(0,))


@pytest.mark.parametrize("seed_f", ["12", "14", "15"])
def test_ambiguity_clusters_in_seed_windows(monkeypatch, seed_f):
    """Reads with 4-7 ambiguity codes inside one seed window: the ring kernel
    hands them to the workgroup kernel, which expands up to five codes over its
    threads (one packed 12-, 14- or 15-mer record each) and restarts unseeded
    beyond that; all stay exact."""
    monkeypatch.setenv("HSFM_SEED_F", seed_f)
    rng = np.random.default_rng(77)
    text, bounds = O.gen_reference(78, 200_000, 3, gc=0.41, repeat_frac=0.2, n_units=20, unit_min=300,
                                   unit_max=1500, iupac_frac=0.005)
    o = O.OracleBidir.build(text=text, bounds=bounds, sample_rate=32, threads=8)
    g = hf.BidirFmIndex.from_bytes(o.to_bytes())
    qs = []
    for i in range(120):
        st = int(rng.integers(0, len(text) - 150))
        q = text[st:st + 150].copy()
        q[q == 0] = 1
        c = int(rng.integers(10, 130))
        namb = 4 + i % 4
        at = c + rng.choice(12, size=namb, replace=False)
        q[at] = random_text(rng, namb, alphabet="NRYSWKMBDHV")
        qs.append(q)
    check(o, g, qs, 15, modes=(0,))
    check(o, g, qs[:40], 19, modes=(1,))


def test_persistent_kernel_deferral_levels():
    """Copies of one unit with scattered ambiguity codes: the sets of a 13-mer
    window hold one interval per distinct copy, outgrowing the 8-interval and
    32-interval levels, so reads end on the wave kernel; all stay exact."""
    rng = np.random.default_rng(21)
    unit = random_text(rng, 400)
    copies = []
    for _ in range(70):
        u = unit.copy()
        mask = rng.random(400) < 0.03
        u[mask] = rng.choice(O.IUPAC_AMBIG, size=int(mask.sum())).astype(np.uint8)
        copies.append(u)
    seqs = [O.decode(np.concatenate(copies[:35])), O.decode(np.concatenate(copies[35:]))]
    o, g = _pair(seqs, sample_rate=8)
    text, _ = O.concatenate(seqs)
    qs = _mem_reads(rng, text, 200, 120, 0.01, 0.0)
    qs += [unit[i:i + 100].copy() for i in range(0, 300, 7)]
    check(o, g, qs, 13)
    assert g.last_stats()["fallback_reads"] > 0


def test_deferred_reads_split_into_parts():
    """Reads the persistent kernel defers (IUPAC-rich text, some reads with a
    run of N) are fewer than the workgroups of the deferred pass, so each is
    split into parts by MEM start; parts whose sets outgrow the LDS send the
    whole read to the global-set pass while the other parts' records are
    dropped.  Both modes stay equal to the oracle."""
    rng = np.random.default_rng(77)
    rich = random_text(rng, 20000, iupac_frac=0.3)
    plain = random_text(rng, 30000)
    o, g = _pair([O.decode(rich), O.decode(plain)], sample_rate=8)
    qs = []
    for i in range(40):
        src = rich if i < 30 else plain
        st = int(rng.integers(0, len(src) - 150))
        q = src[st:st + 150].copy()
        flip = rng.random(150) < 0.01
        q[flip] = random_text(rng, int(flip.sum()), alphabet="ACGT")
        if i % 4 == 0:
            at = int(rng.integers(30, 110))
            q[at:at + 8] = 5  # N
        qs.append(q)
    check(o, g, qs, 13)
    st = g.last_stats()
    assert st["fallback_reads"] > 0 and st["capacity_retries"] > 0


def test_position_buffers_grow_between_calls():
    """A call with few positions leaves small row and position buffers; the
    next call's MEMs hit 60 copies of a unit, so its positions outgrow them
    (more than two per record): the call grows them and redoes the hits and
    resolve.  Both calls equal the oracle, and a third call of the first
    reads still does after the growth."""
    rng = np.random.default_rng(5)
    unit = random_text(rng, 300)
    seqs = [O.decode(np.concatenate([unit] * 60)), O.decode(random_text(rng, 40000))]
    o, g = _pair(seqs, sample_rate=8)
    text, _ = O.concatenate(seqs)
    start = len(unit) * 60 + 1
    few = [text[start + 100 * i:start + 100 * i + 100].copy() for i in range(5)]
    many = [unit[i:i + 120].copy() for i in range(0, 180, 3)]
    check(o, g, few, 20, modes=(0,))
    check(o, g, many, 20, modes=(0,))
    assert sum(len(h.positions) for hits in g.find_smems_gpu(many, 20, o.fwd.seq_boundaries) for h in hits) > 2000
    check(o, g, few, 20, modes=(0,))


def test_wave_sets_beyond_lds_capacity():
    """All-N queries over an IUPAC-rich text: the interval sets exceed the
    1024-interval LDS sets and re-run with global-memory sets."""
    rng = np.random.default_rng(8)
    o, g = _pair([O.decode(random_text(rng, 20000, iupac_frac=0.3))], sample_rate=16)
    qs = [np.full(10, 5, np.uint8), np.concatenate([random_text(rng, 20), np.full(6, 5, np.uint8)])]
    check(o, g, qs, 3, modes=(0,))
    assert g.last_stats()["capacity_retries"] > 0


def test_locate_then_mem_on_one_bidir_index():
    """The forward half of a BidirFmIndex serves locate and MEM search in any order."""
    o, g = _pair(["ACGTTGCAACGGTACGATTACA", "TTGACCAGTAGGAC"], sample_rate=4)
    offs, hits = g.fwd.locate_batch(["ACG", "TTG"])
    eo, eh = o.fwd.locate_batch(["ACG", "TTG"])
    assert np.array_equal(offs, eo) and np.array_equal(hits, eh)
    check(o, g, ["ACGTTGCAAC", "CCAGTAGGA"], 3)


@pytest.mark.parametrize("seg", ["0", "32", "100", "256"])
def test_long_reads_split_into_segments(monkeypatch, seg):
    """Reads of 300-4000 codes split into work items of `seg` codes (0: whole
    reads) give the oracle's MEMs, including MEMs that cross item bounds."""
    rng = np.random.default_rng(41)
    text, bounds = O.gen_reference(42, 200_000, 3, gc=0.41, repeat_frac=0.3, n_units=20, unit_min=300,
                                   unit_max=2000, iupac_frac=0.005)
    o = O.OracleBidir.build(text=text, bounds=bounds, sample_rate=16, threads=8)
    g = hf.BidirFmIndex.from_bytes(o.to_bytes())
    flat, off = O.gen_reads(43, text, bounds, 60, 1500, sigma=0.6, len_min=300, len_max=4000, sub=0.015,
                            both_strands=True)
    qs = [flat[off[i]:off[i + 1]] for i in range(len(off) - 1)]
    qs += [random_text(rng, 700) for _ in range(5)]
    monkeypatch.setenv("HSFM_SEGMENT", seg)
    check(o, g, qs, 19, cap=32)


def test_rate64_multi_contig_repeats_with_n():
    """The learned-seed workload's reference shape at 1.5 Mbp: 24 contigs, 45%
    repeats from 200 families at 5-25% divergence, GC 0.41, 0.2% N, SA sample
    rate 64; 150 bp reads from both strands with 1% substitutions, min_len 21,
    SMEM and MEM lists against the oracle."""
    text, bounds = O.gen_reference(505, 1_500_000, 24, gc=0.41, repeat_frac=0.45, n_units=200,
                                   div_min=0.05, div_max=0.25, iupac_frac=0.002, iupac_codes=[5])
    o = O.OracleBidir.build(text=text, bounds=bounds, sample_rate=64, threads=8)
    g = hf.BidirFmIndex.from_bytes(o.to_bytes())
    flat, off = O.gen_reads(9, text, bounds, 1200, 150, sub=0.01, both_strands=True)
    qs = [flat[off[i]:off[i + 1]] for i in range(len(off) - 1)]
    check(o, g, qs, 21)
