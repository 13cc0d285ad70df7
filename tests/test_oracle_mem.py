"""Pins the oracle's find_smems / find_mems and the device MEM algorithm's CPU
model (checker validation, CPU only)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import (brute_occurs_fn, containment_max, mem_model_min_len_skip, random_text,
                           smem_random_cases)


def _spans(mems):
    return [(m.query_start, m.query_end) for m in mems]


def test_smem_known_answers(golden):
    for case in golden["smem"]["exact"]:
        b = O.OracleBidir.build(case["ref"], sample_rate=1)
        fn = b.find_smems if case["mode"] == "smem" else b.find_mems
        got = fn(case["query"], case["min_len"])
        assert len(got) == len(case["mems"]), case
        for m, (s, e, c) in zip(got, case["mems"]):
            assert (m.query_start, m.query_end) == (s, e)
            if c is not None:
                assert m.match_count == c


def test_smem_brute_force(golden):
    for ref, q in golden["smem"]["brute_force"]:
        b = O.OracleBidir.build([ref], sample_rate=1)
        text, _ = O.concatenate([ref])
        want = [(s, e) for s, e, _ in O.brute_mems(text, q, 1)]
        assert _spans(b.find_smems(q, 1)) == want, (ref, q)
        assert _spans(b.find_mems(q, 1)) == want, (ref, q)


def test_smem_regressions(golden):
    g = golden["smem"]
    q = g["regression_query"]
    for reg in g["regressions"]:
        seqs = [q[a:e] for a, e, _ in reg["refs"]]
        names = [h for _, _, h in reg["refs"]]
        b = O.OracleBidir.build(seqs, headers=names, sample_rate=1)
        for mode in (0, 1):
            mems = b.find_mems_batch([q], min_len=reg["min_len"], mode=mode)[0]
            hit_refs = {names[s] for m in mems for s, _ in b.positions(m)}
            assert hit_refs == set(names), (reg["source"], mode)


def test_smems_equal_containment_filtered_mems_randomized():
    for seqs, names, full in smem_random_cases():
        b = O.OracleBidir.build(seqs, headers=names, sample_rate=1)
        sm = _spans(b.find_smems(full, 15))
        mm = _spans(b.find_mems(full, 15))
        assert sorted(sm) == containment_max(mm)
        # MEM spans never nest, so both modes return the same list
        assert sm == mm


@pytest.mark.parametrize("seed", range(4))
def test_mems_iupac_brute_force_and_smem_equality(seed):
    rng = np.random.default_rng(100 + seed)
    for trial in range(5):
        seqs = [O.decode(random_text(rng, int(rng.integers(30, 160)), iupac_frac=0.08))
                for _ in range(int(rng.integers(1, 3)))]
        text, bounds = O.concatenate(seqs)
        b = O.OracleBidir.build(text=text, bounds=bounds, sample_rate=int(rng.choice([1, 4, 16])))
        for _ in range(6):
            src = text[rng.integers(0, max(1, len(text) - 40)):][:40]
            q = np.where(src == 0, 1, src).astype(np.uint8)
            q = q.copy()
            flip = rng.random(len(q)) < 0.15
            q[flip] = random_text(rng, int(flip.sum()), alphabet="ACGTN")
            min_len = int(rng.integers(1, 6))
            want = O.brute_mems(text, q, min_len)
            sm = b.find_smems(q, min_len)
            mm = b.find_mems(q, min_len)
            assert [(m.query_start, m.query_end, m.match_count) for m in mm] == want
            assert [(m.query_start, m.query_end, m.match_count) for m in sm] == want
            # positions: every occurrence of the span, as a set
            for m in mm:
                pos = sorted(p for _, p in b.positions(m, ref_boundaries=[len(text)]))
                assert pos == O.naive_locate(text, q[m.query_start:m.query_end])


def test_device_algorithm_model_matches_brute_force():
    """CPU model of the device's min_len-skip MEM algorithm == brute-force MEMs."""
    rng = np.random.default_rng(2024)
    for trial in range(60):
        text, _ = O.concatenate([O.decode(random_text(rng, int(rng.integers(20, 120)),
                                                      iupac_frac=0.1 if trial % 2 else 0.0))])
        q = random_text(rng, int(rng.integers(1, 30)), alphabet="ACGTN", iupac_frac=0.05)
        min_len = int(rng.integers(1, 8))
        occ = brute_occurs_fn(text, q)
        got = mem_model_min_len_skip(occ, len(q), min_len)
        want = [(s, e) for s, e, _ in O.brute_mems(text, q, min_len)]
        assert got == want, (O.decode(text), O.decode(q), min_len)


def test_empty_query_and_zero_min_len():
    b = O.OracleBidir.build(["ACGT"], sample_rate=1)
    assert b.find_smems(np.zeros(0, np.uint8), 1) == []
    assert b.find_mems(np.zeros(0, np.uint8), 1) == []
    assert b.find_smems("ACGT", 0) == []


def test_bidir_extend_matches_unidirectional_count():
    rng = np.random.default_rng(8)
    text, bounds = O.concatenate([O.decode(random_text(rng, 400, iupac_frac=0.05))])
    b = O.OracleBidir.build(text=text, bounds=bounds, sample_rate=4)
    n = b.fwd.text_len
    for _ in range(40):
        q = random_text(rng, int(rng.integers(1, 7)))
        iv = (0, n, 0, n)
        for c in q:
            iv = O.extend_right(b.rev, iv, int(c)) if iv else None
        cnt = (iv[1] - iv[0]) if iv else 0
        assert cnt == len([p for p in O.naive_locate(text, q) if all(text[p + k] == q[k] for k in range(len(q)))])
        iv2 = (0, n, 0, n)
        for c in q[::-1]:
            iv2 = O.extend_left(b.fwd, iv2, int(c)) if iv2 else None
        assert (iv2[:2] if iv2 else None) == (iv[:2] if iv else None)
