"""Pins the oracle's count / locate / lookup seeding (checker validation, CPU only)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import random_text


def test_count_known_answers(golden):
    for text, pat, exp in golden["count"]["cases"]:
        idx = O.OracleIndex.build([text], sample_rate=1)
        assert idx.count(pat) == exp, (text, pat)


def test_count_and_locate_sampled(golden):
    g = golden["count_sampled"]
    idx = O.OracleIndex.build([g["text"]], sample_rate=g["rate"])
    assert idx.count(g["pattern"]) == g["count"]
    assert sorted(p for _, p in idx.locate(g["pattern"])) == g["positions"]


def test_locate_known_answers(golden):
    for text, pat, exp in golden["locate"]["cases"]:
        idx = O.OracleIndex.build([text], sample_rate=1)
        assert sorted(p for _, p in idx.locate(pat)) == exp


def test_map_position(golden):
    g = golden["map_position"]
    idx = O.OracleIndex.build(g["seqs"], sample_rate=1)
    for pos, s, o in g["cases"]:
        assert idx.map_position(pos) == (s, o)


def test_multi_sequence_rates(golden):
    g = golden["multi_sequence_rates"]
    base = O.OracleIndex.build(g["seqs"], sample_rate=1)
    for rate in g["rates"]:
        idx = O.OracleIndex.build(g["seqs"], sample_rate=rate)
        for p in g["patterns"]:
            assert sorted(idx.locate(p)) == sorted(base.locate(p)), (rate, p)


def test_iupac_locate_vs_brute_force(golden):
    for seqs, q in golden["iupac_locate"]["cases"]:
        idx = O.OracleIndex.build(seqs, sample_rate=1)
        text, bounds = O.concatenate(seqs)
        want = sorted(O.ref_map(p, bounds) for p in O.naive_locate(text, q))
        assert sorted(idx.locate(q)) == want, (seqs, q)


@pytest.mark.parametrize("tag", [0, 1])
def test_random_iupac_locate_vs_brute_force(tag):
    rng = np.random.default_rng(21 + tag)
    for trial in range(6):
        seqs = [O.decode(random_text(rng, int(rng.integers(20, 400)), iupac_frac=0.05))
                for _ in range(int(rng.integers(1, 4)))]
        text, bounds = O.concatenate(seqs)
        idx = O.OracleIndex.build(text=text, bounds=bounds, sample_rate=int(rng.choice([1, 3, 8, 32])), tag=tag)
        for _ in range(40):
            m = int(rng.integers(1, 9))
            q = random_text(rng, m, alphabet="ACGTN", iupac_frac=0.2)
            want = sorted(O.ref_map(p, bounds) for p in O.naive_locate(text, q, tag))
            got = idx.locate(q)
            assert sorted(got) == want
            assert idx.count(q) == len(want)


def test_backward_search_intervals_sorted_disjoint():
    rng = np.random.default_rng(5)
    text, bounds = O.concatenate([O.decode(random_text(rng, 500, iupac_frac=0.1))])
    idx = O.OracleIndex.build(text=text, bounds=bounds, sample_rate=4)
    for _ in range(50):
        q = random_text(rng, int(rng.integers(1, 6)), alphabet="ACGTN")
        ivs = idx.backward_search(q)
        for (a, b), (c, d) in zip(ivs, ivs[1:]):
            assert a < b < c < d  # merged: disjoint and non-adjacent
    assert idx.backward_search(np.zeros(0, np.uint8)) == [(0, idx.text_len)]


def _lookup_closed_form(text, q, depth, tag):
    """Seeded backward search = exact ACGT match on the last `depth` codes and
    IUPAC match on the rest (query.rs:225-236 + lookup.rs:63-79)."""
    m = len(q)
    if m < depth or any(c not in (1, 2, 3, 4) for c in q[m - depth:]):
        return O.naive_locate(text, q, tag)
    out = []
    for p in O.naive_locate(text, q, tag):
        if all(text[p + k] == q[k] for k in range(m - depth, m)):
            out.append(p)
    return out


def test_lookup_known_answers(golden):
    for text, depth, tag, pats in golden["lookup"]["cases"]:
        with_l = O.OracleIndex.build([text], sample_rate=1, lookup_depth=depth, tag=tag)
        no_l = O.OracleIndex.build([text], sample_rate=1, lookup_depth=0, tag=tag)
        for p in pats:
            assert with_l.count(p) == no_l.count(p), (text, p)
    g = golden["lookup"]["exact_n_zero"]
    assert O.OracleIndex.build([g["text"]], sample_rate=1, tag=1, lookup_depth=3).count(g["pattern"]) == 0


def test_lookup_seeding_closed_form():
    # text ACNT, depth 3: CGT counts 1 without lookup (N matches G) and 0 with it
    idx0 = O.OracleIndex.build(["ACNT"], sample_rate=1)
    idx3 = O.OracleIndex.build(["ACNT"], sample_rate=1, lookup_depth=3)
    assert idx0.count("CGT") == 1
    assert idx3.count("CGT") == 0
    rng = np.random.default_rng(9)
    for trial in range(4):
        t = random_text(rng, 600, iupac_frac=0.08)
        text, bounds = O.concatenate([O.decode(t)])
        depth = int(rng.integers(1, 5))
        idx = O.OracleIndex.build(text=text, bounds=bounds, sample_rate=8, lookup_depth=depth)
        for _ in range(60):
            q = random_text(rng, int(rng.integers(1, 9)), alphabet="ACGT", iupac_frac=0.1)
            want = sorted(_lookup_closed_form(text, q, depth, 0))
            got = sorted(p for _, p in idx.locate(q))
            assert got == want, (depth, O.decode(q))


def test_batched_locate_matches_single():
    rng = np.random.default_rng(4)
    seqs = [O.decode(random_text(rng, 300, iupac_frac=0.02)) for _ in range(3)]
    idx = O.OracleIndex.build(seqs, sample_rate=8, lookup_depth=3)
    qs = [random_text(rng, int(rng.integers(0, 8)), alphabet="ACGTNR") for _ in range(100)]
    offs, hits = idx.locate_batch(qs, threads=3)
    for i, q in enumerate(qs):
        assert [tuple(map(int, h)) for h in hits[offs[i]:offs[i + 1]]] == idx.locate(q)
