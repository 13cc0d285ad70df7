"""CPU pin of the learned index's restart lookup (oracle/lsi.py restart_*,
the model of hsfm_lsi_dev.h and k_lsi_windows / k_lsi_short / k_lsi_bounds):
for base strings drawn from the text, mutated and random, the run of full
entries at the lower bound is exactly the set of compatible text windows, and
for an absent string the neighbour formula over the G / H bounds gives exactly
the longest suffix with a compatible occurrence -- both checked by brute force
over the whole text, under IupacDna (ambiguity codes expand) and ExactDna
(they cut windows), with several contigs and short ones."""
import numpy as np
import pytest

from oracle import lsi as L
from oracle import oracle as O


def _text(seed, n, frac, contigs):
    text, bounds = O.gen_reference(seed, n, contigs, gc=0.45, repeat_frac=0.3, n_units=4, unit_min=20, unit_max=80,
                                   iupac_frac=frac)
    return text


@pytest.mark.parametrize("tag,seed,k", [(0, 1, 6), (0, 2, 8), (1, 3, 6), (1, 4, 7)])
def test_restart_lookup_matches_brute_force(tag, seed, k):
    rng = np.random.default_rng(seed)
    text = _text(seed, 1500, 0.02, 4)
    ent = L.restart_entries(text, k, tag)
    assert any(e[0] & L.SHORT for e in ent)
    if tag == 0:
        assert any(e[0] & L.VARIANT and not e[0] & L.SHORT for e in ent)
    queries = []
    for _ in range(250):  # windows of the text with bases for ambiguity codes, some mutated
        s = int(rng.integers(0, len(text) - k))
        w = [int(c) if 1 <= c <= 4 else int(rng.integers(1, 5)) for c in text[s:s + k]]
        if rng.random() < 0.5:
            w[int(rng.integers(0, k))] = int(rng.integers(1, 5))
        queries.append(w)
    queries += [list(rng.integers(1, 5, k)) for _ in range(150)]
    present = absent = 0
    for w in queries:
        got = L.restart_lookup(ent, k, w)
        exp = L.restart_brute(text, k, w, tag)
        assert got == exp, (w, got, exp)
        present += bool(exp[0])
        absent += not exp[0]
    assert present > 50 and absent > 50


def test_without_bounds_short_entries_would_be_wrong():
    """The G / H bounds matter: using the raw neighbours' common suffix
    (capped by their own reach only) misses longer suffixes behind a short
    neighbour on some absent strings of a text with many short windows."""
    rng = np.random.default_rng(9)
    k = 6
    text = _text(11, 1200, 0.05, 6)
    ent = L.restart_entries(text, k, 1)
    plain = [((e[0] & ~((31 << L.G_SHIFT) | (31 << L.H_SHIFT))) |
              ((e[1] << L.G_SHIFT) | (e[1] << L.H_SHIFT) if e[0] & L.SHORT else 0), e[1], e[2]) for e in ent]
    wrong = 0
    for _ in range(3000):
        w = list(rng.integers(1, 5, k))
        exp = L.restart_brute(text, k, w, 1)
        assert L.restart_lookup(ent, k, w) == exp
        wrong += L.restart_lookup(plain, k, w) != exp
    assert wrong > 0
