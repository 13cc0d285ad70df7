"""GPU check of the reverse-complement helper (both-strand searching) against
the complement fixed by the IUPAC base sets (src/alphabet.rs:149-168: the
complement of a code is the code whose base set is its base set with A<->T and
C<->G), and a located reverse-strand read."""
import numpy as np
import pytest

import haystackfm_amd as hf
from oracle import oracle as O
from tests.helpers import random_text

pytestmark = pytest.mark.gpu

_SWAP = {1: 4, 4: 1, 2: 3, 3: 2}  # A<->T, C<->G


def _comp_code(c: int) -> int:
    if c == 0 or c >= 16:
        return c
    want = sorted(_SWAP[b] for b in O.iupac_bases(c))
    hits = [x for x in range(1, 16) if sorted(O.iupac_bases(x)) == want]
    assert len(hits) == 1
    return hits[0]


def _revcomp(q):
    return np.array([_comp_code(int(c)) for c in q[::-1]], np.uint8)


def test_reverse_complement_matches_base_sets():
    rng = np.random.default_rng(31)
    qs = [rng.integers(0, 16, int(n)).astype(np.uint8) for n in rng.integers(0, 300, 200)]
    qs += [np.zeros(0, np.uint8), np.array([16, 250, 0, 5, 1], np.uint8), np.arange(16, dtype=np.uint8)]
    flat, off = hf.reverse_complement(qs)
    assert len(off) == len(qs) + 1 and int(off[-1]) == sum(len(q) for q in qs)
    for i, q in enumerate(qs):
        assert np.array_equal(flat[int(off[i]):int(off[i + 1])], _revcomp(q)), i
    back, off2 = hf.reverse_complement((flat, off))
    assert np.array_equal(off2, off) and np.array_equal(back, np.concatenate(qs))


def test_reverse_strand_read_locates_at_its_origin():
    rng = np.random.default_rng(32)
    seqs = [O.decode(random_text(rng, 20_000, iupac_frac=0.01)) for _ in range(2)]
    o = O.OracleIndex.build(seqs, sample_rate=8)
    g = hf.FmIndex.from_bytes(o.to_bytes())
    text, _ = O.concatenate(seqs)
    starts = rng.integers(0, 19_900, 50)
    fwd = [text[s:s + 40].copy() for s in starts]
    reads = [_revcomp(f) for f in fwd]  # reads from the reverse strand
    flat, off = hf.reverse_complement(reads)
    offs, hits = g.locate_batch((flat, off))
    for i, s in enumerate(starts):
        got = {tuple(map(int, h)) for h in hits[int(offs[i]):int(offs[i + 1])]}
        assert (0, int(s)) in got
