"""Pins the oracle's alphabet, construction and rank primitives against the
reference's golden data and brute force (checker validation, CPU only)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import random_text


def test_compat_table_matches_golden(golden):
    g = golden["compat"]
    for code in range(16):
        lst = O.compatible(code)
        assert len(lst) == g["len"][code]
        assert lst == g["table"][code][:len(lst)]
        assert all(v == 0 for v in g["table"][code][len(lst):])


def test_compat_is_base_set_overlap():
    # closed form: two codes are compatible iff their base sets intersect
    for a in range(1, 16):
        ba = set(O.iupac_bases(a))
        for b in range(16):
            bb = set(O.iupac_bases(b))
            assert (b in O.compatible(a)) == bool(ba & bb), (a, b)
    # ExactDna: ACGT match themselves only (alphabet.rs:286-294)
    for c in range(16):
        assert O.compatible(c, tag=1) == ([c] if 1 <= c <= 4 else [])


def test_compat_symmetric():
    for a in range(1, 16):
        for b in O.compatible(a):
            assert a in O.compatible(b)


def test_iupac_bases_golden(golden):
    for code, bases in golden["iupac_bases"]["cases"]:
        assert O.decode(O.iupac_bases(O.encode(code)[0])) == bases


def test_encode_rules():
    assert list(O.encode("acgtU")) == [1, 2, 3, 4, 4]
    for bad in "-.XZ":
        with pytest.raises(ValueError):
            O.encode(bad)


def test_c_array_golden(golden):
    for text, expect in golden["c_array"]["cases"]:
        body = text.rstrip("$")
        idx = O.OracleIndex.build([body], sample_rate=1)
        assert idx.c_array == expect


@pytest.mark.parametrize("seed", range(6))
def test_sais_equals_comparison_sort(seed):
    rng = np.random.default_rng(seed)
    for _ in range(60):
        n = int(rng.integers(1, 400))
        alpha = int(rng.integers(1, 17))
        t = rng.integers(0, alpha, n).astype(np.uint8)
        t[-1] = 0
        if rng.random() < 0.5:  # several sentinels
            t[rng.integers(0, n, 3)] = 0
        assert (O.suffix_array(t) == O.suffix_array(t, naive=True)).all()


def _naive_sa(text):
    return sorted(range(len(text)), key=lambda i: bytes(text[i:]))


def test_sais_against_python_sort():
    rng = np.random.default_rng(7)
    for _ in range(30):
        t = rng.integers(0, 6, int(rng.integers(1, 120))).astype(np.uint8)
        t[-1] = 0
        assert list(O.suffix_array(t)) == _naive_sa(t)


@pytest.mark.parametrize("encoding", [0, 1])
@pytest.mark.parametrize("iupac", [0.0, 0.2])
def test_rank_and_lf_step_vs_naive(encoding, iupac):
    rng = np.random.default_rng(11 + encoding)
    seqs = [O.decode(random_text(rng, int(rng.integers(40, 300)), iupac_frac=iupac)) for _ in range(3)]
    text, bounds = O.concatenate(seqs)
    idx = O.OracleIndex.build(text=text, bounds=bounds, sample_rate=4, encoding=encoding)
    sa = O.suffix_array(text)
    n = len(text)
    bwt = np.array([text[(int(s) - 1) % n] for s in sa], np.uint8)
    for c in range(16):
        pref = np.concatenate([[0], np.cumsum(bwt == c)])
        for i in range(0, n + 1, 7):
            assert idx.rank(c, i) == pref[i], (c, i)
        assert idx.rank(c, n) == pref[n]
    C = idx.c_array
    for pos in range(n):
        sym, r = idx.lf_step(pos)
        assert sym == bwt[pos] == idx.symbol_at(pos)
        assert r == int((bwt[:pos] == sym).sum())
        # LF(i) is the row of suffix SA[i]-1 (inverse-BWT closed form)
        lf = C[sym] + r
        assert int(sa[lf]) == (int(sa[pos]) - 1) % n


def test_sampled_sa_and_resolve():
    rng = np.random.default_rng(3)
    seqs = [O.decode(random_text(rng, int(rng.integers(50, 200)))) for _ in range(4)]
    text, bounds = O.concatenate(seqs)
    sa = O.suffix_array(text)
    starts = {0} | {int(b) for b in bounds[:-1]}
    for rate in (1, 2, 5, 32):
        idx = O.OracleIndex.build(text=text, bounds=bounds, sample_rate=rate)
        for row in range(len(text)):
            v = idx.sa_get(row)
            sampled = sa[row] % rate == 0 or sa[row] in starts
            assert (v is not None) == sampled
            if v is not None:
                assert v == sa[row]
            assert idx.resolve_sa(row) == sa[row]


def test_serialize_roundtrip_and_tag_last_byte():
    idx = O.OracleIndex.build(["ACGTACGTACGT"], sample_rate=4)
    b = idx.to_bytes()
    assert b[-1] == 0  # alphabet_tag is the last byte (serialize.rs:138)
    r = O.OracleIndex.from_bytes(b)
    assert r.to_bytes() == b
    for p in ["A", "AC", "ACGT", "GT", "ACGTACGT"]:
        assert r.count(p) == idx.count(p)
        assert sorted(r.locate(p)) == sorted(idx.locate(p))
    bad = bytearray(b)
    bad[-1] = 0xFF
    with pytest.raises(ValueError):
        O.OracleIndex.from_bytes(bytes(bad))
    with pytest.raises(ValueError):
        O.OracleIndex.from_bytes(b[:-5])


def test_serialize_field_layout():
    """bincode-1 field order (serialize.rs:55-67): spot-check fixed offsets."""
    import struct
    idx = O.OracleIndex.build(["ACGT", "TGCA"], sample_rate=2, lookup_depth=2)
    b = idx.to_bytes()
    c = struct.unpack_from("<16I", b, 0)
    assert list(c) == idx.c_array
    L, P = b[64], b[65]
    assert (L, P) == (5, 3)  # {$,A,C,G,T}
    off = 66 + 32
    (nsb,) = struct.unpack_from("<Q", b, off)
    n = idx.text_len
    nb = (n + 63) // 64
    assert nsb == ((nb + 7) // 8) * L
    off += 8 + 4 * nsb
    (nblk,) = struct.unpack_from("<Q", b, off)
    stride = 6 * L + 8 * P
    assert nblk == nb * stride
    off += 8 + nblk
    (st,) = struct.unpack_from("<Q", b, off)
    assert st == stride
    enc, tl = struct.unpack_from("<II", b, off + 8)
    assert (enc, tl) == (0, n)
