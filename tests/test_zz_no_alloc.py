"""The allocation-free learned-seed lookup (hsfm_lsi_seed_device_noalloc)
against the entry point it mirrors (hsfm_lsi_seed_device): the same results
and load counts on two calls in a row, the second with fewer keys, for the
LDS-leaf and the global-leaf lookup kernels."""
import ctypes as C

import numpy as np
import pytest

import haystackfm_amd as hf
from oracle import oracle as O

pytestmark = pytest.mark.gpu


class _Dev:
    def __init__(self, g):
        self.L = hf._lib()
        self.dev = g.fwd._dev[0]._h
        self.ptrs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        hf._check(self.L.hsfm_device_alloc(self.dev, max(nbytes, 8), C.byref(p)), "alloc")
        self.ptrs.append(p)
        return p

    def put(self, arr):
        p = self.alloc(arr.nbytes)
        hf._check(self.L.hsfm_host_to_device(self.dev, p, arr.ctypes.data, arr.nbytes), "h2d")
        return p

    def get(self, p, shape, dt):
        out = np.zeros(shape, dt)
        hf._check(self.L.hsfm_device_to_host(self.dev, out.ctypes.data, p, out.nbytes), "d2h")
        return out

    def free(self):
        for p in self.ptrs:
            self.L.hsfm_device_free(self.dev, p)


@pytest.mark.parametrize("lookup", ["lds", "global"])
def test_noalloc_lookup_matches_seed_device(lookup):
    text, bounds = O.gen_reference(93, 400_000, 3, gc=0.41, repeat_frac=0.3, n_units=30, unit_min=300,
                                   unit_max=3000, iupac_frac=0.002)
    o = O.OracleBidir.build(text=text, bounds=bounds, sample_rate=16, threads=8)
    g = hf.BidirFmIndex.from_bytes(o.to_bytes())
    lsi = hf.LearnedSeedIndex(g, k=21, leaves=4096, hidden=16, train_steps=200, lookup=lookup)
    rng = np.random.default_rng(5)
    flat, _ = O.gen_reads(94, text, bounds, 4000, 150, sub=0.01, both_strands=True)
    q = hf.kmer_keys(flat, 21)
    q = q[q != np.uint64(0xFFFFFFFFFFFFFFFF)]
    q = np.concatenate([q, rng.integers(0, 4 ** 21, 20000, dtype=np.uint64)])
    rng.shuffle(q)
    d = _Dev(g)
    try:
        for n in (len(q), len(q) // 3):  # two calls in a row, the second smaller
            keys = np.ascontiguousarray(q[:n])
            dk = d.put(keys)
            da, db = d.alloc(16 * n), d.alloc(16 * n)
            st = lsi.seed_device(dk.value, n, da.value)
            probes = lsi.seed_device_noalloc(dk.value, n, db.value)
            a = d.get(da, (n, 4), np.uint32)
            b = d.get(db, (n, 4), np.uint32)
            assert np.array_equal(a, b)
            assert probes == st["probes"]
            assert np.array_equal(b, hf.kmer_search(g, keys, 21))
            assert b[:, 3].sum() > n // 4
    finally:
        d.free()
