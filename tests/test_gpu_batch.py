"""GPU checks of the batched (graph-captured) MEM pipeline for uneven reads:
every configuration -- main kernel, resolve fan-out, graphs or eager
launches, windows of one or more batches per graph, windows in flight --
returns exactly what the one-shot entry point returns, batch by batch."""
import numpy as np
import pytest

import haystackfm_amd as hf
from oracle import oracle as O
from tests.helpers import random_text

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def uneven():
    text, bounds = O.gen_reference(21, 400_000, 4, gc=0.41, repeat_frac=0.2, n_units=30, unit_min=300,
                                   unit_max=3000, iupac_frac=0.005)
    o = O.OracleBidir.build(text=text, bounds=bounds, sample_rate=32, threads=8)
    g = hf.BidirFmIndex.from_bytes(o.to_bytes())
    flat, off = O.gen_reads(22, text, bounds, 3000, 150, sigma=1.0, len_min=32, len_max=3000, sub=0.015,
                            both_strands=True)
    return o, g, (flat, off), bounds


def _same(a, b):
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2])


@pytest.mark.parametrize("window", [1, 2, 0])
@pytest.mark.parametrize("main,fan,graph", [(1, 1, True), (0, 2, True), (2, 4, True), (3, 1, True), (1, 2, False)])
def test_batched_matches_one_shot(uneven, main, fan, graph, window):
    o, g, (flat, off), bounds = uneven
    ref = g.find_mems_batch((flat, off), 19, hf.HSFM_MODE_SMEM, bounds, 1024)
    got = g.find_mems_batched((flat, off), 19, hf.HSFM_MODE_SMEM, bounds, 1024, batch_reads=700,
                              main_kernel=main, resolve_fanout=fan, use_graph=graph, window_batches=window)
    _same(got[:3], ref)
    st = got[3]
    assert st["batches"] == 5 and st["main_kernel"] == main and st["resolve_fanout"] == fan
    assert st["records"] == len(ref[1]) and st["positions"] == len(ref[2])
    # windows of 1, 2 and (default) all 5 batches: 5, 3 and 1 graphs
    assert st["window_batches"] == (window or 5) and st["windows"] == {1: 5, 2: 3, 0: 1}[window]


@pytest.mark.parametrize("sets", [1, 2, 3, 4, 8])
def test_batches_in_flight(uneven, sets):
    """1 to 8 windows of two batches in flight (their own buffers, graphs and
    streams): the same results, handed over in batch order."""
    o, g, (flat, off), bounds = uneven
    ref = g.find_mems_batch((flat, off), 19, hf.HSFM_MODE_SMEM, bounds, 1024)
    got = g.find_mems_batched((flat, off), 19, hf.HSFM_MODE_SMEM, bounds, 1024, batch_reads=450,
                              main_kernel=1, resolve_fanout=2, batch_sets=sets, window_batches=2)
    _same(got[:3], ref)
    assert got[3]["batches"] == 7 and got[3]["windows"] == 4


def test_batched_matches_oracle_sample(uneven):
    o, g, (flat, off), bounds = uneven
    n = 150
    sub = (flat[:int(off[n])], off[:n + 1])
    got = g.find_mems_batched(sub, 19, hf.HSFM_MODE_MEM, bounds, 64, batch_reads=64)
    exp = o.find_mems_batch(qflat=sub[0], qoff=sub[1], min_len=19, mode=1, threads=4)
    ho, hits = got[0], got[1]
    for q, em in enumerate(exp):
        spans = sorted({(int(h[0]), int(h[1])) for h in hits[ho[q]:ho[q + 1]]})
        assert spans == sorted({(m.query_start, m.query_end) for m in em})


def test_autotune_picks_and_caches(uneven):
    _, g, (flat, off), bounds = uneven
    ref = g.find_mems_batch((flat, off), 19, hf.HSFM_MODE_SMEM, bounds, 1024)
    got = g.find_mems_batched((flat, off), 19, hf.HSFM_MODE_SMEM, bounds, 1024, batch_reads=1000, autotune=True)
    _same(got[:3], ref)
    st = got[3]
    assert 0 <= st["main_kernel"] < 4 and st["resolve_fanout"] in (1, 2, 4)
    assert all(t > 0 for t in st["tune_ms"])
    again = g.find_mems_batched((flat, off), 19, hf.HSFM_MODE_SMEM, bounds, 1024, batch_reads=1000)
    assert (again[3]["main_kernel"], again[3]["resolve_fanout"]) == (st["main_kernel"], st["resolve_fanout"])
    _same(again[:3], ref)


def test_oversized_batch_runs_host_driven():
    """Reads of N over an IUPAC-rich text make sets beyond the in-graph global
    capacity: their batch re-runs on the host-driven path, still exact."""
    rng = np.random.default_rng(8)
    seqs = [O.decode(random_text(rng, 120_000, iupac_frac=0.3))]
    o = O.OracleBidir.build(seqs, sample_rate=16, threads=8)
    g = hf.BidirFmIndex.from_bytes(o.to_bytes())
    qs = [random_text(rng, 60, iupac_frac=0.01) for _ in range(20)] + [np.full(40, 5, np.uint8)]
    ref = g.find_mems_batch(qs, 13, hf.HSFM_MODE_SMEM, [], 16)
    for window in (1, 0):
        got = g.find_mems_batched(qs, 13, hf.HSFM_MODE_SMEM, [], 16, batch_reads=8, window_batches=window)
        _same(got[:3], ref)
        assert got[3]["fallback_batches"] >= 1


@pytest.mark.parametrize("seed_f", ["14", "15"])
def test_batched_with_longer_seeds(uneven, monkeypatch, seed_f):
    """The batched pipeline seeded from packed 14- or 15-mer records returns
    what the one-shot entry point returns with 12-mer seeds."""
    _, g, (flat, off), bounds = uneven
    monkeypatch.setenv("HSFM_SEED_F", "12")
    ref = g.find_mems_batch((flat, off), 19, hf.HSFM_MODE_SMEM, bounds, 1024)
    monkeypatch.setenv("HSFM_SEED_F", seed_f)
    got = g.find_mems_batched((flat, off), 19, hf.HSFM_MODE_SMEM, bounds, 1024, batch_reads=700, main_kernel=1,
                              resolve_fanout=1)
    _same(got[:3], ref)


def test_callback_into_the_same_index_is_refused(uneven):
    """The index stays locked while a batch callback runs: a call on the same
    index from the callback returns an error instead of deadlocking, and the
    pipeline's results are unaffected."""
    o, g, (flat, off), bounds = uneven
    sub = (flat[:int(off[200])], off[:201])
    errors = []
    L = hf._lib()
    dev = g.fwd._dev[0]._h
    ptrs = []
    for arr in (sub[0], sub[1]):
        p = hf.C.c_void_p()
        hf._check(L.hsfm_device_alloc(dev, max(arr.nbytes, 8), hf.C.byref(p)), "alloc")
        hf._check(L.hsfm_host_to_device(dev, p, arr.ctypes.data, arr.nbytes), "h2d")
        ptrs.append(p)

    def on_batch(first, n, res):
        try:
            g.find_mems_batch(sub, 19, hf.HSFM_MODE_SMEM, (), 16)
        except hf.GpuError as e:
            errors.append(str(e))

    try:
        st = g.find_mems_batched_device(ptrs[0].value, ptrs[1].value, 200, 19, hf.HSFM_MODE_SMEM, 0, 0, 16,
                                        batch_reads=100, main_kernel=1, resolve_fanout=1, on_batch=on_batch)
    finally:
        for p in ptrs:
            L.hsfm_device_free(dev, p)
    assert st["batches"] == 2 and len(errors) == 2
    assert all("batch callback on the same index" in e for e in errors)
    # the index is usable again once the batched call has returned
    ref = g.find_mems_batch(sub, 19, hf.HSFM_MODE_SMEM, (), 16)
    got = g.find_mems_batched(sub, 19, hf.HSFM_MODE_SMEM, (), 16, batch_reads=100, main_kernel=1, resolve_fanout=1)
    _same(got[:3], ref)


def _device_batched(g, dq, do, dr, nq, nrefs, **kw):
    """find_mems_batched_device on device buffers; the batches' results copied
    to the host and joined as find_mems_batch returns them."""
    L = hf._lib()
    dev = g.fwd._dev[0]._h
    parts = []

    def on_batch(first, n, r):
        nh, npos = int(r.num_hits), int(r.num_positions)
        ho = np.zeros(n + 1, np.uint64)
        hits = np.zeros((nh, 8), np.uint32)
        pos = np.zeros((npos, 2), np.uint32)
        for arr, src in ((ho, r.d_hit_offsets), (hits, r.d_hits), (pos, r.d_positions)):
            if arr.nbytes:
                hf._check(L.hsfm_device_to_host(dev, arr.ctypes.data, hf.C.c_void_p(src), arr.nbytes), "d2h")
        parts.append((int(first), ho, hits, pos))

    st = g.find_mems_batched_device(dq, do, nq, 19, hf.HSFM_MODE_SMEM, dr, nrefs, 1024, on_batch=on_batch, **kw)
    parts.sort(key=lambda t: t[0])
    ho_all, hits_all, pos_all, nh, npos = [np.zeros(1, np.uint64)], [], [], 0, 0
    for _, ho, hits, pos in parts:
        ho_all.append(ho[1:] + np.uint64(nh))
        if len(hits):
            po = (hits[:, 4].astype(np.uint64) | (hits[:, 5].astype(np.uint64) << np.uint64(32))) + np.uint64(npos)
            hits = hits.copy()
            hits[:, 4] = (po & np.uint64(0xFFFFFFFF)).astype(np.uint32)
            hits[:, 5] = (po >> np.uint64(32)).astype(np.uint32)
        hits_all.append(hits)
        pos_all.append(pos)
        nh += len(hits)
        npos += len(pos)
    return (np.concatenate(ho_all), np.concatenate(hits_all), np.concatenate(pos_all)), st


def test_graphs_reused_across_calls(uneven):
    """A second call with the same buffers and configuration launches the
    graphs the first one captured; after the read buffer is overwritten with
    other reads of the same lengths, the reused graphs return the new reads'
    results; another batch size captures again."""
    _, g, (flat, off), bounds = uneven
    rng = np.random.default_rng(31)
    other = flat.copy()
    flip = rng.random(len(other)) < 0.05
    other[flip] = rng.integers(1, 5, int(flip.sum())).astype(np.uint8) # A, C, G, T codes
    rb = np.asarray(bounds, np.uint32)
    nq = len(off) - 1
    L = hf._lib()
    dev = g.fwd._dev[0]._h
    ptrs = []
    for arr in (flat, off.astype(np.uint64), rb):
        p = hf.C.c_void_p()
        hf._check(L.hsfm_device_alloc(dev, max(arr.nbytes, 8), hf.C.byref(p)), "alloc")
        hf._check(L.hsfm_host_to_device(dev, p, arr.ctypes.data, arr.nbytes), "h2d")
        ptrs.append(p)
    dq, do, dr = (p.value for p in ptrs)
    kw = dict(batch_reads=500, main_kernel=1, resolve_fanout=2, batch_sets=2, window_batches=2)
    try:
        first, st1 = _device_batched(g, dq, do, dr, nq, len(rb), **kw)
        _same(first, g.find_mems_batch((flat, off), 19, hf.HSFM_MODE_SMEM, bounds, 1024))
        assert st1["windows"] == 3 and st1["graphs_captured"] >= 1
        again, st2 = _device_batched(g, dq, do, dr, nq, len(rb), **kw)
        _same(again, first)
        assert st2["graphs_captured"] == 0 and st2["windows"] == 3
        hf._check(L.hsfm_host_to_device(dev, ptrs[0], other.ctypes.data, other.nbytes), "h2d")
        new, st3 = _device_batched(g, dq, do, dr, nq, len(rb), **kw)
        assert st3["graphs_captured"] == 0
        _same(new, g.find_mems_batch((other, off), 19, hf.HSFM_MODE_SMEM, bounds, 1024))
        kw["batch_reads"] = 700
        hf._check(L.hsfm_host_to_device(dev, ptrs[0], flat.ctypes.data, flat.nbytes), "h2d")
        redo, st4 = _device_batched(g, dq, do, dr, nq, len(rb), **kw)
        assert st4["graphs_captured"] >= 1
        _same(redo, first)
    finally:
        for p in ptrs:
            L.hsfm_device_free(dev, p)
