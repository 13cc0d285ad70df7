"""GPU parity at the stated size of the learned-seed workload.

1 Gbp in 24 contigs (GC 0.41, 45% copies of 2,000 repeat families of
300-6,000 bp at 5-25% divergence, 0.2% N), bidirectional index with SA sample
rate 64 and no lookup table, built on the device; 4,000,000 x 150 bp reads
from both strands with 1% substitutions; SMEM min_len 21, max_hits 1024;
learned seed index k = 21 with 4,096 leaves.

  - learned-seed SMEM over the whole batch equals table-seeded SMEM, record
    for record and position for position;
  - plain seeding (restarts from the full interval) on a slice of 100,000
    reads equals table seeding on the same slice;
  - the oracle's SMEMs (find_smems, resolve, reference map) on a sample spread
    over the batch that holds every read the learned-mode main kernel defers.
"""
import numpy as np
import pytest

import haystackfm_amd as hf
from oracle import oracle as O
from tests.test_gpu_mem import check as check_mems
from workloads import index_build as IB

pytestmark = pytest.mark.gpu

MIN_LEN, CAP = 21, 1024


def _log(*a):
    print("[config5]", *a, flush=True)


def _same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.fixture(scope="module")
def big():
    text, bounds = IB.synth_reference(0xC5EED, 1_000_000_000, 24, gc=0.41, repeat_frac=0.45, n_units=2000,
                                      unit_min=300, unit_max=6000, div_min=0.05, div_max=0.25,
                                      iupac_frac=0.002, iupac_codes=[5])
    _log(f"reference {len(text)} bp")
    idx = hf.build_index_bytes_gpu(text, bounds, sample_rate=64, lookup_depth=0, bidir=True)
    _log(f"index {len(idx) / 1e6:.0f} MB")
    sq, so = IB.synth_reads(0x5EED5, text, bounds, 4_000_000, 150, sub=0.01, both_strands=True, shuffle=True)
    del text
    g = hf.BidirFmIndex.from_bytes(idx)
    g.set_seed_mode(hf.HSFM_SEED_TABLES)
    _log("reads synthesised, index uploaded")
    return idx, g, np.asarray(bounds, np.uint32), sq, so


def test_learned_seeding_whole_batch_and_oracle_sample(big):
    idx, g, bounds, sq, so = big
    nq = len(so) - 1
    ref = g.find_mems_batch((sq, so), MIN_LEN, hf.HSFM_MODE_SMEM, bounds, CAP)
    _log(f"tables: {len(ref[1])} records, {len(ref[2])} positions")
    lsi = hf.LearnedSeedIndex(g, k=21, leaves=4096)
    assert lsi.info()["mem_seeding"] == 1
    try:
        g.set_seed_mode(hf.HSFM_SEED_LEARNED, lsi)
        got = g.find_mems_batch((sq, so), MIN_LEN, hf.HSFM_MODE_SMEM, bounds, CAP)
        st = g.last_stats()
        deferred = g.last_deferred()
        _log(f"learned: deferred {len(deferred)}, accesses {st['occ_probes']}")
        _same(got, ref)
        assert st["guard_hits"] == 0

        n = 100_000
        sub = (sq[:int(so[n])], so[:n + 1])
        g.set_seed_mode(hf.HSFM_SEED_PLAIN)
        plain = g.find_mems_batch(sub, MIN_LEN, hf.HSFM_MODE_SMEM, bounds, CAP)
        g.set_seed_mode(hf.HSFM_SEED_TABLES)
        _same(plain, g.find_mems_batch(sub, MIN_LEN, hf.HSFM_MODE_SMEM, bounds, CAP))
        _log("plain slice equal")

        # the oracle on 200 reads spread over the batch plus every deferred read
        pick = np.union1d(np.linspace(0, nq - 1, 200).astype(np.int64), deferred.astype(np.int64))
        reads = [sq[int(so[i]):int(so[i + 1])] for i in pick]
        o = O.OracleBidir.from_bytes(idx)
        g.set_seed_mode(hf.HSFM_SEED_LEARNED, lsi)
        check_mems(o, g, reads, MIN_LEN, bounds, CAP, modes=(0,))
        _log(f"oracle sample of {len(reads)} reads equal")
    finally:
        g.set_seed_mode(hf.HSFM_SEED_TABLES)
