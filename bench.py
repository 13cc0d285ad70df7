#!/usr/bin/env python3
"""Benchmark: batched SMEM+locate of 150 bp reads against a 100 Mbp index on one MI355X.

Workload (one "step" = one pass of the pipeline over the whole read batch,
inputs already resident in HBM):
  reference  100 Mbp in 8 contigs of uniform random lengths, GC 0.41, 20% copies
             of 200 repeat units (300-6,000 bp) at 5-15% divergence, 0.5% IUPAC
             ambiguity codes (R Y S W K M B D H V N); sa_sample_rate 32,
             lookup depth 12 (forward index); bidirectional index.
  reads      1,000,000 x 150 bp sampled from both strands with 2% substitutions
             and 0.2% 1-3 bp indels, plus 100,000 random reads.
  pipeline   BidirFmIndex::find_smems_gpu: SMEM find -> SA-row resolve (LF walk)
             -> reference-boundary map, min_len 18, max_hits_per_mem 1024.
Also measured in the same run (reported beside the headline): batched locate of
1,000,000 x 150 bp reads (both strands, 1% substitutions) on the same index.

`--learned-1g` runs the learned-seed workload instead: a 1 Gbp reference in 24
contigs (45% copies of 2,000 repeat families at 5-25% divergence, GC 0.41, 0.2%
N; SA sample rate 64), 4,000,000 x 150 bp reads (both strands, 1% substitutions);
the learned seed index (4,096 leaves) against plain exact search on one 21-mer per
read, and the SMEM pipeline (min_len 21) over the reads.  Its JSON line reports
the speedup as `value`.

The GPU leg uses only haystackfm_amd (libhsfm.so behind the C ABI, plus the host
index builder).  The oracle is used only in the cpu_baseline leg: to time the
CPU restatement of the reference path on a bounded sample and to diff the GPU
output of that sample (a run whose sample is not bit-exact reports no value).
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SMEM+locate reads/s (150bp, 100Mbp ref) bit-exact; learned-seed vs plain speedup"
HBM_PEAK_GBS = 8000.0
BYTES_DEFINITION = ("64 B x the random accesses of the kernel's own algorithm, counted on the device: Occ record "
                    "probes (one per interval step, two when the interval's ends lie in different records), seed "
                    "records and entries, learned-index rounds (64-byte units), and the suffix-array and text "
                    "loads that finish single-row sets; this is fewer than the reference algorithm's probes "
                    "(reference_algorithm_probes_per_read), so the bytes per read fall as work is removed")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ref-bp", type=int, default=100_000_000)
    ap.add_argument("--reads", type=int, default=1_000_000)
    ap.add_argument("--random-reads", type=int, default=100_000)
    ap.add_argument("--locate-reads", type=int, default=1_000_000)
    ap.add_argument("--min-len", type=int, default=18)
    ap.add_argument("--max-hits", type=int, default=1024)
    ap.add_argument("--check-reads", type=int, default=2000)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--cache", default=os.environ.get("HSFM_BENCH_CACHE", "/tmp/hsfm_bench_cache"))
    ap.add_argument("--uneven-reads", type=int, default=2_000_000,
                    help="reads of the uneven-length batched leg (0 skips it)")
    ap.add_argument("--lsi-leaves", type=int, default=4096,
                    help="stage-2 leaves of the learned seed index (4096 as the learned-seed workload states)")
    ap.add_argument("--lsi-leaves-alt", type=int, default=1 << 20,
                    help="a second leaf count measured beside it (0: skip)")
    ap.add_argument("--seed-keys", type=int, default=4_000_000,
                    help="read 21-mers for the learned-seed vs plain-search leg (0 skips it)")
    ap.add_argument("--learned-1g", action="store_true",
                    help="run only the 1 Gbp learned-seed workload (host index build takes minutes)")
    ap.add_argument("--no-1g", action="store_true",
                    help="leave the 1 Gbp learned-seed workload out of the default run")
    ap.add_argument("--big-ref-bp", type=int, default=1_000_000_000,
                    help="reference length of the --learned-1g workload")
    ap.add_argument("--big-reads", type=int, default=4_000_000,
                    help="150 bp reads of the --learned-1g workload")
    ap.add_argument("--random-access-calibration", action="store_true",
                    help="only time chains of random dependent 64-byte record loads (libhsfm_calib.so)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 1 GiB device buffer 3 times (scales FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


def workload(args):
    """Reference, bidirectional index bytes and read sets (cached on disk)."""
    from workloads import index_build as IB
    key = hashlib.sha1(f"v2-{args.ref_bp}-{args.reads}-{args.random_reads}-{args.locate_reads}-{args.uneven_reads}"
                       .encode()).hexdigest()[:12]
    os.makedirs(args.cache, exist_ok=True)
    path = os.path.join(args.cache, f"wl_{key}.npz")
    if os.path.exists(path):
        z = np.load(path, allow_pickle=False)
        log(f"workload cache hit {path}")
        return {k: z[k] for k in z.files}
    t0 = time.time()
    text, bounds = IB.synth_reference(0xC0FFEE, args.ref_bp, 8, gc=0.41, repeat_frac=0.20, n_units=200,
                                      unit_min=300, unit_max=6000, div_min=0.05, div_max=0.15,
                                      iupac_frac=0.005)
    log(f"reference {len(text)} bp in {time.time() - t0:.1f}s")
    t0 = time.time()
    import haystackfm_amd as hf
    idx = hf.build_index_bytes_gpu(text, bounds, sample_rate=32, lookup_depth=12, bidir=True)
    log(f"bidirectional index {len(idx) / 1e6:.0f} MB built on the GPU in {time.time() - t0:.1f}s")
    sq, so = IB.synth_reads(0x5EED3, text, bounds, args.reads, 150, sub=0.02, indel=0.002, both_strands=True,
                            n_random=args.random_reads, gc=0.41, shuffle=True)
    lq, lo = IB.synth_reads(0x5EED2, text, bounds, args.locate_reads, 150, sub=0.01, both_strands=True)
    # uneven reads: lognormal(ln 150, 1.0) lengths clipped to [32, 10000], 1.5% substitutions, shuffled
    uq, uo = IB.synth_reads(0x5EED4, text, bounds, args.uneven_reads, 150, sigma=1.0, len_min=32, len_max=10000,
                            sub=0.015, both_strands=True, shuffle=True)
    d = dict(text=text, bounds=bounds, index=np.frombuffer(idx, np.uint8), sq=sq, so=so, lq=lq, lo=lo, uq=uq, uo=uo)
    np.savez(path, **d)
    return d


ORACLE_BUILD = "oracle/liboracle.so (gcc -O3 -march=x86-64-v2)"


def oracle_module():
    """The oracle, compiled with -march=native for this host on first use."""
    global ORACLE_BUILD
    from oracle import oracle as O
    if O._lib is None:
        try:
            O.use_native_build(os.path.join(tempfile.gettempdir(), "hsfm_oracle_native"))
            ORACLE_BUILD = "gcc -O3 -march=native, compiled on this host"
        except (OSError, subprocess.CalledProcessError) as e:
            log(f"native oracle build failed ({e}); using oracle/liboracle.so")
    return O


def cpu_baseline(idx_bytes, sq, so, lq, lo, bounds, args, gpu_smem_sample, gpu_loc_sample):
    """Oracle leg: CPU restatement timed on a bounded sample + parity diff."""
    O = oracle_module()
    ob = O.OracleBidir.from_bytes(idx_bytes)
    cores = args.threads
    res = {}
    # parity: the sample's GPU output must equal the oracle's
    n = min(args.check_reads, len(so) - 1)
    exp = ob.find_mems_batch(qflat=sq[:int(so[n])], qoff=so[:n + 1], min_len=args.min_len, mode=0, threads=cores)
    ho, hits, pos = gpu_smem_sample
    ok = True
    for q in range(n):
        want = []
        for m in exp[q]:
            for lo_, hi_, _, _ in sorted(m.ivs.tolist()):
                size = hi_ - lo_
                k = min(size, args.max_hits)
                rows = np.arange(lo_, lo_ + k, dtype=np.uint32)
                p = ob.fwd.resolve_rows(rows)
                want.append((m.query_start, m.query_end, size, int(size > args.max_hits),
                             [O.ref_map(int(x), bounds) for x in p]))
        got = []
        for h in hits[ho[q]:ho[q + 1]]:
            po = int(h[4]) | (int(h[5]) << 32)
            got.append((int(h[0]), int(h[1]), int(h[2]), int(h[3]), [tuple(map(int, t)) for t in pos[po:po + int(h[6])]]))
        if got != want:
            ok = False
            log(f"SMEM parity mismatch at read {q}: gpu={got[:3]} oracle={want[:3]}")
            break
    nl = min(args.check_reads * 5, len(lo) - 1)
    eo, eh = ob.fwd.locate_batch(qflat=lq[:int(lo[nl])], qoff=lo[:nl + 1], threads=cores)
    go, gh = gpu_loc_sample
    if not (np.array_equal(eo, go) and np.array_equal(eh, gh)):
        ok = False
        log("locate parity mismatch")
    res["parity_ok"] = ok
    res["parity_sample"] = {"smem_reads": n, "locate_reads": nl}
    if args.no_cpu:
        return res
    # timing, as the reference's bench_utils.rs:34-43: one warm-up, then the
    # mean of three timed runs; the sample is sized so that the four runs of
    # each leg and thread count take about cpu-seconds / 4 together
    host = host_info(cores)

    def smem_run(k, th):
        # find_smems + locate of the first min(size, max_hits) rows of each interval,
        # the same work the GPU pipeline does (find -> resolve -> ref map)
        mems = ob.find_mems_batch(qflat=sq[:int(so[k])], qoff=so[:k + 1], min_len=args.min_len, mode=0,
                                  threads=th)
        rows = [np.arange(lo_, lo_ + min(hi_ - lo_, args.max_hits), dtype=np.uint32)
                for ms in mems for m in ms for lo_, hi_, _, _ in m.ivs.tolist()]
        if rows:
            pos = ob.fwd.resolve_rows(np.concatenate(rows), threads=th)
            np.searchsorted(bounds, pos, side="right")

    def loc_run(k, th):
        ob.fwd.locate_batch(qflat=lq[:int(lo[k])], qoff=lo[:k + 1], threads=th)

    def measure(fn, nmax, k0, th):
        budget = args.cpu_seconds / 16.0  # per run: 4 runs x 2 thread counts x 2 legs
        t0 = time.perf_counter()
        fn(k0, th)  # calibration run, also the warm-up
        t = time.perf_counter() - t0
        k = int(min(nmax, max(k0, k0 * budget / max(t, 1e-3))))
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            fn(k, th)
            ts.append(time.perf_counter() - t0)
        return k, k / float(np.mean(ts))

    for leg, fn, nmax, k0, what in (
            ("smem", smem_run, len(so) - 1, 64, "reads of the SMEM batch: oracle find_smems + resolve of the "
                                                "reported rows + ref map"),
            ("locate", loc_run, len(lo) - 1, 1024, "reads of the locate batch: oracle FmIndex::locate")):
        k1, v1 = measure(fn, nmax, k0, 1)
        kn, vn = measure(fn, nmax, k0 * cores, cores)
        res[leg] = {"value": vn, "unit": "reads/s", "cores": cores, "kind": "port",
                    "sample": f"first {kn} {what}, {cores} threads; mean of 3 runs after 1 warm-up",
                    "single_thread": {"value": v1, "unit": "reads/s", "cores": 1,
                                      "sample": f"first {k1} {what}, 1 thread (the reference's query paths are "
                                                "single-threaded); mean of 3 runs after 1 warm-up"},
                    "host": host}
    # algorithmic probes of the reference algorithm on a small sample
    O.probes_enable(True)
    ns = min(300, len(so) - 1)
    O.probes_reset()
    ob.find_mems_batch(qflat=sq[:int(so[ns])], qoff=so[:ns + 1], min_len=args.min_len, mode=0, threads=1)
    p = O.probes_get()
    res["smem_probes_per_read"] = p["occ_probes"] / ns
    nls = min(20000, len(lo) - 1)
    O.probes_reset()
    ob.fwd.locate_batch(qflat=lq[:int(lo[nls])], qoff=lo[:nls + 1], threads=1, counts_only=True)
    p = O.probes_get()
    res["locate_search_probes_per_read"] = p["occ_probes"] / nls
    O.probes_enable(False)
    return res


def host_info(threads):
    """The GPU box's host as the CPU baseline saw it."""
    model = None
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                model = ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return {"cpu_model": model, "host_logical_cpus": os.cpu_count(), "threads_used": threads,
            "threads_note": "the all-core runs use the box's CPU share (16 threads for one GPU), not every "
                            "logical CPU the host shows",
            "oracle_build": ORACLE_BUILD}


def random_record_ceiling(achieved_gbps):
    """The measured rate of random dependent loads in the SMEM kernel's access
    mix on this GPU (committed calibration run: 64-byte records within 1 GiB,
    every sixth load 16 bytes anywhere in a 16 GiB table) and the fraction of it
    `achieved_gbps` reaches; the 4 GiB record rate beside it."""
    path = os.path.join(ROOT, "profiles", "random_access_ceiling.json")
    try:
        with open(path) as f:
            rows = json.load(f)["rows"]
    except (OSError, ValueError, KeyError):
        return None
    mix = big = None
    for r in rows:
        if r.get("lanes") != 196608:
            continue
        if r.get("mode") == "mixed" and r.get("table_bytes") == 1 << 34 and r.get("window_bytes") == 1 << 30:
            mix = float(r["GBps_at_64B"])
        if r.get("mode") == "64B record" and r.get("table_bytes") == 1 << 32 and r.get("window_bytes") == 1 << 32:
            big = float(r["GBps_at_64B"])
    if mix is None:
        return None
    return {"GBps": mix, "frac": round(achieved_gbps / mix, 4),
            "what": "random dependent loads in the SMEM kernel's mix: 64-byte records within 1 GiB, every sixth load "
                    "16 bytes anywhere in 16 GiB (profiles/random_access_ceiling.md)",
            "records_4GiB_GBps": big}


def measured_traffic(kernel, args, nq, nl):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (copy-calibrated), when they were taken on this workload."""
    path = os.path.join(ROOT, "profiles", "traffic_smem_locate_100Mbp.json")
    try:
        with open(path) as f:
            doc = json.load(f)
    except (OSError, ValueError):
        return None
    w = doc.get("workload", {})
    if w.get("reference_bp") != args.ref_bp or w.get("smem_reads") != nq or w.get("locate_reads") != nl:
        return None
    k = doc.get("kernels", {}).get(kernel)
    return k["bytes_per_launch"] if k else None


def fetch_per_access(kernel, args, nq, nl, accesses):
    """The kernel's committed FETCH_SIZE per device-counted access against what
    the same counter reads for one lone random 64-byte record (the random-load
    calibration kernel): how many lone-record fetches each access costs."""
    path = os.path.join(ROOT, "profiles", "traffic_smem_locate_100Mbp.json")
    try:
        with open(path) as f:
            doc = json.load(f)
    except (OSError, ValueError):
        return None
    w = doc.get("workload", {})
    k = doc.get("kernels", {}).get(kernel)
    cal = doc.get("random_record_calibration")
    if (not k or not cal or not accesses or w.get("reference_bp") != args.ref_bp or w.get("smem_reads") != nq
            or w.get("locate_reads") != nl):
        return None
    per = k["fetch_kib"] * 1024.0 / accesses
    return {"fetch_counter_bytes_per_access": round(per, 1),
            "lone_random_record_counter_bytes": cal["fetch_bytes_per_record"],
            "records_fetched_per_access": round(per / cal["fetch_bytes_per_record"], 2),
            "source": "profiles/smem_locate_100Mbp_traffic.md"}


def measured_cache(kernel, args, nq, nl):
    """L2 hit rate and L1-to-L2 read requests per probe of `kernel` from the
    committed rocprofv3 TCC_HIT / TCC_MISS / TCP_TCC_READ_REQ pass, when it was
    taken on this workload (no counter for the Infinity Cache exists on gfx950)."""
    path = os.path.join(ROOT, "profiles", "l2_smem_locate_100Mbp.json")
    try:
        with open(path) as f:
            doc = json.load(f)
    except (OSError, ValueError):
        return None
    w = doc.get("workload", {})
    if w.get("reference_bp") != args.ref_bp or w.get("smem_reads") != nq or w.get("locate_reads") != nl:
        return None
    k = doc.get("kernels", {}).get(kernel)
    return {"l2_hit_rate": k["l2_hit_rate"], "l2_read_requests_per_probe": k["read_req_per_probe"],
            "source": "profiles/l2_hit_rate_100Mbp.md"} if k else None


def measured_traffic_1g(kernel, args, keys):
    """HBM bytes per launch of `kernel` on the 1 Gbp learned-seed workload from
    the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes (copy-calibrated)."""
    path = os.path.join(ROOT, "profiles", "traffic_learned_seed_1Gbp.json")
    try:
        with open(path) as f:
            doc = json.load(f)
    except (OSError, ValueError):
        return None
    w = doc.get("workload", {})
    if (w.get("reference_bp") != args.big_ref_bp or w.get("reads") != args.big_reads
            or w.get("seed_keys") != args.seed_keys or w.get("leaves") != args.lsi_leaves):
        return None
    k = doc.get("kernels", {}).get(kernel)
    return k["bytes_per_launch"] if k else None


def read_kmer_keys(flat, off, k, limit):
    """Keys of the exact-ACGT k-mer windows inside reads (first `limit`)."""
    import haystackfm_amd as hf
    keys = hf.kmer_keys(flat, k)
    start = np.arange(len(keys), dtype=np.int64)
    rid = np.searchsorted(off, start, side="right") - 1
    inside = start + k <= off[rid + 1]
    keys = keys[inside & (keys != np.uint64(0xFFFFFFFFFFFFFFFF))]
    if len(keys) > limit:
        keys = keys[np.linspace(0, len(keys) - 1, limit).astype(np.int64)]
    return keys


def read_window_keys(flat, off, k, limit):
    """One k-mer window per read (at a read-dependent offset) for the first
    `limit` reads whose window is exact ACGT: a cheap key set for large read
    batches, where enumerating every window costs tens of seconds."""
    nr = min(len(off) - 1, limit)
    ln = (off[1:nr + 1] - off[:nr]).astype(np.int64)
    r = np.arange(nr, dtype=np.int64)
    ok = ln >= k
    start = off[:nr].astype(np.int64) + (r * 7919) % np.maximum(ln - k + 1, 1)
    w = flat[start[ok][:, None] + np.arange(k)[None, :]]
    good = ((w >= 1) & (w <= 4)).all(axis=1)
    d = (w[good].astype(np.uint64) - np.uint64(1))
    keys = np.zeros(len(d), np.uint64)
    for t in range(k):
        keys = (keys << np.uint64(2)) | d[:, t]
    return keys


def uneven_leg(bidir, wl, args):
    """Uneven-length reads through the batched pipeline: batches of 64k reads
    handed to the callback in order, one HIP graph per window of batches
    (one persistent kernel over the window's work items), configuration
    picked by the autotuner on the first call; timed calls reuse it."""
    import torch
    import haystackfm_amd as hf
    flat, off = wl["uq"], wl["uo"]
    nq = len(off) - 1
    bounds = wl["bounds"]
    d_q = torch.from_numpy(flat).cuda()
    d_o = torch.from_numpy(off.astype(np.int64)).cuda()
    d_rb = torch.from_numpy(bounds.astype(np.int32)).cuda()
    torch.cuda.synchronize()
    run = lambda tune: bidir.find_mems_batched_device(d_q.data_ptr(), d_o.data_ptr(), nq, 19, hf.HSFM_MODE_SMEM,
                                                      d_rb.data_ptr(), len(bounds), 1024, batch_reads=65536,
                                                      autotune=tune, use_graph=True)
    t0 = time.perf_counter()
    tuned = run(True)
    tune_wall = time.perf_counter() - t0
    walls, stats = [], None
    for _ in range(args.steps):
        t0 = time.perf_counter()
        stats = run(False)
        walls.append(time.perf_counter() - t0)
    w = float(np.mean(walls))
    log(f"uneven: {nq} reads, {w * 1e3:.1f} ms/pass, stats {stats}, tuning {tuned['tune_ms']}")
    # the one-shot call on the same reads, for the cost of batching
    one = lambda: bidir.find_mems_device(d_q.data_ptr(), d_o.data_ptr(), nq, 19, hf.HSFM_MODE_SMEM,
                                         d_rb.data_ptr(), len(bounds), 1024)
    one()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one()
    torch.cuda.synchronize()
    w1 = (time.perf_counter() - t0) / args.steps
    log(f"uneven one-shot: {w1 * 1e3:.1f} ms")
    return {"reads": nq, "mean_len": round(float(off[-1]) / nq, 1), "min_len": 19, "batch_reads": 65536,
            "value": round(nq / w, 1), "unit": "reads/s", "ms_per_pass": round(w * 1e3, 2),
            "one_shot_ms": round(w1 * 1e3, 2), "batched_over_one_shot": round(w / w1, 3),
            "graph_ms": round(stats["total_ms"], 2), "batches": stats["batches"],
            "window_batches": stats["window_batches"], "windows": stats["windows"],
            "fallback_batches": stats["fallback_batches"], "records": stats["records"],
            "positions": stats["positions"],
            "tuned": {"main_kernel": ["thread/read ring 8", "thread/read ring 16", "thread/read ring 32",
                                      "wave/read"][stats["main_kernel"]],
                      "resolve_fanout": stats["resolve_fanout"],
                      "batch0_ms": [round(x, 3) for x in tuned["tune_ms"]],
                      "tuning_wall_s": round(tune_wall, 2)}}


def _digest_mem_result(bidir, nq, r):
    """sha1 of a device-resident MEM result (hit offsets, records, positions)."""
    import ctypes as C
    import haystackfm_amd as hf
    L = hf._lib()
    dev = bidir.fwd._dev[0]._h
    nh, npos, d_ho, d_hits, d_pos = r
    h = hashlib.sha1()
    for ptr, nbytes in ((d_ho, 8 * (nq + 1)), (d_hits, 32 * nh), (d_pos, 8 * npos)):
        buf = np.zeros(max(nbytes, 1), np.uint8)
        if nbytes:
            hf._check(L.hsfm_device_to_host(dev, buf.ctypes.data, C.c_void_p(ptr), nbytes), "d2h")
        h.update(buf[:nbytes].tobytes())
    return h.hexdigest()


def smem_seed_modes_leg(bidir, d_sq, d_so, nq, d_rb, nb, min_len, lsi, args, lsi_alt=None):
    """The SMEM pipeline (find -> resolve -> ref map) over the same reads with
    each restart seeding: plain backward search from the full interval, the
    packed 12/14/15-mer tables, learned 21-mer seeds with plain search as the
    fallback, learned seeds with the tables as the fallback, and (lsi_alt)
    learned seeds from an index with more leaves.  Every mode's result is
    digested and must be identical."""
    import torch
    import haystackfm_amd as hf
    modes = [("plain", hf.HSFM_SEED_PLAIN, None), ("tables", hf.HSFM_SEED_TABLES, None),
             ("learned", hf.HSFM_SEED_LEARNED, lsi), ("learned_tables", hf.HSFM_SEED_LEARNED_TABLES, lsi)]
    if lsi_alt is not None:
        modes.append(("learned_more_leaves", hf.HSFM_SEED_LEARNED, lsi_alt))
    out, digests = {}, set()
    for name, mode, l in modes:
        bidir.set_seed_mode(mode, l)

        def step():
            return bidir.find_mems_device(d_sq.data_ptr(), d_so.data_ptr(), nq, min_len, hf.HSFM_MODE_SMEM,
                                          d_rb.data_ptr(), nb, args.max_hits)
        for _ in range(max(1, args.warmup)):
            r = step()
        digests.add(_digest_mem_result(bidir, nq, r))
        torch.cuda.synchronize()
        kms, probes = [], 0
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
            st = bidir.last_stats()
            kms.append(st["main_kernel_ms"])
            probes = st["occ_probes"]
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / args.steps * 1e3
        out[name] = {"ms_per_step": round(ms, 3), "reads_per_s": round(nq / (ms * 1e-3), 1),
                     "main_kernel_ms": round(float(np.mean(kms)), 3), "loads_per_read": round(probes / nq, 2),
                     "deferred_reads": int(st["fallback_reads"]), "capacity_retries": int(st["capacity_retries"]),
                     "deferred_ms": round(st["fallback_kernel_ms"], 3)}
        log(f"SMEM seeding {name}: {ms:.2f} ms/step, main kernel {np.mean(kms):.2f} ms, {probes / nq:.1f} loads/read, "
            f"stats {st}")
    bidir.set_seed_mode(hf.HSFM_SEED_TABLES)
    same = len(digests) == 1
    out["identical_results"] = same
    if same:
        out["learned_vs_plain"] = round(out["plain"]["ms_per_step"] / out["learned"]["ms_per_step"], 3)
        out["learned_tables_vs_tables"] = round(out["tables"]["ms_per_step"] / out["learned_tables"]["ms_per_step"], 3)
        out["learned_tables_vs_plain"] = round(out["plain"]["ms_per_step"] / out["learned_tables"]["ms_per_step"], 3)
        out["learned_vs_tables"] = round(out["tables"]["ms_per_step"] / out["learned"]["ms_per_step"], 3)
        if "learned_more_leaves" in out:
            out["learned_more_leaves_vs_plain"] = round(out["plain"]["ms_per_step"] /
                                                        out["learned_more_leaves"]["ms_per_step"], 3)
    return out


def learned_seed_leg(bidir, flat, off, args, leaves=None, keys=None, keep=None):
    """Learned seed index vs plain exact search on the same read k-mers: both
    on the device, outputs compared bit for bit, kernel time by HIP events.
    keep (a list) receives the learned index, for the SMEM seeding leg."""
    import torch
    import haystackfm_amd as hf
    k = 21
    if keys is None:
        keys = read_kmer_keys(flat, off, k, args.seed_keys)
    t0 = time.time()
    leaves = leaves or args.lsi_leaves
    lsi = hf.LearnedSeedIndex(bidir, k=k, leaves=leaves, hidden=16, train_steps=400, sample=1 << 16)
    build_s = time.time() - t0
    info = lsi.info()
    if keep is not None:
        keep.append(lsi)
    dq = torch.from_numpy(keys.view(np.int64)).cuda()
    d_plain = torch.empty((len(keys), 4), dtype=torch.int32, device="cuda")
    d_lsi = torch.empty_like(d_plain)
    hf.kmer_search_device(bidir, k, dq.data_ptr(), len(keys), d_plain.data_ptr())
    lsi.seed_device(dq.data_ptr(), len(keys), d_lsi.data_ptr())
    torch.cuda.synchronize()
    plain_ms, lsi_ms, probes = [], [], 0
    # each call is about a millisecond: alternate the two kernels over 20 calls
    # each (or the step count, if larger) so that the ratio is not one sample's
    for _ in range(max(args.steps, 20)):
        plain_ms.append(hf.kmer_search_device(bidir, k, dq.data_ptr(), len(keys), d_plain.data_ptr())["kernel_ms"])
        st = lsi.seed_device(dq.data_ptr(), len(keys), d_lsi.data_ptr())
        lsi_ms.append(st["kernel_ms"])
        probes = st["probes"]
    torch.cuda.synchronize()
    exact = bool(torch.equal(d_plain, d_lsi))
    pm, lm = float(np.mean(plain_ms)), float(np.mean(lsi_ms))
    log(f"learned seed: {len(keys)} keys, plain {pm:.3f} ms, learned {lm:.3f} ms, exact {exact}, info {info}")
    return {"k": k, "keys": int(len(keys)), "found": int(d_lsi[:, 3].sum().item()), "bit_exact": exact,
            "plain_kernel_ms": round(pm, 4), "learned_kernel_ms": round(lm, 4),
            "speedup": round(pm / lm, 3) if exact and lm > 0 else None,
            "learned_probes_per_key": round(probes / max(1, len(keys)), 3),
            "index_keys": int(info["num_keys"]), "leaves": leaves, "hidden": 16,
            "max_err": int(info["max_err"]), "mean_window": round(info["mean_window"], 2),
            "train_mse": [info["mse_initial"], info["mse_final"]],
            "build_s": round(build_s, 2), "enum_ms": round(info["enum_ms"], 1),
            "train_ms": round(info["train_ms"], 1), "fit_ms": round(info["fit_ms"], 1),
            "fence_levels": int(info["fence_levels"]), "entries": int(info["entries"]),
            "variant_entries": int(info["variant_entries"]), "short_entries": int(info["short_entries"]),
            "extra_ms": round(info["extra_ms"], 1), "mem_seeding": bool(info["mem_seeding"])}


def workload_1g(args):
    """The learned-seed workload: 1 Gbp in 24 contigs of uniform random
    lengths, GC 0.41, 45% copies of 2,000 repeat families (300-6,000 bp) at
    5-25% divergence, 0.2% N; bidirectional index with SA sample rate 64 and no
    lookup table; reads of 150 bp from both strands with 1% substitutions."""
    from workloads import index_build as IB
    key = hashlib.sha1(f"1g-v1-{args.big_ref_bp}-{args.big_reads}".encode()).hexdigest()[:12]
    os.makedirs(args.cache, exist_ok=True)
    path = os.path.join(args.cache, f"wl_{key}.npz")
    if os.path.exists(path):
        z = np.load(path, allow_pickle=False)
        log(f"workload cache hit {path}")
        return z["bounds"], z["index"].tobytes(), z["sq"], z["so"]
    os.environ.setdefault("HSFM_BUILD_PROGRESS", "1")
    t0 = time.time()
    text, bounds = IB.synth_reference(0xC5EED, args.big_ref_bp, 24, gc=0.41, repeat_frac=0.45, n_units=2000,
                                      unit_min=300, unit_max=6000, div_min=0.05, div_max=0.25,
                                      iupac_frac=0.002, iupac_codes=[5])
    log(f"reference {len(text)} bp in {time.time() - t0:.1f}s")
    t0 = time.time()
    import haystackfm_amd as hf
    idx = hf.build_index_bytes_gpu(text, bounds, sample_rate=64, lookup_depth=0, bidir=True)
    log(f"bidirectional index {len(idx) / 1e6:.0f} MB built on the GPU in {time.time() - t0:.1f}s")
    sq, so = IB.synth_reads(0x5EED5, text, bounds, args.big_reads, 150, sub=0.01, both_strands=True, shuffle=True)
    log(f"{args.big_reads} reads synthesised")
    del text
    np.savez(path, bounds=bounds, index=np.frombuffer(idx, np.uint8), sq=sq, so=so)
    return bounds, idx, sq, so


def keys_to_codes(keys, k):
    """Base codes (A=1 .. T=4) of 2-bit keys, first base most significant."""
    sh = np.arange(k - 1, -1, -1, dtype=np.uint64) * np.uint64(2)
    return ((keys[:, None] >> sh[None, :]) & np.uint64(3)).astype(np.uint8) + 1


def exact_kmer_interval(O, ob, codes):
    """(1, fwd_lo, fwd_hi, rev_lo) of exact left extensions from the full
    interval, last code first (the oracle's extend_left), or (0,) when absent."""
    n = ob.fwd.text_len
    iv = (0, n, 0, n)
    for c in codes[::-1]:
        iv = O.extend_left(ob.fwd, iv, int(c))
        if iv is None or iv[1] <= iv[0]:
            return (0,)
    return (1, iv[0], iv[1], iv[2])


def main_1g(args):
    print(json.dumps(run_1g(args)), flush=True)


def run_1g(args):
    """Learned seed index vs plain exact search on the 1 Gbp workload, plus the
    SMEM pipeline (min_len 21) over its reads with each restart seeding; the
    JSON object of its line."""
    import torch
    import haystackfm_amd as hf
    bounds, idx, sq, so = workload_1g(args)
    t0 = time.time()
    bidir = hf.BidirFmIndex.from_bytes(idx)
    bidir._handle()
    log(f"index upload {time.time() - t0:.1f}s")
    k = 21
    keys = read_window_keys(sq, so, k, args.seed_keys)
    log(f"{len(keys)} read 21-mers")
    keep = []
    lseed = learned_seed_leg(bidir, sq, so, args, keys=keys, keep=keep)
    keep_alt = []
    lseed_alt = (learned_seed_leg(bidir, sq, so, args, leaves=args.lsi_leaves_alt, keys=keys, keep=keep_alt)
                 if args.lsi_leaves_alt else None)

    # SMEM pipeline over the reads (min_len 21)
    min_len = 21
    dev = torch.device("cuda:0")
    d_sq = torch.from_numpy(sq).to(dev)
    d_so = torch.from_numpy(so.astype(np.int64)).to(dev)
    d_rb = torch.from_numpy(bounds.astype(np.int32)).to(dev)
    nq = len(so) - 1
    torch.cuda.synchronize()

    def smem_step():
        return bidir.find_mems_device(d_sq.data_ptr(), d_so.data_ptr(), nq, min_len, hf.HSFM_MODE_SMEM,
                                      d_rb.data_ptr(), len(bounds), args.max_hits)

    for i in range(args.warmup):
        smem_step()
        log(f"SMEM warmup {i} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kms = []
    for _ in range(args.steps):
        r = smem_step()
        kms.append(bidir.last_stats()["main_kernel_ms"])
    torch.cuda.synchronize()
    smem_s = (time.perf_counter() - t0) / args.steps
    smem_stats = bidir.last_stats()
    log(f"SMEM step {smem_s * 1e3:.1f} ms, {r[0]} MEM records, {r[1]} positions")
    seeding = smem_seed_modes_leg(bidir, d_sq, d_so, nq, d_rb, len(bounds), min_len, keep[0], args,
                                  keep_alt[0] if keep_alt else None)

    # parity samples: SMEM lists and 21-mer intervals against the oracle
    O = oracle_module()
    ob = O.OracleBidir.from_bytes(idx)
    del idx
    n = min(args.check_reads, nq)
    got = bidir.find_mems_batch((sq[:int(so[n])], so[:n + 1]), min_len, hf.HSFM_MODE_SMEM, bounds, args.max_hits)
    exp = ob.find_mems_batch(qflat=sq[:int(so[n])], qoff=so[:n + 1], min_len=min_len, mode=0, threads=args.threads)
    ho, hits, pos = got
    ok = True
    for q in range(n):
        want = []
        for m in exp[q]:
            for lo_, hi_, _, _ in sorted(m.ivs.tolist()):
                size = hi_ - lo_
                nrows = min(size, args.max_hits)
                p = ob.fwd.resolve_rows(np.arange(lo_, lo_ + nrows, dtype=np.uint32))
                want.append((m.query_start, m.query_end, size, int(size > args.max_hits),
                             [O.ref_map(int(x), bounds) for x in p]))
        have = []
        for h in hits[ho[q]:ho[q + 1]]:
            po = int(h[4]) | (int(h[5]) << 32)
            have.append((int(h[0]), int(h[1]), int(h[2]), int(h[3]),
                         [tuple(map(int, t)) for t in pos[po:po + int(h[6])]]))
        if have != want:
            ok = False
            log(f"SMEM parity mismatch at read {q}")
            break
    ks = keys[np.linspace(0, len(keys) - 1, min(len(keys), 2000)).astype(np.int64)]
    d_keys = torch.from_numpy(ks.view(np.int64)).cuda()
    d_out = torch.empty((len(ks), 4), dtype=torch.int32, device="cuda")
    hf.kmer_search_device(bidir, k, d_keys.data_ptr(), len(ks), d_out.data_ptr())
    gout = d_out.cpu().numpy().view(np.uint32)
    codes = keys_to_codes(ks, k)
    for i in range(len(ks)):
        want = exact_kmer_interval(O, ob, codes[i])
        have = (int(gout[i, 3]), int(gout[i, 0]), int(gout[i, 1]), int(gout[i, 2]))
        if have[:len(want)] != want:
            ok = False
            log(f"21-mer interval mismatch at key {i}")
            break
    ok = ok and bool(lseed["bit_exact"]) and (lseed_alt is None or bool(lseed_alt["bit_exact"]))

    # CPU baseline: the oracle's plain backward search of the same 21-mers
    cpu = None
    if not args.no_cpu:
        cores = args.threads
        allc = keys_to_codes(keys, k)

        def run(m):
            flat = np.ascontiguousarray(allc[:m]).reshape(-1)
            off = np.arange(m + 1, dtype=np.uint64) * np.uint64(k)
            t = time.perf_counter()
            ob.fwd.locate_batch(qflat=flat, qoff=off, threads=cores, counts_only=True)
            return time.perf_counter() - t

        m = min(len(keys), 20000 * cores)
        t = run(m)
        m2 = int(min(len(keys), max(m, m * args.cpu_seconds / max(t, 1e-3))))
        t2 = run(m2)
        cpu = {"value": m2 / t2, "unit": "keys/s", "cores": cores, "kind": "port",
               "sample": f"first {m2} of the read 21-mers: oracle backward_search + count, {cores} threads"}
    sp = lseed["speedup"]
    probes = lseed["learned_probes_per_key"] * lseed["keys"]
    lms = lseed["learned_kernel_ms"]
    lroof = None
    if lms > 0:
        a = probes * 64.0 / (lms * 1e-3) / 1e9
        lroof = {"bound": "hbm", "achieved": round(a, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(a / HBM_PEAK_GBS, 4), "traffic": measured_traffic_1g("k_lsi_lookup", args, keys),
                 "kernel": "k_lsi_lookup", "kernel_ms": lms,
                 "bytes_per_key": round(lseed["learned_probes_per_key"] * 64.0, 1)}
    # the learned-seed SMEM step's main kernel: 64 B x its device-counted
    # accesses (learned-index rounds in 64-byte units, interval probes, text
    # and suffix-array loads) over its HIP-event time
    lm = seeding["learned"]
    roof = None
    if lm["main_kernel_ms"] > 0:
        a = lm["loads_per_read"] * nq * 64.0 / (lm["main_kernel_ms"] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(a, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(a / HBM_PEAK_GBS, 4), "traffic": measured_traffic_1g("k_mem_persist_lsi", args, keys),
                "kernel": "k_mem_persist<16,256,true>", "kernel_ms": lm["main_kernel_ms"],
                "bytes_per_read": round(lm["loads_per_read"] * 64.0, 1),
                "bytes_definition": "64 B x the kernel's device-counted random accesses (see the 100 Mbp line)"}
    same = bool(seeding.get("identical_results"))
    out = {
        "metric": "learned-seed vs plain backward-search SMEM speedup (1 Gbp ref, 150 bp reads, min_len 21) bit-exact",
        "value": seeding.get("learned_vs_plain") if ok and same else None,
        "unit": "x",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": seeding["learned"]["ms_per_step"],
        "lookup": {"metric": "learned lookup vs plain 21-step exact search speedup (read 21-mers) bit-exact",
                   "value": sp if ok else None, "unit": "x", "kernel_ms": lms, "roofline": lroof},
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+u64", "data": "synthetic",
        "config": {"workload": f"learned_seed: {args.big_ref_bp / 1e9:g} Gbp 24-contig reference (45% repeats from "
                               "2000 families at 5-25% divergence, GC 0.41, 0.2% N), bidirectional index rate 64; "
                               f"{args.big_reads} x 150 bp reads (1% subs, both strands); {lseed['keys']} read "
                               f"21-mers, {args.lsi_leaves} leaves; SMEM min_len 21, max_hits {args.max_hits}",
                   "reference_bp": int(args.big_ref_bp), "reads": nq, "read_len": 150, "min_len": min_len},
        "roofline": roof,
        "cpu_baseline": cpu,
        "parity": ok,
        "parity_sample": {"smem_reads": n, "kmers": int(len(ks))},
        "learned_seed": lseed,
        "learned_seed_more_leaves": lseed_alt,
        "smem_seeding": seeding,
        "smem": {"value": round(nq / smem_s, 1) if ok else None, "unit": "reads/s", "ms_per_step": round(smem_s * 1e3, 3),
                 "main_kernel_ms": round(float(np.mean(kms)), 3),
                 "stats": {k_: (round(v, 3) if isinstance(v, float) else v) for k_, v in smem_stats.items()}},
    }
    return out


def random_access_calibration():
    """Records per second served to chains of random dependent loads (the
    access pattern of a backward-search step) by libhsfm_calib.so: whole 64-byte
    records, 16 bytes of a record, lo/hi record pairs and the SMEM kernel's mix,
    over tables of 256 MiB to 32 GiB, at the persistent kernels' lane count (and
    4x it for 4 GiB)."""
    import ctypes as C
    lib = C.CDLL(os.path.join(ROOT, "haystackfm_amd", "libhsfm_calib.so"))
    fn = lib.hsfm_calib_random_records
    fn.restype = C.c_int
    fn.argtypes = [C.c_int, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, C.c_int,
                   C.POINTER(C.c_double), C.POINTER(C.c_double)]
    res = []
    G = 1 << 30
    for table, window, lanes, mode in [(4 * G, 4 * G, 196608, 0), (4 * G, 4 * G, 196608, 1), (4 * G, 4 * G, 196608, 2),
                                       (4 * G, 4 * G, 196608, 3), (4 * G, 4 * G, 786432, 0), (4 * G, 4 * G, 786432, 2),
                                       (G // 4, G // 4, 196608, 0), (G, G, 196608, 0), (2 * G, 2 * G, 196608, 0),
                                       (16 * G, 16 * G, 196608, 0), (G, G, 196608, 1), (16 * G, 16 * G, 196608, 1),
                                       (16 * G, G, 196608, 4), (G, G, 196608, 4)]:
        ms, recs = C.c_double(), C.c_double()
        rc = fn(0, table, window, lanes, 200, mode, 3, C.byref(ms), C.byref(recs))
        if rc != 0:
            raise RuntimeError(f"hsfm_calib_random_records failed ({rc})")
        row = {"mode": ["64B record", "16B of a record", "lo/hi record pair", "64B record, quad-cooperative",
                        "mixed"][mode],
               "table_bytes": table, "window_bytes": window,
               "lanes": lanes, "ms": round(ms.value, 3), "G_records_per_s": round(recs.value / ms.value / 1e6, 2),
               "GBps_at_64B": round(recs.value * 64 / ms.value / 1e6, 1)}
        log(row)
        res.append(row)
    print(json.dumps({"random_access_calibration": res}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (reads shard across GPUs by "
                                   "replicating the index, see DESIGN.md)"}))
        sys.exit(2)
    if args.random_access_calibration:
        random_access_calibration()
        return
    import torch
    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.int32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": src.numel() * 4, "copies": 3}), flush=True)
        return
    if args.learned_1g:
        main_1g(args)
        return
    import haystackfm_amd as hf

    wl = workload(args)
    idx_bytes = wl["index"].tobytes()
    t0 = time.time()
    bidir = hf.BidirFmIndex.from_bytes(idx_bytes)
    bidir._handle()
    log(f"index upload {time.time() - t0:.1f}s")
    dev = torch.device("cuda:0")
    d_sq = torch.from_numpy(wl["sq"]).to(dev)
    d_so = torch.from_numpy(wl["so"].astype(np.int64)).to(dev)
    d_lq = torch.from_numpy(wl["lq"]).to(dev)
    d_lo = torch.from_numpy(wl["lo"].astype(np.int64)).to(dev)
    bounds = wl["bounds"]
    d_rb = torch.from_numpy(bounds.astype(np.int32)).to(dev)
    nq = len(wl["so"]) - 1
    nl = len(wl["lo"]) - 1
    torch.cuda.synchronize()

    def smem_step():
        return bidir.find_mems_device(d_sq.data_ptr(), d_so.data_ptr(), nq, args.min_len, hf.HSFM_MODE_SMEM,
                                      d_rb.data_ptr(), len(bounds), args.max_hits)

    def locate_step():
        return bidir.fwd.locate_device(d_lq.data_ptr(), d_lo.data_ptr(), nl)

    for i in range(args.warmup):
        smem_step()
        log(f"warmup {i} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kernel_ms, resolve_ms = [], []
    for i in range(args.steps):
        r = smem_step()
        st = bidir.last_stats()
        kernel_ms.append(st["main_kernel_ms"])
        resolve_ms.append(st["resolve_kernel_ms"])
    torch.cuda.synchronize()
    smem_s = (time.perf_counter() - t0) / args.steps
    smem_stats = bidir.last_stats()
    log(f"SMEM step {smem_s * 1e3:.1f} ms, {r[0]} MEM records, {r[1]} positions, stats {smem_stats}")

    locate_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lk = []
    for i in range(args.steps):
        lr = locate_step()
        lk.append(bidir.fwd.last_stats()["main_kernel_ms"])
    torch.cuda.synchronize()
    loc_s = (time.perf_counter() - t0) / args.steps
    loc_stats = bidir.fwd.last_stats()
    log(f"locate step {loc_s * 1e3:.1f} ms, {lr[0]} hits, stats {loc_stats}")

    # sample outputs for the parity diff (host entry points, same kernels)
    n = min(args.check_reads, nq)
    sample = bidir.find_mems_batch((wl["sq"][:int(wl["so"][n])], wl["so"][:n + 1]), args.min_len,
                                   hf.HSFM_MODE_SMEM, bounds, args.max_hits)
    nls = min(args.check_reads * 5, nl)
    lsample = bidir.fwd.locate_batch((wl["lq"][:int(wl["lo"][nls])], wl["lo"][:nls + 1]))
    lseed = learned_seed_leg(bidir, wl["sq"], wl["so"], args) if args.seed_keys else None
    lseed_alt = (learned_seed_leg(bidir, wl["sq"], wl["so"], args, leaves=args.lsi_leaves_alt)
                 if args.seed_keys and args.lsi_leaves_alt else None)
    uneven = uneven_leg(bidir, wl, args) if args.uneven_reads else None
    cpu = cpu_baseline(idx_bytes, wl["sq"], wl["so"], wl["lq"], wl["lo"], bounds, args, sample, lsample)

    value = nq / smem_s
    kms = float(np.mean(kernel_ms))
    roof = None
    probes = smem_stats.get("occ_probes", 0)
    if probes and kms > 0:
        # algorithmic bytes: the 64 B Occ records (and seed-table sectors) the
        # kernel's algorithm must fetch, counted on the device, per launch
        bytes_per_launch = probes * 64.0
        achieved = bytes_per_launch / (kms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic("k_mem_persist", args, nq, nl),
                "kernel": "k_mem_persist", "kernel_ms": round(kms, 3),
                "bytes_per_read": round(bytes_per_launch / nq, 1),
                "bytes_definition": BYTES_DEFINITION,
                "measured_random_record_ceiling": random_record_ceiling(achieved),
                "cache": measured_cache("k_mem_persist", args, nq, nl),
                "fetch_vs_random_record": fetch_per_access("k_mem_persist", args, nq, nl, probes)}
        if roof["traffic"]:
            roof["traffic_frac"] = round(roof["traffic"] / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    loc_roof = None
    lkms = float(np.mean(lk))
    if loc_stats.get("occ_probes") and lkms > 0:
        a = loc_stats["occ_probes"] * 64.0 / (lkms * 1e-3) / 1e9
        loc_roof = {"bound": "hbm", "achieved": round(a, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(a / HBM_PEAK_GBS, 4), "traffic": measured_traffic("k_locate_search", args, nq, nl),
                    "kernel": "k_locate_search", "kernel_ms": round(lkms, 3),
                    "bytes_per_read": round(loc_stats["occ_probes"] * 64.0 / nl, 1),
                    "bytes_definition": BYTES_DEFINITION,
                    "measured_random_record_ceiling": random_record_ceiling(a),
                    "cache": measured_cache("k_locate_search", args, nq, nl),
                    "fetch_vs_random_record": fetch_per_access("k_locate_search", args, nq, nl,
                                                               loc_stats["occ_probes"])}
        if loc_roof["traffic"]:
            loc_roof["traffic_frac"] = round(loc_roof["traffic"] / (lkms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        if cpu.get("locate_search_probes_per_read"):
            # the survey's definition: the reference algorithm's probes (counted by
            # the oracle) x 64 B per read, over the whole locate step
            b8 = cpu["locate_search_probes_per_read"] * 64.0 * nl
            loc_roof["reference_algorithm_frac"] = round(b8 / loc_s / 1e9 / HBM_PEAK_GBS, 4)
    cpu_line = None
    if "smem" in cpu:
        cpu_line = cpu["smem"]
    out = {
        "metric": METRIC,
        "value": round(value, 1) if cpu["parity_ok"] else None,
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(smem_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "smem_locate: 100 Mbp 8-contig IUPAC reference (20% repeats, 0.5% ambiguity), "
                               "bidirectional index rate 32 lookup 12; 1.0M x 150 bp reads (2% subs, 0.2% indels, "
                               "both strands) + 0.1M random; SMEM min_len 18 -> resolve -> ref map, max_hits 1024",
                   "reference_bp": int(args.ref_bp), "reads": nq, "read_len": 150, "min_len": args.min_len,
                   "max_hits_per_mem": args.max_hits},
        "roofline": roof,
        "cpu_baseline": cpu_line,
        "parity": cpu["parity_ok"],
        "parity_sample": cpu["parity_sample"],
        "locate": {"value": round(nl / loc_s, 1) if cpu["parity_ok"] else None, "unit": "reads/s",
                   "ms_per_step": round(loc_s * 1e3, 3), "reads": nl, "search_kernel_ms": round(float(np.mean(lk)), 3),
                   "cpu_baseline": cpu.get("locate"), "roofline": loc_roof,
                   "stats": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in loc_stats.items()},
                   "reference_algorithm_probes_per_read": cpu.get("locate_search_probes_per_read")},
        "smem_stats": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in smem_stats.items()},
        "reference_algorithm_probes_per_read": cpu.get("smem_probes_per_read"),
        "learned_seed_speedup": None,  # the 1 Gbp learned-seed SMEM speedup, below
        "learned_seed_lookup_speedup": lseed["speedup"] if lseed else None,
        "learned_seed": lseed,
        "learned_seed_more_leaves": lseed_alt,
        "uneven_batched": uneven,
    }
    if not args.no_1g:
        del bidir
        big = run_1g(args)
        out["learned_seed_1g"] = {k: big.get(k) for k in ("config", "value", "ms_per_step", "lookup", "roofline",
                                                          "parity", "parity_sample", "cpu_baseline", "learned_seed",
                                                          "smem_seeding", "smem")}
        out["learned_seed_1g"]["value_metric"] = big.get("metric")
        out["learned_seed_speedup"] = big.get("value")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
