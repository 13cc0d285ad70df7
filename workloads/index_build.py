"""Host-side index construction and seeded workload synthesis (libhsfm_build.so).

build_index_bytes / build_bidir_bytes produce the byte streams of the
reference's FmIndex::build_cpu + to_bytes and BidirFmIndex::build_cpu +
to_bytes (src/fm_index/mod.rs:117-195, serialize.rs:13-26,
bidir_index.rs:53-80,183-192).  synth_reference / synth_reads generate the
seeded synthetic references and read sets of the benchmark workloads.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "libhsfm_build.so")
_lib = None
IUPAC_AMBIG = [6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 5]  # R Y S W K M B D H V N


def build_host_lib(verbose: bool = False) -> str:
    """libhsfm_build.so from hsfm_build.cpp (host code only)."""
    tmp = _PATH + ".tmp"
    cmd = [os.environ.get("CXX", "g++"), "-O3", "-std=c++17", "-fPIC", "-shared", "-march=x86-64-v2",
           "-Wall", "-o", tmp, os.path.join(_HERE, "hsfm_build.cpp"), "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("host library build failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, _PATH)
    return _PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            build_host_lib()
        L = C.CDLL(_PATH)
        u8p, u32p, u64p = C.POINTER(C.c_uint8), C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)
        for name in ("hsfm_build_index", "hsfm_build_bidir"):
            f = getattr(L, name)
            f.argtypes = [u8p, C.c_uint64, u32p, C.c_uint32, C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32,
                          C.c_uint8, C.c_int, C.POINTER(u8p), u64p]
            f.restype = C.c_int
        L.hsfm_build_free.argtypes = [C.c_void_p]
        L.hsfm_synth_reference.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_double, C.c_double, C.c_uint32,
                                           C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double, u8p,
                                           C.c_uint32, C.POINTER(u8p), u64p, C.POINTER(u32p)]
        L.hsfm_synth_reads.argtypes = [C.c_uint64, u8p, C.c_uint64, u32p, C.c_uint32, C.c_uint64, C.c_uint32,
                                       C.c_double, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_int,
                                       C.c_uint64, C.c_double, C.c_int, C.POINTER(u8p), C.POINTER(u64p)]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _build(fn, text, bounds, headers, sample_rate, lookup_depth, encoding, tag, threads) -> bytes:
    text = np.ascontiguousarray(text, np.uint8)
    bounds = np.ascontiguousarray(bounds, np.uint32)
    hb = b"".join(h.encode() + b"\0" for h in headers) if headers is not None else None
    out = C.POINTER(C.c_uint8)()
    n = C.c_uint64()
    rc = fn(_p(text, C.c_uint8), len(text), _p(bounds, C.c_uint32), len(bounds), hb, sample_rate, lookup_depth,
            encoding, tag, threads, C.byref(out), C.byref(n))
    if rc != 0:
        raise ValueError(f"index construction failed ({rc})")
    # a numpy view copies lengths beyond 2 GiB (ctypes.string_at takes an int size)
    data = np.ctypeslib.as_array(out, shape=(n.value,)).tobytes() if n.value else b""
    lib().hsfm_build_free(C.cast(out, C.c_void_p))
    return data


def build_index_bytes(text, bounds, *, headers=None, sample_rate=32, lookup_depth=0, encoding=0, tag=0,
                      threads=8) -> bytes:
    return _build(lib().hsfm_build_index, text, bounds, headers, sample_rate, lookup_depth, encoding, tag, threads)


def build_bidir_bytes(text, bounds, *, headers=None, sample_rate=32, lookup_depth=0, encoding=0, tag=0,
                      threads=8) -> bytes:
    return _build(lib().hsfm_build_bidir, text, bounds, headers, sample_rate, lookup_depth, encoding, tag, threads)


def synth_reference(seed, total, ncontig=1, gc=0.5, repeat_frac=0.0, n_units=0, unit_min=300, unit_max=6000,
                    div_min=0.05, div_max=0.15, iupac_frac=0.0, iupac_codes=None):
    codes = np.array(iupac_codes if iupac_codes is not None else IUPAC_AMBIG, np.uint8)
    tp, n, bp = C.POINTER(C.c_uint8)(), C.c_uint64(), C.POINTER(C.c_uint32)()
    rc = lib().hsfm_synth_reference(seed, total, ncontig, gc, repeat_frac, n_units, unit_min, unit_max, div_min,
                                    div_max, iupac_frac, _p(codes, C.c_uint8), len(codes), C.byref(tp), C.byref(n),
                                    C.byref(bp))
    if rc != 0:
        raise ValueError("reference synthesis failed")
    text = np.ctypeslib.as_array(tp, shape=(n.value,)).copy()
    bounds = np.ctypeslib.as_array(bp, shape=(ncontig,)).copy()
    lib().hsfm_build_free(C.cast(tp, C.c_void_p))
    lib().hsfm_build_free(C.cast(bp, C.c_void_p))
    return text, bounds


def synth_reads(seed, text, bounds, nreads, length, *, sigma=0.0, len_min=0, len_max=0, sub=0.0, indel=0.0,
                both_strands=False, n_random=0, gc=0.5, shuffle=False):
    text = np.ascontiguousarray(text, np.uint8)
    bounds = np.ascontiguousarray(bounds, np.uint32)
    fp, op = C.POINTER(C.c_uint8)(), C.POINTER(C.c_uint64)()
    rc = lib().hsfm_synth_reads(seed, _p(text, C.c_uint8), len(text), _p(bounds, C.c_uint32), len(bounds), nreads,
                                length, sigma, len_min, len_max or length, sub, indel, 1 if both_strands else 0,
                                n_random, gc, 1 if shuffle else 0, C.byref(fp), C.byref(op))
    if rc != 0:
        raise ValueError("read synthesis failed")
    nq = nreads + n_random
    off = np.ctypeslib.as_array(op, shape=(nq + 1,)).copy()
    flat = np.ctypeslib.as_array(fp, shape=(max(int(off[-1]), 1),)).copy()
    lib().hsfm_build_free(C.cast(fp, C.c_void_p))
    lib().hsfm_build_free(C.cast(op, C.c_void_p))
    return flat, off
