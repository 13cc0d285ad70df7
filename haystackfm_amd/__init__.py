"""haystackfm_amd -- MI355X (gfx950) batched query path of haystackfm.

Python mirror of the reference crate's GPU query API (sriram98v/haystackfm):

    FmIndex.from_bytes(data)                          src/fm_index/serialize.rs:32-52
    FmIndex.locate_gpu(queries)                       src/fm_index/query.rs:367-384
    locate_batch_gpu(ctx, index, queries)             src/gpu/locate.rs:52-56
    BidirFmIndex.from_bytes(data)                     src/fm_index/bidir_index.rs:195-210
    BidirFmIndex.find_smems_gpu / find_mems_gpu       src/fm_index/bidir_index.rs:225-276
    MemHit                                            src/gpu/mem_hit.rs:6-19
    Alphabet / IupacDna / ExactDna / AlphabetFns      src/alphabet.rs:203-305
    GpuContext, context_cache.get_or_init             src/gpu/mod.rs:19-66, context_cache.rs:5-19

Every query runs on the GPU through libhsfm.so (hand-written HIP kernels behind
the C ABI in include/hsfm.h).  There is no CPU fallback: without the library
or a GPU, the GPU methods raise GpuError, as the reference does without an
adapter.  Queries are encoded IUPAC codes ($=0, A..T=1..4, N=5, R..V=6..15).
"""
from __future__ import annotations

import ctypes as C
import os
import struct
import threading
from dataclasses import dataclass, field
from typing import Callable, Iterable, Sequence

import numpy as np

__all__ = [
    "FmIndex", "BidirFmIndex", "MemHit", "GpuContext", "FmIndexError", "GpuError",
    "DeserializeError", "Alphabet", "AlphabetFns", "IupacDna", "ExactDna", "DnaSequence",
    "locate_batch_gpu", "context_cache", "encode", "alphabet_fns_from_tag", "lib_path",
    "LearnedSeedIndex", "kmer_keys", "kmer_search", "kmer_search_device", "lsi_gradient", "build_index_bytes_gpu",
    "reverse_complement", "reverse_complement_device",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libhsfm.so")


# --------------------------------------------------------------------------
# errors (src/error.rs)
# --------------------------------------------------------------------------
class FmIndexError(Exception):
    """FmIndexError (src/error.rs:5-30)."""


class GpuError(FmIndexError):
    """FmIndexError::GpuError (src/error.rs:26-29)."""


class DeserializeError(FmIndexError):
    """FmIndexError::DeserializeError (src/error.rs:18-20)."""


# --------------------------------------------------------------------------
# alphabet (src/alphabet.rs)
# --------------------------------------------------------------------------
SENTINEL, A, C_, G, T, N = 0, 1, 2, 3, 4, 5
_SYMBOLS = "$ACGTNRYSWKMBDHV"
# alphabet.rs:174-193 compatible_symbols, in the reference's order
_COMPAT = [
    [], [1, 5, 6, 9, 11, 13, 14, 15], [2, 5, 7, 8, 11, 12, 14, 15], [3, 5, 6, 8, 10, 12, 13, 15],
    [4, 5, 7, 9, 10, 12, 13, 14], list(range(1, 16)), [1, 3, 5, 6, 8, 9, 10, 11, 12, 13, 14, 15],
    [2, 4, 5, 7, 8, 9, 10, 11, 12, 13, 14, 15], [2, 3, 5, 6, 7, 8, 10, 11, 12, 13, 14, 15],
    [1, 4, 5, 6, 7, 9, 10, 11, 12, 13, 14, 15], [3, 4, 5, 6, 7, 8, 9, 10, 12, 13, 14, 15],
    [1, 2, 5, 6, 7, 8, 9, 11, 12, 13, 14, 15], [2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15],
    [1, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15], [1, 2, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15],
    [1, 2, 3, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15],
]


def compatible_symbols(code: int) -> list[int]:
    """alphabet.rs:174-193"""
    return list(_COMPAT[code]) if 0 <= code < 16 else []


def _exact_compatible(code: int) -> list[int]:
    """ExactDna::compatible (alphabet.rs:286-294)"""
    return [code] if 1 <= code <= 4 else []


@dataclass(frozen=True)
class AlphabetFns:
    """AlphabetFns (alphabet.rs:203-212)."""
    compatible_fn: Callable[[int], Sequence[int]]
    core_symbols: tuple
    tag: int

    def compat_masks(self) -> np.ndarray:
        """compat_mask[b] = bitmask of compatible_fn(b) for every byte b (C ABI field)."""
        m = np.zeros(256, np.uint16)
        for b in range(256):
            for s in self.compatible_fn(b):
                if 0 <= s < 16:
                    m[b] |= 1 << s
        return m


class Alphabet:
    """Alphabet trait (alphabet.rs:232-235): subclasses implement fns()."""

    @classmethod
    def fns(cls) -> AlphabetFns:
        raise NotImplementedError


class IupacDna(Alphabet):
    """alphabet.rs:262-272"""

    @classmethod
    def fns(cls) -> AlphabetFns:
        return AlphabetFns(compatible_symbols, (A, C_, G, T), 0)


class ExactDna(Alphabet):
    """alphabet.rs:282-305"""

    @classmethod
    def fns(cls) -> AlphabetFns:
        return AlphabetFns(_exact_compatible, (A, C_, G, T), 1)


def alphabet_fns_from_tag(tag: int):
    """alphabet.rs:240-254: None for unknown tags."""
    return {0: IupacDna.fns(), 1: ExactDna.fns()}.get(tag)


_ENC = np.full(256, 255, np.uint8)
for _i, _ch in enumerate(_SYMBOLS):
    _ENC[ord(_ch)] = _i
    _ENC[ord(_ch.lower())] = _i
_ENC[ord("U")] = T
_ENC[ord("u")] = T


def encode(s: str | bytes) -> np.ndarray:
    """encode_byte over a string (alphabet.rs:54-109); raises on invalid bytes."""
    b = np.frombuffer(s.encode() if isinstance(s, str) else bytes(s), np.uint8)
    out = _ENC[b]
    if (out == 255).any():
        i = int(np.argmax(out == 255))
        raise FmIndexError(f"invalid DNA character {chr(b[i])!r} at position {i}")
    return out


@dataclass
class DnaSequence:
    """DnaSequence (alphabet.rs:311-370): encoded bases + header."""
    bases: np.ndarray
    header: str = ""

    @classmethod
    def from_str(cls, s: str, header: str = "") -> "DnaSequence":
        if not s:
            raise FmIndexError("empty sequence")
        return cls(encode(s), header)

    def as_slice(self) -> np.ndarray:
        return self.bases

    def __len__(self) -> int:
        return len(self.bases)


# --------------------------------------------------------------------------
# C ABI (include/hsfm.h)
# --------------------------------------------------------------------------
class _View(C.Structure):
    _fields_ = [
        ("c_array", C.c_uint32 * 16),
        ("occ_num_lanes", C.c_uint8), ("occ_num_planes", C.c_uint8),
        ("occ_symbol_to_lane", C.c_uint8 * 16), ("occ_lane_to_symbol", C.c_uint8 * 16),
        ("occ_superblock_checkpoints", C.c_void_p), ("occ_superblock_len", C.c_uint64),
        ("occ_block_data", C.c_void_p), ("occ_block_data_len", C.c_uint64),
        ("occ_block_stride", C.c_uint64),
        ("occ_encoding", C.c_uint32), ("occ_text_len", C.c_uint32),
        ("sa_word_data", C.c_void_p), ("sa_word_data_len", C.c_uint64),
        ("sa_vals", C.c_void_p), ("sa_vals_len", C.c_uint64),
        ("sa_sample_rate", C.c_uint32), ("sa_text_len", C.c_uint32),
        ("text_len", C.c_uint32), ("num_sequences", C.c_uint32),
        ("seq_boundaries", C.c_void_p), ("seq_boundaries_len", C.c_uint64),
        ("has_lookup", C.c_uint32), ("lookup_depth", C.c_uint32),
        ("lookup_core_symbols", C.c_void_p), ("lookup_core_len", C.c_uint64),
        ("lookup_intervals", C.c_void_p), ("lookup_intervals_len", C.c_uint64),
        ("alphabet_tag", C.c_uint8),
        ("compat_mask", C.c_uint16 * 256),
    ]


class _LocRes(C.Structure):
    _fields_ = [("num_queries", C.c_uint32), ("offsets", C.POINTER(C.c_uint64)),
                ("hits", C.POINTER(C.c_uint32))]


class _PosRes(C.Structure):
    _fields_ = [("num_queries", C.c_uint32), ("offsets", C.POINTER(C.c_uint64)),
                ("positions", C.POINTER(C.c_uint32))]


class _LocDev(C.Structure):
    _fields_ = [("num_queries", C.c_uint32), ("total_hits", C.c_uint64),
                ("d_offsets", C.c_void_p), ("d_hits", C.c_void_p)]


class _MemHitC(C.Structure):
    _fields_ = [("query_start", C.c_uint32), ("query_end", C.c_uint32), ("match_count", C.c_uint32),
                ("truncated", C.c_uint32), ("pos_offset", C.c_uint64), ("pos_len", C.c_uint32),
                ("_reserved", C.c_uint32)]


class _MemRes(C.Structure):
    _fields_ = [("num_queries", C.c_uint32), ("hit_offsets", C.POINTER(C.c_uint64)),
                ("hits", C.POINTER(_MemHitC)), ("positions", C.POINTER(C.c_uint32)),
                ("num_hits", C.c_uint64), ("num_positions", C.c_uint64)]


class _MemDev(C.Structure):
    _fields_ = [("num_queries", C.c_uint32), ("num_hits", C.c_uint64), ("num_positions", C.c_uint64),
                ("d_hit_offsets", C.c_void_p), ("d_hits", C.c_void_p), ("d_positions", C.c_void_p)]


_BATCH_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(_MemDev))


class _Stats(C.Structure):
    _fields_ = [("fallback_reads", C.c_uint64), ("capacity_retries", C.c_uint64),
                ("guard_hits", C.c_uint64), ("records", C.c_uint64), ("main_kernel_ms", C.c_double),
                ("resolve_kernel_ms", C.c_double), ("occ_probes", C.c_uint64), ("diag", C.c_uint64 * 4),
                ("fallback_kernel_ms", C.c_double), ("sa_full_rows", C.c_uint64)]


class _LsiConfig(C.Structure):
    _fields_ = [("k", C.c_uint32), ("leaves", C.c_uint32), ("hidden", C.c_uint32), ("train_steps", C.c_uint32),
                ("sample", C.c_uint32), ("lr", C.c_float), ("lookup", C.c_uint32)]


class _LsiInfo(C.Structure):
    _fields_ = [("num_keys", C.c_uint64), ("entries", C.c_uint64), ("variant_entries", C.c_uint64),
                ("short_entries", C.c_uint64), ("max_err", C.c_uint32), ("fence_levels", C.c_uint32),
                ("mean_window", C.c_double), ("mse_initial", C.c_double), ("mse_final", C.c_double),
                ("enum_ms", C.c_double), ("train_ms", C.c_double), ("fit_ms", C.c_double),
                ("extra_ms", C.c_double), ("mem_seeding", C.c_uint32), ("pad2", C.c_uint32)]


class _SeedStats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("probes", C.c_uint64)]


class _BatchConfig(C.Structure):
    _fields_ = [("batch_reads", C.c_uint32), ("autotune", C.c_int), ("main_kernel", C.c_int),
                ("resolve_fanout", C.c_int), ("use_graph", C.c_int), ("batch_sets", C.c_int),
                ("window_batches", C.c_uint32)]


class _BatchStats(C.Structure):
    _fields_ = [("batches", C.c_uint32), ("fallback_batches", C.c_uint32), ("main_kernel", C.c_int),
                ("resolve_fanout", C.c_int), ("records", C.c_uint64), ("positions", C.c_uint64),
                ("total_ms", C.c_double), ("tune_total_ms", C.c_double), ("tune_ms", C.c_double * 6),
                ("window_batches", C.c_uint32), ("windows", C.c_uint32), ("graphs_captured", C.c_uint32)]


_lib_obj = None
_lib_lock = threading.Lock()

HSFM_MODE_SMEM, HSFM_MODE_MEM = 0, 1
# restart seeding of the one-shot MEM calls (hsfm_mem_seed_mode)
HSFM_SEED_TABLES, HSFM_SEED_PLAIN, HSFM_SEED_LEARNED, HSFM_SEED_LEARNED_TABLES = 0, 1, 2, 3


def lib_path() -> str:
    return _LIB_PATH


def _lib():
    """Load libhsfm.so (building it in-tree if nothing is built yet)."""
    global _lib_obj
    with _lib_lock:
        if _lib_obj is not None:
            return _lib_obj
        if not os.path.exists(_LIB_PATH):
            from . import build as _build
            _build.build()
        try:
            L = C.CDLL(_LIB_PATH)
        except OSError as e:
            raise GpuError(f"cannot load {_LIB_PATH}: {e}") from e
        vp, u8p, u64p, u32p = C.c_void_p, C.POINTER(C.c_uint8), C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)
        L.hsfm_last_error.restype = C.c_char_p
        L.hsfm_device_open.argtypes = [C.c_int, C.POINTER(vp)]
        L.hsfm_device_close.argtypes = [vp]
        L.hsfm_index_view_from_bytes.argtypes = [u8p, C.c_uint64, C.POINTER(_View)]
        L.hsfm_index_upload.argtypes = [vp, C.POINTER(_View), C.POINTER(_View), C.POINTER(vp)]
        L.hsfm_index_free.argtypes = [vp]
        L.hsfm_locate_batch.argtypes = [vp, vp, vp, C.c_uint32, C.POINTER(C.POINTER(_LocRes))]
        L.hsfm_count_batch.argtypes = [vp, vp, vp, C.c_uint32, vp]
        L.hsfm_locate_positions_batch.argtypes = [vp, vp, vp, C.c_uint32, C.POINTER(C.POINTER(_PosRes))]
        L.hsfm_locate_device.argtypes = [vp, vp, vp, C.c_uint32, C.POINTER(_LocDev)]
        L.hsfm_find_mems_batch.argtypes = [vp, vp, vp, C.c_uint32, C.c_uint32, C.c_int, vp, C.c_uint32,
                                           C.c_uint32, C.POINTER(C.POINTER(_MemRes))]
        L.hsfm_find_mems_device.argtypes = [vp, vp, vp, C.c_uint32, C.c_uint32, C.c_int, vp, C.c_uint32,
                                            C.c_uint32, C.POINTER(_MemDev)]
        L.hsfm_free_result.argtypes = [vp]
        L.hsfm_last_stats.argtypes = [vp, C.POINTER(_Stats)]
        L.hsfm_last_deferred.argtypes = [vp, C.POINTER(C.c_uint32), C.c_uint64, C.POINTER(C.c_uint64)]
        L.hsfm_mem_seed_mode.argtypes = [vp, C.c_int, vp]
        L.hsfm_build_index_device.argtypes = [vp, u8p, C.c_uint64, u32p, C.c_uint32, C.c_char_p, C.c_uint32,
                                              C.c_uint32, C.c_uint32, C.c_uint8, C.c_int, C.POINTER(u8p), u64p]
        L.hsfm_lsi_build.argtypes = [vp, C.POINTER(_LsiConfig), C.POINTER(vp)]
        L.hsfm_lsi_get_info.argtypes = [vp, C.POINTER(_LsiInfo)]
        L.hsfm_lsi_seed_device.argtypes = [vp, vp, C.c_uint64, vp, C.POINTER(_SeedStats)]
        L.hsfm_kmer_search_device.argtypes = [vp, C.c_uint32, vp, C.c_uint64, vp, C.POINTER(_SeedStats)]
        L.hsfm_lsi_seed.argtypes = [vp, vp, C.c_uint64, vp, C.POINTER(_SeedStats)]
        L.hsfm_lsi_seed_device_noalloc.argtypes = [vp, vp, C.c_uint64, vp, C.POINTER(C.c_uint64)]
        L.hsfm_kmer_search.argtypes = [vp, C.c_uint32, vp, C.c_uint64, vp]
        L.hsfm_lsi_grad.argtypes = [vp, vp, C.c_uint32, vp, vp, C.c_uint32, vp, vp]
        L.hsfm_revcomp_device.argtypes = [vp, vp, vp, C.c_uint32, vp]
        L.hsfm_revcomp_batch.argtypes = [vp, vp, vp, C.c_uint32, vp]
        L.hsfm_lsi_free.argtypes = [vp]
        L.hsfm_find_mems_batched.argtypes = [vp, vp, vp, C.c_uint32, C.c_uint32, C.c_int, vp, C.c_uint32,
                                             C.c_uint32, C.POINTER(_BatchConfig), _BATCH_FN, vp,
                                             C.POINTER(_BatchStats)]
        L.hsfm_device_alloc.argtypes = [vp, C.c_uint64, C.POINTER(vp)]
        L.hsfm_device_free.argtypes = [vp, vp]
        L.hsfm_host_to_device.argtypes = [vp, vp, vp, C.c_uint64]
        L.hsfm_device_to_host.argtypes = [vp, vp, vp, C.c_uint64]
        _lib_obj = L
        return L


def _check(rc: int, what: str):
    if rc != 0:
        msg = _lib().hsfm_last_error().decode(errors="replace")
        if rc == -2:
            raise DeserializeError(msg)
        raise GpuError(f"{what}: {msg}")


# --------------------------------------------------------------------------
# device context (src/gpu/mod.rs, src/gpu/context_cache.rs)
# --------------------------------------------------------------------------
class GpuContext:
    """One HIP device + stream; GpuContext::new (src/gpu/mod.rs:26-66)."""

    def __init__(self, ordinal: int = 0):
        h = C.c_void_p()
        _check(_lib().hsfm_device_open(ordinal, C.byref(h)), "GpuContext::new")
        self._h = h
        self.ordinal = ordinal

    @classmethod
    def new(cls, ordinal: int = 0) -> "GpuContext":
        return cls(ordinal)


class _ContextCache:
    """Process-wide context (context_cache.rs:5-19)."""

    def __init__(self):
        self._ctx = None
        self._lock = threading.Lock()

    def get_or_init(self) -> GpuContext:
        with self._lock:
            if self._ctx is None:
                self._ctx = GpuContext(int(os.environ.get("HSFM_DEVICE", "0")))
            return self._ctx


context_cache = _ContextCache()


# --------------------------------------------------------------------------
# queries -> CSR
# --------------------------------------------------------------------------
def _to_csr(queries) -> tuple[np.ndarray, np.ndarray]:
    if isinstance(queries, tuple) and len(queries) == 2 and isinstance(queries[1], np.ndarray):
        flat, off = queries
        return np.ascontiguousarray(flat, np.uint8), np.ascontiguousarray(off, np.uint64)
    arrs = []
    for q in queries:
        if isinstance(q, DnaSequence):
            q = q.bases
        elif isinstance(q, str):
            q = encode(q)
        arrs.append(np.frombuffer(bytes(q), np.uint8) if isinstance(q, (bytes, bytearray)) else np.asarray(q, np.uint8))
    off = np.zeros(len(arrs) + 1, np.uint64)
    if arrs:
        off[1:] = np.cumsum([len(a) for a in arrs], dtype=np.uint64)
    flat = np.concatenate(arrs) if arrs and int(off[-1]) > 0 else np.zeros(1, np.uint8)
    return np.ascontiguousarray(flat, np.uint8), off


# --------------------------------------------------------------------------
# FmIndex bytes (src/fm_index/serialize.rs)
# --------------------------------------------------------------------------
def _parse_headers(data: bytes) -> list[str]:
    """Walk the bincode stream to the seq_headers field (serialize.rs:55-67)."""
    off = 64 + 2 + 32  # c_array, num_lanes, num_planes, two lane maps

    def vec(elem):
        nonlocal off
        (k,) = struct.unpack_from("<Q", data, off)
        off += 8 + k * elem

    vec(4)            # superblock_checkpoints
    vec(1)            # block_data
    off += 8 + 4 + 4  # block_stride, encoding, text_len
    vec(1)            # word_data
    vec(4)            # sa_vals
    off += 4 + 4 + 4 + 4
    vec(4)            # seq_boundaries
    (nh,) = struct.unpack_from("<Q", data, off)
    off += 8
    out = []
    for _ in range(nh):
        (l,) = struct.unpack_from("<Q", data, off)
        off += 8
        out.append(data[off:off + l].decode())
        off += l
    return out


class FmIndex:
    """Device-resident FmIndex built from FmIndex::to_bytes output."""

    def __init__(self, data: bytes, alphabet_fns: AlphabetFns | None = None):
        self._data = bytes(data)
        self._buf = np.frombuffer(self._data, np.uint8) if self._data else np.zeros(1, np.uint8)
        view = _View()
        _check(_lib().hsfm_index_view_from_bytes(self._buf.ctypes.data_as(C.POINTER(C.c_uint8)),
                                                 len(self._data), C.byref(view)), "FmIndex::from_bytes")
        fns = alphabet_fns_from_tag(view.alphabet_tag)
        if fns is None:
            raise DeserializeError(f"unknown alphabet tag {view.alphabet_tag} in serialized FM-index")
        if alphabet_fns is not None:
            fns = alphabet_fns
            m = fns.compat_masks()
            C.memmove(view.compat_mask, m.ctypes.data, 512)
        self.alphabet_fns = fns
        self._view = view
        self.seq_headers = _parse_headers(self._data)
        self.text_len = int(view.text_len)
        self.num_sequences = int(view.num_sequences)
        self._seq_boundaries = np.frombuffer(C.string_at(view.seq_boundaries, 4 * view.seq_boundaries_len),
                                             np.uint32).copy()
        self._dev = None  # (ctx, handle)
        self._pair_rev = None  # reverse index when this is the forward half of a BidirFmIndex

    @classmethod
    def from_bytes(cls, data: bytes) -> "FmIndex":
        """FmIndex::from_bytes (serialize.rs:32-52)."""
        return cls(data)

    def seq_boundaries(self) -> np.ndarray:
        return self._seq_boundaries

    # device handle ---------------------------------------------------------
    def _handle(self, ctx: GpuContext | None = None, rev: "FmIndex | None" = None):
        if self._dev is not None:
            return self._dev[1]
        rev = rev if rev is not None else self._pair_rev
        ctx = ctx or context_cache.get_or_init()
        h = C.c_void_p()
        rv = C.byref(rev._view) if rev is not None else None
        _check(_lib().hsfm_index_upload(ctx._h, C.byref(self._view), rv, C.byref(h)), "index upload")
        self._dev = (ctx, h)
        return h

    def __del__(self):
        d = getattr(self, "_dev", None)
        if d is not None and _lib_obj is not None:
            _lib_obj.hsfm_index_free(d[1])
            self._dev = None

    def last_stats(self) -> dict:
        st = _Stats()
        _check(_lib().hsfm_last_stats(self._handle(), C.byref(st)), "stats")
        return {k: (list(getattr(st, k)) if k == "diag" else getattr(st, k)) for k, _ in _Stats._fields_}

    # queries ---------------------------------------------------------------
    def locate_batch(self, queries, ctx: GpuContext | None = None) -> tuple[np.ndarray, np.ndarray]:
        """CSR result: offsets[nq+1], hits[total, 2] = (seq_idx, pos_in_seq)."""
        flat, off = _to_csr(queries)
        nq = len(off) - 1
        h = self._handle(ctx)
        res = C.POINTER(_LocRes)()
        _check(_lib().hsfm_locate_batch(h, flat.ctypes.data, off.ctypes.data, nq, C.byref(res)), "locate_batch_gpu")
        try:
            r = res.contents
            offs = np.ctypeslib.as_array(r.offsets, shape=(nq + 1,)).copy()
            tot = int(offs[-1])
            hits = (np.ctypeslib.as_array(r.hits, shape=(2 * tot,)).copy().reshape(-1, 2)
                    if tot else np.zeros((0, 2), np.uint32))
        finally:
            _lib().hsfm_free_result(res)
        return offs, hits

    def locate_positions_batch(self, queries, ctx: GpuContext | None = None) -> tuple[np.ndarray, np.ndarray]:
        """CSR result: offsets[nq+1], positions[total] = absolute text offsets
        (FmIndex::locate_positions, query.rs:46-55), ascending SA-row order."""
        flat, off = _to_csr(queries)
        nq = len(off) - 1
        h = self._handle(ctx)
        res = C.POINTER(_PosRes)()
        _check(_lib().hsfm_locate_positions_batch(h, flat.ctypes.data, off.ctypes.data, nq, C.byref(res)),
               "locate_positions_batch")
        try:
            r = res.contents
            offs = np.ctypeslib.as_array(r.offsets, shape=(nq + 1,)).copy()
            tot = int(offs[-1])
            pos = np.ctypeslib.as_array(r.positions, shape=(tot,)).copy() if tot else np.zeros(0, np.uint32)
        finally:
            _lib().hsfm_free_result(res)
        return offs, pos

    def locate_positions_many(self, patterns) -> list[np.ndarray]:
        """FmIndex::locate_positions_many (query.rs:71-208) on the GPU: absolute
        text offsets per pattern."""
        offs, pos = self.locate_positions_batch(patterns)
        return [pos[int(offs[i]):int(offs[i + 1])] for i in range(len(offs) - 1)]

    def locate_positions(self, pattern) -> np.ndarray:
        """FmIndex::locate_positions (query.rs:46-55) of one pattern."""
        return self.locate_positions_many([pattern])[0]

    def locate_gpu(self, queries) -> list[list[tuple[str, int]]]:
        """FmIndex::locate_gpu (query.rs:367-384): (header, pos) per hit."""
        offs, hits = self.locate_batch(queries)
        hdr = self.seq_headers
        return [[(hdr[int(s)], int(p)) for s, p in hits[offs[i]:offs[i + 1]]] for i in range(len(offs) - 1)]

    def count_gpu(self, queries) -> np.ndarray:
        """FmIndex::count (query.rs:9-14) over a batch, on the GPU."""
        flat, off = _to_csr(queries)
        nq = len(off) - 1
        out = np.zeros(max(nq, 1), np.uint32)
        _check(_lib().hsfm_count_batch(self._handle(), flat.ctypes.data, off.ctypes.data, nq, out.ctypes.data),
               "count_gpu")
        return out[:nq]

    def locate_device(self, d_qflat: int, d_qoff: int, nq: int) -> tuple[int, int, int]:
        """Inputs already in device memory (raw pointers); returns
        (total_hits, d_offsets, d_hits) pointing into library-owned device memory."""
        r = _LocDev()
        _check(_lib().hsfm_locate_device(self._handle(), C.c_void_p(d_qflat), C.c_void_p(d_qoff), nq, C.byref(r)),
               "locate_device")
        return int(r.total_hits), int(r.d_offsets or 0), int(r.d_hits or 0)


def locate_batch_gpu(ctx: GpuContext, index: FmIndex, queries) -> list[list[tuple[int, int]]]:
    """locate_batch_gpu (src/gpu/locate.rs:52-56): (seq_idx, pos) per hit, per query."""
    offs, hits = index.locate_batch(queries, ctx)
    return [[(int(s), int(p)) for s, p in hits[offs[i]:offs[i + 1]]] for i in range(len(offs) - 1)]


@dataclass
class MemHit:
    """MemHit (src/gpu/mem_hit.rs:6-19)."""
    query_start: int
    query_end: int
    match_count: int
    positions: list = field(default_factory=list)
    truncated: bool = False


class BidirFmIndex:
    """BidirFmIndex: forward + reverse FmIndex (bidir_index.rs:29-34)."""

    def __init__(self, fwd: FmIndex, rev: FmIndex):
        self.fwd, self.rev = fwd, rev
        fwd._pair_rev = rev  # one device index serves locate (forward) and MEM (both)

    @classmethod
    def from_bytes(cls, data: bytes) -> "BidirFmIndex":
        """bidir_index.rs:195-210: [u32 LE fwd_len][fwd][rev]."""
        if len(data) < 4:
            raise DeserializeError("truncated bidirectional index")
        (fl,) = struct.unpack_from("<I", data, 0)
        if len(data) < 4 + fl:
            raise DeserializeError("truncated forward index")
        return cls(FmIndex(data[4:4 + fl]), FmIndex(data[4 + fl:]))

    def seq_boundaries(self) -> np.ndarray:
        """bidir_index.rs:85-87"""
        return self.fwd.seq_boundaries()

    def text_len(self) -> int:
        return self.fwd.text_len

    def _handle(self, ctx=None):
        return self.fwd._handle(ctx, rev=self.rev)

    def last_stats(self) -> dict:
        st = _Stats()
        _check(_lib().hsfm_last_stats(self._handle(), C.byref(st)), "stats")
        return {k: (list(getattr(st, k)) if k == "diag" else getattr(st, k)) for k, _ in _Stats._fields_}

    def set_seed_mode(self, mode: int, lsi: "LearnedSeedIndex | None" = None):
        """Restart seeding of the one-shot MEM calls: HSFM_SEED_TABLES (default),
        HSFM_SEED_PLAIN, HSFM_SEED_LEARNED (lsi, plain fallback) or
        HSFM_SEED_LEARNED_TABLES (lsi, table fallback).  Results do not depend
        on it."""
        _check(_lib().hsfm_mem_seed_mode(self._handle(), int(mode), lsi._h if lsi is not None else None),
               "mem_seed_mode")
        self._seed_lsi = lsi  # keep the learned index alive while it is attached

    def last_deferred(self) -> np.ndarray:
        """Reads the main kernel of the last one-shot MEM call deferred to the
        workgroup passes (ascending)."""
        n = C.c_uint64()
        _check(_lib().hsfm_last_deferred(self._handle(), None, 0, C.byref(n)), "last_deferred")
        out = np.zeros(max(n.value, 1), np.uint32)
        _check(_lib().hsfm_last_deferred(self._handle(), out.ctypes.data_as(C.POINTER(C.c_uint32)), n.value,
                                         C.byref(n)), "last_deferred")
        return out[:n.value]

    def find_mems_batch(self, queries, min_len: int, mode: int, ref_boundaries=(), max_hits_per_mem: int = 1024):
        """Raw CSR result: (hit_offsets[nq+1], hits[nh, 8] u32 view, positions[np, 2])."""
        flat, off = _to_csr(queries)
        nq = len(off) - 1
        rb = np.ascontiguousarray(np.asarray(ref_boundaries, np.uint32))
        rbp = rb.ctypes.data if len(rb) else None
        res = C.POINTER(_MemRes)()
        _check(_lib().hsfm_find_mems_batch(self._handle(), flat.ctypes.data, off.ctypes.data, nq, int(min_len), mode,
                                           rbp, len(rb), int(max_hits_per_mem), C.byref(res)), "find_mems_gpu")
        try:
            r = res.contents
            ho = np.ctypeslib.as_array(r.hit_offsets, shape=(nq + 1,)).copy()
            nh, npos = int(r.num_hits), int(r.num_positions)
            if nh:
                raw = np.ctypeslib.as_array(C.cast(r.hits, C.POINTER(C.c_uint32)), shape=(nh * 8,)).copy()
                hits = raw.reshape(nh, 8)
            else:
                hits = np.zeros((0, 8), np.uint32)
            pos = (np.ctypeslib.as_array(r.positions, shape=(2 * npos,)).copy().reshape(-1, 2)
                   if npos else np.zeros((0, 2), np.uint32))
        finally:
            _lib().hsfm_free_result(res)
        return ho, hits, pos

    def find_mems_batched(self, queries, min_len: int, mode: int, ref_boundaries=(), max_hits_per_mem: int = 1024,
                          batch_reads: int = 65536, autotune: bool = False, main_kernel: int = -1,
                          resolve_fanout: int = 0, use_graph: bool = True, batch_sets: int = 0,
                          window_batches: int = 0):
        """The batched pipeline (one HIP graph per batch); same raw CSR as
        find_mems_batch, plus the pipeline statistics."""
        flat, off = _to_csr(queries)
        nq = len(off) - 1
        rb = np.ascontiguousarray(np.asarray(ref_boundaries, np.uint32))
        h = self._handle()
        dev = self.fwd._dev[0]._h
        L = _lib()
        bufs = []

        def dalloc(arr):
            p = C.c_void_p()
            _check(L.hsfm_device_alloc(dev, max(arr.nbytes, 8), C.byref(p)), "device_alloc")
            bufs.append(p)
            if arr.nbytes:
                _check(L.hsfm_host_to_device(dev, p, arr.ctypes.data, arr.nbytes), "host_to_device")
            return p

        parts = []

        def on_batch(_user, first, nb, res_p):
            r = res_p.contents
            nh, npos = int(r.num_hits), int(r.num_positions)
            ho = np.zeros(nb + 1, np.uint64)
            _check(L.hsfm_device_to_host(dev, ho.ctypes.data, C.c_void_p(r.d_hit_offsets), ho.nbytes), "d2h")
            hits = np.zeros((nh, 8), np.uint32)
            pos = np.zeros((npos, 2), np.uint32)
            if nh:
                _check(L.hsfm_device_to_host(dev, hits.ctypes.data, C.c_void_p(r.d_hits), hits.nbytes), "d2h")
            if npos:
                _check(L.hsfm_device_to_host(dev, pos.ctypes.data, C.c_void_p(r.d_positions), pos.nbytes), "d2h")
            parts.append((int(first), ho, hits, pos))

        cb = _BATCH_FN(on_batch)
        try:
            dq, do, dr = dalloc(flat), dalloc(off), dalloc(rb)
            cfg = _BatchConfig(batch_reads, int(autotune), main_kernel, resolve_fanout, int(use_graph), batch_sets,
                               window_batches)
            st = _BatchStats()
            _check(L.hsfm_find_mems_batched(h, dq, do, nq, int(min_len), mode, dr if len(rb) else None, len(rb),
                                            int(max_hits_per_mem), C.byref(cfg), cb, None, C.byref(st)),
                   "find_mems_batched")
        finally:
            for p in bufs:
                L.hsfm_device_free(dev, p)
        parts.sort(key=lambda t: t[0])
        ho_all, hits_all, pos_all = [np.zeros(1, np.uint64)], [], []
        nh_run = npos_run = 0
        for _, ho, hits, pos in parts:
            ho_all.append(ho[1:] + np.uint64(nh_run))
            if len(hits):
                po = (hits[:, 4].astype(np.uint64) | (hits[:, 5].astype(np.uint64) << np.uint64(32))) + np.uint64(npos_run)
                hits = hits.copy()
                hits[:, 4] = (po & np.uint64(0xFFFFFFFF)).astype(np.uint32)
                hits[:, 5] = (po >> np.uint64(32)).astype(np.uint32)
            hits_all.append(hits)
            pos_all.append(pos)
            nh_run += len(hits)
            npos_run += len(pos)
        stats = {k: (list(getattr(st, k)) if k == "tune_ms" else getattr(st, k)) for k, _ in _BatchStats._fields_}
        return (np.concatenate(ho_all), np.concatenate(hits_all) if hits_all else np.zeros((0, 8), np.uint32),
                np.concatenate(pos_all) if pos_all else np.zeros((0, 2), np.uint32), stats)

    def find_mems_batched_device(self, d_qflat: int, d_qoff: int, nq: int, min_len: int, mode: int,
                                 d_ref_boundaries: int, n_refs: int, max_hits_per_mem: int = 1024,
                                 batch_reads: int = 65536, autotune: bool = False, main_kernel: int = -1,
                                 resolve_fanout: int = 0, use_graph: bool = True, on_batch=None,
                                 batch_sets: int = 0, window_batches: int = 0) -> dict:
        """Batched pipeline on device-resident reads; on_batch(first, n, result)
        receives each batch's device result (None: results are dropped)."""
        cfg = _BatchConfig(batch_reads, int(autotune), main_kernel, resolve_fanout, int(use_graph), batch_sets,
                           window_batches)
        st = _BatchStats()
        cb = _BATCH_FN(lambda _u, first, n, r: on_batch(first, n, r.contents)) if on_batch else _BATCH_FN()
        _check(_lib().hsfm_find_mems_batched(self._handle(), C.c_void_p(d_qflat), C.c_void_p(d_qoff), nq,
                                             int(min_len), mode, C.c_void_p(d_ref_boundaries or None), n_refs,
                                             int(max_hits_per_mem), C.byref(cfg), cb, None, C.byref(st)),
               "find_mems_batched")
        return {k: (list(getattr(st, k)) if k == "tune_ms" else getattr(st, k)) for k, _ in _BatchStats._fields_}

    def _mem_hits(self, queries, min_len, mode, ref_boundaries, max_hits_per_mem) -> list[list[MemHit]]:
        ho, hits, pos = self.find_mems_batch(queries, min_len, mode, ref_boundaries, max_hits_per_mem)
        out = []
        for q in range(len(ho) - 1):
            lst = []
            for h in hits[ho[q]:ho[q + 1]]:
                po = int(h[4]) | (int(h[5]) << 32)
                pl = int(h[6])
                lst.append(MemHit(int(h[0]), int(h[1]), int(h[2]),
                                  [(int(a), int(b)) for a, b in pos[po:po + pl]], bool(h[3])))
            out.append(lst)
        return out

    def find_smems_gpu(self, queries, min_len: int, ref_boundaries=(), max_hits_per_mem: int = 1024):
        """BidirFmIndex::find_smems_gpu (bidir_index.rs:225-245)."""
        return self._mem_hits(queries, min_len, HSFM_MODE_SMEM, ref_boundaries, max_hits_per_mem)

    def find_mems_gpu(self, queries, min_len: int, ref_boundaries=(), max_hits_per_mem: int = 1024):
        """BidirFmIndex::find_mems_gpu (bidir_index.rs:256-276)."""
        return self._mem_hits(queries, min_len, HSFM_MODE_MEM, ref_boundaries, max_hits_per_mem)

    def find_mems_device(self, d_qflat: int, d_qoff: int, nq: int, min_len: int, mode: int,
                         d_ref_boundaries: int, n_refs: int, max_hits_per_mem: int = 1024):
        """Device-resident variant; returns (num_hits, num_positions, d_hit_offsets, d_hits, d_positions)."""
        r = _MemDev()
        _check(_lib().hsfm_find_mems_device(self._handle(), C.c_void_p(d_qflat), C.c_void_p(d_qoff), nq,
                                            int(min_len), mode, C.c_void_p(d_ref_boundaries or None), n_refs,
                                            int(max_hits_per_mem), C.byref(r)), "find_mems_device")
        return int(r.num_hits), int(r.num_positions), int(r.d_hit_offsets or 0), int(r.d_hits or 0), int(r.d_positions or 0)


# --------------------------------------------------------------------------
# learned seed index
# --------------------------------------------------------------------------
def kmer_keys(codes, k: int) -> np.ndarray:
    """Keys of every exact-ACGT k-mer window of codes (two bits per base, first
    base most significant); windows with another code get key 2^64-1."""
    c = np.asarray(codes, np.uint8)
    m = len(c) - k + 1
    if m <= 0:
        return np.zeros(0, np.uint64)
    ok = (c >= 1) & (c <= 4)
    d = np.where(ok, c.astype(np.uint64) - 1, 0).astype(np.uint64)
    keys = np.zeros(m, np.uint64)
    for t in range(k):
        keys = (keys << np.uint64(2)) | d[t:t + m]
    bad = np.convolve(~ok, np.ones(k, int), "valid") > 0
    keys[bad] = np.uint64(0xFFFFFFFFFFFFFFFF)
    return keys


class LearnedSeedIndex:
    """Two-stage FP32 learned index over the exact-ACGT k-mers of an index.

    seed(keys) returns, per key, (fwd_lo, fwd_hi, rev_lo, found): the interval
    plain exact backward search gives for the k-mer (bidir.rs:112-130)."""

    LOOKUP = {"auto": 0, "global": 1, "lds": 2}

    def __init__(self, index, k: int = 21, leaves: int = 4096, hidden: int = 16, train_steps: int = 400,
                 sample: int = 1 << 16, lr: float = 0.01, lookup: str = "auto"):
        self.index = index
        self.k = k
        cfg = _LsiConfig(k, leaves, hidden, train_steps, sample, lr, self.LOOKUP[lookup])
        h = C.c_void_p()
        _check(_lib().hsfm_lsi_build(index._handle(), C.byref(cfg), C.byref(h)), "lsi_build")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib_obj is not None:
            _lib_obj.hsfm_lsi_free(h)
            self._h = None

    def info(self) -> dict:
        st = _LsiInfo()
        _check(_lib().hsfm_lsi_get_info(self._h, C.byref(st)), "lsi_info")
        return {k: getattr(st, k) for k, _ in _LsiInfo._fields_ if k != "pad2"}

    def seed(self, keys, stats: dict | None = None) -> np.ndarray:
        keys = np.ascontiguousarray(np.asarray(keys, np.uint64))
        out = np.zeros((len(keys), 4), np.uint32)
        st = _SeedStats()
        _check(_lib().hsfm_lsi_seed(self._h, keys.ctypes.data, len(keys), out.ctypes.data, C.byref(st)), "lsi_seed")
        if stats is not None:
            stats.update(kernel_ms=st.kernel_ms, probes=st.probes)
        return out

    def seed_device_noalloc(self, d_keys: int, n: int, d_out: int) -> int:
        """seed_device through the allocation-free entry point; returns the
        call's loads in 64-byte units."""
        probes = C.c_uint64()
        _check(_lib().hsfm_lsi_seed_device_noalloc(self._h, C.c_void_p(d_keys), n, C.c_void_p(d_out),
                                                   C.byref(probes)), "lsi_seed_device_noalloc")
        return probes.value

    def seed_device(self, d_keys: int, n: int, d_out: int) -> dict:
        st = _SeedStats()
        _check(_lib().hsfm_lsi_seed_device(self._h, C.c_void_p(d_keys), n, C.c_void_p(d_out), C.byref(st)),
               "lsi_seed_device")
        return {"kernel_ms": st.kernel_ms, "probes": st.probes}


def kmer_search(index, keys, k: int) -> np.ndarray:
    """Plain exact search of k-mer keys on the device (the learned index's judge)."""
    keys = np.ascontiguousarray(np.asarray(keys, np.uint64))
    out = np.zeros((len(keys), 4), np.uint32)
    _check(_lib().hsfm_kmer_search(index._handle(), k, keys.ctypes.data, len(keys), out.ctypes.data), "kmer_search")
    return out


def kmer_search_device(index, k: int, d_keys: int, n: int, d_out: int) -> dict:
    st = _SeedStats()
    _check(_lib().hsfm_kmer_search_device(index._handle(), k, C.c_void_p(d_keys), n, C.c_void_p(d_out),
                                          C.byref(st)), "kmer_search_device")
    return {"kernel_ms": st.kernel_ms}


def build_index_bytes_gpu(text, bounds, *, headers=None, sample_rate: int = 32, lookup_depth: int = 0,
                          encoding: int = 0, tag: int = 0, bidir: bool = False,
                          ctx: "GpuContext | None" = None) -> bytes:
    """FmIndex::build_cpu + to_bytes (BidirFmIndex's with bidir=True) computed
    on the device: the same byte stream (hsfm_build_index_device)."""
    ctx = ctx or context_cache.get_or_init()
    text = np.ascontiguousarray(text, np.uint8)
    bounds = np.ascontiguousarray(bounds, np.uint32)
    hb = b"".join(h.encode() + b"\0" for h in headers) if headers is not None else None
    out = C.POINTER(C.c_uint8)()
    n = C.c_uint64()
    _check(_lib().hsfm_build_index_device(ctx._h, text.ctypes.data_as(C.POINTER(C.c_uint8)), len(text),
                                          bounds.ctypes.data_as(C.POINTER(C.c_uint32)), len(bounds), hb,
                                          sample_rate, lookup_depth, encoding, tag, 1 if bidir else 0,
                                          C.byref(out), C.byref(n)), "build_index_device")
    try:
        return np.ctypeslib.as_array(out, shape=(n.value,)).tobytes() if n.value else b""
    finally:
        _lib().hsfm_free_result(C.cast(out, C.c_void_p))


def reverse_complement(queries, ctx: "GpuContext | None" = None):
    """Reverse complements of a query batch on the device, for searching the
    other strand (no counterpart in the reference, which searches the given
    strand only): returns (flat, offsets) in the layout of the input."""
    ctx = ctx or context_cache.get_or_init()
    flat, off = _to_csr(queries)
    nq = len(off) - 1
    out = np.zeros(max(int(off[-1]), 1), np.uint8)
    if nq:
        _check(_lib().hsfm_revcomp_batch(ctx._h, flat.ctypes.data, off.ctypes.data, nq, out.ctypes.data),
               "reverse_complement")
    return out[:int(off[-1])], off.copy()


def reverse_complement_device(d_qflat: int, d_qoff: int, nq: int, d_out: int, ctx: "GpuContext | None" = None):
    """reverse_complement on device-resident queries (d_out holds qoff[nq] bytes)."""
    ctx = ctx or context_cache.get_or_init()
    _check(_lib().hsfm_revcomp_device(ctx._h, C.c_void_p(d_qflat), C.c_void_p(d_qoff), nq, C.c_void_p(d_out)),
           "reverse_complement_device")


def lsi_gradient(params: np.ndarray, hidden: int, x: np.ndarray, y: np.ndarray, ctx: "GpuContext | None" = None):
    """Stage-1 loss and gradient computed by the training kernel (for checks)."""
    ctx = ctx or context_cache.get_or_init()
    p = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    g = np.zeros(3 * hidden + 1, np.float32)
    loss = np.zeros(1, np.float32)
    _check(_lib().hsfm_lsi_grad(ctx._h, p.ctypes.data, hidden, x.ctypes.data, y.ctypes.data, len(x),
                                g.ctypes.data, loss.ctypes.data), "lsi_grad")
    return float(loss[0]), g
