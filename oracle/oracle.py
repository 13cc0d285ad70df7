"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: this module is the checker.  tests/, the smoke test
and bench.py's cpu_baseline leg import it; the product package never does.

The C library restates the reference crate's CPU query path (see the file:line
citations in hsfm_oracle.c) and its index construction, so the checker can
hand the product exactly the bytes ``FmIndex::to_bytes`` would produce.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class _LocOut(C.Structure):
    _fields_ = [("offsets", u64p), ("hits", u32p)]


class _MemOut(C.Structure):
    _fields_ = [("mem_off", u64p), ("mems", u32p), ("ivs", u32p), ("nmem", C.c_uint64),
                ("niv", C.c_uint64)]


class _Biv(C.Structure):
    _fields_ = [("fwd_lo", C.c_uint32), ("fwd_hi", C.c_uint32), ("rev_lo", C.c_uint32),
                ("rev_hi", C.c_uint32)]


class _Probe(C.Structure):
    _fields_ = [("occ_probes", C.c_uint64), ("sa_probes", C.c_uint64),
                ("sa_val_reads", C.c_uint64), ("steps", C.c_uint64)]


def build_lib() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def use_native_build(outdir: str) -> str:
    """Compile the oracle with -march=native for the host it runs on (the
    timed CPU baseline) into outdir and use it for every later oracle call.
    Must run before the first oracle call of the process; returns its path."""
    global _LIB_PATH
    if _lib is not None:
        raise RuntimeError("the oracle library is already loaded")
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "liboracle_native.so")
    subprocess.run(["make", "-s", "-B", "-C", _HERE, f"OUT={out}", "CFLAGS=-O3 -march=native -fPIC -std=gnu11 "
                    "-Wall -Wextra -Wno-unused-parameter"], check=True)
    _LIB_PATH = out
    return out


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build_lib()
    L = C.CDLL(_LIB_PATH)
    vp = C.c_void_p
    L.or_encode_byte.argtypes = [C.c_uint8]
    L.or_encode_byte.restype = C.c_int
    L.or_compatible.argtypes = [C.c_uint8, C.c_uint8, u8p]
    L.or_compatible.restype = C.c_uint32
    L.or_compat_mask.argtypes = [C.c_uint8, C.c_uint8]
    L.or_compat_mask.restype = C.c_uint16
    L.or_iupac_bases.argtypes = [C.c_uint8, u8p]
    L.or_iupac_bases.restype = C.c_uint32
    L.or_build.argtypes = [u8p, C.c_uint64, u32p, C.c_uint32, C.c_char_p, C.c_uint32, C.c_uint32,
                           C.c_uint32, C.c_uint8, C.c_int, C.c_int]
    L.or_build.restype = vp
    L.or_build_rev.argtypes = [u8p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint8, C.c_int, C.c_int]
    L.or_build_rev.restype = vp
    L.or_free.argtypes = [vp]
    L.or_to_bytes.argtypes = [vp, C.POINTER(u8p), u64p]
    L.or_from_bytes.argtypes = [u8p, C.c_uint64, C.c_char_p, C.c_int]
    L.or_from_bytes.restype = vp
    L.or_free_buf.argtypes = [vp]
    for n in ("or_text_len", "or_num_lanes", "or_sample_rate", "or_num_sequences", "or_lookup_depth"):
        getattr(L, n).argtypes = [vp]
        getattr(L, n).restype = C.c_uint32
    L.or_seq_boundaries.argtypes = [vp]
    L.or_seq_boundaries.restype = u32p
    L.or_c_array.argtypes = [vp, u32p]
    L.or_rank.argtypes = [vp, C.c_uint8, C.c_uint32]
    L.or_rank.restype = C.c_uint32
    L.or_lf_step.argtypes = [vp, C.c_uint32, u8p, u32p]
    L.or_symbol_at.argtypes = [vp, C.c_uint32]
    L.or_symbol_at.restype = C.c_uint8
    L.or_sa_get.argtypes = [vp, C.c_uint32, u32p]
    L.or_sa_get.restype = C.c_int
    L.or_resolve_sa.argtypes = [vp, C.c_uint32]
    L.or_resolve_sa.restype = C.c_uint32
    L.or_map_position.argtypes = [vp, C.c_uint32, u32p, u32p]
    L.or_map_position.restype = C.c_int
    L.or_backward_search.argtypes = [vp, u8p, C.c_uint32, u32p, C.c_uint32]
    L.or_backward_search.restype = C.c_uint32
    L.or_count.argtypes = [vp, u8p, C.c_uint32]
    L.or_count.restype = C.c_uint32
    L.or_locate_batch.argtypes = [vp, u8p, u64p, C.c_uint32, C.c_int, C.c_int, C.POINTER(_LocOut)]
    L.or_locate_out_free.argtypes = [C.POINTER(_LocOut)]
    L.or_extend_right.argtypes = [vp, C.POINTER(_Biv), C.c_uint8, C.POINTER(_Biv)]
    L.or_extend_right.restype = C.c_int
    L.or_extend_left.argtypes = [vp, C.POINTER(_Biv), C.c_uint8, C.POINTER(_Biv)]
    L.or_extend_left.restype = C.c_int
    L.or_find_mems_batch.argtypes = [vp, vp, u8p, u64p, C.c_uint32, C.c_uint32, C.c_int, C.c_int,
                                     C.POINTER(_MemOut)]
    L.or_mem_out_free.argtypes = [C.POINTER(_MemOut)]
    L.or_resolve_rows.argtypes = [vp, u32p, C.c_uint64, u32p, C.c_int]
    L.or_probe_reset.argtypes = []
    L.or_probe_get.argtypes = [C.POINTER(_Probe)]
    L.or_probe_enable.argtypes = [C.c_int]
    L.or_naive_locate.argtypes = [u8p, C.c_uint64, u8p, C.c_uint32, C.c_uint8, u32p, C.c_uint64]
    L.or_naive_locate.restype = C.c_uint64
    L.or_brute_mems.argtypes = [u8p, C.c_uint64, u8p, C.c_uint32, C.c_uint32, C.c_uint8, u32p,
                                C.c_uint64]
    L.or_brute_mems.restype = C.c_uint64
    L.or_suffix_array.argtypes = [u8p, C.c_uint64, u32p, C.c_int]
    L.or_suffix_array.restype = C.c_int
    L.or_gen_reference.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_double, C.c_double,
                                   C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                   C.c_double, u8p, C.c_uint32, C.POINTER(u8p), u64p,
                                   C.POINTER(u32p)]
    L.or_gen_reads.argtypes = [C.c_uint64, u8p, C.c_uint64, u32p, C.c_uint32, C.c_uint64,
                               C.c_uint32, C.c_double, C.c_uint32, C.c_uint32, C.c_double,
                               C.c_double, C.c_int, C.c_uint64, C.c_double, C.POINTER(u8p),
                               C.POINTER(u64p)]
    _lib = L
    return L


def _p8(a: np.ndarray):
    return a.ctypes.data_as(u8p)


def _p32(a: np.ndarray):
    return a.ctypes.data_as(u32p)


def _p64(a: np.ndarray):
    return a.ctypes.data_as(u64p)


# --------------------------------------------------------------------------
# alphabet helpers (alphabet.rs)
# --------------------------------------------------------------------------
SYMBOLS = "$ACGTNRYSWKMBDHV"
IUPAC_AMBIG = [SYMBOLS.index(c) for c in "RYSWKMBDHVN"]


def encode(s: str) -> np.ndarray:
    L = lib()
    out = np.empty(len(s), dtype=np.uint8)
    for i, ch in enumerate(s.encode()):
        v = L.or_encode_byte(ch)
        if v < 0:
            raise ValueError(f"invalid character {chr(ch)!r} at {i}")
        out[i] = v
    return out


def decode(codes) -> str:
    return "".join(SYMBOLS[int(c)] for c in codes)


def compatible(code: int, tag: int = 0) -> list[int]:
    buf = (C.c_uint8 * 16)()
    k = lib().or_compatible(tag, code, buf)
    return list(buf[:k])


def compat_mask(code: int, tag: int = 0) -> int:
    return int(lib().or_compat_mask(tag, code))


def iupac_bases(code: int) -> list[int]:
    buf = (C.c_uint8 * 4)()
    k = lib().or_iupac_bases(code, buf)
    return list(buf[:k])


def concatenate(seqs) -> tuple[np.ndarray, np.ndarray]:
    """alphabet.rs:375-393: s1 $ s2 $ ... with cumulative exclusive ends."""
    parts, bounds, tot = [], [], 0
    for s in seqs:
        a = encode(s) if isinstance(s, str) else np.asarray(s, dtype=np.uint8)
        parts.append(a)
        parts.append(np.zeros(1, np.uint8))
        tot += len(a) + 1
        bounds.append(tot)
    return np.concatenate(parts) if parts else np.zeros(0, np.uint8), np.array(bounds, np.uint32)


def to_csr(queries) -> tuple[np.ndarray, np.ndarray]:
    arrs = [encode(q) if isinstance(q, str) else np.asarray(q, dtype=np.uint8) for q in queries]
    off = np.zeros(len(arrs) + 1, np.uint64)
    if arrs:
        off[1:] = np.cumsum([len(a) for a in arrs])
    flat = np.concatenate(arrs) if arrs and off[-1] > 0 else np.zeros(1, np.uint8)
    return np.ascontiguousarray(flat, np.uint8), off


def suffix_array(text: np.ndarray, naive: bool = False) -> np.ndarray:
    text = np.ascontiguousarray(text, np.uint8)
    sa = np.empty(len(text), np.uint32)
    if len(text):
        if lib().or_suffix_array(_p8(text), len(text), _p32(sa), 1 if naive else 0) != 0:
            raise ValueError("suffix array construction failed")
    return sa


# --------------------------------------------------------------------------
# index
# --------------------------------------------------------------------------
class OracleIndex:
    """FmIndex restated (fm_index/mod.rs).  Owns a C or_fmi handle."""

    def __init__(self, handle, text=None):
        if not handle:
            raise ValueError("oracle index construction failed")
        self._h = handle
        self.text = text

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.or_free(h)
            self._h = None

    @classmethod
    def build(cls, seqs=None, *, text=None, bounds=None, headers=None, sample_rate=32,
              lookup_depth=0, encoding=0, tag=0, naive_sa=False, threads=1):
        if seqs is not None:
            text, bounds = concatenate(seqs)
        text = np.ascontiguousarray(text, np.uint8)
        bounds = np.ascontiguousarray(bounds, np.uint32)
        hb = None
        if headers is not None:
            hb = b"".join(h.encode() + b"\0" for h in headers)
        h = lib().or_build(_p8(text), len(text), _p32(bounds), len(bounds), hb, sample_rate,
                           lookup_depth, encoding, tag, 1 if naive_sa else 0, threads)
        return cls(h, text)

    @classmethod
    def build_rev(cls, text, *, sample_rate=32, encoding=0, tag=0, naive_sa=False, threads=1):
        text = np.ascontiguousarray(text, np.uint8)
        h = lib().or_build_rev(_p8(text), len(text), sample_rate, encoding, tag,
                               1 if naive_sa else 0, threads)
        return cls(h, None)

    def to_bytes(self) -> bytes:
        p = u8p()
        n = C.c_uint64()
        lib().or_to_bytes(self._h, C.byref(p), C.byref(n))
        out = C.string_at(p, n.value)
        lib().or_free_buf(C.cast(p, C.c_void_p))
        return out

    @classmethod
    def from_bytes(cls, data: bytes):
        buf = np.frombuffer(data, np.uint8) if len(data) else np.zeros(1, np.uint8)
        err = C.create_string_buffer(256)
        h = lib().or_from_bytes(_p8(buf), len(data), err, 256)
        if not h:
            raise ValueError(err.value.decode())
        return cls(h, None)

    # --- accessors
    @property
    def text_len(self) -> int:
        return lib().or_text_len(self._h)

    @property
    def num_lanes(self) -> int:
        return lib().or_num_lanes(self._h)

    @property
    def sample_rate(self) -> int:
        return lib().or_sample_rate(self._h)

    @property
    def lookup_depth(self) -> int:
        return lib().or_lookup_depth(self._h)

    @property
    def seq_boundaries(self) -> np.ndarray:
        n = lib().or_num_sequences(self._h)
        p = lib().or_seq_boundaries(self._h)
        return np.array(p[:n], dtype=np.uint32)

    @property
    def c_array(self) -> list[int]:
        buf = (C.c_uint32 * 16)()
        lib().or_c_array(self._h, buf)
        return list(buf)

    # --- primitives
    def rank(self, c: int, i: int) -> int:
        return lib().or_rank(self._h, c, i)

    def lf_step(self, pos: int) -> tuple[int, int]:
        s = C.c_uint8()
        r = C.c_uint32()
        lib().or_lf_step(self._h, pos, C.byref(s), C.byref(r))
        return s.value, r.value

    def symbol_at(self, pos: int) -> int:
        return lib().or_symbol_at(self._h, pos)

    def sa_get(self, row: int):
        v = C.c_uint32()
        return v.value if lib().or_sa_get(self._h, row, C.byref(v)) else None

    def resolve_sa(self, row: int) -> int:
        return lib().or_resolve_sa(self._h, row)

    def map_position(self, pos: int):
        s, o = C.c_uint32(), C.c_uint32()
        return (s.value, o.value) if lib().or_map_position(self._h, pos, C.byref(s), C.byref(o)) else None

    def backward_search(self, p) -> list[tuple[int, int]]:
        p = encode(p) if isinstance(p, str) else np.ascontiguousarray(p, np.uint8)
        buf = p if len(p) else np.zeros(1, np.uint8)
        cap = 64
        while True:
            out = np.empty(2 * cap, np.uint32)
            k = lib().or_backward_search(self._h, _p8(buf), len(p), _p32(out), cap)
            if k <= cap:
                return [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(k)]
            cap = k

    def count(self, p) -> int:
        p = encode(p) if isinstance(p, str) else np.ascontiguousarray(p, np.uint8)
        buf = p if len(p) else np.zeros(1, np.uint8)
        return lib().or_count(self._h, _p8(buf), len(p))

    def locate_batch(self, queries=None, *, qflat=None, qoff=None, threads=1, counts_only=False):
        """Per query (seq_idx, pos) pairs in FmIndex::locate order (query.rs:22-39).
        Returns (offsets[nq+1] uint64, hits[total,2] uint32)."""
        if queries is not None:
            qflat, qoff = to_csr(queries)
        qflat = np.ascontiguousarray(qflat, np.uint8)
        qoff = np.ascontiguousarray(qoff, np.uint64)
        nq = len(qoff) - 1
        out = _LocOut()
        lib().or_locate_batch(self._h, _p8(qflat if len(qflat) else np.zeros(1, np.uint8)),
                              _p64(qoff), nq, threads, 1 if counts_only else 0, C.byref(out))
        offs = np.ctypeslib.as_array(out.offsets, shape=(nq + 1,)).copy()
        tot = int(offs[-1]) if not counts_only else 0
        hits = (np.ctypeslib.as_array(out.hits, shape=(2 * tot,)).copy().reshape(-1, 2)
                if tot else np.zeros((0, 2), np.uint32))
        lib().or_locate_out_free(C.byref(out))
        return offs, hits

    def locate(self, p) -> list[tuple[int, int]]:
        offs, hits = self.locate_batch([p])
        return [tuple(map(int, h)) for h in hits]

    def resolve_rows(self, rows, threads=1) -> np.ndarray:
        rows = np.ascontiguousarray(rows, np.uint32)
        out = np.empty(len(rows), np.uint32)
        if len(rows):
            lib().or_resolve_rows(self._h, _p32(rows), len(rows), _p32(out), threads)
        return out


def extend_right(rev: OracleIndex, iv, c):
    a = _Biv(*iv)
    o = _Biv()
    return (o.fwd_lo, o.fwd_hi, o.rev_lo, o.rev_hi) if lib().or_extend_right(rev._h, C.byref(a), c, C.byref(o)) else None


def extend_left(fwd: OracleIndex, iv, c):
    a = _Biv(*iv)
    o = _Biv()
    return (o.fwd_lo, o.fwd_hi, o.rev_lo, o.rev_hi) if lib().or_extend_left(fwd._h, C.byref(a), c, C.byref(o)) else None


@dataclass
class OracleMem:
    query_start: int
    query_end: int
    match_count: int
    ivs: np.ndarray = field(repr=False)  # [k, 4] fwd_lo fwd_hi rev_lo rev_hi


class OracleBidir:
    """BidirFmIndex restated (bidir_index.rs + smem.rs)."""

    def __init__(self, fwd: OracleIndex, rev: OracleIndex):
        self.fwd, self.rev = fwd, rev

    @classmethod
    def build(cls, seqs=None, *, text=None, bounds=None, headers=None, sample_rate=32,
              lookup_depth=0, encoding=0, tag=0, naive_sa=False, threads=1):
        if seqs is not None:
            text, bounds = concatenate(seqs)
        fwd = OracleIndex.build(text=text, bounds=bounds, headers=headers, sample_rate=sample_rate,
                                lookup_depth=lookup_depth, encoding=encoding, tag=tag,
                                naive_sa=naive_sa, threads=threads)
        rev = OracleIndex.build_rev(text, sample_rate=sample_rate, encoding=encoding, tag=tag,
                                    naive_sa=naive_sa, threads=threads)
        return cls(fwd, rev)

    @classmethod
    def from_bytes(cls, data: bytes):
        """bidir_index.rs:195-210"""
        fl = int.from_bytes(data[:4], "little")
        return cls(OracleIndex.from_bytes(data[4:4 + fl]), OracleIndex.from_bytes(data[4 + fl:]))

    def to_bytes(self) -> bytes:
        """bidir_index.rs:183-192: [u32 LE fwd_len][fwd][rev]"""
        f, r = self.fwd.to_bytes(), self.rev.to_bytes()
        return len(f).to_bytes(4, "little") + f + r

    def find_mems_batch(self, queries=None, *, qflat=None, qoff=None, min_len=1, mode=0, threads=1):
        """mode 0 = find_smems (smem.rs:71-104), mode 1 = find_mems (smem.rs:113-124)."""
        if queries is not None:
            qflat, qoff = to_csr(queries)
        qflat = np.ascontiguousarray(qflat, np.uint8)
        qoff = np.ascontiguousarray(qoff, np.uint64)
        nq = len(qoff) - 1
        out = _MemOut()
        lib().or_find_mems_batch(self.fwd._h, self.rev._h,
                                 _p8(qflat if len(qflat) else np.zeros(1, np.uint8)), _p64(qoff),
                                 nq, min_len, mode, threads, C.byref(out))
        mo = np.ctypeslib.as_array(out.mem_off, shape=(nq + 1,)).copy()
        nm, ni = out.nmem, out.niv
        mems = np.ctypeslib.as_array(out.mems, shape=(4 * (nm + 1),)).copy().reshape(-1, 4)
        ivs = (np.ctypeslib.as_array(out.ivs, shape=(4 * ni,)).copy().reshape(-1, 4)
               if ni else np.zeros((0, 4), np.uint32))
        lib().or_mem_out_free(C.byref(out))
        res = []
        for q in range(nq):
            lst = []
            for k in range(int(mo[q]), int(mo[q + 1])):
                s, e, cnt, io = (int(x) for x in mems[k])
                lst.append(OracleMem(s, e, cnt, ivs[io:int(mems[k + 1][3])]))
            res.append(lst)
        return res

    def find_smems(self, q, min_len):
        return self.find_mems_batch([q], min_len=min_len, mode=0)[0]

    def find_mems(self, q, min_len):
        return self.find_mems_batch([q], min_len=min_len, mode=1)[0]

    def positions(self, mem: OracleMem, ref_boundaries=None, cap=None):
        """Positions of a Mem's intervals, rows in SA order per interval (bidir_index.rs:157-168).
        With ref_boundaries: (ref_id, offset) via the partition point of ref_map.wgsl:31-59."""
        rows = []
        for lo, hi, _, _ in mem.ivs:
            k = int(hi) - int(lo)
            if cap is not None:
                k = min(k, cap)
            rows.extend(range(int(lo), int(lo) + k))
        pos = self.fwd.resolve_rows(np.array(rows, np.uint32))
        b = self.fwd.seq_boundaries if ref_boundaries is None else np.asarray(ref_boundaries, np.uint32)
        return [ref_map(int(p), b) for p in pos]


def ref_map(pos: int, bounds) -> tuple[int, int]:
    """partition point: first i with bounds[i] > pos; offset = pos - bounds[i-1]."""
    i = int(np.searchsorted(np.asarray(bounds, np.uint32), pos, side="right"))
    start = int(bounds[i - 1]) if i > 0 else 0
    return i, pos - start


# --------------------------------------------------------------------------
# instrumentation, brute force, generators
# --------------------------------------------------------------------------
def probes_enable(on: bool):
    lib().or_probe_enable(1 if on else 0)


def probes_get() -> dict:
    p = _Probe()
    lib().or_probe_get(C.byref(p))
    return {"occ_probes": p.occ_probes, "sa_probes": p.sa_probes, "sa_val_reads": p.sa_val_reads,
            "steps": p.steps}


def probes_reset():
    lib().or_probe_reset()


def naive_locate(text, q, tag=0) -> list[int]:
    text = np.ascontiguousarray(text, np.uint8)
    q = encode(q) if isinstance(q, str) else np.ascontiguousarray(q, np.uint8)
    cap = len(text) + 1
    out = np.empty(cap, np.uint32)
    k = lib().or_naive_locate(_p8(text), len(text), _p8(q if len(q) else np.zeros(1, np.uint8)),
                              len(q), tag, _p32(out), cap)
    return [int(x) for x in out[:k]]


def brute_mems(text, q, min_len, tag=0) -> list[tuple[int, int, int]]:
    text = np.ascontiguousarray(text, np.uint8)
    q = encode(q) if isinstance(q, str) else np.ascontiguousarray(q, np.uint8)
    cap = len(q) * len(q) + 1
    out = np.empty(3 * cap, np.uint32)
    k = lib().or_brute_mems(_p8(text), len(text), _p8(q if len(q) else np.zeros(1, np.uint8)),
                            len(q), min_len, tag, _p32(out), cap)
    return [tuple(int(x) for x in out[3 * i:3 * i + 3]) for i in range(k)]


def gen_reference(seed, total, ncontig=1, gc=0.5, repeat_frac=0.0, n_units=0, unit_min=300,
                  unit_max=6000, div_min=0.05, div_max=0.15, iupac_frac=0.0, iupac_codes=None):
    codes = np.array(iupac_codes if iupac_codes is not None else IUPAC_AMBIG, np.uint8)
    tp, n, bp = u8p(), C.c_uint64(), u32p()
    rc = lib().or_gen_reference(seed, total, ncontig, gc, repeat_frac, n_units, unit_min, unit_max,
                                div_min, div_max, iupac_frac, _p8(codes), len(codes), C.byref(tp),
                                C.byref(n), C.byref(bp))
    if rc != 0:
        raise ValueError("reference generation failed")
    text = np.ctypeslib.as_array(tp, shape=(n.value,)).copy()
    bounds = np.ctypeslib.as_array(bp, shape=(ncontig,)).copy()
    lib().or_free_buf(C.cast(tp, C.c_void_p))
    lib().or_free_buf(C.cast(bp, C.c_void_p))
    return text, bounds


def gen_reads(seed, text, bounds, nreads, length, *, sigma=0.0, len_min=0, len_max=0, sub=0.0,
              indel=0.0, both_strands=False, n_random=0, gc=0.5):
    text = np.ascontiguousarray(text, np.uint8)
    bounds = np.ascontiguousarray(bounds, np.uint32)
    fp, op = u8p(), u64p()
    lib().or_gen_reads(seed, _p8(text), len(text), _p32(bounds), len(bounds), nreads, length, sigma,
                       len_min, len_max or length, sub, indel, 1 if both_strands else 0, n_random,
                       gc, C.byref(fp), C.byref(op))
    nq = nreads + n_random
    off = np.ctypeslib.as_array(op, shape=(nq + 1,)).copy()
    flat = np.ctypeslib.as_array(fp, shape=(max(int(off[-1]), 1),)).copy()
    lib().or_free_buf(C.cast(fp, C.c_void_p))
    lib().or_free_buf(C.cast(op, C.c_void_p))
    return flat, off
