"""C-ABI boundary checks that need no GPU: the library loads, exports every
entry point include/hsfm.h declares, and parses FmIndex::to_bytes streams."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import haystackfm_amd as hf
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "hsfm.h")).read()
    return sorted(set(re.findall(r"\b(hsfm_[a-z_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = hf._lib()
    names = _declared_symbols()
    assert "hsfm_locate_batch" in names and "hsfm_find_mems_batch" in names
    for n in names:
        assert hasattr(lib, n), n


def _view(data: bytes):
    buf = np.frombuffer(data, np.uint8)
    v = hf._View()
    rc = hf._lib().hsfm_index_view_from_bytes(buf.ctypes.data_as(C.POINTER(C.c_uint8)), len(data), C.byref(v))
    return rc, v, buf


def test_view_from_bytes_matches_oracle_fields():
    idx = O.OracleIndex.build(["ACGTNACGTRYK", "TTGCA"], headers=["chr1", "chr2"], sample_rate=4,
                              lookup_depth=3)
    data = idx.to_bytes()
    rc, v, _ = _view(data)
    assert rc == 0
    assert list(v.c_array) == idx.c_array
    assert v.text_len == idx.text_len
    assert v.occ_num_lanes == idx.num_lanes
    assert v.sa_sample_rate == 4
    assert v.has_lookup == 1 and v.lookup_depth == 3 and v.lookup_intervals_len == 64
    assert v.num_sequences == 2
    assert list(np.frombuffer(C.string_at(v.seq_boundaries, 8), np.uint32)) == list(idx.seq_boundaries)
    for b in range(256):
        assert v.compat_mask[b] == (O.compat_mask(b) if b < 16 else 0)
    fm = hf.FmIndex.from_bytes(data)
    assert fm.seq_headers == ["chr1", "chr2"]


def test_view_from_bytes_exact_alphabet():
    data = O.OracleIndex.build(["ACGTNACGT"], sample_rate=1, tag=1).to_bytes()
    rc, v, _ = _view(data)
    assert rc == 0 and v.alphabet_tag == 1
    for b in range(256):
        assert v.compat_mask[b] == (O.compat_mask(b, tag=1) if b < 16 else 0)


def test_bad_tag_and_truncation_are_deserialize_errors():
    data = bytearray(O.OracleIndex.build(["ACGT"], sample_rate=4).to_bytes())
    data[-1] = 0xFF  # serialize.rs:137-145
    rc, _, _ = _view(bytes(data))
    assert rc == -2
    with pytest.raises(hf.DeserializeError):
        hf.FmIndex.from_bytes(bytes(data))
    good = O.OracleIndex.build(["ACGT"], sample_rate=4).to_bytes()
    for cut in (0, 10, 70, len(good) - 1):
        rc, _, _ = _view(good[:cut])
        assert rc == -2, cut
    with pytest.raises(hf.DeserializeError):
        hf.BidirFmIndex.from_bytes(b"\x01\x00")


def test_gpu_methods_fail_loudly_without_device(monkeypatch):
    """No GPU in this container: opening the device must raise GpuError
    rather than fall back to the CPU."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except Exception:
        pass
    fm = hf.FmIndex.from_bytes(O.OracleIndex.build(["ACGT"], sample_rate=1).to_bytes())
    hf.context_cache._ctx = None
    with pytest.raises(hf.GpuError):
        fm.locate_gpu(["ACG"])
