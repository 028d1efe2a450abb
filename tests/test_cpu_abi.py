"""The C-ABI library without a GPU: it loads, exports every symbol that
include/quoin.h declares, answers the host-only queries like the reference
(capi/src/lib.rs, capi/src/typed.rs), and fails loudly instead of silently
computing on the CPU."""
import ctypes
import os
import re
import struct

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def qa():
    import quoin_amd
    return quoin_amd


def header_functions():
    text = open(os.path.join(ROOT, "include", "quoin.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(quoin_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(qa):
    names = header_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(qa.lib, name), name
    assert set(names) == set(qa.EXPORTS)


def test_reference_symbols_present(qa):
    # the 14 symbols of capi/include/quoin.h
    ref = ["quoin_version", "quoin_compress_bound", "quoin_compress", "quoin_decompress",
           "quoin_decompressed_value_count", "quoin_ctx_create", "quoin_ctx_free",
           "quoin_compress_ctx", "quoin_decompress_ctx", "quoin_dtype_width",
           "quoin_typed_compress_bound", "quoin_typed_compress", "quoin_typed_compress_decimal",
           "quoin_typed_decompress"]
    for name in ref:
        assert name in header_functions(), name


def test_host_queries(qa):
    lib = qa.lib
    assert qa.version().startswith("quoin")
    for n in (0, 1, 1000, 1 << 30):
        assert lib.quoin_compress_bound(n) == 9 * n + 1024  # capi/src/lib.rs:37-40
    widths = {0: 1, 1: 1, 2: 2, 3: 4, 4: 8, 5: 1, 6: 2, 7: 4, 8: 8, 9: 4, 10: 8, 11: 4, 12: 8,
              13: 16, 14: 32, 15: 0, -1: 0}
    for d, w in widths.items():
        assert lib.quoin_dtype_width(d) == w, d  # capi/src/typed.rs:51-62
    for d in (0, 3, 10, 13, 14):
        w = max(widths[d], 8)
        assert lib.quoin_typed_compress_bound(d, 1000) == 1000 * (w + 1) + 1024


def test_value_count_from_header(qa):
    hdr = b"FCR1" + bytes([2, 0, 16, 0]) + struct.pack("<Q", 12345)
    buf = np.frombuffer(hdr, dtype=np.uint8)
    assert qa.lib.quoin_decompressed_value_count(ctypes.c_void_p(buf.ctypes.data), 16) == 12345
    bad = np.frombuffer(b"XCR1" + hdr[4:], dtype=np.uint8)
    assert qa.lib.quoin_decompressed_value_count(ctypes.c_void_p(bad.ctypes.data), 16) == \
        qa.ERR_CORRUPT
    assert qa.lib.quoin_decompressed_value_count(ctypes.c_void_p(buf.ctypes.data), 15) == \
        qa.ERR_CORRUPT
    assert qa.lib.quoin_decompressed_value_count(None, 16) == qa.ERR_NULL


def test_null_arguments(qa):
    lib = qa.lib
    got = ctypes.c_size_t()
    assert lib.quoin_compress(None, 1, None, 0, ctypes.byref(got)) == qa.ERR_NULL
    assert lib.quoin_typed_compress(4, 1, None, None, None, 0) == 0
    assert lib.quoin_typed_decompress(4, 1, None, 0, None, None) == qa.ERR_NULL
    assert lib.quoin_typed_compress_decimal(13, 1, None, None, 2, 10, None, 0) == 0


def test_last_error_names_the_null_argument(qa):
    # quoin_last_error is the code behind the last 0 returned on this thread:
    # a null argument must replace whatever an earlier call left
    lib = qa.lib
    v = np.arange(4, dtype=np.int64)
    dst = np.zeros(64, dtype=np.uint8)
    vp, dp = ctypes.c_void_p(v.ctypes.data), ctypes.c_void_p(dst.ctypes.data)
    assert lib.quoin_typed_compress(99, 4, vp, None, dp, 64) == 0  # unknown dtype first
    assert lib.quoin_last_error() == qa.ERR_UNSUPPORTED
    assert lib.quoin_typed_compress(1, 4, None, None, dp, 64) == 0
    assert lib.quoin_last_error() == qa.ERR_NULL
    assert lib.quoin_typed_compress(99, 4, vp, None, dp, 64) == 0
    assert lib.quoin_typed_compress_decimal(13, 4, None, None, 2, 10, dp, 64) == 0
    assert lib.quoin_last_error() == qa.ERR_NULL


def test_decompress_leaves_out_values_on_error(qa):
    # capi/src/lib.rs:80-110: *out_values is written only after a good decode
    got = ctypes.c_size_t(777)
    junk = np.frombuffer(b"not a quoin stream", dtype=np.uint8)
    out = np.zeros(4, dtype=np.float64)
    rc = qa.lib.quoin_decompress(ctypes.c_void_p(junk.ctypes.data), len(junk),
                                 ctypes.c_void_p(out.ctypes.data), 4, ctypes.byref(got))
    assert rc == qa.ERR_CORRUPT and got.value == 777


def test_no_gpu_fails_loudly(qa):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    v = np.arange(100, dtype=np.int64)
    with pytest.raises(qa.QuoinError) as e:
        qa.compress_column(v, None, qa.Config(qa.Level.FASTEST))
    assert e.value.code == qa.ERR_DEVICE
    assert qa.Gpu.__init__ is not None
    with pytest.raises(qa.QuoinError):
        qa.Gpu(0)


def test_every_level_needs_the_device(qa):
    # Balanced/High/Max plan and size their blocks on the device before the
    # host coders run: without a GPU they fail loudly like Fastest and Fast
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    v = np.arange(100, dtype=np.int64)
    for lvl in (qa.Level.BALANCED, qa.Level.HIGH, qa.Level.MAX):
        with pytest.raises(qa.QuoinError) as e:
            qa.compress_column(v, None, qa.Config(lvl))
        assert e.value.code == qa.ERR_DEVICE


def test_unknown_selection_is_refused(qa):
    # a selection other than Full (0) and Sample (1) is refused before any
    # device work; Sample itself needs the device like every level
    import torch
    v = np.arange(100, dtype=np.int64)
    with pytest.raises(qa.QuoinError) as e:
        qa.compress_column(v, None, qa.Config(qa.Level.MAX, selection=7))
    assert e.value.code == qa.ERR_UNSUPPORTED
    if not torch.cuda.is_available():
        with pytest.raises(qa.QuoinError) as e:
            qa.compress_column(v, None, qa.Config(qa.Level.MAX, selection=qa.Selection.SAMPLE))
        assert e.value.code == qa.ERR_DEVICE
