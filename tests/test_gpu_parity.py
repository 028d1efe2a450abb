"""GPU stream parity: the HIP encoder's bytes equal the oracle's, and the HIP
decoder round-trips its own and the oracle's streams bit for bit."""
import numpy as np
import pytest

import _data
import _oracle as orc

pytestmark = pytest.mark.gpu

ENGINE = {"F64": orc.F64, "I64": orc.I64, "U64": orc.U64, "I32": orc.I32, "U32": orc.U32,
          "F32": orc.F32}


@pytest.fixture(scope="module")
def qa():
    import quoin_amd
    return quoin_amd


def engine_dtype(arr):
    return {np.dtype(np.float64): orc.F64, np.dtype(np.int64): orc.I64,
            np.dtype(np.uint64): orc.U64, np.dtype(np.int32): orc.I32,
            np.dtype(np.uint32): orc.U32, np.dtype(np.float32): orc.F32}[arr.dtype]


def check(qa, name, vals, validity=None, level=0, block_size=0):
    print(f"case {name} n={len(vals)} dtype={vals.dtype} nulls={validity is not None} "
          f"level={level} bs={block_size}", flush=True)
    cfg = qa.Config(level=qa.Level(level), block_size=block_size)
    got = qa.compress_column(vals, validity, cfg)
    want = orc.compress(vals, engine_dtype(vals), validity, level, block_size)
    assert got == want, f"{name}: stream differs (gpu {len(got)} B, oracle {len(want)} B)"
    out, vout = qa.decompress_column(got, qa.FROM_NP[vals.dtype], len(vals),
                                     with_validity=validity is not None)
    # the oracle's decode is the reference behaviour (f32 signaling NaNs come
    # back quieted: src/dtype.rs:40-48); everything else is the input itself
    _, want_vals, _ = orc.decompress(want)
    assert out.view(np.uint8).tobytes() == want_vals.view(np.uint8).tobytes(), f"{name}: decode"
    if vals.dtype != np.float32:
        assert out.view(np.uint8).tobytes() == vals.view(np.uint8).tobytes(), f"{name}: round trip"
    if validity is not None:
        n = len(vals)
        exp = validity.copy()
        if n % 8:
            exp[-1] &= (1 << (n % 8)) - 1
        assert vout.tobytes() == exp.tobytes()


SIZES = [0, 1, 2, 15, 16, 1023, 1024, 1025, 32767, 32768, 32769, 70000, 131072 + 5]


@pytest.mark.parametrize("n", SIZES)
def test_int_shapes_fastest(qa, n):
    for name, v in _data.int_columns(n, seed=n).items():
        check(qa, name, v)


@pytest.mark.parametrize("n", [1, 1000, 40000, 140000])
def test_float_shapes_fastest(qa, n):
    for name, v in _data.float_columns(n, seed=n).items():
        check(qa, name, v)


def test_nulls_fastest(qa):
    for density in (0.0, 0.01, 0.5, 1.0):
        for name, v in _data.int_columns(50000, seed=3).items():
            vv, bm = _data.with_nulls(v, density, seed=7)
            check(qa, f"{name}+nulls{density}", vv, bm)


def test_fixed_block_sizes(qa):
    for bs in (1, 100, 1000, 1024, 5000, 131072, 200000):
        v = _data.int_columns(20000, seed=bs)["ts_i64"]
        check(qa, f"bs{bs}", v, block_size=bs)


def test_special_f32(qa):
    check(qa, "special_f32", _data.special_f32())


def test_config1_shapes(qa):
    """Integer lanes at a reduced size (the bench runs 16 Mi per column)."""
    n = 1 << 20
    cols = _data.int_columns(n, seed=0)
    for name in ("ts_i64", "u32_k3", "u32_k11", "u32_k23"):
        v = cols[name]
        if name.startswith("u32"):
            v, bm = _data.with_nulls(v, 0.01, seed=11)
            check(qa, name, v, bm)
        else:
            check(qa, name, v)


def test_corrupt_streams_report_errors(qa):
    """Mutated and truncated streams never fault the decoder: they decode or
    return an error code (tests/robustness.rs:40-52, 77-100)."""
    r = np.random.default_rng(3)
    v = _data.int_columns(5000, seed=5)["u32_k11"]
    vv, bm = _data.with_nulls(v, 0.05, seed=2)
    seeds = [(qa.compress_column(vv, bm, qa.Config(qa.Level.FASTEST)), qa.DType.U32, len(vv)),
             (qa.compress_column(_data.int_columns(5000, 1)["ts_i64"], None,
                                 qa.Config(qa.Level.FASTEST)), qa.DType.I64, 5000)]
    for s, dt, n in seeds:
        for it in range(150):
            b = bytearray(s)
            k = int(r.integers(0, len(b)))
            if it % 3 == 0:
                b = b[:k]
            else:
                b[k] ^= 1 << int(r.integers(0, 8))
            try:
                qa.decompress_column(bytes(b), dt, n, True)
            except qa.QuoinError as e:
                assert e.code in (qa.ERR_CORRUPT, qa.ERR_UNSUPPORTED), e.code
    for it in range(200):
        b = bytes(r.integers(0, 256, int(r.integers(0, 80)), dtype=np.uint8))
        try:
            qa.decompress_column(b, qa.DType.F64, 10)
        except qa.QuoinError as e:
            assert e.code == qa.ERR_CORRUPT


def _blob_range(stream):
    """(start, end) of the validity blob: after the 16-byte header, varint
    length then the blob (src/format.rs:24-94)."""
    pos, shift, vlen = 16, 0, 0
    while True:
        b = stream[pos]
        pos += 1
        vlen |= (b & 0x7F) << shift
        shift += 7
        if not b & 0x80:
            break
    return pos, pos + vlen


def test_validity_blob_damage_matches_oracle(qa):
    """Damage inside the validity blob (runs or raw bitmap): the decoder gives
    the oracle's values and bitmap when the oracle decodes, and CORRUPT when it
    does not (src/validity.rs:73-119)."""
    r = np.random.default_rng(17)
    v = _data.int_columns(6000, seed=7)["u32_k11"]
    n = len(v)
    tags = set()
    for frac, seed in ((0.03, 4), (0.5, 5)):  # runs blob, raw bitmap blob
        vv, bm = _data.with_nulls(v, frac, seed=seed)
        s = orc.compress(vv, orc.U32, bm, orc.FASTEST)
        b0, b1 = _blob_range(s)
        tags.add(s[b0])
        muts = []
        for tag in (0, 1, 2, 0x81):
            muts.append([(b0, tag)])
        for _ in range(120):
            k = int(r.integers(b0 + 1, b1))
            kind = int(r.integers(0, 3))
            val = (s[k] ^ (1 << int(r.integers(0, 8)))) if kind == 0 else (s[k] | 0x80 if kind == 1 else 0)
            muts.append([(k, val)])
        print(f"case validity damage tag={s[b0]} blob={b1 - b0} B mutations={len(muts)}", flush=True)
        ok = bad = 0
        for i, m in enumerate(muts):
            b = bytearray(s)
            for k, val in m:
                b[k] = val
            b = bytes(b)
            try:
                _, want_vals, want_vb = orc.decompress(b)
            except ValueError:
                want_vals = None
            try:
                out, vout = qa.decompress_column(b, qa.DType.U32, n, True)
            except qa.QuoinError as e:
                assert want_vals is None and e.code == qa.ERR_CORRUPT, (i, m, e.code)
                bad += 1
                continue
            assert want_vals is not None, (i, m)
            assert out.tobytes() == want_vals.tobytes(), (i, m)
            exp = np.frombuffer(want_vb, dtype=np.uint8).copy()
            if n % 8:
                exp[-1] &= (1 << (n % 8)) - 1
            assert vout.tobytes() == exp.tobytes(), (i, m)
            ok += 1
        assert ok and bad, (ok, bad)
    assert tags == {0, 1}, tags


def test_empty_column_validity_blob_is_checked(qa):
    """n = 0 with the validity flag: validity::decode still reads the blob's
    tag (src/validity.rs:73-81), so an empty blob or an unknown tag is
    corrupt, as the oracle says, and a raw or runs tag decodes."""
    import struct
    hdr = b"FCR1" + bytes([2, 1, 16, orc.U32]) + struct.pack("<Q", 0)
    for blob, ok in ((b"", False), (b"\x02", False), (b"\x81", False), (b"\x00", True),
                     (b"\x01", True)):
        s = hdr + bytes([len(blob)]) + blob
        print(f"case empty-column blob {blob.hex()}", flush=True)
        try:
            orc.decompress(s)
            want_ok = True
        except ValueError:
            want_ok = False
        assert want_ok == ok, blob
        if ok:
            out, vout = qa.decompress_column(s, qa.DType.U32, 0, True)
            assert len(out) == 0
        else:
            with pytest.raises(qa.QuoinError) as e:
                qa.decompress_column(s, qa.DType.U32, 0, True)
            assert e.value.code == qa.ERR_CORRUPT


def test_too_small_decompress_still_decodes_first(qa):
    """quoin_decompress with a small dst decodes first (capi/src/lib.rs:80-110):
    a corrupt payload is CORRUPT, not BUFFER_TOO_SMALL, and *out_values is
    written only after a good decode."""
    import ctypes
    v = _data.float_columns(5000, seed=3)["walk_f64"]
    s = bytearray(qa.compress_column(v, None, qa.Config(qa.Level.FAST)))
    out = np.zeros(10, dtype=np.float64)
    for damage, want in ((None, qa.ERR_BUFFER_TOO_SMALL), (len(s) - 20, qa.ERR_CORRUPT)):
        b = bytearray(s)
        if damage is not None:
            b = b[:damage]  # truncated: the last frame's payload is cut short
            with pytest.raises(ValueError):
                orc.decompress(bytes(b))
        buf = np.frombuffer(bytes(b), dtype=np.uint8)
        got = ctypes.c_size_t(99)
        rc = qa.lib.quoin_decompress(ctypes.c_void_p(buf.ctypes.data), len(buf),
                                     ctypes.c_void_p(out.ctypes.data), 10, ctypes.byref(got))
        print(f"case too-small damage={damage} rc={rc}", flush=True)
        assert rc == want
        assert got.value == (len(v) if want == qa.ERR_BUFFER_TOO_SMALL else 99)
