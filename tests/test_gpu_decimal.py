"""GPU parity of the decimal container (src/decimal.rs, through capi's
quoin_typed_compress_decimal / quoin_typed_decompress): the container bytes
equal the oracle's at Level::Fast for Decimal32/64/128/256, with and without
nulls, and decode gives back the significands (nulls 0) and the bitmap."""
import numpy as np
import pytest

import _data
import _oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qa():
    import quoin_amd
    return quoin_amd


def _le(vals, w):
    return b"".join(int(v).to_bytes(w, "little", signed=True) for v in vals)


def cases():
    r = np.random.default_rng(51)
    return {
        "prices": [100_000 + i % 500 for i in range(4096)],
        "mixed_sign": [(i - 1000) * 7 for i in range(2000)],
        "wide128": [(i << 90) | (i * 0x1_0000_0001) for i in range(1000)],
        "extremes128": [-(1 << 127), 0, (1 << 127) - 1, -1, 1],
        "empty": [],
        "single": [-123_456_789_012_345_678],
        "random64": [int(x) for x in r.integers(-(1 << 62), 1 << 62, 70000)],
        "cents": [int(x) for x in r.integers(0, 10**7, 40000)],
    }


def _check(qa, name, vals, dtype, wire, ow, validity=None, scale=2, precision=30):
    raw = _le(vals, ow)
    n = len(vals)
    print(f"case decimal {name} dtype={int(dtype)} n={n} nulls={validity is not None}", flush=True)
    got = qa.compress_decimal(raw, dtype, validity, scale, precision)
    wide = _le(vals, 16 if wire == 6 else 32)
    # the typed ABI compresses the limbs at Config::default (capi/src/typed.rs:71-73)
    want = orc.compress_decimal(wide, wire, validity, scale, precision, level=orc.MAX)
    assert got == want, f"{name}: container differs ({len(got)} vs {len(want)} B)"
    back, vb = qa.decompress_decimal(got, dtype, n, with_validity=True)
    expect = list(vals)
    if validity is not None:
        expect = [v if (validity[i >> 3] >> (i & 7)) & 1 else 0 for i, v in enumerate(vals)]
    assert back == _le(expect, ow), f"{name}: decode"
    if validity is not None and n:
        exp_bm = validity.copy()
        if n % 8:
            exp_bm[-1] &= (1 << (n % 8)) - 1
        assert vb == exp_bm.tobytes()
    elif n:
        assert vb == np.packbits(np.ones(n, np.uint8), bitorder="little").tobytes()


def test_decimal128_and_256_match_oracle(qa):
    for name, vals in cases().items():
        _check(qa, name, vals, qa.DType.DECIMAL128, 6, 16)
        _check(qa, name + "_256", vals, qa.DType.DECIMAL256, 7, 32, scale=4, precision=60)
    big = [-(1 << 255), 0, (1 << 255) - 1, -1, 1] + [(i << 192) | (i * 0x9E3779B9) for i in range(300)]
    _check(qa, "wide256", big, qa.DType.DECIMAL256, 7, 32, scale=0, precision=76)


def test_decimal32_64_widen_to_128(qa):
    r = np.random.default_rng(52)
    v64 = [int(x) for x in r.integers(-(1 << 63), 1 << 63, 30000, dtype=np.int64)]
    v32 = [int(x) for x in r.integers(-(1 << 31), 1 << 31, 30000, dtype=np.int64)]
    _check(qa, "d64", v64, qa.DType.DECIMAL64, 6, 8)
    _check(qa, "d32", v32, qa.DType.DECIMAL32, 6, 4)


def test_decimal_nulls(qa):
    for density in (0.01, 0.3, 1.0):
        vals = [50_000 - i % 300 for i in range(20000)]
        _, bm = _data.with_nulls(np.zeros(len(vals), np.int64), density, seed=53)
        _check(qa, f"nulls{density}", vals, qa.DType.DECIMAL128, 6, 16, validity=bm)
        _check(qa, f"nulls{density}_256", vals, qa.DType.DECIMAL256, 7, 32, validity=bm)


def test_decimal_type_mismatch_and_damage(qa):
    vals = [100_000 + i % 500 for i in range(3000)]
    s = qa.compress_decimal(_le(vals, 16), qa.DType.DECIMAL128, None, 2, 18)
    for dt in (qa.DType.DECIMAL256, qa.DType.I64):   # the column type does not match
        with pytest.raises(qa.QuoinError):
            qa.decompress_decimal(s, dt, len(vals)) if dt == qa.DType.DECIMAL256 else \
                qa.decompress_column(s, dt, len(vals))
    lane = qa.compress_column(np.arange(100, dtype=np.int64))
    with pytest.raises(qa.QuoinError):                 # a decimal asked of a lane stream
        qa.decompress_decimal(lane, qa.DType.DECIMAL128, 100)
    with pytest.raises(qa.QuoinError):
        qa.decompress_decimal(s, qa.DType.DECIMAL128, len(vals) + 1)
    for cut in (18, 30, len(s) - 1):
        with pytest.raises(qa.QuoinError):
            qa.decompress_decimal(s[:cut], qa.DType.DECIMAL128, len(vals))
