"""Level::Fast stream parity: the HIP encoder's bytes equal the oracle's for
every codec of the Fast pool, and the HIP decoder round-trips them."""
import numpy as np
import pytest

import _data
import _oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qa():
    import quoin_amd
    return quoin_amd


ODT = {np.dtype(np.float64): orc.F64, np.dtype(np.int64): orc.I64,
       np.dtype(np.uint64): orc.U64, np.dtype(np.int32): orc.I32,
       np.dtype(np.uint32): orc.U32, np.dtype(np.float32): orc.F32}
MODES = {v: k for k, v in orc.MODE.items()}


def enc(qa, name, vals, validity=None, block_size=0, decode=True):
    print(f"case {name} n={len(vals)} dtype={vals.dtype} nulls={validity is not None} "
          f"bs={block_size}", flush=True)
    cfg = qa.Config(level=qa.Level.FAST, block_size=block_size)
    got = qa.compress_column(vals, validity, cfg)
    want = orc.compress(vals, ODT[vals.dtype], validity, orc.FAST, block_size)
    if got != want:
        gf, wf = orc.frames(got) if got[:4] == b"FCR1" else [], orc.frames(want)
        for i, (g, w) in enumerate(zip(gf, wf)):
            if g != w:
                raise AssertionError(
                    f"{name}: frame {i} gpu mode {MODES.get(g[0])} len {len(g[2])} vs oracle "
                    f"{MODES.get(w[0])} len {len(w[2])}")
        raise AssertionError(f"{name}: stream differs ({len(got)} vs {len(want)} B)")
    if decode:
        out, _ = qa.decompress_column(got, qa.FROM_NP[vals.dtype], len(vals),
                                      with_validity=validity is not None)
        _, ov, _ = orc.decompress(want)
        assert out.view(np.uint8).tobytes() == ov.view(np.uint8).tobytes(), f"{name}: decode"
    return got


def modes_of(stream):
    return {MODES.get(m) for m, _, _ in orc.frames(stream)}


DECODE = True


@pytest.mark.parametrize("n", [1, 15, 1000, 40000, 140000])
def test_int_shapes_fast(qa, n):
    for name, v in _data.int_columns(n, seed=n).items():
        enc(qa, name, v, decode=DECODE)


@pytest.mark.parametrize("n", [1, 3000, 40000, 140000])
def test_float_shapes_fast(qa, n):
    for name, v in _data.float_columns(n, seed=n).items():
        enc(qa, name, v, decode=DECODE)


def test_codec_coverage_fast(qa):
    """Columns built so that each Fast codec wins somewhere."""
    r = np.random.default_rng(7)
    n = 70000
    seen = set()
    cases = {
        "dict_scattered": r.choice(r.integers(0, 1 << 60, 40).astype(np.uint64), n),
        "rle_runs": np.repeat(r.integers(0, 1 << 60, n // 50 + 1).astype(np.uint64), 50)[:n],
        "geo_runs_dec": np.repeat(r.integers(0, 10**6, n).astype(np.float64) / 1000,
                                  r.geometric(1 / 8, n))[:n],
        "xorz_small": np.arange(10, dtype=np.int64) * 3,
        "decimal2": np.round(r.normal(1000, 50, n), 2),
        "decimal3_wide": r.integers(0, 10**9, n) / 1000.0,
        "fm_smooth": 1000.0 + np.arange(n) / 1000.0,
        "alprd_unif": r.integers(0, 2**53, n, dtype=np.uint64) * 2.0**-53,
        "alprd_coords": (40.0 + r.random(n) * 1e-3),
        "f32_dec": (r.integers(0, 1000, n) / 10).astype(np.float32),
        "specials": np.array([0.0, -0.0, np.nan, np.inf, -np.inf, 1e308, 5e-324] * 100),
        "dict_card_mid": r.integers(0, 3000, n).astype(np.int64) * 1_000_003,
        "dict_card_hi": r.integers(0, 20000, 131072).astype(np.int64) * 7,
        # distinct counts above the LDS set limit (6144) and within n / 2 + 1:
        # the value-range path, a skewed split that falls back to the global
        # table, and a grown (repeats) block with 32768 distinct values
        "dict_card_ranged": r.integers(0, 9000, n).astype(np.int64) * 13,
        "dict_card_skewed": r.permutation(np.concatenate([
            np.repeat(r.integers(0, 100, 100).astype(np.int64), 260),
            1_000_000 + np.arange(6700, dtype=np.int64) * 3]))[: 32768],
        "dict_grown_repeats": np.repeat(r.permutation(40000).astype(np.int64) * 5, 4)[:131072],
        "alprd_normal_exc": np.concatenate([r.normal(0, 1, n - 50), r.random(50) * 1e300]),
        # ALP with exceptions in most sub-blocks: every 97th value full precision
        "alp_exc": np.where(np.arange(n) % 97 == 5, r.random(n) * 1e3, np.round(r.normal(1000, 50, n), 2)),
        "alp_exc_specials": np.where(np.arange(n) % 331 == 7, np.nan, np.round(r.normal(0, 9, n), 1)),
    }
    for name, v in cases.items():
        s = enc(qa, name, v, decode=DECODE)
        seen |= modes_of(s)
    print("modes seen:", sorted(x for x in seen if x))
    for m in ("DICT", "RLE", "ALP", "ALP_RD", "FOR_BITPACK", "DELTA_BITPACK", "FLOAT_MULT"):
        assert m in seen, m


def test_nulls_fast(qa):
    r = np.random.default_rng(8)
    for density in (0.01, 0.3):
        for name, v in list(_data.float_columns(30000, seed=4).items()) + \
                list(_data.int_columns(30000, seed=4).items()):
            vv, bm = _data.with_nulls(v, density, seed=9)
            enc(qa, f"{name}+nulls{density}", vv, bm, decode=DECODE)


def test_plain_workload_walk(qa):
    """The plain-method workload: 1 Mi f64 values round(100 (1000 + cumsum of
    N(0, 1) steps)) / 100, seed 0, half away from zero; ALP takes the blocks at
    Level::Fast. Streams equal the oracle's at both GPU levels."""
    r = np.random.default_rng(0)
    x = 100 * (1000 + np.cumsum(r.normal(size=1 << 20)))
    v = np.sign(x) * np.floor(np.abs(x) + 0.5) / 100
    s = enc(qa, "walk_1Mi_fast", v, decode=DECODE)
    assert modes_of(s) == {"ALP"}, modes_of(s)
    cfg = qa.Config(level=qa.Level.FASTEST)
    assert qa.compress_column(v, None, cfg) == orc.compress(v, orc.F64, None, orc.FASTEST)


def _tiles(n, makers):
    """Tile k (1024 values) from makers[k % len(makers)]."""
    return np.concatenate([makers[(i // 1024) % len(makers)](min(1024, n - i))
                           for i in range(0, n, 1024)])


def test_float_mult_tiles_with_different_first_scales(qa):
    """FloatMult's block scale is the first scale every value passes
    (float_mult.rs:24-43); here tiles pass first at different scales, so the
    block's scale is above some tiles' first, and in the last case one kind of
    tile fails every scale the other kind passes."""
    r = np.random.default_rng(23)
    n = 40000
    one = lambda m: r.integers(0, 2**30, m) / 10          # first passes at 10
    three = lambda m: r.integers(0, 2**40, m) / 1000      # 1000; ALP bails (w > 32)
    four = lambda m: r.integers(0, 2**40, m) / 10000
    big = lambda m: (9.2e15 + r.integers(0, 10**6, m)).astype(np.float64)  # fails at 1000+
    small3 = lambda m: r.integers(0, 10**6, m) / 1000
    for name, makers, scale in (("fm_1_3", [one, three], 2), ("fm_3_1_1_1", [three, one, one, one], 2),
                                ("fm_1_1_1_4", [one, one, one, four], 3),
                                ("fm_big_small3", [big, small3], None)):
        v = _tiles(n, makers)
        st = enc(qa, name, v, decode=DECODE)
        if scale is not None:
            fr = orc.frames(st)
            assert all(MODES.get(m) == "FLOAT_MULT" and p[1] == scale for m, _, p in fr), name


def test_alp_exceptions_with_nulls_fast(qa):
    """ALP tiles with exceptions gathered through a null bitmap."""
    r = np.random.default_rng(11)
    n = 50000
    v = np.where(np.arange(n) % 89 == 3, r.random(n) * 7, np.round(r.normal(50, 5, n), 2))
    for density in (0.01, 0.3):
        vv, bm = _data.with_nulls(v, density, seed=12)
        s = enc(qa, f"alp_exc+nulls{density}", vv, bm, decode=DECODE)
        assert "ALP" in modes_of(s)


def _varint(v):
    out = bytearray()
    while True:
        b = v & 0x7F
        v >>= 7
        out.append(b | (0x80 if v else 0))
        if not v:
            return bytes(out)


def test_rle_frame_zero_and_long_runs(qa):
    """A hand-built RLE frame (rle.rs:45-75) with zero-length runs, runs that
    span several 1024-value tiles and runs of 1: the device decode equals the
    oracle's and the expansion written out directly."""
    import struct
    r = np.random.default_rng(13)
    n = 100_000
    lens = []
    while sum(lens) < n:
        k = r.integers(0, 4)
        lens.append(int(0 if k == 0 else r.integers(1, 4) if k == 1 else r.integers(1, 40)
                        if k == 2 else r.integers(1000, 5000)))
    lens[-1] -= sum(lens) - n
    vals = r.integers(0, 1 << 63, len(lens), dtype=np.uint64)
    payload = (_varint(len(lens)) + vals.astype("<u8").tobytes()
               + orc.encode_mode("FOR_BITPACK", np.array(lens, dtype=np.uint64)))
    head = orc.compress(np.zeros(n, np.uint64), orc.U64, None, orc.FAST)[:16]
    stream = head + struct.pack("<BII", 53, n, len(payload)) + payload
    want = np.repeat(vals, lens)
    _, ov, _ = orc.decompress(stream)
    assert ov.tobytes() == want.tobytes()
    out, _ = qa.decompress_column(stream, qa.DType.U64, n)
    assert out.tobytes() == want.tobytes()


def test_fixed_block_sizes_fast(qa):
    v = _data.float_columns(20000, seed=2)["walk_f64"]
    for bs in (1, 100, 1000, 1024, 5000, 131072):
        enc(qa, f"walk bs{bs}", v, block_size=bs, decode=DECODE)


def test_dict_above_lds_limit_keeps_the_all_ones_value(qa):
    """A block of about 9,600 distinct 40-bit values (above the LDS set's
    limit, below n / 2 + 1, so Dict is sized by value ranges) holding the
    all-ones value (i64 -1) only late in the block: the dictionary must count
    it. Dict wins here; the stream must equal the oracle's."""
    r = np.random.default_rng(11)
    n = 32768
    words = r.integers(0, 2**40, 10_000, dtype=np.int64)
    v = words[r.integers(0, len(words), n)]
    v[30_000] = -1
    s = enc(qa, "dict_allones_late", v)
    assert "DICT" in modes_of(s), modes_of(s)
    # the same values with the all-ones value early (found before the LDS set overflows)
    v2 = v.copy()
    v2[30_000] = v[0]
    v2[5] = -1
    enc(qa, "dict_allones_early", v2)
