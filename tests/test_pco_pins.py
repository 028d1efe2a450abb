"""pco (vendor/quoin-pco, the Mode::Pco block codec) on host cores, pinned on
the CPU.

The reference ships no pco byte vectors, so the restatement is pinned by the
known answers of pco's own unit tests, function by function, by properties
its stability tests check (every truncated stream is refused), by round trips
of every number type at both compression levels quoin uses, and by the
size guarantee of wrapped/guarantee.rs. pco's bytes on general data beyond
these are "parity unpinned": no run of the reference confirms them.
"""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "quoin_amd", "libquoin_amd.so")

U32, U64, I32, I64, F32, F64 = 1, 2, 3, 4, 5, 6
u64p = C.POINTER(C.c_uint64)
f32p = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("libquoin_amd.so not built")
    lib = C.CDLL(LIB)
    lib.qa_pco_sample_f32.restype = C.c_size_t
    lib.qa_pco_quantize_weights.restype = C.c_uint32
    lib.qa_pco_quant_k_f32.restype = C.c_uint32
    lib.qa_pco_quant_k_f64.restype = C.c_uint32
    lib.qa_pco_log2_approx.restype = C.c_float
    lib.qa_pco_log2_approx.argtypes = [C.c_float]
    lib.qa_pco_optimize_bins_u32.restype = C.c_size_t
    lib.qa_pco_histogram_u32.restype = C.c_size_t
    lib.qa_pco_lookbacks_u32.restype = C.c_size_t
    lib.qa_pco_pair_gcd_f64.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double)]
    lib.qa_pco_pair_gcd_f32.argtypes = [C.c_float, C.c_float, C.POINTER(C.c_float)]
    lib.qa_pco_snap_f64.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    return lib


def compress(lib, nt, level, lane):
    lane = np.ascontiguousarray(lane, dtype=np.uint64)
    cap = len(lane) * 16 + 4096
    dst = (C.c_uint8 * cap)()
    ol = C.c_size_t()
    rc = lib.qa_pco_compress(nt, level, lane.ctypes.data_as(u64p), len(lane), dst, cap, C.byref(ol))
    assert rc == 0, rc
    return bytes(dst[:ol.value])


def decompress(lib, nt, b, n):
    out = np.zeros(max(n, 1), dtype=np.uint64)
    buf = (C.c_uint8 * max(1, len(b))).from_buffer_copy(b if b else b"\0")
    rc = lib.qa_pco_decompress(nt, buf, len(b), n, out.ctypes.data_as(u64p))
    return rc, out[:n]


def describe(lib, nt, level, lane):
    lane = np.ascontiguousarray(lane, dtype=np.uint64)
    info = (C.c_uint64 * 4)()
    assert lib.qa_pco_describe(nt, level, lane.ctypes.data_as(u64p), len(lane), info) == 0
    return list(info)


def f32_ordered(x):
    b = int(np.array([x], np.float32).view(np.uint32)[0])
    return (~b) & 0xFFFFFFFF if b >> 31 else b ^ 0x80000000


def f32_from_ordered(l):
    l &= 0xFFFFFFFF
    b = l ^ 0x80000000 if l & 0x80000000 else (~l) & 0xFFFFFFFF
    return np.array([b], np.uint32).view(np.float32)[0]


def test_xoroshiro_sample_kat(lib):
    # sampling.rs test_choose_sample: 150 numbers -> 13 sampled, smallest three
    # of the kept negatives -147, -142, -119 (here as magnitudes: the float
    # filter keeps |x|, and only 0.0 is dropped by both filters)
    x = np.array([-float(i) for i in range(150)], dtype=np.float32)
    out = (C.c_float * 200)()
    k = lib.qa_pco_sample_f32(x.ctypes.data_as(f32p), 150, out, 200)
    assert k == 13
    assert sorted(out[:k])[-3:] == [119.0, 142.0, 147.0]
    # calc_sample_n: fewer than 10 numbers give no sample
    assert lib.qa_pco_sample_f32(x.ctypes.data_as(f32p), 9, out, 200) == 0


def test_xoroshiro_first_outputs(lib):
    # Xoroshiro128++ (Blackman-Vigna) seeded through SplitMix64, restated
    # independently here
    def splitmix(x):
        x = (x + 0x9E3779B97F4A7C15) & (2**64 - 1)
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
        return x, z ^ (z >> 31)

    def rotl(v, k):
        return ((v << k) | (v >> (64 - k))) & (2**64 - 1)

    x, s0 = splitmix(0)
    x, s1 = splitmix(x)
    want = []
    for _ in range(8):
        want.append((rotl((s0 + s1) & (2**64 - 1), 17) + s0) & (2**64 - 1))
        s1 ^= s0
        s0 = rotl(s0, 49) ^ s1 ^ ((s1 << 21) & (2**64 - 1))
        s1 = rotl(s1, 28)
    out = (C.c_uint64 * 8)()
    lib.qa_pco_xoroshiro(C.c_uint64(0), out, 8)
    assert list(out) == want


def test_ans_quantize_and_spread(lib):
    # ans/encoding.rs test_quantize_weights, ans/spec.rs ans_spec_new
    c = (C.c_uint32 * 2)(77, 100)
    assert lib.qa_pco_quantize_weights(c, 2, C.c_uint64(177), 4) == 4 and list(c) == [7, 9]
    c = (C.c_uint32 * 2)(77, 77)
    assert lib.qa_pco_quantize_weights(c, 2, C.c_uint64(154), 4) == 1 and list(c) == [1, 1]
    w = (C.c_uint32 * 4)(1, 1, 3, 11)
    o = (C.c_uint32 * 16)()
    assert lib.qa_pco_spread(4, w, 4, o) == 0
    assert list(o) == [0, 3, 2, 3, 2, 3, 3, 3, 3, 1, 3, 2, 3, 3, 3, 3]
    for weights, want in (([1], [0]), ([2], [0, 0])):
        w = (C.c_uint32 * len(weights))(*weights)
        o = (C.c_uint32 * len(want))()
        assert lib.qa_pco_spread(len(want).bit_length() - 1, w, len(weights), o) == 0
        assert list(o) == want
    # weights that do not fill the table are refused
    w = (C.c_uint32 * 2)(1, 1)
    assert lib.qa_pco_spread(2, w, 2, (C.c_uint32 * 4)()) == -1


def test_float_mult_pair_gcd(lib):
    # float_mult.rs test_approx_pair_gcd
    g = C.c_double()
    for a, b, want in ((0.0, 0.0, None), (1.0, 0.0, None), (1.0, 1.0, None), (1.0, 2.0, 1.0),
                       (6.0, 3.0, 3.0), (10.01, 0.009999999999999787, 0.009999999999999787)):
        r = lib.qa_pco_pair_gcd_f64(a, b, C.byref(g))
        assert (g.value if r else None) == want, (a, b)
    gf = C.c_float()
    fmax = np.float32(np.finfo(np.float32).max) * np.float32(0.5)
    assert lib.qa_pco_pair_gcd_f32(2.0 ** 100, 3.0, C.byref(gf)) == 0
    assert lib.qa_pco_pair_gcd_f32(float(fmax), float(fmax * np.float32(0.6)), C.byref(gf)) == 1
    assert np.float32(gf.value) == fmax * np.float32(0.2)
    assert lib.qa_pco_pair_gcd_f32(float(fmax), 1e-13, C.byref(gf)) == 0
    third = float(np.float32(1) / np.float32(3))
    assert lib.qa_pco_pair_gcd_f32(third, 0.25, C.byref(gf)) == 1
    twelfth = np.float32(1) / np.float32(12)
    assert abs(f32_ordered(gf.value) - f32_ordered(twelfth)) <= 1


def test_float_mult_snap_and_config(lib):
    # float_mult.rs test_snap (f64) and test_choose_config_and_bid (f32)
    ob, oi = C.c_double(), C.c_double()
    for base, want in ((0.01000333, (0.01, 100.0)), (0.009999666, (0.01, 100.0)),
                       (0.143, (1.0 / 7.0, 7.0)), (0.0105, (0.0105, 1.0 / 0.0105))):
        lib.qa_pco_snap_f64(base, C.byref(ob), C.byref(oi))
        assert (ob.value, oi.value) == want, base
    fb, fi = C.c_float(), C.c_float()
    sev = np.array([((i % 50) - 20) * (np.float32(1.0) / np.float32(7.0)) for i in range(1000)],
                   np.float32)
    ones = np.ones(1000, np.float32)
    noisy = np.array([f32_from_ordered(f32_ordered(np.float32(0.1) * np.float32(i - 100)) +
                                       (-7 + i % 15)) for i in range(1000)], np.float32)
    for arr, want in ((sev, (np.float32(1) / np.float32(7), 7.0)), (ones, (1.0, 1.0)),
                      (noisy, (np.float32(1) / np.float32(10), 10.0))):
        assert lib.qa_pco_float_mult_config_f32(arr.ctypes.data_as(f32p), len(arr), C.byref(fb),
                                                C.byref(fi)) == 1
        assert (np.float32(fb.value), np.float32(fi.value)) == (np.float32(want[0]), np.float32(want[1]))
    bits = C.c_double()
    assert lib.qa_pco_float_mult_bid_f32(sev[:50].ctypes.data_as(f32p), 50, C.byref(fb), C.byref(bits)) == 1
    assert lib.qa_pco_float_mult_bid_f32(ones.ctypes.data_as(f32p), 1000, C.byref(fb), C.byref(bits)) == 0
    junk = np.sin(np.arange(1000, dtype=np.float32)).astype(np.float32)
    assert lib.qa_pco_float_mult_bid_f32(junk.ctypes.data_as(f32p), 1000, C.byref(fb), C.byref(bits)) == 0
    # test_trailing_zeros and test_candidate_euclidean
    tz = np.array([0.0, 3.0, 6.0, 21.0, 2.0 ** 100] * 5, np.float32)
    assert lib.qa_pco_float_mult_config_f32(tz.ctypes.data_as(f32p), len(tz), C.byref(fb), C.byref(fi)) == 1
    assert fb.value == 3.0
    e = np.array([0.0, 2.0 ** -100, 0.0037, 1.0001] * 5 + [np.finfo(np.float32).max], np.float32)
    assert lib.qa_pco_float_mult_config_f32(e.ctypes.data_as(f32p), len(e), C.byref(fb), C.byref(fi)) == 1
    assert abs(fb.value - 1e-4) < 1e-6


def test_int_mult_candidate_base(lib):
    # int_mult.rs test_calc_candidate_gcd
    u = C.c_uint32()
    for smp, want in (([0, 4, 8], None), ([0, 4, 8, 10, 14, 18, 20, 24, 28], 4),
                      ([1, 11, 21, 31, 41, 51, 61, 71, 82], 10),
                      ([1, 11, 22, 31, 41, 51, 61, 71, 82], None)):
        a = np.array(smp, np.uint32)
        r = lib.qa_pco_int_candidate_base(a.ctypes.data_as(u32p), len(a), C.byref(u))
        assert (u.value if r else None) == want, smp


def test_float_quant_best_k(lib):
    # float_quant.rs test_estimate_best_k, test_estimate_best_k_full_precision
    s = np.array([1.0, 1.25, -1.5, 1.75, -0.875, 0.75, 0.625] * 3 + [0.0], np.float32)
    s[-1] = np.array([np.float32(1.0).view(np.uint32) + 1], np.uint32).view(np.float32)[0]
    bits = C.c_double()
    assert lib.qa_pco_quant_k_f32(s.ctypes.data_as(f32p), len(s), C.byref(bits)) == 21
    assert 10.0 < bits.value < 21.0
    d = np.ones(20)
    assert lib.qa_pco_quant_k_f64(d.ctypes.data_as(C.POINTER(C.c_double)), 20, C.byref(bits)) == 52
    assert bits.value == 52.0


def test_mode_choice(lib):
    # float.rs test_choose_mult_mode / test_choose_quant_mode
    f = (np.arange(1000) * 1.5).astype(np.float64).view(np.uint64)
    base = int(np.array([1.5]).view(np.uint64)[0]) ^ (1 << 63)
    assert describe(lib, F64, 8, f)[:2] == [2, base]
    one = np.array([1.0]).view(np.uint64)[0]
    q = (one + (np.arange(1000, dtype=np.uint64) << np.uint64(20))).astype(np.uint64)
    assert describe(lib, F64, 8, q)[:2] == [3, 20]


def test_bin_optimization_and_log2(lib):
    # bin_optimization.rs test_bin_optimization, _enveloped, test_log2_approx
    o = (C.c_uint64 * 50)()
    h = np.array([[100, 1, 16], [100, 33, 48], [100, 49, 64], [100, 65, 74], [50, 75, 79]],
                 np.uint64).ravel()
    nb = lib.qa_pco_optimize_bins_u32(h.ctypes.data_as(u64p), 5, 10, o, 50)
    assert [list(o[5 * i:5 * i + 5]) for i in range(nb)] == [
        [100, 1, 16, 4, 0], [200, 33, 64, 5, 1], [150, 65, 79, 4, 2]]
    h = np.array([[1000, 0, 150], [1000, 200, 200]], np.uint64).ravel()
    nb = lib.qa_pco_optimize_bins_u32(h.ctypes.data_as(u64p), 2, 10, o, 50)
    assert [list(o[5 * i:5 * i + 5]) for i in range(nb)] == [[1000, 0, 150, 8, 0], [1000, 200, 200, 0, 1]]
    for e in range(32):
        assert lib.qa_pco_log2_approx(float(1 << e)) == float(e)
    prev = -np.inf
    for i in range(1, 101):
        a = lib.qa_pco_log2_approx(float(i))
        assert a >= prev and abs(a - np.log2(i)) < 0.0076
        prev = a


def test_histogram_quicksort(lib):
    # histograms.rs test_histogram_quicksort (any shuffle gives these bins)
    rng = np.random.default_rng(1)

    def hist(lat, lg):
        lat = np.ascontiguousarray(lat, dtype=np.uint32)
        o = (C.c_uint64 * 300)()
        nb = lib.qa_pco_histogram_u32(lat.ctypes.data_as(u32p), len(lat), lg, o, 300)
        return [tuple(o[3 * i:3 * i + 3]) for i in range(nb)]

    assert hist([], 2) == [] and hist([8], 0) == [(1, 8, 8)]
    for _ in range(16):
        assert hist(rng.permutation(100), 2) == [(25, 0, 24), (25, 25, 49), (25, 50, 74), (25, 75, 99)]
        a = np.zeros(100, np.uint32); a[0] = 1; rng.shuffle(a)
        assert hist(a, 2) == [(99, 0, 0), (1, 1, 1)]
        a = np.ones(100, np.uint32); a[0] = 0; rng.shuffle(a)
        assert hist(a, 2) == [(1, 0, 0), (99, 1, 1)]
        a = np.full(100, 5, np.uint32); a[0] = 3; a[1:3] = 7; rng.shuffle(a)
        assert hist(a.copy(), 2) == [(1, 3, 3), (97, 5, 5), (2, 7, 7)]
        assert hist(a.copy(), 1) == [(98, 3, 5), (2, 7, 7)]
        a = np.full(100, 5, np.uint32); a[0:2] = 3; a[2] = 7; rng.shuffle(a)
        assert hist(a, 1) == [(2, 3, 3), (98, 5, 7)]


def test_lookbacks(lib):
    # delta/lookback.rs test_lookback_encode_decode
    orig = np.full(100, 100, np.uint32)
    for i, v in ((1, 200), (2, 201), (3, 202), (5, 203), (15, 204), (50, 205)):
        orig[i] = v
    lb = (C.c_uint32 * 100)()
    n = lib.qa_pco_lookbacks_u32(4, 1, orig.ctypes.data_as(u32p), 100, lb)
    assert n == 98 and (lb[0], lb[2], lb[13], lb[48]) == (1, 4, 10, 1)


def shapes(rng, n):
    walk = np.round(100 * (1000 + np.cumsum(rng.standard_normal(n)))) / 100
    yield "walk", F64, walk.view(np.uint64)
    yield "unif", F64, rng.random(n).view(np.uint64)
    ts = (1_700_000_000_000_000_000 + np.cumsum(rng.integers(1, 1001, n))).astype(np.int64)
    yield "ts", I64, ts.view(np.uint64)
    yield "u32", U32, rng.integers(0, 2 ** 11, n).astype(np.uint64)
    yield "i32", I32, rng.integers(-5000, 5000, n).astype(np.int64).view(np.uint64)
    f32 = (rng.random(n).astype(np.float32) * 2 - 1).astype(np.float64)
    yield "f32", F32, f32.view(np.uint64)
    yield "u64mult", U64, (rng.integers(0, 10 ** 6, n) * 7919).astype(np.uint64)
    yield "const", F64, np.full(n, np.array([3.25]).view(np.uint64)[0], np.uint64)
    spec = np.array([np.nan, 0.0, -0.0, np.inf, -np.inf, 5e-324] * (n // 6 + 1))[:n]
    yield "special", F64, spec.view(np.uint64)
    quant = np.array([np.float32(x) for x in rng.random(n)]).astype(np.float64)
    yield "quant", F64, quant.view(np.uint64)


@pytest.mark.parametrize("n", [1, 2, 9, 10, 255, 256, 257, 1000, 32768])
def test_roundtrip_and_guarantee(lib, n):
    rng = np.random.default_rng(n)
    for name, nt, lane in shapes(rng, n):
        for level in (8, 12):
            b = compress(lib, nt, level, lane)
            assert b[:4] == b"pco!" and b[-1] == 0
            rc, back = decompress(lib, nt, b, n)
            assert rc == 0 and np.array_equal(back, lane), (name, level)
            # wrapped/guarantee.rs chunk_size plus the standalone framing
            bits = 32 if nt in (U32, I32, F32) else 64
            meta = (4 + (4 + 5 + 5 + 64 + 32 * 32) + 19 + bits + (6 if bits == 32 else 7) + 7) // 8
            framing = 4 + 1 + 1 + 5 + 2 + 4 + 1
            assert len(b) <= framing + meta + (n * bits + 7) // 8, name


def test_stream_damage_refused(lib):
    # stability.rs: every strict prefix is refused; so are another number
    # type and another count
    rng = np.random.default_rng(7)
    walk = (np.round(100 * (1000 + np.cumsum(rng.standard_normal(3000)))) / 100).view(np.uint64)
    b = compress(lib, F64, 12, walk)
    for i in range(len(b)):
        assert decompress(lib, F64, b[:i], 3000)[0] != 0, i
    assert decompress(lib, I64, b, 3000)[0] != 0
    assert decompress(lib, F64, b, 2999)[0] != 0
    assert decompress(lib, F64, b"", 0)[0] != 0
    bad = bytearray(b)
    bad[0] ^= 1
    assert decompress(lib, F64, bytes(bad), 3000)[0] != 0


def test_empty_block_declines(lib):
    # pcodec.rs empty_block_bails: no candidate for n = 0
    ol = C.c_size_t()
    assert lib.qa_pco_compress(F64, 8, None, 0, None, 0, C.byref(ol)) == 1
