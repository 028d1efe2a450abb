"""The engine's host competition at Levels Balanced / High / Max, pinned on the
CPU against the oracle's block encoder (src/encoder.rs:275-469).

On the GPU path the device supplies each block's candidate sizes for the
modes it emits itself (RAW, CONST, STRIDE, XORZ, ForBitpack, DeltaBitpack,
ALP, RLE), the FloatMult scale, the ALP-RD cut and dictionary, the block
features and (High / Max) the predictor and linear residual streams; host
cores run the rest. Here those inputs come from the oracle (each pinned by
its own tests against the device at Level::Fast), and the winner the engine's
host code picks (its mode, and its payload when host-written) must be the
oracle's frame for the same block, byte for byte.
"""
import ctypes as C
import os

import numpy as np
import pytest

import _oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "quoin_amd", "libquoin_amd.so")
U64MAX = (1 << 64) - 1
# device candidate slots, in the reference's consider order
DEVICE_SLOTS = [("RAW", 0), ("CONST", 1), ("STRIDE", 2), ("XORZ", 3), ("FOR_BITPACK", 4),
                ("DELTA_BITPACK", 5), ("ALP", 7), ("RLE", 10)]
HOST_MODES = {0, 4, 7, 11, 17, 19, 20, 25, 29, 45, 52, 54}
TEXT = b"the quick brown fox jumps over the lazy dog and then some more words here, again"


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("libquoin_amd.so not built")
    lib = C.CDLL(LIB)
    lib.qa_hybrid_compete.argtypes = [
        C.c_int, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32,
        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int32),
        C.POINTER(C.c_int32), C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]
    lib.qa_hybrid_encode_mode.argtypes = [
        C.c_int, C.c_uint64, C.c_int, C.c_uint, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32,
        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
        C.POINTER(C.c_size_t)]
    return lib


def device_inputs(lane, dtype):
    """What the device hands the host stage for one block: the FloatMult
    scale, the ALP-RD cut and dictionary, the five residual streams."""
    fm = orc.encode_mode_lv("FLOAT_MULT", lane, dtype, orc.FAST)
    fm_scale = fm[1] if fm is not None else 7
    rd = orc.encode_mode_lv("ALP_RD", lane, dtype, orc.FAST)
    rd_dict = np.zeros(8, dtype=np.uint64)
    if rd is not None:
        rdv = np.array([rd[0], rd[1], 0], dtype=np.uint32)
        rd_dict[:rd[1]] = np.frombuffer(rd[2:2 + 8 * rd[1]], dtype="<u8")
    else:
        rdv = np.array([0, 0, 1], dtype=np.uint32)
    res = [orc.pred_encode(0, lane, 16), orc.pred_encode(25, lane, 16),
           orc.linear_encode(20, lane), orc.linear_encode(45, lane), orc.linear_encode(7, lane, 0)]
    bufs = [np.frombuffer(r + b"\0", dtype=np.uint8) if r is not None else None for r in res]
    ptrs = (C.c_void_p * 5)(*[b.ctypes.data if b is not None else None for b in bufs])
    rlen = np.array([len(r) if r is not None else 0 for r in res], dtype=np.uint64)
    return fm_scale, rdv, rd_dict, bufs, ptrs, rlen


def encode_mode(lib, mode, values, dtype, level):
    lane = orc.lane_words(values)
    fm_scale, rdv, rd_dict, bufs, ptrs, rlen = device_inputs(lane, dtype)
    lam = {orc.BALANCED: 2, orc.HIGH: 0, orc.MAX: 0}[level]
    cap = 16 * len(lane) + 4096
    out = np.zeros(cap, dtype=np.uint8)
    ln = C.c_size_t()
    rc = lib.qa_hybrid_encode_mode(orc.MODE[mode], lam, 1 if level == orc.MAX else 0,
                                   12 if level == orc.MAX else 8, dtype, lane.ctypes.data,
                                   len(lane), fm_scale, rdv.ctypes.data, rd_dict.ctypes.data, ptrs,
                                   rlen.ctypes.data, out.ctypes.data, cap, C.byref(ln))
    assert rc in (0, 1)
    return bytes(out[:ln.value]) if rc == 1 else None


def compete(lib, values, dtype, level):
    lane = orc.lane_words(values)
    n = len(lane)
    sz = np.full(11, U64MAX, dtype=np.uint64)
    for name, slot in DEVICE_SLOTS:
        p = orc.encode_mode_lv(name, lane, dtype, orc.FAST)
        if p is not None:
            sz[slot] = len(p)
    _, dlow, rep, inc = orc.block_feats(lane, dtype)
    feats = np.array([inc, dlow, rep], dtype=np.int32)
    fm_scale, rdv, rd_dict, bufs, ptrs, rlen = device_inputs(lane, dtype)
    mode, host, ln = C.c_int32(), C.c_int32(), C.c_size_t()
    cap = 16 * n + 4096
    out = np.zeros(cap, dtype=np.uint8)
    rc = lib.qa_hybrid_compete(level, dtype, lane.ctypes.data, n, sz.ctypes.data, feats.ctypes.data,
                               fm_scale, rdv.ctypes.data, rd_dict.ctypes.data, ptrs,
                               rlen.ctypes.data, C.byref(mode), C.byref(host), out.ctypes.data, cap,
                               C.byref(ln))
    assert rc == 0
    payload = bytes(out[:ln.value]) if host.value else None
    return mode.value, host.value, ln.value, payload


def datasets(n, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(n)
    walk = np.round(100 + np.cumsum(rng.normal(0, 0.05, n)), 2)
    return {
        "price_walk_f64": (walk, orc.F64),
        "noise_f64": (rng.normal(0, 1, n), orc.F64),
        "sine_f64": (np.sin(t / 50.0) * 20 + 3, orc.F64),
        "decimals3_f64": (np.round(rng.uniform(0, 50, n), 3), orc.F64),
        "repeats_f64": (np.repeat(rng.normal(0, 1, n // 16 + 1), 16)[:n], orc.F64),
        "timestamps_i64": ((1_700_000_000_000 + t * 1000 + rng.integers(0, 7, n)).astype(np.int64),
                           orc.I64),
        "lowcard_i64": (rng.choice(np.array([3, 17, 99, 1024, -5], dtype=np.int64), n), orc.I64),
        "runs_u32": (np.repeat(rng.integers(0, 1 << 20, n // 40 + 1), 40)[:n].astype(np.uint32),
                     orc.U32),
        "sensor_f32": ((20 + np.cumsum(rng.normal(0, 0.01, n))).astype(np.float32), orc.F32),
        "random_i32": (rng.integers(-(1 << 31), 1 << 31, n).astype(np.int32), orc.I32),
        "ids_u64": ((rng.integers(0, 1 << 40, n) * 8).astype(np.uint64), orc.U64),
        "text_like_u64": (rng.choice(np.frombuffer(TEXT, dtype="<u8"), n), orc.U64),
    }


@pytest.mark.parametrize("level", [orc.BALANCED, orc.HIGH, orc.MAX])
@pytest.mark.parametrize("n", [700, 4096])
def test_host_competition_matches_the_oracle(lib, level, n):
    for name, (vals, dt) in datasets(n, 7 + n).items():
        stream = orc.compress(vals, dt, level=level, block_size=n)
        fr = orc.frames(stream)
        assert len(fr) == 1, name
        want_mode, _, want_payload = fr[0]
        mode, host, ln, payload = compete(lib, vals, dt, level)
        assert mode == want_mode, (name, level, n, mode, want_mode)
        assert ln == len(want_payload), (name, level, n)
        assert host == (mode in HOST_MODES), (name, mode)
        if host:
            assert payload == want_payload, (name, level, n, mode)


@pytest.mark.parametrize("level", [orc.BALANCED, orc.MAX])
def test_host_competition_one_full_block(lib, level):
    n = 32768
    for name in ("price_walk_f64", "timestamps_i64", "sensor_f32", "lowcard_i64"):
        vals, dt = datasets(n, 3)[name]
        want_mode, _, want_payload = orc.frames(orc.compress(vals, dt, level=level, block_size=n))[0]
        mode, host, ln, payload = compete(lib, vals, dt, level)
        assert (mode, ln) == (want_mode, len(want_payload)), name
        if host:
            assert payload == want_payload, name


def test_incompressible_integers_stay_raw(lib):
    rng = np.random.default_rng(5)
    vals = rng.integers(0, 1 << 64, 2048, dtype=np.uint64)
    for level in (orc.BALANCED, orc.HIGH, orc.MAX):
        want = orc.frames(orc.compress(vals, orc.U64, level=level, block_size=2048))[0]
        assert (want[0], len(want[2])) == (5, 8 * 2048)
        mode, host, ln, _ = compete(lib, vals, orc.U64, level)
        assert (mode, host, ln) == (5, 0, 8 * 2048)


HOST_CODED = ["PRED", "PRED_RC", "PRED2", "DELTA2", "DELTA_DP", "ORDERED_DELTA", "FLOAT_MULT",
              "LZ", "BYTE_TRANSPOSE", "ALP_RD", "DICT", "PCO"]


@pytest.mark.parametrize("level", [orc.BALANCED, orc.HIGH, orc.MAX])
def test_every_host_coded_mode_matches_the_oracle(lib, level):
    """Each host-coded candidate's payload, winner or not, with the level's
    lambda and (Max) the LZ cascade inside code_residuals."""
    seen = 0
    for name, (vals, dt) in datasets(2048, 11).items():
        lane = orc.lane_words(vals)
        for mode in HOST_CODED:
            want = orc.encode_mode_lv(mode, lane, dt, level)
            got = encode_mode(lib, mode, vals, dt, level)
            if mode in ("DELTA2", "DELTA_DP", "FLOAT_MULT", "ALP_RD") and dt not in (orc.F64, orc.F32):
                continue  # float-only modes: the competition never asks
            assert got == want, (name, mode, level)
            seen += want is not None
    assert seen > 80
