"""Levels Balanced / High / Max on the GPU path: the engine's streams equal the
oracle's byte for byte (src/encoder.rs:275-469 with the host coders), and the
engine decodes its own streams and the oracle's, including a stream of every
host-coded frame mode."""
import struct

import numpy as np
import pytest

import _data
import _gpu
import _oracle as orc

pytestmark = pytest.mark.gpu

LEVELS = [orc.BALANCED, orc.HIGH, orc.MAX]


@pytest.fixture(scope="module")
def qa():
    import quoin_amd
    return quoin_amd


def engine_dtype(arr):
    return {np.dtype(np.float64): orc.F64, np.dtype(np.int64): orc.I64,
            np.dtype(np.uint64): orc.U64, np.dtype(np.int32): orc.I32,
            np.dtype(np.uint32): orc.U32, np.dtype(np.float32): orc.F32}[arr.dtype]


def check(qa, name, vals, level, validity=None, block_size=0, plog2=16, selection=0):
    print(f"case {name} n={len(vals)} dtype={vals.dtype} nulls={validity is not None} "
          f"level={level} bs={block_size} plog2={plog2} selection={selection}", flush=True)
    cfg = qa.Config(level=qa.Level(level), block_size=block_size, predictor_log2=plog2,
                    selection=qa.Selection(selection))
    got = qa.compress_column(vals, validity, cfg)
    want = orc.compress(vals, engine_dtype(vals), validity, level, block_size,
                        selection=selection, predictor_log2=plog2, threads=16)
    if got != want:
        gf, wf = orc.frames(got) if validity is None else [], \
            orc.frames(want) if validity is None else []
        diff = [(i, g[0], w[0], len(g[2]), len(w[2])) for i, (g, w) in enumerate(zip(gf, wf))
                if g != w][:4]
        raise AssertionError(f"{name}: stream differs (gpu {len(got)} B, oracle {len(want)} B; "
                             f"first differing frames (i, gpu mode, oracle mode, lens): {diff})")
    out, vout = qa.decompress_column(got, qa.FROM_NP[vals.dtype], len(vals),
                                     with_validity=validity is not None)
    _, want_vals, _ = orc.decompress(want)
    assert out.view(np.uint8).tobytes() == want_vals.view(np.uint8).tobytes(), f"{name}: decode"
    return got


@pytest.mark.parametrize("level", LEVELS)
def test_int_columns(qa, level):
    for name, v in _data.int_columns(40000, seed=21).items():
        check(qa, name, v, level)


@pytest.mark.parametrize("level", LEVELS)
def test_float_columns(qa, level):
    for name, v in _data.float_columns(40000, seed=22).items():
        check(qa, name, v, level)


@pytest.mark.parametrize("n", [0, 1, 2, 63, 64, 65, 1000])
def test_small_columns_max(qa, n):
    for name, v in {**_data.int_columns(n, seed=n), **_data.float_columns(n, seed=n)}.items():
        check(qa, name, v, orc.MAX)


def test_grown_blocks_max(qa):
    # repeats / low distinct counts grow blocks past one 32768-value region
    n = 140000
    r = np.random.default_rng(4)
    cols = {
        "runs_i64": np.repeat(r.integers(0, 50, n // 8 + 1), 8)[:n].astype(np.int64),
        "lowcard_f64": r.choice(np.array([1.5, 2.25, -7.0, 1e9]), n),
        "walk_f64": _data.float_columns(n, seed=5)["walk_f64"],
    }
    for name, v in cols.items():
        check(qa, name, v, orc.MAX)


def test_nulls_max(qa):
    for density in (0.0, 0.3, 1.0):
        for name, v in list(_data.int_columns(20000, seed=9).items())[:5]:
            vv, bm = _data.with_nulls(v, density, seed=3)
            check(qa, f"{name}+nulls{density}", vv, orc.MAX, validity=bm)


def test_fixed_blocks_and_predictor_sizes(qa):
    v = _data.float_columns(30000, seed=8)["sine_f64"]
    for bs in (1, 1000, 4096):
        check(qa, f"sine bs{bs}", v, orc.HIGH, block_size=bs)
    for plog2 in (10, 12):
        check(qa, f"sine plog2 {plog2}", v, orc.MAX, plog2=plog2)


def test_batch_of_mixed_columns_max(qa):
    import quoin_amd
    gpu = quoin_amd.Gpu(0)
    cols = []
    for i, (name, v) in enumerate({**_data.int_columns(9000, seed=31),
                                   **_data.float_columns(9000, seed=32)}.items()):
        cols.append((v, _data.with_nulls(v, 0.1, seed=i)[1] if i % 3 == 0 else None))
    cols = [(_data.with_nulls(v, 0.1, seed=i)[0] if bm is not None else v, bm)
            for i, (v, bm) in enumerate(cols)]
    print(f"case batch of {len(cols)} columns at Max", flush=True)
    got = _gpu.device_compress(qa, gpu, cols, qa.Config(qa.Level.MAX))
    for i, ((v, bm), g) in enumerate(zip(cols, got)):
        want = orc.compress(v, engine_dtype(v), bm, orc.MAX, threads=16)
        assert g == want, f"column {i}: stream differs"
    outs = _gpu.device_decompress(qa, gpu, got, [qa.FROM_NP[v.dtype] for v, _ in cols],
                                  [len(v) for v, _ in cols], True)
    for i, ((v, bm), (o, vo)) in enumerate(zip(cols, outs)):
        _, wv, _ = orc.decompress(got[i])
        assert o.view(np.uint8).tobytes() == wv.view(np.uint8).tobytes(), f"column {i}: decode"


HOST_CODED = ["PRED", "PRED_RC", "PRED2", "DELTA2", "DELTA_DP", "ORDERED_DELTA", "FLOAT_MULT",
              "LZ", "BYTE_TRANSPOSE", "ALP_RD", "DICT", "PCO"]
FLOAT_ONLY = {"DELTA2", "DELTA_DP", "FLOAT_MULT", "ALP_RD"}


def one_frame_stream(dtype, n, mode, payload, plog2=16):
    return (b"FCR1" + bytes([2, 0, plog2, dtype]) + struct.pack("<Q", n) +
            bytes([orc.MODE[mode]]) + struct.pack("<II", n, len(payload)) + payload)


@pytest.mark.parametrize("level", LEVELS)
def test_every_host_coded_frame_decodes(qa, level):
    """A one-frame stream of each host-coded mode, as each level writes it
    (entropy forms, lambda, the Max LZ cascade), decodes to the oracle's
    values."""
    cases = {
        "walk_f64": (_data.float_columns(5000, seed=2)["walk_f64"], orc.F64),
        "dec3_f64": (_data.float_columns(5000, seed=2)["dec3_f64"], orc.F64),
        "unif_f32": (_data.float_columns(5000, seed=2)["unif_f32"], orc.F32),
        "ts_i64": (_data.int_columns(5000, seed=2)["ts_i64"], orc.I64),
        "u32_k11": (_data.int_columns(5000, seed=2)["u32_k11"], orc.U32),
        "i32_walk": (_data.int_columns(5000, seed=2)["i32_walk"], orc.I32),
    }
    seen = 0
    for name, (v, dt) in cases.items():
        lane = orc.lane_words(v)
        for mode in HOST_CODED:
            if mode in FLOAT_ONLY and dt not in (orc.F64, orc.F32):
                continue
            p = orc.encode_mode_lv(mode, lane, dt, level)
            if p is None:
                continue
            s = one_frame_stream(dt, len(v), mode, p)
            print(f"case frame {mode} {name} level={level} plen={len(p)}", flush=True)
            _, want, _ = orc.decompress(s)
            out, _ = qa.decompress_column(s, qa.FROM_NP[v.dtype], len(v))
            assert out.view(np.uint8).tobytes() == want.view(np.uint8).tobytes(), (name, mode)
            seen += 1
    assert seen >= 50


def test_reference_default_config_streams_decode(qa):
    """Streams the oracle writes at Config::default (Level::Max, adaptive
    blocks) over mixed columns, with nulls, decode through the engine."""
    for i, (name, v) in enumerate({**_data.int_columns(70000, seed=41),
                                   **_data.float_columns(70000, seed=42)}.items()):
        bm = _data.with_nulls(v, 0.05, seed=i)[1] if i % 2 else None
        vv = _data.with_nulls(v, 0.05, seed=i)[0] if i % 2 else v
        s = orc.compress(vv, engine_dtype(vv), bm, orc.MAX, threads=16)
        print(f"case default-config decode {name} nulls={bm is not None} {len(s)} B", flush=True)
        out, vout = qa.decompress_column(s, qa.FROM_NP[vv.dtype], len(vv),
                                         with_validity=bm is not None)
        _, want, _ = orc.decompress(s)
        assert out.view(np.uint8).tobytes() == want.view(np.uint8).tobytes(), name


@pytest.mark.parametrize("level", [orc.FASTEST, orc.FAST, orc.BALANCED, orc.HIGH, orc.MAX])
def test_sample_selection(qa, level):
    """Selection::Sample (src/encoder.rs:483-641): the winner on each block's
    sample of 8 runs of 64 values, then on the whole block, next to RAW,
    CONST, STRIDE, LZ and pco; at every level, with small and grown blocks."""
    cols = {**_data.int_columns(40000, seed=51), **_data.float_columns(40000, seed=52)}
    for name, v in cols.items():
        check(qa, f"sample {name}", v, level, selection=1)
    check(qa, "sample small", _data.float_columns(300, seed=3)["walk_f64"], level, selection=1)
    vv, bm = _data.with_nulls(_data.int_columns(20000, seed=6)["u32_k11"], 0.2, seed=4)
    check(qa, "sample nulls", vv, level, validity=bm, selection=1)
