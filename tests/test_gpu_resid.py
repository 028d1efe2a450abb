"""GPU parity of the linear residual coders (src/codecs/linear.rs) against the
oracle: OrderedDelta, Delta2 and DeltaDp streams byte for byte, the orders
select_order picks, DeltaDp's bail-outs, decoded values, and the outcome of
corrupt and truncated streams, through quoin_gpu_linear_encode /
quoin_gpu_linear_decode."""
import math

import numpy as np
import pytest

import _oracle as orc
from _data import float_columns, int_columns

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a GPU", allow_module_level=True)

import quoin_amd as qa  # noqa: E402

DEV = torch.device("cuda:0")
KINDS = (qa.LINEAR_ORDERED_DELTA, qa.LINEAR_DELTA2, qa.LINEAR_DELTA_DP)
GUARD = 0xA5


@pytest.fixture(scope="module")
def gpu():
    g = qa.Gpu(0)
    yield g
    g.close()


def _lane(x):
    x = np.asarray(x)
    return np.ascontiguousarray(x.astype(np.float64).view(np.uint64) if x.dtype.kind == "f"
                                else x.astype(np.uint64))


def gpu_encode(gpu, kind, lanes, order=0, dst_off=0):
    """Returns (streams or None per block, orders); checks the bytes around each
    stream are untouched."""
    keep, ins = [], []
    for lane in lanes:
        v = torch.from_numpy(lane.view(np.int64).copy()).to(DEV) if len(lane) else \
            torch.zeros(1, dtype=torch.int64, device=DEV)
        cap = int(qa.lib.quoin_linear_bound(len(lane)))
        o = torch.full((cap + dst_off + 32,), GUARD, dtype=torch.uint8, device=DEV)
        keep.append((v, o))
        ins.append(qa.LinearInput(v.data_ptr(), len(lane), o.data_ptr() + dst_off, cap))
    lens, ords, sts = gpu.linear_encode(kind, ins, order)
    torch.cuda.synchronize()
    out = []
    for (v, o), ln, st in zip(keep, lens, sts):
        h = o.cpu().numpy()
        assert (h[:dst_off] == GUARD).all()
        if st:
            assert kind == qa.LINEAR_DELTA_DP and ln == 0
            out.append(None)
            continue
        assert (h[dst_off + ln:] == GUARD).all(), "write past the stream"
        out.append(h[dst_off:dst_off + ln].tobytes())
    return out, ords


def gpu_decode(gpu, kind, streams, ns):
    keep, ss = [], []
    for s, n in zip(streams, ns):
        t = torch.from_numpy(np.frombuffer(bytes(s) + b"\0" * 16, dtype=np.uint8).copy()).to(DEV)
        o = torch.full((n + 2,), -0x5A5A5A5A5A5A5A5B, dtype=torch.int64, device=DEV)
        keep.append((t, o))
        ss.append(qa.LinearStream(t.data_ptr(), len(s), n, o.data_ptr() + 8))
    sts = gpu.linear_decode(kind, ss)
    torch.cuda.synchronize()
    vals = []
    for (t, o), n in zip(keep, ns):
        h = o.cpu().numpy().view(np.uint64)
        assert h[0] == h[-1] == np.uint64(0xA5A5A5A5A5A5A5A5), "write outside out"
        vals.append(h[1:1 + n])
    return vals, sts


def _check(gpu, kind, lanes, order=0, dst_off=0):
    got, ords = gpu_encode(gpu, kind, lanes, order, dst_off)
    for lane, g, o in zip(lanes, got, ords):
        if kind == qa.LINEAR_ORDERED_DELTA:
            want = orc.linear_encode(kind, lane)
        else:
            want_order = order or orc.linear_select_order(lane)
            assert o == want_order, (o, want_order)
            want = orc.linear_encode(kind, lane, want_order)
        assert g == want, f"kind {kind} n={len(lane)} order {o}: stream differs"
    streams = [g for g in got if g is not None]
    keep_lanes = [l for l, g in zip(lanes, got) if g is not None]
    if streams:
        vals, sts = gpu_decode(gpu, kind, streams, [len(l) for l in keep_lanes])
        for lane, v, st in zip(keep_lanes, vals, sts):
            assert st == qa.OK
            assert v.tolist() == lane.tolist(), f"kind {kind} n={len(lane)}: decode differs"
    return got


def _sine(n, seed=3):
    r = np.random.default_rng(seed)
    i = np.arange(n, dtype=np.float64)
    return _lane(np.sin(2 * np.pi * i / 1000) * 1000 + r.normal(0, 1e-3, n))


def _specials(n, seed=5):
    r = np.random.default_rng(seed)
    pool = np.array([0, 1 << 63, 0x7FF0000000000000, 0xFFF0000000000000, 0x7FF8000000000000,
                     0xFFF8000000000000, 0x7FF0000000000001, 0x7FF4000000000123, 1, 0x000FFFFFFFFFFFFF,
                     0x3FF0000000000000, 0xBFF0000000000000, 0x7FEFFFFFFFFFFFFF], dtype=np.uint64)
    lane = pool[r.integers(0, len(pool), n)]
    walk = _lane(np.cumsum(r.normal(size=n)))
    mix = r.random(n) < 0.5
    lane[mix] = walk[mix]
    return lane


def _lanes():
    f = float_columns(40000, seed=11)
    i = int_columns(40000, seed=12)
    return {
        "sine": _sine(70000),
        "walk": _lane(f["walk_f64"]),
        "dec3": _lane(f["dec3_f64"]),
        "unif": _lane(f["unif_f64"]),
        "ramp": _lane(np.arange(9000) * 0.5),
        "parabola": _lane(0.5 * np.arange(9000.0) ** 2),
        "ints": _lane(np.random.default_rng(1).integers(-2**30, 2**30, 5000).astype(np.float64)),
        "const": _lane(np.full(3000, 7.25)),
        "zeros": np.zeros(2500, dtype=np.uint64),
        "specials": _specials(6000),
        "bits": np.random.default_rng(2).integers(0, 2**63, 4100, dtype=np.uint64) * np.uint64(2),
        "ts_i64": _lane(i["ts_i64"].view(np.uint64)),
        "wide_walk": _lane(i["i64_wide_walk"].view(np.uint64)),
    }


@pytest.mark.parametrize("kind", KINDS)
def test_linear_columns(gpu, kind):
    for name, lane in _lanes().items():
        print(f"[resid] kind {kind} column {name} n={len(lane)}", flush=True)
        _check(gpu, kind, [lane])


@pytest.mark.parametrize("kind", KINDS)
def test_linear_block_sizes_batch(gpu, kind):
    # one launch over blocks of many sizes, around the 2048-value tile and the
    # select_order thresholds, with empty blocks in between
    base = _sine(20000, seed=8)
    sizes = [0, 1, 2, 3, 4, 5, 7, 8, 9, 1023, 1024, 1025, 2047, 2048, 2049, 0, 4096, 6001, 1, 11111]
    lanes, off = [], 0
    for s in sizes:
        lanes.append(base[off:off + s].copy())
        off = (off + 997) % 8000
    print(f"[resid] kind {kind} batch sizes {sizes}", flush=True)
    _check(gpu, kind, lanes)


@pytest.mark.parametrize("kind", KINDS[1:])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_linear_fixed_orders(gpu, kind, order):
    lanes = [_sine(5000, seed=order), _lane(np.arange(3000) * 0.5), _specials(3000, seed=order)]
    print(f"[resid] kind {kind} order {order}", flush=True)
    _check(gpu, kind, lanes, order=order)


@pytest.mark.parametrize("kind", KINDS)
def test_linear_unaligned_destinations(gpu, kind):
    lane = _sine(9000, seed=4)
    for off in (1, 3, 7, 13, 15):
        print(f"[resid] kind {kind} dst offset {off}", flush=True)
        _check(gpu, kind, [lane, lane[:5000]], dst_off=off)


def test_linear_dp_bails_like_the_oracle(gpu):
    k = qa.LINEAR_DELTA_DP
    lanes = [np.zeros(0, dtype=np.uint64),
             _lane([0.0, 1e16, 1.0, 3.0]),
             _lane(np.random.default_rng(3).random(7000) * 1e300),
             _lane(np.arange(7000) * 0.5)]
    lanes[2][6000] = _lane([1e-300])[0]  # a late value that does not add back
    got = _check(gpu, k, lanes)
    assert got[0] is None and got[1] is None and got[3] is not None


def test_linear_select_order_reference_shapes(gpu):
    smooth = _lane([math.sin(i * 0.002) * 100.0 + i * 0.01 for i in range(2000)])
    s, noise = 1, []
    for _ in range(2000):
        s = (s * 6364136223846793005 + 1) & ((1 << 64) - 1)
        noise.append(s >> 2)
    noise = np.array(noise, dtype=np.uint64)
    _, ords = gpu_encode(gpu, qa.LINEAR_DELTA2, [smooth, noise])
    assert ords[0] >= 3 and ords[1] <= 2  # linear.rs:276-291


@pytest.mark.parametrize("kind", KINDS)
def test_linear_decode_errors(gpu, kind):
    # halves of small integers: every residual adds back exactly, so DeltaDp
    # codes the block too
    lane = _lane(np.random.default_rng(9).integers(-1000, 1000, 6000) * 0.5)
    good = orc.linear_encode(kind, lane, 3) if kind != qa.LINEAR_ORDERED_DELTA else \
        orc.linear_encode(kind, lane)
    assert good is not None
    n = len(lane)
    r = np.random.default_rng(kind)
    cases = [(good, n), (good + b"\x00", n), (good[:-1], n), (good, n + 1), (good, n - 1),
             (good[:len(good) // 2], n), (b"", 0), (b"", 1), (bytes([3]), 0), (bytes([0, 0]), 1),
             (bytes([5, 0]), 1), (bytes([1]) + b"\x80" * 10 + b"\x00", 1),
             (bytes([1]) + b"\xff" * 9 + b"\x01", 1), (bytes([2]) + b"\x82\x80\x00", 1)]
    for _ in range(60):
        b = bytearray(good)
        for _ in range(int(r.integers(1, 4))):
            b[int(r.integers(0, len(b)))] ^= 1 << int(r.integers(0, 8))
        cases.append((bytes(b), n))
    print(f"[resid] kind {kind} {len(cases)} corrupt cases", flush=True)
    vals, sts = gpu_decode(gpu, kind, [c for c, _ in cases], [m for _, m in cases])
    for (c, m), v, st in zip(cases, vals, sts):
        rc, want = orc.linear_decode(kind, c, m)
        assert (st == qa.OK) == (rc == 0), (len(c), m, st, rc)
        if rc == 0:
            assert v.tolist() == want.tolist()


def test_linear_large_blocks(gpu):
    # the entropy workload's shape: blocks of 32768 f64 values, sine + noise
    lane = _sine(1 << 20, seed=3)
    blocks = [lane[i:i + 32768] for i in range(0, len(lane), 32768)]
    for kind in KINDS:
        print(f"[resid] kind {kind} 32 blocks of 32768", flush=True)
        _check(gpu, kind, blocks)
