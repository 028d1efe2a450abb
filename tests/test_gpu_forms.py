"""The plain form (one launch sequence per block, quoin_gpu_set_launch_form 1)
and the batched form (one launch per stage for the whole batch) produce the
same bytes and the same decode; both equal the oracle."""
import numpy as np
import pytest

import _data
import _gpu
import _oracle as orc

pytestmark = pytest.mark.gpu

ODT = {np.dtype(np.float64): orc.F64, np.dtype(np.int64): orc.I64,
       np.dtype(np.uint64): orc.U64, np.dtype(np.int32): orc.I32,
       np.dtype(np.uint32): orc.U32, np.dtype(np.float32): orc.F32}


@pytest.fixture(scope="module")
def qa():
    import quoin_amd
    return quoin_amd


def batch(n, seed):
    ic = _data.int_columns(n, seed)
    fc = _data.float_columns(n, seed + 1)
    cols = [(ic["ts_i64"], None), _data.with_nulls(ic["u32_k3"], 0.01, seed),
            _data.with_nulls(ic["u32_k11"], 0.3, seed + 2), (ic["i64_runs"], None),
            (fc["walk_f64"], None), (fc["unif_f64"], None), (fc["unif_f32"], None),
            (np.repeat(np.arange(37, dtype=np.float64) * 1.25, n // 37 + 1)[:n], None)]
    return cols


@pytest.mark.parametrize("level", [0, 1])
@pytest.mark.parametrize("n", [1000, 70000, 300000])
def test_plain_equals_batched(qa, level, n):
    cols = batch(n, seed=n + level)
    cfg = qa.Config(level=qa.Level(level))
    gpu = qa.Gpu(0)
    try:
        print(f"case forms n={n} level={level} batched", flush=True)
        gpu.set_launch_form(0)
        batched = _gpu.device_compress(qa, gpu, cols, cfg)
        print(f"case forms n={n} level={level} plain", flush=True)
        gpu.set_launch_form(1)
        plain = _gpu.device_compress(qa, gpu, cols, cfg)
        for i, ((v, bm), a, b) in enumerate(zip(cols, batched, plain)):
            assert a == b, f"column {i}: plain stream differs"
            if n <= 70000:
                assert a == orc.compress(v, ODT[v.dtype], bm, level), f"column {i}: oracle"
        dts = [qa.FROM_NP[v.dtype] for v, _ in cols]
        ns = [len(v) for v, _ in cols]
        for form in (0, 1):
            print(f"case forms n={n} level={level} decode form {form}", flush=True)
            gpu.set_launch_form(form)
            outs = _gpu.device_decompress(qa, gpu, plain, dts, ns, True)
            for i, ((v, bm), (o, vo)) in enumerate(zip(cols, outs)):
                _, want, _ = orc.decompress(batched[i])
                assert o.view(np.uint8).tobytes() == want.view(np.uint8).tobytes(), \
                    f"column {i}: decode form {form}"
                if bm is not None:
                    exp = bm.copy()
                    if n % 8:
                        exp[-1] &= (1 << (n % 8)) - 1
                    assert vo == exp.tobytes()
                else:
                    full = np.packbits(np.ones(n, np.uint8), bitorder="little")
                    assert vo == full.tobytes()
    finally:
        gpu.close()


def test_many_columns_alternating_kinds(qa):
    """A batch whose columns cycle through the tile decoders' kinds (FoR into
    4-byte integers, Delta / FloatMult / ALP into 8-byte lanes, RLE and RAW
    f32 through the general decoder), so each decoder takes many separate
    runs of tiles in one launch; streams and decodes equal the oracle's."""
    r = np.random.default_rng(21)
    cols = []
    for i in range(24):
        n = int(r.integers(1, 90000)) if i % 5 else int(r.integers(1, 3000))
        k = i % 6
        if k == 0:
            v = (r.integers(0, 2**32, n, dtype=np.uint64) >> np.uint64(21)).astype(np.uint32)
            cols.append(_data.with_nulls(v, 0.05, i))
        elif k == 1:
            cols.append((np.cumsum(r.integers(1, 1000, n)).astype(np.int64), None))
        elif k == 2:
            cols.append((r.integers(0, 10**9, n) / 1000.0, None))
        elif k == 3:
            cols.append((np.repeat(r.integers(0, 256, n) / 8.0, r.geometric(1 / 8, n))[:n], None))
        elif k == 4:
            cols.append((np.round(r.normal(100, 9, n), 2), None))
        else:
            cols.append((((r.integers(0, 2**24, n) * 2.0**-23) - 1).astype(np.float32), None))
    cfg = qa.Config(level=qa.Level.FAST)
    gpu = qa.Gpu(0)
    try:
        print(f"case alternating kinds, {len(cols)} columns", flush=True)
        streams = _gpu.device_compress(qa, gpu, cols, cfg)
        for i, ((v, bm), st) in enumerate(zip(cols, streams)):
            assert st == orc.compress(v, ODT[v.dtype], bm, orc.FAST), f"column {i}: oracle"
        outs = _gpu.device_decompress(qa, gpu, streams, [qa.FROM_NP[v.dtype] for v, _ in cols],
                                      [len(v) for v, _ in cols], False)
        for i, (st, (o, _)) in enumerate(zip(streams, outs)):
            _, want, _ = orc.decompress(st)
            assert o.view(np.uint8).tobytes() == want.view(np.uint8).tobytes(), f"column {i}"
    finally:
        gpu.close()


def test_launch_form_rejects_unknown(qa):
    gpu = qa.Gpu(0)
    try:
        for form in (3, -1):
            with pytest.raises(qa.QuoinError):
                gpu.set_launch_form(form)
    finally:
        gpu.close()


@pytest.mark.parametrize("level", [0, 1])
def test_graph_replays_equal_eager(qa, level):
    """Launch form 0 captures a batch shape met on two calls in a row into a
    hipGraph and replays it while the shape repeats. Five compress calls of
    one shape with new values each time, with calls of another shape, host
    reads and synchronizations between them, give the streams of a form-2
    handle (always kernel by kernel) and the oracle's; decodes replayed the
    same way return the oracle's values and bitmaps, and the tile stage's
    timing events are recorded on replays too."""
    cfg = qa.Config(level=qa.Level(level))
    n = 70000
    graph, eager = qa.Gpu(0), qa.Gpu(0)
    try:
        eager.set_launch_form(2)
        other = batch(3000, seed=7 + level)[:3]
        for rep in range(5):
            cols = batch(n, seed=300 + 10 * level + rep)  # one shape, new values
            print(f"case graph compress level={level} rep={rep}", flush=True)
            got = _gpu.device_compress(qa, graph, cols, cfg)
            assert got == _gpu.device_compress(qa, eager, cols, cfg), f"rep {rep}: eager"
            if rep >= 2:
                _, kms, _ = graph.last_timing()
                assert kms > 0, "tile-stage events of a replay"
            if rep in (2, 4):
                for i, ((v, bm), st) in enumerate(zip(cols, got)):
                    assert st == orc.compress(v, ODT[v.dtype], bm, level), f"rep {rep} column {i}"
            if rep == 3:  # another shape between two replays
                small = _gpu.device_compress(qa, graph, other, cfg)
                for i, ((v, bm), st) in enumerate(zip(other, small)):
                    assert st == orc.compress(v, ODT[v.dtype], bm, level), f"other column {i}"
        counts = graph.graph_counts()
        print("graph counts after compress:", counts, flush=True)
        assert counts["compress_captures"] >= 1 and counts["compress_launches"] >= 4, counts
        dts = [qa.FROM_NP[v.dtype] for v, _ in cols]
        ns = [len(v) for v, _ in cols]
        want = [orc.decompress(st) for st in got]
        for rep in range(4):
            print(f"case graph decompress level={level} rep={rep}", flush=True)
            outs = _gpu.device_decompress(qa, graph, got, dts, ns, True)
            for i, ((v, bm), (o, vo)) in enumerate(zip(cols, outs)):
                assert o.view(np.uint8).tobytes() == want[i][1].view(np.uint8).tobytes(), \
                    f"rep {rep} column {i}"
                assert o.tobytes() == v.tobytes(), f"rep {rep} column {i}: input"
                if bm is not None:
                    assert vo == bm.tobytes(), f"rep {rep} column {i}: bitmap"
            if rep == 1:  # another shape between two replays
                _gpu.device_decompress(qa, graph, small, [qa.FROM_NP[v.dtype] for v, _ in other],
                                       [len(v) for v, _ in other], False)
        counts = graph.graph_counts()
        print("graph counts after decompress:", counts, flush=True)
        assert counts["decompress_captures"] >= 1 and counts["decompress_launches"] >= 2, counts
    finally:
        graph.close()
        eager.close()


def test_autotuned_shapes_give_the_same_bytes(qa):
    """A batch large enough for the autotuner (over 4096 tiles of each tile
    stage) tunes k_dec_fl<4>, k_dec_fl<8> and k_emit_fl on its first call; the
    streams and decodes equal those at the fixed shape and the oracle's."""
    r = np.random.default_rng(61)
    n = 5 << 20
    ts = (1_700_000_000_000_000_000 + np.cumsum(r.integers(1, 1001, n))).astype(np.int64)
    u = (r.integers(0, 2**32, n, dtype=np.uint64) >> np.uint64(21)).astype(np.uint32)
    u, bm = _data.with_nulls(u, 0.01, 62)
    cols = [(ts, None), (u, bm)]
    cfg = qa.Config(level=qa.Level.FASTEST)
    tuned = qa.Gpu(0)
    fixed = qa.Gpu(0)
    try:
        fixed.set_autotune(False)
        print("case autotune: tuned handle", flush=True)
        a = _gpu.device_compress(qa, tuned, cols, cfg)
        print("case autotune: fixed handle", flush=True)
        b = _gpu.device_compress(qa, fixed, cols, cfg)
        assert a == b
        for (v, m), st in zip(cols, a):
            assert st == orc.compress(v, ODT[v.dtype], m, orc.FASTEST)
        dts = [qa.FROM_NP[v.dtype] for v, _ in cols]
        oa = _gpu.device_decompress(qa, tuned, a, dts, [n, n], True)
        ob = _gpu.device_decompress(qa, fixed, a, dts, [n, n], True)
        for (v, _), (x, vx), (y, vy) in zip(cols, oa, ob):
            assert x.tobytes() == y.tobytes() == v.tobytes() and vx == vy
        shapes = tuned.tuned_shapes()
        print("tuned shapes:", shapes, flush=True)
        assert all(s in (1, 2, 4, 8) for s in shapes.values()), shapes
        assert all(s == 0 for s in fixed.tuned_shapes().values())
    finally:
        tuned.close()
        fixed.close()


def _validity_shapes():
    """u32 columns whose validity blobs take every path of the bitmap decoder:
    runs spanning many 4096-bit windows, hundreds of runs per window, a raw
    bitmap, an all-null column, columns shorter than one 64-word window (their
    windows straddle columns), runs ending on word edges and at n."""
    r = np.random.default_rng(91)
    out = []

    def col(n, valid):
        v = r.integers(0, 1 << 20, n, dtype=np.uint64).astype(np.uint32)
        v[~valid] = 0
        out.append((v, np.packbits(valid, bitorder="little")))

    n = 600000
    valid = np.ones(n, bool)
    valid[r.choice(n, 5, replace=False)] = False
    col(n, valid)                                   # sparse: runs of ~10^5
    n = 70001
    valid = np.ones(n, bool)
    valid[1::2] = False
    col(n, valid)                                   # alternating: raw bitmap blob
    n = 200000
    valid = np.ones(n, bool)
    for s in range(0, n, 23):
        valid[s:s + 1 + (s % 3)] = False
    col(n, valid)                                   # ~350 runs per window
    col(5000, np.zeros(5000, bool))                 # all null
    valid = np.ones(100, bool)
    valid[[3, 4, 50, 99]] = False
    col(100, valid)                                 # shorter than a window
    n = 64 * 300
    valid = np.ones(n, bool)
    valid[64 * 10:64 * 20] = False
    valid[-64:] = False
    col(n, valid)                                   # word edges, a null run ending at n
    return out


def test_validity_shapes_in_one_batch(qa):
    """Every blob shape in one batch (so bitmap windows straddle columns):
    streams equal the oracle's, and the decode returns the oracle's values and
    bitmaps (src/validity.rs:40-119)."""
    cols = _validity_shapes()
    gpu = qa.Gpu(0)
    try:
        print(f"case validity shapes {len(cols)} columns", flush=True)
        got = _gpu.device_compress(qa, gpu, cols, qa.Config(level=qa.Level.FASTEST))
        for i, ((v, bm), g) in enumerate(zip(cols, got)):
            assert g == orc.compress(v, orc.U32, bm, orc.FASTEST), f"column {i}: stream"
        outs = _gpu.device_decompress(qa, gpu, got, [qa.DType.U32] * len(cols),
                                      [len(v) for v, _ in cols], True)
        for i, ((v, bm), (o, vo)) in enumerate(zip(cols, outs)):
            _, want, want_vb = orc.decompress(got[i])
            assert o.tobytes() == want.tobytes(), f"column {i}: values"
            exp = np.frombuffer(want_vb, dtype=np.uint8).copy()
            if len(v) % 8:
                exp[-1] &= (1 << (len(v) % 8)) - 1
            assert vo == exp.tobytes(), f"column {i}: bitmap"
            assert vo == bm.tobytes(), f"column {i}: bitmap vs input"
    finally:
        gpu.close()


def test_handles_in_parallel_threads(qa):
    """Three handles used from three host threads at once, each with its own
    streams, events and pinned staging: every stream equals the oracle's and
    every decode round-trips, on repeated calls."""
    import threading
    cfg = qa.Config(level=qa.Level.FASTEST)
    jobs = []
    for k in range(3):
        n = 150000 + 1000 * k
        ic = _data.int_columns(n, seed=40 + k)
        cols = [(ic["ts_i64"], None), _data.with_nulls(ic["u32_k11"], 0.02, seed=50 + k)]
        want = [orc.compress(v, ODT[v.dtype], bm, orc.FASTEST) for v, bm in cols]
        jobs.append((n, cols, want))
    errors = []

    def work(k):
        n, cols, want = jobs[k]
        try:
            g = qa.Gpu(0)
            try:
                for rep in range(3):
                    got = _gpu.device_compress(qa, g, cols, cfg)
                    assert got == want, f"thread {k} call {rep}: streams"
                    outs = _gpu.device_decompress(qa, g, got, [qa.FROM_NP[v.dtype] for v, _ in cols],
                                                  [n, n], True)
                    for (v, _), (o, _) in zip(cols, outs):
                        assert o.tobytes() == v.tobytes(), f"thread {k} call {rep}: decode"
            finally:
                g.close()
        except Exception as e:  # reported by the main thread
            errors.append((k, repr(e)))

    print("case three handles in three threads", flush=True)
    ts = [threading.Thread(target=work, args=(k,)) for k in range(3)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in ts), "a thread did not finish"
    assert not errors, errors
