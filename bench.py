"""Benchmark of the MI355X quoin block path on the integer-lanes workload.

Workload (one GPU): 16 Mi i64 monotone timestamps (start 1.7e18, deltas uniform
[1, 1000]) plus three 16 Mi u32 columns uniform in [0, 2^k), k in {3, 11, 23},
each with a Bernoulli 1 % null bitmap. All blocks of all four columns go through
one batched launch sequence. A step = compress the four columns + decompress the
four streams, device-resident (inputs already in HBM). The first call is checked
bit for bit: every stream byte against the CPU oracle at the same level, and
every round trip in full. Further legs: the plain form, serving, rANS decode,
the residual coders, the mixed-float workload and Config::default (Level::Max).

Prints one JSON line: the metric, its value and the configuration, the
roofline of the dominant kernels and the CPU baseline. --gpus > 1 exits: the
benchmark runs on one GPU.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

GENERATOR = ("xoshiro256** seeded by SplitMix64, Box-Muller normals, Lemire bounded integers "
             "(bench_gen.c); checksums: 64-bit FNV-1a over little-endian words")
_G = None


def _gen():
    """The column generator (bench_gen.so, built by __graft_entry__.build())."""
    global _G
    if _G is None:
        path = os.path.join(ROOT, "bench_gen.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built: run __graft_entry__.build()")
        g = ctypes.CDLL(path)
        u64, dbl, vp = ctypes.c_uint64, ctypes.c_double, ctypes.c_void_p
        for name in ("walk_f64", "timestamps_i64", "decimal_f64", "full_f64", "dict_runs_f64",
                     "sine_f64", "unif_f32", "raw"):
            getattr(g, "qg_" + name).argtypes = [u64, u64, vp]
        g.qg_kbits_u32.argtypes = [u64, u64, ctypes.c_uint32, vp]
        g.qg_validity.argtypes = [u64, u64, dbl, vp, vp]
        g.qg_loguniform.argtypes = [u64, u64, dbl, dbl, vp]
        g.qg_checksum.argtypes = [vp, u64]
        g.qg_checksum.restype = u64
        _G = g
    return _G


def gen(name, seed, n, dtype):
    """n values of bench_gen.c's qg_<name> column from `seed`."""
    out = np.empty(n, dtype=dtype)
    getattr(_gen(), "qg_" + name)(seed, n, out.ctypes.data)
    return out


def gen_kbits(seed, n, k):
    out = np.empty(n, dtype=np.uint32)
    _gen().qg_kbits_u32(seed, n, k, out.ctypes.data)
    return out


def gen_validity(seed, n, p_valid, values=None):
    """A Bernoulli(p_valid) bitmap (LSB first); null slots of `values` (4-byte
    elements) are zeroed."""
    bm = np.empty((n + 7) // 8, dtype=np.uint8)
    _gen().qg_validity(seed, n, p_valid, bm.ctypes.data, values.ctypes.data if values is not None else None)
    return bm


def checksum(a):
    a = np.ascontiguousarray(a)
    return "%016x" % _gen().qg_checksum(a.ctypes.data, a.nbytes)


def column_checksums(cols):
    """name -> checksum of the values (and of the bitmap, when there is one)."""
    out = {}
    for c in cols:
        name, v = c[0], c[1]
        bm = c[2] if len(c) > 2 else None
        out[name] = checksum(v) + ("/" + checksum(bm) if bm is not None else "")
    return out


def make_workload(n, seed=100):
    """The integer-lanes columns: n i64 timestamps (t_0 = 1.7e18, steps uniform
    in [1, 1000]) and three n u32 columns of uniform k-bit values, k in
    {3, 11, 23}, each with a Bernoulli(0.99) validity bitmap (null slots 0).
    Column j's values come from seed + j, its bitmap from seed + 10 + j."""
    cols = [("ts_i64", gen("timestamps_i64", seed, n, np.int64), None)]
    for j, k in enumerate((3, 11, 23), 1):
        v = gen_kbits(seed + j, n, k)
        bm = gen_validity(seed + 10 + j, n, 0.99, v)
        cols.append((f"u32_k{k}", v, bm))
    return cols


def measured_traffic(stage, args=None, name=None):
    """HBM bytes per call of a stage, from the committed counter summary
    (profiles/traffic_<name>.json: rocprofv3 FETCH_SIZE and WRITE_SIZE passes,
    scaled by the same counters on a plain copy of known size), or None when
    no summary matches this configuration. The integer-lanes summaries are
    per level (name from args) and must match the column length."""
    fname = name or f"intlanes_{args.level}"
    path = os.path.join(ROOT, "profiles", f"traffic_{fname}.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    st = t.get("stages", {}).get(stage)
    if not st or (name is None and t.get("values_per_column") != args.n):
        return None
    return st.get("hbm_bytes_per_call")


def host_cores():
    """Threads the CPU baseline uses on this host (the box's CPU share is 16
    for one GPU) and the CPU model."""
    n = os.cpu_count() or 1
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return min(n, 16), n, model


def oracle_round_trip(cols, level, threads):
    """Seconds of one oracle compress + decompress of every column, input in
    memory to output in memory (the C calls only: no Python copies of the
    decoded values). cols: [(values, engine dtype, bitmap or None)]."""
    import ctypes
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as orc
    lib = orc.lib
    t0 = time.perf_counter()
    for v, dt, bm in cols:
        out, ln = ctypes.POINTER(ctypes.c_uint8)(), ctypes.c_size_t()
        rc = lib.orc_compress_column_cfg(dt, len(v), orc._ptr(v), orc._ptr(bm), int(level), 0, 0, 16,
                                         threads, ctypes.byref(out), ctypes.byref(ln))
        assert rc == 0, rc
        dtv, n, vals = ctypes.c_int(), ctypes.c_size_t(), ctypes.c_void_p()
        vb = ctypes.POINTER(ctypes.c_uint8)()
        rc = lib.orc_decompress_column_mt(out, ln.value, threads, ctypes.byref(dtv), ctypes.byref(n),
                                          ctypes.byref(vals), ctypes.byref(vb))
        assert rc == 0, rc
        lib.orc_free(out)
        lib.orc_free(vals)
        if vb:
            lib.orc_free(vb)
    return time.perf_counter() - t0


def cpu_baseline(cols, level, runs=5, one_thread_values=4 << 20, label=""):
    """The oracle (the C restatement of the reference encoder and decoder,
    blocks fanned out over threads like the reference's rayon pool) timed on
    this host: all cores (up to 16) over the whole workload, median of `runs`;
    and one thread over the first `one_thread_values` values of each column,
    median of 3. cols: [(values, engine dtype, bitmap or None)]."""
    threads, ncpu, model = host_cores()
    raw = sum(v.nbytes for v, _, _ in cols)
    oracle_round_trip([(v[:4096], dt, bm[:512] if bm is not None else None) for v, dt, bm in cols],
                      level, threads)  # warm-up
    times = sorted(oracle_round_trip(cols, level, threads) for _ in range(runs))
    sample = [(v[:one_thread_values], dt, bm[:(one_thread_values + 7) // 8] if bm is not None else None)
              for v, dt, bm in cols]
    raw1 = sum(v.nbytes for v, _, _ in sample)
    t1 = sorted(oracle_round_trip(sample, level, 1) for _ in range(3))
    med, med1 = times[len(times) // 2], t1[1]
    sums = [checksum(v) + ("/" + checksum(bm) if bm is not None else "") for v, _, bm in cols]
    return {"value": round(raw / med / 1e9, 4), "unit": "GB/s", "cores": threads, "kind": "port",
            "input_checksums": sums,
            "one_thread_gbs": round(raw1 / med1 / 1e9, 4), "nproc": ncpu, "cpu_model": model,
            "runs": runs, "median_s": round(med, 4),
            "sample": f"{label}the whole workload ({raw / 1e6:.1f} MB), compress+decompress "
                      f"at level {int(level)}, oracle C over {threads} threads, median of {runs} "
                      f"runs; one_thread_gbs: first {one_thread_values} values of each column "
                      f"({raw1 / 1e6:.1f} MB), 1 thread, median of 3"}


def serve_pool(qa, gpu, cfg, dev, count, seed=400):
    """Requests of the serving workload: uncompressed sizes log-uniform in
    [4 KB, 64 MB], types f64/f32/i64/i32 in turn with the generators of the
    other workloads (f64: 2-decimal walk, uniform doubles, 256-value runs;
    f32 uniform [-1, 1); i64 timestamps; i32 uniform k-bit), half compress and
    half decompress (of a stream made beforehand). Device buffers per request.
    Sizes come from `seed`, request i's values from seed + 1 + i."""
    import torch
    sizes = np.empty(count, dtype=np.float64)
    _gen().qg_loguniform(seed, count, float(4 << 10), float(64 << 20), sizes.ctypes.data)
    pool = []
    for i in range(count):
        nbytes = int(sizes[i])
        kind = ["f64", "f32", "i64", "i32"][i % 4]
        w = 4 if kind in ("f32", "i32") else 8
        n = max(nbytes // w, 1)
        rs = seed + 1 + i  # request i's values
        if kind == "f64":
            v = gen(("walk_f64", "full_f64", "dict_runs_f64")[i // 4 % 3], rs, n, np.float64)
        elif kind == "f32":
            v = gen("unif_f32", rs, n, np.float32)
        elif kind == "i64":
            v = gen("timestamps_i64", rs, n, np.int64)
        else:
            v = gen_kbits(rs, n, [3, 11, 23][i // 4 % 3]).view(np.int32)
        v = np.ascontiguousarray(v)
        dt = qa.FROM_NP[v.dtype]
        t = torch.from_numpy(v.view(np.uint8).copy()).to(dev)
        cap = qa.lib.quoin_typed_compress_bound(int(dt), n)
        s = torch.empty(cap, dtype=torch.uint8, device=dev)
        col = qa.DeviceColumn(int(dt), 0, n, t.data_ptr(), None, s.data_ptr(), cap)
        req = {"kind": "c" if i % 2 == 0 else "d", "bytes": v.nbytes, "t": t, "s": s,
               "col": col, "busy": False}
        if req["kind"] == "d":
            ln = gpu.compress([col], cfg)[0]
            host = s[:ln].cpu().numpy().copy()
            src = torch.from_numpy(host.copy()).to(dev)
            out = torch.empty(v.nbytes, dtype=torch.uint8, device=dev)
            req.update(host=host, src=src, out=out,
                       st=qa.DeviceStream(int(dt), 0, n, src.data_ptr(), host.ctypes.data, ln,
                                          out.data_ptr(), None))
        pool.append(req)
    torch.cuda.synchronize()
    return pool


def serve_run(qa, srv, pool, cfg, rate, count, seed=2):
    """Open loop at `rate` requests/s (Poisson; rate None = all at once).
    Returns (GB/s, latencies in ms, wall seconds)."""
    import threading
    import queue
    # exponential gaps -ln(u) / rate, u uniform in (0, 1] from the generator
    gaps = None
    if rate:
        raw = gen("raw", seed, count, np.uint64)
        gaps = -np.log(((raw >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0**-53) / rate
    q = queue.Queue()
    lat = []

    def collect():
        while True:
            item = q.get()
            if item is None:
                return
            req, t = item
            rc, _, ms = srv.wait(t)
            if rc != qa.OK:
                raise RuntimeError(f"served job failed: {rc}")
            lat.append((ms, req["bytes"]))
            req["busy"] = False

    th = threading.Thread(target=collect)
    th.start()
    done_bytes = 0
    t0 = time.perf_counter()
    next_t = t0
    k = 0
    for j in range(count):
        if rate:
            next_t += gaps[j]
            d = next_t - time.perf_counter()
            if d > 0:
                time.sleep(d)
        # the next request whose buffers are free
        for _ in range(len(pool)):
            req = pool[k % len(pool)]
            k += 1
            if not req["busy"]:
                break
        else:
            while req["busy"]:
                time.sleep(1e-4)
        req["busy"] = True
        if req["kind"] == "c":
            t = srv.submit_compress(req["col"], cfg)
        else:
            t = srv.submit_decompress(req["st"])
        done_bytes += req["bytes"]
        q.put((req, t))
    q.put(None)
    th.join()
    wall = time.perf_counter() - t0
    return done_bytes / wall / 1e9, lat, wall


def serve_bench(qa, dev, budget_gb=64, slots=3, pool_size=192, per_point=800):
    cfg = qa.Config(level=qa.Level.FAST)
    gpu = qa.Gpu(0)
    pool = serve_pool(qa, gpu, cfg, dev, pool_size)
    gpu.close()
    srv = qa.Server(0, budget_gb << 30, slots)
    try:
        serve_run(qa, srv, pool, cfg, None, len(pool))  # warm the slots' scratch
        sat_gbs, lat, wall = serve_run(qa, srv, pool, cfg, None, per_point)
        sat_rps = per_point / wall
        points = []
        for frac in (0.25, 0.5, 0.75, 0.9):
            j0, b0 = srv.stats()
            gbs, lat, wall = serve_run(qa, srv, pool, cfg, frac * sat_rps, per_point,
                                      seed=500 + int(frac * 100))
            j1, b1 = srv.stats()
            ms = [m for m, _ in lat]
            sm = [m for m, b in lat if b <= 1 << 20] or [0.0]
            lg = [m for m, b in lat if b > 1 << 20] or [0.0]
            points.append({"offered_rps": round(frac * sat_rps, 1), "achieved_gbs": round(gbs, 3),
                           "p50_ms": round(float(np.percentile(ms, 50)), 3),
                           "p99_ms": round(float(np.percentile(ms, 99)), 3),
                           "small_le_1mb": {"p50_ms": round(float(np.percentile(sm, 50)), 3),
                                            "p99_ms": round(float(np.percentile(sm, 99)), 3)},
                           "large": {"p50_ms": round(float(np.percentile(lg, 50)), 3),
                                     "p99_ms": round(float(np.percentile(lg, 99)), 3)},
                           "jobs": j1 - j0, "batches": b1 - b0})
    finally:
        srv.close()
    hi = points[-1]
    return {"workload": f"{pool_size} requests, 4 KB-64 MB log-uniform, f64/f32/i64/i32, "
                        f"half compress half decompress, Level::Fast, Poisson arrivals",
            "hbm_budget_gb": budget_gb, "slots": slots,
            "saturation_gbs": round(sat_gbs, 3), "saturation_rps": round(sat_rps, 1),
            "p50_ms": hi["p50_ms"], "p99_ms": hi["p99_ms"], "at_offered_rps": hi["offered_rps"],
            "p99_ms_small_le_1mb": hi["small_le_1mb"]["p99_ms"],
            "scheduling": "slot 0 serves requests of up to 1 MiB only; each slot's arenas are "
                          "capped at budget / slots",
            "points": points}


def entropy_lanes(n, seed=300):
    """The entropy workload's two columns as f64 lanes: n f64 values
    sin(2 pi i / 1000) * 1e3 + N(0, 1e-3) (from seed) and n f32 values
    ((u32 >> 8) 2^-23) - 1 (from seed + 1) widened to f64."""
    f64 = gen("sine_f64", seed, n, np.float64)
    f32 = gen("unif_f32", seed + 1, n, np.float32)
    return f64, f32.astype(np.float64)


def linear_blocks(qa, gpu, torch, V, block, kind, order=0):
    """Residual streams of every `block`-value block of the device f64 lane V in
    one quoin_gpu_linear_encode call (src/codecs/linear.rs). Returns the slotted
    output tensor, the LinearInput list, and each block's (offset, length)."""
    nb = V.numel() // block
    cap = int(qa.lib.quoin_linear_bound(block))
    out = torch.empty(nb * cap, dtype=torch.uint8, device=V.device)
    vp, op = V.data_ptr(), out.data_ptr()
    ins = [qa.LinearInput(vp + 8 * b * block, block, op + b * cap, cap) for b in range(nb)]
    lens, _, sts = gpu.linear_encode(kind, ins, order)
    spans = [(b * cap, lens[b]) for b in range(nb) if not sts[b]]
    return out, ins, spans


def resid_bench(qa, gpu, dev, n, steps, warmup, block=32768, seed=300):
    """The linear residual coders (src/codecs/linear.rs) and the FCM / DFCM
    predictors (src/codecs/pred.rs) on n f64 values
    sin(2 pi i / 1000) * 1e3 + N(0, 1e-3) in 32768-value blocks, every block of
    a kind in one call: encode = select_order + the residual streams, decode =
    the varint parse + the recurrence, checked against the input. GB/s are of
    raw f64 bytes; the roofline is Delta2 encode's (n * 8 + stream bytes per
    call over its kernels' time)."""
    import torch
    f64, _ = entropy_lanes(n, seed)
    V = torch.from_numpy(f64.view(np.int64)).to(dev)
    nb = n // block
    raw = nb * block * 8
    res = {"workload": f"{nb} blocks of {block} f64 values, sine + noise", "kinds": {}}
    samples = []
    for name, kind in (("delta2", qa.LINEAR_DELTA2), ("delta_dp", qa.LINEAR_DELTA_DP),
                       ("ordered_delta", qa.LINEAR_ORDERED_DELTA)):
        out, ins, _ = linear_blocks(qa, gpu, torch, V, block, kind)
        lens, ords, sts = gpu.linear_encode(kind, ins)
        D = torch.zeros_like(V)
        ss = [qa.LinearStream(ins[b].dst, lens[b], block, D.data_ptr() + 8 * b * block)
              for b in range(nb) if not sts[b]]
        dst = gpu.linear_decode(kind, ss)
        torch.cuda.synchronize()
        assert all(x == qa.OK for x in dst), f"{name}: decode failed"
        ok = [b for b in range(nb) if not sts[b]]
        if ok:
            sel = torch.tensor(ok, device=dev)
            assert torch.equal(D.view(nb, block)[sel], V[:nb * block].view(nb, block)[sel]), \
                f"{name}: round trip mismatch"
        for _ in range(warmup):
            gpu.linear_encode(kind, ins)
        ek = 0.0
        for _ in range(steps):
            gpu.linear_encode(kind, ins)
            ek += gpu.last_timing()[1]
        for _ in range(warmup):
            gpu.linear_decode(kind, ss)
        dt = dk = 0.0
        for _ in range(steps):
            gpu.linear_decode(kind, ss)
            t, k, _ = gpu.last_timing()
            dt += t
            dk += k
        ek, dt, dk = ek / steps, dt / steps, dk / steps
        coded = int(sum(lens))
        res["kinds"][name] = {
            "stream_bytes": coded, "declined_blocks": int(sum(sts)),
            "orders": {str(o): ords.count(o) for o in sorted(set(ords))} if kind != 7 else None,
            "encode_ms": round(ek, 4), "encode_gbs": round(raw / (ek * 1e-3) / 1e9, 1),
            "decode_ms": round(dt, 4), "decode_gbs": round(raw / (dt * 1e-3) / 1e9, 1),
            "recurrence_ms": round(dk, 4)}
        if name == "delta2":
            alg = raw + coded
            res["roofline"] = {"bound": "hbm", "kernel": "linear encode: k_lin_order + k_lin_enc",
                               "kernel_ms": round(ek, 4), "algorithmic_bytes": alg,
                               "achieved": round(alg / (ek * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                               "unit": "GB/s",
                               "frac": round(alg / (ek * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)}
            samples = [f64[b * block:(b + 1) * block].view(np.uint64) for b in range(0, nb, 64)]
        del out, D
    # FCM / DFCM (src/codecs/pred.rs), predictor_log2 16, on the same blocks
    cap = int(qa.lib.quoin_linear_bound(block))
    for name, kind in (("fcm", qa.PRED_FCM), ("dfcm", qa.PRED_DFCM)):
        out = torch.empty(nb * cap, dtype=torch.uint8, device=dev)
        ins = [qa.LinearInput(V.data_ptr() + 8 * b * block, block, out.data_ptr() + b * cap, cap)
               for b in range(nb)]
        lens = gpu.pred_encode(kind, ins, 16)
        D = torch.zeros_like(V)
        ss = [qa.LinearStream(ins[b].dst, lens[b], block, D.data_ptr() + 8 * b * block) for b in range(nb)]
        st = gpu.pred_decode(kind, ss, 16)
        torch.cuda.synchronize()
        assert all(x == qa.OK for x in st) and torch.equal(D[:nb * block], V[:nb * block]), \
            f"{name}: round trip mismatch"
        for _ in range(warmup):
            gpu.pred_encode(kind, ins, 16)
        ek = 0.0
        for _ in range(steps):
            gpu.pred_encode(kind, ins, 16)
            ek += gpu.last_timing()[1]
        dt = dk = 0.0
        for _ in range(max(steps // 5, 1)):
            gpu.pred_decode(kind, ss, 16)
            t, k, _ = gpu.last_timing()
            dt += t
            dk += k
        ek, dt, dk = ek / steps, dt / max(steps // 5, 1), dk / max(steps // 5, 1)
        res["kinds"][name] = {
            "stream_bytes": int(sum(lens)), "predictor_log2": 16,
            "encode_ms": round(ek, 4), "encode_gbs": round(raw / (ek * 1e-3) / 1e9, 1),
            "decode_ms": round(dt, 4), "decode_gbs": round(raw / (dt * 1e-3) / 1e9, 2),
            "recurrence_ms": round(dk, 4)}
        del out, D
    del V
    return res, samples


def resid_cpu_baseline(samples):
    """Oracle Delta2 encode (select_order included) and decode, single-threaded C,
    over every 64th block."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as orc
    t0 = time.perf_counter()
    enc = [orc.linear_encode(20, b) for b in samples]
    te = time.perf_counter() - t0
    t0 = time.perf_counter()
    for e, b in zip(enc, samples):
        orc.linear_decode(20, e, len(b))
    td = time.perf_counter() - t0
    raw = sum(len(b) for b in samples) * 8
    return {"value": round(raw / te / 1e9, 4), "decode_value": round(raw / td / 1e9, 4),
            "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"{len(samples)} of the 32768-value blocks, Delta2 encode / decode, "
                      "oracle C, 1 thread"}


def entropy_bench(qa, gpu, dev, n, steps, warmup, block=32768, seed=300):
    """rANS decode on the entropy workload (entropy_lanes): each 32768-value
    block's Delta2 residual bytes, from the device residual coder, are rANS-coded
    by the GPU coder, then every block stream is decoded in one launch; decoded
    bytes are checked against the residuals."""
    import torch
    res, bspans = [], []
    lanes = entropy_lanes(n, seed)
    input_sums = {"sine_f64": checksum(lanes[0]), "unif_f32 (as f32)": checksum(lanes[1].astype(np.float32))}
    for lane in lanes:
        V = torch.from_numpy(lane.view(np.int64)).to(dev)
        out, _, sp = linear_blocks(qa, gpu, torch, V, block, qa.LINEAR_DELTA2)
        res.append(torch.cat([out[o:o + ln] for o, ln in sp]))
        bspans.append([ln for _, ln in sp])
        del V, out
    syms = torch.cat(res)
    spans, off = [], 0
    for lns in bspans:
        for ln in lns:
            spans.append((off, ln))
            off += ln
    caps = [qa.lib.quoin_rans_bound(ln) for _, ln in spans]
    coff = np.concatenate([[0], np.cumsum(caps)]).tolist()
    coded = torch.empty(coff[-1], dtype=torch.uint8, device=dev)
    dec = torch.empty(syms.numel(), dtype=torch.uint8, device=dev)
    sp, cp, dp = syms.data_ptr(), coded.data_ptr(), dec.data_ptr()
    ins = [qa.RansInput(sp + o, ln, cp + coff[k], caps[k]) for k, (o, ln) in enumerate(spans)]
    torch.cuda.synchronize()
    lens = gpu.rans_compress(ins)
    t0 = time.perf_counter()
    for _ in range(2):
        gpu.rans_compress(ins)
    torch.cuda.synchronize()
    enc_s = (time.perf_counter() - t0) / 2
    sts = [qa.RansStream(cp + coff[k], lens[k], ln, dp + o) for k, (o, ln) in enumerate(spans)]
    dl, st = gpu.rans_decompress(sts)
    torch.cuda.synchronize()
    assert all(x == qa.OK for x in st) and torch.equal(dec, syms), "rANS round trip mismatch"
    for _ in range(warmup):
        gpu.rans_decompress(sts)
    torch.cuda.synchronize()
    kern = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        gpu.rans_decompress(sts)
        kern += gpu.last_timing()[1]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    nsym = syms.numel()
    cbytes = int(sum(lens))
    k_ms = kern / steps
    alg = cbytes + nsym
    sample = [bytes(coded[coff[k]:coff[k] + lens[k]].cpu().numpy()) for k in range(0, len(spans), 64)]
    sample_len = [spans[k][1] for k in range(0, len(spans), 64)]
    del coded, dec
    layer = entropy_layer_bench(qa, gpu, syms.device, syms, spans)
    return {
        "entropy_layer": layer,
        "workload": "entropy decode: 32Mi f64 sine+noise and 32Mi f32 uniform[-1,1), Delta2 "
                    "residual bytes of each 32768-value block, rANS-coded (one stream per block)",
        "inputs": {"generator": GENERATOR, "checksums": input_sums},
        "streams": len(spans), "symbols": nsym, "coded_bytes": cbytes,
        "decode_gbs": round(nsym / (k_ms * 1e-3) / 1e9, 2),
        "decode_wall_gbs": round(nsym * steps / wall / 1e9, 2),
        "encode_gbs": round(nsym / enc_s / 1e9, 2),
        "roofline": {"bound": "hbm", "kernel": "k_rans_dec", "kernel_ms": round(k_ms, 4),
                     "traffic": measured_traffic("rans_decode", name="entropy_rans"),
                     "algorithmic_bytes": alg,
                     "achieved": round(alg / (k_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(alg / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)},
    }, (sample, sample_len)


def entropy_layer_bench(qa, gpu, dev, syms, spans, every=16):
    """The residual entropy layer (src/entropy/mod.rs code_residuals /
    decode_residuals) on every `every`-th block's Delta2 residual bytes of the
    entropy workload, at Level::Max (lambda 0 + LZ cascade: the range coder and
    LZ on host cores, rANS on the GPU) and Level::Balanced (lambda 2:
    rANS-only). Decode is checked against the residuals."""
    import torch
    pick = list(range(0, len(spans), every))
    caps = [int(qa.lib.quoin_residual_bound(spans[k][1])) for k in pick]
    coff = np.concatenate([[0], np.cumsum(caps)]).tolist()
    coded = torch.empty(coff[-1], dtype=torch.uint8, device=dev)
    dec = torch.empty(sum(spans[k][1] for k in pick), dtype=torch.uint8, device=dev)
    doff = np.concatenate([[0], np.cumsum([spans[k][1] for k in pick])]).tolist()
    sp, cp, dp = syms.data_ptr(), coded.data_ptr(), dec.data_ptr()
    ins = [qa.RansInput(sp + spans[k][0], spans[k][1], cp + coff[j], caps[j])
           for j, k in enumerate(pick)]
    raw = int(sum(spans[k][1] for k in pick))
    out = {"streams": len(pick), "residual_bytes": raw, "hbm_workers": "rANS on the GPU; range "
           "coder and LZ cascade on host cores", "levels": {}}
    for name, lam, lz in (("max", 0, True), ("balanced", 2, False)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lens = gpu.code_residuals(ins, lam, lz)
        torch.cuda.synchronize()
        te = time.perf_counter() - t0
        tags = [int(coded[coff[j]].item()) for j in range(len(pick))]
        sts = [qa.RansStream(cp + coff[j], lens[j], spans[k][1], dp + doff[j])
               for j, k in enumerate(pick)]
        t0 = time.perf_counter()
        dl, st = gpu.decode_residuals(sts)
        torch.cuda.synchronize()
        td = time.perf_counter() - t0
        ok = all(x == qa.OK for x in st)
        for j, k in enumerate(pick):
            o, ln = spans[k]
            ok = ok and torch.equal(dec[doff[j]:doff[j] + ln], syms[o:o + ln])
        assert ok, f"entropy layer round trip mismatch at {name}"
        out["levels"][name] = {
            "coded_bytes": int(sum(lens)),
            "tags": {t: tags.count(c) for t, c in (("rc", 0), ("rans", 2), ("lz", 3)) if c in tags},
            "host_routed_streams": sum(1 for t in tags if t != 2),
            "encode_gbs": round(raw / te / 1e9, 3), "decode_gbs": round(raw / td / 1e9, 3)}
    return out


def rans_cpu_baseline(sample, sample_len):
    """Oracle rANS decode (single-threaded C) over every 64th block stream."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as orc
    t0 = time.perf_counter()
    for c, ln in zip(sample, sample_len):
        orc.rans_decompress(c, ln)
    dt = time.perf_counter() - t0
    return {"value": round(sum(sample_len) / dt / 1e9, 4), "unit": "GB/s", "cores": 1,
            "kind": "port", "sample": f"{len(sample)} of the block streams ({sum(sample_len)} symbols), "
                                      "rANS decode, oracle C, 1 thread"}


MODE_NAMES = {0: "pred", 1: "const", 2: "stride", 3: "xorz", 4: "lz", 5: "raw",
              7: "ordered_delta", 9: "alp", 11: "byte_transpose", 17: "float_mult", 19: "dict",
              20: "delta2", 25: "pred2", 29: "pred_rc", 45: "delta_dp", 50: "for_bitpack",
              51: "delta_bitpack", 52: "alp_rd", 53: "rle", 54: "pco"}
# winners whose payload the device emits at the entropy levels; the others are
# written by the host coders (qa_hybrid.cpp)
DEVICE_EMITTED = {1, 2, 3, 5, 9, 50, 51, 53}


def float_columns(total=64 << 20, ncols=40, seed=200):
    """The mixed-float workload: `ncols` f64 columns whose lengths are
    exp(U[ln 65536, ln 8388608]) rescaled to sum to `total` and kept in that
    range. Column i is, by i % 4: 0, 1 decimals k / 1000 with k uniform in
    [0, 10^9) (scale 10^3 in [0, 1e6)); 2 full-precision uniform [0, 1) doubles
    (u64 >> 11) 2^-53; 3 a 256-value dictionary of such decimals, written in
    runs of Geometric(1/8) length (mean 8), each run's value uniform over the
    256. Lengths come from `seed`, column i's values from seed + 1 + i."""
    lo, hi = 65536, 8388608
    L = np.empty(ncols, dtype=np.float64)
    _gen().qg_loguniform(seed, ncols, float(lo), float(hi), L.ctypes.data)
    for _ in range(50):
        L = np.clip(L * total / L.sum(), lo, hi)
    L = np.floor(L).astype(np.int64)
    for i in np.argsort(L):  # the rounding remainder goes to columns with room
        d = int(total - L.sum())
        if d == 0:
            break
        L[i] = min(max(L[i] + d, lo), hi)
    cols = []
    for i, n in enumerate(L.tolist()):
        kind = ("decimal", "decimal", "full", "dict")[i % 4]
        fn = {"decimal": "decimal_f64", "full": "full_f64", "dict": "dict_runs_f64"}[kind]
        cols.append((kind, gen(fn, seed + 1 + i, n, np.float64)))
    return cols


def float_bench(qa, gpu, dev, steps, warmup, check_oracle=True):
    """The mixed-float workload (float_columns) at Level::Fast: all blocks of
    all 40 columns in one batched compress call and one batched decompress
    call, device-resident. The first call is checked: every column's round
    trip in full, and the stream bytes of the shortest column of each kind
    against the oracle. Reports GB/s of uncompressed f64, the per-mode block
    wins, and the tile decode stage (k_dec_tiles / k_dec_fl) against 8 TB/s."""
    import torch
    cols = float_columns()
    cfg = qa.Config(level=qa.Level.FAST)
    dv, ds, do, dcols = [], [], [], []
    raw = 0
    for kind, v in cols:
        t = torch.from_numpy(v.view(np.uint8)).to(dev)
        cap = qa.lib.quoin_typed_compress_bound(int(qa.DType.F64), len(v))
        s = torch.empty(cap, dtype=torch.uint8, device=dev)
        dv.append(t)
        ds.append(s)
        do.append(torch.empty(v.nbytes, dtype=torch.uint8, device=dev))
        dcols.append(qa.DeviceColumn(int(qa.DType.F64), 0, len(v), t.data_ptr(), None,
                                     s.data_ptr(), cap))
        raw += v.nbytes
    torch.cuda.synchronize()
    qa.reset_mode_win_counts()
    lens = gpu.compress(dcols, cfg)
    wins = qa.mode_win_counts()
    host = [ds[i][:lens[i]].cpu().numpy() for i in range(len(cols))]
    dsts = [qa.DeviceStream(int(qa.DType.F64), 0, len(v), ds[i].data_ptr(), host[i].ctypes.data,
                            lens[i], do[i].data_ptr(), None) for i, (_, v) in enumerate(cols)]
    gpu.decompress(dsts)
    torch.cuda.synchronize()
    for i in range(len(cols)):
        assert torch.equal(do[i], dv[i]), f"mixed-float round trip mismatch on column {i}"
    checked = []
    if check_oracle:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as orc
        threads = host_cores()[0]
        for i, (_, v) in enumerate(cols):
            exp = orc.compress(v, orc.F64, None, orc.FAST, threads=threads)
            assert host[i].tobytes() == exp, f"mixed-float stream differs from the oracle on column {i}"
            checked.append(i)
    dcols = (qa.DeviceColumn * len(dcols))(*dcols)
    dsts = (qa.DeviceStream * len(dsts))(*dsts)
    for _ in range(warmup):
        gpu.compress(dcols, cfg)
        gpu.decompress(dsts)
    torch.cuda.synchronize()
    dc = dd = kd = kc = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        gpu.compress(dcols, cfg)
        t, k, _ = gpu.last_timing()
        dc += t
        kc += k
        gpu.decompress(dsts)
        t, k, _ = gpu.last_timing()
        dd += t
        kd += k
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    dc, dd, kd, kc = dc / steps, dd / steps, kd / steps, kc / steps
    comp = int(sum(lens))
    alg = comp + raw
    return {
        "workload": "mixed float: 64Mi f64 in 40 columns, lengths log-uniform 64Ki-8Mi; 20 decimal "
                    "(scale 10^3), 10 full-precision uniform [0,1), 10 256-value dictionary runs "
                    "(geometric, mean 8); Level::Fast, one batched launch sequence per direction",
        "column_lengths": [len(v) for _, v in cols], "uncompressed_bytes": raw,
        "compressed_bytes": comp, "ratio": round(raw / comp, 3),
        "round_trip_gbs": round(raw * steps / wall / 1e9, 2),
        "compress_gbs": round(raw / (dc * 1e-3) / 1e9, 2),
        "decompress_gbs": round(raw / (dd * 1e-3) / 1e9, 2),
        "device_ms_compress": round(dc, 3), "device_ms_decompress": round(dd, 3),
        "wins": {MODE_NAMES.get(m, str(m)): int(wins[m]) for m in range(64) if wins[m]},
        "inputs": {"generator": GENERATOR,
                   "checksums": {f"{i}:{kind}": checksum(v) for i, (kind, v) in enumerate(cols)}},
        "oracle_checked_columns": checked,
        "roofline": {"bound": "hbm", "kernel": "tile decode: k_dec_tiles + k_dec_fl",
                     "kernel_ms": round(kd, 4), "algorithmic_bytes": alg,
                     "traffic": measured_traffic("tile_decode", name="mixed_float_fast"),
                     "achieved": round(alg / (kd * 1e-3) / 1e9, 1) if kd > 0 else 0.0,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(alg / (kd * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if kd > 0 else 0.0,
                     "decode_sequence_gbs": round(alg / (dd * 1e-3) / 1e9, 1)},
        "encode_roofline": {"bound": "hbm", "kernel": "tile encode: k_emit_fl + k_emit_tiles",
                            "kernel_ms": round(kc, 4),
                            "traffic": measured_traffic("tile_encode", name="mixed_float_fast"),
                            "achieved": round(alg / (kc * 1e-3) / 1e9, 1) if kc > 0 else 0.0,
                            "encode_sequence_gbs": round(alg / (dc * 1e-3) / 1e9, 1)},
    }, cols


def float_cpu_baseline(cols):
    """The oracle at Level::Fast over the 40 mixed-float columns."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as orc
    return cpu_baseline([(v, orc.F64, None) for _, v in cols], orc.FAST, runs=3,
                        one_thread_values=256 << 10, label="mixed float: ")


def default_workload(seed=0):
    """Columns of the Config::default leg: the plain-method column (1 Mi f64
    2-decimal random walk, seed 0), and the first 1 Mi values of each
    integer-lanes column and 2 Mi of each entropy column (the same
    generators and seeds as those legs)."""
    n = 1 << 20
    cols = [("walk_f64 (1Mi, plain method)", gen("walk_f64", seed, n, np.float64), None)]
    for name, v, bm in make_workload(1 << 20):
        cols.append((name + " (1Mi)", v, bm))
    m = 2 << 20
    f64, f32 = entropy_lanes(m)
    cols.append(("sine_f64 (2Mi)", f64, None))
    cols.append(("unif_f32 (2Mi)", f32.astype(np.float32), None))
    return cols


def default_bench(qa, gpu, dev, steps, check_oracle=True):
    """Config::default (Level::Max) over default_workload(): the device plans
    and sizes, writes the residual streams and emits its own modes; host cores
    run the rest of the competition. One batched compress call and one batched
    decompress call per step, device-resident. The first call's streams are
    checked against the oracle at Level::Max byte for byte, and every round
    trip in full. Reports GB/s, the per-mode wins, and the split of blocks
    between device-emitted and host-written payloads."""
    import torch
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as orc
    cols = default_workload()
    cfg = qa.Config()  # Config::default
    odt = {np.dtype(np.float64): orc.F64, np.dtype(np.int64): orc.I64,
           np.dtype(np.uint32): orc.U32, np.dtype(np.float32): orc.F32}
    dv, ds, do, dvo, dcols = [], [], [], [], []
    raw = 0
    for name, v, bm in cols:
        dt = qa.FROM_NP[v.dtype]
        t = torch.from_numpy(v.view(np.uint8).copy()).to(dev)
        b = torch.from_numpy(bm.copy()).to(dev) if bm is not None else None
        cap = qa.lib.quoin_typed_compress_bound(int(dt), len(v))
        st = torch.empty(cap, dtype=torch.uint8, device=dev)
        dv.append((t, b))
        ds.append(st)
        do.append(torch.empty(v.nbytes, dtype=torch.uint8, device=dev))
        dvo.append(torch.empty(max((len(v) + 7) // 8, 1), dtype=torch.uint8, device=dev)
                   if bm is not None else None)
        dcols.append(qa.DeviceColumn(int(dt), 0, len(v), t.data_ptr(),
                                     b.data_ptr() if b is not None else None, st.data_ptr(), cap))
        raw += v.nbytes
    torch.cuda.synchronize()
    qa.reset_mode_win_counts()
    t0 = time.perf_counter()
    lens = gpu.compress(dcols, cfg)
    first_c = time.perf_counter() - t0
    wins = qa.mode_win_counts()
    host = [ds[i][:lens[i]].cpu().numpy() for i in range(len(cols))]
    dsts = [qa.DeviceStream(int(qa.FROM_NP[v.dtype]), 0, len(v), ds[i].data_ptr(), host[i].ctypes.data,
                            lens[i], do[i].data_ptr(), dvo[i].data_ptr() if dvo[i] is not None else None)
            for i, (_, v, _) in enumerate(cols)]
    gpu.decompress(dsts)
    torch.cuda.synchronize()
    for i, (name, v, bm) in enumerate(cols):
        assert do[i].cpu().numpy().tobytes() == v.view(np.uint8).tobytes(), f"round trip: {name}"
    checked = []
    if check_oracle:
        threads = host_cores()[0]
        for i, (name, v, bm) in enumerate(cols):
            exp = orc.compress(v, odt[v.dtype], bm, orc.MAX, threads=threads)
            assert host[i].tobytes() == exp, f"Config::default stream differs from the oracle: {name}"
            checked.append(name)
    dcols = (qa.DeviceColumn * len(dcols))(*dcols)
    dsts = (qa.DeviceStream * len(dsts))(*dsts)
    tc = td = 0.0
    for _ in range(steps):
        t0 = time.perf_counter()
        gpu.compress(dcols, cfg)
        t1 = time.perf_counter()
        gpu.decompress(dsts)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        tc += t1 - t0
        td += t2 - t1
    comp = int(sum(lens))
    blocks = int(sum(wins))
    dev_blocks = int(sum(wins[m] for m in DEVICE_EMITTED))
    per_col = []
    for i, (name, v, bm) in enumerate(cols):
        per_col.append({"column": name, "values": len(v), "bytes": int(v.nbytes),
                        "compressed": int(lens[i]), "ratio": round(v.nbytes / max(lens[i], 1), 3)})
    return {
        "workload": "Config::default (Level::Max, adaptive blocks): 1Mi f64 2-decimal walk (plain "
                    "method), the integer-lanes columns at 1Mi each (1% nulls on the u32 columns), "
                    "2Mi f64 sine+noise and 2Mi f32 uniform; one batched call per direction",
        "columns": per_col, "uncompressed_bytes": raw, "compressed_bytes": comp,
        "ratio": round(raw / comp, 3),
        "compress_gbs": round(raw * steps / tc / 1e9, 4),
        "decompress_gbs": round(raw * steps / td / 1e9, 4),
        "round_trip_gbs": round(raw * steps / (tc + td) / 1e9, 4),
        "first_compress_s": round(first_c, 3),
        "wins": {MODE_NAMES.get(m, str(m)): int(wins[m]) for m in range(64) if wins[m]},
        "routing": {"blocks": blocks, "device_emitted": dev_blocks, "host_written": blocks - dev_blocks,
                    "host_share": round((blocks - dev_blocks) / max(blocks, 1), 4)},
        "inputs": {"generator": GENERATOR, "checksums": column_checksums(cols)},
        "oracle_checked_columns": checked,
    }, [(v, odt[v.dtype], bm) for _, v, bm in cols]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=16 << 20)
    ap.add_argument("--level", default="fastest")
    ap.add_argument("--cpu-sample", type=int, default=4 << 20,
                    help="values per column of the 1-thread CPU figure")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=2,
                    help="steps of the plain form (one launch sequence per block); 0 skips it")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop leg")
    ap.add_argument("--no-entropy", action="store_true", help="skip the rANS decode leg")
    ap.add_argument("--entropy-n", type=int, default=32 << 20,
                    help="values per column of the rANS decode leg")
    ap.add_argument("--no-resid", action="store_true", help="skip the residual coder leg")
    ap.add_argument("--no-float", action="store_true", help="skip the mixed-float leg")
    ap.add_argument("--no-default", action="store_true",
                    help="skip the Config::default (Level::Max) leg")
    ap.add_argument("--default-steps", type=int, default=2)
    ap.add_argument("--resid-n", type=int, default=32 << 20,
                    help="f64 values of the residual coder leg")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="bytes of a plain elementwise copy run after the measurement "
                         "(for scaling HBM traffic counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU in this build; --gpus {args.gpus} is not supported "
              f"(blocks are independent: they would shard across GPUs by block index range)", file=sys.stderr)
        sys.exit(2)

    import torch
    import quoin_amd as qa

    level = {"fastest": qa.Level.FASTEST, "fast": qa.Level.FAST, "balanced": qa.Level.BALANCED,
             "high": qa.Level.HIGH, "max": qa.Level.MAX}[args.level]
    cfg = qa.Config(level=level)
    dev = torch.device("cuda:0")
    cols = make_workload(args.n)
    gpu = qa.Gpu(0)

    dtag = {np.dtype(np.int64): qa.DType.I64, np.dtype(np.uint32): qa.DType.U32}
    dvals, dbms, dstreams, douts, dvouts, dcols = [], [], [], [], [], []
    raw_bytes = 0
    for name, v, bm in cols:
        t = torch.from_numpy(v.view(np.uint8).copy()).to(dev)
        dvals.append(t)
        b = torch.from_numpy(bm.copy()).to(dev) if bm is not None else None
        dbms.append(b)
        cap = qa.lib.quoin_typed_compress_bound(int(dtag[v.dtype]), len(v))
        s = torch.empty(cap, dtype=torch.uint8, device=dev)
        dstreams.append(s)
        douts.append(torch.empty(v.nbytes, dtype=torch.uint8, device=dev))
        dvouts.append(torch.empty(max((len(v) + 7) // 8, 1), dtype=torch.uint8, device=dev)
                      if bm is not None else None)
        dcols.append(qa.DeviceColumn(int(dtag[v.dtype]), 0, len(v), t.data_ptr(),
                                     b.data_ptr() if b is not None else None, s.data_ptr(), cap))
        raw_bytes += v.nbytes
    torch.cuda.synchronize()

    # correctness on the first call (outside the timed region)
    lens = gpu.compress(dcols, cfg)
    host_streams = [dstreams[i][:lens[i]].cpu().numpy() for i in range(len(cols))]
    dsts = []
    for i, (name, v, bm) in enumerate(cols):
        dsts.append(qa.DeviceStream(int(dtag[v.dtype]), 0, len(v), dstreams[i].data_ptr(),
                                    host_streams[i].ctypes.data, lens[i], douts[i].data_ptr(),
                                    dvouts[i].data_ptr() if dvouts[i] is not None else None))
    gpu.decompress(dsts)
    torch.cuda.synchronize()
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as orc
    odt = {np.dtype(np.int64): orc.I64, np.dtype(np.uint32): orc.U32}
    for i, (name, v, bm) in enumerate(cols):
        # every stream byte for byte against the oracle at the same level
        exp = orc.compress(v, odt[v.dtype], bm, int(level), threads=host_cores()[0])
        assert host_streams[i].tobytes() == exp, f"stream differs from the oracle on {name}"
        got = douts[i].cpu().numpy()
        assert got.tobytes() == v.view(np.uint8).tobytes(), f"round trip mismatch on {name}"
        if bm is not None:
            vb = dvouts[i].cpu().numpy()[: len(bm)]
            exp = bm.copy()
            if len(v) % 8:
                exp[-1] &= (1 << (len(v) % 8)) - 1
            assert vb.tobytes() == exp.tobytes(), f"validity mismatch on {name}"
    comp_bytes = sum(lens)

    # the descriptor arrays are built once, as a C caller reusing them would
    dcols = (qa.DeviceColumn * len(dcols))(*dcols)
    dsts = (qa.DeviceStream * len(dsts))(*dsts)
    for _ in range(args.warmup):
        gpu.compress(dcols, cfg)
        gpu.decompress(dsts)
    torch.cuda.synchronize()
    tc = td = 0.0
    dev_c = dev_d = 0.0
    kern_c = kern_d = 0.0  # HIP events on the library's stream around its tile kernels
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        gpu.compress(dcols, cfg)
        t1 = time.perf_counter()
        tot, kms, _ = gpu.last_timing()
        dev_c += tot
        kern_c += kms
        gpu.decompress(dsts)
        t2 = time.perf_counter()
        tot, kms, _ = gpu.last_timing()
        dev_d += tot
        kern_d += kms
        tc += t1 - t0
        td += t2 - t1
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start

    # the plain form on the same columns: one launch sequence per block
    plain = None
    if args.plain_steps > 0:
        gpu.set_launch_form(1)
        gpu.compress(dcols, cfg)
        gpu.decompress(dsts)
        torch.cuda.synchronize()
        tp = time.perf_counter()
        for _ in range(args.plain_steps):
            gpu.compress(dcols, cfg)
            gpu.decompress(dsts)
        torch.cuda.synchronize()
        plain_wall = time.perf_counter() - tp
        gpu.set_launch_form(0)
        plain = raw_bytes * args.plain_steps / plain_wall / 1e9
    ms_step = wall / args.steps * 1e3
    value = raw_bytes * args.steps / wall / 1e9  # GB/s of uncompressed data, round trip
    c_gbs = raw_bytes * args.steps / tc / 1e9
    d_gbs = raw_bytes * args.steps / td / 1e9
    dev_c_ms = dev_c / args.steps
    dev_d_ms = dev_d / args.steps

    # Roofline of the dominant decode kernels: the tile decode stage
    # (k_dec_fl<4> for the FoR u32 columns, k_dec_fl<8> for the Delta i64
    # column, k_dec_tiles for other modes: none here), timed by the library's
    # HIP events around the stage on its own stream. Algorithmic bytes per call
    # = the stream bytes it reads (an upper bound of the frame payloads) + the
    # typed values it writes + the decoded validity words it reads for the
    # streams with nulls.
    bitmap_bytes = sum(((len(v) + 63) // 64) * 8 for _, v, bm in cols if bm is not None)
    dec_alg = comp_bytes + raw_bytes + bitmap_bytes
    kd_ms = kern_d / args.steps
    dec_gbs = dec_alg / (kd_ms * 1e-3) / 1e9 if kd_ms > 0 else 0.0
    roofline = {"bound": "hbm", "kernel": "tile decode: k_dec_fl<4> + k_dec_fl<8>",
                "achieved": round(dec_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(dec_gbs / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic("tile_decode", args), "algorithmic_bytes": dec_alg,
                "kernel_ms": round(kd_ms, 4),
                "decode_sequence_gbs": round((comp_bytes + raw_bytes) / (dev_d_ms * 1e-3) / 1e9, 1)}
    # the encode side's tile stage (k_emit_fl, and k_emit_tiles over its list:
    # empty here): reads the (compacted) lane, writes payloads
    kc_ms = kern_c / args.steps
    enc_alg = raw_bytes + comp_bytes
    enc_roofline = {"bound": "hbm", "kernel": "tile encode: k_emit_fl + k_emit_tiles",
                    "kernel_ms": round(kc_ms, 4),
                    "achieved": round(enc_alg / (kc_ms * 1e-3) / 1e9, 1) if kc_ms > 0 else 0.0,
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "algorithmic_bytes": enc_alg,
                    "traffic": measured_traffic("tile_encode", args)}
    enc_roofline["frac"] = round(enc_roofline["achieved"] / HBM_PEAK_GBS, 4)

    if args.calibrate_copy:
        # a plain copy of known size for scaling FETCH_SIZE / WRITE_SIZE counters
        src = torch.ones(args.calibrate_copy // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()

    serving = None
    if not args.no_serve:
        serving = serve_bench(qa, dev)

    entropy = ent_sample = None
    if not args.no_entropy:
        entropy, ent_sample = entropy_bench(qa, gpu, dev, args.entropy_n, 10, 2)

    resid = resid_sample = None
    if not args.no_resid:
        resid, resid_sample = resid_bench(qa, gpu, dev, args.resid_n, 10, 2)

    floatmix = float_cols = None
    if not args.no_float:
        floatmix, float_cols = float_bench(qa, gpu, dev, 5, 1)

    default = default_cols = None
    if not args.no_default:
        default, default_cols = default_bench(qa, gpu, dev, args.default_steps)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline([(v, odt[v.dtype], bm) for _, v, bm in cols], level, runs=5,
                           one_thread_values=args.cpu_sample, label="integer lanes: ")
        if default is not None:
            default["cpu_baseline"] = cpu_baseline(default_cols, qa.Level.MAX, runs=3,
                                                   one_thread_values=256 << 10,
                                                   label="Config::default leg: ")
        if entropy is not None:
            entropy["cpu_baseline"] = rans_cpu_baseline(*ent_sample)
        if resid is not None:
            resid["cpu_baseline"] = resid_cpu_baseline(resid_sample)
        if floatmix is not None:
            floatmix["cpu_baseline"] = float_cpu_baseline(float_cols)

    out = {
        "metric": "compress+decompress GB/s (bit-exact), batched",
        "value": round(value, 3), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64 lanes (i64/u32 columns)",
        "data": "synthetic",
        "config": {"workload": "integer lanes: 16Mi i64 timestamps + 3x16Mi u32 k=3/11/23, "
                               "1% nulls on the u32 columns, one batched launch sequence",
                   "level": args.level, "values_per_column": args.n,
                   "uncompressed_bytes": raw_bytes, "compressed_bytes": comp_bytes,
                   "compress_gbs": round(c_gbs, 2), "decompress_gbs": round(d_gbs, 2),
                   "device_ms_compress": round(dev_c_ms, 3),
                   "device_ms_decompress": round(dev_d_ms, 3),
                   "plain_form_gbs": round(plain, 3) if plain else None,
                   "batched_over_plain": round(value / plain, 2) if plain else None,
                   "waves_per_workgroup": gpu.tuned_shapes(),
                   "inputs": {"generator": GENERATOR, "checksums": column_checksums(cols)}},
        "roofline": roofline,
        "encode_roofline": enc_roofline,
        "serving": serving,
        "entropy_decode": entropy,
        "residual_coders": resid,
        "mixed_float": floatmix,
        "config_default": default,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    gpu.close()


if __name__ == "__main__":
    main()
