"""Benchmark of the MI355X quoin block path on the integer-lanes workload.

Workload (one GPU): 16 Mi i64 monotone timestamps (start 1.7e18, deltas uniform
[1, 1000]) plus three 16 Mi u32 columns uniform in [0, 2^k), k in {3, 11, 23},
each with a Bernoulli 1 % null bitmap. All blocks of all four columns go through
one batched launch sequence. A step = compress the four columns + decompress the
four streams, device-resident (inputs already in HBM). The first step is checked
bit for bit (stream bytes against the CPU oracle on the first step's sample,
round trip in full).

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


def make_workload(n, seed=0):
    r = np.random.default_rng(seed)
    ts = (1_700_000_000_000_000_000 + np.cumsum(r.integers(1, 1001, n))).astype(np.int64)
    cols = [("ts_i64", ts, None)]
    for k in (3, 11, 23):
        v = (r.integers(0, 2**32, n, dtype=np.uint64) >> np.uint64(32 - k)).astype(np.uint32)
        valid = r.random(n) >= 0.01
        v[~valid] = 0
        bm = np.packbits(valid.astype(np.uint8), bitorder="little")
        cols.append((f"u32_k{k}", v, bm))
    return cols


def measured_traffic(stage, args):
    """HBM bytes per call of a tile stage on this workload, from the committed
    counter summary (profiles/traffic_<workload>.json: rocprofv3 FETCH_SIZE
    and WRITE_SIZE passes, scaled by the same counters on a plain copy of
    known size), or None when no summary matches this configuration."""
    path = os.path.join(ROOT, "profiles", f"traffic_intlanes_{args.level}.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    st = t.get("stages", {}).get(stage)
    if t.get("values_per_column") != args.n or not st:
        return None
    return st.get("hbm_bytes_per_call")


def cpu_baseline(cols, level, sample_n):
    """Oracle (single-threaded C restatement) over a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as orc
    dmap = {np.dtype(np.int64): orc.I64, np.dtype(np.uint32): orc.U32}
    raw = 0
    t0 = time.perf_counter()
    for name, v, bm in cols:
        vs = v[:sample_n]
        bms = bm[: (sample_n + 7) // 8] if bm is not None else None
        s = orc.compress(vs, dmap[v.dtype], bms, level)
        orc.decompress(s)
        raw += vs.nbytes
    dt = time.perf_counter() - t0
    return raw / dt / 1e9, raw


def serve_pool(qa, gpu, cfg, dev, count, seed=1):
    """Requests of the serving workload: uncompressed sizes log-uniform in
    [4 KB, 64 MB], types f64/f32/i64/i32 in turn with the generators of the
    other workloads (f64: 2-decimal walk, uniform doubles, 256-value runs;
    f32 uniform [-1, 1); i64 timestamps; i32 uniform k-bit), half compress and
    half decompress (of a stream made beforehand). Device buffers per request."""
    import torch
    r = np.random.default_rng(seed)
    pool = []
    for i in range(count):
        nbytes = int(np.exp(r.uniform(np.log(4 << 10), np.log(64 << 20))))
        kind = ["f64", "f32", "i64", "i32"][i % 4]
        w = 4 if kind in ("f32", "i32") else 8
        n = max(nbytes // w, 1)
        if kind == "f64":
            g = i // 4 % 3
            if g == 0:
                v = np.round(100 * (1000 + np.cumsum(r.normal(size=n)))) / 100
            elif g == 1:
                v = r.integers(0, 2**53, n, dtype=np.uint64) * 2.0**-53
            else:
                runs = r.geometric(1 / 8, n // 4 + 2)
                v = np.repeat(r.integers(0, 256, len(runs)) * 0.5, runs)[:n]
                if len(v) < n:
                    v = np.resize(v, n)
        elif kind == "f32":
            v = ((r.integers(0, 2**24, n) * 2.0**-23) - 1).astype(np.float32)
        elif kind == "i64":
            v = (1_700_000_000_000_000_000 + np.cumsum(r.integers(1, 1001, n))).astype(np.int64)
        else:
            k = [3, 11, 23][i // 4 % 3]
            v = (r.integers(0, 2**32, n, dtype=np.uint64) >> np.uint64(32 - k)).astype(np.int32)
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
    r = np.random.default_rng(seed)
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
            lat.append(ms)
            req["busy"] = False

    th = threading.Thread(target=collect)
    th.start()
    done_bytes = 0
    t0 = time.perf_counter()
    next_t = t0
    k = 0
    for _ in range(count):
        if rate:
            next_t += r.exponential(1.0 / rate)
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


def serve_bench(qa, dev, budget_gb=64, slots=2, pool_size=48, per_point=800):
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
            gbs, lat, wall = serve_run(qa, srv, pool, cfg, frac * sat_rps, per_point, seed=int(frac * 100))
            j1, b1 = srv.stats()
            points.append({"offered_rps": round(frac * sat_rps, 1), "achieved_gbs": round(gbs, 3),
                           "p50_ms": round(float(np.percentile(lat, 50)), 3),
                           "p99_ms": round(float(np.percentile(lat, 99)), 3),
                           "jobs": j1 - j0, "batches": b1 - b0})
    finally:
        srv.close()
    hi = points[-1]
    return {"workload": f"{pool_size} requests, 4 KB-64 MB log-uniform, f64/f32/i64/i32, "
                        f"half compress half decompress, Level::Fast, Poisson arrivals",
            "hbm_budget_gb": budget_gb, "slots": slots,
            "saturation_gbs": round(sat_gbs, 3), "saturation_rps": round(sat_rps, 1),
            "p50_ms": hi["p50_ms"], "p99_ms": hi["p99_ms"], "at_offered_rps": hi["offered_rps"],
            "points": points}


def entropy_lanes(n, seed=3):
    """The entropy workload's two columns as f64 lanes: n f64 values
    sin(2 pi i / 1000) * 1e3 + N(0, 1e-3) and n f32 values ((u32 >> 8) 2^-23) - 1
    widened to f64."""
    r = np.random.default_rng(seed)
    i = np.arange(n, dtype=np.float64)
    f64 = np.sin(2 * np.pi * i / 1000) * 1000 + r.normal(0, 1e-3, n)
    u = r.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    f32 = (u >> 8).astype(np.float32) * np.float32(2.0**-23) - np.float32(1)
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


def resid_bench(qa, gpu, dev, n, steps, warmup, block=32768, seed=3):
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


def entropy_bench(qa, gpu, dev, n, steps, warmup, block=32768, seed=3):
    """rANS decode on the entropy workload (entropy_lanes): each 32768-value
    block's Delta2 residual bytes, from the device residual coder, are rANS-coded
    by the GPU coder, then every block stream is decoded in one launch; decoded
    bytes are checked against the residuals."""
    import torch
    res, bspans = [], []
    for lane in entropy_lanes(n, seed):
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
        "streams": len(spans), "symbols": nsym, "coded_bytes": cbytes,
        "decode_gbs": round(nsym / (k_ms * 1e-3) / 1e9, 2),
        "decode_wall_gbs": round(nsym * steps / wall / 1e9, 2),
        "encode_gbs": round(nsym / enc_s / 1e9, 2),
        "roofline": {"bound": "hbm", "kernel": "k_rans_dec", "kernel_ms": round(k_ms, 4),
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


MODE_NAMES = {1: "const", 2: "stride", 3: "xorz", 5: "raw", 9: "alp", 17: "float_mult",
              19: "dict", 50: "for_bitpack", 51: "delta_bitpack", 52: "alp_rd", 53: "rle"}


def float_columns(total=64 << 20, ncols=40, seed=4):
    """The mixed-float workload: `ncols` f64 columns whose lengths are
    exp(U[ln 65536, ln 8388608]) rescaled to sum to `total` and kept in that
    range. Column i is, by i % 4: 0, 1 decimals k / 1000 with k uniform in
    [0, 10^9) (scale 10^3 in [0, 1e6)); 2 full-precision uniform [0, 1) doubles
    (u64 >> 11) 2^-53; 3 a 256-value dictionary of such decimals, written in
    runs of Geometric(1/8) length (mean 8), each run's value uniform over the
    256."""
    r = np.random.default_rng(seed)
    lo, hi = 65536, 8388608
    L = np.exp(r.uniform(np.log(lo), np.log(hi), ncols))
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
        if kind == "decimal":
            v = r.integers(0, 10**9, n) / 1000.0
        elif kind == "full":
            v = (r.integers(0, 2**64, n, dtype=np.uint64) >> np.uint64(11)) * 2.0**-53
        else:
            vals = r.integers(0, 10**9, 256) / 1000.0
            runs = r.geometric(1 / 8, n // 4 + 64)
            while runs.sum() < n:
                runs = np.concatenate([runs, r.geometric(1 / 8, n // 8 + 64)])
            v = np.repeat(vals[r.integers(0, 256, len(runs))], runs)[:n]
        cols.append((kind, np.ascontiguousarray(v, dtype=np.float64)))
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
        for kind in ("decimal", "full", "dict"):
            i = min((j for j in range(len(cols)) if cols[j][0] == kind), key=lambda j: len(cols[j][1]))
            exp = orc.compress(cols[i][1], orc.F64, None, orc.FAST)
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
        "oracle_checked_columns": checked,
        "roofline": {"bound": "hbm", "kernel": "tile decode: k_dec_tiles + k_dec_fl",
                     "kernel_ms": round(kd, 4), "algorithmic_bytes": alg,
                     "achieved": round(alg / (kd * 1e-3) / 1e9, 1) if kd > 0 else 0.0,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(alg / (kd * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if kd > 0 else 0.0,
                     "decode_sequence_gbs": round(alg / (dd * 1e-3) / 1e9, 1)},
        "encode_roofline": {"bound": "hbm", "kernel": "tile encode: k_emit_fl + k_emit_tiles",
                            "kernel_ms": round(kc, 4),
                            "achieved": round(alg / (kc * 1e-3) / 1e9, 1) if kc > 0 else 0.0,
                            "encode_sequence_gbs": round(alg / (dc * 1e-3) / 1e9, 1)},
    }, cols


def float_cpu_baseline(cols, budget_values=1 << 20):
    """Oracle compress + decompress at Level::Fast (single-threaded C) over the
    first `budget_values` values of each kind's shortest column."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as orc
    sample = []
    for kind in ("decimal", "full", "dict"):
        i = min((j for j in range(len(cols)) if cols[j][0] == kind), key=lambda j: len(cols[j][1]))
        sample.append(cols[i][1][:budget_values])
    t0 = time.perf_counter()
    for v in sample:
        orc.decompress(orc.compress(v, orc.F64, None, orc.FAST))
    dt = time.perf_counter() - t0
    rawb = sum(v.nbytes for v in sample)
    return {"value": round(rawb / dt / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"the shortest column of each kind ({rawb / 1e6:.1f} MB), compress+decompress "
                      "at Level::Fast, oracle C, 1 thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=16 << 20)
    ap.add_argument("--level", default="fastest")
    ap.add_argument("--cpu-sample", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=2,
                    help="steps of the plain form (one launch sequence per block); 0 skips it")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop leg")
    ap.add_argument("--no-entropy", action="store_true", help="skip the rANS decode leg")
    ap.add_argument("--entropy-n", type=int, default=32 << 20,
                    help="values per column of the rANS decode leg")
    ap.add_argument("--no-resid", action="store_true", help="skip the residual coder leg")
    ap.add_argument("--no-float", action="store_true", help="skip the mixed-float leg")
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

    level = {"fastest": qa.Level.FASTEST, "fast": qa.Level.FAST}[args.level]
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
    for i, (name, v, bm) in enumerate(cols):
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

    cpu = None
    if not args.no_cpu:
        cpu_gbs, sample_raw = cpu_baseline(cols, int(level), args.cpu_sample)
        cpu = {"value": round(cpu_gbs, 4), "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": f"first {args.cpu_sample} values of each of the 4 columns "
                         f"({sample_raw / 1e6:.1f} MB), compress+decompress, oracle C, 1 thread"}
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
                   "waves_per_workgroup": gpu.tuned_shapes()},
        "roofline": roofline,
        "encode_roofline": enc_roofline,
        "serving": serving,
        "entropy_decode": entropy,
        "residual_coders": resid,
        "mixed_float": floatmix,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    gpu.close()


if __name__ == "__main__":
    main()
