"""Serving loop (quoin_server_*): each job's result is the same alone and in a
continuous-batching mix, admission by the HBM budget serializes without
changing results, and a job larger than the budget is refused."""
import numpy as np
import pytest

import _data
import _gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qa():
    import quoin_amd
    return quoin_amd


def columns(seed):
    r = np.random.default_rng(seed)
    out = []
    for i, n in enumerate([1, 700, 5000, 40000, 131072 + 11, 250000]):
        ic = _data.int_columns(n, seed + i)
        fc = _data.float_columns(n, seed + i)
        pick = [(ic["ts_i64"], None), _data.with_nulls(ic["u32_k11"], 0.05, seed + i),
                (ic["i32_mixed"], None), (fc["walk_f64"], None), (fc["unif_f32"], None),
                (fc["unif_f64"], None)]
        out.append(pick[(seed + i) % len(pick)])
        out.append(pick[(seed + 3 * i + 1) % len(pick)])
    r.shuffle(out)
    return out


class Bufs:
    """Device buffers of one job (input column, output stream, decode output)."""

    def __init__(self, qa, v, bm):
        import torch
        dev = torch.device("cuda:0")
        self.v, self.bm = v, bm
        self.dt = qa.FROM_NP[v.dtype]
        self.t = torch.from_numpy(v.view(np.uint8).copy()).to(dev) if v.size else \
            torch.empty(1, dtype=torch.uint8, device=dev)
        self.b = torch.from_numpy(bm.copy()).to(dev) if bm is not None else None
        self.cap = qa.lib.quoin_typed_compress_bound(int(self.dt), len(v))
        self.s = torch.zeros(self.cap, dtype=torch.uint8, device=dev)
        w = np.dtype(qa.NP[self.dt]).itemsize
        self.o = torch.zeros(max(len(v) * w, 1), dtype=torch.uint8, device=dev)
        self.vo = torch.zeros(max((len(v) + 7) // 8, 1), dtype=torch.uint8, device=dev)

    def column(self, qa):
        return qa.DeviceColumn(int(self.dt), 0, len(self.v), self.t.data_ptr(),
                               self.b.data_ptr() if self.b is not None else None,
                               self.s.data_ptr(), self.cap)

    def stream(self, qa, host):
        self.host = np.frombuffer(host, dtype=np.uint8).copy()
        import torch
        self.ds = torch.from_numpy(self.host.copy()).to(torch.device("cuda:0"))
        return qa.DeviceStream(int(self.dt), 0, len(self.v), self.ds.data_ptr(),
                               self.host.ctypes.data, len(host), self.o.data_ptr(),
                               self.vo.data_ptr())


def alone(qa, cols, cfgs):
    gpu = qa.Gpu(0)
    try:
        streams = [_gpu.device_compress(qa, gpu, [c], cfg)[0] for c, cfg in zip(cols, cfgs)]
        decoded = [_gpu.device_decompress(qa, gpu, [s], [qa.FROM_NP[v.dtype]], [len(v)], True)[0]
                   for s, (v, _) in zip(streams, cols)]
    finally:
        gpu.close()
    return streams, decoded


def served(qa, cols, cfgs, streams, budget=0, slots=2):
    """Submit every compress job and every decompress job (of the given
    streams) in one burst; returns the server's outputs and its stats."""
    import torch
    bufs = [Bufs(qa, v, bm) for v, bm in cols]
    srv = qa.Server(0, budget, slots)
    try:
        tickets = []
        for i, (b, cfg) in enumerate(zip(bufs, cfgs)):
            tickets.append(("c", i, srv.submit_compress(b.column(qa), cfg)))
            tickets.append(("d", i, srv.submit_decompress(b.stream(qa, streams[i]))))
        lens = {}
        for kind, i, t in tickets:
            rc, n, ms = srv.wait(t)
            assert rc == qa.OK, f"job {kind}{i}: rc {rc}"
            assert ms >= 0
            if kind == "c":
                lens[i] = n
        torch.cuda.synchronize()
        jobs, batches = srv.stats()
    finally:
        srv.close()
    out_s = [bufs[i].s[:lens[i]].cpu().numpy().tobytes() for i in range(len(bufs))]
    out_d = []
    for b in bufs:
        w = np.dtype(qa.NP[b.dt]).itemsize
        out_d.append((b.o[:len(b.v) * w].cpu().numpy().view(qa.NP[b.dt]),
                      b.vo[:(len(b.v) + 7) // 8].cpu().numpy().tobytes()))
    return out_s, out_d, jobs, batches


def compare(cols, ref_s, ref_d, got_s, got_d):
    for i, (v, bm) in enumerate(cols):
        assert got_s[i] == ref_s[i], f"job {i} ({v.dtype}, n={len(v)}): stream differs in the mix"
        assert got_d[i][0].view(np.uint8).tobytes() == ref_d[i][0].view(np.uint8).tobytes(), \
            f"job {i}: decode differs in the mix"
        assert got_d[i][1] == ref_d[i][1], f"job {i}: validity differs in the mix"


def test_results_same_alone_and_in_the_mix(qa):
    cols = columns(7)
    cfgs = [qa.Config(qa.Level(i % 2)) for i in range(len(cols))]
    print(f"case serve mix jobs={2 * len(cols)}", flush=True)
    ref_s, ref_d = alone(qa, cols, cfgs)
    got_s, got_d, jobs, batches = served(qa, cols, cfgs, ref_s)
    compare(cols, ref_s, ref_d, got_s, got_d)
    assert jobs == 2 * len(cols)
    assert batches < jobs, "jobs that waited together should share launches"


def test_tight_budget_serializes_without_changing_results(qa):
    cols = columns(11)[:6]
    cfgs = [qa.Config(qa.Level.FAST)] * len(cols)
    ref_s, ref_d = alone(qa, cols, cfgs)
    # each slot's share (budget / slots) has room for the largest single job only
    big = max(len(v) * v.dtype.itemsize * 2 + 1024 + len(v) * (v.dtype.itemsize + 100)
              for v, _ in cols)
    print(f"case serve budget={3 * big}", flush=True)
    got_s, got_d, jobs, batches = served(qa, cols, cfgs, ref_s, budget=3 * big, slots=3)
    compare(cols, ref_s, ref_d, got_s, got_d)


def test_job_over_budget_is_refused(qa):
    v = np.arange(100000, dtype=np.int64)
    b = Bufs(qa, v, None)
    srv = qa.Server(0, 1 << 20, 1)
    try:
        with pytest.raises(qa.QuoinError) as e:
            srv.submit_compress(b.column(qa), qa.Config(qa.Level.FAST))
        assert e.value.code == qa.ERR_BUFFER_TOO_SMALL
    finally:
        srv.close()


def test_submit_and_wait_on_two_threads_match_the_oracle(qa):
    """One thread submits while another waits (ctypes releases the GIL for
    both), with Fast and Max jobs mixed; every served stream is the oracle's
    and every served decode is the oracle's decode. A second wait on a ticket
    already waited for is refused."""
    import queue
    import threading

    import torch

    import _oracle as orc
    odt = {np.dtype(np.float64): orc.F64, np.dtype(np.int64): orc.I64,
           np.dtype(np.uint64): orc.U64, np.dtype(np.int32): orc.I32,
           np.dtype(np.uint32): orc.U32, np.dtype(np.float32): orc.F32}
    cols = columns(23)[:10]
    levels = [qa.Level.MAX if i % 2 else qa.Level.FAST for i in range(len(cols))]
    print(f"case serve two threads jobs={len(cols)}", flush=True)
    bufs = [Bufs(qa, v, bm) for v, bm in cols]
    srv = qa.Server(0, 0, 2)
    q, results, errors = queue.Queue(), {}, []

    def waiter():
        try:
            while True:
                item = q.get()
                if item is None:
                    return
                i, t = item
                results[i] = srv.wait(t)
        except Exception as e:  # surfaced below
            errors.append(e)

    th = threading.Thread(target=waiter)
    th.start()
    try:
        tickets = []
        for i, b in enumerate(bufs):
            t = srv.submit_compress(b.column(qa), qa.Config(levels[i]))
            tickets.append(t)
            q.put((i, t))
        q.put(None)
        th.join(timeout=240)
        assert not th.is_alive() and not errors, errors
        torch.cuda.synchronize()
        assert srv.wait(tickets[0])[0] == qa.ERR_NULL  # claimed by the waiter already
    finally:
        srv.close()
    for i, (v, bm) in enumerate(cols):
        rc, n, _ = results[i]
        assert rc == qa.OK, (i, rc)
        got = bufs[i].s[:n].cpu().numpy().tobytes()
        want = orc.compress(v, odt[v.dtype], bm, int(levels[i]), threads=16)
        assert got == want, f"job {i} (level {int(levels[i])}): stream differs from the oracle"


def test_small_requests_do_not_wait_behind_large_ones(qa):
    """With two or more slots, slot 0 takes only requests of up to 1 MiB of
    values: a small job submitted behind a burst of large ones completes
    first, and every result is still the same as alone."""
    r = np.random.default_rng(8)
    large = [(r.integers(0, 1 << 40, 4 << 20).astype(np.int64), None) for _ in range(6)]
    small = [(r.integers(0, 1000, 2000).astype(np.int64), None)]
    cols = large + small
    cfgs = [qa.Config(qa.Level.FAST)] * len(cols)
    print("case serve size classes", flush=True)
    ref_s, _ = alone(qa, cols, cfgs)
    bufs = [Bufs(qa, v, bm) for v, bm in cols]
    srv = qa.Server(0, 0, 2)
    try:
        tickets = [srv.submit_compress(b.column(qa), cfg) for b, cfg in zip(bufs, cfgs)]
        res = [srv.wait(t) for t in tickets]
        # the small job alone on the same (warm) server, best of three
        alone_ms = min(srv.wait(srv.submit_compress(bufs[-1].column(qa), cfgs[-1]))[2]
                       for _ in range(3))
    finally:
        srv.close()
    for i, (rc, n, ms) in enumerate(res):
        assert rc == qa.OK, (i, rc)
        assert bufs[i].s[:n].cpu().numpy().tobytes() == ref_s[i], f"job {i} differs from alone"
    # Queued behind the burst, the small job would finish about its own time
    # after the large ones; on its own slot it runs beside them (their kernels
    # share the GPU, so it may still end a little after them).
    small_ms, large_ms = res[-1][2], max(ms for _, _, ms in res[:-1])
    assert small_ms < large_ms + 0.5 * alone_ms, (small_ms, alone_ms, [ms for _, _, ms in res])
