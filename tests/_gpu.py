"""Device-resident round trips through the quoin_gpu_* C ABI (torch tensors
hold the device buffers)."""
import numpy as np


def device_compress(qa, gpu, cols, cfg):
    """cols: list of (values ndarray, validity bitmap or None). Returns the
    host stream bytes of each column."""
    import torch
    dev = torch.device("cuda:0")
    keep, dcols = [], []
    for v, bm in cols:
        dt = qa.FROM_NP[v.dtype]
        t = torch.from_numpy(v.view(np.uint8).copy()).to(dev) if v.size else \
            torch.empty(1, dtype=torch.uint8, device=dev)
        b = torch.from_numpy(bm.copy()).to(dev) if bm is not None else None
        cap = qa.lib.quoin_typed_compress_bound(int(dt), len(v))
        s = torch.empty(cap, dtype=torch.uint8, device=dev)
        keep += [t, b, s]
        dcols.append(qa.DeviceColumn(int(dt), 0, len(v), t.data_ptr(),
                                     b.data_ptr() if b is not None else None, s.data_ptr(), cap))
    lens = gpu.compress(dcols, cfg)
    torch.cuda.synchronize()
    return [keep[3 * i + 2][:lens[i]].cpu().numpy().tobytes() for i in range(len(cols))]


def device_decompress(qa, gpu, streams, dtypes, ns, with_validity):
    """Returns [(values ndarray, validity bytes or None)] per stream."""
    import torch
    dev = torch.device("cuda:0")
    keep, dst, hosts = [], [], []
    for s, dt, n in zip(streams, dtypes, ns):
        h = np.frombuffer(s, dtype=np.uint8).copy()
        hosts.append(h)
        d = torch.from_numpy(h.copy()).to(dev)
        w = np.dtype(qa.NP[dt]).itemsize
        o = torch.empty(max(n * w, 1), dtype=torch.uint8, device=dev)
        vo = torch.empty(max((n + 7) // 8, 1), dtype=torch.uint8, device=dev) \
            if with_validity else None
        keep += [d, o, vo]
        dst.append(qa.DeviceStream(int(dt), 0, n, d.data_ptr(), h.ctypes.data, len(s),
                                   o.data_ptr(), vo.data_ptr() if vo is not None else None))
    gpu.decompress(dst)
    torch.cuda.synchronize()
    out = []
    for i, (dt, n) in enumerate(zip(dtypes, ns)):
        w = np.dtype(qa.NP[dt]).itemsize
        vals = keep[3 * i + 1][:n * w].cpu().numpy().view(qa.NP[dt])
        vo = keep[3 * i + 2]
        out.append((vals, vo[:(n + 7) // 8].cpu().numpy().tobytes() if vo is not None else None))
    return out
