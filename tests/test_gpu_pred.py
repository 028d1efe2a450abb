"""GPU parity of the FCM / DFCM predictors (src/codecs/pred.rs) against the
oracle: residual streams byte for byte from quoin_gpu_pred_encode (context
keys radix-sorted on the device), values and error outcomes from
quoin_gpu_pred_decode (the table recurrence on the device)."""
import numpy as np
import pytest

import _oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a GPU", allow_module_level=True)

import quoin_amd as qa  # noqa: E402

DEV = torch.device("cuda:0")


@pytest.fixture(scope="module")
def gpu():
    g = qa.Gpu(0)
    yield g
    g.close()


def blocks():
    r = np.random.default_rng(41)
    walk = np.round(100 * (1000 + np.cumsum(r.normal(size=40000)))) / 100
    return [
        np.array([], dtype=np.uint64),
        np.array([7], dtype=np.uint64),
        np.array([5, 5, 5], dtype=np.uint64),
        np.array([10, 20, 30, 40], dtype=np.uint64),
        r.integers(0, 16, 2049, dtype=np.uint64),                 # crosses an encode tile
        walk.view(np.uint64),                                     # f64 bits of a random walk
        (np.arange(32768) * 3 + 1_700_000_000_000).astype(np.uint64),   # a ramp: DFCM territory
        np.repeat(r.integers(0, 1 << 60, 700, dtype=np.uint64), 47)[:32768],
        r.integers(0, 1 << 64, 5000, dtype=np.uint64),
    ]


def encode(gpu, kind, bl, log2):
    keep, ins = [], []
    for v in bl:
        t = torch.from_numpy(v.copy()).to(DEV) if len(v) else torch.zeros(1, dtype=torch.int64, device=DEV)
        cap = int(qa.lib.quoin_linear_bound(len(v)))
        o = torch.zeros(cap, dtype=torch.uint8, device=DEV)
        keep.append((t, o))
        ins.append(qa.LinearInput(t.data_ptr(), len(v), o.data_ptr(), cap))
    lens = gpu.pred_encode(kind, ins, log2)
    torch.cuda.synchronize()
    return [o[:ln].cpu().numpy().tobytes() for (_, o), ln in zip(keep, lens)]


def decode(gpu, kind, payloads, ns, log2):
    keep, sts = [], []
    for pl, n in zip(payloads, ns):
        t = torch.from_numpy(np.frombuffer(pl, dtype=np.uint8).copy()).to(DEV) if pl else \
            torch.zeros(1, dtype=torch.uint8, device=DEV)
        o = torch.zeros(max(n, 1), dtype=torch.int64, device=DEV)
        keep.append((t, o))
        sts.append(qa.LinearStream(t.data_ptr(), len(pl), n, o.data_ptr()))
    st = gpu.pred_decode(kind, sts, log2)
    torch.cuda.synchronize()
    return st, [o[:n].cpu().numpy().view(np.uint64) for (_, o), n in zip(keep, ns)]


@pytest.mark.parametrize("log2", [16, 10, 3])
@pytest.mark.parametrize("kind", [qa.PRED_FCM, qa.PRED_DFCM])
def test_pred_encode_decode_match_oracle(gpu, kind, log2):
    bl = blocks()
    print(f"case pred kind={kind} log2={log2} blocks={len(bl)}", flush=True)
    got = encode(gpu, kind, bl, log2)
    for i, (v, g) in enumerate(zip(bl, got)):
        assert g == orc.pred_encode(kind, v, log2), f"block {i} (n={len(v)})"
    st, back = decode(gpu, kind, got, [len(v) for v in bl], log2)
    for i, (v, s, b) in enumerate(zip(bl, st, back)):
        assert s == qa.OK and b.tolist() == v.tolist(), f"block {i}"


def test_pred_decode_errors_match_oracle(gpu):
    v = blocks()[5]
    for kind in (qa.PRED_FCM, qa.PRED_DFCM):
        e = orc.pred_encode(kind, v, 16)
        cases = [(e, len(v)), (e + b"\x00", len(v)), (e[:-1], len(v)), (e, len(v) + 1),
                 (b"\x80" * 10 + b"\x01", 1), (b"", 0), (b"", 1)]
        print(f"case pred errors kind={kind}", flush=True)
        st, back = decode(gpu, kind, [c[0] for c in cases], [c[1] for c in cases], 16)
        for i, ((pl, n), s, b) in enumerate(zip(cases, st, back)):
            rc, want = orc.pred_decode(kind, pl, n, 16)
            assert (s == qa.OK) == (rc == 0), (kind, i, s, rc)
            if rc == 0:
                assert b.tolist() == want.tolist(), (kind, i)
