"""quoin_gpu_linear_encode_noalloc (the linear residual encoder with its
scratch in a fixed device array, for callers that cannot allocate) against
quoin_gpu_linear_encode, through the C ABI: the same streams, lengths, orders
and statuses on two calls in a row, the second with a smaller input, and the
size limit refused before any device work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a GPU", allow_module_level=True)

import quoin_amd as qa  # noqa: E402

DEV = torch.device("cuda:0")
GUARD = 0xA5
KINDS = (qa.LINEAR_ORDERED_DELTA, qa.LINEAR_DELTA2, qa.LINEAR_DELTA_DP)


@pytest.fixture(scope="module")
def gpu():
    g = qa.Gpu(0)
    yield g
    g.close()


def lanes_of(sizes, seed):
    r = np.random.default_rng(seed)
    out = []
    for i, n in enumerate(sizes):
        if i % 3 == 0:
            x = np.sin(np.arange(n) / 40.0) * 1000 + r.normal(0, 1e-3, n)
        elif i % 3 == 1:
            x = np.round(np.cumsum(r.normal(size=n)), 2)
        else:
            x = r.integers(-(1 << 40), 1 << 40, n).astype(np.float64)
        out.append(np.ascontiguousarray(x.view(np.uint64)))
    return out


def encode(gpu, fn, kind, lanes, order):
    keep, ins = [], []
    for lane in lanes:
        v = torch.from_numpy(lane.view(np.int64).copy()).to(DEV) if len(lane) else \
            torch.zeros(1, dtype=torch.int64, device=DEV)
        cap = int(qa.lib.quoin_linear_bound(len(lane)))
        o = torch.full((cap + 32,), GUARD, dtype=torch.uint8, device=DEV)
        keep.append((v, o))
        ins.append(qa.LinearInput(v.data_ptr(), len(lane), o.data_ptr(), cap))
    lens, ords, sts = fn(kind, ins, order)
    torch.cuda.synchronize()
    streams = []
    for (v, o), ln, st in zip(keep, lens, sts):
        if st:  # DeltaDp declined the block (the reference's None): no stream
            assert kind == qa.LINEAR_DELTA_DP and ln == 0
            streams.append(None)
            continue
        h = o.cpu().numpy()
        assert (h[ln:] == GUARD).all(), "write past the stream"
        streams.append(h[:ln].tobytes())
    return streams, lens, ords, sts


@pytest.mark.parametrize("kind", KINDS)
def test_noalloc_matches_the_allocating_entry_point(gpu, kind):
    calls = [lanes_of([70000, 1, 0, 4096, 2049, 300000], seed=1),
             lanes_of([5000, 3, 2048], seed=2)]  # the second call is smaller
    for order in (0, 2):
        for i, lanes in enumerate(calls):
            print(f"case noalloc kind={kind} order={order} call={i} "
                  f"n={[len(x) for x in lanes]}", flush=True)
            want = encode(gpu, gpu.linear_encode, kind, lanes, order)
            got = encode(gpu, gpu.linear_encode_noalloc, kind, lanes, order)
            assert got == want


def test_noalloc_refuses_calls_beyond_its_scratch(gpu):
    # checked on the host before any device work: no buffer is touched
    ins = [qa.LinearInput(None, 0, None, 0)] * 1025
    with pytest.raises(qa.QuoinError) as e:
        gpu.linear_encode_noalloc(qa.LINEAR_ORDERED_DELTA, ins)
    assert e.value.code == qa.ERR_BUFFER_TOO_SMALL
    big = 65536 * 2048 + 1  # one value past 65536 tiles; the pointers are never read
    ins = [qa.LinearInput(1 << 20, big, 1 << 20, int(qa.lib.quoin_linear_bound(big)))]
    with pytest.raises(qa.QuoinError) as e:
        gpu.linear_encode_noalloc(qa.LINEAR_ORDERED_DELTA, ins)
    assert e.value.code == qa.ERR_BUFFER_TOO_SMALL
