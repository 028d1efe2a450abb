"""GPU parity of the rANS coder (src/entropy/rans.rs) against the oracle:
coded streams byte for byte, decoded bytes, and the error/success outcome of
corrupt and truncated payloads, through quoin_gpu_rans_compress /
quoin_gpu_rans_decompress."""
import struct

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


def _dev(b, off=0):
    """device copy of bytes b starting `off` bytes into a fresh allocation"""
    t = torch.zeros(len(b) + off + 16, dtype=torch.uint8, device=DEV)
    if len(b):
        t[off:off + len(b)] = torch.from_numpy(np.frombuffer(bytes(b), dtype=np.uint8).copy()).to(DEV)
    return t, t.data_ptr() + off


def gpu_compress(gpu, datas):
    keep, ins = [], []
    for d in datas:
        t, p = _dev(d)
        cap = max(qa.lib.quoin_rans_bound(len(d)), 1)
        o = torch.zeros(cap, dtype=torch.uint8, device=DEV)
        keep.append((t, o))
        ins.append(qa.RansInput(p, len(d), o.data_ptr(), cap))
    lens = gpu.rans_compress(ins)
    torch.cuda.synchronize()
    out = []
    for (t, o), ln in zip(keep, lens):
        out.append(o[:ln].cpu().numpy().tobytes() if ln else None)
    return out


def gpu_decompress(gpu, payloads, max_lens, tagged=False, src_off=0, out_off=0):
    keep, sts = [], []
    for pl, ml in zip(payloads, max_lens):
        t, p = _dev(pl, src_off)
        o = torch.zeros(ml + out_off + 1, dtype=torch.uint8, device=DEV)
        keep.append((t, o))
        sts.append(qa.RansStream(p, len(pl), ml, o.data_ptr() + out_off))
    lens, st = gpu.rans_decompress(sts, tagged=tagged)
    torch.cuda.synchronize()
    res = []
    for (t, o), ln, s in zip(keep, lens, st):
        res.append((s, o[out_off:out_off + ln].cpu().numpy().tobytes() if s == qa.OK else None))
    return res


def oracle_dec(payload, max_len):
    try:
        return qa.OK, orc.rans_decompress(payload, max_len)
    except ValueError:
        return qa.ERR_CORRUPT, None


def _lcg_bytes(seed, n, f):
    s, out = seed, bytearray()
    for _ in range(n):
        s = (s * 6364136223846793005 + 1) & (2**64 - 1)
        out.append(f(s))
    return bytes(out)


def _inputs():
    r = np.random.default_rng(11)
    cases = [b"\x00", bytes([42] * 1000), bytes([1, 2, 3, 4, 5]),
             _lcg_bytes(1, 20000, lambda s: (s >> 60) & (0xFF if s & 0xF == 0 else 0x07)),
             bytes(i % 256 for i in range(256 * 40)),
             _lcg_bytes(0x9E3779B9, 30000, lambda s: s >> 56),
             bytes([0] * 9000 + [i % 255 + 1 for i in range(1000)]),
             bytes([0] * 5000 + [1] * 5000 + [2, 3, 4]), bytes([0, 1, 2])]
    for n in (1, 2, 3, 4, 5, 7, 63, 64, 65, 255, 256, 257, 1023, 1024, 1025, 4097):
        k = int(r.integers(1, 257))
        cases.append(r.integers(0, k, n).astype(np.uint8).tobytes())
    # binary sources from 1 to ~0.05 bits per symbol: the stream position moves
    # slowly and round boundaries land on every offset of the payload window
    for p, n in ((0.5, 40_000), (0.7, 41_111), (0.9, 45_000), (0.97, 60_003), (0.995, 99_999)):
        cases.append((r.random(n) > p).astype(np.uint8).tobytes())
    # geometric-ish residual bytes and a large stream crossing many windows
    cases.append(np.minimum(r.geometric(0.3, 200_000), 255).astype(np.uint8).tobytes())
    cases.append(r.integers(0, 256, 300_001).astype(np.uint8).tobytes())
    cases.append((r.integers(0, 4, 150_000) * 60).astype(np.uint8).tobytes())
    return cases


def test_rans_encode_parity(gpu):
    cases = _inputs() + [b""]
    got = gpu_compress(gpu, cases)
    for i, (d, g) in enumerate(zip(cases, got)):
        want = orc.rans_compress(d)
        assert g == want, f"case {i} (n={len(d)}): stream differs"


def test_rans_decode_parity(gpu):
    cases = _inputs()
    payloads = [orc.rans_compress(d) for d in cases]
    res = gpu_decompress(gpu, payloads, [len(d) for d in cases])
    for i, (d, (st, out)) in enumerate(zip(cases, res)):
        assert st == qa.OK, f"case {i}"
        assert out == d, f"case {i} (n={len(d)}): bytes differ"


def test_rans_roundtrip_batch_and_alignment(gpu):
    r = np.random.default_rng(12)
    datas = [np.minimum(r.geometric(float(p), int(n)), 255).astype(np.uint8).tobytes()
             for p, n in zip(r.uniform(0.05, 0.9, 37), r.integers(1, 90_000, 37))]
    enc = gpu_compress(gpu, datas)
    for off_s, off_o in ((0, 0), (1, 3), (7, 1), (13, 8)):
        res = gpu_decompress(gpu, enc, [len(d) for d in datas], src_off=off_s, out_off=off_o)
        for d, (st, out) in zip(datas, res):
            assert st == qa.OK and out == d


def test_rans_decode_truncations_and_flips(gpu):
    d = np.minimum(np.random.default_rng(13).geometric(0.2, 3000), 255).astype(np.uint8).tobytes()
    c = orc.rans_compress(d)
    payloads = [c[:k] for k in range(len(c))]
    r = np.random.default_rng(14)
    for _ in range(300):
        b = bytearray(c)
        for _ in range(int(r.integers(1, 4))):
            b[int(r.integers(0, len(b)))] ^= 1 << int(r.integers(0, 8))
        payloads.append(bytes(b))
    for max_len in (len(d), len(d) - 1, 4 * len(d)):
        res = gpu_decompress(gpu, payloads, [max_len] * len(payloads))
        for i, (p, (st, out)) in enumerate(zip(payloads, res)):
            ost, oout = oracle_dec(p, max_len)
            assert st == ost, f"payload {i} max_len {max_len}: status {st} vs oracle {ost}"
            if st == qa.OK:
                assert out == oout, f"payload {i}: bytes differ"


def _v(x):
    out = bytearray()
    while True:
        b = x & 0x7F
        x >>= 7
        if x:
            out.append(b | 0x80)
        else:
            out.append(b)
            return bytes(out)


def test_rans_decode_corner_cases(gpu):
    L = 1 << 23
    states = struct.pack("<4I", L, L, L, L)
    st0 = struct.pack("<4I", L + 2048, L, L, L)
    z = struct.pack("<4I", 0, L, L, L)
    zz = struct.pack("<4I", 0, 0, 5, 1)
    dup = _v(2) + bytes([5]) + _v(2048) + bytes([5]) + _v(2048)
    cases = [
        (_v(8) + _v(1) + bytes([9]) + _v((1 << 32) + 4096) + states, 8),  # f narrowed to u32
        (_v(1) + dup + st0 + b"\x01\x02", 1),                            # slot past cum[256]
        (_v(1) + dup + st0 + b"\x01", 1),                                # ... truncated
        (_v(1) + dup + z + b"\x01\x02\x03\x04", 1),                      # x' = 0: four bytes
        (_v(1) + dup + z + b"\x01\x02\x03", 1),                          # ... truncated
        (_v(4) + dup + zz + b"\x00\x00\x07\x01\x02\x03" * 3 + b"\x09" * 9, 4),  # zero runs
        (_v(7) + dup + zz + bytes(range(1, 40)), 7),
        (_v(1) + _v(1) + bytes([1]) + _v(4095) + states, 1),             # sum != M
        (_v(1) + _v(0) + states, 1), (_v(1) + _v(257) + states, 1),
        (_v(1) + _v(1) + bytes([1]) + _v(0) + states, 1),
        (_v(1) + _v(1) + bytes([1]) + _v(4097) + states, 1),
        (_v(9) + _v(1) + bytes([9]) + _v(4096) + states, 8),             # n above the bound
        (_v(0) + _v(1) + bytes([9]) + _v(4096) + states, 0),             # n = 0, valid
        (_v(0) + _v(1) + bytes([9]) + _v(4096) + states[:15], 0),        # n = 0, short states
        (b"", 4), (b"\x80" * 11, 4),                                     # empty, varint overflow
        (_v(3) + _v(1) + bytes([9]) + b"\x80" * 9 + b"\x01" + states, 3),
    ]
    payloads = [c for c, _ in cases]
    res = gpu_decompress(gpu, payloads, [m for _, m in cases])
    for i, ((p, m), (st, out)) in enumerate(zip(cases, res)):
        ost, oout = oracle_dec(p, m)
        assert st == ost, f"case {i}: status {st} vs oracle {ost}"
        if st == qa.OK:
            assert out == oout, f"case {i}: bytes differ"


def test_rans_tagged(gpu):
    d = bytes([3, 1, 4, 1, 5, 9, 2, 6] * 50)
    c = orc.rans_compress(d)
    payloads = [b"\x02" + c, b"\x00" + c, b"\x01" + c, b"\x03" + c, b"\x07" + c, b""]
    res = gpu_decompress(gpu, payloads, [len(d)] * len(payloads), tagged=True)
    assert res[0] == (qa.OK, d)
    assert [s for s, _ in res[1:]] == [qa.ERR_UNSUPPORTED] * 3 + [qa.ERR_CORRUPT] * 2


def test_rans_bound_checks(gpu):
    t, p = _dev(b"\x01" * 100)
    o = torch.zeros(100, dtype=torch.uint8, device=DEV)
    with pytest.raises(qa.QuoinError) as e:
        gpu.rans_compress([qa.RansInput(p, 100, o.data_ptr(), 100)])
    assert e.value.code == qa.ERR_BUFFER_TOO_SMALL
