"""GPU parity of the residual entropy layer (src/entropy/mod.rs) against the
oracle: quoin_gpu_code_residuals writes the same [tag] ++ coded bytes as
code_residuals at every level's lambda (rANS on the GPU; the range coder and the
LZ cascade on host cores), and quoin_gpu_decode_residuals gives back the
residuals, or the same success/failure as the oracle on damaged streams."""
import numpy as np
import pytest

import _oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a GPU", allow_module_level=True)

import quoin_amd as qa  # noqa: E402

DEV = torch.device("cuda:0")
M64 = (1 << 64) - 1


@pytest.fixture(scope="module")
def gpu():
    g = qa.Gpu(0)
    yield g
    g.close()


def _dev(b):
    t = torch.zeros(len(b) + 16, dtype=torch.uint8, device=DEV)
    if len(b):
        t[:len(b)] = torch.from_numpy(np.frombuffer(bytes(b), dtype=np.uint8).copy()).to(DEV)
    return t


def code(gpu, datas, lam, allow_lz):
    keep, ins = [], []
    for d in datas:
        t = _dev(d)
        cap = qa.lib.quoin_residual_bound(len(d))
        o = torch.zeros(cap, dtype=torch.uint8, device=DEV)
        keep.append((t, o))
        ins.append(qa.RansInput(t.data_ptr(), len(d), o.data_ptr(), cap))
    lens = gpu.code_residuals(ins, lam, allow_lz)
    torch.cuda.synchronize()
    return [o[:ln].cpu().numpy().tobytes() for (_, o), ln in zip(keep, lens)]


def decode(gpu, payloads, max_lens):
    keep, sts = [], []
    for pl, ml in zip(payloads, max_lens):
        t = _dev(pl)
        o = torch.zeros(ml + 1, dtype=torch.uint8, device=DEV)
        keep.append((t, o))
        sts.append(qa.RansStream(t.data_ptr(), len(pl), ml, o.data_ptr()))
    lens, st = gpu.decode_residuals(sts)
    torch.cuda.synchronize()
    return [(s, o[:ln].cpu().numpy().tobytes() if s == qa.OK else None)
            for (_, o), ln, s in zip(keep, lens, st)]


def _lcg(n, seed):
    s, out = seed, bytearray()
    for _ in range(n):
        s = (s * 6364136223846793005 + 1442695040888963407) & M64
        out.append((s >> 33) & 0xFF)
    return bytes(out)


def corpus():
    r = np.random.default_rng(31)
    block = bytes(((i * 2654435761) >> 16) & 0xFF for i in range(1024))
    walk = np.cumsum(r.integers(-3, 4, 30000)).astype(np.int64)
    resid = np.diff(walk, prepend=0).astype(np.int8).tobytes()
    return {
        "empty": b"",
        "one": b"\x07",
        "short63": bytes(r.integers(0, 256, 63, dtype=np.uint8)),
        "short65": bytes(r.integers(0, 256, 65, dtype=np.uint8)),
        "repeats": block * 50,                       # the LZ cascade wins at Max
        "noise": _lcg(4000, 0x12345678),
        "skewed": bytes(r.choice([0, 0, 0, 0, 0, 0, 1, 255], 40000).astype(np.uint8)),
        "alphabet": bytes(range(256)) * 16,
        "residuals": resid,
        "cycled": np.array([float(i & 15) for i in range(6000)]).tobytes(),
        "constant": b"\x00" * 9000,
    }


# (lambda, allow_lz) of Max, High, Balanced, Fast, Fastest (src/lib.rs:124-147)
LEVELS = [(0, True), (0, False), (2, False), (4, False), (16, False), (1, False), (3, True)]


@pytest.mark.parametrize("lam,allow_lz", LEVELS)
def test_code_residuals_matches_oracle(gpu, lam, allow_lz):
    cs = corpus()
    print(f"case code_residuals lambda={lam} allow_lz={allow_lz}", flush=True)
    got = code(gpu, list(cs.values()), lam, allow_lz)
    for (name, d), g in zip(cs.items(), got):
        want = orc.code_residuals(d, lam, allow_lz)
        assert g == want, f"{name}: tag {g[:1].hex()} len {len(g)} vs oracle tag {want[:1].hex()} len {len(want)}"
    tags = {g[0] for g in got}
    if lam == 0 and allow_lz:
        assert {0, 2, 3} <= tags, tags          # every coder takes some stream
    if lam >= 2 and not allow_lz:
        assert tags == {0, 2}, tags             # the range coder only for the empty stream
    print(f"case decode_residuals lambda={lam} allow_lz={allow_lz}", flush=True)
    back = decode(gpu, got, [len(d) for d in cs.values()])
    for (name, d), (st, b) in zip(cs.items(), back):
        assert st == qa.OK and b == d, name


def test_decode_residuals_oracle_streams_and_errors(gpu):
    cs = corpus()
    payloads, maxes, expect = [], [], []
    for name, d in cs.items():
        for lam, lz in ((0, True), (2, False)):
            payloads.append(orc.code_residuals(d, lam, lz))
            maxes.append(len(d))
            expect.append(d)
    # legacy TAG_TANS streams (tans.rs; decoded, no longer produced)
    for name in ("skewed", "alphabet", "residuals", "one"):
        d = cs[name]
        t = orc.tans_compress(d)
        if t is not None:
            payloads.append(b"\x01" + t)
            maxes.append(len(d))
            expect.append(d)
    rep = cs["repeats"]
    lzs = orc.code_residuals(rep, 0, True)
    assert lzs[0] == 3
    bad = [
        (b"", 10),                               # no tag
        (b"\x01\x00", 10),                       # tANS, header cut short
        (b"\x07\x00", 10),                       # unknown tag
        (lzs, len(rep) - 1),                     # cascade length over the bound
        (b"\x00" + orc.rc_compress(rep), len(rep) - 1),
        (b"\x02" + orc.rans_compress(rep), len(rep) - 1),
        (lzs[:len(lzs) // 2], len(rep)),         # truncated cascade
        (b"\x03\x05\x00" + b"\x00" * 6, 100),    # cascade whose LZ stream says 0 bytes, not 5
    ]
    for pl, ml in bad:
        payloads.append(pl)
        maxes.append(ml)
        expect.append(None)
    print(f"case decode_residuals {len(payloads)} oracle / damaged streams", flush=True)
    got = decode(gpu, payloads, maxes)
    for i, ((st, b), pl, ml, e) in enumerate(zip(got, payloads, maxes, expect)):
        try:
            want = orc.decode_residuals(pl, ml)
            werr = None
        except ValueError as ex:
            want, werr = None, ex.args[1]
        if werr is None:
            assert st == qa.OK and b == want, i
            if e is not None:
                assert b == e, i
        else:
            want_st = qa.ERR_UNSUPPORTED if werr == -17 else qa.ERR_CORRUPT
            assert st == want_st, (i, st, werr)


def test_code_residuals_buffer_too_small(gpu):
    d = bytes(range(200))
    t = _dev(d)
    cap = qa.lib.quoin_residual_bound(len(d)) - 1
    o = torch.zeros(cap, dtype=torch.uint8, device=DEV)
    with pytest.raises(qa.QuoinError) as e:
        gpu.code_residuals([qa.RansInput(t.data_ptr(), len(d), o.data_ptr(), cap)], 0, True)
    assert e.value.code == qa.ERR_BUFFER_TOO_SMALL
