"""The engine's host coders (the range coder, rANS restated for host cores, LZ,
the entropy tag layer of src/entropy/mod.rs:45-178), pinned on the CPU
against the oracle: code_residuals gives the oracle's bytes at every lambda
with and without the LZ cascade, and decode_residuals gives the oracle's
outcome (the same bytes, or a refusal) on good, mutated and truncated
streams of every tag."""
import ctypes as C
import os

import numpy as np
import pytest

import _oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "quoin_amd", "libquoin_amd.so")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("libquoin_amd.so not built")
    lib = C.CDLL(LIB)
    lib.qa_host_code_residuals.argtypes = [C.c_void_p, C.c_size_t, C.c_uint64, C.c_int,
                                           C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]
    lib.qa_host_decode_residuals.argtypes = [C.c_void_p, C.c_size_t, C.c_uint64, C.c_void_p,
                                             C.c_size_t, C.POINTER(C.c_size_t)]
    return lib


def code(lib, data, lam, allow_lz):
    d = np.frombuffer(bytes(data) + b"\0", dtype=np.uint8)
    cap = 2 * len(data) + 4096
    out = np.zeros(cap, dtype=np.uint8)
    ln = C.c_size_t()
    assert lib.qa_host_code_residuals(d.ctypes.data, len(data), lam, int(allow_lz),
                                      out.ctypes.data, cap, C.byref(ln)) == 0
    return bytes(out[:ln.value])


def decode(lib, payload, max_len):
    p = np.frombuffer(bytes(payload) + b"\0", dtype=np.uint8)
    cap = max_len + 16
    out = np.zeros(cap, dtype=np.uint8)
    ln = C.c_size_t()
    rc = lib.qa_host_decode_residuals(p.ctypes.data, len(payload), max_len, out.ctypes.data, cap,
                                      C.byref(ln))
    return None if rc != 0 else bytes(out[:ln.value])


def oracle_decode(payload, max_len):
    try:
        return orc.decode_residuals(payload, max_len)
    except ValueError:
        return None


def streams(seed):
    r = np.random.default_rng(seed)
    return {
        "empty": b"",
        "one": b"\x07",
        "zeros": bytes(500),
        "skewed": bytes(np.minimum(r.geometric(0.3, 3000), 255).astype(np.uint8)),
        "uniform": bytes(r.integers(0, 256, 2000, dtype=np.uint8)),
        "repeats": bytes(np.tile(r.integers(0, 256, 37, dtype=np.uint8), 60)),
        "text": b"the quick brown fox jumps over the lazy dog " * 40,
        "varints": bytes(np.concatenate([r.integers(0, 128, 1500), r.integers(128, 256, 200)])
                         .astype(np.uint8)),
    }


@pytest.mark.parametrize("lam", [0, 2, 4, 16])
@pytest.mark.parametrize("allow_lz", [False, True])
def test_code_residuals_matches_the_oracle(lib, lam, allow_lz):
    for name, s in streams(lam + 7 * allow_lz).items():
        got = code(lib, s, lam, allow_lz)
        assert got == orc.code_residuals(s, lam, allow_lz), (name, lam, allow_lz)
        assert decode(lib, got, len(s)) == s, name


def test_damaged_streams_get_the_oracles_outcome(lib):
    r = np.random.default_rng(11)
    tags = set()
    for lam, allow_lz in ((0, True), (0, False), (16, False), (2, True)):
        for name, s in streams(3).items():
            c = code(lib, s, lam, allow_lz)
            if not c:
                continue
            tags.add(c[0])
            for it in range(40):
                b = bytearray(c)
                k = int(r.integers(0, len(b)))
                if it % 4 == 0:
                    b = b[:k]
                elif it % 4 == 1 and len(b) > 1:
                    b[k] ^= 1 << int(r.integers(0, 8))
                else:
                    b[k] = int(r.integers(0, 256))
                b = bytes(b)
                assert decode(lib, b, len(s)) == oracle_decode(b, len(s)), (name, lam, it)
    assert {0, 2, 3} <= tags, tags
    # the legacy tANS tag is still decoded (mod.rs:163-178)
    t = b"\x01" + orc.tans_compress(b"abracadabra" * 30)
    assert decode(lib, t, 330) == oracle_decode(t, 330) == b"abracadabra" * 30
    for cut in range(1, len(t)):
        assert decode(lib, t[:cut], 330) == oracle_decode(t[:cut], 330)
