"""The benchmark's column generator (bench_gen.c): xoshiro256** seeded by
SplitMix64, pinned against the published SplitMix64 output and an independent
restatement of both generators (Blackman and Vigna 2018), plus the workload
properties the benchmark states. CPU only."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
M = (1 << 64) - 1


def _lib():
    so = os.path.join(ROOT, "bench_gen.so")
    if not os.path.exists(so):
        subprocess.check_call(["gcc", "-std=c11", "-O2", "-shared", "-fPIC",
                               os.path.join(ROOT, "bench_gen.c"), "-lm", "-o", so])
    g = ctypes.CDLL(so)
    g.qg_checksum.restype = ctypes.c_uint64
    g.qg_checksum.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    return g


G = _lib()


def _raw(seed, n):
    out = np.empty(n, dtype=np.uint64)
    G.qg_raw(ctypes.c_uint64(seed), ctypes.c_uint64(n), ctypes.c_void_p(out.ctypes.data))
    return [int(x) for x in out]


def _col(name, seed, n, dtype, *extra):
    out = np.empty(n, dtype=dtype)
    getattr(G, "qg_" + name)(ctypes.c_uint64(seed), ctypes.c_uint64(n), *extra,
                             ctypes.c_void_p(out.ctypes.data))
    return out


def _splitmix(x):
    x = (x + 0x9E3779B97F4A7C15) & M
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return x, z ^ (z >> 31)


def _xoshiro(seed, n):
    x, s = seed, []
    for _ in range(4):
        x, z = _splitmix(x)
        s.append(z)
    rotl = lambda v, k: ((v << k) | (v >> (64 - k))) & M
    out = []
    for _ in range(n):
        out.append((rotl((s[1] * 5) & M, 7) * 9) & M)
        t = (s[1] << 17) & M
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 45)
    return out


def test_splitmix64_published_first_output():
    # SplitMix64 from state 0: first output 0xE220A8397B1DCDAF
    assert _splitmix(0)[1] == 0xE220A8397B1DCDAF


@pytest.mark.parametrize("seed", [0, 1, 12345, 2**64 - 1])
def test_xoshiro256ss_matches_restatement(seed):
    assert _raw(seed, 64) == _xoshiro(seed, 64)


def test_checksum_is_fnv1a_over_words():
    assert G.qg_checksum(None, 0) == 0xCBF29CE484222325
    b = np.arange(13, dtype=np.uint8)
    h = 0xCBF29CE484222325
    w = int.from_bytes(bytes(b[:8]), "little")
    h = ((h ^ w) * 0x100000001B3) & M
    for x in b[8:]:
        h = ((h ^ int(x)) * 0x100000001B3) & M
    assert G.qg_checksum(ctypes.c_void_p(b.ctypes.data), 13) == h


def test_workload_properties():
    n = 200_000
    ts = _col("timestamps_i64", 5, n, np.int64)
    assert ts[0] == 1_700_000_000_000_000_000
    d = np.diff(ts)
    assert d.min() >= 1 and d.max() <= 1000
    k = _col("kbits_u32", 6, n, np.uint32, ctypes.c_uint32(11))
    assert k.max() < 2**11 and k.max() > 2**11 - 16
    vals = np.ones(n, dtype=np.uint32)
    bm = np.empty((n + 7) // 8, dtype=np.uint8)
    G.qg_validity(ctypes.c_uint64(7), ctypes.c_uint64(n), ctypes.c_double(0.99),
                  ctypes.c_void_p(bm.ctypes.data), ctypes.c_void_p(vals.ctypes.data))
    valid = np.unpackbits(bm, bitorder="little")[:n].astype(bool)
    assert abs(valid.mean() - 0.99) < 0.002
    assert (vals[~valid] == 0).all() and (vals[valid] == 1).all()
    dec = _col("decimal_f64", 8, n, np.float64)
    assert dec.min() >= 0 and dec.max() < 1e6
    assert np.array_equal(np.round(dec * 1000) / 1000, dec)
    full = _col("full_f64", 9, n, np.float64)
    assert full.min() >= 0 and full.max() < 1
    runs = _col("dict_runs_f64", 10, n, np.float64)
    assert len(np.unique(runs)) == 256
    mean_run = n / (1 + np.count_nonzero(np.diff(runs)))
    assert 7.5 < mean_run < 8.5
    f32 = _col("unif_f32", 11, n, np.float32)
    assert f32.min() >= -1 and f32.max() < 1
    walk = _col("walk_f64", 0, n, np.float64)
    assert np.array_equal(np.round(walk * 100) / 100, walk)
    steps = np.diff(walk)
    assert abs(steps.std() - 1.0) < 0.02 and abs(steps.mean()) < 0.02
    sine = _col("sine_f64", 12, n, np.float64)
    i = np.arange(n, dtype=np.float64)
    noise = sine - np.sin(2 * np.pi * i / 1000) * 1000
    assert abs(noise.std() - 1e-3) < 5e-5


def test_columns_depend_only_on_seed_and_n():
    a = _col("walk_f64", 3, 1000, np.float64)
    b = _col("walk_f64", 3, 5000, np.float64)
    assert np.array_equal(a, b[:1000])
    assert not np.array_equal(a, _col("walk_f64", 4, 1000, np.float64))
