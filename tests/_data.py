"""Seeded synthetic columns shared by the parity tests (numpy PCG64)."""
import numpy as np


def rng(seed=0):
    return np.random.default_rng(seed)


def with_nulls(values, density, seed=1):
    """Arrow bitmap with Bernoulli(1-density) validity; null slots zeroed."""
    r = rng(seed)
    valid = r.random(len(values)) >= density
    v = values.copy()
    v[~valid] = 0
    bm = np.packbits(valid.astype(np.uint8), bitorder="little")
    return v, bm


def int_columns(n, seed=0):
    r = rng(seed)
    ts = (1_700_000_000_000_000_000 + np.cumsum(r.integers(1, 1001, n))).astype(np.int64)
    cols = {
        "ts_i64": ts,
        "u32_k3": (r.integers(0, 2**32, n, dtype=np.uint64) >> 29).astype(np.uint32),
        "u32_k11": (r.integers(0, 2**32, n, dtype=np.uint64) >> 21).astype(np.uint32),
        "u32_k23": (r.integers(0, 2**32, n, dtype=np.uint64) >> 9).astype(np.uint32),
        "i32_mixed": (r.integers(-1000, 1000, n)).astype(np.int32),
        "u64_full": r.integers(0, 2**63, n, dtype=np.uint64) * 2 + 1,
        "i64_ramp": np.arange(n, dtype=np.int64) * 3 - 7,
        "i64_const": np.full(n, -5, dtype=np.int64),
        "i64_runs": np.repeat(r.integers(0, 50, n // 8 + 1), 8)[:n].astype(np.int64),
        # FoR wider than 32 bits (the 16-lane FastLanes form)
        "i64_w40": r.integers(0, 2**40, n).astype(np.int64),
        "i64_w57": r.integers(0, 2**57, n, dtype=np.int64),
        # Delta with zig-zag deltas of about 33 bits
        "i64_wide_walk": np.cumsum(r.integers(-2**31, 2**31, n)).astype(np.int64),
        # Delta into a 4-byte column
        "i32_walk": (7 + np.cumsum(r.integers(-3, 4, n))).astype(np.int32),
    }
    return cols


def float_columns(n, seed=0):
    r = rng(seed)
    walk = np.round(100 * (1000 + np.cumsum(r.normal(size=n)))) / 100
    cols = {
        "walk_f64": walk,
        "dec3_f64": r.integers(0, 10**9, n) / 1000.0,
        "unif_f64": (r.integers(0, 2**53, n, dtype=np.uint64) * 2.0**-53),
        "unif_f32": ((r.integers(0, 2**24, n) * 2.0**-23) - 1).astype(np.float32),
        "sine_f64": np.sin(np.arange(n) * 2 * np.pi / 1000) * 1e3 + r.normal(0, 1e-3, n),
    }
    return cols


def special_f32():
    bits = np.array([0, 0x80000000, 1, 0x807FFFFF, 0x00400000, 0x7F800000, 0xFF800000,
                     0x7FC00000, 0x7FA00001, 0xFFC12345, 0x3F800000, 0x7F7FFFFF, 0x00800000],
                    dtype=np.uint32)
    return bits.view(np.float32)
