"""ctypes bindings for the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: the oracle is the checker the HIP path is compared
against; nothing in the product imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

F64, I64, U64, I32, U32, F32 = 0, 1, 2, 3, 4, 5
FASTEST, FAST, BALANCED, HIGH, MAX = 0, 1, 2, 3, 4
NP_DTYPE = {F64: np.float64, I64: np.int64, U64: np.uint64, I32: np.int32, U32: np.uint32,
            F32: np.float32}

MODE = dict(PRED=0, CONST=1, STRIDE=2, XORZ=3, LZ=4, RAW=5, ORDERED_DELTA=7, ALP=9,
            BYTE_TRANSPOSE=11, FLOAT_MULT=17, DICT=19, DELTA2=20, PRED2=25, PRED_RC=29,
            DELTA_DP=45, FOR_BITPACK=50, DELTA_BITPACK=51, ALP_RD=52, RLE=53, PCO=54)


def _load():
    if not os.path.exists(_LIB):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    lib = ctypes.CDLL(_LIB)
    u8p = ctypes.POINTER(ctypes.c_uint8)
    lib.orc_compress_column.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t,
                                        ctypes.POINTER(u8p), ctypes.POINTER(ctypes.c_size_t)]
    lib.orc_decompress_column.argtypes = [ctypes.c_void_p, ctypes.c_size_t,
                                          ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_size_t),
                                          ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(u8p)]
    lib.orc_free.argtypes = [ctypes.c_void_p]
    lib.orc_compress_column_cfg.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(u8p), ctypes.POINTER(ctypes.c_size_t)]
    lib.orc_encode_mode_lv.restype = u8p
    lib.orc_encode_mode_lv.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]
    lib.orc_decode_mode_lv.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    lib.orc_estimate_order1_bytes.restype = ctypes.c_size_t
    lib.orc_estimate_order1_bytes.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    lib.orc_encode_mode.restype = u8p
    lib.orc_encode_mode.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_size_t)]
    lib.orc_decode_mode.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                    ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    lib.orc_rans_compress.restype = u8p
    lib.orc_rans_compress.argtypes = [ctypes.c_void_p, ctypes.c_size_t,
                                      ctypes.POINTER(ctypes.c_size_t)]
    lib.orc_rans_decompress.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
                                        ctypes.POINTER(u8p), ctypes.POINTER(ctypes.c_size_t)]
    for f in ("orc_rc_compress", "orc_lz_compress", "orc_tans_compress"):
        getattr(lib, f).restype = u8p
        getattr(lib, f).argtypes = [ctypes.c_void_p, ctypes.c_size_t,
                                    ctypes.POINTER(ctypes.c_size_t)]
    lib.orc_rc_decompress.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
                                      ctypes.POINTER(u8p), ctypes.POINTER(ctypes.c_size_t)]
    lib.orc_tans_decompress.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
                                        ctypes.POINTER(u8p), ctypes.POINTER(ctypes.c_size_t)]
    lib.orc_lz_decompress.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
                                      ctypes.POINTER(u8p)]
    lib.orc_crc32c_u64.restype = ctypes.c_uint32
    lib.orc_crc32c_u64.argtypes = [ctypes.c_uint32, ctypes.c_uint64]
    lib.orc_pred_encode.restype = u8p
    lib.orc_pred_encode.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_size_t)]
    lib.orc_pred_decode.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
                                    ctypes.c_int, ctypes.c_void_p]
    lib.orc_compress_decimal.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_int8, ctypes.c_uint8,
                                         ctypes.c_int, ctypes.POINTER(u8p),
                                         ctypes.POINTER(ctypes.c_size_t)]
    lib.orc_decompress_decimal.argtypes = [ctypes.c_void_p, ctypes.c_size_t,
                                           ctypes.POINTER(ctypes.c_int),
                                           ctypes.POINTER(ctypes.c_size_t),
                                           ctypes.POINTER(ctypes.c_int8),
                                           ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(u8p),
                                           ctypes.POINTER(u8p)]
    lib.orc_code_residuals.restype = u8p
    lib.orc_code_residuals.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64,
                                       ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]
    lib.orc_decode_residuals.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
                                         ctypes.POINTER(u8p), ctypes.POINTER(ctypes.c_size_t)]
    lib.orc_linear_select_order.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    lib.orc_linear_encode.restype = u8p
    lib.orc_linear_encode.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_size_t)]
    lib.orc_linear_decode.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                      ctypes.c_size_t, ctypes.c_void_p]
    lib.orc_plan_blocks.restype = ctypes.c_size_t
    lib.orc_plan_blocks.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
                                    ctypes.c_void_p]
    for f in ("orc_pack32", "orc_unpack32", "orc_pack64", "orc_unpack64"):
        getattr(lib, f).argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
    lib.orc_mode_wins.argtypes = [ctypes.c_void_p]
    lib.orc_decompress_column_mt.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p]
    lib.orc_block_feats.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    return lib


lib = _load()


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def compress(values, dtype, validity=None, level=FAST, block_size=0, selection=0,
             predictor_log2=16, threads=1):
    """selection 0 = Full, 1 = Sample; threads = host threads over blocks."""
    values = np.ascontiguousarray(values, dtype=NP_DTYPE[dtype])
    if validity is not None:
        validity = np.ascontiguousarray(validity, dtype=np.uint8)
    out = ctypes.POINTER(ctypes.c_uint8)()
    n = ctypes.c_size_t()
    rc = lib.orc_compress_column_cfg(dtype, len(values), _ptr(values), _ptr(validity), level,
                                     block_size, selection, predictor_log2, threads,
                                     ctypes.byref(out), ctypes.byref(n))
    if rc != 0:
        raise RuntimeError(f"oracle compress rc={rc}")
    b = ctypes.string_at(out, n.value)
    lib.orc_free(out)
    return b


def decompress(stream, threads=1):
    """Returns (dtype, values ndarray, validity ndarray|None) or raises.
    threads > 1 decodes the frames over that many threads."""
    buf = np.frombuffer(stream, dtype=np.uint8)
    dt = ctypes.c_int()
    n = ctypes.c_size_t()
    vals = ctypes.c_void_p()
    vb = ctypes.POINTER(ctypes.c_uint8)()
    if threads > 1:
        rc = lib.orc_decompress_column_mt(_ptr(buf) if len(buf) else None, len(buf), threads,
                                          ctypes.byref(dt), ctypes.byref(n), ctypes.byref(vals),
                                          ctypes.byref(vb))
    else:
        rc = lib.orc_decompress_column(_ptr(buf) if len(buf) else None, len(buf),
                                       ctypes.byref(dt), ctypes.byref(n), ctypes.byref(vals),
                                       ctypes.byref(vb))
    if rc != 0:
        raise ValueError(f"oracle decompress rc={rc}")
    npdt = NP_DTYPE[dt.value]
    nbytes = n.value * np.dtype(npdt).itemsize
    arr = np.frombuffer(ctypes.string_at(vals, nbytes), dtype=npdt).copy()
    lib.orc_free(vals)
    validity = None
    if vb:
        validity = np.frombuffer(ctypes.string_at(vb, (n.value + 7) // 8), dtype=np.uint8).copy()
        lib.orc_free(vb)
    return dt.value, arr, validity


def encode_mode(mode, lane, dtype=U64):
    lane = np.ascontiguousarray(lane, dtype=np.uint64)
    n = ctypes.c_size_t()
    p = lib.orc_encode_mode(MODE[mode] if isinstance(mode, str) else mode, _ptr(lane), len(lane),
                            dtype, ctypes.byref(n))
    if not p:
        return None
    b = ctypes.string_at(p, n.value)
    lib.orc_free(p)
    return b


def decode_mode(mode, payload, n, dtype=U64):
    buf = np.frombuffer(payload, dtype=np.uint8)
    out = np.zeros(max(n, 1), dtype=np.uint64)
    rc = lib.orc_decode_mode(MODE[mode] if isinstance(mode, str) else mode,
                             _ptr(buf) if len(buf) else None, len(buf), n, dtype, _ptr(out))
    if rc != 0:
        raise ValueError(f"decode rc={rc}")
    return out[:n]


def rans_compress(data):
    data = np.ascontiguousarray(np.frombuffer(bytes(data), dtype=np.uint8))
    n = ctypes.c_size_t()
    p = lib.orc_rans_compress(_ptr(data) if len(data) else None, len(data), ctypes.byref(n))
    if not p:
        return None
    b = ctypes.string_at(p, n.value)
    lib.orc_free(p)
    return b


def rans_decompress(payload, max_len):
    buf = np.frombuffer(payload, dtype=np.uint8)
    out = ctypes.POINTER(ctypes.c_uint8)()
    n = ctypes.c_size_t()
    rc = lib.orc_rans_decompress(_ptr(buf), len(buf), max_len, ctypes.byref(out), ctypes.byref(n))
    if rc != 0:
        raise ValueError(f"rans rc={rc}")
    b = ctypes.string_at(out, n.value)
    lib.orc_free(out)
    return b


def _bytes_in(data):
    data = np.ascontiguousarray(np.frombuffer(bytes(data), dtype=np.uint8))
    return data, (_ptr(data) if len(data) else None)


def _take(p, n):
    b = ctypes.string_at(p, n)
    lib.orc_free(p)
    return b


def rc_compress(data):
    """entropy/rc.rs compress_bytes: varint(len) ++ range-coded bytes."""
    data, dp = _bytes_in(data)
    n = ctypes.c_size_t()
    p = lib.orc_rc_compress(dp, len(data), ctypes.byref(n))
    return _take(p, n.value)


def rc_decompress(payload, max_len):
    buf, bp = _bytes_in(payload)
    out = ctypes.POINTER(ctypes.c_uint8)()
    n = ctypes.c_size_t()
    rc = lib.orc_rc_decompress(bp, len(buf), max_len, ctypes.byref(out), ctypes.byref(n))
    if rc != 0:
        raise ValueError(f"rc rc={rc}")
    return _take(out, n.value)


def tans_compress(data):
    """entropy/tans.rs compress_bytes; None where the reference returns None."""
    data, dp = _bytes_in(data)
    n = ctypes.c_size_t()
    p = lib.orc_tans_compress(dp, len(data), ctypes.byref(n))
    return _take(p, n.value) if p else None


def tans_decompress(payload, max_len):
    buf, bp = _bytes_in(payload)
    out = ctypes.POINTER(ctypes.c_uint8)()
    n = ctypes.c_size_t()
    rc = lib.orc_tans_decompress(bp, len(buf), max_len, ctypes.byref(out), ctypes.byref(n))
    if rc != 0:
        raise ValueError(f"tans rc={rc}", rc)
    return _take(out, n.value)


def lz_compress(data):
    """codecs/lz.rs lz_compress."""
    data, dp = _bytes_in(data)
    n = ctypes.c_size_t()
    p = lib.orc_lz_compress(dp, len(data), ctypes.byref(n))
    return _take(p, n.value)


def lz_decompress(stream, expected):
    buf, bp = _bytes_in(stream)
    out = ctypes.POINTER(ctypes.c_uint8)()
    rc = lib.orc_lz_decompress(bp, len(buf), expected, ctypes.byref(out))
    if rc != 0:
        raise ValueError(f"lz rc={rc}")
    return _take(out, expected)


def compress_decimal(values, wire, validity=None, scale=0, precision=38, level=FAST):
    """decimal.rs container: values = bytes of n significands, 16 (wire 6,
    Decimal128) or 32 (wire 7, Decimal256) little-endian bytes each."""
    w = 16 if wire == 6 else 32
    buf = np.frombuffer(bytes(values), dtype=np.uint8)
    n = len(buf) // w
    vb = np.ascontiguousarray(validity, dtype=np.uint8) if validity is not None else None
    out = ctypes.POINTER(ctypes.c_uint8)()
    ln = ctypes.c_size_t()
    rc = lib.orc_compress_decimal(wire, n, _ptr(buf) if n else None, _ptr(vb), scale, precision,
                                  level, ctypes.byref(out), ctypes.byref(ln))
    if rc != 0:
        raise RuntimeError(f"oracle decimal compress rc={rc}")
    return _take(out, ln.value)


def decompress_decimal(stream):
    """(wire, values bytes, validity bitmap or None, scale, precision); raises
    ValueError(rc) on error."""
    buf, bp = _bytes_in(stream)
    wire, n = ctypes.c_int(), ctypes.c_size_t()
    sc, pr = ctypes.c_int8(), ctypes.c_uint8()
    vals, vb = ctypes.POINTER(ctypes.c_uint8)(), ctypes.POINTER(ctypes.c_uint8)()
    rc = lib.orc_decompress_decimal(bp, len(buf), ctypes.byref(wire), ctypes.byref(n),
                                    ctypes.byref(sc), ctypes.byref(pr), ctypes.byref(vals),
                                    ctypes.byref(vb))
    if rc != 0:
        raise ValueError(f"decimal rc={rc}", rc)
    w = 16 if wire.value == 6 else 32
    v = _take(vals, n.value * w)
    bm = _take(vb, (n.value + 7) // 8) if vb else None
    return wire.value, v, bm, sc.value, pr.value


def crc32c_u64(crc, v):
    """hash.rs crc32c_u64_sw."""
    return int(lib.orc_crc32c_u64(crc, v))


def pred_encode(kind, lane, log2):
    """codecs/pred.rs encode (kind 0) / dfcm_encode (kind 25)."""
    lane = np.ascontiguousarray(lane, dtype=np.uint64)
    n = ctypes.c_size_t()
    p = lib.orc_pred_encode(kind, _ptr(lane) if len(lane) else None, len(lane), log2, ctypes.byref(n))
    return _take(p, n.value)


def pred_decode(kind, payload, n, log2):
    """(rc, values) of pred.rs decode / dfcm_decode."""
    buf, bp = _bytes_in(payload)
    out = np.zeros(max(n, 1), dtype=np.uint64)
    rc = lib.orc_pred_decode(kind, bp, len(buf), n, log2, _ptr(out))
    return rc, out[:n]


def code_residuals(data, lam, allow_lz):
    """entropy/mod.rs code_residuals: [tag] ++ coded."""
    data, dp = _bytes_in(data)
    n = ctypes.c_size_t()
    p = lib.orc_code_residuals(dp, len(data), lam, 1 if allow_lz else 0, ctypes.byref(n))
    return _take(p, n.value)


def decode_residuals(payload, max_len):
    """entropy/mod.rs decode_residuals; raises ValueError(rc) on error."""
    buf, bp = _bytes_in(payload)
    out = ctypes.POINTER(ctypes.c_uint8)()
    n = ctypes.c_size_t()
    rc = lib.orc_decode_residuals(bp, len(buf), max_len, ctypes.byref(out), ctypes.byref(n))
    if rc != 0:
        raise ValueError(f"decode_residuals rc={rc}", rc)
    return _take(out, n.value)


def block_feats(lane, dtype):
    """(exp_range, distinct_low, repeats, incompressible) of one block
    (encoder.rs:655-750)."""
    lane = np.ascontiguousarray(lane, dtype=np.uint64)
    out = np.zeros(4, dtype=np.uint32)
    lib.orc_block_feats(_ptr(lane) if len(lane) else None, len(lane), dtype, _ptr(out))
    return tuple(int(x) for x in out)


def encode_mode_lv(mode, lane, dtype, level, plog2=16):
    """One mode's payload at a level (encoder.rs:543-581), None when it
    declines."""
    lane = np.ascontiguousarray(lane, dtype=np.uint64)
    n = ctypes.c_size_t()
    p = lib.orc_encode_mode_lv(MODE[mode] if isinstance(mode, str) else mode,
                               _ptr(lane) if len(lane) else None, len(lane), plog2, dtype, level,
                               ctypes.byref(n))
    if not p:
        return None
    b = ctypes.string_at(p, n.value)
    lib.orc_free(p)
    return b


def lane_words(values):
    """The u64 lane words of a typed column (src/lib.rs:389-403): f32 widened
    to f64 bits, i32 sign-extended, u32 zero-extended."""
    v = np.asarray(values)
    if v.dtype == np.float32:
        return v.astype(np.float64).view(np.uint64)
    if v.dtype == np.int32:
        return v.astype(np.int64).view(np.uint64)
    if v.dtype == np.uint32:
        return v.astype(np.uint64)
    return np.ascontiguousarray(v).view(np.uint64)


def linear_select_order(lane):
    lane = np.ascontiguousarray(lane, dtype=np.uint64)
    return lib.orc_linear_select_order(_ptr(lane), len(lane))


def linear_encode(kind, lane, order=None):
    """Residual stream of `lane` (u64 words) for kind 7 / 20 / 45; order None =
    select_order. None when DeltaDp bails."""
    lane = np.ascontiguousarray(lane, dtype=np.uint64)
    if order is None:
        order = linear_select_order(lane)
    n = ctypes.c_size_t()
    p = lib.orc_linear_encode(kind, _ptr(lane) if len(lane) else None, len(lane), order,
                              ctypes.byref(n))
    if not p:
        return None
    b = ctypes.string_at(p, n.value)
    lib.orc_free(p)
    return b


def linear_decode(kind, payload, n):
    """(rc, values): rc 0 on success."""
    buf = np.frombuffer(bytes(payload) + b"\0", dtype=np.uint8)
    out = np.zeros(max(n, 1), dtype=np.uint64)
    rc = lib.orc_linear_decode(kind, _ptr(buf), len(payload), n, _ptr(out))
    return rc, out[:n]


def plan_blocks(lane, block_size=0):
    lane = np.ascontiguousarray(lane, dtype=np.uint64)
    ranges = np.zeros(2 * (len(lane) + 1), dtype=np.uint64)
    nb = lib.orc_plan_blocks(_ptr(lane), len(lane), block_size, _ptr(ranges))
    return [(int(ranges[2 * i]), int(ranges[2 * i + 1])) for i in range(nb)]


def mode_wins():
    out = np.zeros(64, dtype=np.uint64)
    lib.orc_mode_wins(_ptr(out))
    return out


def frames(stream):
    """Parse (mode, n, payload) frames of a stream without validity."""
    import struct
    pos = 16
    if stream[5] & 1:
        v, shift = 0, 0
        while True:
            b = stream[pos]
            pos += 1
            v |= (b & 0x7F) << shift
            shift += 7
            if not b & 0x80:
                break
        pos += v
    out = []
    while pos < len(stream):
        mode = stream[pos]
        n, plen = struct.unpack_from("<II", stream, pos + 1)
        out.append((mode, n, stream[pos + 9: pos + 9 + plen]))
        pos += 9 + plen
    return out
