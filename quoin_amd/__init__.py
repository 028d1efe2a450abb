"""quoin_amd — Python binding of the MI355X quoin block path (libquoin_amd.so).

Mirrors the reference's public API shape (src/lib.rs:97-547): ``Level``,
``Config``, ``compress``/``decompress`` for f64, ``compress_column`` /
``decompress_column`` for typed columns with an Arrow validity bitmap, and
``mode_win_counts``. ``Gpu`` exposes the device-resident batched entry points
(device pointers, e.g. from torch tensors) that the GPU build adds.

Everything runs through the HIP library; when it is missing or no GPU is
usable, calls raise instead of falling back to any CPU path.
"""
import ctypes
import enum
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libquoin_amd.so")

OK, ERR_NULL, ERR_BUFFER_TOO_SMALL, ERR_CORRUPT, ERR_PANIC = 0, -1, -2, -3, -4
ERR_UNSUPPORTED, ERR_DEVICE = -5, -6


class QuoinError(RuntimeError):
    def __init__(self, code, what=""):
        super().__init__(f"quoin error {code}{': ' + what if what else ''}")
        self.code = code


class Level(enum.IntEnum):
    """src/lib.rs:97-118"""
    FASTEST = 0
    FAST = 1
    BALANCED = 2
    HIGH = 3
    MAX = 4


# Arrow-native dtype tags (capi/include/quoin.h:97-115)
class DType(enum.IntEnum):
    BOOL = 0
    I8 = 1
    I16 = 2
    I32 = 3
    I64 = 4
    U8 = 5
    U16 = 6
    U32 = 7
    U64 = 8
    F32 = 9
    F64 = 10
    DECIMAL32 = 11
    DECIMAL64 = 12
    DECIMAL128 = 13
    DECIMAL256 = 14


DECIMAL_WIDTH = {DType.DECIMAL32: 4, DType.DECIMAL64: 8, DType.DECIMAL128: 16, DType.DECIMAL256: 32}


NP = {DType.BOOL: np.uint8, DType.I8: np.int8, DType.I16: np.int16, DType.I32: np.int32,
      DType.I64: np.int64, DType.U8: np.uint8, DType.U16: np.uint16, DType.U32: np.uint32,
      DType.U64: np.uint64, DType.F32: np.float32, DType.F64: np.float64}
FROM_NP = {np.dtype(np.int8): DType.I8, np.dtype(np.int16): DType.I16,
           np.dtype(np.int32): DType.I32, np.dtype(np.int64): DType.I64,
           np.dtype(np.uint8): DType.U8, np.dtype(np.uint16): DType.U16,
           np.dtype(np.uint32): DType.U32, np.dtype(np.uint64): DType.U64,
           np.dtype(np.float32): DType.F32, np.dtype(np.float64): DType.F64}


class Selection(enum.IntEnum):
    """src/lib.rs:150-161"""
    FULL = 0
    SAMPLE = 1


@dataclass
class Config:
    """src/lib.rs:194-245. Config::default is Level::Max, adaptive blocks,
    Selection::Full, predictor_log2 16 (clamped to [10, 16])."""
    level: Level = Level.MAX
    block_size: int = 0  # 0 = adaptive
    selection: Selection = Selection.FULL
    predictor_log2: int = 16


class _CConfig(ctypes.Structure):
    _fields_ = [("level", ctypes.c_int32), ("block_size", ctypes.c_uint32),
                ("selection", ctypes.c_int32), ("predictor_log2", ctypes.c_uint32)]


class DeviceColumn(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int32), ("reserved", ctypes.c_uint32), ("n", ctypes.c_uint64),
                ("values", ctypes.c_void_p), ("validity", ctypes.c_void_p),
                ("dst", ctypes.c_void_p), ("dst_cap", ctypes.c_uint64)]


class DeviceStream(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int32), ("reserved", ctypes.c_uint32), ("n", ctypes.c_uint64),
                ("src", ctypes.c_void_p), ("host_src", ctypes.c_void_p),
                ("src_len", ctypes.c_uint64), ("values_out", ctypes.c_void_p),
                ("validity_out", ctypes.c_void_p)]


class Job(ctypes.Structure):
    """QuoinJob (include/quoin.h): kind 0 compress (cfg, column), 1 decompress (stream)."""
    _fields_ = [("kind", ctypes.c_int32), ("reserved", ctypes.c_uint32), ("cfg", _CConfig),
                ("column", DeviceColumn), ("stream", DeviceStream)]


JOB_COMPRESS, JOB_DECOMPRESS = 0, 1


class RansInput(ctypes.Structure):
    """QuoinRansInput (include/quoin.h): device bytes to code and their output."""
    _fields_ = [("src", ctypes.c_void_p), ("n", ctypes.c_uint64), ("dst", ctypes.c_void_p),
                ("dst_cap", ctypes.c_uint64)]


class RansStream(ctypes.Structure):
    """QuoinRansStream (include/quoin.h): a device rANS payload and its output."""
    _fields_ = [("src", ctypes.c_void_p), ("src_len", ctypes.c_uint64), ("max_len", ctypes.c_uint64),
                ("out", ctypes.c_void_p)]


class LinearInput(ctypes.Structure):
    """QuoinLinearInput (include/quoin.h): a device block of u64 lane words and
    the buffer its residual stream goes to."""
    _fields_ = [("vals", ctypes.c_void_p), ("n", ctypes.c_uint64), ("dst", ctypes.c_void_p),
                ("dst_cap", ctypes.c_uint64)]


class LinearStream(ctypes.Structure):
    """QuoinLinearStream (include/quoin.h): a device residual stream and its output."""
    _fields_ = [("src", ctypes.c_void_p), ("src_len", ctypes.c_uint64), ("n", ctypes.c_uint64),
                ("out", ctypes.c_void_p)]


LINEAR_ORDERED_DELTA, LINEAR_DELTA2, LINEAR_DELTA_DP = 7, 20, 45


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run python quoin_amd/build.py")
    lib = ctypes.CDLL(LIB_PATH)
    sz, vp, u8p = ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p
    lib.quoin_version.restype = ctypes.c_char_p
    lib.quoin_compress_bound.restype = sz
    lib.quoin_compress_bound.argtypes = [sz]
    lib.quoin_typed_compress_bound.restype = sz
    lib.quoin_typed_compress_bound.argtypes = [ctypes.c_int32, sz]
    lib.quoin_dtype_width.restype = sz
    lib.quoin_dtype_width.argtypes = [ctypes.c_int32]
    lib.quoin_typed_compress_cfg.restype = sz
    lib.quoin_typed_compress_cfg.argtypes = [ctypes.POINTER(_CConfig), ctypes.c_int32, sz, vp,
                                             u8p, u8p, sz]
    lib.quoin_typed_compress.restype = sz
    lib.quoin_typed_compress.argtypes = [ctypes.c_int32, sz, vp, u8p, u8p, sz]
    lib.quoin_typed_decompress.argtypes = [ctypes.c_int32, sz, u8p, sz, vp, u8p]
    lib.quoin_typed_compress_decimal.restype = sz
    lib.quoin_typed_compress_decimal.argtypes = [ctypes.c_int32, sz, vp, u8p, ctypes.c_int8,
                                                 ctypes.c_uint8, u8p, sz]
    lib.quoin_compress.argtypes = [vp, sz, u8p, sz, ctypes.POINTER(sz)]
    lib.quoin_compress_cfg.argtypes = [ctypes.POINTER(_CConfig), vp, sz, u8p, sz,
                                       ctypes.POINTER(sz)]
    lib.quoin_decompress.argtypes = [u8p, sz, vp, sz, ctypes.POINTER(sz)]
    lib.quoin_decompressed_value_count.restype = ctypes.c_ssize_t
    lib.quoin_decompressed_value_count.argtypes = [u8p, sz]
    lib.quoin_mode_win_counts.argtypes = [vp]
    lib.quoin_gpu_create.restype = vp
    lib.quoin_gpu_create.argtypes = [ctypes.c_int, ctypes.c_uint64]
    lib.quoin_gpu_destroy.argtypes = [vp]
    lib.quoin_gpu_compress.argtypes = [vp, ctypes.POINTER(_CConfig), ctypes.POINTER(DeviceColumn),
                                       sz, ctypes.POINTER(ctypes.c_uint64), vp]
    lib.quoin_gpu_decompress.argtypes = [vp, ctypes.POINTER(DeviceStream), sz, vp]
    lib.quoin_gpu_set_launch_form.argtypes = [vp, ctypes.c_int]
    lib.quoin_gpu_last_timing.argtypes = [vp, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_char_p)]
    lib.quoin_server_create.restype = vp
    lib.quoin_server_create.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32,
                                        ctypes.c_uint32]
    lib.quoin_server_submit.argtypes = [vp, ctypes.POINTER(Job), ctypes.POINTER(ctypes.c_uint64)]
    lib.quoin_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_double)]
    lib.quoin_server_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64),
                                       ctypes.POINTER(ctypes.c_uint64)]
    lib.quoin_server_destroy.argtypes = [vp]
    lib.quoin_rans_bound.restype = sz
    lib.quoin_rans_bound.argtypes = [sz]
    lib.quoin_gpu_rans_compress.argtypes = [vp, ctypes.POINTER(RansInput), sz,
                                            ctypes.POINTER(ctypes.c_uint64), vp]
    lib.quoin_gpu_rans_decompress.argtypes = [vp, ctypes.POINTER(RansStream), sz, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_uint64),
                                              ctypes.POINTER(ctypes.c_int32), vp]
    lib.quoin_gpu_pred_encode.argtypes = [vp, ctypes.c_int32, ctypes.c_uint32,
                                          ctypes.POINTER(LinearInput), sz,
                                          ctypes.POINTER(ctypes.c_uint64), vp]
    lib.quoin_gpu_pred_decode.argtypes = [vp, ctypes.c_int32, ctypes.c_uint32,
                                          ctypes.POINTER(LinearStream), sz,
                                          ctypes.POINTER(ctypes.c_int32), vp]
    lib.quoin_gpu_set_autotune.argtypes = [vp, ctypes.c_int]
    lib.quoin_gpu_tuned_shapes.argtypes = [vp, ctypes.c_void_p]
    lib.quoin_gpu_graph_counts.argtypes = [vp, ctypes.c_void_p]
    lib.quoin_residual_bound.restype = sz
    lib.quoin_residual_bound.argtypes = [sz]
    lib.quoin_gpu_code_residuals.argtypes = [vp, ctypes.c_uint64, ctypes.c_int,
                                             ctypes.POINTER(RansInput), sz,
                                             ctypes.POINTER(ctypes.c_uint64), vp]
    lib.quoin_gpu_decode_residuals.argtypes = [vp, ctypes.POINTER(RansStream), sz,
                                               ctypes.POINTER(ctypes.c_uint64),
                                               ctypes.POINTER(ctypes.c_int32), vp]
    lib.quoin_linear_bound.restype = ctypes.c_uint64
    lib.quoin_linear_bound.argtypes = [ctypes.c_uint64]
    lib.quoin_gpu_linear_encode.argtypes = [vp, ctypes.c_int32, ctypes.c_int32,
                                            ctypes.POINTER(LinearInput), sz,
                                            ctypes.POINTER(ctypes.c_uint64),
                                            ctypes.POINTER(ctypes.c_int32),
                                            ctypes.POINTER(ctypes.c_int32), vp]
    lib.quoin_gpu_linear_encode_noalloc.argtypes = lib.quoin_gpu_linear_encode.argtypes
    lib.quoin_gpu_linear_decode.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(LinearStream), sz,
                                            ctypes.POINTER(ctypes.c_int32), vp]
    return lib


lib = _load()

# every symbol include/quoin.h declares (checked by the CPU test suite)
EXPORTS = [
    "quoin_version", "quoin_compress_bound", "quoin_compress", "quoin_decompress",
    "quoin_decompressed_value_count", "quoin_ctx_create", "quoin_ctx_free", "quoin_compress_ctx",
    "quoin_decompress_ctx", "quoin_dtype_width", "quoin_typed_compress_bound",
    "quoin_typed_compress", "quoin_typed_compress_decimal", "quoin_typed_decompress",
    "quoin_typed_compress_cfg", "quoin_compress_cfg", "quoin_mode_win_counts",
    "quoin_reset_mode_win_counts", "quoin_last_error", "quoin_gpu_create", "quoin_gpu_destroy", "quoin_gpu_compress",
    "quoin_gpu_decompress", "quoin_gpu_set_launch_form", "quoin_gpu_last_timing",
    "quoin_server_create", "quoin_server_submit", "quoin_server_wait", "quoin_server_stats",
    "quoin_server_destroy", "quoin_rans_bound", "quoin_gpu_rans_compress",
    "quoin_gpu_rans_decompress", "quoin_linear_bound", "quoin_gpu_linear_encode",
    "quoin_gpu_linear_decode", "quoin_residual_bound", "quoin_gpu_code_residuals",
    "quoin_gpu_decode_residuals", "quoin_gpu_pred_encode", "quoin_gpu_pred_decode",
    "quoin_gpu_set_autotune", "quoin_gpu_tuned_shapes", "quoin_gpu_linear_encode_noalloc",
    "quoin_gpu_graph_counts",
]


def version():
    return lib.quoin_version().decode()


def _cfg(config):
    config = config or Config()
    return _CConfig(int(config.level), int(config.block_size), int(config.selection),
                    int(config.predictor_log2))


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None and a.size else None


def compress_column(values, validity=None, config=None, dtype=None):
    """Typed column -> stream bytes (src/lib.rs:429-455)."""
    values = np.ascontiguousarray(values)
    dtype = DType(dtype) if dtype is not None else FROM_NP[values.dtype]
    values = values.astype(NP[dtype], copy=False)
    if validity is not None:
        validity = np.ascontiguousarray(validity, dtype=np.uint8)
    n = len(values)
    cap = lib.quoin_typed_compress_bound(int(dtype), n)
    bs = (config or Config()).block_size
    if bs:
        cap += -(-n // bs) * 9  # 9 framing bytes per fixed-size block
    dst = np.empty(cap, dtype=np.uint8)
    cfg = _cfg(config)
    got = lib.quoin_typed_compress_cfg(ctypes.byref(cfg), int(dtype), n,
                                       _ptr(values) or ctypes.c_void_p(dst.ctypes.data),
                                       _ptr(validity), ctypes.c_void_p(dst.ctypes.data), cap)
    if got == 0:
        raise QuoinError(lib.quoin_last_error(), "compress failed")
    return dst[:got].tobytes()


def compress_decimal(values, dtype, validity=None, scale=0, precision=38):
    """Decimal column -> stream bytes (src/decimal.rs container; capi
    quoin_typed_compress_decimal). values: bytes of n little-endian
    significands of DECIMAL_WIDTH[dtype] bytes each."""
    dtype = DType(dtype)
    raw = np.frombuffer(bytes(values), dtype=np.uint8)
    n = len(raw) // DECIMAL_WIDTH[dtype]
    if validity is not None:
        validity = np.ascontiguousarray(validity, dtype=np.uint8)
    cap = 40 * n + 8192  # four U64 limb streams at their bound, and the container
    dst = np.empty(cap, dtype=np.uint8)
    got = lib.quoin_typed_compress_decimal(int(dtype), n, _ptr(raw) or ctypes.c_void_p(dst.ctypes.data),
                                           _ptr(validity), scale, precision,
                                           ctypes.c_void_p(dst.ctypes.data), cap)
    if got == 0:
        raise QuoinError(lib.quoin_last_error(), "decimal compress failed")
    return dst[:got].tobytes()


def decompress_decimal(stream, dtype, n, with_validity=False):
    """Decimal stream -> (significand bytes, validity bytes or None)."""
    dtype = DType(dtype)
    src = np.frombuffer(stream, dtype=np.uint8)
    out = np.empty(max(n * DECIMAL_WIDTH[dtype], 1), dtype=np.uint8)
    vo = np.empty(max((n + 7) // 8, 1), dtype=np.uint8) if with_validity else None
    rc = lib.quoin_typed_decompress(int(dtype), n, ctypes.c_void_p(src.ctypes.data), len(src),
                                    ctypes.c_void_p(out.ctypes.data), _ptr(vo))
    if rc != OK:
        raise QuoinError(rc)
    return out[:n * DECIMAL_WIDTH[dtype]].tobytes(), (vo[:(n + 7) // 8].tobytes() if vo is not None else None)


def decompress_column(stream, dtype, n, with_validity=False):
    """Stream -> (values, validity|None) (src/lib.rs:462-508)."""
    dtype = DType(dtype)
    src = np.frombuffer(stream, dtype=np.uint8)
    out = np.empty(max(n, 1), dtype=NP[dtype])
    vout = np.empty(max((n + 7) // 8, 1), dtype=np.uint8) if with_validity else None
    rc = lib.quoin_typed_decompress(int(dtype), n, ctypes.c_void_p(src.ctypes.data), len(src),
                                    ctypes.c_void_p(out.ctypes.data),
                                    ctypes.c_void_p(vout.ctypes.data) if vout is not None else None)
    if rc != OK:
        raise QuoinError(rc)
    return out[:n], (vout[:(n + 7) // 8] if vout is not None else None)


def compress(values, config=None):
    """f64 convenience (src/lib.rs:514-516)."""
    return compress_column(np.asarray(values, dtype=np.float64), None, config)


def decompress(stream):
    """f64 convenience (src/lib.rs:522-527)."""
    src = np.frombuffer(stream, dtype=np.uint8)
    n = lib.quoin_decompressed_value_count(ctypes.c_void_p(src.ctypes.data), len(src))
    if n < 0:
        raise QuoinError(int(n))
    out = np.empty(max(n, 1), dtype=np.float64)
    got = ctypes.c_size_t()
    rc = lib.quoin_decompress(ctypes.c_void_p(src.ctypes.data), len(src),
                              ctypes.c_void_p(out.ctypes.data), n, ctypes.byref(got))
    if rc != OK:
        raise QuoinError(rc)
    return out[:n]


def mode_win_counts():
    out = np.zeros(64, dtype=np.uint64)
    lib.quoin_mode_win_counts(ctypes.c_void_p(out.ctypes.data))
    return out


def reset_mode_win_counts():
    lib.quoin_reset_mode_win_counts()


PRED_FCM, PRED_DFCM = 0, 25  # frame modes Pred / Pred2 (src/mode.rs:11, 31)


class Gpu:
    """Handle over the device-resident batched API (include/quoin.h)."""

    def __init__(self, device=0, hbm_budget=0):
        self.h = lib.quoin_gpu_create(device, hbm_budget)
        if not self.h:
            raise QuoinError(ERR_DEVICE, "no usable GPU")

    def close(self):
        if self.h:
            lib.quoin_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_launch_form(self, form):
        rc = lib.quoin_gpu_set_launch_form(self.h, form)
        if rc != OK:
            raise QuoinError(rc)

    def compress(self, cols, config=None, stream=None):
        """cols: list (or prebuilt ctypes array) of DeviceColumn; returns stream
        lengths."""
        arr = cols if isinstance(cols, ctypes.Array) else (DeviceColumn * len(cols))(*cols)
        lens = (ctypes.c_uint64 * len(cols))()
        cfg = _cfg(config)
        rc = lib.quoin_gpu_compress(self.h, ctypes.byref(cfg), arr, len(cols), lens, stream)
        if rc != OK:
            raise QuoinError(rc)
        return [int(x) for x in lens]

    def decompress(self, streams, stream=None):
        """streams: list (or prebuilt ctypes array) of DeviceStream."""
        arr = streams if isinstance(streams, ctypes.Array) else (DeviceStream * len(streams))(*streams)
        rc = lib.quoin_gpu_decompress(self.h, arr, len(streams), stream)
        if rc != OK:
            raise QuoinError(rc)

    def rans_compress(self, inputs, stream=None):
        """inputs: list of RansInput; returns the coded lengths (src/entropy/rans.rs:52)."""
        arr = (RansInput * len(inputs))(*inputs)
        lens = (ctypes.c_uint64 * len(inputs))()
        rc = lib.quoin_gpu_rans_compress(self.h, arr, len(inputs), lens, stream)
        if rc != OK:
            raise QuoinError(rc)
        return [int(x) for x in lens]

    def rans_decompress(self, streams, tagged=False, stream=None):
        """streams: list of RansStream; returns (decoded lengths, per-stream status)
        (src/entropy/rans.rs:110). Failed streams are reported, not raised."""
        arr = (RansStream * len(streams))(*streams)
        lens = (ctypes.c_uint64 * len(streams))()
        st = (ctypes.c_int32 * len(streams))()
        rc = lib.quoin_gpu_rans_decompress(self.h, arr, len(streams), 1 if tagged else 0, lens, st,
                                           stream)
        if rc not in (OK, ERR_CORRUPT, ERR_UNSUPPORTED):
            raise QuoinError(rc)
        return [int(x) for x in lens], [int(x) for x in st]

    def code_residuals(self, inputs, lam, allow_lz=False, stream=None):
        """[tag] ++ coded for a batch of device byte streams
        (src/entropy/mod.rs:121-149). inputs: list of RansInput with dst_cap >=
        residual_bound(n). lam: Level::lambda (Max/High 0, Balanced 2, Fast 4,
        Fastest 16); allow_lz: Level::Max. Returns the coded lengths."""
        arr = (RansInput * len(inputs))(*inputs)
        lens = (ctypes.c_uint64 * len(inputs))()
        rc = lib.quoin_gpu_code_residuals(self.h, lam, 1 if allow_lz else 0, arr, len(inputs),
                                          lens, stream)
        if rc != OK:
            raise QuoinError(rc)
        return [int(x) for x in lens]

    def decode_residuals(self, streams, stream=None):
        """Inverse of code_residuals (mod.rs:151-178) for a batch of RansStream.
        Returns (decoded lengths, per-stream status); failed streams are
        reported, not raised."""
        arr = (RansStream * len(streams))(*streams)
        lens = (ctypes.c_uint64 * len(streams))()
        st = (ctypes.c_int32 * len(streams))()
        rc = lib.quoin_gpu_decode_residuals(self.h, arr, len(streams), lens, st, stream)
        if rc not in (OK, ERR_CORRUPT, ERR_UNSUPPORTED):
            raise QuoinError(rc)
        return [int(x) for x in lens], [int(x) for x in st]

    def pred_encode(self, kind, inputs, predictor_log2=16, stream=None):
        """FCM (kind PRED_FCM) / DFCM (kind PRED_DFCM) residual streams of a batch
        of blocks (src/codecs/pred.rs:23-37, 59-75). inputs: list of LinearInput.
        Returns the stream lengths."""
        k = len(inputs)
        arr = (LinearInput * k)(*inputs)
        lens = (ctypes.c_uint64 * k)()
        rc = lib.quoin_gpu_pred_encode(self.h, kind, predictor_log2, arr, k, lens, stream)
        if rc != OK:
            raise QuoinError(rc)
        return [int(x) for x in lens]

    def pred_decode(self, kind, streams, predictor_log2=16, stream=None):
        """Values of a batch of FCM / DFCM residual streams (pred.rs:39-57,
        77-97). Returns the per-block status (OK or ERR_CORRUPT)."""
        k = len(streams)
        arr = (LinearStream * k)(*streams)
        st = (ctypes.c_int32 * k)()
        rc = lib.quoin_gpu_pred_decode(self.h, kind, predictor_log2, arr, k, st, stream)
        if rc not in (OK, ERR_CORRUPT):
            raise QuoinError(rc)
        return [int(x) for x in st]

    def linear_encode(self, kind, inputs, order=0, stream=None):
        """Residual streams of a batch of blocks (src/codecs/linear.rs:31-45,
        125-159, 191-212). inputs: list of LinearInput. Returns (lengths, orders,
        statuses); status 1 = DeltaDp declined the block (the reference's None)."""
        k = len(inputs)
        arr = (LinearInput * k)(*inputs)
        lens = (ctypes.c_uint64 * k)()
        ords = (ctypes.c_int32 * k)()
        st = (ctypes.c_int32 * k)()
        rc = lib.quoin_gpu_linear_encode(self.h, kind, order, arr, k, lens, ords, st, stream)
        if rc != OK:
            raise QuoinError(rc)
        return [int(x) for x in lens], [int(x) for x in ords], [int(x) for x in st]

    def linear_encode_noalloc(self, kind, inputs, order=0, stream=None):
        """linear_encode with no device allocation: the scratch is one fixed
        device array per GPU (at most 1024 blocks and 65536 tiles of 2048
        values per call)."""
        k = len(inputs)
        arr = (LinearInput * k)(*inputs)
        lens = (ctypes.c_uint64 * k)()
        ords = (ctypes.c_int32 * k)()
        st = (ctypes.c_int32 * k)()
        rc = lib.quoin_gpu_linear_encode_noalloc(self.h, kind, order, arr, k, lens, ords, st,
                                                 stream)
        if rc != OK:
            raise QuoinError(rc)
        return [int(x) for x in lens], [int(x) for x in ords], [int(x) for x in st]

    def linear_decode(self, kind, streams, stream=None):
        """Values of a batch of residual streams (linear.rs:47-66, 161-180,
        214-236). Returns the per-block status (OK or ERR_CORRUPT)."""
        k = len(streams)
        arr = (LinearStream * k)(*streams)
        st = (ctypes.c_int32 * k)()
        rc = lib.quoin_gpu_linear_decode(self.h, kind, arr, k, st, stream)
        if rc not in (OK, ERR_CORRUPT):
            raise QuoinError(rc)
        return [int(x) for x in st]

    def set_autotune(self, on):
        rc = lib.quoin_gpu_set_autotune(self.h, 1 if on else 0)
        if rc != OK:
            raise QuoinError(rc)

    def tuned_shapes(self):
        """Tuned wavefronts per workgroup of k_dec_fl<4>, k_dec_fl<8>, k_emit_fl
        (0 = not tuned yet)."""
        out = (ctypes.c_uint32 * 3)()
        lib.quoin_gpu_tuned_shapes(self.h, out)
        return {"k_dec_fl<4>": out[0], "k_dec_fl<8>": out[1], "k_emit_fl": out[2]}

    def graph_counts(self):
        """Launch form 0's hipGraphs so far: captures and graph launches of the
        compress and decompress sequences."""
        out = (ctypes.c_uint64 * 4)()
        lib.quoin_gpu_graph_counts(self.h, out)
        return {"compress_captures": out[0], "compress_launches": out[1],
                "decompress_captures": out[2], "decompress_launches": out[3]}

    def last_timing(self):
        t, k = ctypes.c_double(), ctypes.c_double()
        name = ctypes.c_char_p()
        lib.quoin_gpu_last_timing(self.h, ctypes.byref(t), ctypes.byref(k), ctypes.byref(name))
        return t.value, k.value, (name.value or b"").decode()


class Server:
    """Serving loop over one device (include/quoin.h, quoin_server_*): jobs are
    admitted by an HBM byte budget and continuously batched per free slot."""

    def __init__(self, device=0, hbm_budget=0, slots=2, max_batch=0):
        self.h = lib.quoin_server_create(device, hbm_budget, slots, max_batch)
        if not self.h:
            raise QuoinError(ERR_DEVICE, "no usable GPU")

    def submit_compress(self, column, config=None):
        job = Job(JOB_COMPRESS, 0, _cfg(config), column, DeviceStream())
        return self._submit(job)

    def submit_decompress(self, stream):
        job = Job(JOB_DECOMPRESS, 0, _CConfig(0, 0), DeviceColumn(), stream)
        return self._submit(job)

    def _submit(self, job):
        t = ctypes.c_uint64()
        rc = lib.quoin_server_submit(self.h, ctypes.byref(job), ctypes.byref(t))
        if rc != OK:
            raise QuoinError(rc)
        return t.value

    def wait(self, ticket):
        """-> (return code, stream length for compress, latency in ms)"""
        n, ms = ctypes.c_uint64(), ctypes.c_double()
        rc = lib.quoin_server_wait(self.h, ticket, ctypes.byref(n), ctypes.byref(ms))
        return rc, n.value, ms.value

    def stats(self):
        j, b = ctypes.c_uint64(), ctypes.c_uint64()
        lib.quoin_server_stats(self.h, ctypes.byref(j), ctypes.byref(b))
        return j.value, b.value

    def close(self):
        if self.h:
            lib.quoin_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
