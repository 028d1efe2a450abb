// qa_common.h — structures shared by the host engine and the gfx950 kernels of
// the quoin block path, plus small device helpers.
//
// Layout in HBM (one batch = many columns):
//   lane[]   : every column lowered to its u64 lane (src/lib.rs:389-403) and,
//              with nulls, compacted to its valid values (src/validity.rs:122),
//              each column starting at a 1024-aligned offset.
//   regions  : per column, ceil(n / R) region slots of R values (R = 32768 for
//              adaptive planning, the fixed block size otherwise). A block
//              starts at a region slot (blocks are always region-aligned:
//              src/encoder.rs:142-178), so per-block state is indexed by the
//              slot where the block starts.
//   tiles    : per region, T = ceil(R / 1024) tile slots of 1024 values — the
//              FastLanes sub-block (src/bitpack.rs:22). A 128 Ki block spans
//              4 regions = 128 consecutive tile slots.
#pragma once
#include <stdint.h>

#define QA_TILE 1024
#define QA_BASE_Q 32768
#define QA_MAX_Q 131072

// engine dtype wire ids (src/dtype.rs:70-80)
enum { QA_F64 = 0, QA_I64 = 1, QA_U64 = 2, QA_I32 = 3, QA_U32 = 4, QA_F32 = 5 };
enum { QA_FASTEST = 0, QA_FAST = 1 };

// mode ids (src/mode.rs:9-85)
enum {
    QA_M_CONST = 1, QA_M_STRIDE = 2, QA_M_XORZ = 3, QA_M_RAW = 5, QA_M_ALP = 9,
    QA_M_FLOAT_MULT = 17, QA_M_DICT = 19, QA_M_FOR = 50, QA_M_DELTA = 51, QA_M_ALP_RD = 52,
    QA_M_RLE = 53,
    // a block whose payload the host competition wrote: the device lays out
    // its frame and leaves the mode byte and payload to k_host_frames
    QA_M_HOST = 0xFF
};

// error flags raised by kernels (host maps them to QUOIN_ERR_*)
// QA_ERR_STALLED: a look-back wait passed its bound (a device fault, not a
// property of the stream)
enum { QA_ERR_CORRUPT = 1, QA_ERR_UNSUPPORTED = 2, QA_ERR_OVERFLOW = 4, QA_ERR_STALLED = 8 };
// spin bound of the look-back waits (each round sleeps and reloads)
#define QA_SPIN_MAX (1u << 22)

struct DCol {
    const void *values;       // typed input, device
    const uint8_t *validity;  // device bitmap or null
    uint8_t *dst;             // device stream output
    uint64_t dst_cap;
    uint64_t n;               // logical values
    uint64_t lane_off;        // lane[] offset (values)
    uint64_t tile_base;       // first global tile slot
    uint64_t vword_base;      // first validity 64-bit word slot (global)
    uint64_t val_base;        // first logical value (global, columns concatenated)
    uint32_t nvwords;         // ceil(n / 64) when a bitmap is given, else 0
    uint32_t dtype;           // engine wire id
    uint32_t in_width;        // 4 or 8
    uint32_t region_size;     // R
    uint32_t tpr;             // tiles per region
    uint32_t region_base;     // first global region slot
    uint32_t nregions;        // region slots (upper bound from n)
    uint32_t adaptive;
    uint32_t level;
    uint32_t in_kind;         // QA_IN_*: input element type when narrower than the lane
    uint32_t plog2;           // the header's predictor_log2 (Config::predictor_log2)
};

// narrow Arrow inputs widened on load (capi/src/typed.rs:147-153)
enum { QA_IN_NATIVE = 0, QA_IN_I8 = 1, QA_IN_I16 = 2, QA_IN_U8 = 3, QA_IN_U16 = 4 };

// element form of a decoded column's output (src/lib.rs:492-500)
enum { QA_OUT_U8 = 0, QA_OUT_U16, QA_OUT_U32, QA_OUT_F32, QA_OUT_U64 };
#ifdef __HIPCC__
__host__ __device__
#endif
inline uint32_t qa_out_form(uint32_t dtype, uint32_t out_kind) {
    if (out_kind == QA_IN_I8 || out_kind == QA_IN_U8) return QA_OUT_U8;
    if (out_kind == QA_IN_I16 || out_kind == QA_IN_U16) return QA_OUT_U16;
    if (dtype == QA_F32) return QA_OUT_F32;
    if (dtype == QA_I32 || dtype == QA_U32) return QA_OUT_U32;
    return QA_OUT_U64;
}

// How a column's values are read as u64 lane words (src/lib.rs:389-403): the
// kernels widen on load, so a column without nulls is read in place and a
// column with nulls is compacted in its own element width.
enum { QA_LK_U64 = 0, QA_LK_I32, QA_LK_U32, QA_LK_F32, QA_LK_I8, QA_LK_I16, QA_LK_U8, QA_LK_U16 };

struct DColState {
    const void *lanep;    // lane base: the input itself, or the compacted copy
    uint32_t lkind;       // QA_LK_*
    uint32_t lwidth;      // bytes per element
    uint64_t nv;          // compacted (valid) value count
    uint64_t vblob_len;   // validity blob bytes (0 without nulls)
    uint64_t vblob_rle;   // RLE form bytes (without tag)
    uint64_t frames_off;  // first frame offset within the stream
    uint64_t stream_len;  // total stream bytes
    uint32_t has_nulls;
    uint32_t nblocks;
    uint32_t err;
    uint32_t pad;
};

// candidate slots of the competition, in the reference's consider order
enum { C_RAW = 0, C_CONST, C_STRIDE, C_XORZ, C_FOR, C_DELTA, C_FM, C_ALP, C_ALPRD, C_DICT, C_RLE,
       C_N };

struct DBlock {
    uint64_t start;        // value offset in the column's lane
    uint64_t payload_len;  // winner payload bytes
    uint64_t frame_off;    // frame offset within the column stream
    uint64_t v0, v1;       // first two values (CONST / STRIDE / DELTA base)
    uint64_t sz[C_N];      // candidate payload sizes (UINT64_MAX = bailed)
    uint64_t score;        // score of the current winner (size + decode penalty)
    uint32_t len;          // values; 0 = no block starts at this slot
    uint32_t col;
    uint32_t ntiles;
    uint32_t mode;
    uint32_t exp_range;
    uint8_t distinct_low, repeats, incompressible, compressible;
};

// Per-tile statistics. They are computed before block planning, so every
// quantity that depends on the block boundary is kept split: the tile's first
// element is recorded on its own and the block decision re-applies the block
// start rules (delta base, XORZ predecessor 0, CONST/STRIDE anchors).
struct DTile {
    uint64_t umin, umax;
    int64_t smin, smax;
    uint64_t dzmin, dzmax;   // zig-zag deltas of elements 1..cnt-1 (delta_bitpack.rs:32-46)
    uint64_t dz_first;       // zig-zag delta of element 0 against the column predecessor
    uint64_t v_first, v_last, step;
    uint64_t xorz_rest;      // varint bytes of v_i ^ v_{i-1}, i = 1..cnt-1 (xorz.rs:8-16)
    uint64_t off;            // sub-block offset (encode: in the payload; decode: in the stream)
    uint64_t tail;           // decode: delta prefix carried into the tile
    uint32_t chg_rest;       // i in 1..cnt-1 with v_i != v_{i-1}
    uint32_t exp_min, exp_max;
    uint32_t cnt;
    uint32_t w;
    uint8_t arith, allsame, active, pad;
};

// Per-tile statistics of the float-value codecs (Float family, Level::Fast).
struct DTileF {
    int64_t kmin[7], kmax[7];  // FloatMult k = round(x * s) per scale (float_mult.rs:45-64)
    int64_t dmin;              // ALP digit minimum (filler included)
    uint64_t rd_cmin, rd_cmax; // ALP-RD codes range
    uint64_t rd_rmin, rd_rmax; // ALP-RD rights range
    uint32_t rd_exc;           // ALP-RD exceptions in the tile
    uint32_t rd_exc_bytes;     // their varint position bytes + 8 each
    uint32_t alp_exc;
    uint8_t fm_pass;           // bit s: every value round-trips at scale s
    uint8_t alp_e, alp_f, alp_w;
    uint8_t alp_bail;
    uint8_t fm_fail;           // bit s: some value does not round-trip at scale s
    uint8_t pad1, pad2;        // (scales in neither mask were not evaluated)
    uint64_t alp_off, rd_exc_off, rd_coff, rd_roff;  // emit offsets
    uint64_t x_off;            // XORZ byte offset of the tile
    uint64_t rle_base;         // runs starting before this tile inside its block
    // chunk slots (chunk c of a block lives in the block's tile slot c):
    uint64_t ch_min, ch_off;   // RLE length chunk / Dict value chunk: FoR min and offset
    uint64_t dv_min, dv_off;
    uint32_t ch_w, dv_w;
    uint32_t rd_wc, rd_wr;     // ALP-RD codes / rights widths
    uint32_t dict_w, fm_w;
    uint64_t dict_min;         // smallest code (rank) in the tile
};

// Probe partials of a full 32768-value region of an adaptive column, for the
// 131072-value block it may open or join: the hash buckets of every 32nd value
// (a 16384-bit set per region slot, stored separately), and over every 64th
// value after the region's first: min, max and the largest zig-zag step from
// the value before. v0 / dz0: the region's first value and its step.
struct DRegionG {
    uint64_t gmn, gmx, gdz, v0, dz0;
    uint32_t full, pad;
};

// Per-block state of the Level::Fast codecs.
struct DBlockX {
    uint64_t rd_dict[8];
    uint64_t rd_nexc, rd_excb, rd_clen, rd_rlen;
    uint64_t card;            // Dict cardinality (or > max when bailed)
    uint64_t dict_vlen;       // Dict values blob bytes
    uint64_t dict_clen;       // Dict codes blob bytes
    uint64_t nruns, rle_llen; // RLE runs and lengths blob bytes
    uint32_t rd_left, rd_dl, rd_bail;
    uint32_t fm_scale;        // first scale passing for every value (7 = none)
    uint32_t dict_state;      // 0 exact in LDS, 1 bailed, 2 exact in global scratch
    uint32_t dict_vtag;       // 0 raw values, 1 delta-bitpacked
    uint32_t rle_bail, pad;
};

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
#ifdef __HIPCC__
#include <hip/hip_runtime.h>

__device__ __forceinline__ uint64_t qa_zigzag(uint64_t n) {
    return (n << 1) ^ (uint64_t)((int64_t)n >> 63);
}
__device__ __forceinline__ uint64_t qa_unzigzag(uint64_t z) { return (z >> 1) ^ (0ull - (z & 1)); }
__device__ __forceinline__ uint32_t qa_width(uint64_t range) {
    return range == 0 ? 0u : 64u - (uint32_t)__clzll((long long)range);
}
__device__ __forceinline__ uint32_t qa_varint_len(uint64_t v) {
    // 1 + floor(bits/7) for v > 0
    uint32_t bits = v ? 64u - (uint32_t)__clzll((long long)v) : 1u;
    return (bits + 6) / 7;
}

// Unaligned little-endian loads from a byte stream. The stream buffer always
// carries >= 8 bytes of tail padding so the second aligned word is in bounds.
__device__ __forceinline__ uint32_t qa_ld32u(const uint8_t *p) {
    uintptr_t a = (uintptr_t)p;
    const uint32_t *w = (const uint32_t *)(a & ~(uintptr_t)3);
    uint32_t s = (uint32_t)(a & 3);
    uint32_t lo = w[0];
    if (s == 0) return lo;
    uint32_t hi = w[1];
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> (8 * s));
}
__device__ __forceinline__ uint64_t qa_ld64u(const uint8_t *p) {
    return (uint64_t)qa_ld32u(p) | ((uint64_t)qa_ld32u(p + 4) << 32);
}
__device__ __forceinline__ void qa_st8(uint8_t *p, uint8_t v) { *p = v; }
__device__ __forceinline__ void qa_st32u(uint8_t *p, uint32_t v) {
    p[0] = (uint8_t)v;
    p[1] = (uint8_t)(v >> 8);
    p[2] = (uint8_t)(v >> 16);
    p[3] = (uint8_t)(v >> 24);
}
__device__ __forceinline__ void qa_st64u(uint8_t *p, uint64_t v) {
    qa_st32u(p, (uint32_t)v);
    qa_st32u(p + 4, (uint32_t)(v >> 32));
}
// write a LEB128 varint (src/varint.rs:9-21); returns bytes written
__device__ __forceinline__ uint32_t qa_put_varint(uint8_t *p, uint64_t v) {
    uint32_t k = 0;
    for (;;) {
        uint8_t b = (uint8_t)(v & 0x7f);
        v >>= 7;
        if (v) {
            p[k++] = b | 0x80;
        } else {
            p[k++] = b;
            return k;
        }
    }
}

// f32 <-> f64 lane conversions as integer operations with x86 semantics
// (src/lib.rs:398 `f as f64`, src/codecs/raw.rs:16): exact widening, signaling
// NaNs are quieted, subnormals are preserved (no FTZ on any path).
__device__ __forceinline__ uint64_t qa_f32_to_lane(uint32_t b) {
    uint64_t sign = (uint64_t)(b >> 31) << 63;
    uint32_t e = (b >> 23) & 0xFF, m = b & 0x7FFFFF;
    if (e == 0xFF) {
        if (m) m |= 0x400000;  // quiet
        return sign | (0x7FFull << 52) | ((uint64_t)m << 29);
    }
    if (e == 0) {
        if (m == 0) return sign;
        // subnormal f32: value = m * 2^-149; normalize
        // leading one at bit p = 23 - lz; value = 1.f * 2^(p - 149)
        uint32_t lz = (uint32_t)__clz((int)m) - 8;
        uint64_t mm = (uint64_t)((m << lz) & 0x7FFFFFu);
        uint64_t ee = 897u - lz;
        return sign | (ee << 52) | (mm << 29);
    }
    return sign | ((uint64_t)(e - 127 + 1023) << 52) | ((uint64_t)m << 29);
}
// ---------------------------------------------------------------------------
// Cross-lane operations on DPP (data-parallel primitives of the GFX9 ISA):
// quad_perm / row_half_mirror / row_mirror combine inside a 16-lane row,
// row_bcast15 / row_bcast31 carry a row's last lane into the next row(s).
// Lanes a step does not update keep `old` (the operation's identity). Every
// lane of the wavefront must be active.
// ---------------------------------------------------------------------------
#define QA_DPP_QP_1032 0xb1  // quad_perm [1,0,3,2]
#define QA_DPP_QP_2301 0x4e  // quad_perm [2,3,0,1]
#define QA_DPP_ROW_SHR(n) (0x110 + (n))
#define QA_DPP_ROW_MIRROR 0x140
#define QA_DPP_ROW_HALF_MIRROR 0x141
#define QA_DPP_ROW_BCAST15 0x142
#define QA_DPP_ROW_BCAST31 0x143
#define QA_DPP_WAVE_SHR1 0x138  // lane i takes lane i - 1 (lane 0 keeps old)

template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ uint32_t qa_dpp32(uint32_t old, uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, ROWMASK, 0xF, false);
}
template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ uint64_t qa_dpp64(uint64_t old, uint64_t v) {
    const uint32_t lo = qa_dpp32<CTRL, ROWMASK>((uint32_t)old, (uint32_t)v);
    const uint32_t hi = qa_dpp32<CTRL, ROWMASK>((uint32_t)(old >> 32), (uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t qa_readlane64(uint64_t v, int lane) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}
// reduction over the wavefront, the result in every lane (read from lane 63)
template <typename Op>
__device__ __forceinline__ uint64_t qa_wave_reduce64(uint64_t v, Op op, uint64_t id) {
    v = op(v, qa_dpp64<QA_DPP_QP_1032>(id, v));
    v = op(v, qa_dpp64<QA_DPP_QP_2301>(id, v));
    v = op(v, qa_dpp64<QA_DPP_ROW_HALF_MIRROR>(id, v));
    v = op(v, qa_dpp64<QA_DPP_ROW_MIRROR>(id, v));
    v = op(v, qa_dpp64<QA_DPP_ROW_BCAST15, 0xA>(id, v));
    v = op(v, qa_dpp64<QA_DPP_ROW_BCAST31, 0xC>(id, v));
    return qa_readlane64(v, 63);
}
// inclusive prefix sum over the wavefront (lane order)
__device__ __forceinline__ uint64_t qa_wave_scan_add64(uint64_t v) {
    v += qa_dpp64<QA_DPP_ROW_SHR(1)>(0, v);
    v += qa_dpp64<QA_DPP_ROW_SHR(2)>(0, v);
    v += qa_dpp64<QA_DPP_ROW_SHR(4)>(0, v);
    v += qa_dpp64<QA_DPP_ROW_SHR(8)>(0, v);
    v += qa_dpp64<QA_DPP_ROW_BCAST15, 0xA>(0, v);
    v += qa_dpp64<QA_DPP_ROW_BCAST31, 0xC>(0, v);
    return v;
}

// inclusive prefix maximum over the wavefront (lane order), unsigned
__device__ __forceinline__ uint32_t qa_wave_scan_max32(uint32_t v) {
    v = max(v, qa_dpp32<QA_DPP_ROW_SHR(1)>(0, v));
    v = max(v, qa_dpp32<QA_DPP_ROW_SHR(2)>(0, v));
    v = max(v, qa_dpp32<QA_DPP_ROW_SHR(4)>(0, v));
    v = max(v, qa_dpp32<QA_DPP_ROW_SHR(8)>(0, v));
    v = max(v, qa_dpp32<QA_DPP_ROW_BCAST15, 0xA>(0, v));
    v = max(v, qa_dpp32<QA_DPP_ROW_BCAST31, 0xC>(0, v));
    return v;
}

// index of the calling wavefront in its workgroup, known to be uniform (so
// what it indexes is read with scalar loads and scalar address arithmetic)
__device__ __forceinline__ int qa_wave_id() {
    return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// LDS written by one wavefront and read back by its other lanes
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// lane word i of a column stored as kind k
__device__ __forceinline__ uint64_t qa_lane_get(const void *p, uint32_t k, uint64_t i) {
    switch (k) {
    case QA_LK_I32: return (uint64_t)(int64_t)((const int32_t *)p)[i];
    case QA_LK_U32: return (uint64_t)((const uint32_t *)p)[i];
    case QA_LK_F32: return qa_f32_to_lane(((const uint32_t *)p)[i]);
    case QA_LK_I8: return (uint64_t)(int64_t)((const int8_t *)p)[i];
    case QA_LK_I16: return (uint64_t)(int64_t)((const int16_t *)p)[i];
    case QA_LK_U8: return (uint64_t)((const uint8_t *)p)[i];
    case QA_LK_U16: return (uint64_t)((const uint16_t *)p)[i];
    default: return ((const uint64_t *)p)[i];
    }
}

// narrowing of a lane that came from an f32 (exact)
__device__ __forceinline__ uint32_t qa_lane_to_f32(uint64_t x) {
    uint32_t sign = (uint32_t)(x >> 63) << 31;
    uint32_t e = (uint32_t)(x >> 52) & 0x7FF;
    uint64_t m = x & 0xFFFFFFFFFFFFFull;
    if (e == 0x7FF) {
        if (m) return sign | 0x7F800000u | 0x400000u | (uint32_t)(m >> 29);
        return sign | 0x7F800000u;
    }
    if (e == 0) return sign;  // f64 subnormals are far below the f32 range
    int32_t ee = (int32_t)e - 1023 + 127;
    if (ee >= 0xFF) return sign | 0x7F800000u;
    if (ee <= 0) {
        // f32 subnormal: exact when the lane came from an f32
        uint64_t full = m | (1ull << 52);
        uint32_t sh = (uint32_t)(1 - ee) + 29;
        return sign | (uint32_t)(full >> sh);
    }
    return sign | ((uint32_t)ee << 23) | (uint32_t)(m >> 29);
}
#endif
