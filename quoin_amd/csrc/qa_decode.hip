// qa_decode.hip — gfx950 kernels decoding a batch of quoin streams.
//
// The host scans the 9-byte frame headers (src/decoder.rs:65-94) and hands the
// frame table over; everything sized by the data runs here:
//   1 index: walk each frame's sub-block headers (widths), validate bounds
//   2 DeltaBitpack: per-tile delta sums, then a per-frame carry scan
//     (src/codecs/delta_bitpack.rs:48-64: a wrapping prefix sum)
//   3 per-tile unpack (FastLanes, src/bitpack.rs:78-161) + frame-of-reference
//   4 scatter compact values to logical positions for streams with nulls
//     (src/validity.rs:130-146) and narrow to the column type (src/lib.rs:492-506)
#include "qa_kernels.h"

#define WAVE 64

__device__ __forceinline__ uint64_t lane_bytes_d(uint32_t dt) {
    return (dt == QA_I32 || dt == QA_U32 || dt == QA_F32) ? 4 : 8;
}

// value k of a FastLanes block packed at byte address p with width w
__device__ __forceinline__ uint64_t unpack_one(const uint8_t *p, uint32_t w, uint32_t k) {
    if (w == 0) return 0;
    if (w <= 32) {
        uint32_t r = k >> 5, l = k & 31;
        uint32_t bit = r * w, word = bit >> 5, sh = bit & 31;
        uint32_t lo = qa_ld32u(p + 4ull * (word * 32 + l));
        uint32_t mask = w >= 32 ? 0xFFFFFFFFu : ((1u << w) - 1);
        uint32_t v = lo >> sh;
        if (sh + w > 32) v |= qa_ld32u(p + 4ull * ((word + 1) * 32 + l)) << (32 - sh);
        return v & mask;
    }
    uint32_t r = k >> 4, l = k & 15;
    uint32_t bit = r * w, word = bit >> 6, sh = bit & 63;
    uint64_t lo = qa_ld64u(p + 8ull * (word * 16 + l));
    uint64_t mask = w >= 64 ? ~0ull : ((1ull << w) - 1);
    uint64_t v = lo >> sh;
    if (sh + w > 64) v |= qa_ld64u(p + 8ull * ((word + 1) * 16 + l)) << (64 - sh);
    return v & mask;
}

__device__ __forceinline__ int frame_by_tile(const DecArgs &a, uint64_t t) {
    int lo = 0, hi = (int)a.nframes - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (a.frames[mid].tile_base <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// stage 1: sub-block index (one thread per frame)
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool rd_varint(const uint8_t *p, uint64_t end, uint64_t *pos,
                                          uint64_t *out) {
    uint64_t r = 0;
    uint32_t sh = 0;
    for (;;) {
        if (*pos >= end) return false;
        uint8_t b = p[(*pos)++];
        r |= (uint64_t)(b & 0x7f) << sh;
        if (!(b & 0x80)) {
            *out = r;
            return true;
        }
        sh += 7;
        if (sh >= 64) return false;
    }
}

__global__ void k_dec_index(DecArgs a) {
    uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= a.nframes) return;
    const DFrame &fr = a.frames[f];
    const DDecCol &col = a.cols[fr.col];
    const uint8_t *s = col.src;
    uint64_t pos = fr.payload_off, end = fr.payload_off + fr.plen;
    uint32_t err = 0;
    switch (fr.mode) {
    case QA_M_RAW:
        if (fr.plen != fr.n * lane_bytes_d(col.dtype)) err = QA_ERR_CORRUPT;
        break;
    case QA_M_CONST:
        if (fr.plen != 8) err = QA_ERR_CORRUPT;
        break;
    case QA_M_STRIDE:
        if (fr.plen != 16) err = QA_ERR_CORRUPT;
        break;
    case QA_M_FOR:
    case QA_M_DELTA: {
        if (fr.mode == QA_M_DELTA) {
            if (fr.plen < 8) {
                err = QA_ERR_CORRUPT;
                break;
            }
            pos += 8;
        }
        uint64_t n;
        if (!rd_varint(s, end, &pos, &n) || n != fr.n) {
            err = QA_ERR_CORRUPT;
            break;
        }
        for (uint32_t t = 0; t < fr.ntiles; t++) {
            if (pos + 9 > end) {
                err = QA_ERR_CORRUPT;
                break;
            }
            uint32_t w = s[pos];
            if (w > 64) {
                err = QA_ERR_CORRUPT;
                break;
            }
            DTile &ts = a.tiles[fr.tile_base + t];
            ts.w = w;
            ts.off = pos;
            pos += 9;
            if (w) {
                if (pos + 128ull * w > end) {
                    err = QA_ERR_CORRUPT;
                    break;
                }
                pos += 128ull * w;
            }
        }
        break;
    }
    default: err = QA_ERR_UNSUPPORTED;
    }
    if (err) atomicOr(a.err, err);
}

// ---------------------------------------------------------------------------
// stage 2: DeltaBitpack carries
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wsum(T v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += (T)__shfl_xor(v, d, 64);
    return v;
}

__global__ void __launch_bounds__(256) k_dec_delta_sums(DecArgs a) {
    if (*a.err) return;
    const int lane = threadIdx.x & 63;
    uint64_t t = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= a.ntiles) return;
    int f = frame_by_tile(a, t);
    const DFrame &fr = a.frames[f];
    if (fr.mode != QA_M_DELTA) return;
    const DDecCol &col = a.cols[fr.col];
    DTile &ts = a.tiles[t];
    uint32_t lt = (uint32_t)(t - fr.tile_base);
    uint32_t cnt = min((uint32_t)QA_TILE, fr.n - lt * QA_TILE);
    const uint8_t *h = col.src + ts.off;
    uint64_t mn = qa_ld64u(h + 1);
    uint64_t s = 0;
    for (uint32_t k = lane; k < cnt; k += WAVE) s += qa_unzigzag(mn + unpack_one(h + 9, ts.w, k));
    s = wsum(s);
    if (lane == 0) ts.tail = s;
}

__global__ void k_dec_delta_scan(DecArgs a) {
    uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= a.nframes || *a.err) return;
    const DFrame &fr = a.frames[f];
    if (fr.mode != QA_M_DELTA) return;
    uint64_t acc = qa_ld64u(a.cols[fr.col].src + fr.payload_off);  // base
    for (uint32_t t = 0; t < fr.ntiles; t++) {
        DTile &ts = a.tiles[fr.tile_base + t];
        uint64_t s = ts.tail;
        ts.tail = acc;
        acc += s;
    }
}

// ---------------------------------------------------------------------------
// stage 3: per-tile decode
// ---------------------------------------------------------------------------
__device__ __forceinline__ void store_typed(const DDecCol &col, uint64_t idx, uint64_t x) {
    switch (col.out_kind) {
    case QA_IN_I8:
    case QA_IN_U8: ((uint8_t *)col.out)[idx] = (uint8_t)x; return;
    case QA_IN_I16:
    case QA_IN_U16: ((uint16_t *)col.out)[idx] = (uint16_t)x; return;
    default: break;
    }
    switch (col.dtype) {
    case QA_F32: ((uint32_t *)col.out)[idx] = qa_lane_to_f32(x); break;
    case QA_I32:
    case QA_U32: ((uint32_t *)col.out)[idx] = (uint32_t)x; break;
    default: ((uint64_t *)col.out)[idx] = x;
    }
}

__device__ __forceinline__ void put_value(const DDecCol &col, uint64_t *lane_scratch,
                                          uint64_t idx, uint64_t x) {
    if (col.nv != col.n) {  // nulls: compact scratch, scattered later
        lane_scratch[col.lane_off + idx] = x;
        return;
    }
    store_typed(col, idx, x);
}

__device__ __forceinline__ uint64_t raw_value(const uint8_t *p, uint32_t dt, uint64_t i) {
    if (dt == QA_F32) return qa_f32_to_lane(qa_ld32u(p + 4 * i));
    if (lane_bytes_d(dt) == 4) return qa_ld32u(p + 4 * i);
    return qa_ld64u(p + 8 * i);
}

__global__ void __launch_bounds__(256) k_dec_tiles(DecArgs a) {
    if (*a.err) return;
    const int lane = threadIdx.x & 63;
    uint64_t t = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= a.ntiles) return;
    int f = frame_by_tile(a, t);
    const DFrame &fr = a.frames[f];
    const DDecCol &col = a.cols[fr.col];
    const DTile &ts = a.tiles[t];
    uint32_t lt = (uint32_t)(t - fr.tile_base);
    uint32_t cnt = min((uint32_t)QA_TILE, fr.n - lt * QA_TILE);
    uint64_t base = fr.out_off + (uint64_t)lt * QA_TILE;
    const uint8_t *pl = col.src + fr.payload_off;
    switch (fr.mode) {
    case QA_M_RAW:
        for (uint32_t k = lane; k < cnt; k += WAVE)
            put_value(col, a.lane, base + k, raw_value(pl, col.dtype, (uint64_t)lt * QA_TILE + k));
        break;
    case QA_M_CONST: {
        uint64_t v = qa_ld64u(pl);
        for (uint32_t k = lane; k < cnt; k += WAVE) put_value(col, a.lane, base + k, v);
        break;
    }
    case QA_M_STRIDE: {
        uint64_t v0 = qa_ld64u(pl), st = qa_ld64u(pl + 8);
        for (uint32_t k = lane; k < cnt; k += WAVE)
            put_value(col, a.lane, base + k, v0 + ((uint64_t)lt * QA_TILE + k) * st);
        break;
    }
    case QA_M_FOR: {
        const uint8_t *h = col.src + ts.off;
        uint64_t mn = qa_ld64u(h + 1);
        for (uint32_t k = lane; k < cnt; k += WAVE)
            put_value(col, a.lane, base + k, mn + unpack_one(h + 9, ts.w, k));
        break;
    }
    case QA_M_DELTA: {
        // each lane owns 16 consecutive values: local sums, wave scan, replay
        const uint8_t *h = col.src + ts.off;
        uint64_t mn = qa_ld64u(h + 1);
        uint32_t k0 = lane * 16;
        uint64_t d[16];
        uint64_t s = 0;
#pragma unroll
        for (int j = 0; j < 16; j++) {
            uint32_t k = k0 + j;
            d[j] = k < cnt ? qa_unzigzag(mn + unpack_one(h + 9, ts.w, k)) : 0;
            s += d[j];
        }
        uint64_t inc = s;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            uint64_t y = __shfl_up(inc, o, 64);
            if (lane >= o) inc += y;
        }
        uint64_t acc = ts.tail + inc - s;
#pragma unroll
        for (int j = 0; j < 16; j++) {
            uint32_t k = k0 + j;
            acc += d[j];
            if (k < cnt) put_value(col, a.lane, base + k, acc);
        }
        break;
    }
    default: break;
    }
}

// ---------------------------------------------------------------------------
// stage 4: nulls — rank prefix over the bitmap, scatter + narrow
// ---------------------------------------------------------------------------
__device__ __forceinline__ int dcol_by_vword(const DecArgs &a, uint64_t j) {
    int lo = 0, hi = (int)a.ncols - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (a.cols[mid].vword_base <= j) lo = mid; else hi = mid - 1;
    }
    return lo;
}
__device__ __forceinline__ uint64_t dbm_word(const uint8_t *bm, uint64_t n, uint64_t j) {
    // decoded bitmaps are library-owned and padded to whole 64-bit words
    uint64_t w = ((const uint64_t *)bm)[j];
    uint64_t bits = n - j * 64;
    if (bits < 64) w &= (1ull << bits) - 1;
    return w;
}

__global__ void k_dec_vcount(DecArgs a) {
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.nvwords) return;
    int c = dcol_by_vword(a, j);
    const DDecCol &col = a.cols[c];
    a.vpre[j] = (uint64_t)__popcll(dbm_word(col.bitmap, col.n, j - col.vword_base));
}

__global__ void k_dec_scatter(DecArgs a) {
    if (*a.err) return;
    uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.nvals) return;
    int lo = 0, hi = (int)a.ncols - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (a.cols[mid].val_base <= g) lo = mid; else hi = mid - 1;
    }
    const DDecCol &col = a.cols[lo];
    uint64_t i = g - col.val_base;
    if (i >= col.n) return;
    if (col.validity_out && (i & 7) == 0) {
        uint8_t byte;
        if (col.bitmap) {
            byte = col.bitmap[i >> 3];
        } else {
            byte = 0xFF;
        }
        uint64_t rem = col.n - i;
        if (rem < 8) byte &= (uint8_t)((1u << rem) - 1);
        col.validity_out[i >> 3] = byte;
    }
    if (col.nv == col.n) return;  // decoded in place
    uint64_t j = i >> 6;
    uint64_t w = dbm_word(col.bitmap, col.n, j);
    uint32_t b = (uint32_t)(i & 63);
    uint64_t x = 0;
    if ((w >> b) & 1) {
        uint64_t rank = a.vpre[col.vword_base + j] - a.vpre[col.vword_base] +
                        (uint64_t)__popcll(w & ((1ull << b) - 1));
        x = a.lane[col.lane_off + rank];
    }
    store_typed(col, i, x);
}

static inline unsigned nbk(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }
#define QA_LAUNCH(k, grid, ...)                                   \
    do {                                                          \
        dim3 g_ = (grid);                                         \
        if (g_.x) hipLaunchKernelGGL(k, g_, __VA_ARGS__);         \
    } while (0)

void qa_decode_launch(const DecArgs &a, hipStream_t s) {
    if (a.nframes) {
        QA_LAUNCH(k_dec_index, dim3(nbk(a.nframes, 64)), dim3(64), 0, s, a);
        QA_LAUNCH(k_dec_delta_sums, dim3(nbk(a.ntiles, 4)), dim3(256), 0, s, a);
        QA_LAUNCH(k_dec_delta_scan, dim3(nbk(a.nframes, 64)), dim3(64), 0, s, a);
        QA_LAUNCH(k_dec_tiles, dim3(nbk(a.ntiles, 4)), dim3(256), 0, s, a);
    }
    if (a.nvwords) {
        QA_LAUNCH(k_dec_vcount, dim3(nbk(a.nvwords, 256)), dim3(256), 0, s, a);
        qa_exclusive_scan(a.vpre, a.nvwords, a.vpre, a.scan_tmp, s);
    }
    if (a.nvals) QA_LAUNCH(k_dec_scatter, dim3(nbk(a.nvals, 256)), dim3(256), 0, s, a);
}
