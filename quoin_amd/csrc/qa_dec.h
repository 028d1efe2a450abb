// qa_dec.h — device helpers shared by the decode kernel files.
#pragma once
#include "qa_kernels.h"

#define WAVE 64

__device__ __forceinline__ bool fast_mode(uint32_t m) {
    return m == QA_M_XORZ || m == QA_M_FLOAT_MULT || m == QA_M_ALP || m == QA_M_ALP_RD ||
           m == QA_M_DICT || m == QA_M_RLE;
}
__device__ __forceinline__ uint64_t lane_bytes_d(uint32_t dt) {
    return (dt == QA_I32 || dt == QA_U32 || dt == QA_F32) ? 4 : 8;
}

#define QA_SEARCH(cnt, keyexpr, key)                              \
    int lo = 0, hi = (int)(cnt) - 1;                              \
    while (lo < hi) {                                             \
        int mid = (lo + hi + 1) >> 1;                             \
        if ((keyexpr) <= (key)) lo = mid; else hi = mid - 1;      \
    }                                                             \
    return lo;

__device__ __forceinline__ int frame_by_tile(const DecArgs &a, uint64_t t) {
    return (int)a.tiles[t].frame;
}
__device__ __forceinline__ int dcol_by_vword(const DecArgs &a, uint64_t j) {
    QA_SEARCH(a.ncols, a.cols[mid].vword_base, j)
}
__device__ __forceinline__ int dcol_by_vbyte(const DecArgs &a, uint64_t g) {
    QA_SEARCH(a.ncols, a.cols[mid].vbyte_base, g)
}
__device__ __forceinline__ int dcol_by_avbyte(const DecArgs &a, uint64_t g) {
    QA_SEARCH(a.ncols, a.cols[mid].avbyte_base, g)
}
__device__ __forceinline__ bool frame_used(const DecArgs &a, uint32_t f) {
    const DFrame &fr = a.frames[f];
    return f < a.dst[fr.col].fused_end;
}

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

// ---------------------------------------------------------------------------
// LDS staging of packed bytes: nwords u32 words starting at an arbitrary byte
// address, realigned. The aligned word holding the last needed byte is inside
// the allocation (aligned words never cross it).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void stage_words(const uint8_t *src, uint32_t nbytes, uint32_t *lds,
                                            int lane) {
    uintptr_t ad = (uintptr_t)src;
    const uint32_t *al = (const uint32_t *)(ad & ~(uintptr_t)3);
    const uint32_t m = (uint32_t)(ad & 3);
    const uint32_t nwords = (nbytes + 3) / 4;
    // aligned words that hold bytes [src, src + nbytes)
    const uint32_t nal = (m + nbytes + 3) / 4;
    // eight loads in flight per lane before the first LDS store
    for (uint32_t i0 = 0; i0 < nwords; i0 += 8 * WAVE) {
        uint32_t lo[8], hi[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const uint32_t i = i0 + u * WAVE + lane;
            lo[u] = i < nwords ? al[i] : 0;
            hi[u] = (m && i + 1 < nal) ? al[i + 1] : 0;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const uint32_t i = i0 + u * WAVE + lane;
            if (i < nwords)
                lds[i] = m ? (uint32_t)((((uint64_t)hi[u] << 32) | lo[u]) >> (8 * m)) : lo[u];
        }
    }
}
// FoR sub-block value k at stream sub-block header h (direct global reads;
// sub-block offsets are validated by the index walk), src/for_bitpack.rs:108-157
__device__ __forceinline__ uint64_t for_value_direct(const uint8_t *h, uint32_t k) {
    uint32_t w = h[0];
    uint64_t mn = qa_ld64u(h + 1);
    if (w == 0) return mn;
    if (w <= 32) {
        uint32_t r = k >> 5, l = k & 31, bit = r * w, word = bit >> 5, sh = bit & 31;
        uint32_t v = qa_ld32u(h + 9 + 4ull * (word * 32 + l)) >> sh;
        if (sh + w > 32) v |= qa_ld32u(h + 9 + 4ull * ((word + 1) * 32 + l)) << (32 - sh);
        return mn + (v & (w >= 32 ? 0xFFFFFFFFu : ((1u << w) - 1)));
    }
    uint32_t r = k >> 4, l = k & 15, bit = r * w, word = bit >> 6, sh = bit & 63;
    uint64_t v = qa_ld64u(h + 9 + 8ull * (word * 16 + l)) >> sh;
    if (sh + w > 64) v |= qa_ld64u(h + 9 + 8ull * ((word + 1) * 16 + l)) << (64 - sh);
    return mn + (v & (w >= 64 ? ~0ull : ((1ull << w) - 1)));
}
__device__ __forceinline__ uint64_t unpack_lds(const uint32_t *wd, uint32_t w, uint32_t k) {
    if (w == 0) return 0;
    if (w <= 32) {
        uint32_t r = k >> 5, l = k & 31;
        uint32_t bit = r * w, word = bit >> 5, sh = bit & 31;
        uint32_t mask = w >= 32 ? 0xFFFFFFFFu : ((1u << w) - 1);
        uint32_t v = wd[word * 32 + l] >> sh;
        if (sh + w > 32) v |= wd[(word + 1) * 32 + l] << (32 - sh);
        return v & mask;
    }
    uint32_t r = k >> 4, l = k & 15;
    uint32_t bit = r * w, word = bit >> 6, sh = bit & 63;
    uint64_t mask = w >= 64 ? ~0ull : ((1ull << w) - 1);
    uint32_t i0 = 2 * (word * 16 + l);
    uint64_t lo = ((uint64_t)wd[i0 + 1] << 32) | wd[i0];
    uint64_t v = lo >> sh;
    if (sh + w > 64) {
        uint32_t i1 = 2 * ((word + 1) * 16 + l);
        uint64_t hi = ((uint64_t)wd[i1 + 1] << 32) | wd[i1];
        v |= hi << (64 - sh);
    }
    return v & mask;
}

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

// bitmap word j of an all-null stream: its 64 values written as zeros (no
// tile writes them)
__device__ __forceinline__ void dec_nullfill_one(const DecArgs &a, uint64_t j) {
    if (j >= a.nvwords || *a.err) return;
    const int ci = dcol_by_vword(a, j);
    const DDecCol &c = a.cols[ci];
    if (a.dst[ci].nv != 0) return;
    uint64_t i0 = (j - c.vword_base) * 64;
    for (uint64_t i = i0; i < i0 + 64 && i < c.n; i++) store_typed(c, i, 0);
}
// validity_out byte g of the streams without a bitmap: all valid, trailing
// bits cleared (src/lib.rs:503-506)
__device__ __forceinline__ void dec_allvalid_one(const DecArgs &a, uint64_t g) {
    if (g >= a.navbytes || *a.err) return;
    const DDecCol &col = a.cols[dcol_by_avbyte(a, g)];
    uint64_t b = g - col.avbyte_base, rem = col.n - b * 8;
    col.validity_out[b] = rem >= 8 ? 0xFF : (uint8_t)((1u << rem) - 1);
}

// position of the r-th set bit (0-based) of w; r < popcount(w)
__device__ __forceinline__ uint32_t select64(uint64_t w, uint32_t r) {
    uint32_t pos = 0, c = (uint32_t)__popc((uint32_t)w);
    if (r >= c) { r -= c; w >>= 32; pos += 32; }
    c = (uint32_t)__popc((uint32_t)w & 0xFFFFu);
    if (r >= c) { r -= c; w >>= 16; pos += 16; }
    c = (uint32_t)__popc((uint32_t)w & 0xFFu);
    if (r >= c) { r -= c; w >>= 8; pos += 8; }
    for (int k = 0; k < 8; k++, pos++, w >>= 1)
        if (w & 1) {
            if (r == 0) return pos;
            r--;
        }
    return pos;
}

// Word w of a stream's bitmap, holding valid values [P, P + popcount): if the
// start of a tile of an aligned stream is among them, its logical position.
__device__ __forceinline__ void tile_start_in_word(const DecArgs &a, const DDecCol &col,
                                                   uint64_t w, uint64_t P, uint64_t word) {
    if (!col.tl_aligned || !word) return;
    const uint64_t c = (P + QA_TILE - 1) & ~(uint64_t)(QA_TILE - 1);
    const uint32_t cnt = (uint32_t)__popcll(word);
    if (c == 0 || c >= P + cnt) return;
    const uint64_t t = col.tile0 + c / QA_TILE;
    if (t < col.tile1) a.tlpos[t] = w * 64 + select64(word, (uint32_t)(c - P));
}

// logical position of the c-th valid value of a stream with nulls (c < nv),
// by a binary search of the words' valid prefix and a select in the word;
// n (and a flagged error) when the bitmap does not hold it
__device__ __forceinline__ uint64_t tile_lpos(const DecArgs &a, const DDecCol &col, uint64_t c) {
    if (c == 0) return 0;  // the first tile also takes the leading nulls
    if (col.tl_aligned) return a.tlpos[col.tile0 + c / QA_TILE];
    const uint64_t *vp = a.vpre + col.vword_base;
    uint64_t lo = 0, hi = col.nvwords - 1;
    while (lo < hi) {
        const uint64_t mid = (lo + hi + 1) >> 1;
        if (vp[mid] <= c) lo = mid; else hi = mid - 1;
    }
    const uint64_t r = c - vp[lo], w = col.bitmap[lo];
    if (r >= (uint64_t)__popcll(w)) {
        atomicOr(a.err, QA_ERR_CORRUPT);
        return col.n;
    }
    return lo * 64 + select64(w, (uint32_t)r);
}


// Output element of the column (its narrow Arrow type, src/lib.rs:492-506).
__device__ __forceinline__ int out_form(const DDecCol &col) {
    return (int)qa_out_form(col.dtype, col.out_kind);
}
// Decoded values are written once and not read back by the decoder: stores
// are non-temporal (streaming) so they do not displace the payloads in L2.
template <int F>
__device__ __forceinline__ void st_out(void *out, uint64_t i, uint64_t x) {
    if (F == QA_OUT_U8) __builtin_nontemporal_store((uint8_t)x, (uint8_t *)out + i);
    else if (F == QA_OUT_U16) __builtin_nontemporal_store((uint16_t)x, (uint16_t *)out + i);
    else if (F == QA_OUT_U32) __builtin_nontemporal_store((uint32_t)x, (uint32_t *)out + i);
    else if (F == QA_OUT_F32) __builtin_nontemporal_store(qa_lane_to_f32(x), (uint32_t *)out + i);
    else __builtin_nontemporal_store(x, (uint64_t *)out + i);
}

// Tile values x[q] (value k = lane + 64 q) of a stream without nulls go to
// out[base + k].
template <int F>
__device__ __forceinline__ void emit_dense_f(void *out, uint64_t base, uint32_t cnt,
                                             const uint64_t (&x)[QA_TILE / WAVE], int lane) {
#pragma unroll
    for (int q = 0; q < QA_TILE / WAVE; q++) {
        uint32_t k = lane + WAVE * q;
        if (k < cnt) st_out<F>(out, base + k, x[q]);
    }
}
__device__ __forceinline__ void emit_dense(void *out, uint32_t oform, uint64_t base, uint32_t cnt,
                                           const uint64_t (&x)[QA_TILE / WAVE], int lane) {
    switch (oform) {
    case QA_OUT_U8: emit_dense_f<QA_OUT_U8>(out, base, cnt, x, lane); break;
    case QA_OUT_U16: emit_dense_f<QA_OUT_U16>(out, base, cnt, x, lane); break;
    case QA_OUT_U32: emit_dense_f<QA_OUT_U32>(out, base, cnt, x, lane); break;
    case QA_OUT_F32: emit_dense_f<QA_OUT_F32>(out, base, cnt, x, lane); break;
    default: emit_dense_f<QA_OUT_U64>(out, base, cnt, x, lane); break;
    }
}

// Values vv[0..cnt) of a stream with nulls go to the set bits of the bitmap
// in [L0, L1) in order; the other positions get zero (src/validity.rs:130-146).
// The wavefront walks the 64-position windows of bitmap words (L0 rounded
// down), so every store is one aligned run of 64 outputs: lane i loads word i
// of the range once (64 windows per load), a window is one readlane (uniform)
// with the positions outside [L0, L1) masked off, each lane's rank inside it
// is one mbcnt, and the count before the window carries in a scalar register.
// A rank past cnt (a bitmap with more set bits than values) stores zero.
// LDS positions of the tile decoders' transposed values (k_dec_fl): lane L
// writes its values 16 L .. 16 L + 15 as 16-byte chunks, chunk c at chunk slot
// c ^ (a few bits of L), so the lanes of one write spread over all banks. SW =
// 4 for 4-byte values (4 chunks per lane), 8 for 8-byte values (8 chunks), 0
// for an unswizzled array.
template <int SW>
__device__ __forceinline__ uint32_t lds_pos(uint32_t k) {
    if (SW == 4) return (k & ~15u) | ((((k >> 2) & 3u) ^ ((k >> 5) & 3u)) << 2) | (k & 3u);
    if (SW == 8) return (k & ~15u) | ((((k >> 1) & 7u) ^ ((k >> 4) & 7u)) << 1) | (k & 1u);
    return k;
}

template <int F, typename V, int SW = 0>
__device__ __forceinline__ void emit_gather_t(void *out, const uint64_t *bm, uint64_t L0,
                                              uint64_t L1, const V *vv, uint32_t cnt, int lane,
                                              bool have_first = false, uint64_t first = 0) {
    // have_first: the caller already loaded this lane's word of the first
    // window (gather_first_word), early, so its latency overlapped the unpack
    if (L1 <= L0) return;
    const uint64_t w0 = L0 >> 6, w1 = (L1 - 1) >> 6;  // inclusive word range
    // positions relative to the first window (a tile's range spans a few
    // thousand positions at most), so the per-window work is 32-bit
    const uint64_t base = w0 << 6;
    const uint32_t l0 = (uint32_t)(L0 - base), l1 = (uint32_t)(L1 - base);
    const uint32_t nw = (uint32_t)(w1 - w0 + 1);
    const uint64_t lbit = 1ull << lane;
    uint32_t R = 0;  // values before the window
    for (uint32_t c0 = 0; c0 < nw; c0 += WAVE) {
        const uint32_t wi = c0 + (uint32_t)lane;
        const uint64_t myw = c0 == 0 && have_first ? first : bm[w0 + (wi < nw ? wi : nw - 1)];
        const uint32_t nwin = nw - c0 < (uint32_t)WAVE ? nw - c0 : (uint32_t)WAVE;
        for (uint32_t q = 0; q < nwin; q++) {
            const uint32_t p = (c0 + q) << 6;
            uint64_t W = qa_readlane64(myw, (int)q);
            if (p < l0) W &= ~0ull << (l0 - p);
            if (l1 - p < 64) W &= (1ull << (l1 - p)) - 1;
            const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(W >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)W, 0u));
            const uint32_t idx = R + below;
            const uint64_t x = (W & lbit) && idx < cnt ? (uint64_t)vv[lds_pos<SW>(idx)] : 0;
            const uint32_t pos = p + (uint32_t)lane;
            if (pos >= l0 && pos < l1) st_out<F>(out, base + pos, x);
            R += (uint32_t)__popcll(W);
        }
    }
}
// this lane's bitmap word of the first 64-word window of [L0, L1), the load
// emit_gather_t makes first
__device__ __forceinline__ uint64_t gather_first_word(const uint64_t *bm, uint64_t L0, uint64_t L1,
                                                      int lane) {
    if (L1 <= L0) return 0;
    const uint64_t w0 = L0 >> 6, w1 = (L1 - 1) >> 6;
    const uint32_t nw = (uint32_t)(w1 - w0 + 1), wi = (uint32_t)lane;
    return bm[w0 + (wi < nw ? wi : nw - 1)];
}

template <int F>
__device__ __forceinline__ void emit_gather_f(void *out, const uint64_t *bm, uint64_t L0,
                                              uint64_t L1, const uint64_t *vv, uint32_t cnt,
                                              int lane) {
    emit_gather_t<F, uint64_t>(out, bm, L0, L1, vv, cnt, lane);
}
__device__ __forceinline__ void emit_gather(void *out, uint32_t oform, const uint64_t *bm,
                                            uint64_t L0, uint64_t L1, const uint64_t *vv,
                                            uint32_t cnt, int lane) {
    switch (oform) {
    case QA_OUT_U8: emit_gather_f<QA_OUT_U8>(out, bm, L0, L1, vv, cnt, lane); break;
    case QA_OUT_U16: emit_gather_f<QA_OUT_U16>(out, bm, L0, L1, vv, cnt, lane); break;
    case QA_OUT_U32: emit_gather_f<QA_OUT_U32>(out, bm, L0, L1, vv, cnt, lane); break;
    case QA_OUT_F32: emit_gather_f<QA_OUT_F32>(out, bm, L0, L1, vv, cnt, lane); break;
    default: emit_gather_f<QA_OUT_U64>(out, bm, L0, L1, vv, cnt, lane); break;
    }
}

// ---------------------------------------------------------------------------
// Lane-contiguous FastLanes unpack of the 32-lane layout (w <= 32,
// src/bitpack.rs:78-103): lane L gets the tile's values k = 16 L + i (i < 16).
// They all lie in row r = L >> 1, in FastLanes lanes 16 (L & 1) + i, so each is
// one funnel shift of the row-words j = (r w) >> 5 and j + 1 at the lane-wide
// shift (r w) & 31: two runs of 16 consecutive staged words per lane.
// Row-word j + 1 may lie past the packed words (read from the staging buffer,
// which is larger, and masked off).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void unpack16_u32(const uint32_t *wd, uint32_t w, int lane,
                                             uint32_t (&v)[16]) {
    if (w == 0) {
#pragma unroll
        for (int i = 0; i < 16; i++) v[i] = 0;
        return;
    }
    const uint32_t bit = (uint32_t)(lane >> 1) * w, j = bit >> 5, sh = bit & 31;
    const uint32_t mask = w >= 32 ? 0xFFFFFFFFu : ((1u << w) - 1);
    const uint4 *lo = (const uint4 *)(wd + j * 32 + (lane & 1) * 16);
    const uint4 *hi = (const uint4 *)(wd + (j + 1) * 32 + (lane & 1) * 16);
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const uint4 a = lo[c], b = hi[c];
        v[4 * c + 0] = __builtin_amdgcn_alignbit(b.x, a.x, sh) & mask;
        v[4 * c + 1] = __builtin_amdgcn_alignbit(b.y, a.y, sh) & mask;
        v[4 * c + 2] = __builtin_amdgcn_alignbit(b.z, a.z, sh) & mask;
        v[4 * c + 3] = __builtin_amdgcn_alignbit(b.w, a.w, sh) & mask;
    }
}

// The same for the 16-lane layout (32 < w <= 64, src/bitpack.rs:136-161): lane
// L gets row r = L, i.e. FastLanes lanes i of row r: 64-bit row-words
// j = (r w) >> 6 and j + 1 (read only when the value straddles, so never past
// the 16 w packed words) at the lane-wide shift (r w) & 63.
__device__ __forceinline__ void unpack16_u64(const uint32_t *wd, uint32_t w, int lane,
                                             uint64_t (&v)[16]) {
    const uint64_t *w64 = (const uint64_t *)wd;
    const uint32_t bit = (uint32_t)lane * w, j = bit >> 6, sh = bit & 63;
    const uint64_t mask = w >= 64 ? ~0ull : ((1ull << w) - 1);
    const bool straddle = sh + w > 64;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint64_t lo = w64[j * 16 + i];
        const uint64_t hi = straddle ? w64[(j + 1) * 16 + i] : 0;
        v[i] = (sh ? (lo >> sh) | (hi << (64 - sh)) : lo) & mask;
    }
}

// The same unpacks from a swizzled staging (stage_words_fl): a 32-bit row
// (w <= 32) keeps its eight 16-byte chunks at chunk ^ (row & 7), a 64-bit row
// (w > 32) its sixteen 8-byte words at word ^ (row & 15). The lanes of one
// read hold different rows, so they then reach different LDS banks (read in
// place, each lane's row-words sit 128 bytes apart and share banks).
__device__ __forceinline__ uint32_t fl_swz32(uint32_t i) {
    return (i & ~31u) | ((((i >> 2) & 7u) ^ ((i >> 5) & 7u)) << 2) | (i & 3u);
}
__device__ __forceinline__ uint32_t fl_swz64(uint32_t i) {  // i: 32-bit word index
    const uint32_t q = i >> 1, j = q >> 4;
    return (((j << 4) | ((q & 15u) ^ (j & 15u))) << 1) | (i & 1u);
}
// stage_words with each word at its swizzled position (W64: the 64-bit layout)
template <bool W64>
__device__ __forceinline__ void stage_words_fl(const uint8_t *src, uint32_t nbytes, uint32_t *lds,
                                               int lane) {
    uintptr_t ad = (uintptr_t)src;
    const uint32_t *al = (const uint32_t *)(ad & ~(uintptr_t)3);
    const uint32_t m = (uint32_t)(ad & 3);
    const uint32_t nwords = (nbytes + 3) / 4;
    const uint32_t nal = (m + nbytes + 3) / 4;
    for (uint32_t i0 = 0; i0 < nwords; i0 += 8 * WAVE) {
        uint32_t lo[8], hi[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const uint32_t i = i0 + u * WAVE + lane;
            lo[u] = i < nwords ? al[i] : 0;
            hi[u] = (m && i + 1 < nal) ? al[i + 1] : 0;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const uint32_t i = i0 + u * WAVE + lane;
            if (i < nwords)
                lds[W64 ? fl_swz64(i) : fl_swz32(i)] =
                    m ? (uint32_t)((((uint64_t)hi[u] << 32) | lo[u]) >> (8 * m)) : lo[u];
        }
    }
}
__device__ __forceinline__ void unpack16_u32_fl(const uint32_t *wd, uint32_t w, int lane,
                                                uint32_t (&v)[16]) {
    if (w == 0) {
#pragma unroll
        for (int i = 0; i < 16; i++) v[i] = 0;
        return;
    }
    const uint32_t bit = (uint32_t)(lane >> 1) * w, j = bit >> 5, sh = bit & 31;
    const uint32_t mask = w >= 32 ? 0xFFFFFFFFu : ((1u << w) - 1);
    const uint32_t h4 = ((uint32_t)lane & 1) * 4;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const uint4 a = *(const uint4 *)(wd + j * 32 + (((h4 + c) ^ (j & 7)) << 2));
        const uint4 b = *(const uint4 *)(wd + (j + 1) * 32 + (((h4 + c) ^ ((j + 1) & 7)) << 2));
        v[4 * c + 0] = __builtin_amdgcn_alignbit(b.x, a.x, sh) & mask;
        v[4 * c + 1] = __builtin_amdgcn_alignbit(b.y, a.y, sh) & mask;
        v[4 * c + 2] = __builtin_amdgcn_alignbit(b.z, a.z, sh) & mask;
        v[4 * c + 3] = __builtin_amdgcn_alignbit(b.w, a.w, sh) & mask;
    }
}
__device__ __forceinline__ void unpack16_u64_fl(const uint32_t *wd, uint32_t w, int lane,
                                                uint64_t (&v)[16]) {
    const uint64_t *w64 = (const uint64_t *)wd;
    const uint32_t bit = (uint32_t)lane * w, j = bit >> 6, sh = bit & 63;
    const uint64_t mask = w >= 64 ? ~0ull : ((1ull << w) - 1);
    const bool straddle = sh + w > 64;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint64_t lo = w64[j * 16 + ((uint32_t)i ^ (j & 15))];
        const uint64_t hi = straddle ? w64[(j + 1) * 16 + ((uint32_t)i ^ ((j + 1) & 15))] : 0;
        v[i] = (sh ? (lo >> sh) | (hi << (64 - sh)) : lo) & mask;
    }
}

// Values y[i] (value k = 16 lane + i) of a stream without nulls to
// out[base + k]: 16-byte stores when the lane's 16 values are all in the tile
// and the destination is aligned, element stores otherwise.
template <int F>
__device__ __forceinline__ void emit_dense16_f(void *out, uint64_t base, uint32_t cnt,
                                               const uint64_t (&y)[16], int lane) {
    const uint32_t k0 = 16u * (uint32_t)lane;
    if (F == QA_OUT_U32 || F == QA_OUT_F32) {
        uint32_t *o = (uint32_t *)out + base + k0;
        if (k0 + 16 <= cnt && ((uintptr_t)o & 15) == 0) {
            uint32_t u[16];
#pragma unroll
            for (int i = 0; i < 16; i++) u[i] = F == QA_OUT_F32 ? qa_lane_to_f32(y[i]) : (uint32_t)y[i];
#pragma unroll
            for (int c = 0; c < 4; c++)
                ((uint4 *)o)[c] = make_uint4(u[4 * c], u[4 * c + 1], u[4 * c + 2], u[4 * c + 3]);
            return;
        }
    } else if (F == QA_OUT_U64) {
        uint64_t *o = (uint64_t *)out + base + k0;
        if (k0 + 16 <= cnt && ((uintptr_t)o & 15) == 0) {
#pragma unroll
            for (int c = 0; c < 8; c++)
                ((uint4 *)o)[c] = make_uint4((uint32_t)y[2 * c], (uint32_t)(y[2 * c] >> 32),
                                             (uint32_t)y[2 * c + 1], (uint32_t)(y[2 * c + 1] >> 32));
            return;
        }
    }
#pragma unroll
    for (int i = 0; i < 16; i++)
        if (k0 + i < cnt) st_out<F>(out, base + k0 + i, y[i]);
}
__device__ __forceinline__ void emit_dense16(void *out, uint32_t oform, uint64_t base, uint32_t cnt,
                                             const uint64_t (&y)[16], int lane) {
    switch (oform) {
    case QA_OUT_U8: emit_dense16_f<QA_OUT_U8>(out, base, cnt, y, lane); break;
    case QA_OUT_U16: emit_dense16_f<QA_OUT_U16>(out, base, cnt, y, lane); break;
    case QA_OUT_U32: emit_dense16_f<QA_OUT_U32>(out, base, cnt, y, lane); break;
    case QA_OUT_F32: emit_dense16_f<QA_OUT_F32>(out, base, cnt, y, lane); break;
    default: emit_dense16_f<QA_OUT_U64>(out, base, cnt, y, lane); break;
    }
}

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v, int) { return qa_wave_scan_add64(v); }


__device__ __forceinline__ bool wave_any_d(bool p) { return __ballot(p) != 0; }

// ---------------------------------------------------------------------------
// Sub-block walks by one wavefront. Each sub-block header gives the position
// of the next one, so the format's walk is sequential. Every round the lanes
// test whether the next 64 sub-blocks repeat the current header's stride key
// (the width, and for ALP the exception count): the first lane whose header
// differs ends the confirmed run, and the walk resumes there. A run of equal
// widths (the usual case) costs one round trip per 64 sub-blocks. All lanes
// return the same value.
// ---------------------------------------------------------------------------

// for_bitpack payload [pos, end) that must hold n values (for_bitpack.rs:108-157).
// Records sub-block offsets (which = 0: .off and .w, which = 1: .tail). Returns
// the position after the walk or ~0 on a malformed payload.
static __device__ uint64_t for_walk_wave(const uint8_t *s, uint64_t pos, uint64_t end, uint64_t n,
                                         DTileD *tl, int which, int lane) {
    uint64_t m;
    if (!rd_varint(s, end, &pos, &m) || m != n) return ~0ull;
    const uint64_t nsub = (n + QA_TILE - 1) / QA_TILE;
    for (uint64_t t = 0; t < nsub;) {
        if (pos + 9 > end) return ~0ull;
        const uint32_t w0 = s[pos];
        if (w0 > 64) return ~0ull;
        const uint64_t stride = 9 + 128ull * w0;
        if (stride > end - pos) return ~0ull;
        const uint64_t nin = nsub - t < (uint64_t)WAVE ? nsub - t : (uint64_t)WAVE;
        const uint64_t pj = pos + (uint64_t)lane * stride;
        bool ok = false;
        if ((uint64_t)lane < nin && pj + stride <= end) ok = s[pj] == w0;
        const uint64_t inmask = nin >= 64 ? ~0ull : ((1ull << nin) - 1);
        const uint64_t fail = __ballot(!ok) & inmask;
        const uint64_t cnf = fail ? (uint64_t)__builtin_ctzll(fail) : nin;  // >= 1: lane 0 passes
        if ((uint64_t)lane < cnf) {
            if (which == 0) {
                tl[t + lane].off = pj;
                tl[t + lane].w = w0;
            } else {
                tl[t + lane].tail = pj;
            }
        }
        t += cnf;
        pos += cnf * stride;
    }
    return pos;
}

// ALP sub-blocks [e][f][exc u16][min i64][w u8][128 w B][exc x 10 B] of a
// payload holding ntiles of them (alp.rs:168-248). Returns the position after
// the walk or ~0.
__device__ __forceinline__ bool alp_hdr_ok(const uint8_t *h) {
    uint32_t e = h[0], f = h[1], w = h[12];
    return e <= 18 && f <= e && w <= 32;
}
static __device__ uint64_t alp_walk_wave(const uint8_t *s, uint64_t pos, uint64_t end,
                                         uint64_t ntiles, DTileD *tl, int lane) {
    for (uint64_t t = 0; t < ntiles;) {
        if (pos + 13 > end || !alp_hdr_ok(s + pos)) return ~0ull;
        const uint32_t w0 = s[pos + 12];
        const uint32_t x0 = s[pos + 2] | ((uint32_t)s[pos + 3] << 8);
        const uint64_t stride = 13 + 128ull * w0 + 10ull * x0;
        if (stride > end - pos) return ~0ull;
        const uint64_t nin = ntiles - t < (uint64_t)WAVE ? ntiles - t : (uint64_t)WAVE;
        const uint64_t pj = pos + (uint64_t)lane * stride;
        bool ok = false;
        if ((uint64_t)lane < nin && pj + stride <= end) {
            const uint8_t *h = s + pj;
            ok = h[12] == w0 && (h[2] | ((uint32_t)h[3] << 8)) == x0 && alp_hdr_ok(h);
        }
        const uint64_t inmask = nin >= 64 ? ~0ull : ((1ull << nin) - 1);
        const uint64_t fail = __ballot(!ok) & inmask;
        const uint64_t cnf = fail ? (uint64_t)__builtin_ctzll(fail) : nin;
        if ((uint64_t)lane < cnf) {
            tl[t + lane].off = pj;
            tl[t + lane].w = w0;
        }
        t += cnf;
        pos += cnf * stride;
    }
    return pos;
}
