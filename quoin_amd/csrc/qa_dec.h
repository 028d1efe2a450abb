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
    QA_SEARCH(a.nframes, a.frames[mid].tile_base, t)
}
__device__ __forceinline__ int dcol_by_vword(const DecArgs &a, uint64_t j) {
    QA_SEARCH(a.ncols, a.cols[mid].vword_base, j)
}
__device__ __forceinline__ int dcol_by_vbyte(const DecArgs &a, uint64_t g) {
    QA_SEARCH(a.ncols, a.cols[mid].vbyte_base, g)
}
__device__ __forceinline__ int dcol_by_val(const DecArgs &a, uint64_t g) {
    QA_SEARCH(a.ncols, a.cols[mid].val_base, g)
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
    uint32_t m = (uint32_t)(ad & 3);
    uint32_t nwords = (nbytes + 3) / 4;
    // aligned words that hold bytes [src, src + nbytes)
    uint32_t nal = (m + nbytes + 3) / 4;
    if (m == 0) {
        for (uint32_t i = lane; i < nwords; i += WAVE) lds[i] = al[i];
    } else {
        for (uint32_t i = lane; i < nwords; i += WAVE) {
            uint32_t lo = al[i];
            uint32_t hi = i + 1 < nal ? al[i + 1] : 0;
            lds[i] = (uint32_t)((((uint64_t)hi << 32) | lo) >> (8 * m));
        }
    }
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
// LDS written by one wavefront and read back by its other lanes
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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

__device__ __forceinline__ void put_value(const DDecCol &col, bool nulls, uint64_t *lane_scratch,
                                          uint64_t idx, uint64_t x) {
    if (nulls) {  // compact scratch, scattered to logical positions later
        lane_scratch[col.lane_off + idx] = x;
        return;
    }
    store_typed(col, idx, x);
}

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v, int lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint64_t y = __shfl_up(v, o, 64);
        if (lane >= o) v += y;
    }
    return v;
}


__device__ __forceinline__ bool wave_any_d(bool p) { return __ballot(p) != 0; }
