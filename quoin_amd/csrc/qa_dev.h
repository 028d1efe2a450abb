// qa_dev.h — device helpers shared by the encode kernel files: column lookups,
// lane access, wavefront reductions, LDS-to-global byte stores, the FastLanes
// pack, and the scoring rules of the competition.
#pragma once
#include "qa_kernels.h"

#define WAVE 64
#define NONE32 0xFFFFFFFFu
#define RF_DLOW 1u
#define RF_REP 2u
#define RF_INC 4u

// ---------------------------------------------------------------------------
// lookups and lane access
// ---------------------------------------------------------------------------
#define QA_COL_SEARCH(field, key)                                   \
    int lo = 0, hi = nc - 1;                                        \
    while (lo < hi) {                                               \
        int mid = (lo + hi + 1) >> 1;                               \
        if (c[mid].field <= (key)) lo = mid; else hi = mid - 1;     \
    }                                                               \
    return lo;

__device__ __forceinline__ int col_by_region(const DCol *c, int nc, uint32_t r) {
    QA_COL_SEARCH(region_base, r)
}
__device__ __forceinline__ int col_by_tile(const DCol *c, int nc, uint64_t t) {
    QA_COL_SEARCH(tile_base, t)
}
// zero-word columns share the next column's base and lose the tie
__device__ __forceinline__ int col_by_vword(const DCol *c, int nc, uint64_t j) {
    QA_COL_SEARCH(vword_base, j)
}
__device__ __forceinline__ uint64_t lane_bytes_of(uint32_t dt) {
    return (dt == QA_I32 || dt == QA_U32 || dt == QA_F32) ? 4 : 8;
}
__device__ __forceinline__ bool is_float(uint32_t dt) { return dt == QA_F64 || dt == QA_F32; }
__device__ __forceinline__ bool is_signed(uint32_t dt) { return dt == QA_I32 || dt == QA_I64; }
__device__ __forceinline__ uint64_t lane_at(const DColState &st, uint64_t i) {
    return qa_lane_get(st.lanep, st.lkind, i);
}
// lane word i with the storage kind fixed at compile time (QA_LK_ANY: by lk)
#define QA_LK_ANY 99
template <int LK>
__device__ __forceinline__ uint64_t lane_get_t(const void *p, uint32_t lk, uint64_t i) {
    if (LK == QA_LK_U64) return ((const uint64_t *)p)[i];
    if (LK == QA_LK_U32) return ((const uint32_t *)p)[i];
    if (LK == QA_LK_I32) return (uint64_t)(int64_t)((const int32_t *)p)[i];
    if (LK == QA_LK_F32) return qa_f32_to_lane(((const uint32_t *)p)[i]);
    return qa_lane_get(p, lk, i);
}

// lane_get_t through global-address-space pointers: the loads are global_load,
// not flat_load (no LDS aperture check, vmcnt only)
typedef const __attribute__((address_space(1))) uint32_t qa_gu32;
typedef const __attribute__((address_space(1))) int32_t qa_gi32;
typedef const __attribute__((address_space(1))) uint64_t qa_gu64;
typedef uint32_t qa_u32x4 __attribute__((ext_vector_type(4)));
typedef uint64_t qa_u64x2 __attribute__((ext_vector_type(2)));
template <int LK>
__device__ __forceinline__ uint64_t lane_get_g(const void *p, uint32_t lk, uint64_t i) {
    if (LK == QA_LK_U64) return ((const qa_gu64 *)p)[i];
    if (LK == QA_LK_U32) return ((const qa_gu32 *)p)[i];
    if (LK == QA_LK_I32) return (uint64_t)(int64_t)((const qa_gi32 *)p)[i];
    if (LK == QA_LK_F32) return qa_f32_to_lane(((const qa_gu32 *)p)[i]);
    return qa_lane_get(p, lk, i);
}

// bitmap word j of a column (bits >= n cleared); byte loads: the caller's
// bitmap is ceil(n/8) bytes with no padding or alignment promise
__device__ __forceinline__ uint64_t bm_word(const uint8_t *bm, uint64_t n, uint64_t j) {
    uint64_t nbytes = (n + 7) >> 3, b0 = j * 8;
    uint64_t w = 0;
    if (b0 + 8 <= nbytes) {
#pragma unroll
        for (int k = 0; k < 8; k++) w |= (uint64_t)bm[b0 + k] << (8 * k);
    } else {
        for (int k = 0; k < 8; k++)
            if (b0 + k < nbytes) w |= (uint64_t)bm[b0 + k] << (8 * k);
    }
    uint64_t bits = n - j * 64;
    if (bits < 64) w &= (1ull << bits) - 1;
    return w;
}

// wavefront reductions (DPP, result in every lane); every lane must be active
struct OpMinU {
    __device__ uint64_t operator()(uint64_t a, uint64_t b) const { return a < b ? a : b; }
    static constexpr uint64_t id = ~0ull;
};
struct OpMaxU {
    __device__ uint64_t operator()(uint64_t a, uint64_t b) const { return a > b ? a : b; }
    static constexpr uint64_t id = 0;
};
struct OpMinS {
    __device__ uint64_t operator()(uint64_t a, uint64_t b) const {
        return (int64_t)a < (int64_t)b ? a : b;
    }
    static constexpr uint64_t id = 0x7FFFFFFFFFFFFFFFull;
};
struct OpMaxS {
    __device__ uint64_t operator()(uint64_t a, uint64_t b) const {
        return (int64_t)a > (int64_t)b ? a : b;
    }
    static constexpr uint64_t id = 0x8000000000000000ull;
};
struct OpAdd {
    __device__ uint64_t operator()(uint64_t a, uint64_t b) const { return a + b; }
    static constexpr uint64_t id = 0;
};
struct OpAnd {
    __device__ uint64_t operator()(uint64_t a, uint64_t b) const { return a & b; }
    static constexpr uint64_t id = ~0ull;
};
template <typename T, typename Op>
__device__ __forceinline__ T wave_reduce(T v, Op op) {
    return (T)qa_wave_reduce64((uint64_t)(int64_t)v, op, Op::id);
}

// Copy nbytes held in LDS (word-aligned at 0) to an arbitrary global byte
// address: aligned dword stores inside, byte stores at the two edges only, so
// neighbouring writers' bytes are never touched. lds must hold one extra word.
__device__ __forceinline__ void wave_store_bytes(uint8_t *dst, const uint32_t *lds, uint32_t nbytes,
                                                 int lane) {
    const uint8_t *sb = (const uint8_t *)lds;
    uint32_t m = (uint32_t)((uintptr_t)dst & 3);
    uint32_t head = (4 - m) & 3;
    if (head > nbytes) head = nbytes;
    if ((uint32_t)lane < head) dst[lane] = sb[lane];
    uint32_t rem = nbytes - head, nw = rem >> 2;
    __attribute__((address_space(1))) uint32_t *dw =
        (__attribute__((address_space(1))) uint32_t *)(dst + head);
    // four words per lane per step: the LDS reads of a step overlap
    for (uint32_t i0 = 0; i0 < nw; i0 += 4 * WAVE) {
        uint32_t o[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t i = i0 + u * WAVE + lane;
            o[u] = 0;
            if (i < nw) {
                const uint64_t pair = ((uint64_t)(head ? lds[i + 1] : 0) << 32) | lds[i];
                o[u] = (uint32_t)(pair >> (8 * head));
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t i = i0 + u * WAVE + lane;
            if (i < nw) dw[i] = o[u];
        }
    }
    uint32_t tb = head + 4 * nw;
    if (tb + lane < nbytes) dst[tb + lane] = sb[tb + lane];
}

// FastLanes pack (src/bitpack.rs:47-133) from residuals in LDS into LDS words:
// output word (k, l) gathers the rows whose bit range meets word k of lane l
__device__ __forceinline__ void pack_lds(const uint64_t *res, uint32_t w, uint32_t *out, int lane) {
    if (w <= 32) {
        for (uint32_t idx = lane; idx < 32 * w; idx += WAVE) {
            uint32_t k = idx >> 5, l = idx & 31;
            uint32_t lo_bit = 32 * k;
            uint32_t r0 = lo_bit / w, r1 = min(31u, (lo_bit + 31) / w);
            uint32_t word = 0;
            for (uint32_t r = r0; r <= r1; r++) {
                uint32_t v = (uint32_t)res[r * 32 + l];
                int off = (int)(r * w) - (int)lo_bit;
                word |= off >= 0 ? (off < 32 ? v << off : 0u) : v >> (-off);
            }
            out[idx] = word;
        }
    } else {
        for (uint32_t idx = lane; idx < 16 * w; idx += WAVE) {
            uint32_t k = idx >> 4, l = idx & 15;
            uint32_t lo_bit = 64 * k;
            uint32_t r0 = lo_bit / w, r1 = min(63u, (lo_bit + 63) / w);
            uint64_t word = 0;
            for (uint32_t r = r0; r <= r1; r++) {
                uint64_t v = res[r * 16 + l];
                int off = (int)(r * w) - (int)lo_bit;
                word |= off >= 0 ? (off < 64 ? v << off : 0ull) : v >> (-off);
            }
            out[2 * idx] = (uint32_t)word;
            out[2 * idx + 1] = (uint32_t)(word >> 32);
        }
    }
}

__device__ __forceinline__ uint64_t weight_of(uint32_t mode) {
    switch (mode) {  // src/encoder.rs:79-89
    case QA_M_RAW: case QA_M_CONST: case QA_M_STRIDE: return 0;
    case QA_M_XORZ: case QA_M_FOR: case QA_M_DELTA: case QA_M_RLE: return 1;
    case QA_M_ALP: case QA_M_ALP_RD: case QA_M_DICT: return 2;
    default: return 6;  // FLOAT_MULT
    }
}
__device__ __forceinline__ uint64_t penalty_of(uint32_t mode, uint64_t lambda, uint64_t dec) {
    return (lambda * weight_of(mode) * dec) >> 8;  // src/encoder.rs:94-97
}

__device__ __forceinline__ uint64_t tile_for_range(const DTile &s, bool sg) {
    return sg ? (uint64_t)s.smax - (uint64_t)s.smin : s.umax - s.umin;
}
// zig-zag delta range of a tile inside a block (element 0 of the block has
// delta 0 against the base v0)
__device__ __forceinline__ void tile_dz(const DTile &s, bool block_first, uint64_t *mn,
                                        uint64_t *mx) {
    uint64_t f = block_first ? 0 : s.dz_first;
    *mn = s.cnt > 1 ? (s.dzmin < f ? s.dzmin : f) : f;
    *mx = s.cnt > 1 ? (s.dzmax > f ? s.dzmax : f) : f;
}

