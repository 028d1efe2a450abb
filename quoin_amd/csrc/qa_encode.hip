// qa_encode.hip — gfx950 kernels of the batched block competition (encode).
//
// Two passes over the values:
//   pass 1 (k_region_stats): one workgroup per region slot computes every
//     block-independent statistic of its 1024-value tiles and the region's
//     probe features (src/encoder.rs:696-750) and incompressibility samples
//     (src/encoder.rs:655-683).
//   pass 2 (k_emit_tiles): one wavefront per tile writes the winner's payload.
// Between them: block planning (src/encoder.rs:154-178), features of grown
// blocks, the per-block candidate sizes and the ordered Best::consider replay
// (src/encoder.rs:99-140, 275-469), the validity blob (src/validity.rs:40-70)
// and the frame layout (src/encoder.rs:184-214, 752-759).
#include "qa_dev.h"


// ---------------------------------------------------------------------------
// stage 1: validity popcounts / transitions, lane setup, compaction
// ---------------------------------------------------------------------------
__global__ void k_vcount(EncArgs a) {
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.nvwords) return;
    int c = col_by_vword(a.cols, a.ncols, j);
    const DCol &col = a.cols[c];
    uint64_t lj = j - col.vword_base;
    uint64_t w = bm_word(col.validity, col.n, lj);
    a.vpre[j] = (uint64_t)__popcll(w);
    // transitions for the validity RLE: runs alternate starting from a 0-run,
    // so bit -1 counts as 0 (src/validity.rs:40-60)
    uint64_t prev_top = lj ? (bm_word(col.validity, col.n, lj - 1) >> 63) : 0;
    uint64_t t = w ^ ((w << 1) | prev_top);
    uint64_t bits = col.n - lj * 64;
    if (bits < 64) t &= (1ull << bits) - 1;
    a.tpre[j] = (uint64_t)__popcll(t);
}

__device__ __forceinline__ uint32_t lkind_of(const DCol &col) {
    switch (col.in_kind) {
    case QA_IN_I8: return QA_LK_I8;
    case QA_IN_I16: return QA_LK_I16;
    case QA_IN_U8: return QA_LK_U8;
    case QA_IN_U16: return QA_LK_U16;
    default: break;
    }
    switch (col.dtype) {
    case QA_I32: return QA_LK_I32;
    case QA_U32: return QA_LK_U32;
    case QA_F32: return QA_LK_F32;
    default: return QA_LK_U64;
    }
}

// One thread per column: the column's state, written whole (every field a
// later stage accumulates into starts at zero), and the batch's mode-win
// counters cleared; no kernel before this one touches either.
__global__ void k_col_counts(EncArgs a) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < 64) a.wins[c] = 0;
    if (c >= (int)a.ncols) return;
    const DCol &col = a.cols[c];
    DColState st{};
    if (col.nvwords) {
        uint64_t nv = a.vpre[col.vword_base + col.nvwords] - a.vpre[col.vword_base];
        st.nv = nv;
        st.has_nulls = nv < col.n;
        a.colbase[c] = a.tpre[col.vword_base];
    } else {
        st.nv = col.n;
        st.has_nulls = 0;
        a.colbase[c] = a.nvwords ? a.tpre[col.vword_base] : 0;
    }
    st.lkind = lkind_of(col);
    st.lwidth = col.in_width;
    // without nulls the kernels read the caller's values in place
    st.lanep = st.has_nulls ? (const void *)((const uint8_t *)a.lane + col.lane_off * 8)
                            : col.values;
    a.cst[c] = st;
    if (c == (int)a.ncols - 1) {
        uint64_t T = a.nvwords ? a.tpre[a.nvwords] : 0;
        a.colbase[a.ncols] = T;
        a.colbase[a.ncols + 1] = T + a.ncols;  // run slots in use
    }
}

// compaction of the valid values, in the input's own element width
// (src/validity.rs:122-127). One wavefront per 64 bitmap words: lane i reads
// word i, its column and its rank base; the wavefront then moves the valid
// slots of 8 words per step (8 independent loads in flight per lane).
__device__ __forceinline__ uint64_t ld_w(const uint8_t *p, uint64_t i, uint32_t w) {
    switch (w) {
    case 1: return p[i];
    case 2: return ((const uint16_t *)p)[i];
    case 4: return ((const uint32_t *)p)[i];
    default: return ((const uint64_t *)p)[i];
    }
}
__device__ __forceinline__ void st_w(uint8_t *p, uint64_t i, uint32_t w, uint64_t v) {
    switch (w) {
    case 1: p[i] = (uint8_t)v; break;
    case 2: ((uint16_t *)p)[i] = (uint16_t)v; break;
    case 4: ((uint32_t *)p)[i] = (uint32_t)v; break;
    default: ((uint64_t *)p)[i] = v;
    }
}

// the 64 words of one wavefront when every word's column has element type T:
// 32 words' loads in flight before their stores (on this ISA loads and stores
// share one completion counter, so batching them is what keeps loads queued)
template <typename T>
__device__ __forceinline__ void compact_words(uint64_t w, uint64_t rank0, uint64_t pos0,
                                              uint64_t src, uint64_t dst, int lane) {
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int i0 = 0; i0 < WAVE; i0 += 32) {
        T v[32];
#pragma unroll
        for (int u = 0; u < 32; u++) {
            const int i = i0 + u;
            const uint64_t wi = qa_readlane64(w, i);
            const T *sp = (const T *)qa_readlane64(src, i);
            const uint64_t pos = qa_readlane64(pos0, i) + lane;
            v[u] = ((wi >> lane) & 1) ? sp[pos] : (T)0;
        }
#pragma unroll
        for (int u = 0; u < 32; u++) {
            const int i = i0 + u;
            const uint64_t wi = qa_readlane64(w, i);
            if ((wi >> lane) & 1) {
                T *dp = (T *)qa_readlane64(dst, i);
                dp[qa_readlane64(rank0, i) + (uint64_t)__popcll(wi & below)] = v[u];
            }
        }
    }
}

__global__ void __launch_bounds__(256) k_compact(EncArgs a) {
    const int lane = threadIdx.x & 63;
    const uint64_t j0 = ((uint64_t)blockIdx.x * 4 + qa_wave_id()) * WAVE;
    if (j0 >= a.nvwords) return;
    const uint64_t j = j0 + lane;
    uint64_t w = 0, rank0 = 0, pos0 = 0;
    uint64_t src = 0, dst = 0;
    uint32_t width = 0;
    if (j < a.nvwords) {
        const int c = col_by_vword(a.cols, a.ncols, j);
        const DCol &col = a.cols[c];
        if (a.cst[c].has_nulls) {
            w = bm_word(col.validity, col.n, j - col.vword_base);
            rank0 = a.vpre[j] - a.vpre[col.vword_base];
            pos0 = (j - col.vword_base) * 64;
            src = (uint64_t)col.values;
            dst = (uint64_t)((uint8_t *)a.lane + col.lane_off * 8);
            width = col.in_width;
        }
    }
    // one element width across the wavefront's words (0 = nothing to move)
    const uint64_t nz = __ballot(width != 0);
    if (!nz) return;
    const uint32_t w0 = (uint32_t)__builtin_amdgcn_readlane((int)width, __builtin_ctzll(nz));
    if (__ballot(width != 0 && width != w0) == 0) {
        switch (w0) {
        case 1: compact_words<uint8_t>(w, rank0, pos0, src, dst, lane); return;
        case 2: compact_words<uint16_t>(w, rank0, pos0, src, dst, lane); return;
        case 4: compact_words<uint32_t>(w, rank0, pos0, src, dst, lane); return;
        default: compact_words<uint64_t>(w, rank0, pos0, src, dst, lane); return;
        }
    }
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int i0 = 0; i0 < WAVE; i0 += 8) {
        uint64_t v[8], rk[8], dp[8];
        uint32_t wd[8];
        bool on[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int i = i0 + u;  // word i of the 64: its fields are uniform (scalar reads)
            const uint64_t wi = qa_readlane64(w, i);
            wd[u] = (uint32_t)__builtin_amdgcn_readlane((int)width, i);
            dp[u] = qa_readlane64(dst, i);
            on[u] = (wi >> lane) & 1;  // bits past the column end are clear
            rk[u] = qa_readlane64(rank0, i) + (uint64_t)__popcll(wi & below);
            const uint64_t pos = qa_readlane64(pos0, i) + lane;
            const uint8_t *sp = (const uint8_t *)qa_readlane64(src, i);
            v[u] = on[u] ? ld_w(sp, pos, wd[u]) : 0;
        }
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (on[u]) st_w((uint8_t *)dp[u], rk[u], wd[u], v[u]);
    }
}

// ---------------------------------------------------------------------------
// stage 2: fused tile statistics + region probe, one workgroup per region
// ---------------------------------------------------------------------------
// Each wavefront takes tiles of the region. Per tile it keeps what the
// candidates of the Fastest/Fast pools need:
//   min/max (unsigned; signed for I32/I64, dtype.rs:121-123)    FoR, Dict
//   zig-zag delta min/max of elements 1..cnt-1                   DeltaBitpack
//   first/last value, step, all-equal, arithmetic                CONST, STRIDE
//   changes between neighbours                                   RLE, repeats probe
//   XORZ varint bytes (Fast only)                                XORZ
// The probe's exponent range only gates the entropy levels' byte transpose
// (encoder.rs:414-417) and is not computed. The previous element of each lane
// comes from its neighbour lane.
// sample accumulators of a full region, taken from the tile registers:
// p* every 16th value after the first (the region probe), g* every 64th value
// after the first (the grown-block probe partials)
struct ProbeAcc {
    uint64_t pmn, pmx, pdz, gmn, gmx, gdz;
};

// 16 consecutive lane values from index i0 (global loads): vector loads when
// the address is 16-byte aligned, else one load per value
template <int LK>
__device__ __forceinline__ void load16(const void *lp, uint32_t lk, uint64_t i0, uint64_t (&v)[16]) {
    if (LK == QA_LK_U32 || LK == QA_LK_I32 || LK == QA_LK_F32) {
        const qa_gu32 *p = (const qa_gu32 *)lp + i0;
        if (((uintptr_t)p & 15) == 0) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const qa_u32x4 w = ((const __attribute__((address_space(1))) qa_u32x4 *)p)[j];
                const uint32_t e[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    if (LK == QA_LK_U32) v[4 * j + u] = e[u];
                    else if (LK == QA_LK_I32) v[4 * j + u] = (uint64_t)(int64_t)(int32_t)e[u];
                    else v[4 * j + u] = qa_f32_to_lane(e[u]);
                }
            }
            return;
        }
    } else if (LK == QA_LK_U64) {
        const qa_gu64 *p = (const qa_gu64 *)lp + i0;
        if (((uintptr_t)p & 15) == 0) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const qa_u64x2 w = ((const __attribute__((address_space(1))) qa_u64x2 *)p)[j];
                v[2 * j] = w.x;
                v[2 * j + 1] = w.y;
            }
            return;
        }
    }
#pragma unroll
    for (int u = 0; u < 16; u++) v[u] = lane_get_g<LK>(lp, lk, i0 + u);
}

#define QA_DPP_ROW_ROR1 0x121  // lane i of a row takes lane (i - 1) mod 16

// reduction over a row of 16 lanes, the result in each of them
template <typename Op>
__device__ __forceinline__ uint64_t row_reduce64(uint64_t v, Op op) {
    v = op(v, qa_dpp64<QA_DPP_QP_1032>(Op::id, v));
    v = op(v, qa_dpp64<QA_DPP_QP_2301>(Op::id, v));
    v = op(v, qa_dpp64<QA_DPP_ROW_HALF_MIRROR>(Op::id, v));
    v = op(v, qa_dpp64<QA_DPP_ROW_MIRROR>(Op::id, v));
    return v;
}
__device__ __forceinline__ uint32_t row_sum32(uint32_t v) {
    v += qa_dpp32<QA_DPP_QP_1032>(0, v);
    v += qa_dpp32<QA_DPP_QP_2301>(0, v);
    v += qa_dpp32<QA_DPP_ROW_HALF_MIRROR>(0, v);
    v += qa_dpp32<QA_DPP_ROW_MIRROR>(0, v);
    return v;
}

// u32 lanes keep their values, min, max and equality tests in 32 bits
template <int LK> struct LaneVal { using T = uint64_t; };
template <> struct LaneVal<QA_LK_U32> { using T = uint32_t; };

template <int CTRL, typename T>
__device__ __forceinline__ T dpp_t(T v) {
    if constexpr (sizeof(T) == 4) return qa_dpp32<CTRL>(0, v);
    else return qa_dpp64<CTRL>(0, v);
}
template <bool MAX, typename T>
__device__ __forceinline__ T row_minmax(T v) {
    T o = dpp_t<QA_DPP_QP_1032>(v);
    v = MAX ? (o > v ? o : v) : (o < v ? o : v);
    o = dpp_t<QA_DPP_QP_2301>(v);
    v = MAX ? (o > v ? o : v) : (o < v ? o : v);
    o = dpp_t<QA_DPP_ROW_HALF_MIRROR>(v);
    v = MAX ? (o > v ? o : v) : (o < v ? o : v);
    o = dpp_t<QA_DPP_ROW_MIRROR>(v);
    v = MAX ? (o > v ? o : v) : (o < v ? o : v);
    return v;
}

// 16 consecutive values of a u32 lane from index i0
template <int LK>
__device__ __forceinline__ void load16(const void *lp, uint32_t lk, uint64_t i0, uint32_t (&v)[16]) {
    const qa_gu32 *p = (const qa_gu32 *)lp + i0;
    if (((uintptr_t)p & 15) == 0) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const qa_u32x4 w = ((const __attribute__((address_space(1))) qa_u32x4 *)p)[j];
            v[4 * j] = w.x;
            v[4 * j + 1] = w.y;
            v[4 * j + 2] = w.z;
            v[4 * j + 3] = w.w;
        }
        return;
    }
#pragma unroll
    for (int u = 0; u < 16; u++) v[u] = p[u];
}

// A wavefront takes four tiles at a time, one per row of 16 lanes. Lane L of
// a row walks four chunks of its tile: chunk c holds tile values 256 c + 16 L
// .. 256 c + 16 L + 15, so neighbours are registers of the same lane (the first
// one is lane L - 1's last, for lane 0 lane 15's last of the previous chunk, or
// the value before the tile) and the probe's sample positions are fixed
// register slots: every 8th value of the region is slot 0 or 8, every 16th
// slot 0, every 32nd slot 0 of even lanes, every 64th slot 0 of every fourth
// lane. Per-tile totals are reductions over the row (four DPP steps), one set
// per four tiles.
template <int LK, bool SG, bool FASTL, bool REG>
__device__ __forceinline__ uint32_t region_tiles(const DCol &col, const void *lp, uint32_t lk,
                                                 uint64_t rs, uint64_t len, DTile *tl,
                                                 uint64_t *seen8, uint64_t *seen32,
                                                 ProbeAcc &pa) {
    const int lane = threadIdx.x & 63, wv = qa_wave_id(), nwv = blockDim.x >> 6;
    const uint32_t L = lane & 15, row = lane >> 4;
    const uint32_t ntl = (uint32_t)((len + QA_TILE - 1) / QA_TILE);
    uint32_t eq_region = 0;
    for (uint32_t t0 = 4 * wv; t0 < col.tpr; t0 += 4 * nwv) {
        const uint32_t t = t0 + row;
        const bool act = t < ntl;  // a row past the region's values reads its first value
        const uint64_t k0 = act ? rs + (uint64_t)t * QA_TILE : rs;  // column index of the tile start
        const uint32_t cnt = act ? (uint32_t)min((uint64_t)QA_TILE, rs + len - k0) : 1u;
        using VT = typename LaneVal<LK>::T;
        const VT before = k0 ? (VT)lane_get_g<LK>(lp, lk, k0 - 1) : 0;
        const VT vf = (VT)lane_get_g<LK>(lp, lk, k0);
        VT umin = (VT)~0ull, umax = 0;
        uint64_t dzmin = ~0ull, dzmax = 0, xb = 0;
        uint64_t smin = 0x7FFFFFFFFFFFFFFFull, smax = 0x8000000000000000ull;
        uint32_t eqs = 0;  // equal neighbours inside the tile, + (1 << 16) across tiles
        VT carry = before;  // lane 15: the last value of the previous chunk
#pragma unroll 1
        for (uint32_t c = 0; c < 4; c++) {
            const uint32_t kb = 256 * c + 16 * L;  // this lane's first tile index
            VT v[16];
            if (kb + 16 <= cnt) {
                load16<LK>(lp, lk, k0 + kb, v);
            } else {
#pragma unroll
                for (int u = 0; u < 16; u++)  // past cnt: vf, neutral for min and max
                    v[u] = kb + u < cnt ? (VT)lane_get_g<LK>(lp, lk, k0 + kb + u) : vf;
            }
            const VT pv0 = dpp_t<QA_DPP_ROW_ROR1>(L == 15 ? carry : v[15]);
            carry = v[15];
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const uint32_t k = kb + i;
                const bool in = k < cnt;
                const VT xv = v[i];
                const VT pvv = i ? v[i - 1] : pv0;
                const uint64_t x = xv, p = pvv;
                umin = xv < umin ? xv : umin;
                umax = xv > umax ? xv : umax;
                if (SG) {
                    smin = (int64_t)x < (int64_t)smin ? x : smin;
                    smax = (int64_t)x > (int64_t)smax ? x : smax;
                }
                const uint64_t z = qa_zigzag(x - p);
                const bool pair = in && k > 0;  // neighbours inside the tile
                if (pair) {
                    dzmin = z < dzmin ? z : dzmin;
                    dzmax = z > dzmax ? z : dzmax;
                    if (FASTL) xb += qa_varint_len(x ^ p);
                    eqs += xv == pvv;
                }
                if (i == 0 && k == 0 && k0 > rs && x == p) eqs += 1u << 16;  // across tiles of the region
                if (REG && act) {  // every tile of a full region is full
                    const bool first = t == 0 && k == 0;  // region position 0
                    if (i == 0 || i == 8) {
                        const uint64_t h = x * 0x9E3779B97F4A7C15ull;
                        const uint32_t bucket = (uint32_t)((h >> 50) & 0x3FFF);
                        atomicOr((unsigned long long *)&seen8[bucket >> 6], 1ull << (bucket & 63));
                        if (i == 0 && (L & 1) == 0)
                            atomicOr((unsigned long long *)&seen32[bucket >> 6], 1ull << (bucket & 63));
                    }
                    if (i == 0 && !first) {
                        pa.pmn = x < pa.pmn ? x : pa.pmn;
                        pa.pmx = x > pa.pmx ? x : pa.pmx;
                        pa.pdz = z > pa.pdz ? z : pa.pdz;
                        if ((L & 3) == 0) {
                            pa.gmn = x < pa.gmn ? x : pa.gmn;
                            pa.gmx = x > pa.gmx ? x : pa.gmx;
                            pa.gdz = z > pa.gdz ? z : pa.gdz;
                        }
                    }
                }
            }
        }
        umin = row_minmax<false>(umin);
        umax = row_minmax<true>(umax);
        if (SG) {
            smin = row_reduce64(smin, OpMinS());
            smax = row_reduce64(smax, OpMaxS());
        }
        dzmin = row_reduce64(dzmin, OpMinU());
        dzmax = row_reduce64(dzmax, OpMaxU());
        if (FASTL) xb = row_reduce64(xb, OpAdd());
        eqs = row_sum32(eqs);
        const uint32_t eqt = eqs & 0xFFFF;  // at most 1023
        const uint32_t chg = cnt ? cnt - 1 - eqt : 0;  // changes between neighbours in the tile
        // all equal <=> min == max; arithmetic <=> every zig-zag delta equal
        const bool same = umin == umax, arith = cnt <= 1 || dzmin == dzmax;
        if (L == 0 && t < col.tpr) {
            DTile &ts = tl[t];
            const uint64_t vl = lane_get_g<LK>(lp, lk, k0 + cnt - 1);
            const uint64_t step = cnt > 1 ? lane_get_g<LK>(lp, lk, k0 + 1) - (uint64_t)vf : 0;
            if (!act) {
                ts.active = 0;
            } else {
                ts.umin = umin;
                ts.umax = umax;
                ts.smin = (int64_t)smin;
                ts.smax = (int64_t)smax;
                ts.dzmin = dzmin;
                ts.dzmax = dzmax;
                ts.xorz_rest = xb;
                ts.chg_rest = chg;
                ts.exp_min = 0;
                ts.exp_max = 0;
                ts.v_first = vf;
                ts.v_last = vl;
                ts.step = step;
                ts.dz_first = k0 ? qa_zigzag((uint64_t)vf - (uint64_t)before) : 0;
                ts.cnt = cnt;
                ts.arith = (uint8_t)arith;
                ts.allsame = (uint8_t)same;
                ts.active = 1;
            }
        }
        // the four rows' equal-neighbour counts (in-tile + across), active rows only
        const uint32_t eqp = act ? eqt + (eqs >> 16) : 0u;
        eq_region += (uint32_t)__builtin_amdgcn_readlane((int)eqp, 0) +
                     (uint32_t)__builtin_amdgcn_readlane((int)eqp, 16) +
                     (uint32_t)__builtin_amdgcn_readlane((int)eqp, 32) +
                     (uint32_t)__builtin_amdgcn_readlane((int)eqp, 48);
    }
    return eq_region;
}

// distinct estimate (encoder.rs:728-743) and incompressibility samples
// (encoder.rs:655-683) of the values [s0, s0 + n): seen is a 16384-bit LDS set
template <int LK>
__device__ __forceinline__ void probe_samples(const void *lp, uint32_t lk, uint64_t s0, uint64_t n,
                                              uint64_t *seen, uint64_t *mn, uint64_t *mx,
                                              uint64_t *dz) {
    // loads are issued in groups so several are in flight per lane
    const uint64_t stride = n / 4096 > 1 ? n / 4096 : 1, bd = blockDim.x;
    for (uint64_t k0 = (uint64_t)threadIdx.x * stride; k0 < n; k0 += 8 * bd * stride) {
        uint64_t v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const uint64_t k = k0 + u * bd * stride;
            v[u] = k < n ? lane_get_t<LK>(lp, lk, s0 + k) : 0;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            if (k0 + u * bd * stride >= n) break;
            uint64_t h = v[u] * 0x9E3779B97F4A7C15ull;
            uint32_t bucket = (uint32_t)((h >> 50) & 0x3FFF);
            atomicOr((unsigned long long *)&seen[bucket >> 6], 1ull << (bucket & 63));
        }
    }
    const uint64_t step = n / 2048 > 1 ? n / 2048 : 1;
    uint64_t a = *mn, b = *mx, d = *dz;
    for (uint64_t i0 = step * (1 + threadIdx.x); i0 < n; i0 += 4 * step * bd) {
        uint64_t v[4], pv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint64_t i = i0 + u * step * bd;
            v[u] = i < n ? lane_get_t<LK>(lp, lk, s0 + i) : 0;
            pv[u] = i < n ? lane_get_t<LK>(lp, lk, s0 + i - 1) : 0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (i0 + u * step * bd >= n) break;
            const uint64_t x = v[u];
            a = x < a ? x : a;
            b = x > b ? x : b;
            const uint64_t z = qa_zigzag(x - pv[u]);
            d = z > d ? z : d;
        }
    }
    *mn = a;
    *mx = b;
    *dz = d;
}

struct RegionShared {
    uint64_t smin, smax, sdz, gmn, gmx, gdz;
    uint32_t eq, dsum;
};

template <int LK, bool SG, bool FASTL>
__device__ __forceinline__ void region_stats_body(const EncArgs &a, uint32_t r, const DCol &col,
                                                  const void *lp, uint32_t lk, uint64_t rs,
                                                  uint64_t len, DTile *tl, uint64_t *seen,
                                                  uint64_t *seen32, RegionShared *sh, bool reg) {
    const int lane = threadIdx.x & 63;
    ProbeAcc pa{~0ull, 0, 0, ~0ull, 0, 0};
    uint32_t eq;
    if (reg) {
        eq = region_tiles<LK, SG, FASTL, true>(col, lp, lk, rs, len, tl, seen, seen32, pa);
    } else {
        eq = region_tiles<LK, SG, FASTL, false>(col, lp, lk, rs, len, tl, seen, seen32, pa);
        probe_samples<LK>(lp, lk, rs, len, seen, &pa.pmn, &pa.pmx, &pa.pdz);
    }
    if (lane == 0) atomicAdd(&sh->eq, eq);
    pa.pmn = wave_reduce(pa.pmn, OpMinU());
    pa.pmx = wave_reduce(pa.pmx, OpMaxU());
    pa.pdz = wave_reduce(pa.pdz, OpMaxU());
    if (lane == 0) {
        atomicMin((unsigned long long *)&sh->smin, pa.pmn);
        atomicMax((unsigned long long *)&sh->smax, pa.pmx);
        atomicMax((unsigned long long *)&sh->sdz, pa.pdz);
        if (reg) {  // only lane 0 took the every-64th samples
            atomicMin((unsigned long long *)&sh->gmn, pa.gmn);
            atomicMax((unsigned long long *)&sh->gmx, pa.gmx);
            atomicMax((unsigned long long *)&sh->gdz, pa.gdz);
        }
    }
}

#define QA_LANE_DISPATCH(F, lk, sg, fastl, ...)                                     \
    do {                                                                            \
        if (fastl) {                                                                \
            switch (lk) {                                                           \
            case QA_LK_U64:                                                         \
                if (sg) F<QA_LK_U64, true, true>(__VA_ARGS__);                      \
                else F<QA_LK_U64, false, true>(__VA_ARGS__);                        \
                break;                                                              \
            case QA_LK_U32: F<QA_LK_U32, false, true>(__VA_ARGS__); break;          \
            case QA_LK_I32: F<QA_LK_I32, true, true>(__VA_ARGS__); break;           \
            case QA_LK_F32: F<QA_LK_F32, false, true>(__VA_ARGS__); break;          \
            default:                                                                \
                if (sg) F<QA_LK_ANY, true, true>(__VA_ARGS__);                      \
                else F<QA_LK_ANY, false, true>(__VA_ARGS__);                        \
            }                                                                       \
        } else {                                                                    \
            switch (lk) {                                                           \
            case QA_LK_U64:                                                         \
                if (sg) F<QA_LK_U64, true, false>(__VA_ARGS__);                     \
                else F<QA_LK_U64, false, false>(__VA_ARGS__);                       \
                break;                                                              \
            case QA_LK_U32: F<QA_LK_U32, false, false>(__VA_ARGS__); break;         \
            case QA_LK_I32: F<QA_LK_I32, true, false>(__VA_ARGS__); break;          \
            case QA_LK_F32: F<QA_LK_F32, false, false>(__VA_ARGS__); break;         \
            default:                                                                \
                if (sg) F<QA_LK_ANY, true, false>(__VA_ARGS__);                     \
                else F<QA_LK_ANY, false, false>(__VA_ARGS__);                       \
            }                                                                       \
        }                                                                           \
    } while (0)

__global__ void __launch_bounds__(512) k_region_stats(EncArgs a) {
    __shared__ uint64_t seen[256], seen32[256];
    __shared__ RegionShared sh;
    const uint32_t r = blockIdx.x;
    const int c = col_by_region(a.cols, a.ncols, r);
    const DCol &col = a.cols[c];
    const DColState &st = a.cst[c];
    const uint64_t nv = st.nv;
    const void *lp = st.lanep;
    const uint32_t lk = st.lkind;
    const uint64_t rs = (uint64_t)(r - col.region_base) * col.region_size;
    const int lane = threadIdx.x & 63;
    DTile *tl = a.tiles + col.tile_base + (uint64_t)(r - col.region_base) * col.tpr;
    if (rs >= nv) {
        for (uint32_t t = threadIdx.x; t < col.tpr; t += blockDim.x) tl[t].active = 0;
        if (threadIdx.x == 0) a.rgrown[r].full = 0;
        return;
    }
    const uint64_t len = nv - rs < col.region_size ? nv - rs : col.region_size;
    // a full adaptive region samples from the tile registers (every 8th value
    // is the region probe's stride, 16th its step) and records the partials
    // of the 131072-value block probe (every 32nd / 64th value)
    const bool reg = col.adaptive && len == QA_BASE_Q;
    for (int k = threadIdx.x; k < 256; k += blockDim.x) {
        seen[k] = 0;
        seen32[k] = 0;
    }
    if (threadIdx.x == 0) {
        sh.eq = 0;
        sh.dsum = 0;
        sh.smin = ~0ull;
        sh.smax = 0;
        sh.sdz = 0;
        sh.gmn = ~0ull;
        sh.gmx = 0;
        sh.gdz = 0;
    }
    __syncthreads();
    const bool sg = is_signed(col.dtype), fastl = col.level != QA_FASTEST;
    QA_LANE_DISPATCH(region_stats_body, lk, sg, fastl, a, r, col, lp, lk, rs, len, tl, seen, seen32,
                     &sh, reg);
    __syncthreads();
    uint32_t d = 0;
    for (int k = threadIdx.x; k < 256; k += blockDim.x) d += (uint32_t)__popcll(seen[k]);
    d = (uint32_t)wave_reduce((uint64_t)d, OpAdd());
    if (lane == 0) atomicAdd(&sh.dsum, d);
    if (reg)
        for (int k = threadIdx.x; k < 256; k += blockDim.x) a.gsets[(uint64_t)r * 256 + k] = seen32[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint64_t v0 = qa_lane_get(lp, lk, rs);
        const uint64_t mn = sh.smin < v0 ? sh.smin : v0, mx = sh.smax > v0 ? sh.smax : v0;
        bool dl = sh.dsum < 2048;
        uint64_t c2 = (uint64_t)(sh.eq + 1) * 2;  // consec_eq counts element 0 against itself
        if (c2 > 0xFFFFFFFFull) c2 = 0xFFFFFFFFull;
        bool rp = c2 > (uint32_t)len;
        bool inc = false;
        if (!is_float(col.dtype) && !dl && !rp && len >= 256) {
            uint32_t thr = (uint32_t)(lane_bytes_of(col.dtype) * 8) - 4;
            inc = qa_width(mx - mn) >= thr && qa_width(sh.sdz) >= thr;
        }
        a.rfeat[r] = (dl ? RF_DLOW : 0) | (rp ? RF_REP : 0) | (inc ? RF_INC : 0);
        DRegionG &g = a.rgrown[r];
        g.full = reg;
        g.gmn = sh.gmn;
        g.gmx = sh.gmx;
        g.gdz = sh.gdz;
        g.v0 = v0;
        g.dz0 = rs ? qa_zigzag(v0 - qa_lane_get(lp, lk, rs - 1)) : 0;
    }
}

// ---------------------------------------------------------------------------
// stage 3: block planning (src/encoder.rs:154-178), one workgroup per column
// ---------------------------------------------------------------------------
#define PLAN_CHUNK 4096

// The walk over an adaptive column's regions as a scan: s = regions left to
// skip before the next block start. At a region, s = 0 makes it a start and
// sets s = 3 if its probe grows the block (four regions), else 0; s > 0
// decrements. A region's transition is a map of the four states, packed as
// 4 x 2 bits; maps compose associatively, so the states come from a prefix
// scan of the maps.
#define PLAN_ID 0xE4u  // s -> s
#define PLAN_T0 0x90u  // 0 -> 0 (a one-region block starts), 1 -> 0, 2 -> 1, 3 -> 2
#define PLAN_T1 0x93u  // 0 -> 3 (a grown block starts), 1 -> 0, 2 -> 1, 3 -> 2
__device__ __forceinline__ uint32_t plan_compose(uint32_t first, uint32_t then) {
    uint32_t out = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) out |= ((then >> (2 * ((first >> (2 * q)) & 3))) & 3) << (2 * q);
    return out;
}

// A block that starts at region lr ends at min(start + 131072, nv) when its
// probe says distinct_low || repeats and the column is adaptive (regions are
// 32768 values there), else at min(start + R, nv); the next block starts at
// region min(lr + 4 or 1, nreg). Each thread takes 16 regions of a chunk of
// 4096: it composes their maps, a scan over the 256 threads gives each its
// entry state, and it marks its starts; owners are found by a backward scan
// of at most three regions (a block spans at most four).
__global__ void __launch_bounds__(256) k_plan(EncArgs a) {
    __shared__ uint64_t featw[PLAN_CHUNK / 64], startw[PLAN_CHUNK / 64];
    __shared__ uint32_t next_s, nb_s, last_s;
    __shared__ uint32_t fsc[256], cl_nb, cl_last, cl_next;
    const int c = blockIdx.x;
    const DCol &col = a.cols[c];
    const uint64_t nv = a.cst[c].nv;
    const uint32_t R = col.region_size;
    const uint32_t nreg = (uint32_t)((nv + R - 1) / R);  // regions holding values
    const bool adaptive = col.adaptive;
    const uint32_t grow = QA_MAX_Q / QA_BASE_Q;
    if (threadIdx.x == 0) {
        next_s = 0;
        nb_s = 0;
        last_s = NONE32;
    }
    for (uint32_t base = 0; base < col.nregions; base += PLAN_CHUNK) {
        const uint32_t cn = min((uint32_t)PLAN_CHUNK, col.nregions - base);
        const uint32_t cw = (cn + 63) / 64;
        __syncthreads();
        // the last block start before this chunk (thread 0 updates it below,
        // after the next barrier)
        const uint32_t prev_last = last_s;
        // feature and start words (the loop bound is a multiple of 64, so each
        // wavefront's lanes cover one aligned word per step)
        for (uint32_t k = threadIdx.x; k < cw * 64; k += blockDim.x) {
            const uint32_t lr = base + k;
            const bool f = k < cn && lr < nreg && (a.rfeat[col.region_base + lr] & 3);
            const bool st = k < cn && lr < nreg && !adaptive;  // fixed size: every region
            const uint64_t fm = __ballot(f), sm = __ballot(st);
            if ((threadIdx.x & 63) == 0) {
                featw[k >> 6] = fm;
                startw[k >> 6] = sm;
            }
        }
        __syncthreads();
        const uint32_t lim = min(base + cn, nreg);
        if (!adaptive) {
            if (threadIdx.x == 0 && next_s < lim) {
                nb_s += lim - next_s;
                last_s = lim - 1;
                next_s = lim;
            }
        } else {
            const uint32_t t = threadIdx.x, k0 = 16 * t;
            uint32_t F = PLAN_ID;
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const uint32_t k = k0 + i;
                if (base + k < lim)
                    F = plan_compose(F, ((featw[k >> 6] >> (k & 63)) & 1) ? PLAN_T1 : PLAN_T0);
            }
            if (t == 0) {
                cl_nb = 0;
                cl_last = 0;
                cl_next = next_s;
            }
            // inclusive scan of the maps in region order (Hillis-Steele over the threads)
            fsc[t] = F;
            __syncthreads();
            for (uint32_t d = 1; d < 256; d <<= 1) {
                const uint32_t other = t >= d ? fsc[t - d] : PLAN_ID;
                __syncthreads();
                if (t >= d) fsc[t] = plan_compose(other, fsc[t]);
                __syncthreads();
            }
            // the state at the chunk's first region: next_s is the first start at or after it
            const uint32_t s0 = next_s > base ? next_s - base : 0;
            uint32_t st = ((t ? fsc[t - 1] : PLAN_ID) >> (2 * s0)) & 3;
            uint32_t bits = 0;
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const uint32_t k = k0 + i;
                if (base + k < lim) {
                    const bool f = (featw[k >> 6] >> (k & 63)) & 1;
                    if (st == 0) bits |= 1u << i;
                    st = st == 0 ? (f ? grow - 1 : 0u) : st - 1;
                }
            }
            if (bits) {
                atomicOr((unsigned long long *)&startw[k0 >> 6], (unsigned long long)bits << (k0 & 63));
                atomicAdd(&cl_nb, (uint32_t)__popc(bits));
                atomicMax(&cl_last, k0 + 31 - (uint32_t)__clz(bits));
            }
            if (lim > base && k0 <= lim - 1 - base && lim - 1 - base < k0 + 16)
                cl_next = lim + st;  // the first start after the chunk's last region
            __syncthreads();
            if (t == 0) {
                nb_s += cl_nb;
                if (cl_nb) last_s = base + cl_last;
                next_s = cl_next;
            }
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < cn; k += blockDim.x) {
            const uint32_t lr = base + k, slot = col.region_base + lr;
            DBlock &b = a.blocks[slot];
            if (lr >= nreg) {
                a.owner[slot] = NONE32;
                b.len = 0;
                continue;
            }
            uint32_t o = NONE32;
            for (uint32_t d = 0; d < grow && o == NONE32; d++) {
                if (d > k) break;
                const uint32_t kk = k - d;
                if ((startw[kk >> 6] >> (kk & 63)) & 1) o = col.region_base + base + kk;
            }
            if (o == NONE32) o = col.region_base + prev_last;  // block begun in an earlier chunk
            a.owner[slot] = o;
            if (o == slot) {
                const uint64_t rs = (uint64_t)lr * R;
                const bool g = adaptive && ((featw[k >> 6] >> (k & 63)) & 1);
                const uint64_t end = g ? min(rs + (uint64_t)QA_MAX_Q, nv) : min(rs + (uint64_t)R, nv);
                b.start = rs;
                b.len = (uint32_t)(end - rs);
                b.col = (uint32_t)c;
                b.ntiles = (uint32_t)((end - rs + QA_TILE - 1) / QA_TILE);
            } else {
                b.len = 0;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) a.cst[c].nblocks = nb_s;
}

// ---------------------------------------------------------------------------
// stage 4: features of blocks grown past one region (131072 values)
// ---------------------------------------------------------------------------
template <int LK, bool SG, bool FASTL>
__device__ __forceinline__ void grown_samples(const void *lp, uint32_t lk, uint64_t s0, uint64_t n,
                                              uint64_t *seen, uint64_t *mn, uint64_t *mx,
                                              uint64_t *dz) {
    probe_samples<LK>(lp, lk, s0, n, seen, mn, mx, dz);
}

__global__ void __launch_bounds__(256) k_grown_feats(EncArgs a) {
    __shared__ uint64_t seen[256];
    __shared__ uint64_t smin, smax, sdz;
    __shared__ uint32_t dsum, eq_s;
    const uint32_t slot = a.r0 + blockIdx.x;
    if (slot >= a.r1) return;
    DBlock &b = a.blocks[slot];
    const uint32_t blen = b.len;
    if (blen == 0) return;
    const uint32_t bc = b.col;
    const DCol &col = a.cols[bc];
    const uint32_t f = a.rfeat[slot];
    if (blen <= col.region_size) {  // a one-region block: its region's probe
        if (threadIdx.x == 0) {
            b.distinct_low = (f & RF_DLOW) != 0;
            b.repeats = (f & RF_REP) != 0;
            b.incompressible = (f & RF_INC) != 0;
            b.exp_range = 0;
        }
        return;
    }
    const void *lp = a.cst[bc].lanep;
    const uint32_t lk = a.cst[bc].lkind;
    const uint64_t n = blen, s0 = b.start;
    const uint32_t bnt = b.ntiles;
    // a block of four full regions combines their recorded probe partials
    // (its stride 32 and step 64 are multiples of the regions' own)
    bool comb = n == QA_MAX_Q && col.region_size == QA_BASE_Q && col.adaptive;
    for (int j = 0; j < 4 && comb; j++) comb = a.rgrown[slot + j].full != 0;
    for (int k = threadIdx.x; k < 256; k += blockDim.x) seen[k] = 0;
    if (threadIdx.x == 0) {
        uint64_t v0 = qa_lane_get(lp, lk, s0);
        smin = v0;
        smax = v0;
        sdz = 0;
        dsum = 0;
        eq_s = 0;
    }
    __syncthreads();
    uint64_t mn = ~0ull, mx = 0, dz = 0;
    if (comb) {
        for (int k = threadIdx.x; k < 256; k += blockDim.x) {
            const uint64_t *gs = a.gsets + (uint64_t)slot * 256 + k;
            seen[k] = gs[0] | gs[256] | gs[512] | gs[768];
        }
        if (threadIdx.x < 4) {
            const DRegionG &g = a.rgrown[slot + threadIdx.x];
            mn = g.gmn;
            mx = g.gmx;
            dz = g.gdz;
            if (threadIdx.x > 0) {  // a later region's first value is a sample too
                mn = g.v0 < mn ? g.v0 : mn;
                mx = g.v0 > mx ? g.v0 : mx;
                dz = g.dz0 > dz ? g.dz0 : dz;
            }
        }
    } else {
        QA_LANE_DISPATCH(grown_samples, lk, false, false, lp, lk, s0, n, seen, &mn, &mx, &dz);
    }
    // one LDS atomic per wavefront (all 256 threads on one address serialize)
    mn = wave_reduce(mn, OpMinU());
    mx = wave_reduce(mx, OpMaxU());
    dz = wave_reduce(dz, OpMaxU());
    if ((threadIdx.x & 63) == 0) {
        atomicMin((unsigned long long *)&smin, mn);
        atomicMax((unsigned long long *)&smax, mx);
        atomicMax((unsigned long long *)&sdz, dz);
    }
    // the consecutive-equal count combines from the tiles
    const DTile *tl = a.tiles + col.tile_base + (uint64_t)(slot - col.region_base) * col.tpr;
    uint32_t eq = 0;
    for (uint32_t t = threadIdx.x; t < bnt; t += blockDim.x) {
        const DTile &ts = tl[t];
        eq += ts.cnt - 1 - ts.chg_rest;
        if (t > 0) eq += ts.v_first == tl[t - 1].v_last;
    }
    eq = (uint32_t)wave_reduce((uint64_t)eq, OpAdd());
    if ((threadIdx.x & 63) == 0) atomicAdd(&eq_s, eq);
    __syncthreads();
    uint32_t d = 0;
    for (int k = threadIdx.x; k < 256; k += blockDim.x) d += (uint32_t)__popcll(seen[k]);
    d = (uint32_t)wave_reduce((uint64_t)d, OpAdd());
    if ((threadIdx.x & 63) == 0) atomicAdd(&dsum, d);
    __syncthreads();
    if (threadIdx.x == 0) {
        bool dl = dsum < 2048;
        uint64_t c2 = (uint64_t)(eq_s + 1) * 2;
        if (c2 > 0xFFFFFFFFull) c2 = 0xFFFFFFFFull;
        bool rp = c2 > (uint32_t)n;
        bool inc = false;
        if (!is_float(col.dtype) && !dl && !rp && n >= 256) {
            uint32_t thr = (uint32_t)(lane_bytes_of(col.dtype) * 8) - 4;
            inc = qa_width(smax - smin) >= thr && qa_width(sdz) >= thr;
        }
        b.distinct_low = dl;
        b.repeats = rp;
        b.incompressible = inc;
        b.exp_range = 0;
    }
}

// ---------------------------------------------------------------------------
// stage 5: candidate sizes and the ordered replay of Best::consider, one
// wavefront per block
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// stage 6: validity blob + frame layout
// ---------------------------------------------------------------------------
__global__ void k_vtrans_pos(EncArgs a) {
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.nvwords) return;
    int c = col_by_vword(a.cols, a.ncols, j);
    const DCol &col = a.cols[c];
    uint64_t lj = j - col.vword_base;
    uint64_t w = bm_word(col.validity, col.n, lj);
    uint64_t prev_top = lj ? (bm_word(col.validity, col.n, lj - 1) >> 63) : 0;
    uint64_t t = w ^ ((w << 1) | prev_top);
    uint64_t bits = col.n - lj * 64;
    if (bits < 64) t &= (1ull << bits) - 1;
    uint64_t k = a.tpre[j];
    while (t) {
        uint32_t b = (uint32_t)__ffsll((long long)t) - 1;
        a.tpos[k++] = lj * 64 + b;
        t &= t - 1;
    }
}

// run slot s = global transition index + column index; run k of a column ends
// at its transition k (or at n for the last run)
__device__ __forceinline__ int col_by_run(const EncArgs &a, uint64_t s) {
    int lo = 0, hi = (int)a.ncols - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (a.colbase[mid] + (uint64_t)mid <= s) lo = mid; else hi = mid - 1;
    }
    return lo;
}
__global__ void k_vrun_sizes(EncArgs a) {
    uint64_t total = a.colbase[a.ncols + 1];
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < total;
         s += (uint64_t)gridDim.x * blockDim.x) {
        int c = col_by_run(a, s);
        const DCol &col = a.cols[c];
        uint64_t first = a.colbase[c] + c, T = a.colbase[c + 1] - a.colbase[c];
        uint64_t k = s - first;
        if (!col.nvwords || !a.cst[c].has_nulls) {
            a.rsz[s] = 0;
            continue;
        }
        uint64_t st = k == 0 ? 0 : a.tpos[a.colbase[c] + k - 1];
        uint64_t en = k == T ? col.n : a.tpos[a.colbase[c] + k];
        a.rsz[s] = qa_varint_len(en - st);
    }
}

__global__ void __launch_bounds__(256) k_col_layout(EncArgs a) {
    // one workgroup per column: validity blob choice, then frame offsets
    const int c = blockIdx.x;
    const DCol &col = a.cols[c];
    DColState &st = a.cst[c];
    if (c == 0 && threadIdx.x == 0) a.tlist[0] = 0;  // k_emit_tiles' list, filled after the layout
    __shared__ uint64_t frames_off_s;
    __shared__ uint64_t buf[256];
    // the column's blocks per winning mode, added to the batch's win counters
    // once per mode (one counter atomic per block would serialize on a few
    // addresses); the host stages count no wins
    __shared__ uint32_t mcount[64];
    if (threadIdx.x < 64) mcount[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        uint64_t vlen = 0;
        if (st.has_nulls) {
            uint64_t first = a.colbase[c] + c, T = a.colbase[c + 1] - a.colbase[c];
            uint64_t rle = a.rsz[first + T + 1] - a.rsz[first];
            uint64_t raw = (col.n + 7) / 8;
            st.vblob_rle = rle;
            vlen = 1 + (rle < raw ? rle : raw);
        }
        st.vblob_len = vlen;
        frames_off_s = 16 + (st.has_nulls ? qa_varint_len(vlen) + vlen : 0);
        st.frames_off = frames_off_s;
    }
    __syncthreads();
    uint64_t carry = frames_off_s;
    for (uint32_t base = 0; base < col.nregions; base += blockDim.x) {
        uint32_t lr = base + threadIdx.x;
        uint64_t v = 0;
        if (lr < col.nregions) {
            const DBlock &b = a.blocks[col.region_base + lr];
            if (b.len) {
                v = 9 + b.payload_len;
                if (!a.hybrid) atomicAdd(&mcount[b.mode & 63], 1u);
            }
        }
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int d = 1; d < 256; d <<= 1) {
            uint64_t o = threadIdx.x >= (unsigned)d ? buf[threadIdx.x - d] : 0;
            __syncthreads();
            buf[threadIdx.x] += o;
            __syncthreads();
        }
        uint64_t incl = buf[threadIdx.x];
        if (lr < col.nregions) a.blocks[col.region_base + lr].frame_off = carry + incl - v;
        uint64_t tot = buf[255];
        __syncthreads();
        carry += tot;
    }
    if (threadIdx.x == 0) {
        st.stream_len = carry;
        if (carry > col.dst_cap) st.err |= QA_ERR_OVERFLOW;
    }
    __syncthreads();
    if (threadIdx.x < 64 && mcount[threadIdx.x])
        atomicAdd(&a.wins[threadIdx.x], (unsigned long long)mcount[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// stage 7: emission
// ---------------------------------------------------------------------------
__device__ __forceinline__ void emit_header_one(const EncArgs &a, uint64_t ci) {
    if (ci >= a.ncols) return;
    const int c = (int)ci;
    const DCol &col = a.cols[c];
    const DColState &st = a.cst[c];
    if (st.err) return;
    uint8_t *d = col.dst;
    // src/format.rs:56-63
    d[0] = 'F'; d[1] = 'C'; d[2] = 'R'; d[3] = '1';
    d[4] = 2;
    d[5] = st.has_nulls ? 1 : 0;
    d[6] = (uint8_t)(col.plog2 ? col.plog2 : 16);
    d[7] = (uint8_t)col.dtype;
    qa_st64u(d + 8, col.n);
    if (st.has_nulls) {
        uint32_t k = qa_put_varint(d + 16, st.vblob_len);
        uint64_t raw = (col.n + 7) / 8;
        d[16 + k] = st.vblob_rle < raw ? 1 : 0;
    }
}

// run slots s0, s0 + stride, ... of the device-side run count
__device__ __forceinline__ void emit_vruns_from(const EncArgs &a, uint64_t s0, uint64_t stride) {
    uint64_t total = a.colbase[a.ncols + 1];
    for (uint64_t s = s0; s < total; s += stride) {
        int c = col_by_run(a, s);
        const DCol &col = a.cols[c];
        const DColState &st = a.cst[c];
        if (!col.nvwords || !st.has_nulls || st.err) continue;
        uint64_t raw = (col.n + 7) / 8;
        if (!(st.vblob_rle < raw)) continue;
        uint64_t first = a.colbase[c] + c, T = a.colbase[c + 1] - a.colbase[c];
        uint64_t k = s - first;
        uint64_t stp = k == 0 ? 0 : a.tpos[a.colbase[c] + k - 1];
        uint64_t en = k == T ? col.n : a.tpos[a.colbase[c] + k];
        uint64_t off = 16 + qa_varint_len(st.vblob_len) + 1 + (a.rsz[s] - a.rsz[first]);
        qa_put_varint(col.dst + off, en - stp);
    }
}

// raw validity blob: the bitmap bytes verbatim (src/validity.rs:63-68)
// the raw bitmap blob [0][ceil(n / 8) bytes] of columns whose runs blob is not
// smaller; one thread per bitmap word (8 bytes)
__device__ __forceinline__ void emit_vraw_one(const EncArgs &a, uint64_t j) {
    if (j >= a.nvwords) return;
    int c = col_by_vword(a.cols, a.ncols, j);
    const DCol &col = a.cols[c];
    const DColState &st = a.cst[c];
    if (!col.nvwords || !st.has_nulls || st.err) return;
    const uint64_t raw = (col.n + 7) / 8;
    if (st.vblob_rle < raw) return;
    const uint64_t b0 = (j - col.vword_base) * 8;
    uint8_t *d = col.dst + 16 + qa_varint_len(st.vblob_len) + 1;
    for (uint64_t bi = b0; bi < b0 + 8 && bi < raw; bi++) d[bi] = col.validity[bi];
}

__device__ __forceinline__ void emit_frame_one(const EncArgs &a, uint64_t i) {
    if (i >= (uint64_t)(a.r1 - a.r0)) return;
    const uint32_t slot = a.r0 + (uint32_t)i;
    const DBlock &b = a.blocks[slot];
    if (b.len == 0) return;
    const DCol &col = a.cols[b.col];
    if (a.cst[b.col].err) return;
    uint8_t *d = col.dst + b.frame_off;
    d[0] = (uint8_t)b.mode;  // src/encoder.rs:752-759
    qa_st32u(d + 1, b.len);
    qa_st32u(d + 5, (uint32_t)b.payload_len);
    uint8_t *p = d + 9;
    switch (b.mode) {
    case QA_M_CONST: qa_st64u(p, b.v0); break;
    case QA_M_STRIDE:
        qa_st64u(p, b.v0);
        qa_st64u(p + 8, b.v1 - b.v0);
        break;
    case QA_M_FOR: qa_put_varint(p, b.len); break;
    case QA_M_DELTA:
        qa_st64u(p, b.v0);
        qa_put_varint(p + 8, b.len);
        break;
    default: break;
    }
}

// One wavefront per tile, four independent tiles per workgroup (LDS is
// per-wavefront, so only wavefront-level synchronization is used). FastLanes
// pack (src/bitpack.rs:47-133) straight from the lane: in the u32 form lane l
// of the 32 packs rows [16 h, 16 h + 16) (h = half of the wavefront), in the
// u64 form lane l of the 16 packs rows [16 q, 16 q + 16) (q = quarter); each
// residual is ORed into its one or two LDS words. The body is instantiated
// per lane storage kind so the loads are straight-line code.
template <int LK>
__device__ __forceinline__ void emit_tile_t(const TileEmit &e, uint64_t *o64w, int lane) {
    uint32_t *outw = (uint32_t *)o64w;
    const void *lp = e.lp;
    const uint32_t lk = e.lk, cnt = e.cnt;
    const uint64_t k0 = e.k0;
    if (e.mode == QA_M_RAW) {
        const uint32_t lb = (uint32_t)lane_bytes_of(e.dtype);
        uint64_t xs[QA_TILE / WAVE];
#pragma unroll
        for (int i = 0; i < QA_TILE / WAVE; i++) {
            const uint32_t k = i * WAVE + lane;
            xs[i] = k < cnt ? lane_get_t<LK>(lp, lk, k0 + k) : 0;
        }
#pragma unroll
        for (int i = 0; i < QA_TILE / WAVE; i++) {
            const uint32_t k = i * WAVE + lane;
            if (k >= cnt) break;
            const uint64_t x = xs[i];
            if (e.dtype == QA_F32) outw[k] = qa_lane_to_f32(x);
            else if (lb == 4) outw[k] = (uint32_t)x;
            else {
                outw[2 * k] = (uint32_t)x;
                outw[2 * k + 1] = (uint32_t)(x >> 32);
            }
        }
        wave_lds_sync();
        wave_store_bytes(e.dst, outw, cnt * lb, lane);
        return;
    }
    const bool delta = e.mode == QA_M_DELTA;
    const uint64_t mn = e.mn;
    const uint32_t w = e.w;
    uint8_t *q = e.dst;
    if (lane == 0) q[0] = (uint8_t)w;
    if (lane >= 1 && lane < 9) q[lane] = (uint8_t)(mn >> (8 * (lane - 1)));
    if (w == 0) return;
    for (uint32_t i = lane; i < 32 * w; i += WAVE) outw[i] = 0;
    wave_lds_sync();
    // residual of tile value k: FoR x - min; Delta zigzag(x - prev) - min with
    // the block's first value as its own predecessor (delta_bitpack.rs:23-46);
    // zero past cnt (padding to 1024, for_bitpack.rs:60-72). The lane's 16
    // values (and predecessors) are loaded before any is used.
    const bool u32form = w <= 32;
    const uint32_t l = u32form ? (lane & 31) : (lane & 15);
    const uint32_t r0 = u32form ? 16 * (lane >> 5) : 16 * (lane >> 4);
    const uint32_t L = u32form ? 32 : 16;
    uint64_t xs[16], ps[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t k = (r0 + i) * L + l;
        xs[i] = k < cnt ? lane_get_t<LK>(lp, lk, k0 + k) : 0;
        ps[i] = (delta && k < cnt && (!e.first || k != 0)) ? lane_get_t<LK>(lp, lk, k0 + k - 1)
                                                           : xs[i];
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t k = (r0 + i) * L + l;
        xs[i] = k >= cnt ? 0 : (delta ? qa_zigzag(xs[i] - ps[i]) - mn : xs[i] - mn);
    }
    if (u32form) {
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const uint32_t r = r0 + i;
            const uint32_t v = (uint32_t)xs[i];
            const uint32_t bit = r * w, word = bit >> 5, sh = bit & 31;
            atomicOr(&outw[word * 32 + l], v << sh);
            if (sh + w > 32) atomicOr(&outw[(word + 1) * 32 + l], v >> (32 - sh));
        }
    } else {
        unsigned long long *o64 = (unsigned long long *)o64w;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const uint32_t r = r0 + i;
            const uint64_t v = xs[i];
            const uint32_t bit = r * w, word = bit >> 6, sh = bit & 63;
            atomicOr(&o64[word * 16 + l], (unsigned long long)(v << sh));
            if (sh + w > 64)
                atomicOr(&o64[(word + 1) * 16 + l], (unsigned long long)(v >> (64 - sh)));
        }
    }
    wave_lds_sync();
    wave_store_bytes(q + 9, outw, 128 * w, lane);
}

// one thread per tile: the emit descriptor (block, frame offset, minimum);
// called by whole wavefronts (the list append takes one atomic per wavefront)
__device__ __forceinline__ void tile_desc_one(const EncArgs &a, uint64_t i) {
    const uint64_t t = a.t0 + i;
    if (t >= a.t1) return;
    TileEmit e{};
    const DTile &ts = a.tiles[t];
    const int c = col_by_tile(a.cols, a.ncols, t);
    const DCol &col = a.cols[c];
    const DColState &st = a.cst[c];
    const uint64_t lt = t - col.tile_base;
    const uint32_t lr = (uint32_t)(lt / col.tpr);
    const uint32_t o = lr < col.nregions ? a.owner[col.region_base + lr] : NONE32;
    if (o != NONE32 && !st.err && a.blocks[o].len) e.bmode = a.blocks[o].mode;
    if (ts.active && !st.err && o != NONE32) {
        const DBlock &b = a.blocks[o];
        const uint32_t bt = (col.region_base + lr - o) * col.tpr + (uint32_t)(lt % col.tpr);
        const uint32_t mode = b.mode;
        uint8_t *p = col.dst + b.frame_off + 9;
        e.lp = st.lanep;
        e.k0 = b.start + (uint64_t)bt * QA_TILE;
        e.cnt = ts.cnt;
        e.lk = (uint8_t)st.lkind;
        e.dtype = col.dtype;
        e.first = bt == 0;
        if (mode == QA_M_RAW) {
            e.mode = QA_M_RAW;
            e.dst = p + (uint64_t)bt * QA_TILE * lane_bytes_of(col.dtype);
        } else if (mode == QA_M_FOR || mode == QA_M_DELTA) {
            e.mode = (uint8_t)mode;
            e.w = (uint8_t)ts.w;
            e.fast = ts.w <= 32 ? 1u : 0u;
            e.dst = p + ts.off;
            if (mode == QA_M_DELTA) {
                uint64_t dmx;
                tile_dz(ts, bt == 0, &e.mn, &dmx);
            } else {
                e.mn = is_signed(col.dtype) ? (uint64_t)ts.smin : ts.umin;
            }
        } else if (mode == QA_M_FLOAT_MULT && a.fast) {
            // a FloatMult sub-block is a FoR sub-block of k = round(x * scale)
            // (float_mult.rs:66-98); wider ones stay with k_emit_fast_tiles
            const DTileF &tf = a.tilesf[t];
            const uint32_t sc = a.blocksx[o].fm_scale;
            if (tf.fm_w <= 32) {
                e.mode = QA_M_FLOAT_MULT;
                e.w = (uint8_t)tf.fm_w;
                e.fast = 1;
                e.dst = p + ts.off;
                e.mn = (uint64_t)tf.kmin[sc];
                e.scale = sc;
            }
        }
    }
    a.temit[t] = e;
    // the tiles k_emit_fl does not take go on k_emit_tiles' list (one atomic
    // per wavefront)
    const bool gen = e.mode && !e.fast;
    const uint64_t m = __ballot(gen);
    if (m) {
        const int lane = threadIdx.x & 63, lead = __ffsll((long long)m) - 1;
        uint32_t base = 0;
        if (lane == lead) base = atomicAdd(a.tlist, (uint32_t)__popcll(m));
        base = (uint32_t)__builtin_amdgcn_readlane((int)base, lead);
        const uint32_t r = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (gen) a.tlist[16 + base + r] = (uint32_t)t;
    }
}

static __device__ __forceinline__ uint64_t flat_tid() {
    return (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
}
__global__ void k_emit_header(EncArgs a) { emit_header_one(a, flat_tid()); }
__global__ void k_emit_vruns(EncArgs a) {
    emit_vruns_from(a, flat_tid(), (uint64_t)gridDim.x * blockDim.x);
}
__global__ void k_emit_vraw(EncArgs a) { emit_vraw_one(a, flat_tid()); }
__global__ void k_emit_frames(EncArgs a) { emit_frame_one(a, flat_tid()); }
__global__ void k_tile_desc(EncArgs a) { tile_desc_one(a, flat_tid()); }

// Once the frame layout is known, the stream header, the validity blob, the
// frame headers and the tile emit descriptors depend on nothing of each
// other, so they share one launch of 256-thread workgroups, split by range:
// column headers, validity runs (grid-stride over the device-side run count),
// raw validity words, frame headers, tile descriptors.
struct MetaSplit {
    uint32_t h, r, w, f;  // workgroups of the first four parts; the rest: tiles
};
__global__ void __launch_bounds__(256) k_emit_meta(EncArgs a, MetaSplit m) {
    uint32_t b = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    if (b < m.h) return emit_header_one(a, (uint64_t)b * 256 + tid);
    b -= m.h;
    if (b < m.r) return emit_vruns_from(a, (uint64_t)b * 256 + tid, (uint64_t)m.r * 256);
    b -= m.r;
    if (b < m.w) return emit_vraw_one(a, (uint64_t)b * 256 + tid);
    b -= m.w;
    if (b < m.f) return emit_frame_one(a, (uint64_t)b * 256 + tid);
    b -= m.f;
    tile_desc_one(a, (uint64_t)b * 256 + tid);
}

__constant__ double c_emit_fm_scales[7] = {10.0, 100.0, 1000.0, 10000.0, 100000.0, 1000000.0,
                                           10000000.0};

// FoR / DeltaBitpack / FloatMult tiles with w <= 32 (the 32-lane FastLanes form,
// src/bitpack.rs:47-75): lane l of half h of the wavefront takes rows
// [16 h, 16 h + 16) of FastLanes column l, i.e. tile values (16 h + i) 32 + l.
//
// Residual of tile value k: FoR x - min; Delta zigzag(x - prev) - min with the
// block's first value as its own predecessor (delta_bitpack.rs:23-46);
// FloatMult round(x * scale) - min (float_mult.rs:66-98); zero past cnt
// (padding to 1024, for_bitpack.rs:60-72). Loads past cnt read the tile's last
// value instead (clamped index, no branch) and are masked afterwards.
template <int LK>
__device__ __forceinline__ void fl_residuals(const TileEmit &e, int lane, uint32_t (&res)[16]) {
    const void *lp = e.lp;
    const uint32_t lk = e.lk, cnt = e.cnt, l = lane & 31, r0 = 16 * (lane >> 5);
    const uint64_t k0 = e.k0, mn = e.mn;
    const bool delta = e.mode == QA_M_DELTA, fm = e.mode == QA_M_FLOAT_MULT;
    uint64_t xs[16], ps[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t k = (r0 + i) * 32 + l, kk = k < cnt ? k : cnt - 1;
        xs[i] = lane_get_g<LK>(lp, lk, k0 + kk);
        if (delta) ps[i] = lane_get_g<LK>(lp, lk, k0 + kk - ((e.first && kk == 0) ? 0 : 1));
    }
    if (fm) {
        const double sc = c_emit_fm_scales[e.scale];
#pragma unroll
        for (int i = 0; i < 16; i++)  // k = round(x * s), half away from zero, no contraction
            xs[i] = (uint64_t)(int64_t)round(__dmul_rn(__longlong_as_double((long long)xs[i]), sc));
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t k = (r0 + i) * 32 + l;
        const uint32_t r = (uint32_t)(delta ? qa_zigzag(xs[i] - ps[i]) - mn : xs[i] - mn);
        res[i] = k < cnt ? r : 0u;
    }
}

// Pack at width W (a compile-time constant, so every shift and word index is
// fixed): word (k, l) of the form holds bits [32 k, 32 k + 32) of column l's
// rows laid end to end at W bits each. Half h holds bits [16 h W, 16 (h + 1) W)
// of each column, its words L[j] are computed in registers and written to LDS
// once, with no atomics. For odd W half 1's bits start at bit 16 of word
// (W - 1) / 2, the one word both halves share: half 1's part of it crosses to
// half 0 through a lane swap.
template <int W, int I, int NJ>
__device__ __forceinline__ void fl_pack_rows(const uint32_t (&res)[16], uint32_t (&L)[NJ]) {
    if constexpr (I < 16) {
        constexpr int bit = I * W, j = bit >> 5, sh = bit & 31;
        if constexpr (sh == 0) L[j] = res[I];  // first row in this word
        else L[j] |= res[I] << sh;
        if constexpr (sh + W > 32) L[j + 1] = res[I] >> (32 - sh);
        fl_pack_rows<W, I + 1, NJ>(res, L);
    }
}

template <int W, int J, int NJ>
__device__ __forceinline__ void fl_store_even(const uint32_t (&L)[NJ], uint32_t *outw, uint32_t base) {
    if constexpr (J < NJ) {
        outw[base + J * 32] = L[J];
        fl_store_even<W, J + 1, NJ>(L, outw, base);
    }
}

template <int W, int J, int NJ>
__device__ __forceinline__ void fl_shift_odd(const uint32_t (&L)[NJ], uint32_t (&M)[NJ], bool h) {
    if constexpr (J < NJ) {
        uint32_t prev = 0;
        if constexpr (J > 0) prev = L[J - 1];
        M[J] = h ? __builtin_amdgcn_alignbit(L[J], prev, 16) : L[J];
        fl_shift_odd<W, J + 1, NJ>(L, M, h);
    }
}

template <int W, int T, int NJ>
__device__ __forceinline__ void fl_store_odd(const uint32_t (&M)[NJ], uint32_t *outw, uint32_t base,
                                             bool h) {
    if constexpr (T + 1 < NJ) {
        outw[base + T * 32] = h ? M[T + 1] : M[T];
        fl_store_odd<W, T + 1, NJ>(M, outw, base, h);
    }
}

template <int W>
__device__ __forceinline__ void fl_pack_w(const uint32_t (&res)[16], uint32_t *outw, int lane) {
    constexpr int NJ = (16 * W + 31) / 32;
    const uint32_t l = lane & 31;
    const bool h = lane >= 32;
    uint32_t L[NJ];
    fl_pack_rows<W, 0, NJ>(res, L);
    if constexpr (W % 2 == 0) {
        fl_store_even<W, 0, NJ>(L, outw, (h ? W / 2 : 0) * 32 + l);
    } else {
        // half 1, shifted up 16 bits: M[j] = bits of words (W - 1) / 2 + j
        uint32_t M[NJ];
        fl_shift_odd<W, 0, NJ>(L, M, h);
        const uint32_t x = (uint32_t)__shfl_xor((int)(h ? M[0] : M[NJ - 1]), 32);
        // half 0 writes words 0 .. NJ - 2, half 1 words (W + 1) / 2 .. W - 1
        fl_store_odd<W, 0, NJ>(M, outw, (h ? (W + 1) / 2 : 0) * 32 + l, h);
        if (!h) outw[(NJ - 1) * 32 + l] = M[NJ - 1] | x;  // the shared word
    }
}

__device__ __forceinline__ void fl_pack(uint32_t w, const uint32_t (&res)[16], uint32_t *outw, int lane) {
    switch (w) {
#define QA_FL_PACK(n) \
    case n: fl_pack_w<n>(res, outw, lane); break;
    QA_FL_PACK(1) QA_FL_PACK(2) QA_FL_PACK(3) QA_FL_PACK(4) QA_FL_PACK(5) QA_FL_PACK(6)
    QA_FL_PACK(7) QA_FL_PACK(8) QA_FL_PACK(9) QA_FL_PACK(10) QA_FL_PACK(11) QA_FL_PACK(12)
    QA_FL_PACK(13) QA_FL_PACK(14) QA_FL_PACK(15) QA_FL_PACK(16) QA_FL_PACK(17) QA_FL_PACK(18)
    QA_FL_PACK(19) QA_FL_PACK(20) QA_FL_PACK(21) QA_FL_PACK(22) QA_FL_PACK(23) QA_FL_PACK(24)
    QA_FL_PACK(25) QA_FL_PACK(26) QA_FL_PACK(27) QA_FL_PACK(28) QA_FL_PACK(29) QA_FL_PACK(30)
    QA_FL_PACK(31) QA_FL_PACK(32)
#undef QA_FL_PACK
    default: break;
    }
}

// One wavefront per tile, WPB wavefronts per workgroup, 4 KiB of LDS per
// wavefront for the packed words; then the sub-block [w][min][128 w bytes].
template <int WPB>
__global__ void __launch_bounds__(64 * WPB) k_emit_fl(EncArgs a) {
    __shared__ __align__(16) uint32_t outs[WPB][QA_TILE];
    const int lane = threadIdx.x & 63, wv = qa_wave_id();
    const uint64_t t = a.t0 + (uint64_t)blockIdx.x * WPB + wv;
    if (t >= a.t1) return;
    const TileEmit e = a.temit[t];
    if (!e.fast) return;
    const uint32_t w = e.w;
    uint8_t *q = e.dst;
    if (lane == 0) q[0] = (uint8_t)w;
    if (lane >= 1 && lane < 9) q[lane] = (uint8_t)(e.mn >> (8 * (lane - 1)));
    if (w == 0 || e.cnt == 0) return;  // cnt >= 1 on every active tile; the clamp below needs it
    uint32_t res[16];
    switch (e.lk) {
    case QA_LK_U64: fl_residuals<QA_LK_U64>(e, lane, res); break;
    case QA_LK_U32: fl_residuals<QA_LK_U32>(e, lane, res); break;
    case QA_LK_I32: fl_residuals<QA_LK_I32>(e, lane, res); break;
    case QA_LK_F32: fl_residuals<QA_LK_F32>(e, lane, res); break;
    default: fl_residuals<QA_LK_ANY>(e, lane, res);
    }
    uint32_t *outw = outs[wv];
    fl_pack(w, res, outw, lane);
    wave_lds_sync();
    wave_store_bytes(q + 9, outw, 128 * w, lane);
}

// The other tiles (RAW, and FoR / Delta with w > 32), from the list k_tile_desc
// builds: persistent wavefronts, four per workgroup (LDS is per-wavefront, so
// only wavefront-level synchronization is used). FastLanes pack
// (src/bitpack.rs:47-133) straight from the lane: in the u32 form lane l of the
// 32 packs rows [16 h, 16 h + 16) (h = half of the wavefront), in the u64 form
// lane l of the 16 packs rows [16 q, 16 q + 16) (q = quarter); each residual is
// ORed into its one or two LDS words.
__global__ void __launch_bounds__(256) k_emit_tiles(EncArgs a) {
    __shared__ uint64_t outs[4][QA_TILE + 1];
    const int lane = threadIdx.x & 63, wv = qa_wave_id();
    const uint32_t n = *a.tlist;
    for (uint32_t i = blockIdx.x * 4 + wv; i < n; i += gridDim.x * 4) {
        const TileEmit e = a.temit[a.tlist[16 + i]];
        switch (e.lk) {
        case QA_LK_U64: emit_tile_t<QA_LK_U64>(e, outs[wv], lane); break;
        case QA_LK_U32: emit_tile_t<QA_LK_U32>(e, outs[wv], lane); break;
        case QA_LK_I32: emit_tile_t<QA_LK_I32>(e, outs[wv], lane); break;
        case QA_LK_F32: emit_tile_t<QA_LK_F32>(e, outs[wv], lane); break;
        default: emit_tile_t<QA_LK_ANY>(e, outs[wv], lane);
        }
        wave_lds_sync();  // the next tile reuses the wavefront's LDS
    }
}

__global__ void k_tlist_clear(EncArgs a) {
    if (threadIdx.x == 0) a.tlist[0] = 0;
}

// ---------------------------------------------------------------------------
// launcher
// ---------------------------------------------------------------------------
static inline unsigned nb(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }
// launch unless the grid is empty (a zero-size grid is a launch error)
#define QA_LAUNCH(k, grid, ...)                                   \
    do {                                                          \
        dim3 g_ = (grid);                                         \
        if (g_.x) qa_launch(k, g_, __VA_ARGS__);         \
    } while (0)

// column-level stages: validity counts, compaction, region statistics, plan
static void encode_counts_launch(const EncArgs &a, hipStream_t s) {
    if (a.nvwords) {
        QA_LAUNCH(k_vcount, dim3(nb(a.nvwords, 256)), dim3(256), 0, s, a);
        qa_exclusive_scan2(a.vpre, a.vpre, a.tpre, a.tpre, a.nvwords, a.scan_tmp, s);
    }
    QA_LAUNCH(k_col_counts, dim3(nb(a.ncols > 64 ? a.ncols : 64, 64)), dim3(64), 0, s, a);
}

static void encode_stats_launch(const EncArgs &a, hipStream_t s) {
    if (a.nvwords) QA_LAUNCH(k_compact, dim3(nb(a.nvwords, 4 * WAVE)), dim3(256), 0, s, a);
    QA_LAUNCH(k_region_stats, dim3(a.nregions), dim3(512), 0, s, a);
    QA_LAUNCH(k_plan, dim3(a.ncols), dim3(256), 0, s, a);
}

// the validity blob's runs: transition positions, varint sizes, their prefix
// (reads only the counts and transition prefix; writes tpos and rsz, which
// nothing reads before k_col_layout)
static void encode_vruns_launch(const EncArgs &a, hipStream_t s) {
    if (!a.nvwords) return;
    const uint64_t nruns_max = a.ntrans_max + a.ncols;
    QA_LAUNCH(k_vtrans_pos, dim3(nb(a.nvwords, 256)), dim3(256), 0, s, a);
    unsigned g = nb(nruns_max, 256) < 4096 ? nb(nruns_max, 256) : 4096;
    QA_LAUNCH(k_vrun_sizes, dim3(g), dim3(256), 0, s, a);
    qa_exclusive_scan_dn(a.rsz, nruns_max, a.colbase + a.ncols + 1, a.rsz, a.scan_tmp, s);
}

static void encode_layout_rest_launch(const EncArgs &a, hipStream_t s) {
    const uint64_t nruns_max = a.ntrans_max + a.ncols;
    QA_LAUNCH(k_col_layout, dim3(a.ncols), dim3(256), 0, s, a);
    QA_LAUNCH(k_emit_header, dim3(nb(a.ncols, 64)), dim3(64), 0, s, a);
    if (a.nvwords) {
        unsigned g = nb(nruns_max, 256) < 4096 ? nb(nruns_max, 256) : 4096;
        QA_LAUNCH(k_emit_vruns, dim3(g), dim3(256), 0, s, a);
        QA_LAUNCH(k_emit_vraw, dim3(nb(a.nvwords, 256)), dim3(256), 0, s, a);
    }
}

void qa_encode_plan_launch(const EncArgs &a, hipStream_t s) {
    encode_counts_launch(a, s);
    encode_stats_launch(a, s);
}

// per-block stages over region slots [r0, r1) and tiles [t0, t1): features of
// grown blocks, candidate sizes, the competition
void qa_encode_blocks_launch(const EncArgs &a, hipStream_t s) {
    QA_LAUNCH(k_grown_feats, dim3(a.r1 - a.r0), dim3(256), 0, s, a);
    qa_fast_sizes_launch(a, s);
    qa_fast_final_launch(a, s);
}

// column-level stages: validity blob, frame layout, header
void qa_encode_layout_launch(const EncArgs &a, hipStream_t s) {
    encode_vruns_launch(a, s);
    encode_layout_rest_launch(a, s);
}

// the payloads of tiles [t0, t1): k_emit_fl, k_emit_tiles, the Fast pool's
static void encode_payload_launch(const EncArgs &a, hipStream_t s) {
    // k_emit_fl at the handle's tuned wavefronts per workgroup (tuned here on
    // the first batch with enough tiles: the stage is idempotent)
    auto emit_fl = [&](uint32_t wpb) {
        const uint64_t n = a.t1 - a.t0;
        switch (wpb) {
        case 1: QA_LAUNCH(k_emit_fl<1>, dim3(nb(n, 1)), dim3(64), 0, s, a); break;
        case 2: QA_LAUNCH(k_emit_fl<2>, dim3(nb(n, 2)), dim3(128), 0, s, a); break;
        case 8: QA_LAUNCH(k_emit_fl<8>, dim3(nb(n, 8)), dim3(512), 0, s, a); break;
        default: QA_LAUNCH(k_emit_fl<4>, dim3(nb(n, 4)), dim3(256), 0, s, a);
        }
    };
    if (a.tune && !a.tune[QA_TUNE_EMIT] && a.t1 - a.t0 >= QA_TUNE_MIN_TILES)
        a.tune[QA_TUNE_EMIT] = qa_tune_pick(emit_fl, s);
    qa_mark(a.kev0, s);
    emit_fl(a.tune && a.tune[QA_TUNE_EMIT] ? a.tune[QA_TUNE_EMIT] : 4);
    {
        // persistent over the list: at most four groups per CU
        const unsigned g = nb(a.t1 - a.t0, 4);
        QA_LAUNCH(k_emit_tiles, dim3(g < 1024 ? g : 1024), dim3(256), 0, s, a);
    }
    qa_mark(a.kev1, s);
    qa_fast_emit_launch(a, s);
}

// per-block stages over [r0, r1) / [t0, t1): frame headers and payloads
void qa_encode_emit_launch(const EncArgs &a, hipStream_t s) {
    QA_LAUNCH(k_emit_frames, dim3(nb(a.r1 - a.r0, 64)), dim3(64), 0, s, a);
    qa_launch(k_tlist_clear, dim3(1), dim3(64), 0, s, a);
    QA_LAUNCH(k_tile_desc, dim3(nb(a.t1 - a.t0, 256)), dim3(256), 0, s, a);
    encode_payload_launch(a, s);
}

// the batched form: every block of every column in one launch per stage
// The batched sequence. With a side stream, the validity blob's run chain
// (latency-bound small kernels) runs there, forked after the transition
// counts and joined before the frame layout, beside compaction, statistics,
// planning and the competition; both streams are at the default priority.
void qa_encode_launch(const EncArgs &a, hipStream_t s, int, hipStream_t side, hipEvent_t fork,
                      hipEvent_t join) {
    encode_counts_launch(a, s);
    // a fork that cannot be set up runs the chain on the main stream; a join
    // that cannot be enqueued waits for the side stream on the host
    bool forked = side && a.nvwords && hipEventRecord(fork, s) == hipSuccess &&
                  hipStreamWaitEvent(side, fork, 0) == hipSuccess;
    bool joined = false;
    if (forked) {
        encode_vruns_launch(a, side);
        joined = hipEventRecord(join, side) == hipSuccess;
    }
    encode_stats_launch(a, s);
    qa_encode_blocks_launch(a, s);
    if (forked) {
        if (!joined || hipStreamWaitEvent(s, join, 0) != hipSuccess) hipStreamSynchronize(side);
    } else {
        encode_vruns_launch(a, s);
    }
    // the frame layout (it also clears k_emit_tiles' list), then the header,
    // validity blob, frame headers and tile descriptors in one launch
    QA_LAUNCH(k_col_layout, dim3(a.ncols), dim3(256), 0, s, a);
    MetaSplit m{};
    m.h = nb(a.ncols, 256);
    if (a.nvwords) {
        const uint64_t nruns_max = a.ntrans_max + a.ncols;
        m.r = nb(nruns_max, 256) < 4096 ? nb(nruns_max, 256) : 4096;
        m.w = nb(a.nvwords, 256);
    }
    m.f = nb(a.r1 - a.r0, 256);
    QA_LAUNCH(k_emit_meta, dim3(m.h + m.r + m.w + m.f + nb(a.t1 - a.t0, 256)), dim3(256), 0, s, a,
              m);
    encode_payload_launch(a, s);
}
