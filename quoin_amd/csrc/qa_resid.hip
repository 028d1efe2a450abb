// qa_resid.hip — the linear residual coders of src/codecs/linear.rs on the
// device, batched over independent blocks:
//   OrderedDelta (mode 7)  varint(zigzag(v - pred)), pred in {0, p1, 2 p1 - p2}
//                          wrapping (linear.rs:31-66);
//   Delta2 (mode 20)       [order] ++ varint(bits(v) ^ bits(pred)) with an f64
//                          polynomial prediction (linear.rs:125-180);
//   DeltaDp (mode 45)      [order] ++ varint(bits(r) ^ bits(r_prev)), r = v - pred,
//                          verified to add back exactly (linear.rs:191-236).
//
// Encode is a map from each value and the four before it, so it runs on
// 2048-value tiles: k_lin_order picks each block's order (select_order: four
// finite-difference rounds in LDS, then four lanes each run one order's
// left-to-right sum), k_lin_enc<false> sizes every tile's varints, a prefix sum
// places the tiles, and k_lin_enc<true> recomputes the residuals, stages the
// tile's bytes in LDS at the alignment of their destination and writes them
// with 16-byte stores.
//
// Decode splits the varint parse from the recurrence. The parse runs on
// 4096-byte tiles: terminator counts, a prefix sum, then every terminator byte
// assembles its value from the bytes before it. OrderedDelta is then two
// wrapping prefix sums (one workgroup per block); Delta2 and DeltaDp are f64
// recurrences, one lane per block.
//
// f64 results follow x86-64 SSE2 scalar semantics: a NaN result is the first
// NaN operand made quiet, else the default NaN 0xFFF8000000000000. The
// prediction runs on the hardware first; only a NaN prediction (never the
// case without NaN or infinite history) is recomputed step by step with that
// rule. The file is compiled without FP contraction.
#include <rocprim/device/device_radix_sort.hpp>

#include "qa_kernels.h"

#include <type_traits>

#define LIN_OD 7u
#define LIN_D2 20u
#define LIN_DP 45u
#define LE_TILE 2048  // values per encode tile
#define LE_WG 256
#define LE_PER 8
#define LD_TILE 4096  // stream bytes per decode tile
#define LD_WG 256
#define LD_PER 16
#define LIN_DNAN 0xFFF8000000000000ull

namespace {

__device__ __forceinline__ double b2d(uint64_t b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ uint64_t d2b(double x) { return (uint64_t)__double_as_longlong(x); }
__device__ __forceinline__ bool is_nan(double x) { return x != x; }

// x86-64 SSE2 result of op(a, b) whose hardware result is r
__device__ __forceinline__ double sse_fix(double r, double a, double b) {
    if (!is_nan(r)) return r;
    if (is_nan(a)) return b2d(d2b(a) | (1ull << 51));
    if (is_nan(b)) return b2d(d2b(b) | (1ull << 51));
    return b2d(LIN_DNAN);
}
__device__ __forceinline__ double sse_add(double a, double b) {
    const double r = a + b;
    return is_nan(r) ? sse_fix(r, a, b) : r;
}
__device__ __forceinline__ double sse_sub(double a, double b) {
    const double r = a - b;
    return is_nan(r) ? sse_fix(r, a, b) : r;
}

// Newton coefficients COEFFS[eff][k] (linear.rs:75)
__device__ __forceinline__ double lin_coef(uint32_t eff, uint32_t k) {
    if (k == 0) return (double)eff;
    if (k == 1) return eff == 2 ? -1.0 : eff == 3 ? -3.0 : -6.0;
    if (k == 2) return eff == 3 ? 1.0 : 4.0;
    return -1.0;
}

// predict_f64 (linear.rs:83-94) step by step with the SSE2 NaN rule
__device__ __noinline__ double pred_sse(uint32_t eff, double h0, double h1, double h2, double h3) {
    const double h[4] = {h0, h1, h2, h3};
    double p = 0.0;
    for (uint32_t k = 0; k < eff; k++) {
        const double c = lin_coef(eff, k);
        const double m = sse_fix(c * h[k], c, h[k]);
        p = sse_fix(p + m, p, m);
    }
    return p;
}

// predict_f64: pred = 0.0; pred += c_k h_k for k < eff, h0 the latest value.
// Without a NaN anywhere the hardware steps are the SSE2 steps; NaN is
// absorbing, so a non-NaN result proves no step needed the rule.
// Each term is computed and then selected (eff differs between the lanes'
// blocks), so the recurrence's step has no branch; a term that is not
// selected cannot reach the result.
__device__ __forceinline__ double lin_pred(uint32_t eff, double h0, double h1, double h2,
                                           double h3) {
    double p = 0.0 + (double)eff * h0;
    const double p2 = p + lin_coef(eff, 1) * h1;
    p = eff >= 2 ? p2 : p;
    const double p3 = p + lin_coef(eff, 2) * h2;
    p = eff >= 3 ? p3 : p;
    const double p4 = p - h3;
    p = eff >= 4 ? p4 : p;
    p = eff == 0 ? 0.0 : p;
    if (__builtin_expect(is_nan(p), 0)) p = pred_sse(eff, h0, h1, h2, h3);
    return p;
}

// lin_pred at a constant effective order E: the same operations in the same
// order, with the coefficients as constants
template <int E>
__device__ __forceinline__ double lin_pred_c(double h0, double h1, double h2, double h3) {
    if constexpr (E == 0) {
        return 0.0;
    } else {
        double p = 0.0 + (double)E * h0;
        if constexpr (E >= 2) p = p + (E == 2 ? -1.0 : E == 3 ? -3.0 : -6.0) * h1;
        if constexpr (E >= 3) p = p + (E == 3 ? 1.0 : 4.0) * h2;
        if constexpr (E >= 4) p = p - h3;
        if (__builtin_expect(is_nan(p), 0)) p = pred_sse(E, h0, h1, h2, h3);
        return p;
    }
}

// the block of a batch owning global tile t: the last b with tile0 <= t
__device__ __forceinline__ uint32_t owner_block(const LinD *d, uint32_t nblk, uint64_t t) {
    uint32_t lo = 0, hi = nblk;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (d[mid].tile0 <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

// workgroup exclusive scan of one value per thread (256 threads); *total gets
// the sum. Ends with a barrier, so LDS may be reused after it.
__device__ __forceinline__ uint64_t wg_excl_scan(uint64_t v, uint64_t *total) {
    __shared__ uint64_t ws[LE_WG / 64 + 1];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t inc = qa_wave_scan_add64(v);
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t s = 0;
        for (uint32_t k = 0; k < LE_WG / 64; k++) {
            const uint64_t x = ws[k];
            ws[k] = s;
            s += x;
        }
        ws[LE_WG / 64] = s;
    }
    __syncthreads();
    const uint64_t r = inc - v + ws[w];
    *total = ws[LE_WG / 64];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// encode
// ---------------------------------------------------------------------------

// select_order (linear.rs:125-143) and the per-block result slot. One
// workgroup per block.
template <uint32_t KIND>
__global__ void __launch_bounds__(256) k_lin_order(const LinD *d, LinR *r) {
    __shared__ double df[4][1024 + 2];  // padded: the four summing lanes hit distinct banks
    __shared__ double mag[4];
    const uint32_t b = blockIdx.x, tid = threadIdx.x;
    const uint64_t n = d[b].n;
    uint32_t order = d[b].order;
    if (KIND != LIN_OD && order == 0) {
        if (n < 8) {
            order = 2;
        } else {
            const uint32_t win = n < 1024 ? (uint32_t)n : 1024u;
            const uint64_t *v = d[b].vals + (n - win) / 2;
            // diff_in_place (linear.rs:105-111): d[i] <- d[i+1] - d[i]
            for (uint32_t i = tid; i + 1 < win; i += 256) df[0][i] = b2d(v[i + 1]) - b2d(v[i]);
            __syncthreads();
            for (uint32_t o = 1; o < 4; o++) {
                for (uint32_t i = tid; i + 1 + o < win; i += 256)
                    df[o][i] = df[o - 1][i + 1] - df[o - 1][i];
                __syncthreads();
            }
            if (tid < 4) {
                // mean_abs (linear.rs:113-119): a left-to-right sum, then / len
                const uint32_t len = win - 1 - tid;
                double s = 0.0;
#pragma unroll 8
                for (uint32_t i = 0; i < win - 1; i++) {
                    const double x = fabs(df[tid][i]);
                    s = i < len ? s + x : s;
                }
                mag[tid] = s / (double)len;
            }
            __syncthreads();
            order = 1;
            double best = __builtin_inf();
            for (uint32_t o = 0; o < 4; o++) {
                const double m = mag[o];
                if (isfinite(m) && m < best) {
                    best = m;
                    order = o + 1;
                }
            }
        }
    }
    if (tid == 0) {
        LinR x;
        x.len = 0;
        x.status = 0;
        x.order = KIND == LIN_OD ? 0u : order;
        if (n == 0) {
            x.len = KIND == LIN_D2 ? 1 : 0;  // Delta2 of nothing is [order]
            x.status = KIND == LIN_DP ? 1 : 0;  // dp_encode returns None (linear.rs:192-194)
        }
        r[b] = x;
        if (KIND == LIN_D2 || (KIND == LIN_DP && n > 0)) d[b].dst[0] = (uint8_t)order;
    }
}

// Residuals of one 2048-value tile. WRITE = false: tile byte counts into
// tsz[t] (and DeltaDp's verification); WRITE = true: tsz holds their exclusive
// prefix over the batch and the bytes are written.
// A tile's residual words (thread j holds words j0 .. j0 + LE_PER) as LEB128
// varints (varint.rs:9-21). Sizing pass (!WRITE): the tile's byte count goes
// to tsz[t]. Writing pass: tsz is the exclusive prefix over the tiles; the
// bytes are staged in LDS at the alignment of their destination and written
// with 16-byte stores, and the block's first tile records its stream length
// (hdr header bytes before the varints).
template <bool WRITE>
__device__ __forceinline__ void tile_varints(const LinD &db, LinR &rb, uint64_t t, uint32_t cnt,
                                             uint32_t j0, const uint64_t (&res)[LE_PER],
                                             uint64_t len, uint64_t *tsz, uint64_t hdr) {
    const uint64_t tile0 = db.tile0, n = db.n;
    if (!WRITE) {
        uint64_t tot;
        wg_excl_scan(len, &tot);
        if (threadIdx.x == 0) tsz[t] = tot;
        return;
    } else {
        __shared__ __align__(16) uint8_t stg[LE_TILE * 10 + 32];
        uint64_t tot;
        const uint64_t ex = wg_excl_scan(len, &tot);
        uint8_t *g0 = db.dst + hdr + (tsz[t] - tsz[tile0]);
        const uint32_t a0 = (uint32_t)((uintptr_t)g0 & 15);
        if (j0 < cnt) {
            uint32_t p = a0 + (uint32_t)ex;
#pragma unroll
            for (int m = 0; m < LE_PER; m++) {
                if (j0 + m >= cnt) break;
                uint64_t v = res[m];
                while (v >= 0x80) {
                    stg[p++] = (uint8_t)(v | 0x80);
                    v >>= 7;
                }
                stg[p++] = (uint8_t)v;
            }
        }
        __syncthreads();
        uint8_t *gb = g0 - a0;  // 16-byte aligned
        const uint32_t end = a0 + (uint32_t)tot;
        const uint32_t nch = (end + 15) / 16;
        for (uint32_t c = threadIdx.x; c < nch; c += LE_WG) {
            const uint32_t lo = c * 16, hi = lo + 16;
            if (lo >= a0 && hi <= end) {
                *(uint4 *)(gb + lo) = *(const uint4 *)(stg + lo);
            } else {
                for (uint32_t q = max(lo, a0); q < min(hi, end); q++) gb[q] = stg[q];
            }
        }
        if (t == tile0 && threadIdx.x == 0) {
            const uint64_t nt = (n + LE_TILE - 1) / LE_TILE;
            rb.len = tsz[tile0 + nt] - tsz[tile0] + hdr;
        }
    }
}

template <uint32_t KIND, bool WRITE>
__global__ void __launch_bounds__(LE_WG) k_lin_enc(const LinD *d, LinR *r, uint32_t nblk,
                                                   uint64_t *tsz) {
    __shared__ uint64_t sv[LE_TILE + 8];  // sv[k]: value s0 - 5 + k
    const uint64_t t = blockIdx.x;
    const uint32_t b = owner_block(d, nblk, t);
    if (KIND == LIN_DP && WRITE && r[b].status) return;  // the block bailed
    const uint64_t *vals = d[b].vals;
    const uint64_t n = d[b].n, tile0 = d[b].tile0;
    const uint64_t s0 = (t - tile0) * LE_TILE;
    const uint32_t cnt = (uint32_t)min((uint64_t)LE_TILE, n - s0);
    const uint32_t order = r[b].order;
    for (uint32_t k = threadIdx.x; k < cnt + 5; k += LE_WG) {
        const int64_t i = (int64_t)s0 - 5 + (int64_t)k;
        sv[k] = i >= 0 ? vals[i] : 0;
    }
    __syncthreads();

    const uint32_t j0 = threadIdx.x * LE_PER;
    uint64_t res[LE_PER];
    uint64_t len = 0;
    bool bad = false;
    if (j0 < cnt) {
        uint64_t w[LE_PER + 5];
#pragma unroll
        for (int m = 0; m < LE_PER + 5; m++) w[m] = sv[j0 + m];
        const uint64_t ib = s0 + j0;  // block index of w[5]
        if (KIND == LIN_OD) {
#pragma unroll
            for (int m = 0; m < LE_PER; m++) {
                const uint64_t i = ib + m;
                const uint64_t pred = i == 0 ? 0 : i == 1 ? w[m + 4] : 2 * w[m + 4] - w[m + 3];
                res[m] = qa_zigzag(w[m + 5] - pred);
            }
        } else {
            uint64_t prev = 0;
            if (KIND == LIN_DP && ib > 0) {  // residual of the value before, for the chain
                const uint64_t i = ib - 1;
                const uint32_t eff = i < 4 ? min(order, (uint32_t)i) : order;
                const double p = lin_pred(eff, b2d(w[3]), b2d(w[2]), b2d(w[1]), b2d(w[0]));
                prev = d2b(sse_sub(b2d(w[4]), p));
            }
#pragma unroll
            for (int m = 0; m < LE_PER; m++) {
                const uint64_t i = ib + m;
                const uint32_t eff = i < 4 ? min(order, (uint32_t)i) : order;
                const double p =
                    lin_pred(eff, b2d(w[m + 4]), b2d(w[m + 3]), b2d(w[m + 2]), b2d(w[m + 1]));
                if (KIND == LIN_D2) {
                    res[m] = w[m + 5] ^ d2b(p);
                } else {
                    const double rr = sse_sub(b2d(w[m + 5]), p);
                    const bool ok = d2b(sse_add(p, rr)) == w[m + 5];
                    bad |= (j0 + m < cnt) && !ok;
                    res[m] = d2b(rr) ^ prev;
                    prev = d2b(rr);
                }
            }
        }
#pragma unroll
        for (int m = 0; m < LE_PER; m++)
            if (j0 + m < cnt) len += qa_varint_len(res[m]);
    }

    if (!WRITE && KIND == LIN_DP) {
        const uint64_t any = __ballot(bad);
        if (any && (threadIdx.x & 63) == 0) atomicOr(&r[b].status, 1);
    }
    tile_varints<WRITE>(d[b], r[b], t, cnt, j0, res, len, tsz, KIND == LIN_OD ? 0 : 1);
}

// ---------------------------------------------------------------------------
// decode
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t content_len(const LinD &x, uint64_t hdr) {
    return x.len > hdr ? x.len - hdr : 0;
}

// terminator bytes (high bit clear) of each 4096-byte tile
template <uint32_t KIND>
__global__ void __launch_bounds__(LD_WG) k_lin_vcount(const LinD *d, uint32_t nblk, uint64_t *cnt) {
    const uint64_t t = blockIdx.x;
    const uint32_t b = owner_block(d, nblk, t);
    const uint64_t hdr = KIND == LIN_OD ? 0 : 1;
    const uint8_t *c = d[b].src + hdr;
    const uint64_t b0 = (t - d[b].tile0) * LD_TILE;
    const uint32_t m = (uint32_t)min((uint64_t)LD_TILE, content_len(d[b], hdr) - b0);
    uint64_t k = 0;
    for (uint32_t q = threadIdx.x; q < m; q += LD_WG) k += c[b0 + q] < 0x80;
    uint64_t tot;
    wg_excl_scan(k, &tot);
    if (threadIdx.x == 0) cnt[t] = tot;
}

// every terminator assembles its varint (read_u64, varint.rs:24-39) from the
// bytes before it and stores it at its rank: out[j] = residual word j
template <uint32_t KIND>
__global__ void __launch_bounds__(LD_WG) k_lin_vparse(const LinD *d, LinR *r, uint32_t nblk,
                                                      const uint64_t *pre) {
    __shared__ __align__(16) uint8_t sb[16 + LD_TILE];  // sb[16 + q]: content byte b0 + q
    const uint64_t t = blockIdx.x;
    const uint32_t b = owner_block(d, nblk, t);
    const uint64_t hdr = KIND == LIN_OD ? 0 : 1;
    const uint8_t *c = d[b].src + hdr;
    const uint64_t n = d[b].n, tile0 = d[b].tile0;
    uint64_t *out = d[b].out;
    const uint64_t b0 = (t - tile0) * LD_TILE;
    const uint32_t m = (uint32_t)min((uint64_t)LD_TILE, content_len(d[b], hdr) - b0);
    for (uint32_t q = threadIdx.x; q < m; q += LD_WG) sb[16 + q] = c[b0 + q];
    if (threadIdx.x < 16) {
        const int64_t g = (int64_t)b0 - 16 + (int64_t)threadIdx.x;
        sb[threadIdx.x] = g >= 0 ? c[g] : 0;
    }
    __syncthreads();
    const uint32_t q0 = threadIdx.x * LD_PER;
    uint32_t term = 0;  // bit k: byte q0 + k ends a varint
    if (q0 < m) {
        const uint4 u = *(const uint4 *)(sb + 16 + q0);
        const uint32_t wv[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int k = 0; k < 16; k++)
            if (q0 + k < m && !((wv[k >> 2] >> (8 * (k & 3) + 7)) & 1)) term |= 1u << k;
    }
    uint64_t tot;
    uint64_t j = wg_excl_scan((uint64_t)__popc(term), &tot) + (pre[t] - pre[tile0]);
    bool bad = false;
    while (term) {
        const uint32_t k = __ffs(term) - 1;
        term &= term - 1;
        const int32_t q = (int32_t)(q0 + k);
        // continuation bytes before the terminator, back to the content start
        uint32_t cn = 0;
        while (cn < 10 && (int64_t)b0 + q - (int64_t)cn - 1 >= 0 && sb[16 + q - cn - 1] >= 0x80) cn++;
        if (cn == 10) {
            bad = true;  // leb128 overflow (varint.rs:35-37)
        } else {
            uint64_t x = 0;
            for (uint32_t i = 0; i <= cn; i++)
                x |= (uint64_t)(sb[16 + q - cn + i] & 0x7f) << (7 * i);
            if (j < n) out[j] = x;
        }
        j++;
    }
    if (__ballot(bad) && (threadIdx.x & 63) == 0) atomicOr(&r[b].status, 1);
}

// per block: the order byte, the varint count and the last byte
template <uint32_t KIND>
__global__ void __launch_bounds__(64) k_lin_dcheck(const LinD *d, LinR *r, uint32_t nblk,
                                                   const uint64_t *pre) {
    const uint32_t b = blockIdx.x * 64 + threadIdx.x;
    if (b >= nblk) return;
    const uint64_t hdr = KIND == LIN_OD ? 0 : 1;
    const uint8_t *src = d[b].src;
    const uint64_t len = d[b].len, n = d[b].n, tile0 = d[b].tile0;
    int32_t bad = 0;
    uint32_t order = 0;
    if (KIND != LIN_OD) {
        if (len == 0) {
            bad = 1;  // Truncated (linear.rs:162)
        } else {
            order = src[0];
            if (order < 1 || order > 4) bad = 1;  // linear.rs:164-166
        }
    }
    const uint64_t cl = content_len(d[b], hdr);
    const uint64_t nt = (cl + LD_TILE - 1) / LD_TILE;
    if (pre[tile0 + nt] - pre[tile0] != n) bad = 1;  // too few varints, or trailing bytes
    if (cl && src[hdr + cl - 1] >= 0x80) bad = 1;     // an unfinished varint
    if (bad) atomicOr(&r[b].status, 1);
    r[b].order = order;
}

// OrderedDelta (linear.rs:47-66): with e_i = unzigzag(x_i), a_0 = 0 and
// a_k = e_1 + ... + e_k, v_i = e_0 + a_0 + ... + a_i (all wrapping). One
// workgroup per block, 4096 values a round.
__global__ void __launch_bounds__(LE_WG) k_lin_od(const LinD *d, const LinR *r) {
    const uint32_t b = blockIdx.x;
    if (r[b].status) return;
    const uint64_t n = d[b].n;
    uint64_t *out = d[b].out;
    if (n == 0) return;
    const uint64_t e0 = qa_unzigzag(out[0]);
    __syncthreads();  // every thread has read out[0] before it is rewritten
    uint64_t ca = 0, cv = e0;  // a and v at the end of the previous round
    for (uint64_t c0 = 0; c0 < n; c0 += LE_WG * 16) {
        const uint64_t i0 = c0 + threadIdx.x * 16;
        uint64_t x[16];
        uint64_t s = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const uint64_t i = i0 + k;
            const uint64_t e = (i < n && i > 0) ? qa_unzigzag(out[i]) : 0;
            s += e;
            x[k] = s;  // local inclusive prefix of e
        }
        uint64_t tot1;
        const uint64_t ea = wg_excl_scan(s, &tot1) + ca;
        uint64_t s2 = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const uint64_t a = i0 + k < n ? x[k] + ea : 0;
            s2 += a;
            x[k] = s2;  // local inclusive prefix of a
        }
        uint64_t tot2;
        const uint64_t ev = wg_excl_scan(s2, &tot2) + cv;
#pragma unroll
        for (int k = 0; k < 16; k++)
            if (i0 + k < n) out[i0 + k] = x[k] + ev;
        ca += tot1;
        cv += tot2;
    }
}

// Delta2 (linear.rs:161-180) and DeltaDp (linear.rs:214-236): the
// recurrence over a block's residual words, one lane per block.
template <uint32_t KIND>
__global__ void __launch_bounds__(64) k_lin_seq(const LinD *d, const LinR *r, uint32_t nblk) {
    const uint32_t b = blockIdx.x * 64 + threadIdx.x;
    if (b >= nblk || r[b].status) return;
    const uint64_t n = d[b].n;
    const uint32_t order = r[b].order;
    uint64_t *out = d[b].out;
    double h0 = 0, h1 = 0, h2 = 0, h3 = 0;
    uint64_t prev = 0;
    // one step with the prediction p
    auto step_p = [&](uint64_t x, double p) -> uint64_t {
        uint64_t bits;
        double v;
        if (KIND == LIN_D2) {
            bits = x ^ d2b(p);
            v = b2d(bits);
        } else {
            const uint64_t rb = x ^ prev;
            prev = rb;
            v = sse_add(p, b2d(rb));
            bits = d2b(v);
        }
        h3 = h2;
        h2 = h1;
        h1 = h0;
        h0 = v;
        return bits;
    };
    const uint64_t nw = n < 4 ? n : 4;
    for (uint64_t i = 0; i < nw; i++)
        out[i] = step_p(out[i], lin_pred(min(order, (uint32_t)i), h0, h1, h2, h3));
    // the rest at the block's order: a constant when every block of the
    // wavefront has the same one. The residual words are loaded SEQ_AHEAD
    // values before they are used, so memory latency overlaps the f64
    // dependency chain; loads at a clamped index (n > 4 whenever the loop
    // runs), so the step has no branch.
    constexpr int SEQ_AHEAD = 32;
    auto run = [&](auto E) {
        constexpr int e = decltype(E)::value;
        auto pred = [&]() -> double {
            if constexpr (e < 0) return lin_pred(order, h0, h1, h2, h3);
            else return lin_pred_c<e>(h0, h1, h2, h3);
        };
        uint64_t i = 4;
        uint64_t x[SEQ_AHEAD];
#pragma unroll
        for (int k = 0; k < SEQ_AHEAD; k++) x[k] = i + k < n ? out[i + k] : 0;
        for (; i + SEQ_AHEAD <= n; i += SEQ_AHEAD) {
#pragma unroll
            for (int k = 0; k < SEQ_AHEAD; k++) {
                out[i + k] = step_p(x[k], pred());
                const uint64_t j = i + SEQ_AHEAD + k;
                const uint64_t v = out[j < n ? j : n - 1];
                x[k] = j < n ? v : 0;
            }
        }
        for (int k = 0; k < SEQ_AHEAD; k++)
            if (i + k < n) out[i + k] = step_p(x[k], pred());
    };
    const uint32_t o0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)order);
    if (__ballot(order != o0) == 0) {
        switch (o0) {
        case 0: run(std::integral_constant<int, 0>{}); break;
        case 1: run(std::integral_constant<int, 1>{}); break;
        case 2: run(std::integral_constant<int, 2>{}); break;
        case 3: run(std::integral_constant<int, 3>{}); break;
        case 4: run(std::integral_constant<int, 4>{}); break;
        default: run(std::integral_constant<int, -1>{});
        }
    } else {
        run(std::integral_constant<int, -1>{});
    }
}

template <uint32_t KIND>
void enc_launch(const LinD *d, LinR *r, uint32_t nblk, uint64_t ntiles, uint64_t *tsz,
                uint64_t *tpre, uint64_t *tmp, hipStream_t s) {
    qa_launch(k_lin_order<KIND>, dim3(nblk), dim3(256), 0, s, d, r);
    if (!ntiles) return;
    qa_launch((k_lin_enc<KIND, false>), dim3((unsigned)ntiles), dim3(LE_WG), 0, s, d, r,
                       nblk, tsz);
    qa_exclusive_scan(tsz, ntiles, tpre, tmp, s);
    qa_launch((k_lin_enc<KIND, true>), dim3((unsigned)ntiles), dim3(LE_WG), 0, s, d, r,
                       nblk, tpre);
}

template <uint32_t KIND>
void dec_launch(const LinD *d, LinR *r, uint32_t nblk, uint64_t ntiles, uint64_t *cnt,
                uint64_t *pre, uint64_t *tmp, hipStream_t s, void *kev0, void *kev1) {
    if (ntiles)
        qa_launch(k_lin_vcount<KIND>, dim3((unsigned)ntiles), dim3(LD_WG), 0, s, d, nblk,
                           cnt);
    qa_exclusive_scan(cnt, ntiles, pre, tmp, s);
    if (ntiles)
        qa_launch(k_lin_vparse<KIND>, dim3((unsigned)ntiles), dim3(LD_WG), 0, s, d, r,
                           nblk, pre);
    const unsigned g64 = (nblk + 63) / 64;
    qa_launch(k_lin_dcheck<KIND>, dim3(g64), dim3(64), 0, s, d, r, nblk, pre);
    if (kev0) (void)hipEventRecord((hipEvent_t)kev0, s);
    if constexpr (KIND == LIN_OD)
        qa_launch(k_lin_od, dim3(nblk), dim3(LE_WG), 0, s, d, r);
    else
        qa_launch(k_lin_seq<KIND>, dim3(g64), dim3(64), 0, s, d, r, nblk);
    if (kev1) (void)hipEventRecord((hipEvent_t)kev1, s);
}

// ---------------------------------------------------------------------------
// FCM / DFCM predictors (src/codecs/pred.rs)
// ---------------------------------------------------------------------------
// FCM (mode 0): r_i = v_i ^ T[ctx_i], ctx_0 = 0, ctx_i = crc(v_{i-1}) & mask,
// where T[ctx] holds the value of the last earlier position with the same
// context (0 if none). DFCM (mode 25): r_i = v_i ^ (v_{i-1} + T[ctx_i]) with
// ctx_i = crc(d_{i-1}) & mask, d_j = v_j - v_{j-1} (v_{-1} = 0) and T holding
// the last earlier d at that context. Encode is therefore a "last earlier
// occurrence of the same key" query: each value's key (block << 16 | ctx) and
// flat index are radix-sorted (stable), and a value's predecessor in sorted
// order, if it has the same key, is that occurrence. Decode is the table
// recurrence of pred.rs, one lane per block with its table in HBM.
#define PRED_FCM 0u
#define PRED_DFCM 25u
#define PRED_SEED 0x9e3779b9u  // hash.rs:10
#define PRED_NONE 0xFFFFFFFFu

// CRC32C (reflected polynomial 0x82F63B78) over the 8 little-endian bytes of
// v, no inversion (hash.rs:15-30), one table lookup per byte
__device__ __forceinline__ void crc_table(uint32_t *t) {
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) {
        uint32_t c = i;
        for (int k = 0; k < 8; k++) c = (c >> 1) ^ (0x82F63B78u & (0u - (c & 1u)));
        t[i] = c;
    }
}
__device__ __forceinline__ uint32_t crc32c_u64(const uint32_t *t, uint32_t crc, uint64_t v) {
#pragma unroll
    for (int k = 0; k < 8; k++) crc = (crc >> 8) ^ t[(crc ^ (uint32_t)(v >> (8 * k))) & 0xFFu];
    return crc;
}

// The same CRC32C, eight bytes at a time (slicing by 8): t[k][i] is the CRC
// of byte i followed by k zero bytes, so the eight lookups are independent
// (the decode recurrence waits on one step's CRC before its next table read).
__device__ __forceinline__ void crc_tables8(uint32_t (*t)[256]) {
    crc_table(t[0]);
    __syncthreads();
    for (int k = 1; k < 8; k++) {
        for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) {
            const uint32_t c = t[k - 1][i];
            t[k][i] = (c >> 8) ^ t[0][c & 0xFFu];
        }
        __syncthreads();
    }
}
__device__ __forceinline__ uint32_t crc32c_u64_s8(const uint32_t (*t)[256], uint32_t crc, uint64_t v) {
    const uint32_t lo = (uint32_t)v ^ crc, hi = (uint32_t)(v >> 32);
    return (t[7][lo & 0xFFu] ^ t[6][(lo >> 8) & 0xFFu]) ^ (t[5][(lo >> 16) & 0xFFu] ^ t[4][lo >> 24]) ^
           (t[3][hi & 0xFFu] ^ t[2][(hi >> 8) & 0xFFu]) ^ (t[1][(hi >> 16) & 0xFFu] ^ t[0][hi >> 24]);
}

// block index of a flat position (tile t, slot k) of the encode tiles
__device__ __forceinline__ uint64_t pred_local(const LinD &db, uint64_t f) {
    return (f / LE_TILE - db.tile0) * LE_TILE + f % LE_TILE;
}

template <uint32_t PK>
__global__ void __launch_bounds__(LE_WG) k_pred_keys(const LinD *d, uint32_t nblk, uint32_t mask,
                                                     uint32_t *keys, uint32_t *idx) {
    __shared__ uint32_t ct[256];
    crc_table(ct);
    __syncthreads();
    const uint64_t t = blockIdx.x;
    const uint32_t b = owner_block(d, nblk, t);
    const uint64_t *vals = d[b].vals;
    const uint64_t n = d[b].n, s0 = (t - d[b].tile0) * LE_TILE;
    const uint32_t cnt = (uint32_t)min((uint64_t)LE_TILE, n - s0);
    for (uint32_t k = threadIdx.x; k < LE_TILE; k += LE_WG) {
        const uint64_t f = t * LE_TILE + k;
        uint32_t key = PRED_NONE;
        if (k < cnt) {
            const uint64_t i = s0 + k;
            uint32_t ctx = 0;
            if (i > 0) {
                uint64_t h = vals[i - 1];
                if (PK == PRED_DFCM) h -= i > 1 ? vals[i - 2] : 0;
                ctx = crc32c_u64(ct, PRED_SEED, h) & mask;
            }
            key = (b << 16) | ctx;
        }
        keys[f] = key;
        idx[f] = (uint32_t)f;
    }
}

// each sorted position: the residual of its value from the previous one with
// the same key
template <uint32_t PK>
__global__ void __launch_bounds__(256) k_pred_res(const LinD *d, const uint32_t *keys,
                                                  const uint32_t *idx, uint64_t total,
                                                  uint64_t *res) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= total) return;
    const uint32_t key = keys[p];
    if (key == PRED_NONE) return;
    const LinD &db = d[key >> 16];
    const uint64_t f = idx[p], i = pred_local(db, f);
    const uint64_t *v = db.vals;
    const bool has = p > 0 && keys[p - 1] == key;
    const uint64_t j = has ? pred_local(db, idx[p - 1]) : 0;
    uint64_t pred;
    if (PK == PRED_FCM) pred = has ? v[j] : 0;
    else pred = (i > 0 ? v[i - 1] : 0) + (has ? v[j] - (j > 0 ? v[j - 1] : 0) : 0);
    res[f] = v[i] ^ pred;
}

template <bool WRITE>
__global__ void __launch_bounds__(LE_WG) k_pred_enc(const LinD *d, LinR *r, uint32_t nblk,
                                                    uint64_t *tsz, const uint64_t *resg) {
    const uint64_t t = blockIdx.x;
    const uint32_t b = owner_block(d, nblk, t);
    const uint64_t n = d[b].n, s0 = (t - d[b].tile0) * LE_TILE;
    const uint32_t cnt = (uint32_t)min((uint64_t)LE_TILE, n - s0);
    const uint32_t j0 = threadIdx.x * LE_PER;
    uint64_t res[LE_PER];
    uint64_t len = 0;
#pragma unroll
    for (int m = 0; m < LE_PER; m++) {
        res[m] = j0 + m < cnt ? resg[t * LE_TILE + j0 + m] : 0;
        if (j0 + m < cnt) len += qa_varint_len(res[m]);
    }
    tile_varints<WRITE>(d[b], r[b], t, cnt, j0, res, len, tsz, 0);
}

__global__ void k_pred_zero(uint64_t *p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// pred.rs decode / dfcm_decode over a block's residual words (already parsed
// into out), one lane per block, table at tables + (b << log2)
template <uint32_t PK>
__global__ void __launch_bounds__(64) k_pred_seq(const LinD *d, const LinR *r, uint32_t nblk,
                                                 uint32_t log2, uint64_t *tables) {
    __shared__ uint32_t ct[8][256];
    crc_tables8(ct);
    const uint32_t b = blockIdx.x * 64 + threadIdx.x;
    if (b >= nblk || r[b].status) return;
    const uint32_t mask = (1u << log2) - 1;
    uint64_t *tab = tables + ((uint64_t)b << log2);
    uint64_t *out = d[b].out;
    const uint64_t n = d[b].n;
    uint32_t ctx = 0;
    uint64_t last = 0;
    auto step = [&](uint64_t x) -> uint64_t {
        if (PK == PRED_FCM) {
            const uint64_t v = x ^ tab[ctx];
            tab[ctx] = v;
            ctx = crc32c_u64_s8(ct, PRED_SEED, v) & mask;
            return v;
        } else {
            const uint64_t v = x ^ (last + tab[ctx]);
            const uint64_t delta = v - last;
            tab[ctx] = delta;
            ctx = crc32c_u64_s8(ct, PRED_SEED, delta) & mask;
            last = v;
            return v;
        }
    };
    // residual words are loaded AHEAD values before they are used
    constexpr int AHEAD = 32;
    uint64_t x[AHEAD];
#pragma unroll
    for (int k = 0; k < AHEAD; k++) x[k] = (uint64_t)k < n ? out[k] : 0;
    uint64_t i = 0;
    for (; i + AHEAD <= n; i += AHEAD) {
#pragma unroll
        for (int k = 0; k < AHEAD; k++) {
            out[i + k] = step(x[k]);
            x[k] = i + AHEAD + k < n ? out[i + AHEAD + k] : 0;
        }
    }
    for (int k = 0; k < AHEAD; k++)
        if (i + k < n) out[i + k] = step(x[k]);
}

template <uint32_t PK>
void pred_enc_launch(const LinD *d, LinR *r, uint32_t nblk, uint64_t ntiles, uint32_t log2,
                     uint32_t *keys, uint32_t *idx, uint32_t *keys2, uint32_t *idx2, uint64_t *res,
                     void *sort_tmp, size_t sort_bytes, uint64_t *tsz, uint64_t *tpre,
                     uint64_t *tmp, hipStream_t s) {
    if (!ntiles) return;
    const uint64_t total = ntiles * LE_TILE;
    qa_launch(k_pred_keys<PK>, dim3((unsigned)ntiles), dim3(LE_WG), 0, s, d, nblk,
                       (1u << log2) - 1, keys, idx);
    size_t bytes = sort_bytes;
    (void)rocprim::radix_sort_pairs(sort_tmp, bytes, keys, keys2, idx, idx2, (size_t)total, 0, 32, s);
    qa_launch(k_pred_res<PK>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, d,
                       keys2, idx2, total, res);
    qa_launch((k_pred_enc<false>), dim3((unsigned)ntiles), dim3(LE_WG), 0, s, d, r, nblk, tsz,
                       res);
    qa_exclusive_scan(tsz, ntiles, tpre, tmp, s);
    qa_launch((k_pred_enc<true>), dim3((unsigned)ntiles), dim3(LE_WG), 0, s, d, r, nblk, tpre,
                       res);
}

}  // namespace

size_t qa_pred_sort_bytes(uint64_t total) {
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, bytes, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                    (uint32_t *)nullptr, (uint32_t *)nullptr, (size_t)total, 0, 32);
    return bytes;
}

void qa_pred_encode_launch(uint32_t kind, const LinD *d, LinR *r, uint32_t nblk, uint64_t ntiles,
                           uint32_t log2, uint32_t *keys, uint32_t *idx, uint32_t *keys2,
                           uint32_t *idx2, uint64_t *res, void *sort_tmp, size_t sort_bytes,
                           uint64_t *tsz, uint64_t *tpre, uint64_t *tmp, hipStream_t s) {
    if (kind == PRED_FCM)
        pred_enc_launch<PRED_FCM>(d, r, nblk, ntiles, log2, keys, idx, keys2, idx2, res, sort_tmp,
                                  sort_bytes, tsz, tpre, tmp, s);
    else
        pred_enc_launch<PRED_DFCM>(d, r, nblk, ntiles, log2, keys, idx, keys2, idx2, res, sort_tmp,
                                   sort_bytes, tsz, tpre, tmp, s);
}

void qa_pred_decode_launch(uint32_t kind, const LinD *d, LinR *r, uint32_t nblk, uint64_t ntiles,
                           uint32_t log2, uint64_t *tables, uint64_t *cnt, uint64_t *pre,
                           uint64_t *tmp, hipStream_t s, void *kev0, void *kev1) {
    // the varint parse and its checks are OrderedDelta's (no header byte)
    if (ntiles)
        qa_launch(k_lin_vcount<LIN_OD>, dim3((unsigned)ntiles), dim3(LD_WG), 0, s, d, nblk,
                           cnt);
    qa_exclusive_scan(cnt, ntiles, pre, tmp, s);
    if (ntiles)
        qa_launch(k_lin_vparse<LIN_OD>, dim3((unsigned)ntiles), dim3(LD_WG), 0, s, d, r,
                           nblk, pre);
    const unsigned g64 = (nblk + 63) / 64;
    qa_launch(k_lin_dcheck<LIN_OD>, dim3(g64), dim3(64), 0, s, d, r, nblk, pre);
    const uint64_t words = (uint64_t)nblk << log2;
    const uint64_t zg = (words + 255) / 256;
    qa_launch(k_pred_zero, dim3((unsigned)(zg < 65536 ? zg : 65536)), dim3(256), 0, s,
                       tables, words);
    if (kev0) (void)hipEventRecord((hipEvent_t)kev0, s);
    if (kind == PRED_FCM)
        qa_launch(k_pred_seq<PRED_FCM>, dim3(g64), dim3(64), 0, s, d, r, nblk, log2, tables);
    else
        qa_launch(k_pred_seq<PRED_DFCM>, dim3(g64), dim3(64), 0, s, d, r, nblk, log2, tables);
    if (kev1) (void)hipEventRecord((hipEvent_t)kev1, s);
}

uint64_t qa_lin_enc_tiles(uint64_t n) { return (n + LE_TILE - 1) / LE_TILE; }
uint64_t qa_lin_dec_tiles(uint64_t bytes) { return (bytes + LD_TILE - 1) / LD_TILE; }

void qa_lin_encode_launch(uint32_t kind, const LinD *d, LinR *r, uint32_t nblk, uint64_t ntiles,
                          uint64_t *tsz, uint64_t *tpre, uint64_t *tmp, hipStream_t s) {
    if (kind == LIN_OD) enc_launch<LIN_OD>(d, r, nblk, ntiles, tsz, tpre, tmp, s);
    else if (kind == LIN_D2) enc_launch<LIN_D2>(d, r, nblk, ntiles, tsz, tpre, tmp, s);
    else enc_launch<LIN_DP>(d, r, nblk, ntiles, tsz, tpre, tmp, s);
}

void qa_lin_decode_launch(uint32_t kind, const LinD *d, LinR *r, uint32_t nblk, uint64_t ntiles,
                          uint64_t *cnt, uint64_t *pre, uint64_t *tmp, hipStream_t s, void *kev0,
                          void *kev1) {
    if (kind == LIN_OD) dec_launch<LIN_OD>(d, r, nblk, ntiles, cnt, pre, tmp, s, kev0, kev1);
    else if (kind == LIN_D2) dec_launch<LIN_D2>(d, r, nblk, ntiles, cnt, pre, tmp, s, kev0, kev1);
    else dec_launch<LIN_DP>(d, r, nblk, ntiles, cnt, pre, tmp, s, kev0, kev1);
}
