// qa_fast.hip — candidate sizing for the Level::Fast pool and the block
// decisions (src/encoder.rs:275-469 with entropy off).
//
//   k_float_tiles   FloatMult per-scale checks up to a tile's first passing
//                   scale (src/codecs/float_mult.rs:45-64) and ALP's (e, f)
//                   search + digit range per sub-block
//                   (src/codecs/alp.rs:49-166); one wavefront per tile, FP64
//                   with no contraction
//   k_fm_resolve    FloatMult: the block's first scale every tile passes,
//                   evaluating tiles only at the scales the block needs
//   k_alprd_cut     ALP-RD cut search on the block sample (alp_rd.rs:80-107):
//                   sorted once in LDS, every cut's histogram is a run scan
//   k_alprd_hist    full-block left-part histogram and its top 8
//                   (alp_rd.rs:48-69): LDS hash, exact global-hash fallback
//   k_alprd_tiles   codes / rights ranges and exceptions per tile
//   k_block_decide  sizes RAW..ALP-RD and the first part of the ordered replay
//   k_dict_set      Dict cardinality (dict.rs:87-107): exact LDS hash set, a
//                   bucket-count lower bound that certifies the bail, exact
//                   global-hash fallback; sorted dictionary + values blob size
//   k_dict_tiles    code (rank) range per tile
//   k_rle_starts / k_rle_chunks   run starts and length chunks (rle.rs:10-43)
//   k_block_final   Dict / RLE consideration and the winner's layout
#include "qa_dev.h"

__constant__ double c_exp10[19] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8, 1e9,
                                   1e10, 1e11, 1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18};
__constant__ double c_frac10[19] = {1e0,   1e-1,  1e-2,  1e-3,  1e-4,  1e-5,  1e-6,
                                    1e-7,  1e-8,  1e-9,  1e-10, 1e-11, 1e-12, 1e-13,
                                    1e-14, 1e-15, 1e-16, 1e-17, 1e-18};
__constant__ double c_fm_scales[7] = {10.0, 100.0, 1000.0, 10000.0, 100000.0, 1000000.0,
                                      10000000.0};
#define ALP_MAGIC 6755399441055744.0
#define ALP_UPPER 9.223372036854776e18

__device__ __forceinline__ uint64_t f2b(double d) { return (uint64_t)__double_as_longlong(d); }
__device__ __forceinline__ double b2f(uint64_t b) { return __longlong_as_double((long long)b); }

// Rust `f64 as i64` for the finite values reaching it: truncate, saturate
__device__ __forceinline__ int64_t sat_i64(double x) {
    if (x >= 9223372036854775808.0) return INT64_MAX;
    if (x < -9223372036854775808.0) return INT64_MIN;
    return (int64_t)x;
}
// ALP's digit of a value (alp.rs:33-45) and its round-trip check, without an
// integer round trip: the rounded digit r is an integer-valued double with
// |r| <= 2^63, so (double)(i64)r == r (at 2^63 the saturated i64 converts back
// to 2^63) and the verify product is the same; the digit is returned as that
// double (order-preserving, converted by sat_i64 where an integer is needed).
// m1 = 10^e, m2 = 10^-f, m3 = 10^f, m4 = 10^-e.
__device__ __forceinline__ bool alp_digit_r(uint64_t bits, double m1, double m2, double m3, double m4,
                                            double *r) {
    const double v = b2f(bits);
    const double tmp = __dmul_rn(__dmul_rn(v, m1), m2);
    if (!isfinite(tmp) || tmp >= ALP_UPPER || tmp <= -ALP_UPPER) return false;
    const double rr = __dsub_rn(__dadd_rn(tmp, ALP_MAGIC), ALP_MAGIC);
    if (f2b(__dmul_rn(__dmul_rn(rr, m3), m4)) != bits) return false;
    *r = rr;
    return true;
}
__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0; }

// FloatMult's k = round(x * s) and its round-trip check (float_mult.rs:50-58),
// without an integer round trip: round(x * s) is an integer-valued
// double of magnitude < 9e18 < 2^63, so converting it to i64 and back is exact
__device__ __forceinline__ bool fm_k_r(uint64_t bits, double sc, double *k) {
    const double kk = round(__dmul_rn(b2f(bits), sc));
    if (!isfinite(kk) || fabs(kk) >= 9.0e18) return false;
    if (f2b(__ddiv_rn(kk, sc)) != bits) return false;
    *k = kk;
    return true;
}

// FloatMult at scale s over a tile held in registers (k = lane + 64 q): every
// value round-trips (float_mult.rs:45-64), with a wave-wide early exit on the
// first failure; on success the k range goes to the tile record
__device__ __forceinline__ bool fm_tile(const uint64_t (&xv)[QA_TILE / WAVE], uint32_t cnt, int s,
                                        int lane, DTileF *tf) {
    const double sc = c_fm_scales[s];
    double fmn = 1e300, fmx = -1e300;  // k as exact doubles: min / max commute with the conversion
#pragma unroll
    for (int q = 0; q < QA_TILE / WAVE; q++) {
        if ((uint32_t)(WAVE * q) >= cnt) break;
        const uint32_t k = lane + WAVE * q;
        bool bad = false;
        if (k < cnt) {
            double kk;
            if (fm_k_r(xv[q], sc, &kk)) {
                fmn = kk < fmn ? kk : fmn;
                fmx = kk > fmx ? kk : fmx;
            } else {
                bad = true;
            }
        }
        if (wave_any(bad)) return false;
    }
    int64_t kmn = fmn == 1e300 ? INT64_MAX : (int64_t)fmn, kmx = fmx == -1e300 ? INT64_MIN : (int64_t)fmx;
    kmn = wave_reduce(kmn, OpMinS());
    kmx = wave_reduce(kmx, OpMaxS());
    if (lane == 0) {
        tf->kmin[s] = kmn;
        tf->kmax[s] = kmx;
    }
    return true;
}

// locate a tile slot: column, owner block slot, tile index inside the block
__device__ __forceinline__ bool tile_owner(const EncArgs &a, uint64_t t, int *cp, uint32_t *op,
                                           uint32_t *btp) {
    int c = col_by_tile(a.cols, a.ncols, t);
    const DCol &col = a.cols[c];
    uint64_t lt = t - col.tile_base;
    uint32_t lr = (uint32_t)(lt / col.tpr);
    if (lr >= col.nregions) return false;
    uint32_t o = a.owner[col.region_base + lr];
    if (o == NONE32) return false;
    *cp = c;
    *op = o;
    *btp = (col.region_base + lr - o) * col.tpr + (uint32_t)(lt % col.tpr);
    return true;
}

// ---------------------------------------------------------------------------
// FloatMult + ALP per tile
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_float_tiles(EncArgs a) {
    __shared__ uint64_t samp[4][32];
    const int lane = threadIdx.x & 63, wv = qa_wave_id();
    const uint64_t t = a.t0 + (uint64_t)blockIdx.x * 4 + wv;
    if (t >= a.t1) return;
    const DTile &ts = a.tiles[t];
    if (!ts.active) return;
    int c;
    uint32_t o, bt;
    if (!tile_owner(a, t, &c, &o, &bt)) return;
    const DCol &col = a.cols[c];
    if (!is_float(col.dtype) || col.level == QA_FASTEST) return;
    const DColState &st = a.cst[c];
    const DBlock &b = a.blocks[o];
    const uint64_t k0 = b.start + (uint64_t)bt * QA_TILE;
    const uint32_t cnt = ts.cnt;
    DTileF &tf = a.tilesf[t];
    // the tile's values in registers (k = lane + 64 q), all loads in flight at once
    uint64_t xv[QA_TILE / WAVE];
#pragma unroll
    for (int q = 0; q < QA_TILE / WAVE; q++) {
        const uint32_t k = lane + WAVE * q;
        xv[q] = k < cnt ? lane_at(st, k0 + k) : 0;
    }
    // FloatMult: scales in order up to the tile's first passing one; the block
    // needs the first scale every tile passes, and k_fm_resolve evaluates the
    // scales a tile did not reach only where the block needs them
    uint8_t pass = 0, fail = 0;
    for (int s = 0; s < 7; s++) {
        if (fm_tile(xv, cnt, s, lane, &tf)) {
            pass = (uint8_t)(1u << s);
            break;
        }
        fail |= (uint8_t)(1u << s);
    }
    // ALP: (e, f) on a 32-value sample, 190 pairs spread over the lanes
    const uint32_t stride = cnt / 32 > 1 ? cnt / 32 : 1;
    uint32_t sn = 0;
    for (uint32_t i = 0; i < cnt && sn < 32; i += stride) sn++;
    if (lane < (int)sn) samp[wv][lane] = lane_at(st, k0 + (uint64_t)lane * stride);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    uint32_t best_key = 0xFFFFFFFFu;
    for (uint32_t combo = lane; combo < 190; combo += WAVE) {
        uint32_t e = 0;
        while ((e + 1) * (e + 2) / 2 <= combo) e++;
        uint32_t f = combo - e * (e + 1) / 2;
        const double m1 = c_exp10[e], m2 = c_frac10[f], m3 = c_exp10[f], m4 = c_frac10[e];
        uint32_t exc = 0;
        double flo = 1e300, fhi = -1e300;
        for (uint32_t j = 0; j < sn; j++) {
            double r;
            if (alp_digit_r(samp[wv][j], m1, m2, m3, m4, &r)) {
                flo = r < flo ? r : flo;
                fhi = r > fhi ? r : fhi;
            } else {
                exc++;
            }
        }
        const int64_t lo = exc == sn ? INT64_MAX : sat_i64(flo), hi = exc == sn ? INT64_MIN : sat_i64(fhi);
        uint32_t width;
        if (exc == sn || hi <= lo)
            width = exc == sn ? 64 : 0;
        else
            width = qa_width((uint64_t)hi - (uint64_t)lo);
        uint32_t cost = width * sn + exc * 80;
        uint32_t key = (cost << 16) | (combo << 8) | exc;  // first minimum in (e, f) order
        best_key = key < best_key ? key : best_key;
    }
    best_key = (uint32_t)wave_reduce((uint64_t)best_key, OpMinU());
    const uint32_t bcombo = (best_key >> 8) & 0xFF, bexc = best_key & 0xFF;
    uint32_t be = 0;
    while ((be + 1) * (be + 2) / 2 <= bcombo) be++;
    const uint32_t bf = bcombo - be * (be + 1) / 2;
    bool bail = bexc * 2 > sn;
    uint32_t nexc = 0, w = 0;
    int64_t dmin = 0;
    if (!bail) {
        const double m1 = c_exp10[be], m2 = c_frac10[bf], m3 = c_exp10[bf], m4 = c_frac10[be];
        double flo = 1e300, fhi = -1e300;
        uint32_t first_valid = 0xFFFFFFFFu;
        int64_t fv_digit = 0;
#pragma unroll
        for (int q = 0; q < QA_TILE / WAVE; q++) {
            const uint32_t kb = WAVE * q;
            if (kb >= cnt) break;
            const uint32_t k = kb + lane;
            bool valid = false;
            double r = 0;
            if (k < cnt) valid = alp_digit_r(xv[q], m1, m2, m3, m4, &r);
            if (k < cnt && !valid) nexc++;
            if (valid) {
                flo = r < flo ? r : flo;
                fhi = r > fhi ? r : fhi;
            }
            uint64_t bal = __ballot(valid);
            if (first_valid == 0xFFFFFFFFu && bal) {
                int src = __ffsll((long long)bal) - 1;
                first_valid = kb + src;
                fv_digit = sat_i64(b2f(__shfl(f2b(r), src, 64)));
            }
        }
        int64_t lo = flo == 1e300 ? INT64_MAX : sat_i64(flo), hi = fhi == -1e300 ? INT64_MIN : sat_i64(fhi);
        nexc = (uint32_t)wave_reduce((uint64_t)nexc, OpAdd());
        lo = wave_reduce(lo, OpMinS());
        hi = wave_reduce(hi, OpMaxS());
        int64_t filler = first_valid == 0xFFFFFFFFu ? 0 : fv_digit;
        if (nexc > 0 || cnt < QA_TILE) {  // exceptions and padding take the filler
            lo = filler < lo ? filler : lo;
            hi = filler > hi ? filler : hi;
        }
        dmin = lo;
        w = qa_width((uint64_t)hi - (uint64_t)lo);
        if (w > 32) bail = true;
    }
    if (lane == 0) {
        tf.fm_pass = pass;
        tf.fm_fail = fail;
        tf.alp_e = (uint8_t)be;
        tf.alp_f = (uint8_t)bf;
        tf.alp_w = (uint8_t)w;
        tf.alp_exc = nexc;
        tf.dmin = dmin;
        tf.alp_bail = bail;
    }
}

// ---------------------------------------------------------------------------
// FloatMult: the block's scale (float_mult.rs:24-43, the first scale at which
// every value round-trips). k_float_tiles left each tile's first passing scale
// and the scales below it as failed; a scale no tile failed is the candidate,
// and only the tiles that did not reach it are evaluated there. When every
// tile passes the candidate, the AND of the tiles' pass masks has it as its
// lowest bit (each lower scale failed on some tile); when every scale failed
// on some tile, the AND is empty.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool fast_float_block(const EncArgs &a, const DBlock &b);

__global__ void __launch_bounds__(256) k_fm_resolve(EncArgs a) {
    __shared__ uint32_t sh_fail, sh_open;
    const uint32_t slot = a.r0 + blockIdx.x;
    if (slot >= a.r1) return;
    const DBlock &b = a.blocks[slot];
    if (!fast_float_block(a, b)) return;
    const DCol &col = a.cols[b.col];
    const DColState &st = a.cst[b.col];
    const uint64_t tb = col.tile_base + (uint64_t)(slot - col.region_base) * col.tpr;
    const DTile *tl = a.tiles + tb;
    DTileF *tf = a.tilesf + tb;
    const int lane = threadIdx.x & 63, wv = qa_wave_id();
    for (int round = 0; round < 8; round++) {
        if (threadIdx.x == 0) {
            sh_fail = 0;
            sh_open = 0;
        }
        __syncthreads();
        uint32_t f = 0;
        for (uint32_t t = threadIdx.x; t < b.ntiles; t += blockDim.x) f |= tf[t].fm_fail;
        if (f) atomicOr(&sh_fail, f);
        __syncthreads();
        const uint32_t fail = sh_fail & 0x7F;
        if (fail == 0x7F) return;
        const int s = __ffs((int)(~fail & 0x7F)) - 1;
        for (uint32_t t = (uint32_t)wv; t < b.ntiles; t += blockDim.x / WAVE) {
            DTileF *x = tf + t;
            if (((x->fm_pass | x->fm_fail) >> s) & 1) continue;
            const uint32_t cnt = tl[t].cnt;
            const uint64_t k0 = b.start + (uint64_t)t * QA_TILE;
            uint64_t xv[QA_TILE / WAVE];
#pragma unroll
            for (int q = 0; q < QA_TILE / WAVE; q++) {
                const uint32_t k = lane + WAVE * q;
                xv[q] = k < cnt ? lane_at(st, k0 + k) : 0;
            }
            const bool ok = fm_tile(xv, cnt, s, lane, x);
            if (lane == 0) {
                if (ok) x->fm_pass |= (uint8_t)(1u << s);
                else x->fm_fail |= (uint8_t)(1u << s);
                sh_open = 1;
            }
        }
        __syncthreads();
        const uint32_t open = sh_open;
        __syncthreads();  // every thread has read the flags before they are reset
        if (!open) return;  // every tile passes s
    }
}

// ---------------------------------------------------------------------------
// ALP-RD: cut search on the sorted sample
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t rd_code_bits(uint32_t dl) {  // alp_rd.rs:72-78
    if (dl <= 1) return 1;
    uint32_t b = 32 - __clz((int)(dl - 1));
    return b > 1 ? b : 1;
}

// Bitonic sort of n_pow2 (a power of two >= 2) keys in LDS, ascending, by the
// whole workgroup (its size a multiple of 64). Thread t takes compare pairs
// t, t + blockDim, ...: pair p joins elements lo and lo + j. For j <= 64 the
// 64 pairs of a wavefront lie in 128 consecutive elements, so those stages
// need only the wavefront's own LDS order; wider ones take barriers.
template <typename T>
__device__ void lds_bitonic_sort(T *s, uint32_t n_pow2) {
    for (uint32_t k = 2; k <= n_pow2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            const bool wide = j >= 128;
            if (wide) __syncthreads();
            for (uint32_t p = threadIdx.x; p < n_pow2 / 2; p += blockDim.x) {
                const uint32_t lo = ((p & ~(j - 1)) << 1) | (p & (j - 1)), hi = lo + j;
                const bool up = (lo & k) == 0;
                const T x = s[lo], y = s[hi];
                if ((x > y) == up) {
                    s[lo] = y;
                    s[hi] = x;
                }
            }
            if (wide) __syncthreads();
            else wave_lds_sync();
        }
    }
    __syncthreads();
}

__device__ __forceinline__ bool fast_float_block(const EncArgs &a, const DBlock &b) {
    const DCol &col = a.cols[b.col];
    return b.len && is_float(col.dtype) && col.level != QA_FASTEST;
}

__global__ void __launch_bounds__(1024) k_alprd_cut(EncArgs a) {
    // the samples' top 32 bits: sorted, they order every cut's left part
    // (at most 24 bits)
    __shared__ uint32_t s[4096];
    __shared__ double bits_of[24];
    const uint32_t slot = a.r0 + blockIdx.x;
    if (slot >= a.r1) return;
    const DBlock &b = a.blocks[slot];
    if (!fast_float_block(a, b)) return;
    DBlockX &x = a.blocksx[slot];
    const DColState &st = a.cst[b.col];
    const uint64_t n = b.len;
    const uint64_t step = n / 2048 > 1 ? n / 2048 : 1;
    const uint32_t sn = (uint32_t)((n + step - 1) / step);
    uint32_t np2 = 2;
    while (np2 < sn) np2 <<= 1;
    for (uint32_t i = threadIdx.x; i < np2; i += blockDim.x)
        s[i] = i < sn ? (uint32_t)(lane_at(st, b.start + i * step) >> 32) : ~0u;
    __syncthreads();
    lds_bitonic_sort(s, np2);
    // Every cut's left parts are sorted too, so its histogram is the runs of
    // equal keys: wavefront w takes cuts w + 1 and w + 17. Run starts come
    // from a ballot per 64 samples; the cost needs only the distinct count and
    // the sum of the 8 largest counts (ties do not change the sum). Each lane
    // keeps its own top 8 (descending, in registers) of the runs it closes;
    // the wavefront's top 8 are then taken by 8 rounds of a wavefront maximum.
    const int lane = threadIdx.x & 63, wv = qa_wave_id();
    auto keep = [](uint32_t (&top)[8], uint32_t v) {
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const uint32_t hi = v > top[q] ? v : top[q];
            v = v > top[q] ? top[q] : v;
            top[q] = hi;
        }
    };
    for (uint32_t left = (uint32_t)wv + 1; left <= 24; left += 16) {
        const uint32_t rbw = 64 - left, ksh = rbw - 32;
        uint32_t top[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        uint32_t distinct = 0, open = 0;  // start of the run still open
        for (uint32_t c = 0; c < sn; c += WAVE) {
            const uint32_t i = c + (uint32_t)lane;
            const bool st = i < sn && (i == 0 || (s[i] >> ksh) != (s[i - 1] >> ksh));
            const uint64_t m = __ballot(st);
            distinct += (uint32_t)__builtin_popcountll(m);
            // a run starting here and ending in this chunk; lane 63 (whose run,
            // if any, stays open) takes the open run that this chunk's first
            // start closes
            uint32_t len = 0;
            if (st && lane < 63) {
                const uint64_t later = m & (~0ull << (lane + 1));
                if (later) len = (uint32_t)__builtin_ctzll(later) - (uint32_t)lane;
            }
            if (lane == 63 && m && c > 0) len = c + (uint32_t)__builtin_ctzll(m) - open;
            if (m) open = c + 63 - (uint32_t)__builtin_clzll(m);
            keep(top, len);
        }
        keep(top, lane == 63 ? sn - open : 0u);  // the last run ends at the sample's end
        uint64_t covered = 0;
        for (int r = 0; r < 8; r++) {
            const uint32_t mx = (uint32_t)wave_reduce((uint64_t)top[0], OpMaxU());
            covered += mx;
            const uint64_t who = __ballot(top[0] == mx);
            if (lane == __builtin_ctzll(who)) {
#pragma unroll
                for (int q = 0; q < 7; q++) top[q] = top[q + 1];
                top[7] = 0;
            }
        }
        const uint32_t dl = distinct < 8 ? distinct : 8;
        const uint64_t exc = sn - covered;
        const uint64_t num = (uint64_t)sn * (rbw + rd_code_bits(dl)) + exc * 96;
        if (lane == 0) bits_of[left - 1] = __ddiv_rn((double)num, (double)sn);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t best = 0;
        double bb = __builtin_inf();
        for (uint32_t c = 0; c < 24; c++)
            if (bits_of[c] < bb) {
                bb = bits_of[c];
                best = c + 1;
            }
        x.rd_bail = (best == 0 || bb >= 63.0) ? 1 : 0;
        x.rd_left = best;
    }
}

// ---------------------------------------------------------------------------
// ALP-RD: full-block histogram of the left parts and its top 8
// ---------------------------------------------------------------------------
#define RDH_CAP 4096
#define RDH_LIMIT 3072
#define HEMPTY (~0ull)

__device__ __forceinline__ uint64_t mix64(uint64_t x) {  // SplitMix64 finalizer
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

// top-8 by (count desc, key asc) over a (keys, counts) table; keys < 2^24,
// empty slots hold all-ones
template <typename K>
__device__ void top8_select(K *keys, uint32_t *cnt, uint32_t cap, uint64_t *out, uint32_t *outn) {
    __shared__ unsigned long long best;
    uint32_t got = 0;
    for (int r = 0; r < 8; r++) {
        if (threadIdx.x == 0) best = 0;
        __syncthreads();
        unsigned long long my = 0;
        for (uint32_t i = threadIdx.x; i < cap; i += blockDim.x) {
            if (keys[i] == (K)HEMPTY || cnt[i] == 0) continue;
            unsigned long long p = ((unsigned long long)cnt[i] << 24) | (0xFFFFFFull - keys[i]);
            my = p > my ? p : my;
        }
        my = wave_reduce((uint64_t)my, OpMaxU());
        if ((threadIdx.x & 63) == 0) atomicMax(&best, my);
        __syncthreads();
        unsigned long long w = best;
        if (w == 0) break;
        uint64_t key = 0xFFFFFFull - (w & 0xFFFFFFull);
        for (uint32_t i = threadIdx.x; i < cap; i += blockDim.x)
            if (keys[i] == key) cnt[i] = 0;
        if (threadIdx.x == 0) out[got] = key;
        got++;
        __syncthreads();
    }
    *outn = got;
}

__global__ void __launch_bounds__(1024) k_alprd_hist(EncArgs a) {
    __shared__ uint64_t keys[RDH_CAP];
    __shared__ uint32_t cnt[RDH_CAP];
    __shared__ uint32_t ndist, overflow;
    const uint32_t slot = a.r0 + blockIdx.x;
    if (slot >= a.r1) return;
    const DBlock &b = a.blocks[slot];
    if (!fast_float_block(a, b)) return;
    DBlockX &x = a.blocksx[slot];
    if (x.rd_bail) return;
    const DCol &col = a.cols[b.col];
    const DColState &st = a.cst[b.col];
    const uint64_t n = b.len;
    const uint32_t rbw = 64 - x.rd_left;
    for (uint32_t i = threadIdx.x; i < RDH_CAP; i += blockDim.x) {
        keys[i] = HEMPTY;
        cnt[i] = 0;
    }
    if (threadIdx.x == 0) {
        ndist = 0;
        overflow = 0;
    }
    __syncthreads();
    for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) {
        if (overflow) break;
        uint64_t key = lane_at(st, b.start + i) >> rbw;
        uint32_t h = (uint32_t)(mix64(key) >> 52);
        for (uint32_t probe = 0; probe < RDH_CAP; probe++) {
            uint64_t cur = keys[h];
            if (cur == key) {
                atomicAdd(&cnt[h], 1u);
                break;
            }
            if (cur == HEMPTY) {
                if (atomicAdd(&ndist, 1u) >= RDH_LIMIT) {
                    overflow = 1;
                    break;
                }
                unsigned long long old =
                    atomicCAS((unsigned long long *)&keys[h], HEMPTY, (unsigned long long)key);
                if (old == HEMPTY || old == key) {
                    atomicAdd(&cnt[h], 1u);
                    if (old == key) atomicSub(&ndist, 1u);
                    break;
                }
                atomicSub(&ndist, 1u);
            }
            h = (h + 1) & (RDH_CAP - 1);
        }
    }
    __syncthreads();
    uint32_t dl = 0;
    if (!overflow) {
        top8_select(keys, cnt, RDH_CAP, x.rd_dict, &dl);
    } else {
        // exact recount in the block's global table (next_pow2(2n) <= 4n slots)
        uint32_t cap = 1;
        while (cap < 2 * n) cap <<= 1;
        uint64_t *gk = a.hkeys + 4 * (col.lane_off + b.start);
        uint32_t *gc = a.hcnt + 4 * (col.lane_off + b.start);
        for (uint32_t i = threadIdx.x; i < cap; i += blockDim.x) {
            gk[i] = HEMPTY;
            gc[i] = 0;
        }
        __threadfence_block();
        __syncthreads();
        for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) {
            uint64_t key = lane_at(st, b.start + i) >> rbw;
            uint32_t h = (uint32_t)(mix64(key) & (cap - 1));
            for (;;) {
                unsigned long long old =
                    atomicCAS((unsigned long long *)&gk[h], HEMPTY, (unsigned long long)key);
                if (old == HEMPTY || old == key) {
                    atomicAdd(&gc[h], 1u);
                    break;
                }
                h = (h + 1) & (cap - 1);
            }
        }
        __threadfence_block();
        __syncthreads();
        top8_select(gk, gc, cap, x.rd_dict, &dl);
    }
    if (threadIdx.x == 0) x.rd_dl = dl;
}

// per tile: codes / rights ranges and the exceptions (alp_rd.rs:145-163)
__global__ void __launch_bounds__(256) k_alprd_tiles(EncArgs a) {
    const int lane = threadIdx.x & 63;
    const uint64_t t = a.t0 + (uint64_t)blockIdx.x * 4 + qa_wave_id();
    if (t >= a.t1) return;
    const DTile &ts = a.tiles[t];
    if (!ts.active) return;
    int c;
    uint32_t o, bt;
    if (!tile_owner(a, t, &c, &o, &bt)) return;
    const DBlock &b = a.blocks[o];
    if (!fast_float_block(a, b)) return;
    const DBlockX &x = a.blocksx[o];
    if (x.rd_bail) return;
    const DColState &st = a.cst[c];
    const uint32_t rbw = 64 - x.rd_left;
    const uint64_t mask = (1ull << rbw) - 1;
    const uint64_t p0 = (uint64_t)bt * QA_TILE;  // block-relative position of the tile
    const uint32_t cnt = ts.cnt;
    // the dictionary in registers (entries are distinct left parts < 2^24, so
    // ~0 in an unused slot never matches), the tile's values loaded together
    uint64_t dict[8];
#pragma unroll
    for (int q = 0; q < 8; q++) dict[q] = (uint32_t)q < x.rd_dl ? x.rd_dict[q] : ~0ull;
    uint64_t vv[QA_TILE / WAVE];
#pragma unroll
    for (int j = 0; j < QA_TILE / WAVE; j++) {
        const uint32_t k = lane + WAVE * j;
        vv[j] = k < cnt ? lane_at(st, b.start + p0 + k) : 0;
    }
    uint64_t cmn = ~0ull, cmx = 0, rmn = ~0ull, rmx = 0, eb = 0, ne = 0;
#pragma unroll
    for (int j = 0; j < QA_TILE / WAVE; j++) {
        const uint32_t k = lane + WAVE * j;
        if (k >= cnt) break;
        const uint64_t v = vv[j];
        const uint64_t left = v >> rbw;
        uint64_t code = 0;
        bool found = false;
#pragma unroll
        for (int q = 7; q >= 0; q--)  // the one match, if any
            if (dict[q] == left) {
                code = q;
                found = true;
            }
        if (!found) {
            ne++;
            eb += qa_varint_len(p0 + k) + 8;
        }
        cmn = code < cmn ? code : cmn;
        cmx = code > cmx ? code : cmx;
        const uint64_t rr = v & mask;
        rmn = rr < rmn ? rr : rmn;
        rmx = rr > rmx ? rr : rmx;
    }
    cmn = wave_reduce(cmn, OpMinU());
    cmx = wave_reduce(cmx, OpMaxU());
    rmn = wave_reduce(rmn, OpMinU());
    rmx = wave_reduce(rmx, OpMaxU());
    eb = wave_reduce(eb, OpAdd());
    ne = wave_reduce(ne, OpAdd());
    if (lane == 0) {
        DTileF &tf = a.tilesf[t];
        tf.rd_cmin = cmn;
        tf.rd_cmax = cmx;
        tf.rd_rmin = rmn;
        tf.rd_rmax = rmx;
        tf.rd_exc = (uint32_t)ne;
        tf.rd_exc_bytes = (uint32_t)eb;
    }
}

// ---------------------------------------------------------------------------
// block decision, part 1: RAW .. ALP-RD in the reference's order
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool consider(uint64_t sz, uint32_t m, uint64_t lambda, uint64_t dec,
                                         uint64_t *best_score, uint64_t *best_len,
                                         uint32_t *mode) {
    if (sz == ~0ull) return false;
    uint64_t sc = sz + penalty_of(m, lambda, dec);
    if (sc < *best_score) {  // strict: earlier candidates win ties (encoder.rs:126-139)
        *best_score = sc;
        *best_len = sz;
        *mode = m;
        return true;
    }
    return false;
}

__global__ void __launch_bounds__(256) k_block_decide(EncArgs a) {
    const int lane = threadIdx.x & 63;
    const uint32_t slot = a.r0 + blockIdx.x * 4 + qa_wave_id();
    if (slot >= a.r1) return;
    DBlock &b = a.blocks[slot];
    if (b.len == 0) return;
    const DCol &col = a.cols[b.col];
    DTile *tl = a.tiles + col.tile_base + (uint64_t)(slot - col.region_base) * col.tpr;
    DTileF *tf = a.tilesf ? a.tilesf + col.tile_base + (uint64_t)(slot - col.region_base) * col.tpr
                          : nullptr;
    const uint64_t n = b.len;
    const bool fl = is_float(col.dtype), sg = is_signed(col.dtype);
    const bool reduced = col.level == QA_FASTEST;
    const uint64_t lambda = reduced ? 16 : 4;
    const uint64_t lb = lane_bytes_of(col.dtype);
    const uint64_t dec = n * lb, raw = n * lb;
    const uint64_t v0 = tl[0].v_first;
    const uint64_t v1 = n > 1 ? v0 + tl[0].step : 0;
    const uint64_t stride = v1 - v0;
    uint32_t cst = 1, str = 1;
    uint64_t xz = 0, fsz = 0, dsz = 0, chg = 0;
    for (uint32_t t = lane; t < b.ntiles; t += WAVE) {
        const DTile &s = tl[t];
        cst &= s.allsame && s.v_first == v0;
        uint64_t k0 = (uint64_t)t * QA_TILE;
        str &= (s.cnt == 1 || (s.arith && s.step == stride)) && s.v_first == v0 + k0 * stride;
        xz += s.xorz_rest + qa_varint_len(s.v_first ^ (t ? tl[t - 1].v_last : 0));
        fsz += 9 + 128ull * qa_width(tile_for_range(s, sg));
        uint64_t dmn, dmx;
        tile_dz(s, t == 0, &dmn, &dmx);
        dsz += 9 + 128ull * qa_width(dmx - dmn);
        chg += s.chg_rest + (t ? (s.v_first != tl[t - 1].v_last) : 0);
    }
    cst = wave_reduce(cst, OpAnd());
    str = wave_reduce(str, OpAnd());
    xz = wave_reduce(xz, OpAdd());
    fsz = wave_reduce(fsz, OpAdd()) + qa_varint_len(n);
    dsz = wave_reduce(dsz, OpAdd()) + 8 + qa_varint_len(n);
    chg = wave_reduce(chg, OpAdd());
    // float-value codecs
    uint64_t fmsz = ~0ull, alpsz = ~0ull, rdsz = ~0ull;
    uint32_t fm_scale = 7;
    if (fl && !reduced && tf) {
        uint32_t pass = 0x7F, abail = 0;
        uint64_t asz = 0, nexc = 0, excb = 0, cl = 0, rl = 0;
        for (uint32_t t = lane; t < b.ntiles; t += WAVE) {
            const DTileF &f = tf[t];
            pass &= f.fm_pass;
            abail |= f.alp_bail;
            asz += 13 + 128ull * f.alp_w + 10ull * f.alp_exc;
            nexc += f.rd_exc;
            excb += f.rd_exc_bytes;
            cl += 9 + 128ull * qa_width(f.rd_cmax - f.rd_cmin);
            rl += 9 + 128ull * qa_width(f.rd_rmax - f.rd_rmin);
        }
        pass = wave_reduce(pass, OpAnd());
        abail = (uint32_t)wave_reduce((uint64_t)abail, OpMaxU());
        asz = wave_reduce(asz, OpAdd());
        nexc = wave_reduce(nexc, OpAdd());
        excb = wave_reduce(excb, OpAdd());
        cl = wave_reduce(cl, OpAdd()) + qa_varint_len(n);
        rl = wave_reduce(rl, OpAdd()) + qa_varint_len(n);
        if (pass) {
            fm_scale = __ffs((int)pass) - 1;
            uint64_t s = 0;
            for (uint32_t t = lane; t < b.ntiles; t += WAVE)
                s += 9 + 128ull * qa_width((uint64_t)tf[t].kmax[fm_scale] - (uint64_t)tf[t].kmin[fm_scale]);
            fmsz = 2 + qa_varint_len(n) + wave_reduce(s, OpAdd());
        }
        if (!abail) alpsz = qa_varint_len(n) + asz;
        const DBlockX &x = a.blocksx[slot];
        if (!x.rd_bail && nexc * 4 <= n)
            rdsz = 2 + 8ull * x.rd_dl + qa_varint_len(nexc) + excb + 1 + qa_varint_len(cl) + cl + 1 +
                   rl;
        if (lane == 0) {
            a.blocksx[slot].fm_scale = fm_scale;
            a.blocksx[slot].rd_nexc = nexc;
            a.blocksx[slot].rd_excb = excb;
            a.blocksx[slot].rd_clen = cl;
            a.blocksx[slot].rd_rlen = rl;
        }
    }
    uint32_t mode = QA_M_RAW;
    uint64_t best_len = raw, best_score = raw;
    uint64_t sz[C_N];
    for (int k = 0; k < C_N; k++) sz[k] = ~0ull;
    sz[C_RAW] = raw;
    // Selection::Sample (hybrid 2) sizes every block whatever its features
    if (a.hybrid == 2 || !(!fl && b.incompressible)) {
        sz[C_CONST] = cst ? 8 : ~0ull;
        sz[C_STRIDE] = (n >= 2 && str) ? 16 : ~0ull;
        sz[C_XORZ] = reduced ? ~0ull : xz;
        sz[C_FOR] = fsz;
        sz[C_DELTA] = dsz;
        sz[C_FM] = fmsz;
        sz[C_ALP] = alpsz;
        sz[C_ALPRD] = rdsz;
        const uint32_t order[9] = {QA_M_RAW, QA_M_CONST, QA_M_STRIDE, QA_M_XORZ, QA_M_FOR,
                                   QA_M_DELTA, QA_M_FLOAT_MULT, QA_M_ALP, QA_M_ALP_RD};
        for (int k = 1; k < 9; k++) consider(sz[k], order[k], lambda, dec, &best_score, &best_len, &mode);
    }
    if (lane == 0) {
        b.v0 = v0;
        b.v1 = v1;
        for (int k = 0; k < C_N; k++) b.sz[k] = sz[k];
        // entropy levels: the host decides the gate after its own candidates
        b.compressible = a.hybrid == 2 ? 1
                         : a.hybrid ? !(!fl && b.incompressible)
                                    : best_len * 20 < (n * 8) * 19 && !(!fl && b.incompressible);
        b.mode = mode;
        b.payload_len = best_len;
        b.score = best_score;
        if (!reduced) {
            DBlockX &x = a.blocksx[slot];
            x.rle_bail = chg > n / 2 + 1;
            x.nruns = chg + 1;
            x.dict_state = 1;  // set by k_dict_set when it runs
        }
    }
    // RLE: runs starting before each tile (a wave-wide scan)
    if (!reduced && b.compressible && chg <= n / 2 + 1) {
        uint64_t carry = 0;
        for (uint32_t base = 0; base < b.ntiles; base += WAVE) {
            uint32_t t = base + lane;
            uint64_t starts = 0;
            if (t < b.ntiles)
                starts = tl[t].chg_rest + (t ? (tl[t].v_first != tl[t - 1].v_last) : 1);
            const uint64_t inc = qa_wave_scan_add64(starts);
            if (t < b.ntiles) tf[t].rle_base = carry + inc - starts;
            carry += __shfl(inc, 63, 64);
        }
    }
}

// ---------------------------------------------------------------------------
// Dict: cardinality, sorted dictionary, values blob size
// ---------------------------------------------------------------------------
#define DH_CAP 8192
#define DH_LIMIT 6144

__device__ __forceinline__ bool dict_candidate(const EncArgs &a, const DBlock &b) {
    return b.len && b.compressible && a.cols[b.col].level != QA_FASTEST && !a.hybrid;
}

// delta_bitpack size of the sorted dictionary (dict.rs:40-61); chunk c's
// width and minimum go to the block's tile slot c
__device__ void dict_values_size(const EncArgs &a, uint32_t slot, const uint64_t *d, uint64_t card,
                                 DTileF *tf, uint64_t *vlen, uint32_t *vtag) {
    __shared__ unsigned long long tot;
    if (threadIdx.x == 0) tot = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = qa_wave_id(), nwv = blockDim.x >> 6;
    uint64_t nch = (card + QA_TILE - 1) / QA_TILE;
    for (uint64_t ch = wv; ch < nch; ch += nwv) {
        uint64_t mn = ~0ull, mx = 0;
        uint64_t i0 = ch * QA_TILE, i1 = min(card, i0 + QA_TILE);
        for (uint64_t i = i0 + lane; i < i1; i += WAVE) {
            uint64_t z = i ? qa_zigzag(d[i] - d[i - 1]) : 0;
            mn = z < mn ? z : mn;
            mx = z > mx ? z : mx;
        }
        mn = wave_reduce(mn, OpMinU());
        mx = wave_reduce(mx, OpMaxU());
        if (lane == 0) {
            uint32_t w = qa_width(mx - mn);
            tf[ch].dv_w = w;
            tf[ch].dv_min = mn;
            atomicAdd(&tot, 9ull + 128ull * w);
        }
    }
    __syncthreads();
    uint64_t delta = 8 + qa_varint_len(card) + tot;
    uint64_t raw = 8 * card;
    *vtag = delta < raw ? 1 : 0;
    *vlen = delta < raw ? delta : raw;
}

// the loads of these loops are straight-line code per lane storage kind, so
// the eight of a batch are in flight together
template <int LK>
__device__ __forceinline__ void bucket_loop(const void *lp, uint32_t lk, uint64_t s0, uint64_t n,
                                            uint64_t *bits, uint32_t *has_max) {
    const uint64_t bd = blockDim.x;
    bool hm = false;
    for (uint64_t i0 = threadIdx.x; i0 < n; i0 += 8 * bd) {
        uint64_t v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const uint64_t i = i0 + u * bd;
            v[u] = i < n ? lane_get_t<LK>(lp, lk, s0 + i) : 0;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            if (i0 + u * bd >= n) break;
            hm |= v[u] == HEMPTY;
            const uint32_t bk = (uint32_t)(mix64(v[u]) >> 45);
            atomicOr((unsigned long long *)&bits[bk >> 6], 1ull << (bk & 63));
        }
    }
    if (hm) *has_max = 1;
}

// slot of value v in the 8192-slot LDS table (any function of v gives the
// same set; the dictionary comes out sorted)
__device__ __forceinline__ uint32_t dh_hash(uint64_t v) {
    return ((uint32_t)(v ^ (v >> 32)) * 0x9E3779B1u) >> 19;
}
// insert v (not the all-ones value) into the LDS set; *overflow is raised when
// the set reaches DH_LIMIT or exceeds max_card
__device__ __forceinline__ void dh_insert(uint64_t *keys, uint64_t v, uint32_t *card_l,
                                          uint32_t *overflow, uint64_t max_card) {
    uint32_t h = dh_hash(v);
    for (uint32_t probe = 0; probe < DH_CAP; probe++) {
        uint64_t cur = keys[h];
        if (cur == v) return;
        if (cur == HEMPTY) {
            unsigned long long old =
                atomicCAS((unsigned long long *)&keys[h], HEMPTY, (unsigned long long)v);
            if (old == HEMPTY) {
                uint32_t c = atomicAdd(card_l, 1u) + 1;
                if (c >= DH_LIMIT || c > max_card) *overflow = 1;
                return;
            }
            if (old == v) return;
        }
        // once the set has overflowed it is discarded: stop probing (a full
        // table would otherwise be probed through for every later value)
        if (*(volatile uint32_t *)overflow) return;
        h = (h + 1) & (DH_CAP - 1);
    }
}

// exact insertion of values [s0, s0 + n) into the LDS set (dict phase 1)
template <int LK>
__device__ __forceinline__ void insert_loop(const void *lp, uint32_t lk, uint64_t s0, uint64_t n,
                                            uint64_t *keys, uint32_t *card_l, uint32_t *overflow,
                                            uint32_t *has_max, uint64_t max_card) {
    const uint64_t bd = blockDim.x;
    for (uint64_t i0 = threadIdx.x; i0 < n; i0 += 8 * bd) {
        if (*(volatile uint32_t *)overflow) break;
        uint64_t vs[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const uint64_t i = i0 + u * bd;
            vs[u] = i < n ? lane_get_t<LK>(lp, lk, s0 + i) : 0;
        }
        for (int u = 0; u < 8; u++) {
            if (i0 + u * bd >= n) break;
            const uint64_t v = vs[u];
            if (v == HEMPTY) {
                *has_max = 1;
                continue;
            }
            dh_insert(keys, v, card_l, overflow, max_card);
        }
    }
}

// Lower bound on the distinct count of values [s0, s0 + n): the number of
// occupied buckets among 2^19 hash buckets (distinct buckets need distinct
// values). bits: 8192 LDS words. Every thread returns the bound.
// (it reads every value, so it also reports whether the all-ones value occurs:
// *has_max, raised once per block by any thread)
__device__ uint64_t dict_bucket_bound(const DColState &st, uint64_t s0, uint64_t n, uint64_t *bits,
                                      uint32_t *has_max) {
    __shared__ unsigned long long lb;
    for (uint32_t i = threadIdx.x; i < DH_CAP; i += blockDim.x) bits[i] = 0;
    if (threadIdx.x == 0) lb = 0;
    __syncthreads();
    const void *lp = st.lanep;
    const uint32_t lk = st.lkind;
    switch (lk) {
    case QA_LK_U64: bucket_loop<QA_LK_U64>(lp, lk, s0, n, bits, has_max); break;
    case QA_LK_U32: bucket_loop<QA_LK_U32>(lp, lk, s0, n, bits, has_max); break;
    case QA_LK_I32: bucket_loop<QA_LK_I32>(lp, lk, s0, n, bits, has_max); break;
    case QA_LK_F32: bucket_loop<QA_LK_F32>(lp, lk, s0, n, bits, has_max); break;
    default: bucket_loop<QA_LK_ANY>(lp, lk, s0, n, bits, has_max);
    }
    __syncthreads();
    uint64_t pc = 0;
    for (uint32_t i = threadIdx.x; i < DH_CAP; i += blockDim.x) pc += __popcll(bits[i]);
    pc = wave_reduce(pc, OpAdd());
    if ((threadIdx.x & 63) == 0) atomicAdd(&lb, pc);
    __syncthreads();
    const uint64_t r = lb;
    __syncthreads();
    return r;
}

#define QA_LANE_SWITCH(F, lk, ...)                                   \
    do {                                                             \
        switch (lk) {                                                \
        case QA_LK_U64: F<QA_LK_U64>(__VA_ARGS__); break;            \
        case QA_LK_U32: F<QA_LK_U32>(__VA_ARGS__); break;            \
        case QA_LK_I32: F<QA_LK_I32>(__VA_ARGS__); break;            \
        case QA_LK_F32: F<QA_LK_F32>(__VA_ARGS__); break;            \
        default: F<QA_LK_ANY>(__VA_ARGS__);                          \
        }                                                            \
    } while (0)

// out[*pos++] = v for the lanes that keep theirs, in any order: one LDS atomic
// per wavefront (one per lane serializes on the counter)
template <typename T>
__device__ __forceinline__ void wave_append(bool keep, T v, T *out, unsigned int *pos) {
    const uint64_t m = __ballot(keep);
    if (!m) return;
    const int lane = threadIdx.x & 63, lead = __builtin_ctzll(m);
    uint32_t base = 0;
    if (lane == lead) base = atomicAdd(pos, (uint32_t)__popcll(m));
    base = (uint32_t)__builtin_amdgcn_readlane((int)base, lead);
    const uint32_t r =
        __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    if (keep) out[base + r] = v;
}

// A block grown past one region (131072 values) is too much for one
// workgroup: each of its region slots builds the exact distinct set of its
// part in LDS and leaves it in dparts (dpart[slot] = count | has-max << 31, or
// all ones when the part alone reaches the LDS limit). k_dict_set merges them.
__global__ void __launch_bounds__(1024) k_dict_part(EncArgs a) {
    __shared__ uint64_t keys[DH_CAP];
    __shared__ uint32_t card_l, overflow, has_max, wpos;
    const uint32_t slot = a.r0 + blockIdx.x;
    if (slot >= a.r1) return;
    const uint32_t o = a.owner[slot];
    if (o == NONE32) return;
    const DBlock &b = a.blocks[o];
    if (!dict_candidate(a, b)) return;
    const DCol &col = a.cols[b.col];
    if (b.len <= col.region_size) return;  // one-region blocks: k_dict_set alone
    const DColState &st = a.cst[b.col];
    const uint64_t ps = (uint64_t)(slot - col.region_base) * col.region_size;
    const uint64_t pe = min(ps + (uint64_t)col.region_size, b.start + (uint64_t)b.len);
    for (uint32_t i = threadIdx.x; i < DH_CAP; i += blockDim.x) keys[i] = HEMPTY;
    if (threadIdx.x == 0) {
        card_l = 0;
        overflow = 0;
        has_max = 0;
        wpos = 0;
    }
    __syncthreads();
    QA_LANE_SWITCH(insert_loop, st.lkind, st.lanep, st.lkind, ps, pe - ps, keys, &card_l, &overflow,
                   &has_max, ~0ull);
    __syncthreads();
    if (overflow) {
        if (threadIdx.x == 0) a.dpart[slot] = NONE32;
        return;
    }
    uint64_t *out = a.dparts + (uint64_t)slot * DH_LIMIT;
    for (uint32_t i = threadIdx.x; i < DH_CAP; i += blockDim.x)
        wave_append(keys[i] != HEMPTY, keys[i], out, &wpos);
    if (threadIdx.x == 0) a.dpart[slot] = card_l | (has_max << 31);
}

// Move the occupied slots of the LDS set (card_l of them) to the front, in
// slot order, pad with the all-ones value to a power of two of at least
// card_l + has_max (the all-ones value then lands at index card_l), and sort
// that prefix. Each thread reads its 8 slots before anyone writes.
__device__ void lds_compact_sort(uint64_t *keys, uint32_t card_l, uint32_t has_max) {
    __shared__ uint32_t wsum[16];
    const uint32_t per = DH_CAP / blockDim.x;  // 8 at 1024 threads
    const int lane = threadIdx.x & 63, wv = qa_wave_id();
    uint64_t v[8];
    uint32_t c = 0;
    for (uint32_t j = 0; j < per; j++) {
        v[j] = keys[threadIdx.x * per + j];
        c += v[j] != HEMPTY;
    }
    const uint32_t incl = (uint32_t)qa_wave_scan_add64(c);
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    uint32_t base = incl - c;
    for (int q = 0; q < wv; q++) base += wsum[q];
    for (uint32_t j = 0; j < per; j++)
        if (v[j] != HEMPTY) keys[base++] = v[j];
    uint32_t np2 = 2;
    while (np2 < card_l + has_max) np2 <<= 1;
    __syncthreads();
    for (uint32_t i = card_l + threadIdx.x; i < np2; i += blockDim.x) keys[i] = HEMPTY;
    __syncthreads();
    lds_bitonic_sort(keys, np2);
}

// insertion of the values of [s0, s0 + n) that fall in [lo, hi) (the
// all-ones value excluded) into the LDS set, overflowing at DH_LIMIT
template <int LK>
__device__ __forceinline__ void range_insert_loop(const void *lp, uint32_t lk, uint64_t s0,
                                                  uint64_t n, uint64_t lo, uint64_t hi,
                                                  uint64_t *keys, uint32_t *cnt, uint32_t *ovf) {
    const uint64_t bd = blockDim.x;
    for (uint64_t i0 = threadIdx.x; i0 < n; i0 += 8 * bd) {
        if (*(volatile uint32_t *)ovf) break;
        uint64_t vs[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const uint64_t i = i0 + u * bd;
            vs[u] = i < n ? lane_get_t<LK>(lp, lk, s0 + i) : HEMPTY;
        }
        for (int u = 0; u < 8; u++) {
            const uint64_t v = vs[u];
            if (v != HEMPTY && v >= lo && v < hi) dh_insert(keys, v, cnt, ovf, ~0ull);
        }
    }
}

// Exact sorted dictionary of a block whose cardinality is above the LDS set's
// limit: the values are split into P ranges at splitters taken from a sorted
// sample of 4096 values (about 2048 distinct values per range on average);
// each range's exact set is built and sorted in LDS and appended to sorted[]
// in range order, which is value order. Returns 0 (done: *card_out), 1 (more
// than max_card distinct values) or 2 (a range reached the LDS limit).
__device__ int dict_ranges(const DColState &st, uint64_t s0, uint64_t n, uint64_t max_card,
                           uint32_t has_max, uint64_t *keys, uint64_t *sorted,
                           uint64_t *card_out) {
    __shared__ uint64_t split[66];
    __shared__ uint32_t cnt_s, ovf_s;
    const void *lp = st.lanep;
    const uint32_t lk = st.lkind;
    const uint32_t ns = (uint32_t)(n < 4096 ? n : 4096);
    const uint64_t stride = n / ns;
    for (uint32_t i = threadIdx.x; i < 4096; i += blockDim.x)
        keys[i] = i < ns ? qa_lane_get(lp, lk, s0 + (uint64_t)i * stride) : HEMPTY;
    __syncthreads();
    lds_bitonic_sort(keys, 4096);
    uint32_t P = (uint32_t)((max_card + 2047) / 2048);
    if (P < 2) P = 2;
    if (P > 64) P = 64;
    if (threadIdx.x <= P)
        split[threadIdx.x] = threadIdx.x == 0 ? 0 : (threadIdx.x == P ? HEMPTY : keys[threadIdx.x * ns / P]);
    __syncthreads();
    uint64_t base = 0;
    for (uint32_t pr = 0; pr < P; pr++) {
        const uint64_t lo = split[pr], hi = split[pr + 1];
        if (lo >= hi) continue;
        for (uint32_t i = threadIdx.x; i < DH_CAP; i += blockDim.x) keys[i] = HEMPTY;
        if (threadIdx.x == 0) {
            cnt_s = 0;
            ovf_s = 0;
        }
        __syncthreads();
        QA_LANE_SWITCH(range_insert_loop, lk, lp, lk, s0, n, lo, hi, keys, &cnt_s, &ovf_s);
        __syncthreads();
        const uint32_t c = cnt_s;
        if (ovf_s) return 2;
        if (base + c + has_max > max_card) return 1;
        lds_compact_sort(keys, c, 0);
        for (uint32_t i = threadIdx.x; i < c; i += blockDim.x) sorted[base + i] = keys[i];
        base += c;
        __syncthreads();
    }
    if (has_max && threadIdx.x == 0) sorted[base] = HEMPTY;
    *card_out = base + has_max;
    __threadfence_block();
    __syncthreads();
    return 0;
}

__global__ void __launch_bounds__(1024) k_dict_set(EncArgs a) {
    __shared__ uint64_t keys[DH_CAP];
    __shared__ uint32_t card_l, overflow, has_max;
    __shared__ uint32_t hmax_b;  // the all-ones value occurs (seen by a full bucket pass)
    const uint32_t slot = a.r0 + blockIdx.x;
    if (slot >= a.r1) return;
    const DBlock &b = a.blocks[slot];
    if (!dict_candidate(a, b)) return;
    DBlockX &x = a.blocksx[slot];
    const DCol &col = a.cols[b.col];
    const DColState &st = a.cst[b.col];
    const uint64_t n = b.len, max_card = n / 2 + 1;
    DTileF *tf = a.tilesf + col.tile_base + (uint64_t)(slot - col.region_base) * col.tpr;
    uint64_t *sorted = a.dsorted + col.lane_off + b.start;
    // a grown block whose parts all fit: merge their sets (at most 4 x 6143
    // keys, below n / 2 + 1, so no bail is possible from the count alone)
    const uint32_t nparts = (uint32_t)((n + col.region_size - 1) / col.region_size);
    bool parts_ok = nparts > 1;
    for (uint32_t j = 0; j < nparts && parts_ok; j++) parts_ok = a.dpart[slot + j] != NONE32;
    // a block whose probe saw many distinct values first tries the cheap
    // certificate that its cardinality exceeds n / 2 + 1 (dict.rs bails then)
    bool lb_known = false;
    if (threadIdx.x == 0) hmax_b = 0;  // ordered by the bound's first barrier
    if (!b.distinct_low && !parts_ok) {
        const uint64_t lb = dict_bucket_bound(st, b.start, n, keys, &hmax_b);
        if (lb > max_card) {
            if (threadIdx.x == 0) {
                x.dict_state = 1;
                x.card = 0;
                x.dict_vlen = 0;
                x.dict_vtag = 0;
            }
            return;
        }
        lb_known = true;
    }
    for (uint32_t i = threadIdx.x; i < DH_CAP; i += blockDim.x) keys[i] = HEMPTY;
    if (threadIdx.x == 0) {
        card_l = 0;
        overflow = 0;
        has_max = 0;
    }
    __syncthreads();
    // phase 1: exact set in LDS (the all-ones value is tracked by a flag),
    // from the parts' sets or from the values
    if (parts_ok) {
        for (uint32_t j = 0; j < nparts; j++) {
            const uint32_t pc = a.dpart[slot + j];
            if (threadIdx.x == 0 && (pc >> 31)) has_max = 1;
            const uint64_t *pk = a.dparts + (uint64_t)(slot + j) * DH_LIMIT;
            const uint32_t cnt = pc & 0x7FFFFFFFu;
            for (uint32_t i = threadIdx.x; i < cnt; i += blockDim.x) {
                if (*(volatile uint32_t *)&overflow) break;
                dh_insert(keys, pk[i], &card_l, &overflow, max_card);
            }
        }
    } else {
        QA_LANE_SWITCH(insert_loop, st.lkind, st.lanep, st.lkind, b.start, n, keys, &card_l,
                       &overflow, &has_max, max_card);
    }
    __syncthreads();
    uint64_t card = 0;
    bool exact = false, bail = false;
    if (!overflow) {
        card = card_l + has_max;
        exact = true;
        bail = card > max_card;
    } else if (card_l + has_max > max_card) {
        bail = true;  // the LDS set already holds more distinct values than allowed
    } else {
        // phase 2: a lower bound on the cardinality from 2^19 hash buckets
        // (already known to be within the limit when it ran first)
        const uint64_t lb = lb_known ? 0 : dict_bucket_bound(st, b.start, n, keys, &hmax_b);
        // phase 1 stopped at the overflow, so its flag may have missed the
        // all-ones value; the bucket pass saw every value
        __syncthreads();
        const uint32_t hm = has_max | hmax_b;
        int ranged = 2;
        if (lb > max_card) {
            bail = true;
        } else {
            // phase 3: exact sets of value ranges in LDS
            uint64_t rc = 0;
            ranged = dict_ranges(st, b.start, n, max_card, hm, keys, sorted, &rc);
            if (ranged == 0) card = rc;
            if (ranged == 1) bail = true;
        }
        if (!bail && ranged == 2) {
            // phase 3b (a range overflowed): exact set in the block's global table
            uint32_t cap = 1;
            while (cap < n + 4) cap <<= 1;
            uint64_t *gk = a.hkeys + 4 * (col.lane_off + b.start);
            for (uint32_t i = threadIdx.x; i < cap; i += blockDim.x) gk[i] = HEMPTY;
            if (threadIdx.x == 0) {
                card_l = 0;
                overflow = 0;
            }
            __threadfence_block();
            __syncthreads();
            for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) {
                if (overflow) break;
                uint64_t v = lane_at(st, b.start + i);
                if (v == HEMPTY) continue;
                uint32_t h = (uint32_t)(mix64(v) & (cap - 1));
                for (;;) {
                    unsigned long long old =
                        atomicCAS((unsigned long long *)&gk[h], HEMPTY, (unsigned long long)v);
                    if (old == HEMPTY) {
                        if (atomicAdd(&card_l, 1u) + 1 + hm > max_card) overflow = 1;
                        break;
                    }
                    if (old == v) break;
                    h = (h + 1) & (cap - 1);
                }
            }
            __threadfence_block();
            __syncthreads();
            card = card_l + hm;
            bail = overflow || card > max_card;
            if (!bail) {
                // compact, pad to a power of two with the all-ones value, sort
                // in global memory (this path only runs for cardinalities
                // between the LDS limit and n/2 + 1)
                __shared__ unsigned int wpos;
                if (threadIdx.x == 0) wpos = 0;
                __syncthreads();
                uint64_t *buf = a.hkeys + 4 * (col.lane_off + b.start) + cap;
                uint32_t np2 = 1;
                while (np2 < card) np2 <<= 1;
                for (uint32_t i = threadIdx.x; i < cap; i += blockDim.x)
                    wave_append(gk[i] != HEMPTY, gk[i], buf, &wpos);
                __syncthreads();
                for (uint32_t i = threadIdx.x + wpos; i < np2; i += blockDim.x) buf[i] = HEMPTY;
                __threadfence_block();
                __syncthreads();
                for (uint32_t k = 2; k <= np2; k <<= 1) {
                    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                        for (uint32_t i = threadIdx.x; i < np2; i += blockDim.x) {
                            uint32_t ixj = i ^ j;
                            if (ixj > i) {
                                bool up = (i & k) == 0;
                                uint64_t p = buf[i], q = buf[ixj];
                                if ((p > q) == up) {
                                    buf[i] = q;
                                    buf[ixj] = p;
                                }
                            }
                        }
                        __threadfence_block();
                        __syncthreads();
                    }
                }
                for (uint32_t i = threadIdx.x; i < card; i += blockDim.x) sorted[i] = buf[i];
                __threadfence_block();
                __syncthreads();
            }
        }
    }
    if (exact && !bail) {
        lds_compact_sort(keys, card_l, has_max);
        for (uint32_t i = threadIdx.x; i < card; i += blockDim.x) sorted[i] = keys[i];
        __threadfence_block();
        __syncthreads();
    }
    uint64_t vlen = 0;
    uint32_t vtag = 0;
    if (!bail) dict_values_size(a, slot, sorted, card, tf, &vlen, &vtag);
    if (threadIdx.x == 0) {
        x.dict_state = bail ? 1 : 0;
        x.card = card;
        x.dict_vlen = vlen;
        x.dict_vtag = vtag;
    }
}

__device__ __forceinline__ uint64_t rank_of(const uint64_t *d, uint64_t card, uint64_t v) {
    uint64_t lo = 0, hi = card;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (d[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// code range per tile: ranks of the tile's smallest and largest value
__global__ void k_dict_tiles(EncArgs a) {
    const uint64_t t = a.t0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.t1) return;
    const DTile &ts = a.tiles[t];
    if (!ts.active) return;
    int c;
    uint32_t o, bt;
    if (!tile_owner(a, t, &c, &o, &bt)) return;
    const DBlock &b = a.blocks[o];
    if (!dict_candidate(a, b) || a.blocksx[o].dict_state != 0) return;
    const DCol &col = a.cols[c];
    const uint64_t *d = a.dsorted + col.lane_off + b.start;
    uint64_t card = a.blocksx[o].card;
    uint64_t r0 = rank_of(d, card, ts.umin), r1 = rank_of(d, card, ts.umax);
    a.tilesf[t].dict_w = qa_width(r1 - r0);
    a.tilesf[t].dict_min = r0;
}

// ---------------------------------------------------------------------------
// RLE: run starts, then min / max length per chunk of 1024 runs
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool rle_candidate(const EncArgs &a, const DBlock &b, uint32_t o) {
    return b.len && b.compressible && a.cols[b.col].level != QA_FASTEST && !a.blocksx[o].rle_bail;
}

__global__ void __launch_bounds__(256) k_rle_starts(EncArgs a) {
    const int lane = threadIdx.x & 63;
    const uint64_t t = a.t0 + (uint64_t)blockIdx.x * 4 + qa_wave_id();
    if (t >= a.t1) return;
    const DTile &ts = a.tiles[t];
    if (!ts.active) return;
    int c;
    uint32_t o, bt;
    if (!tile_owner(a, t, &c, &o, &bt)) return;
    const DBlock &b = a.blocks[o];
    if (!rle_candidate(a, b, o)) return;
    const DCol &col = a.cols[c];
    const DColState &st = a.cst[c];
    const uint64_t p0 = (uint64_t)bt * QA_TILE;
    uint32_t *rs = a.rstart + col.lane_off + b.start;
    uint64_t base = a.tilesf[t].rle_base;
    for (uint32_t kb = 0; kb < ts.cnt; kb += WAVE) {
        uint32_t k = kb + lane;
        bool start = false;
        if (k < ts.cnt) {
            uint64_t p = p0 + k;
            start = p == 0 || lane_at(st, b.start + p) != lane_at(st, b.start + p - 1);
        }
        uint64_t bal = __ballot(start);
        if (start) rs[base + __popcll(bal & ((1ull << lane) - 1))] = (uint32_t)(p0 + k);
        base += __popcll(bal);
    }
}

__global__ void __launch_bounds__(256) k_rle_chunks(EncArgs a) {
    __shared__ unsigned long long tot;
    const uint32_t slot = a.r0 + blockIdx.x;
    if (slot >= a.r1) return;
    const DBlock &b = a.blocks[slot];
    if (!rle_candidate(a, b, slot)) return;
    DBlockX &x = a.blocksx[slot];
    const DCol &col = a.cols[b.col];
    const uint32_t *rs = a.rstart + col.lane_off + b.start;
    DTileF *tf = a.tilesf + col.tile_base + (uint64_t)(slot - col.region_base) * col.tpr;
    const int lane = threadIdx.x & 63, wv = qa_wave_id();
    const uint64_t R = x.nruns;
    if (threadIdx.x == 0) tot = 0;
    __syncthreads();
    for (uint64_t ch = wv; ch * QA_TILE < R; ch += 4) {
        uint64_t mn = ~0ull, mx = 0;
        for (uint64_t k = ch * QA_TILE + lane; k < min(R, (ch + 1) * QA_TILE); k += WAVE) {
            uint64_t len = (k + 1 < R ? rs[k + 1] : b.len) - rs[k];
            mn = len < mn ? len : mn;
            mx = len > mx ? len : mx;
        }
        mn = wave_reduce(mn, OpMinU());
        mx = wave_reduce(mx, OpMaxU());
        if (lane == 0) {
            uint32_t w = qa_width(mx - mn);
            tf[ch].ch_w = w;
            tf[ch].ch_min = mn;
            atomicAdd(&tot, 9ull + 128ull * w);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) x.rle_llen = qa_varint_len(R) + tot;
}

// ---------------------------------------------------------------------------
// block decision, part 2: Dict and RLE, then the winner's layout
// ---------------------------------------------------------------------------
template <typename F>
__device__ __forceinline__ void wave_offsets(uint32_t count, uint64_t carry, int lane, F size_of,
                                             uint64_t *out_total, DTileF *tf, int which,
                                             DTile *tl) {
    for (uint32_t base = 0; base < count; base += WAVE) {
        uint32_t t = base + lane;
        uint64_t sz = t < count ? size_of(t) : 0, inc = sz;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            uint64_t y = __shfl_up(inc, o, 64);
            if (lane >= o) inc += y;
        }
        if (t < count) {
            uint64_t off = carry + inc - sz;
            switch (which) {
            case 0: tl[t].off = off; break;
            case 1: tf[t].x_off = off; break;
            case 2: tf[t].alp_off = off; break;
            case 3: tf[t].rd_exc_off = off; break;
            case 4: tf[t].rd_coff = off; break;
            case 5: tf[t].rd_roff = off; break;
            case 6: tf[t].dv_off = off; break;
            case 7: tf[t].ch_off = off; break;
            }
        }
        carry += __shfl(inc, 63, 64);
    }
    *out_total = carry;
}

__global__ void __launch_bounds__(256) k_block_final(EncArgs a) {
    const int lane = threadIdx.x & 63;
    const uint32_t slot = a.r0 + blockIdx.x * 4 + qa_wave_id();
    if (slot >= a.r1) return;
    DBlock &b = a.blocks[slot];
    if (b.len == 0) return;
    const DCol &col = a.cols[b.col];
    DTile *tl = a.tiles + col.tile_base + (uint64_t)(slot - col.region_base) * col.tpr;
    DTileF *tf = a.tilesf ? a.tilesf + col.tile_base + (uint64_t)(slot - col.region_base) * col.tpr
                          : nullptr;
    const uint64_t n = b.len;
    const bool sg = is_signed(col.dtype), reduced = col.level == QA_FASTEST;
    const uint64_t lambda = reduced ? 16 : 4, dec = n * lane_bytes_of(col.dtype);
    uint32_t mode = b.mode;
    uint64_t best_len = b.payload_len, best_score = b.score;
    if (!reduced && b.compressible) {
        DBlockX &x = a.blocksx[slot];
        uint64_t dsz = ~0ull, rsz = ~0ull;
        if (x.dict_state == 0) {
            uint64_t cl = 0;
            for (uint32_t t = lane; t < b.ntiles; t += WAVE) cl += 9 + 128ull * tf[t].dict_w;
            cl = wave_reduce(cl, OpAdd()) + qa_varint_len(n);
            x.dict_clen = cl;
            dsz = qa_varint_len(x.card) + 1 + qa_varint_len(x.dict_vlen) + x.dict_vlen + 1 + cl;
        }
        if (!x.rle_bail) rsz = qa_varint_len(x.nruns) + 8 * x.nruns + x.rle_llen;
        consider(dsz, QA_M_DICT, lambda, dec, &best_score, &best_len, &mode);
        consider(rsz, QA_M_RLE, lambda, dec, &best_score, &best_len, &mode);
        if (lane == 0) {
            b.sz[C_DICT] = dsz;
            b.sz[C_RLE] = rsz;
        }
    }
    if (a.force) {
        const HForce &f = a.force[slot];
        mode = f.host ? QA_M_HOST : f.mode;
        best_len = f.len;
    }
    if (lane == 0) {
        b.mode = mode;
        b.payload_len = best_len;
    }
    uint64_t tot;
    const uint64_t vn = qa_varint_len(n);
    switch (mode) {
    case QA_M_FOR:
    case QA_M_DELTA: {
        bool dl = mode == QA_M_DELTA;
        for (uint32_t t = lane; t < b.ntiles; t += WAVE) {
            uint64_t mn, mx;
            if (dl) tile_dz(tl[t], t == 0, &mn, &mx);
            tl[t].w = qa_width(dl ? mx - mn : tile_for_range(tl[t], sg));
        }
        wave_offsets(b.ntiles, dl ? 8 + vn : vn, lane,
                     [&](uint32_t t) { return 9 + 128ull * tl[t].w; }, &tot, tf, 0, tl);
        break;
    }
    case QA_M_XORZ:
        wave_offsets(b.ntiles, 0, lane, [&](uint32_t t) {
            return tl[t].xorz_rest + qa_varint_len(tl[t].v_first ^ (t ? tl[t - 1].v_last : 0));
        }, &tot, tf, 1, tl);
        break;
    case QA_M_FLOAT_MULT: {
        uint32_t s = a.blocksx[slot].fm_scale;
        for (uint32_t t = lane; t < b.ntiles; t += WAVE)
            tf[t].fm_w = qa_width((uint64_t)tf[t].kmax[s] - (uint64_t)tf[t].kmin[s]);
        wave_offsets(b.ntiles, 2 + vn, lane, [&](uint32_t t) { return 9 + 128ull * tf[t].fm_w; },
                     &tot, tf, 0, tl);
        break;
    }
    case QA_M_ALP:
        wave_offsets(b.ntiles, vn, lane, [&](uint32_t t) {
            return 13 + 128ull * tf[t].alp_w + 10ull * tf[t].alp_exc;
        }, &tot, tf, 2, tl);
        break;
    case QA_M_ALP_RD: {
        const DBlockX &x = a.blocksx[slot];
        uint64_t e0 = 2 + 8ull * x.rd_dl + qa_varint_len(x.rd_nexc);
        wave_offsets(b.ntiles, e0, lane, [&](uint32_t t) { return (uint64_t)tf[t].rd_exc_bytes; },
                     &tot, tf, 3, tl);
        uint64_t cs = e0 + x.rd_excb + 1 + qa_varint_len(x.rd_clen);
        for (uint32_t t = lane; t < b.ntiles; t += WAVE) {
            tf[t].rd_wc = qa_width(tf[t].rd_cmax - tf[t].rd_cmin);
            tf[t].rd_wr = qa_width(tf[t].rd_rmax - tf[t].rd_rmin);
        }
        wave_offsets(b.ntiles, cs + vn, lane, [&](uint32_t t) { return 9 + 128ull * tf[t].rd_wc; },
                     &tot, tf, 4, tl);
        uint64_t rs0 = cs + x.rd_clen + 1;
        wave_offsets(b.ntiles, rs0 + vn, lane, [&](uint32_t t) { return 9 + 128ull * tf[t].rd_wr; },
                     &tot, tf, 5, tl);
        break;
    }
    case QA_M_DICT: {
        const DBlockX &x = a.blocksx[slot];
        uint64_t v0 = qa_varint_len(x.card) + 1 + qa_varint_len(x.dict_vlen);
        if (x.dict_vtag == 1) {
            uint32_t nch = (uint32_t)((x.card + QA_TILE - 1) / QA_TILE);
            wave_offsets(nch, v0 + 8 + qa_varint_len(x.card), lane,
                         [&](uint32_t c) { return 9 + 128ull * tf[c].dv_w; }, &tot, tf, 6, tl);
        }
        uint64_t cs = v0 + x.dict_vlen + 1;
        wave_offsets(b.ntiles, cs + vn, lane, [&](uint32_t t) { return 9 + 128ull * tf[t].dict_w; },
                     &tot, tf, 0, tl);
        break;
    }
    case QA_M_RLE: {
        const DBlockX &x = a.blocksx[slot];
        uint32_t nch = (uint32_t)((x.nruns + QA_TILE - 1) / QA_TILE);
        uint64_t ls = qa_varint_len(x.nruns) + 8 * x.nruns + qa_varint_len(x.nruns);
        wave_offsets(nch, ls, lane, [&](uint32_t c) { return 9 + 128ull * tf[c].ch_w; }, &tot, tf,
                     7, tl);
        break;
    }
    default: break;
    }
}

// ---------------------------------------------------------------------------
static inline unsigned nbf(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }
#define QA_LAUNCH(k, grid, ...)                                   \
    do {                                                          \
        dim3 g_ = (grid);                                         \
        if (g_.x) qa_launch(k, g_, __VA_ARGS__);         \
    } while (0)

// before k_block_decide: the float-value codecs' statistics; then the decision
void qa_fast_sizes_launch(const EncArgs &a, hipStream_t s) {
    if (a.fast) {
        QA_LAUNCH(k_float_tiles, dim3(nbf(a.t1 - a.t0, 4)), dim3(256), 0, s, a);
        QA_LAUNCH(k_fm_resolve, dim3(a.r1 - a.r0), dim3(256), 0, s, a);
        QA_LAUNCH(k_alprd_cut, dim3(a.r1 - a.r0), dim3(1024), 0, s, a);
        QA_LAUNCH(k_alprd_hist, dim3(a.r1 - a.r0), dim3(1024), 0, s, a);
        QA_LAUNCH(k_alprd_tiles, dim3(nbf(a.t1 - a.t0, 4)), dim3(256), 0, s, a);
    }
    QA_LAUNCH(k_block_decide, dim3(nbf(a.r1 - a.r0, 4)), dim3(256), 0, s, a);
}

// Dict / RLE sizing, then the final decision and layout
void qa_fast_final_launch(const EncArgs &a, hipStream_t s) {
    if (a.fast) {
        QA_LAUNCH(k_dict_part, dim3(a.r1 - a.r0), dim3(1024), 0, s, a);
        QA_LAUNCH(k_dict_set, dim3(a.r1 - a.r0), dim3(1024), 0, s, a);
        QA_LAUNCH(k_dict_tiles, dim3(nbf(a.t1 - a.t0, 256)), dim3(256), 0, s, a);
        QA_LAUNCH(k_rle_starts, dim3(nbf(a.t1 - a.t0, 4)), dim3(256), 0, s, a);
        QA_LAUNCH(k_rle_chunks, dim3(a.r1 - a.r0), dim3(256), 0, s, a);
    }
    QA_LAUNCH(k_block_final, dim3(nbf(a.r1 - a.r0, 4)), dim3(256), 0, s, a);
}

// Levels Balanced / High / Max, once the host competition has filled
// a.force: the chosen modes' layout (the sizing stages ran before)
void qa_fast_force_launch(const EncArgs &a, hipStream_t s) {
    QA_LAUNCH(k_block_final, dim3(nbf(a.r1 - a.r0, 4)), dim3(256), 0, s, a);
}
