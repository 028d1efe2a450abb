// qa_fast_dec.hip — decoders of the Level::Fast pool.
//
//   k_dec_index_fast  one wavefront per frame: payload prefixes and sub-block
//                     walks of FloatMult (float_mult.rs:100-133), ALP
//                     (alp.rs:168-248), ALP-RD (alp_rd.rs:208-285), Dict
//                     (dict.rs:154-207) and RLE (rle.rs:45-75), every length
//                     bounds-checked against the frame
//   k_dec_frame_wg    one workgroup (512 threads) per frame: XORZ varints + prefix XOR
//                     (xorz.rs:18-31) into the frame's value slots, Dict
//                     values (raw or delta-bitpacked), RLE run lengths -> run
//                     starts
// The per-tile decoders of every mode are in qa_dec_tiles.hip.
#include "qa_dec.h"

// All lanes parse the payload prefixes together (same addresses, uniform
// control flow); the sub-block walks use the whole wavefront.
__global__ void __launch_bounds__(256) k_dec_index_fast(DecArgs a) {
    const int lane = threadIdx.x & 63;
    uint32_t f = a.f0 + blockIdx.x * 4 + qa_wave_id();
    if (f >= a.f1 || *a.err || !frame_used(a, f)) return;
    const DFrame &fr = a.frames[f];
    if (!fast_mode(fr.mode) || fr.mode == QA_M_XORZ) return;
    const DDecCol &col = a.cols[fr.col];
    const uint8_t *s = col.src;
    DTileD *tl = a.tiles + fr.tile_base;
    DFrameX &x = a.fx[f];
    const uint64_t n = fr.n;
    uint64_t pos = fr.payload_off, end = fr.payload_off + fr.plen;
    uint32_t err = 0;
    switch (fr.mode) {
    case QA_M_FLOAT_MULT: {
        if (fr.plen < 2) { err = QA_ERR_CORRUPT; break; }
        uint8_t tag = s[pos], sc = s[pos + 1];
        if (sc >= 7) { err = QA_ERR_CORRUPT; break; }
        if (tag == 1) { err = QA_ERR_UNSUPPORTED; break; }  // entropy-coded k deltas
        if (tag != 0) { err = QA_ERR_CORRUPT; break; }
        if (lane == 0) x.scale = sc;
        if (for_walk_wave(s, pos + 2, end, n, tl, 0, lane) == ~0ull) err = QA_ERR_CORRUPT;
        break;
    }
    case QA_M_ALP: {
        uint64_t m;
        if (!rd_varint(s, end, &pos, &m) || m != n) { err = QA_ERR_CORRUPT; break; }
        if (alp_walk_wave(s, pos, end, fr.ntiles, tl, lane) == ~0ull) err = QA_ERR_CORRUPT;
        break;
    }
    case QA_M_ALP_RD: {
        if (pos + 2 > end) { err = QA_ERR_CORRUPT; break; }
        uint32_t left = s[pos], dl = s[pos + 1];
        if (left < 1 || left > 63 || dl > 8) { err = QA_ERR_CORRUPT; break; }
        pos += 2;
        if (pos + 8ull * dl > end) { err = QA_ERR_CORRUPT; break; }
        if (lane == 0)
            for (uint32_t q = 0; q < dl; q++) x.dict[q] = qa_ld64u(s + pos + 8 * q);
        pos += 8ull * dl;
        uint64_t nexc;
        if (!rd_varint(s, end, &pos, &nexc) || nexc > n) { err = QA_ERR_CORRUPT; break; }
        if (lane == 0) {
            x.left = left;
            x.dl = dl;
            x.nexc = nexc;
            x.exc_off = pos;
        }
        uint32_t sorted = 1;
        uint64_t last = 0;
        for (uint64_t j = 0; j < nexc && !err; j++) {
            uint64_t p;
            if (!rd_varint(s, end, &pos, &p) || pos + 8 > end || p >= n) { err = QA_ERR_CORRUPT; break; }
            if (lane == 0) {
                a.scr[fr.scr + j] = p;
                a.scr2[fr.scr + j] = qa_ld64u(s + pos);
            }
            pos += 8;
            if (j && p <= last) sorted = 0;
            last = p;
        }
        if (err) break;
        if (lane == 0) x.sorted = sorted;
        if (pos >= end) { err = QA_ERR_CORRUPT; break; }
        uint8_t ctag = s[pos++];
        uint64_t clen;
        if (!rd_varint(s, end, &pos, &clen) || clen > end - pos) { err = QA_ERR_CORRUPT; break; }
        uint64_t cend = pos + clen;
        if (cend >= end) { err = QA_ERR_CORRUPT; break; }
        uint8_t rtag = s[cend];
        if (ctag == 1 || rtag == 1) { err = QA_ERR_UNSUPPORTED; break; }
        if (ctag != 0 || rtag != 0) { err = QA_ERR_CORRUPT; break; }
        if (for_walk_wave(s, pos, cend, n, tl, 0, lane) == ~0ull ||
            for_walk_wave(s, cend + 1, end, n, tl, 1, lane) == ~0ull)
            err = QA_ERR_CORRUPT;
        break;
    }
    case QA_M_DICT: {
        uint64_t card, vlen;
        if (!rd_varint(s, end, &pos, &card) || card > n) { err = QA_ERR_CORRUPT; break; }
        if (pos >= end) { err = QA_ERR_CORRUPT; break; }
        uint8_t vtag = s[pos++];
        if (!rd_varint(s, end, &pos, &vlen) || vlen > end - pos) { err = QA_ERR_CORRUPT; break; }
        if (lane == 0) {
            x.card = card;
            x.vtag = vtag;
            x.aux_off = pos;
        }
        uint64_t vend = pos + vlen;
        if (vtag == 0) {
            if (vlen != card * 8) { err = QA_ERR_CORRUPT; break; }
        } else if (vtag == 1) {
            if (vlen < 8 || for_walk_wave(s, pos + 8, vend, card, tl, 1, lane) == ~0ull) {
                err = QA_ERR_CORRUPT;
                break;
            }
        } else if (vtag == 2) {
            err = QA_ERR_UNSUPPORTED;  // transposed + entropy-coded values
            break;
        } else {
            err = QA_ERR_CORRUPT;
            break;
        }
        if (vend >= end) { err = QA_ERR_CORRUPT; break; }
        uint8_t ctag = s[vend];
        if (ctag == 1) { err = QA_ERR_UNSUPPORTED; break; }
        if (ctag != 0 || for_walk_wave(s, vend + 1, end, n, tl, 0, lane) == ~0ull)
            err = QA_ERR_CORRUPT;
        break;
    }
    case QA_M_RLE: {
        uint64_t R;
        if (!rd_varint(s, end, &pos, &R) || R > n) { err = QA_ERR_CORRUPT; break; }
        if ((end - pos) / 8 < R) { err = QA_ERR_CORRUPT; break; }
        if (lane == 0) {
            x.card = R;
            x.aux_off = pos;
        }
        if (for_walk_wave(s, pos + 8 * R, end, R, tl, 1, lane) == ~0ull) err = QA_ERR_CORRUPT;
        break;
    }
    default: break;
    }
    if (err && lane == 0) atomicOr(a.err, err);
}

// ---------------------------------------------------------------------------
// workgroup-wide scans (up to 1024 threads)
// ---------------------------------------------------------------------------
template <typename T, typename Op>
__device__ __forceinline__ T wg_incl_scan(T v, Op op, T *tot) {
    __shared__ T wpart[16];
    const int lane = threadIdx.x & 63, wv = qa_wave_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T y = __shfl_up(v, o, 64);
        if (lane >= o) v = op(v, y);
    }
    if (lane == 63) wpart[wv] = v;
    __syncthreads();
    T pre = T(0);
    for (int q = 0; q < wv; q++) pre = op(pre, wpart[q]);
    T all = T(0);
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) all = op(all, wpart[q]);
    __syncthreads();
    *tot = all;
    return op(pre, v);
}
struct OpXor64 { __device__ uint64_t operator()(uint64_t a, uint64_t b) const { return a ^ b; } };
struct OpAdd64 { __device__ uint64_t operator()(uint64_t a, uint64_t b) const { return a + b; } };

// FoR sub-block value k at stream sub-block header h

__global__ void __launch_bounds__(512) k_dec_frame_wg(DecArgs a) {
    const uint32_t f = a.f0 + blockIdx.x;
    if (*a.err || f >= a.f1 || !frame_used(a, f)) return;
    const DFrame &fr = a.frames[f];
    if (fr.mode != QA_M_XORZ && fr.mode != QA_M_DICT && fr.mode != QA_M_RLE) return;
    const DDecCol &col = a.cols[fr.col];
    const uint8_t *s = col.src;
    const DFrameX &x = a.fx[f];
    uint64_t *scr = a.scr + fr.scr;
    const uint64_t n = fr.n;
    __shared__ uint64_t carry_s;
    __shared__ uint32_t bad;
    __shared__ __align__(16) uint32_t wst[8][1088];  // staged sub-block words, per wavefront
    if (threadIdx.x == 0) {
        carry_s = 0;
        bad = 0;
    }
    __syncthreads();
    if (fr.mode == QA_M_XORZ) {
        // varint k ends at the k-th byte whose top bit is clear
        const uint8_t *p = s + fr.payload_off;
        const uint64_t len = fr.plen;
        for (uint64_t b0 = 0; b0 < len; b0 += blockDim.x) {
            uint64_t i = b0 + threadIdx.x;
            bool endb = i < len && !(p[i] & 0x80);
            uint64_t tot;
            uint64_t inc = wg_incl_scan((uint64_t)endb, OpAdd64(), &tot);
            if (endb) {
                uint64_t idx = carry_s + inc - 1;
                uint64_t j = i;
                while (j > 0 && (p[j - 1] & 0x80) && i - j < 10) j--;
                if (i - j + 1 > 10) {
                    bad = 1;  // more than ten bytes: shift overflow (varint.rs:33-35)
                } else if (idx < n) {
                    uint64_t v = 0;
                    for (uint64_t q = j; q <= i; q++) v |= (uint64_t)(p[q] & 0x7f) << (7 * (q - j));
                    scr[idx] = v;
                }
            }
            __syncthreads();
            if (threadIdx.x == 0) carry_s += tot;
            __syncthreads();
        }
        // exactly n varints and no trailing bytes
        if (threadIdx.x == 0 && (carry_s != n || (len && (p[len - 1] & 0x80)))) bad = 1;
        __syncthreads();
        if (bad) {
            if (threadIdx.x == 0) atomicOr(a.err, QA_ERR_CORRUPT);
            return;
        }
        // values in place of the varints; each slot is read and written by
        // the same thread
        uint64_t acc = 0;
        for (uint64_t b0 = 0; b0 < n; b0 += blockDim.x) {
            uint64_t i = b0 + threadIdx.x;
            uint64_t v = i < n ? scr[i] : 0, tot;
            uint64_t inc = wg_incl_scan(v, OpXor64(), &tot);
            if (i < n) scr[i] = acc ^ inc;
            acc ^= tot;
        }
        return;
    }
    // Dict values (vtag 0 raw, 1 delta-bitpacked) / RLE lengths -> run starts
    const bool dict = fr.mode == QA_M_DICT;
    if (dict && x.vtag == 0) {
        for (uint64_t i = threadIdx.x; i < x.card; i += blockDim.x)
            scr[i] = qa_ld64u(s + x.aux_off + 8 * i);
        return;
    }
    const uint64_t m = x.card;  // dictionary size or number of runs
    DTileD *tl = a.tiles + fr.tile_base;
    uint64_t acc = dict ? qa_ld64u(s + x.aux_off) : 0;  // delta base / first run start
    // RLE: run k covers [st, st + len); the value tiles whose first position it
    // covers record k, the last run starting at or before that position
    auto mark_tiles = [&](uint64_t k, uint64_t st, uint64_t len) {
        if (dict || !len) return;
        const uint64_t e = st + len < n ? st + len : n;
        for (uint64_t t = (st + QA_TILE - 1) / QA_TILE; t * QA_TILE < e; t++) tl[t].run0 = k;
    };
    // (512 threads: the 16 entries a thread holds fit its registers without
    // scratch; a 1024-thread form spilled and was not kept)
    // Wavefront q of the workgroup unpacks sub-block q of each group of 8
    // (1024 entries) through LDS, so lane L holds entries 16 L .. 16 L + 15 of
    // it: every thread then holds 16 consecutive entries, in entry order, and
    // one workgroup scan of the thread sums places them.
    const int lane = threadIdx.x & 63, wv = qa_wave_id();
    uint32_t *wd = wst[wv];
    uint64_t tot = 0;
    for (uint64_t g0 = 0; g0 < m; g0 += 8ull * QA_TILE) {
        const uint64_t sb = g0 / QA_TILE + wv, i0 = sb * QA_TILE + 16ull * lane;
        uint64_t v[16], sum = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) v[k] = 0;
        if (sb * QA_TILE < m) {
            const uint8_t *h = s + tl[sb].tail;
            const uint32_t w = h[0];
            const uint64_t mn = qa_ld64u(h + 1);
            if (w <= 32) {
                wave_lds_sync();  // the group before is done with the staged words
                stage_words(h + 9, 128 * w, wd, lane);
                wave_lds_sync();
                uint32_t u[16];
                unpack16_u32(wd, w, lane, u);
#pragma unroll
                for (int k = 0; k < 16; k++) v[k] = mn + u[k];
            } else {
#pragma unroll
                for (int k = 0; k < 16; k++) v[k] = for_value_direct(h, 16u * lane + k);
            }
            // an entry is a dictionary delta, or a run length (a run past n
            // reads as n + 1, an error below); entries past m are 0
#pragma unroll
            for (int k = 0; k < 16; k++) {
                const uint64_t r = v[k];
                v[k] = i0 + k >= m ? 0 : dict ? qa_unzigzag(r) : (r > n ? n + 1 : r);
                sum += v[k];
            }
        }
        const uint64_t inc = wg_incl_scan(sum, OpAdd64(), &tot);
        uint64_t run = acc + inc - sum;  // before this thread's first entry
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const uint64_t i = i0 + k;
            if (i < m && sb * QA_TILE < m) {
                if (dict) {
                    run += v[k];
                    scr[i] = run;  // values
                } else {
                    scr[i] = run;  // run starts
                    mark_tiles(i, run, v[k]);
                    run += v[k];
                }
            }
        }
        acc += tot;
    }
    // rle.rs:60-74: the runs must cover exactly n values
    if (!dict && threadIdx.x == 0 && acc != n) atomicOr(a.err, QA_ERR_CORRUPT);
}

static inline unsigned nbd(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

void qa_fast_dec_index_launch(const DecArgs &a, hipStream_t s) {
    if (!a.fast || a.f1 <= a.f0) return;
    qa_launch(k_dec_index_fast, dim3(nbd(a.f1 - a.f0, 4)), dim3(256), 0, s, a);
    qa_launch(k_dec_frame_wg, dim3(a.f1 - a.f0), dim3(512), 0, s, a);
}
