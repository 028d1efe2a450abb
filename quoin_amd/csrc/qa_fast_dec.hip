// qa_fast_dec.hip — decoders of the Level::Fast pool.
//
//   k_dec_index_fast  one wavefront per frame: payload prefixes and sub-block
//                     walks of FloatMult (float_mult.rs:100-133), ALP
//                     (alp.rs:168-248), ALP-RD (alp_rd.rs:208-285), Dict
//                     (dict.rs:154-207) and RLE (rle.rs:45-75), every length
//                     bounds-checked against the frame
//   k_dec_frame_wg    one workgroup per frame: XORZ varints + prefix XOR
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
// workgroup-wide scans (1024 threads)
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

__global__ void __launch_bounds__(1024) k_dec_frame_wg(DecArgs a) {
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
    const DTileD *tl = a.tiles + fr.tile_base;
    uint64_t acc = dict ? qa_ld64u(s + x.aux_off) : 0;  // delta base / first run start
    for (uint64_t b0 = 0; b0 < m; b0 += blockDim.x) {
        uint64_t i = b0 + threadIdx.x, v = 0;
        if (i < m) {
            uint64_t r = for_value_direct(s + tl[i / QA_TILE].tail, (uint32_t)(i % QA_TILE));
            if (dict) {
                v = qa_unzigzag(r);
            } else {
                v = r > n ? n + 1 : r;  // a run past n is an error below
            }
        }
        uint64_t tot;
        uint64_t inc = wg_incl_scan(v, OpAdd64(), &tot);
        if (i < m) scr[i] = dict ? acc + inc : acc + inc - v;  // values / run starts
        acc += tot;
    }
    // rle.rs:60-74: the runs must cover exactly n values
    if (!dict && threadIdx.x == 0 && acc != n) atomicOr(a.err, QA_ERR_CORRUPT);
}

static inline unsigned nbd(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

void qa_fast_dec_index_launch(const DecArgs &a, hipStream_t s) {
    if (!a.fast || a.f1 <= a.f0) return;
    qa_launch(k_dec_index_fast, dim3(nbd(a.f1 - a.f0, 4)), dim3(256), 0, s, a);
    qa_launch(k_dec_frame_wg, dim3(a.f1 - a.f0), dim3(1024), 0, s, a);
}
