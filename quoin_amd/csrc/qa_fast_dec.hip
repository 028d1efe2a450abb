// qa_fast_dec.hip — decoders of the Level::Fast pool.
//
//   k_dec_index_fast  one thread per frame: payload prefixes and sub-block
//                     walks of FloatMult (float_mult.rs:100-133), ALP
//                     (alp.rs:168-248), ALP-RD (alp_rd.rs:208-285), Dict
//                     (dict.rs:154-207) and RLE (rle.rs:45-75), every length
//                     bounds-checked against the frame
//   k_dec_frame_wg    one workgroup per frame: XORZ varints + prefix XOR
//                     (xorz.rs:18-31), Dict values (raw or delta-bitpacked),
//                     RLE run lengths -> run starts
//   k_dec_tiles_fast  one wavefront per 1024-value tile
#include "qa_dec.h"

__constant__ double d_exp10[19] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8, 1e9,
                                   1e10, 1e11, 1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18};
__constant__ double d_frac10[19] = {1e0,   1e-1,  1e-2,  1e-3,  1e-4,  1e-5,  1e-6,
                                    1e-7,  1e-8,  1e-9,  1e-10, 1e-11, 1e-12, 1e-13,
                                    1e-14, 1e-15, 1e-16, 1e-17, 1e-18};
__constant__ double d_fm_scales[7] = {10.0, 100.0, 1000.0, 10000.0, 100000.0, 1000000.0,
                                      10000000.0};

// Walk a for_bitpack payload [pos, end) that must hold n values
// (for_bitpack.rs:108-157); records sub-block offsets (which = 0: DTile.off
// with .w, which = 1: DTile.tail). Returns the position after the walk or
// ~0 on a malformed payload.
__device__ uint64_t for_walk(const uint8_t *s, uint64_t pos, uint64_t end, uint64_t n, DTile *tl,
                             int which) {
    uint64_t m;
    if (!rd_varint(s, end, &pos, &m) || m != n) return ~0ull;
    uint64_t nsub = (n + QA_TILE - 1) / QA_TILE;
    for (uint64_t t = 0; t < nsub; t++) {
        if (pos + 9 > end) return ~0ull;
        uint32_t w = s[pos];
        if (w > 64) return ~0ull;
        if (which == 0) {
            tl[t].off = pos;
            tl[t].w = w;
        } else {
            tl[t].tail = pos;
        }
        pos += 9;
        if (w) {
            if (pos + 128ull * w > end) return ~0ull;
            pos += 128ull * w;
        }
    }
    return pos;
}

__global__ void k_dec_index_fast(DecArgs a) {
    uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= a.nframes || *a.err || !frame_used(a, f)) return;
    const DFrame &fr = a.frames[f];
    if (!fast_mode(fr.mode) || fr.mode == QA_M_XORZ) return;
    const DDecCol &col = a.cols[fr.col];
    const uint8_t *s = col.src;
    DTile *tl = a.tiles + fr.tile_base;
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
        x.scale = sc;
        if (for_walk(s, pos + 2, end, n, tl, 0) == ~0ull) err = QA_ERR_CORRUPT;
        break;
    }
    case QA_M_ALP: {
        uint64_t m;
        if (!rd_varint(s, end, &pos, &m) || m != n) { err = QA_ERR_CORRUPT; break; }
        for (uint32_t t = 0; t < fr.ntiles; t++) {
            if (pos + 13 > end) { err = QA_ERR_CORRUPT; break; }
            uint32_t e = s[pos], ff = s[pos + 1], w = s[pos + 12];
            uint32_t exc = s[pos + 2] | ((uint32_t)s[pos + 3] << 8);
            if (e > 18 || ff > e || w > 32) { err = QA_ERR_CORRUPT; break; }
            tl[t].off = pos;
            tl[t].w = w;
            pos += 13 + 128ull * w;
            if (pos > end || 10ull * exc > end - pos) { err = QA_ERR_CORRUPT; break; }
            pos += 10ull * exc;
        }
        break;
    }
    case QA_M_ALP_RD: {
        if (pos + 2 > end) { err = QA_ERR_CORRUPT; break; }
        uint32_t left = s[pos], dl = s[pos + 1];
        if (left < 1 || left > 63 || dl > 8) { err = QA_ERR_CORRUPT; break; }
        pos += 2;
        if (pos + 8ull * dl > end) { err = QA_ERR_CORRUPT; break; }
        for (uint32_t q = 0; q < dl; q++) x.dict[q] = qa_ld64u(s + pos + 8 * q);
        pos += 8ull * dl;
        uint64_t nexc;
        if (!rd_varint(s, end, &pos, &nexc) || nexc > n) { err = QA_ERR_CORRUPT; break; }
        x.left = left;
        x.dl = dl;
        x.nexc = nexc;
        x.exc_off = pos;
        uint32_t sorted = 1;
        uint64_t last = 0;
        for (uint64_t j = 0; j < nexc && !err; j++) {
            uint64_t p;
            if (!rd_varint(s, end, &pos, &p) || pos + 8 > end || p >= n) { err = QA_ERR_CORRUPT; break; }
            a.scr[fr.scr + j] = p;
            a.scr2[fr.scr + j] = qa_ld64u(s + pos);
            pos += 8;
            if (j && p <= last) sorted = 0;
            last = p;
        }
        if (err) break;
        x.sorted = sorted;
        if (pos >= end) { err = QA_ERR_CORRUPT; break; }
        uint8_t ctag = s[pos++];
        uint64_t clen;
        if (!rd_varint(s, end, &pos, &clen) || clen > end - pos) { err = QA_ERR_CORRUPT; break; }
        uint64_t cend = pos + clen;
        if (cend >= end) { err = QA_ERR_CORRUPT; break; }
        uint8_t rtag = s[cend];
        if (ctag == 1 || rtag == 1) { err = QA_ERR_UNSUPPORTED; break; }
        if (ctag != 0 || rtag != 0) { err = QA_ERR_CORRUPT; break; }
        if (for_walk(s, pos, cend, n, tl, 0) == ~0ull ||
            for_walk(s, cend + 1, end, n, tl, 1) == ~0ull)
            err = QA_ERR_CORRUPT;
        break;
    }
    case QA_M_DICT: {
        uint64_t card, vlen;
        if (!rd_varint(s, end, &pos, &card) || card > n) { err = QA_ERR_CORRUPT; break; }
        if (pos >= end) { err = QA_ERR_CORRUPT; break; }
        uint8_t vtag = s[pos++];
        if (!rd_varint(s, end, &pos, &vlen) || vlen > end - pos) { err = QA_ERR_CORRUPT; break; }
        x.card = card;
        x.vtag = vtag;
        x.aux_off = pos;
        uint64_t vend = pos + vlen;
        if (vtag == 0) {
            if (vlen != card * 8) { err = QA_ERR_CORRUPT; break; }
        } else if (vtag == 1) {
            if (vlen < 8 || for_walk(s, pos + 8, vend, card, tl, 1) == ~0ull) {
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
        if (ctag != 0 || for_walk(s, vend + 1, end, n, tl, 0) == ~0ull) err = QA_ERR_CORRUPT;
        break;
    }
    case QA_M_RLE: {
        uint64_t R;
        if (!rd_varint(s, end, &pos, &R) || R > n) { err = QA_ERR_CORRUPT; break; }
        if ((end - pos) / 8 < R) { err = QA_ERR_CORRUPT; break; }
        x.card = R;
        x.aux_off = pos;
        if (for_walk(s, pos + 8 * R, end, R, tl, 1) == ~0ull) err = QA_ERR_CORRUPT;
        break;
    }
    default: break;
    }
    if (err) atomicOr(a.err, err);
}

// ---------------------------------------------------------------------------
// workgroup-wide scans (1024 threads)
// ---------------------------------------------------------------------------
template <typename T, typename Op>
__device__ __forceinline__ T wg_incl_scan(T v, Op op, T *tot) {
    __shared__ T wpart[16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
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
__device__ __forceinline__ uint64_t for_value_at(const uint8_t *h, uint32_t k) {
    uint32_t w = h[0];
    uint64_t mn = qa_ld64u(h + 1);
    if (w == 0) return mn;
    // direct unaligned extraction (sub-block offsets are validated)
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

__global__ void __launch_bounds__(1024) k_dec_frame_wg(DecArgs a) {
    const uint32_t f = blockIdx.x;
    if (*a.err || f >= a.nframes || !frame_used(a, f)) return;
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
        const bool nulls = a.dst[fr.col].nv != col.n;
        uint64_t acc = 0;
        for (uint64_t b0 = 0; b0 < n; b0 += blockDim.x) {
            uint64_t i = b0 + threadIdx.x;
            uint64_t v = i < n ? scr[i] : 0, tot;
            uint64_t inc = wg_incl_scan(v, OpXor64(), &tot);
            if (i < n) put_value(col, nulls, a.lane, fr.out_off + i, acc ^ inc);
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
    const DTile *tl = a.tiles + fr.tile_base;
    uint64_t acc = dict ? qa_ld64u(s + x.aux_off) : 0;  // delta base / first run start
    for (uint64_t b0 = 0; b0 < m; b0 += blockDim.x) {
        uint64_t i = b0 + threadIdx.x, v = 0;
        if (i < m) {
            uint64_t r = for_value_at(s + tl[i / QA_TILE].tail, (uint32_t)(i % QA_TILE));
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

// ---------------------------------------------------------------------------
__device__ __forceinline__ bool alp_value(int64_t d, uint32_t e, uint32_t f, uint64_t *out) {
    double v = __dmul_rn(__dmul_rn((double)d, d_exp10[f]), d_frac10[e]);
    *out = (uint64_t)__double_as_longlong(v);
    return true;
}

__global__ void __launch_bounds__(256) k_dec_tiles_fast(DecArgs a) {
    __shared__ uint32_t wbuf[4][2 * QA_TILE + 4];
    __shared__ uint64_t vals[4][QA_TILE];
    if (*a.err) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint64_t t = (uint64_t)blockIdx.x * 4 + wv;
    if (t >= a.ntiles) return;
    int f = frame_by_tile(a, t);
    if (!frame_used(a, f)) return;
    const DFrame &fr = a.frames[f];
    if (!fast_mode(fr.mode) || fr.mode == QA_M_XORZ) return;
    const DDecCol &col = a.cols[fr.col];
    const DFrameX &x = a.fx[f];
    const bool nulls = a.dst[fr.col].nv != col.n;
    const DTile &ts = a.tiles[t];
    const uint32_t lt = (uint32_t)(t - fr.tile_base);
    const uint32_t cnt = min((uint32_t)QA_TILE, fr.n - lt * QA_TILE);
    const uint64_t base = fr.out_off + (uint64_t)lt * QA_TILE, p0 = (uint64_t)lt * QA_TILE;
    const uint8_t *s = col.src;
    uint32_t *wd = wbuf[wv];
    uint64_t *vv = vals[wv];
    switch (fr.mode) {
    case QA_M_FLOAT_MULT: {
        const uint8_t *h = s + ts.off;
        uint64_t mn = qa_ld64u(h + 1);
        stage_words(h + 9, 128 * ts.w, wd, lane);
        wave_lds_sync();
        const double sc = d_fm_scales[x.scale];
        for (uint32_t k = lane; k < cnt; k += WAVE) {
            int64_t kk = (int64_t)(mn + unpack_lds(wd, ts.w, k));
            double q = __ddiv_rn((double)kk, sc);  // float_mult.rs:132
            put_value(col, nulls, a.lane, base + k, (uint64_t)__double_as_longlong(q));
        }
        break;
    }
    case QA_M_ALP: {
        const uint8_t *h = s + ts.off;
        uint32_t e = h[0], ff = h[1], exc = h[2] | ((uint32_t)h[3] << 8);
        int64_t mn = (int64_t)qa_ld64u(h + 4);
        stage_words(h + 13, 128 * ts.w, wd, lane);
        wave_lds_sync();
        for (uint32_t k = lane; k < cnt; k += WAVE) {
            uint64_t v;
            alp_value((int64_t)((uint64_t)mn + unpack_lds(wd, ts.w, k)), e, ff, &v);
            vv[k] = v;
        }
        wave_lds_sync();
        if (lane == 0) {  // exceptions patch in list order (alp.rs:224-244)
            const uint8_t *ep = h + 13 + 128ull * ts.w;
            for (uint32_t j = 0; j < exc; j++) {
                uint32_t pp = ep[10 * j] | ((uint32_t)ep[10 * j + 1] << 8);
                if (pp >= cnt) {
                    atomicOr(a.err, QA_ERR_CORRUPT);
                    break;
                }
                vv[pp] = qa_ld64u(ep + 10 * j + 2);
            }
        }
        wave_lds_sync();
        for (uint32_t k = lane; k < cnt; k += WAVE) put_value(col, nulls, a.lane, base + k, vv[k]);
        break;
    }
    case QA_M_ALP_RD: {
        const uint32_t rbw = 64 - x.left;
        const uint8_t *hc = s + ts.off, *hr = s + ts.tail;
        uint32_t wc = hc[0], wr = hr[0];
        uint64_t mc = qa_ld64u(hc + 1), mr = qa_ld64u(hr + 1);
        stage_words(hr + 9, 128 * wr, wd, lane);
        wave_lds_sync();
        bool badc = false;
        for (uint32_t k = lane; k < cnt; k += WAVE) {
            uint64_t code = wc ? for_value_at(hc, k) : mc;
            uint64_t right = mr + unpack_lds(wd, wr, k);
            if (code >= x.dl) {
                badc = true;
                code = 0;
            }
            vv[k] = x.dl ? ((x.dict[code] << (rbw & 63)) | right) : right;
        }
        if (wave_any_d(badc) && lane == 0) atomicOr(a.err, QA_ERR_CORRUPT);
        wave_lds_sync();
        // exceptions inside this tile: out = (left << right_bw) | right
        if (x.sorted) {
            uint64_t lo = 0, hi = x.nexc;
            while (lo < hi) {
                uint64_t mid = (lo + hi) >> 1;
                if (a.scr[fr.scr + mid] < p0) lo = mid + 1; else hi = mid;
            }
            for (uint64_t j = lo + lane; j < x.nexc; j += WAVE) {
                uint64_t pp = a.scr[fr.scr + j];
                if (pp >= p0 + cnt) break;
                uint64_t k = pp - p0;
                vv[k] = (a.scr2[fr.scr + j] << (rbw & 63)) | (mr + unpack_lds(wd, wr, (uint32_t)k));
            }
        } else if (lane == 0) {
            for (uint64_t j = 0; j < x.nexc; j++) {
                uint64_t pp = a.scr[fr.scr + j];
                if (pp < p0 || pp >= p0 + cnt) continue;
                uint64_t k = pp - p0;
                vv[k] = (a.scr2[fr.scr + j] << (rbw & 63)) | (mr + unpack_lds(wd, wr, (uint32_t)k));
            }
        }
        wave_lds_sync();
        for (uint32_t k = lane; k < cnt; k += WAVE) put_value(col, nulls, a.lane, base + k, vv[k]);
        break;
    }
    case QA_M_DICT: {
        const uint8_t *h = s + ts.off;
        uint64_t mn = qa_ld64u(h + 1);
        stage_words(h + 9, 128 * ts.w, wd, lane);
        wave_lds_sync();
        bool bad = false;
        for (uint32_t k = lane; k < cnt; k += WAVE) {
            uint64_t code = mn + unpack_lds(wd, ts.w, k);
            uint64_t v = 0;
            if (code >= x.card) bad = true; else v = a.scr[fr.scr + code];
            put_value(col, nulls, a.lane, base + k, v);
        }
        if (wave_any_d(bad) && lane == 0) atomicOr(a.err, QA_ERR_CORRUPT);
        break;
    }
    case QA_M_RLE: {
        const uint64_t R = x.card;
        for (uint32_t k = lane; k < cnt; k += WAVE) {
            uint64_t p = p0 + k, lo = 0, hi = R;  // last run with start <= p
            while (hi - lo > 1) {
                uint64_t mid = (lo + hi) >> 1;
                if (a.scr[fr.scr + mid] <= p) lo = mid; else hi = mid;
            }
            put_value(col, nulls, a.lane, base + k, qa_ld64u(s + x.aux_off + 8 * lo));
        }
        break;
    }
    default: break;
    }
}

static inline unsigned nbd(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

void qa_fast_dec_index_launch(const DecArgs &a, hipStream_t s) {
    if (!a.fast || !a.nframes) return;
    hipLaunchKernelGGL(k_dec_index_fast, dim3(nbd(a.nframes, 64)), dim3(64), 0, s, a);
    hipLaunchKernelGGL(k_dec_frame_wg, dim3(a.nframes), dim3(1024), 0, s, a);
}
void qa_fast_dec_tiles_launch(const DecArgs &a, hipStream_t s) {
    if (!a.fast || !a.ntiles) return;
    hipLaunchKernelGGL(k_dec_tiles_fast, dim3(nbd(a.ntiles, 4)), dim3(256), 0, s, a);
}
