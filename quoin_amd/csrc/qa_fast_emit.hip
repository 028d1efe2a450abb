// qa_fast_emit.hip — payload emission of the Level::Fast codecs.
//
//   k_emit_fast_frames  per block: the payload prefixes (tags, varints, the
//                       ALP-RD dictionary) of XORZ/FloatMult/ALP/ALP-RD/Dict/RLE
//   k_emit_fast_tiles   per 1024-value tile: XORZ varints (xorz.rs:8-16),
//                       FloatMult k sub-blocks (float_mult.rs:66-98), ALP
//                       sub-blocks (alp.rs:108-166), ALP-RD exceptions + codes +
//                       rights (alp_rd.rs:132-206), Dict codes (dict.rs:87-152)
//   k_emit_fast_chunks  per chunk of 1024 dictionary values / runs: Dict values
//                       (raw or delta-bitpacked) and RLE values + lengths
//                       (rle.rs:10-43)
#include "qa_dev.h"

__constant__ double e_exp10[19] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8, 1e9,
                                   1e10, 1e11, 1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18};
__constant__ double e_frac10[19] = {1e0,   1e-1,  1e-2,  1e-3,  1e-4,  1e-5,  1e-6,
                                    1e-7,  1e-8,  1e-9,  1e-10, 1e-11, 1e-12, 1e-13,
                                    1e-14, 1e-15, 1e-16, 1e-17, 1e-18};
__constant__ double e_fm_scales[7] = {10.0, 100.0, 1000.0, 10000.0, 100000.0, 1000000.0,
                                      10000000.0};

__device__ __forceinline__ uint64_t ef2b(double d) { return (uint64_t)__double_as_longlong(d); }
__device__ __forceinline__ double eb2f(uint64_t b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ int64_t esat(double x) {
    if (x >= 9223372036854775808.0) return INT64_MAX;
    if (x < -9223372036854775808.0) return INT64_MIN;
    return (int64_t)x;
}
__device__ __forceinline__ bool e_alp_digit(uint64_t bits, uint32_t e, uint32_t f, int64_t *d) {
    double v = eb2f(bits);
    double tmp = __dmul_rn(__dmul_rn(v, e_exp10[e]), e_frac10[f]);
    if (!isfinite(tmp) || tmp >= 9.223372036854776e18 || tmp <= -9.223372036854776e18) return false;
    int64_t di = esat(__dsub_rn(__dadd_rn(tmp, 6755399441055744.0), 6755399441055744.0));
    if (ef2b(__dmul_rn(__dmul_rn((double)di, e_exp10[f]), e_frac10[e])) != bits) return false;
    *d = di;
    return true;
}
__device__ __forceinline__ int64_t e_fm_k(uint64_t bits, int s) {
    return (int64_t)round(__dmul_rn(eb2f(bits), e_fm_scales[s]));
}
__device__ __forceinline__ void lds_sync_wave() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
}

__device__ __forceinline__ bool owner_of(const EncArgs &a, uint64_t t, int *cp, uint32_t *op,
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

// [w][min] then the packed residuals (res in LDS, padded with zeros)
__device__ __forceinline__ void emit_subblock(uint8_t *q, uint64_t mn, uint32_t w,
                                              const uint64_t *res, uint32_t *outw, int lane) {
    if (lane == 0) q[0] = (uint8_t)w;
    if (lane >= 1 && lane < 9) q[lane] = (uint8_t)(mn >> (8 * (lane - 1)));
    if (w == 0) return;
    pack_lds(res, w, outw, lane);
    lds_sync_wave();
    wave_store_bytes(q + 9, outw, 128 * w, lane);
}

// ---------------------------------------------------------------------------
__global__ void k_emit_fast_frames(EncArgs a) {
    uint32_t slot = a.r0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= a.r1) return;
    const DBlock &b = a.blocks[slot];
    if (b.len == 0 || a.cst[b.col].err) return;
    const DCol &col = a.cols[b.col];
    uint8_t *p = col.dst + b.frame_off + 9;
    const uint64_t n = b.len;
    switch (b.mode) {
    case QA_M_FLOAT_MULT:
        p[0] = 0;  // FM_BITPACK
        p[1] = (uint8_t)a.blocksx[slot].fm_scale;
        qa_put_varint(p + 2, n);
        break;
    case QA_M_ALP: qa_put_varint(p, n); break;
    case QA_M_ALP_RD: {
        const DBlockX &x = a.blocksx[slot];
        p[0] = (uint8_t)x.rd_left;
        p[1] = (uint8_t)x.rd_dl;
        for (uint32_t q = 0; q < x.rd_dl; q++) qa_st64u(p + 2 + 8 * q, x.rd_dict[q]);
        uint64_t o = 2 + 8ull * x.rd_dl;
        o += qa_put_varint(p + o, x.rd_nexc);
        o += x.rd_excb;
        p[o++] = 0;  // T_BITPACK codes
        o += qa_put_varint(p + o, x.rd_clen);
        qa_put_varint(p + o, n);
        o += x.rd_clen;
        p[o++] = 0;  // T_BITPACK rights
        qa_put_varint(p + o, n);
        break;
    }
    case QA_M_DICT: {
        const DBlockX &x = a.blocksx[slot];
        uint64_t o = qa_put_varint(p, x.card);
        p[o++] = (uint8_t)x.dict_vtag;
        o += qa_put_varint(p + o, x.dict_vlen);
        if (x.dict_vtag == 1) {
            const uint64_t *d = a.dsorted + col.lane_off + b.start;
            qa_st64u(p + o, d[0]);
            qa_put_varint(p + o + 8, x.card);
        }
        o += x.dict_vlen;
        p[o++] = 0;  // CODES_BITPACK
        qa_put_varint(p + o, n);
        break;
    }
    case QA_M_RLE: {
        const DBlockX &x = a.blocksx[slot];
        uint64_t o = qa_put_varint(p, x.nruns);
        qa_put_varint(p + o + 8 * x.nruns, x.nruns);
        break;
    }
    default: break;
    }
}

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_emit_fast_tiles(EncArgs a) {
    __shared__ uint64_t res[QA_TILE];
    __shared__ uint32_t outw[2 * QA_TILE + 4];
    __shared__ uint32_t bytes[(11 * QA_TILE + 16) / 4 + 4];
    const int lane = threadIdx.x;
    const uint64_t t = a.t0 + blockIdx.x;
    if (t >= a.t1) return;
    {  // the emit descriptor knows the block's mode: leave after one load
        // (k_emit_fl wrote the FloatMult tiles with w <= 32)
        const uint32_t bm = a.temit[t].bmode;
        if (bm != QA_M_XORZ && bm != QA_M_FLOAT_MULT && bm != QA_M_ALP && bm != QA_M_ALP_RD &&
            bm != QA_M_DICT)
            return;
        if (a.temit[t].fast) return;
    }
    const DTile &ts = a.tiles[t];
    if (!ts.active) return;
    int c;
    uint32_t o, bt;
    if (!owner_of(a, t, &c, &o, &bt)) return;
    const DBlock &b = a.blocks[o];
    const DCol &col = a.cols[c];
    const DColState &st = a.cst[c];
    if (st.err) return;
    const uint32_t mode = b.mode;
    if (mode != QA_M_XORZ && mode != QA_M_FLOAT_MULT && mode != QA_M_ALP &&
        mode != QA_M_ALP_RD && mode != QA_M_DICT)
        return;
    const DTileF &tf = a.tilesf[t];
    const uint64_t p0 = (uint64_t)bt * QA_TILE, k0 = b.start + p0;
    const uint32_t cnt = ts.cnt;
    uint8_t *p = col.dst + b.frame_off + 9;
    uint8_t *bb = (uint8_t *)bytes;

    if (mode == QA_M_XORZ) {
        uint32_t carry = 0;
        for (uint32_t kb = 0; kb < cnt; kb += WAVE) {
            uint32_t k = kb + lane;
            uint64_t x = 0;
            uint32_t len = 0;
            if (k < cnt) {
                uint64_t prev = p0 + k ? lane_at(st, k0 + k - 1) : 0;
                x = lane_at(st, k0 + k) ^ prev;
                len = qa_varint_len(x);
            }
            uint32_t inc = len;
#pragma unroll
            for (int q = 1; q < 64; q <<= 1) {
                uint32_t y = __shfl_up(inc, q, 64);
                if (lane >= q) inc += y;
            }
            if (k < cnt) qa_put_varint(bb + carry + inc - len, x);
            carry += __shfl(inc, 63, 64);
        }
        lds_sync_wave();
        wave_store_bytes(p + tf.x_off, bytes, carry, lane);
        return;
    }
    if (mode == QA_M_FLOAT_MULT) {
        const int s = (int)a.blocksx[o].fm_scale;
        const uint64_t mn = (uint64_t)tf.kmin[s];
        for (uint32_t k = lane; k < QA_TILE; k += WAVE)
            res[k] = k < cnt ? (uint64_t)e_fm_k(lane_at(st, k0 + k), s) - mn : 0;
        lds_sync_wave();
        emit_subblock(p + ts.off, mn, tf.fm_w, res, outw, lane);
        return;
    }
    if (mode == QA_M_ALP) {
        const uint32_t e = tf.alp_e, f = tf.alp_f, w = tf.alp_w;
        // filler = first valid digit (alp.rs:125-133)
        int64_t filler = 0;
        bool have = false;
        for (uint32_t kb = 0; kb < cnt && !have; kb += WAVE) {
            uint32_t k = kb + lane;
            int64_t d = 0;
            bool v = k < cnt && e_alp_digit(lane_at(st, k0 + k), e, f, &d);
            uint64_t bal = __ballot(v);
            if (bal) {
                int src = __ffsll((long long)bal) - 1;
                filler = __shfl(d, src, 64);
                have = true;
            }
        }
        uint8_t *q = p + tf.alp_off;
        uint32_t ecarry = 0;
        for (uint32_t kb = 0; kb < QA_TILE; kb += WAVE) {
            uint32_t k = kb + lane;
            int64_t d = filler;
            bool exc = false;
            uint64_t bits = 0;
            if (k < cnt) {
                bits = lane_at(st, k0 + k);
                if (!e_alp_digit(bits, e, f, &d)) {
                    d = filler;
                    exc = true;
                }
            }
            res[k] = (uint64_t)d - (uint64_t)tf.dmin;
            uint64_t bal = __ballot(exc);
            if (exc) {
                uint32_t idx = ecarry + (uint32_t)__popcll(bal & ((1ull << lane) - 1));
                uint8_t *r = bb + 10 * idx;
                r[0] = (uint8_t)k;
                r[1] = (uint8_t)(k >> 8);
                for (int z = 0; z < 8; z++) r[2 + z] = (uint8_t)(bits >> (8 * z));
            }
            ecarry += (uint32_t)__popcll(bal);
        }
        if (lane == 0) {
            q[0] = (uint8_t)e;
            q[1] = (uint8_t)f;
            q[2] = (uint8_t)tf.alp_exc;
            q[3] = (uint8_t)(tf.alp_exc >> 8);
        }
        if (lane >= 4 && lane < 12) q[lane] = (uint8_t)((uint64_t)tf.dmin >> (8 * (lane - 4)));
        if (lane == 12) q[12] = (uint8_t)w;
        lds_sync_wave();
        if (w) {
            pack_lds(res, w, outw, lane);
            lds_sync_wave();
            wave_store_bytes(q + 13, outw, 128 * w, lane);
        }
        wave_store_bytes(q + 13 + 128ull * w, bytes, 10 * ecarry, lane);
        return;
    }
    if (mode == QA_M_ALP_RD) {
        const DBlockX &x = a.blocksx[o];
        const uint32_t rbw = 64 - x.rd_left;
        const uint64_t mask = (1ull << rbw) - 1;
        // exceptions: varint(position) + u64 left part, in position order
        uint32_t carry = 0;
        for (uint32_t kb = 0; kb < QA_TILE; kb += WAVE) {
            uint32_t k = kb + lane;
            uint64_t code = 0, right = 0, left = 0;
            bool exc = false;
            uint32_t len = 0;
            if (k < cnt) {
                uint64_t v = lane_at(st, k0 + k);
                left = v >> rbw;
                right = v & mask;
                bool found = false;
                for (uint32_t q = 0; q < x.rd_dl; q++)
                    if (!found && x.rd_dict[q] == left) {
                        code = q;
                        found = true;
                    }
                exc = !found;
                if (exc) len = qa_varint_len(p0 + k) + 8;
            }
            res[k] = code;
            uint32_t inc = len;
#pragma unroll
            for (int q = 1; q < 64; q <<= 1) {
                uint32_t y = __shfl_up(inc, q, 64);
                if (lane >= q) inc += y;
            }
            if (exc) {
                uint8_t *r = bb + carry + inc - len;
                uint32_t vl = qa_put_varint(r, p0 + k);
                for (int z = 0; z < 8; z++) r[vl + z] = (uint8_t)(left >> (8 * z));
            }
            carry += __shfl(inc, 63, 64);
        }
        lds_sync_wave();
        wave_store_bytes(p + tf.rd_exc_off, bytes, carry, lane);
        // codes sub-block
        for (uint32_t k = lane; k < QA_TILE; k += WAVE) res[k] = k < cnt ? res[k] - tf.rd_cmin : 0;
        lds_sync_wave();
        emit_subblock(p + tf.rd_coff, tf.rd_cmin, tf.rd_wc, res, outw, lane);
        lds_sync_wave();
        // rights sub-block
        for (uint32_t k = lane; k < QA_TILE; k += WAVE)
            res[k] = k < cnt ? (lane_at(st, k0 + k) & mask) - tf.rd_rmin : 0;
        lds_sync_wave();
        emit_subblock(p + tf.rd_roff, tf.rd_rmin, tf.rd_wr, res, outw, lane);
        return;
    }
    // Dict codes: the rank of each value in the sorted dictionary
    {
        const DBlockX &x = a.blocksx[o];
        const uint64_t *d = a.dsorted + col.lane_off + b.start;
        const uint64_t card = x.card;
        for (uint32_t k = lane; k < QA_TILE; k += WAVE) {
            uint64_t r = 0;
            if (k < cnt) {
                uint64_t v = lane_at(st, k0 + k), lo = 0, hi = card;
                while (lo < hi) {
                    uint64_t mid = (lo + hi) >> 1;
                    if (d[mid] < v) lo = mid + 1; else hi = mid;
                }
                r = lo - tf.dict_min;
            }
            res[k] = r;
        }
        lds_sync_wave();
        emit_subblock(p + ts.off, tf.dict_min, tf.dict_w, res, outw, lane);
    }
}

// ---------------------------------------------------------------------------
// chunk c of a block uses the block's tile slot c for its layout
__global__ void __launch_bounds__(64) k_emit_fast_chunks(EncArgs a) {
    __shared__ uint64_t res[QA_TILE];
    __shared__ uint32_t outw[2 * QA_TILE + 4];
    const int lane = threadIdx.x;
    const uint64_t t = a.t0 + blockIdx.x;
    if (t >= a.t1) return;
    {
        const uint32_t bm = a.temit[t].bmode;
        if (bm != QA_M_DICT && bm != QA_M_RLE) return;
    }
    int c;
    uint32_t o, bt;
    if (!owner_of(a, t, &c, &o, &bt)) return;
    const DBlock &b = a.blocks[o];
    if (b.len == 0) return;
    const DCol &col = a.cols[c];
    const DColState &st = a.cst[c];
    if (st.err) return;
    if (b.mode != QA_M_DICT && b.mode != QA_M_RLE) return;
    const DBlockX &x = a.blocksx[o];
    const DTileF &tf = a.tilesf[t];
    uint8_t *p = col.dst + b.frame_off + 9;
    const uint64_t ch0 = (uint64_t)bt * QA_TILE;
    if (b.mode == QA_M_DICT) {
        if (ch0 >= x.card) return;
        const uint64_t *d = a.dsorted + col.lane_off + b.start;
        const uint32_t cnt = (uint32_t)min((uint64_t)QA_TILE, x.card - ch0);
        uint64_t v0 = qa_varint_len(x.card) + 1 + qa_varint_len(x.dict_vlen);
        if (x.dict_vtag == 0) {  // VAL_RAW: the sorted values verbatim
            for (uint32_t k = lane; k < cnt; k += WAVE) {
                uint64_t v = d[ch0 + k];
                outw[2 * k] = (uint32_t)v;
                outw[2 * k + 1] = (uint32_t)(v >> 32);
            }
            lds_sync_wave();
            wave_store_bytes(p + v0 + 8 * ch0, outw, 8 * cnt, lane);
            return;
        }
        for (uint32_t k = lane; k < QA_TILE; k += WAVE) {
            uint64_t i = ch0 + k;
            res[k] = k < cnt ? (i ? qa_zigzag(d[i] - d[i - 1]) : 0) - tf.dv_min : 0;
        }
        lds_sync_wave();
        emit_subblock(p + tf.dv_off, tf.dv_min, tf.dv_w, res, outw, lane);
        return;
    }
    // RLE
    if (ch0 >= x.nruns) return;
    const uint32_t *rs = a.rstart + col.lane_off + b.start;
    const uint32_t cnt = (uint32_t)min((uint64_t)QA_TILE, x.nruns - ch0);
    const uint64_t vo = qa_varint_len(x.nruns);
    for (uint32_t k = lane; k < cnt; k += WAVE) {
        uint64_t v = lane_at(st, b.start + rs[ch0 + k]);
        outw[2 * k] = (uint32_t)v;
        outw[2 * k + 1] = (uint32_t)(v >> 32);
    }
    lds_sync_wave();
    wave_store_bytes(p + vo + 8 * ch0, outw, 8 * cnt, lane);
    lds_sync_wave();
    for (uint32_t k = lane; k < QA_TILE; k += WAVE) {
        uint64_t i = ch0 + k;
        res[k] = k < cnt ? ((i + 1 < x.nruns ? rs[i + 1] : b.len) - rs[i]) - tf.ch_min : 0;
    }
    lds_sync_wave();
    emit_subblock(p + tf.ch_off, tf.ch_min, tf.ch_w, res, outw, lane);
}

static inline unsigned nbe(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

void qa_fast_emit_launch(const EncArgs &a, hipStream_t s) {
    if (!a.fast) return;
    if (a.r1 > a.r0)
        qa_launch(k_emit_fast_frames, dim3(nbe(a.r1 - a.r0, 64)), dim3(64), 0, s, a);
    if (a.t1 > a.t0) {
        qa_launch(k_emit_fast_tiles, dim3((unsigned)(a.t1 - a.t0)), dim3(64), 0, s, a);
        qa_launch(k_emit_fast_chunks, dim3((unsigned)(a.t1 - a.t0)), dim3(64), 0, s, a);
    }
}

// ---------------------------------------------------------------------------
// Levels Balanced / High / Max: the lane words the host competition reads,
// and the frames whose payloads it wrote
// ---------------------------------------------------------------------------
__device__ __forceinline__ int col_by_lane(const DCol *c, int nc, uint64_t j) {
    QA_COL_SEARCH(lane_off, j)
}

// out[lane_off + i] = lane word i of its column for i < nv (grid-stride over
// the padded lane range; the padding is left as it is)
__global__ void __launch_bounds__(256) k_lane_words(EncArgs a, uint64_t total, uint64_t *out) {
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < total;
         j += (uint64_t)gridDim.x * blockDim.x) {
        const int c = col_by_lane(a.cols, (int)a.ncols, j);
        const uint64_t i = j - a.cols[c].lane_off;
        const DColState &st = a.cst[c];
        if (i < st.nv) out[j] = lane_at(st, i);
    }
}

void qa_lane_words_launch(const EncArgs &a, uint64_t total, uint64_t *out, hipStream_t s) {
    if (!total || !a.ncols) return;
    const uint64_t g = nbe(total, 256);
    qa_launch(k_lane_words, dim3((unsigned)(g < 8192 ? g : 8192)), dim3(256), 0, s, a,
                       total, out);
}

// one workgroup per host-written frame: its mode byte over the placeholder
// k_emit_frames wrote, then the payload (frame_off comes from the layout)
__global__ void __launch_bounds__(256) k_host_frames(EncArgs a, const HScatter *sc,
                                                     const uint8_t *src) {
    const HScatter h = sc[blockIdx.x];
    const DBlock &b = a.blocks[h.slot];
    if (a.cst[b.col].err) return;
    uint8_t *d = a.cols[b.col].dst + b.frame_off;
    if (threadIdx.x == 0) d[0] = (uint8_t)h.mode;
    const uint8_t *p = src + h.src_off;
    for (uint64_t i = threadIdx.x; i < h.len; i += blockDim.x) d[9 + i] = p[i];
}

void qa_host_frames_launch(const EncArgs &a, const HScatter *sc, uint32_t ns, const uint8_t *src,
                           hipStream_t s) {
    if (ns) qa_launch(k_host_frames, dim3(ns), dim3(256), 0, s, a, sc, src);
}
