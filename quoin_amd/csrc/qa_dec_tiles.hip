// qa_dec_tiles.hip — per-tile decoders of every GPU mode, one wavefront per
// 1024-value tile, writing straight into the typed output.
//
// Each lane holds the tile's values k = lane + 64 q (q < 16) in registers:
//   RAW        (raw.rs:30-49)             staged through LDS, f32 widened
//   CONST      (const_block.rs:13-20)
//   STRIDE     (stride.rs:22-35)          v0 + i * step, wrapping
//   FOR        (for_bitpack.rs:108-157)   FastLanes unpack from LDS + min
//   DELTA      (delta_bitpack.rs:48-64)   unpack, unzigzag, wavefront scan
//                                         seeded with the tile carry
//   FLOAT_MULT (float_mult.rs:100-133)    k / scale in IEEE f64
//   ALP        (alp.rs:168-248)           (d * 10^f) * 10^-e in IEEE f64,
//                                         exceptions patched in list order
//   ALP_RD     (alp_rd.rs:208-285)        (dict[code] << right_bw) | right,
//                                         exceptions patched in list order
//   DICT       (dict.rs:154-207)          code -> sorted dictionary value
//   RLE        (rle.rs:45-75)             run search over the run starts
//   XORZ       (xorz.rs:18-31)            values prefix-XORed per frame by
//                                         k_dec_frame_wg
// Streams without nulls store the values at their positions. Streams with
// nulls put them in LDS and the wavefront walks the bitmap words of the tile's
// logical range (src/validity.rs:130-146).
#include "qa_dec.h"

typedef uint64_t dec_u64x2 __attribute__((ext_vector_type(2)));
typedef uint32_t dec_u32x4 __attribute__((ext_vector_type(4)));

__constant__ double d_exp10[19] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8, 1e9,
                                   1e10, 1e11, 1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18};
__constant__ double d_frac10[19] = {1e0,   1e-1,  1e-2,  1e-3,  1e-4,  1e-5,  1e-6,
                                    1e-7,  1e-8,  1e-9,  1e-10, 1e-11, 1e-12, 1e-13,
                                    1e-14, 1e-15, 1e-16, 1e-17, 1e-18};
__constant__ double d_fm_scales[7] = {10.0, 100.0, 1000.0, 10000.0, 100000.0, 1000000.0,
                                      10000000.0};

#define QV (QA_TILE / WAVE)


// Patch value v at tile index k (uniform across the wavefront) into the lane
// that holds it. Static register indexing keeps x[] in registers.
__device__ __forceinline__ void patch_reg(uint64_t (&x)[QV], uint32_t k, uint64_t v, int lane) {
    const uint32_t q = k >> 6;
    const bool hit = (int)(k & 63) == lane;
#pragma unroll
    for (int i = 0; i < QV; i++) {
        const bool sel = hit & ((uint32_t)i == q);
        x[i] = sel ? v : x[i];  // a select, not an indexed store
    }
}

// the tile that wavefront index g of a tile decoder's launch takes: the
// launch range [t0, t1), or the g-th tile of the launch's run table (g is
// uniform, so the table is read with scalar loads)
__device__ __forceinline__ bool dec_tile_at(const DecArgs &a, uint64_t g, uint64_t *t) {
    if (!a.krun) {
        *t = a.t0 + g;
        return *t < a.t1;
    }
    if (g >= a.ktotal) return false;
    uint32_t lo = 0, hi = a.nkrun;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.krun[mid].pre <= g) lo = mid; else hi = mid;
    }
    *t = a.krun[lo].t0 + (g - a.krun[lo].pre);
    return true;
}

// one thread per tile: its decode descriptor
__device__ __forceinline__ void dec_desc_one(const DecArgs &a, uint64_t t) {
    if (t >= a.t1) return;
    TileDec d{};
    const DTileD &ts = a.tiles[t];
    const uint32_t f = ts.frame;
    const DFrame &fr = a.frames[f];
    const DDecCol &col = a.cols[fr.col];
    const DDecState &st = a.dst[fr.col];
    if (f < st.fused_end && !*a.err) {
        const uint32_t lt = (uint32_t)(t - fr.tile_base);
        d.src = col.src;
        d.out = col.out;
        d.bitmap = col.bitmap;
        d.off = ts.off;
        d.tail = ts.tail;  // DeltaBitpack: the carry (k_dec_delta_sums, k_dec_delta_scan)
        // FloatMult: the divisor's bits, so the tile decoders read no frame state
        if (fr.mode == QA_M_FLOAT_MULT && a.fx)
            d.tail = (uint64_t)__double_as_longlong(d_fm_scales[a.fx[f].scale]);
        d.payload = fr.payload_off;
        d.p0 = (uint64_t)lt * QA_TILE;
        d.base = fr.out_off + d.p0;
        d.frame = f;
        d.cnt = min((uint32_t)QA_TILE, fr.n - lt * QA_TILE);
        d.w = ts.w;
        d.mode = fr.mode;
        d.dtype = col.dtype;
        d.oform = (uint32_t)out_form(col);
        d.nulls = st.nv != col.n;
        d.fl = tile_kind(d.mode, d.oform);
        if (d.nulls) {
            // logical positions [this tile's start, the next tile's, or n):
            // the starts of its first value and of the first value after it
            d.L0 = tile_lpos(a, col, d.base);
            d.L1 = d.base + d.cnt < st.nv ? tile_lpos(a, col, d.base + d.cnt) : col.n;
        }
    }
    a.tdec[t] = d;
}

__global__ void k_dec_desc(DecArgs a) {
    dec_desc_one(a, a.t0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// The descriptors, the zeros of all-null streams and the validity_out of
// streams without a bitmap in one launch: workgroups [0, nd) take 256 tiles,
// [nd, nd + nn) 256 bitmap words, the rest 256 validity_out bytes.
__global__ void k_dec_desc_fill(DecArgs a, uint32_t nd, uint32_t nn) {
    const uint32_t b = blockIdx.x;
    if (b < nd) return dec_desc_one(a, a.t0 + (uint64_t)b * 256 + threadIdx.x);
    if (b < nd + nn) return dec_nullfill_one(a, (uint64_t)(b - nd) * 256 + threadIdx.x);
    dec_allvalid_one(a, (uint64_t)(b - nd - nn) * 256 + threadIdx.x);
}

__global__ void __launch_bounds__(256) k_dec_tiles(DecArgs a) {
    __shared__ __align__(16) uint64_t lds[4][QA_TILE];
    // (an error flagged before k_dec_desc left every descriptor empty; no
    // kernel between it and this one flags errors)
    const int lane = threadIdx.x & 63, wv = qa_wave_id();
    uint64_t t;
    if (!dec_tile_at(a, (uint64_t)blockIdx.x * 4 + wv, &t)) return;
    const TileDec td = a.tdec[t];
    if (!td.mode || td.fl) return;
    const uint32_t f = td.frame;
    const bool nulls = td.nulls;
    const uint32_t cnt = td.cnt;
    const uint64_t p0 = td.p0;
    const uint8_t *s = td.src;
    uint32_t *wd = (uint32_t *)lds[wv];
    uint64_t *vv = lds[wv];
    uint64_t x[QV];
    bool bad = false;
    struct {
        uint64_t off, tail, w;
    } ts{td.off, td.tail, td.w};

    switch (td.mode) {
    case QA_M_RAW: {
        const uint32_t lb = (uint32_t)lane_bytes_d(td.dtype);
        stage_words(s + td.payload + p0 * lb, cnt * lb, wd, lane);
        wave_lds_sync();
#pragma unroll
        for (int q = 0; q < QV; q++) {
            uint32_t k = lane + WAVE * q;
            uint64_t v = 0;
            if (k < cnt) {
                if (lb == 4) {
                    v = wd[k];
                    if (td.dtype == QA_F32) v = qa_f32_to_lane((uint32_t)v);
                } else {
                    v = ((uint64_t)wd[2 * k + 1] << 32) | wd[2 * k];
                }
            }
            x[q] = v;
        }
        break;
    }
    case QA_M_CONST: {
        const uint64_t v = qa_ld64u(s + td.payload);
#pragma unroll
        for (int q = 0; q < QV; q++) x[q] = v;
        break;
    }
    case QA_M_STRIDE: {
        const uint64_t v0 = qa_ld64u(s + td.payload), step = qa_ld64u(s + td.payload + 8);
#pragma unroll
        for (int q = 0; q < QV; q++) x[q] = v0 + (p0 + lane + WAVE * q) * step;
        break;
    }
    case QA_M_FOR:
    case QA_M_DELTA:
    case QA_M_FLOAT_MULT:
    case QA_M_DICT: {
        const uint8_t *h = s + ts.off;
        const uint32_t w = ts.w;
        const uint64_t mn = qa_ld64u(h + 1);
        stage_words(h + 9, 128 * w, wd, lane);
        wave_lds_sync();
#pragma unroll
        for (int q = 0; q < QV; q++) {
            uint32_t k = lane + WAVE * q;
            x[q] = k < cnt ? mn + unpack_lds(wd, w, k) : 0;
        }
        if (td.mode == QA_M_DELTA) {
            // wrapping prefix sum of the unzigzagged deltas, seeded with the
            // tile carry (a wavefront scan per 64 values)
            uint64_t carry = ts.tail;
#pragma unroll
            for (int q = 0; q < QV; q++) {
                const uint32_t k = lane + WAVE * q;
                const uint64_t inc = qa_wave_scan_add64(k < cnt ? qa_unzigzag(x[q]) : 0);
                x[q] = carry + inc;
                carry += qa_readlane64(inc, 63);
            }
        } else if (td.mode == QA_M_FLOAT_MULT) {
            const double sc = __longlong_as_double((long long)td.tail);  // k_dec_desc
#pragma unroll
            for (int q = 0; q < QV; q++)
                x[q] = (uint64_t)__double_as_longlong(__ddiv_rn((double)(int64_t)x[q], sc));
        } else if (td.mode == QA_M_DICT) {
            const uint64_t card = a.fx[f].card;
            const uint64_t *dv = a.scr + a.frames[f].scr;
#pragma unroll
            for (int q = 0; q < QV; q++) {
                uint32_t k = lane + WAVE * q;
                uint64_t code = x[q], v = 0;
                if (k < cnt) {
                    if (code >= card) bad = true; else v = dv[code];
                }
                x[q] = v;
            }
        }
        break;
    }
    case QA_M_ALP: {
        const uint8_t *h = s + ts.off;
        const uint32_t w = ts.w, e = h[0], ff = h[1], exc = h[2] | ((uint32_t)h[3] << 8);
        const uint64_t mn = qa_ld64u(h + 4);
        stage_words(h + 13, 128 * w, wd, lane);
        wave_lds_sync();
        const double m1 = d_exp10[ff], m2 = d_frac10[e];
#pragma unroll
        for (int q = 0; q < QV; q++) {
            uint32_t k = lane + WAVE * q;
            int64_t d = k < cnt ? (int64_t)(mn + unpack_lds(wd, w, k)) : 0;
            x[q] = (uint64_t)__double_as_longlong(__dmul_rn(__dmul_rn((double)d, m1), m2));
        }
        // exceptions in list order (alp.rs:224-244): chunks of 64 read in
        // parallel, applied one by one so a repeated position keeps the last
        const uint8_t *ep = h + 13 + 128ull * w;
        for (uint32_t j0 = 0; j0 < exc; j0 += WAVE) {
            uint32_t j = j0 + lane, pp = 0;
            uint64_t v = 0;
            if (j < exc) {
                pp = ep[10 * j] | ((uint32_t)ep[10 * j + 1] << 8);
                v = qa_ld64u(ep + 10 * j + 2);
            }
            uint32_t m = min((uint32_t)WAVE, exc - j0);
            for (uint32_t i = 0; i < m; i++) {
                uint32_t pk = (uint32_t)__shfl((int)pp, (int)i, 64);
                uint64_t pv = __shfl(v, (int)i, 64);
                if (pk >= cnt) {
                    bad = true;
                    break;
                }
                patch_reg(x, pk, pv, lane);
            }
            if (__ballot(bad)) break;
        }
        break;
    }
    case QA_M_ALP_RD: {
        const DFrameX &fx = a.fx[f];
        const uint32_t rbw = (64 - fx.left) & 63, dl = fx.dl;
        const uint8_t *hc = s + ts.off, *hr = s + ts.tail;
        const uint32_t wc = hc[0], wr = hr[0];
        const uint64_t mc = qa_ld64u(hc + 1), mr = qa_ld64u(hr + 1);
        stage_words(hr + 9, 128 * wr, wd, lane);
        wave_lds_sync();
#pragma unroll
        for (int q = 0; q < QV; q++) {
            uint32_t k = lane + WAVE * q;
            uint64_t v = 0;
            if (k < cnt) {
                uint64_t code = wc ? for_value_direct(hc, k) : mc;
                uint64_t right = mr + unpack_lds(wd, wr, k);
                if (code >= dl) {
                    bad = true;
                    code = 0;
                }
                v = dl ? ((fx.dict[code] << rbw) | right) : right;
            }
            x[q] = v;
        }
        // exceptions: out = (left << right_bw) | right, in list order
        const uint64_t fscr = a.frames[f].scr;
        const uint64_t *epos = a.scr + fscr, *eleft = a.scr2 + fscr;
        uint64_t jb = 0, je = fx.nexc;
        if (fx.sorted) {  // only this tile's slice
            uint64_t lo = 0, hi = je;
            while (lo < hi) {
                uint64_t mid = (lo + hi) >> 1;
                if (epos[mid] < p0) lo = mid + 1; else hi = mid;
            }
            jb = lo;
            lo = jb, hi = je;
            while (lo < hi) {
                uint64_t mid = (lo + hi) >> 1;
                if (epos[mid] < p0 + cnt) lo = mid + 1; else hi = mid;
            }
            je = lo;
        }
        for (uint64_t j0 = jb; j0 < je; j0 += WAVE) {
            uint64_t j = j0 + lane, pp = ~0ull, v = 0;
            if (j < je) {
                pp = epos[j];
                if (pp >= p0 && pp < p0 + cnt)
                    v = (eleft[j] << rbw) | (mr + unpack_lds(wd, wr, (uint32_t)(pp - p0)));
            }
            uint32_t m = (uint32_t)min((uint64_t)WAVE, je - j0);
            for (uint32_t i = 0; i < m; i++) {
                uint64_t pk = __shfl(pp, (int)i, 64);
                uint64_t pv = __shfl(v, (int)i, 64);
                if (pk >= p0 && pk < p0 + cnt) patch_reg(x, (uint32_t)(pk - p0), pv, lane);
            }
        }
        break;
    }
    case QA_M_RLE: {
        // run starts (prefix of the run lengths, non-decreasing, covering the
        // frame exactly: checked by k_dec_frame_wg) -> the value at position p
        // is the last run with start <= p (rle.rs:45-75).
        const uint64_t R = a.fx[f].card, aux = a.fx[f].aux_off;
        const uint64_t *rs = a.scr + a.frames[f].scr;
        // r0 = the last run with start <= p0 (the run covering p0, recorded
        // while the starts were summed; zero-length runs cover nothing)
        const uint64_t r0 = a.tiles[t].run0;
        // each later run starting inside the tile marks its first position
        // with its index - r0; a prefix maximum then gives every position its
        // run. Marks sit at lds_pos<4> and values at lds_pos<8> (16-byte chunks
        // swizzled per lane), so the lane-contiguous reads and writes below hit
        // distinct banks.
        uint32_t *mark = wd;
#pragma unroll
        for (int q = 0; q < QV; q++) mark[lane + WAVE * q] = 0;
        wave_lds_sync();
        for (uint64_t c0 = r0 + 1; c0 < R; c0 += WAVE) {
            const uint64_t j = c0 + lane;
            const uint64_t st = j < R ? rs[j] : ~0ull;
            const bool in = st < p0 + cnt;  // and st > p0: r0 is the last start <= p0
            if (in) atomicMax(&mark[lds_pos<4>((uint32_t)(st - p0))], (uint32_t)(j - r0));
            if (!(__ballot(in) >> 63)) break;
        }
        wave_lds_sync();
        uint32_t run[16], m = 0;
        {
            const uint32_t swz = ((uint32_t)lane >> 1) & 3;  // lds_pos<4> of 16 lane + 4 c
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const uint4 q4 = ((const uint4 *)(mark + 16 * lane))[c ^ swz];
                const uint32_t e[4] = {q4.x, q4.y, q4.z, q4.w};
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    m = max(m, e[u]);
                    run[4 * c + u] = m;
                }
            }
        }
        const uint32_t incl = qa_wave_scan_max32(m);
        const uint32_t before = qa_dpp32<QA_DPP_WAVE_SHR1>(0, incl);
        // the tile's runs are r0 .. r0 + nr - 1: their values are staged in LDS
        // by coalesced loads (zero-length runs can make more than fit: then
        // each value is read directly)
        const uint32_t nr = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63) + 1;
        const bool staged = R && nr <= QA_TILE && r0 + nr <= R;
        uint64_t y[16];
        wave_lds_sync();  // every lane has read its marks
        if (staged) {
            stage_words(s + aux + 8 * r0, 8 * nr, wd, lane);
            wave_lds_sync();
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const uint32_t k = 16 * lane + i;
                y[i] = k < cnt ? vv[max(run[i], before)] : 0;
            }
            wave_lds_sync();  // every lane has read the staged values
        } else {
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const uint32_t k = 16 * lane + i;
                y[i] = k < cnt && R ? qa_ld64u(s + aux + 8 * (r0 + max(run[i], before))) : 0;
            }
        }
        {
            const uint32_t swz = (uint32_t)lane & 7;  // lds_pos<8> of 16 lane + 2 c
#pragma unroll
            for (int c = 0; c < 8; c++)
                ((uint4 *)(vv + 16 * lane))[c ^ swz] =
                    make_uint4((uint32_t)y[2 * c], (uint32_t)(y[2 * c] >> 32), (uint32_t)y[2 * c + 1],
                               (uint32_t)(y[2 * c + 1] >> 32));
        }
        wave_lds_sync();
#pragma unroll
        for (int q = 0; q < QV; q++) x[q] = vv[lds_pos<8>(lane + WAVE * q)];
        break;
    }
    case QA_M_XORZ: {
        const uint64_t *xv = a.scr + a.frames[f].scr + p0;
#pragma unroll
        for (int q = 0; q < QV; q++) {
            uint32_t k = lane + WAVE * q;
            x[q] = k < cnt ? xv[k] : 0;
        }
        break;
    }
    default: return;
    }
    if (__ballot(bad)) {
        if (lane == 0) atomicOr(a.err, QA_ERR_CORRUPT);
        return;
    }

    if (!nulls) {
        emit_dense(td.out, td.oform, td.base, cnt, x, lane);
        return;
    }
    wave_lds_sync();  // every lane is done with the staged words
#pragma unroll
    for (int q = 0; q < QV; q++) {
        uint32_t k = lane + WAVE * q;
        if (k < cnt) vv[k] = x[q];
    }
    wave_lds_sync();
    emit_gather(td.out, td.oform, td.bitmap, td.L0, td.L1, vv, cnt, lane);
}

static inline unsigned nbt(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }


// ---------------------------------------------------------------------------
// FoR / DeltaBitpack tiles into 4- or 8-byte integers, the bulk of integer
// columns (tile_kind in qa_kernels.h: k_dec_fl<4> takes FoR into 4-byte
// integers, k_dec_fl<8> FoR and Delta into 8-byte integers and Delta into
// 4-byte ones). Values are unpacked lane-contiguously (k = 16 lane + i: runs of
// staged row-words and a funnel shift each), Delta is an in-lane prefix plus
// one wavefront scan, and the values go back through LDS to k = lane + 64 q so
// every store writes whole cache lines. Into 4-byte integers only the low 32
// bits of a FoR value matter, so k_dec_fl<4> computes in 32 bits with 4.25 KiB
// of LDS per wavefront: up to 32 wavefronts per CU keep more tiles' loads in
// flight. (FoR of a 4-byte column has w <= 32; a stream with a wider sub-block
// still decodes, by direct reads.)
// ---------------------------------------------------------------------------
template <int OW, int WPB>
__global__ void __launch_bounds__(64 * WPB) k_dec_fl(DecArgs a) {
    // staged words (32 w, plus a row-word read past the last one) and the
    // transposed values; WPB wavefronts per workgroup (the autotuned launch shape)
    constexpr int LW = OW == 4 ? 1088 : 2 * QA_TILE;
    __shared__ __align__(16) uint32_t lds[WPB][LW];
    const int lane = threadIdx.x & 63, wv = qa_wave_id();
    uint64_t t;
    if (!dec_tile_at(a, (uint64_t)blockIdx.x * WPB + wv, &t)) return;
    const TileDec td = a.tdec[t];
    if (td.fl != OW) return;
    const uint32_t cnt = td.cnt, w = td.w;
    uint32_t *wd = lds[wv];
    // streams with nulls: the first bitmap window's words, loaded with the stream
    const uint64_t bw0 = td.nulls ? gather_first_word(td.bitmap, td.L0, td.L1, lane) : 0;
    const uint8_t *h = td.src + td.off;
    const uint64_t mn = qa_ld64u(h + 1);
    const uint32_t k0 = 16u * (uint32_t)lane;
    if (OW == 4) {
        uint32_t y[16];
        if (w <= 32) {
            stage_words_fl<false>(h + 9, 128 * w, wd, lane);
            wave_lds_sync();
            uint32_t v[16];
            unpack16_u32_fl(wd, w, lane, v);
#pragma unroll
            for (int i = 0; i < 16; i++) y[i] = (uint32_t)mn + v[i];
        } else {
#pragma unroll
            for (int i = 0; i < 16; i++) y[i] = k0 + i < cnt ? (uint32_t)for_value_direct(h, k0 + i) : 0u;
        }
        wave_lds_sync();  // every lane is done with the staged words
        const uint32_t swz = ((uint32_t)lane >> 1) & 3;  // chunk slot swizzle (lds_pos<4>)
#pragma unroll
        for (int c = 0; c < 4; c++)
            ((uint4 *)(wd + k0))[c ^ swz] = make_uint4(y[4 * c], y[4 * c + 1], y[4 * c + 2], y[4 * c + 3]);
    } else {
        // FoR-family sub-blocks: [w][min u64][128 w B]; ALP: [e][f][exc u16][min][w][128 w B]
        const bool alp = td.mode == QA_M_ALP;
        const uint64_t base = alp ? qa_ld64u(h + 4) : mn;
        if (w <= 32) stage_words_fl<false>(h + (alp ? 13 : 9), 128 * w, wd, lane);
        else stage_words_fl<true>(h + (alp ? 13 : 9), 128 * w, wd, lane);
        wave_lds_sync();
        uint64_t y[16];
        if (w <= 32) {
            uint32_t v[16];
            unpack16_u32_fl(wd, w, lane, v);
#pragma unroll
            for (int i = 0; i < 16; i++) y[i] = base + v[i];
        } else {
            uint64_t v[16];
            unpack16_u64_fl(wd, w, lane, v);
#pragma unroll
            for (int i = 0; i < 16; i++) y[i] = base + v[i];
        }
        if (td.mode == QA_M_DELTA) {
            // wrapping prefix sum of the unzigzagged deltas: in-lane over the 16
            // values, then one wavefront scan of the lane totals, seeded with
            // the tile carry (delta_bitpack.rs:48-64)
            uint64_t run = 0;
#pragma unroll
            for (int i = 0; i < 16; i++) {
                run += k0 + i < cnt ? qa_unzigzag(y[i]) : 0;
                y[i] = run;
            }
            const uint64_t inc = qa_wave_scan_add64(run);
            const uint64_t off = td.tail + inc - run;
#pragma unroll
            for (int i = 0; i < 16; i++) y[i] += off;
        } else if (td.mode == QA_M_FLOAT_MULT) {
            // k / scale in IEEE f64 (float_mult.rs:100-133)
            const double sc = __longlong_as_double((long long)td.tail);  // k_dec_desc
#pragma unroll
            for (int i = 0; i < 16; i++)
                y[i] = (uint64_t)__double_as_longlong(__ddiv_rn((double)(int64_t)y[i], sc));
        } else if (alp) {
            // (d * 10^f) * 10^-e in IEEE f64, two roundings (alp.rs:168-248)
            const double m1 = d_exp10[h[1]], m2 = d_frac10[h[0]];
#pragma unroll
            for (int i = 0; i < 16; i++)
                y[i] = (uint64_t)__double_as_longlong(__dmul_rn(__dmul_rn((double)(int64_t)y[i], m1), m2));
        }
        wave_lds_sync();  // every lane is done with the staged words
        const uint32_t swz = (uint32_t)lane & 7;  // chunk slot swizzle (lds_pos<8>)
#pragma unroll
        for (int c = 0; c < 8; c++)
            ((uint4 *)(wd + 2 * k0))[c ^ swz] = make_uint4((uint32_t)y[2 * c], (uint32_t)(y[2 * c] >> 32),
                                                           (uint32_t)y[2 * c + 1], (uint32_t)(y[2 * c + 1] >> 32));
        if (alp) {
            // exceptions in list order (alp.rs:224-244), validated sizes: chunks
            // of 64 read in parallel, written one by one so a repeated position
            // keeps the last
            const uint32_t exc = h[2] | ((uint32_t)h[3] << 8);
            const uint8_t *ep = h + 13 + 128ull * w;
            uint64_t *vv = (uint64_t *)wd;
            bool bad = false;
            if (exc) wave_lds_sync();
            for (uint32_t j0 = 0; j0 < exc; j0 += WAVE) {
                const uint32_t j = j0 + lane;
                uint32_t pp = 0;
                uint64_t v = 0;
                if (j < exc) {
                    pp = ep[10 * j] | ((uint32_t)ep[10 * j + 1] << 8);
                    v = qa_ld64u(ep + 10 * j + 2);
                    bad |= pp >= cnt;
                }
                if (__ballot(bad)) break;
                const uint32_t m = min((uint32_t)WAVE, exc - j0);
                for (uint32_t i = 0; i < m; i++)
                    if ((uint32_t)lane == i) vv[lds_pos<8>(pp)] = v;
            }
            if (__ballot(bad)) {
                if (lane == 0) atomicOr(a.err, QA_ERR_CORRUPT);
                return;
            }
        }
    }
    wave_lds_sync();
    const bool o32 = OW == 4 || td.oform == QA_OUT_U32;
    if (!td.nulls) {
        // full tiles: 16 bytes a lane per store, consecutive lanes on
        // consecutive 16 bytes (a store covers 1 KiB)
        const bool full = cnt == QA_TILE && ((uintptr_t)((const uint8_t *)td.out +
                                                         td.base * (o32 ? 4 : 8)) & 15) == 0;
        if (OW == 4 && full) {
            dec_u32x4 *o = (dec_u32x4 *)((uint32_t *)td.out + td.base);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t k = 4u * (uint32_t)lane + 256u * q;
                __builtin_nontemporal_store(*(const dec_u32x4 *)(wd + lds_pos<4>(k)), o + (k >> 2));
            }
        } else if (OW == 8 && !o32 && full) {
            dec_u64x2 *o = (dec_u64x2 *)((uint64_t *)td.out + td.base);
            const uint64_t *vv = (const uint64_t *)wd;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const uint32_t k = 2u * (uint32_t)lane + 128u * q;
                __builtin_nontemporal_store(*(const dec_u64x2 *)(vv + lds_pos<8>(k)), o + (k >> 1));
            }
        } else if (OW == 4) {
            uint32_t *o = (uint32_t *)td.out + td.base + lane;
#pragma unroll
            for (int q = 0; q < QV; q++)
                if (lane + WAVE * q < cnt)
                    __builtin_nontemporal_store(wd[lds_pos<4>(lane + WAVE * q)], o + WAVE * q);
        } else if (o32) {
            uint32_t *o = (uint32_t *)td.out + td.base + lane;
            const uint64_t *vv = (const uint64_t *)wd;
#pragma unroll
            for (int q = 0; q < QV; q++)
                if (lane + WAVE * q < cnt)
                    __builtin_nontemporal_store((uint32_t)vv[lds_pos<8>(lane + WAVE * q)], o + WAVE * q);
        } else {
            uint64_t *o = (uint64_t *)td.out + td.base + lane;
            const uint64_t *vv = (const uint64_t *)wd;
#pragma unroll
            for (int q = 0; q < QV; q++)
                if (lane + WAVE * q < cnt)
                    __builtin_nontemporal_store(vv[lds_pos<8>(lane + WAVE * q)], o + WAVE * q);
        }
        return;
    }
    if (OW == 4)
        emit_gather_t<QA_OUT_U32, uint32_t, 4>(td.out, td.bitmap, td.L0, td.L1, wd, cnt, lane, true, bw0);
    else if (o32)
        emit_gather_t<QA_OUT_U32, uint64_t, 8>(td.out, td.bitmap, td.L0, td.L1, (const uint64_t *)wd, cnt,
                                               lane, true, bw0);
    else
        emit_gather_t<QA_OUT_U64, uint64_t, 8>(td.out, td.bitmap, td.L0, td.L1, (const uint64_t *)wd, cnt,
                                               lane, true, bw0);
}

void qa_dec_desc_launch(const DecArgs &a, hipStream_t s) {
    if (a.t1 <= a.t0) return;
    qa_launch(k_dec_desc, dim3(nbt(a.t1 - a.t0, 256)), dim3(256), 0, s, a);
}

void qa_dec_desc_fill_launch(const DecArgs &a, hipStream_t s) {
    const uint32_t nd = a.t1 > a.t0 ? nbt(a.t1 - a.t0, 256) : 0;
    const uint32_t nn = a.nvwords ? nbt(a.nvwords, 256) : 0;
    const uint32_t nv = a.navbytes ? nbt(a.navbytes, 256) : 0;
    if (nd + nn + nv) qa_launch(k_dec_desc_fill, dim3(nd + nn + nv), dim3(256), 0, s, a, nd, nn);
}

// k_dec_fl<OW> over n tiles at WPB wavefronts per workgroup
template <int OW>
static void launch_fl(const DecArgs &b, uint64_t n, uint32_t wpb, hipStream_t s) {
    switch (wpb) {
    case 1: qa_launch((k_dec_fl<OW, 1>), dim3(nbt(n, 1)), dim3(64), 0, s, b); break;
    case 2: qa_launch((k_dec_fl<OW, 2>), dim3(nbt(n, 2)), dim3(128), 0, s, b); break;
    case 8: qa_launch((k_dec_fl<OW, 8>), dim3(nbt(n, 8)), dim3(512), 0, s, b); break;
    default: qa_launch((k_dec_fl<OW, 4>), dim3(nbt(n, 4)), dim3(256), 0, s, b);
    }
}

void qa_dec_tiles_launch(const DecArgs &a, hipStream_t s) {
    if (a.t1 <= a.t0) return;
    // launches of one tile decoder over [t0, t1) or its run table, at WPB
    // wavefronts per workgroup
    auto fl = [&](uint32_t kind, DecArgs b, uint64_t n, uint32_t wpb) {
        if (kind == 4) launch_fl<4>(b, n, wpb, s);
        else launch_fl<8>(b, n, wpb, s);
    };
    // the handle's tuned shape per decoder; untuned (0) decoders with enough
    // tiles are tuned here first, on this same idempotent stage
    uint32_t *tune = a.tune;
    for (uint32_t kind : {4u, 8u}) {
        const int ti = kind == 4 ? QA_TUNE_DEC4 : QA_TUNE_DEC8;
        if (!tune || tune[ti]) continue;
        uint64_t tiles = 0;
        DecArgs b = a;
        if (a.krun_all) {
            const int i = kind == 4 ? 0 : 1;
            tiles = a.kr_tot[i];
            b.krun = a.krun_all + a.kr_off[i];
            b.nkrun = a.kr_n[i];
            b.ktotal = tiles;
        } else {
            tiles = a.t1 - a.t0;  // the whole range, the other kinds' tiles skipped
        }
        if (tiles < QA_TUNE_MIN_TILES) continue;
        tune[ti] = qa_tune_pick([&](uint32_t w) { fl(kind, b, tiles, w); }, s);
    }
    qa_mark(a.kev0, s);
    const uint32_t w4 = tune && tune[QA_TUNE_DEC4] ? tune[QA_TUNE_DEC4] : 4;
    const uint32_t w8 = tune && tune[QA_TUNE_DEC8] ? tune[QA_TUNE_DEC8] : 4;
    auto launch = [&](uint32_t kind, uint64_t t0, uint64_t t1) {
        DecArgs b = a;
        b.t0 = t0;
        b.t1 = t1;
        if (kind == 4) fl(4, b, t1 - t0, w4);
        else if (kind == 8) fl(8, b, t1 - t0, w8);
        else qa_launch(k_dec_tiles, dim3(nbt(t1 - t0, 4)), dim3(256), 0, s, b);
    };
    if (a.krun_all) {
        const uint32_t kinds[3] = {4, 8, 0};
        for (int i = 0; i < 3; i++) {
            if (!a.kr_tot[i]) continue;
            DecArgs b = a;
            b.krun = a.krun_all + a.kr_off[i];
            b.nkrun = a.kr_n[i];
            b.ktotal = a.kr_tot[i];
            if (kinds[i] == 0) qa_launch(k_dec_tiles, dim3(nbt(b.ktotal, 4)), dim3(256), 0, s, b);
            else fl(kinds[i], b, b.ktotal, kinds[i] == 4 ? w4 : w8);
        }
    } else if (a.runs) {
        for (uint32_t r = 0; r < a.nruns; r++) {
            const uint64_t t0 = a.runs[r].t0 > a.t0 ? a.runs[r].t0 : a.t0;
            const uint64_t t1 = a.runs[r].t1 < a.t1 ? a.runs[r].t1 : a.t1;
            if (t1 > t0) launch(a.runs[r].kind, t0, t1);
        }
    } else {
        launch(4, a.t0, a.t1);
        launch(8, a.t0, a.t1);
        launch(0, a.t0, a.t1);
    }
    qa_mark(a.kev1, s);
}
