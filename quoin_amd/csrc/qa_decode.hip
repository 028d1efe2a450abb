// qa_decode.hip — gfx950 kernels decoding a batch of quoin streams.
//
// The host reads only the 16-byte header and the 9-byte frame headers
// (src/format.rs:65-94, src/decoder.rs:65-94); every data-sized step runs here:
//   1 validity blob -> bitmap + validity_out (src/validity.rs:73-119): RLE
//     varints are split in parallel (a byte ends a varint iff its top bit is
//     clear); one paired scan over the bytes gives each run its index and its
//     start; each bitmap word ORs the valid runs that meet it
//   2 valid count and the frames that cover it (the reference scans frames
//     until it has counted the valid values)
//   3 tile map and sub-block index: each frame's sub-block headers (widths,
//     offsets), bounds checked; one wavefront per frame confirms runs of
//     equal-width sub-blocks per round trip
//   4 streams with nulls: the logical position of each tile's first value
//     (select over the bitmap)
//   5 DeltaBitpack: per-tile delta sums, then a per-frame carry scan
//     (src/codecs/delta_bitpack.rs:48-64: a wrapping prefix sum)
//   6 per-tile decode straight into the typed output (qa_dec_tiles.hip)
//   7 all-null streams and validity_out of streams without a bitmap
#include "qa_dec.h"

// ---------------------------------------------------------------------------
// stage 1: validity blob -> bitmap
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint8_t vtag(const DDecCol &c) {
    return c.vblob_len ? c.src[c.vblob_off] : 0xFF;
}

// per blob byte: 1 if it ends a varint (its top bit is clear), and that
// varint's value (src/varint.rs:24-39), read back over the bytes before it;
// both are scanned together next
__global__ void k_vb_ends(DecArgs a) {
    uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.nvbytes) return;
    const DDecCol &c = a.cols[dcol_by_vbyte(a, g)];
    const uint8_t *p = c.src + c.vblob_off + 1;
    const uint64_t k = g - c.vbyte_base;
    const bool end = vtag(c) == 1 && !(p[k] & 0x80);
    uint64_t v = 0;
    if (end) {
        uint64_t j = k;
        while (j > 0 && (p[j - 1] & 0x80)) j--;
        const uint64_t len = k - j + 1;
        if (len > 10) {
            v = c.n + 1;  // overflowing varint: an error if the decoder reaches it
        } else {
            for (uint64_t q = 0; q < len; q++) v |= (uint64_t)(p[j + q] & 0x7f) << (7 * q);
            if (v > c.n) v = c.n + 1;  // keeps the prefix sums from wrapping
        }
    }
    a.vbpre[g] = end;
    a.vbval[g] = v;
}

// each varint-ending byte: its run's start (the prefix of the values before
// it) at the run's index; the runs' total after the last one. The run that
// reaches the n-th value (start < n <= end) sets how many runs the reference
// loop reads (the smallest m with P[m] >= n) and must end exactly at n.
__global__ void k_vb_runs(DecArgs a) {
    uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.nvbytes) return;
    if (g == 0) a.vruns[a.vbpre[a.nvbytes]] = a.vbval[a.nvbytes];
    const int ci = dcol_by_vbyte(a, g);
    const DDecCol &c = a.cols[ci];
    if (vtag(c) != 1 || (c.src[c.vblob_off + 1 + (g - c.vbyte_base)] & 0x80)) return;
    const uint64_t k = a.vbpre[g];
    a.vruns[k] = a.vbval[g];
    const uint64_t base = a.vbval[c.vbyte_base];
    const uint64_t st = a.vbval[g] - base, en = a.vbval[g + 1] - base;
    if (st < c.n && en >= c.n) {
        a.dst[ci].nruns = k - a.vbpre[c.vbyte_base] + 1;
        if (en != c.n) atomicOr(a.err, QA_ERR_CORRUPT);  // a run overruns
    }
}

// word w of an RLE validity blob's bitmap: the valid runs meeting bits
// [64 w, 64 w + 64), starting from run k0, the last one starting at or before
// bit 64 w (runs alternate starting from a null run, src/validity.rs:40-60);
// rs[k] - base is the start of run k
__device__ __forceinline__ uint64_t vb_word_from(const uint64_t *rs, uint64_t base, uint64_t nr,
                                                 uint64_t k0, uint64_t w) {
    const uint64_t bit0 = w * 64, bit1 = bit0 + 64;
    uint64_t out = 0;
    for (uint64_t k = k0; k < nr; k++) {
        const uint64_t st = rs[k] - base;
        if (st >= bit1) break;
        if (k & 1) {
            const uint64_t en = rs[k + 1] - base;
            const uint64_t s0 = st > bit0 ? st : bit0, e0 = en < bit1 ? en : bit1;
            if (e0 > s0) {
                const uint64_t cntb = e0 - s0, sh = s0 - bit0;
                out |= (cntb == 64 ? ~0ull : ((1ull << cntb) - 1)) << sh;
            }
        }
    }
    return out;
}

// The bitmap words, the decoded validity_out bytes (trailing bits cleared,
// src/lib.rs:503-506) and each word's popcount (into vpre, scanned next).
// One wavefront per 64 consecutive words. A wavefront inside one RLE stream
// finds the last run starting at or before its first bit by a 64-ary search;
// each later run starting inside its 4096 bits marks the first word whose
// bit 0 it reaches, and a prefix maximum over the 64 words gives every lane
// the last run starting at or before its word, so long runs cost nothing
// extra. Other wavefronts (raw bitmap blobs, a wavefront spanning two
// streams) search per word.
__global__ void __launch_bounds__(256) k_vb_words(DecArgs a) {
    __shared__ uint32_t marks[4][WAVE];
    const int lane = threadIdx.x & 63, wv = qa_wave_id();
    const uint64_t j0 = ((uint64_t)blockIdx.x * 4 + wv) * WAVE;
    if (j0 >= a.nvwords) return;
    const uint64_t j = j0 + lane;
    if (*a.err) {
        if (j < a.nvwords) a.vpre[j] = 0;
        return;
    }
    const int cw = dcol_by_vword(a, j0);
    const DDecCol &cwc = a.cols[cw];
    uint64_t out = 0;
    int ci = cw;
    if (j0 + WAVE <= cwc.vword_base + cwc.nvwords && vtag(cwc) == 1) {
        const uint64_t r0 = a.vbpre[cwc.vbyte_base], nr = a.dst[cw].nruns;
        const uint64_t *rs = a.vruns + r0;
        const uint64_t base = rs[0];
        const uint64_t B0 = (j0 - cwc.vword_base) * 64;
        // lo = the last run with start <= B0 (run 0 starts at 0)
        uint64_t lo = 0, hi = nr;
        while (hi - lo > 1) {
            const uint64_t span = hi - lo;
            const uint64_t pi = lo + 1 + (span - 1) * (uint64_t)lane / WAVE;
            const uint64_t le = __ballot(pi < hi && rs[pi] - base <= B0);
            const uint32_t c = (uint32_t)__builtin_popcountll(le);
            if (!c) {
                hi = lo + 1;
            } else {
                const uint64_t plo = lo + 1 + (span - 1) * (uint64_t)(c - 1) / WAVE;
                hi = c < WAVE ? lo + 1 + (span - 1) * (uint64_t)c / WAVE : hi;
                lo = plo;
            }
        }
        uint32_t *mk = marks[wv];
        mk[lane] = 0;
        wave_lds_sync();
        const uint64_t Bend = B0 + 64 * WAVE;
        for (uint64_t k0 = lo + 1; k0 < nr; k0 += WAVE) {
            const uint64_t k = k0 + lane;
            const uint64_t st = k < nr ? rs[k] - base : ~0ull;
            const bool in = st < Bend;  // and st > B0: lo is the last start <= B0
            if (in) {
                const uint64_t wl = (st - B0 + 63) >> 6;  // first word with bit 0 >= st
                if (wl < WAVE) atomicMax(&mk[wl], (uint32_t)(k - lo));
            }
            if (!(__ballot(in) >> 63)) break;
        }
        wave_lds_sync();
        const uint32_t m = qa_wave_scan_max32(mk[lane]);
        out = vb_word_from(rs, base, nr, lo + m, j - cwc.vword_base);
    } else {
        if (j >= a.nvwords) return;
        ci = dcol_by_vword(a, j);
        const DDecCol &c = a.cols[ci];
        const uint64_t w = j - c.vword_base;
        if (vtag(c) == 0) {
            // a blob shorter than ceil(n / 8) is flagged by k_dec_colfin; no
            // read goes past it
            const uint8_t *p = c.src + c.vblob_off + 1;
            const uint64_t nbytes = (c.n + 7) / 8, have = c.vblob_len - 1;
            for (int k = 0; k < 8; k++)
                if (w * 8 + k < nbytes && w * 8 + k < have) out |= (uint64_t)p[w * 8 + k] << (8 * k);
        } else {
            const uint64_t r0 = a.vbpre[c.vbyte_base], nr = a.dst[ci].nruns;
            const uint64_t *rs = a.vruns + r0;
            const uint64_t base = rs[0], bit0 = w * 64;
            // last run k with start P[k] <= bit0
            uint64_t lo = 0, hi = nr ? nr - 1 : 0;
            while (lo < hi) {
                uint64_t mid = (lo + hi + 1) >> 1;
                if (rs[mid] - base <= bit0) lo = mid; else hi = mid - 1;
            }
            out = vb_word_from(rs, base, nr, lo, w);
        }
    }
    const DDecCol &c = a.cols[ci];
    const uint64_t w = j - c.vword_base, bits = c.n - w * 64;
    if (bits < 64) out &= (1ull << bits) - 1;
    c.bitmap[w] = out;
    a.vpre[j] = (uint64_t)__popcll(out);
    if (c.validity_out) {
        uint64_t nbytes = (c.n + 7) / 8, b0 = w * 8;
        uint8_t *vo = c.validity_out + b0;
        if (b0 + 8 <= nbytes && ((uintptr_t)vo & 7) == 0) {
            *(uint64_t *)vo = out;
        } else {
            for (int k = 0; k < 8 && b0 + k < nbytes; k++) vo[k] = (uint8_t)(out >> (8 * k));
        }
    }
}

// valid count and the frames the reference decoder would read for it
__global__ void k_dec_colfin(DecArgs a) {
    int ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= (int)a.ncols) return;
    const DDecCol &c = a.cols[ci];
    DDecState &s = a.dst[ci];
    uint64_t nv = c.n;
    if (c.nvwords) {
        // the blob: a tag of 0 (raw bitmap, at least ceil(n / 8) bytes) or 1
        // (runs, which must reach n: k_vb_runs set nruns if they do)
        const uint8_t tag = vtag(c);
        bool bad = c.vblob_len == 0 || (tag != 0 && tag != 1);
        if (!bad && tag == 0) bad = c.vblob_len - 1 < (c.n + 7) / 8;
        if (!bad && tag == 1) bad = s.nruns == 0;  // n > 0 here: nvwords > 0
        if (bad) atomicOr(a.err, QA_ERR_CORRUPT);
        nv = a.vpre[c.vword_base + c.nvwords] - a.vpre[c.vword_base];
    }
    s.nv = nv;
    // frames are read while fewer than nv values are counted: the first frame
    // starting at or past nv ends the read (out_off is non-decreasing)
    uint32_t lo = c.fbegin, hi = c.fend;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (a.frames[mid].out_off < nv) lo = mid + 1; else hi = mid;
    }
    uint32_t fe = lo;
    s.fused_end = fe;
    uint64_t got = fe > c.fbegin ? a.frames[fe - 1].out_off + a.frames[fe - 1].n : 0;
    if (got != nv) atomicOr(a.err, QA_ERR_CORRUPT);
}

// ---------------------------------------------------------------------------
// stage 3: tile map and sub-block index (one wavefront per frame)
// ---------------------------------------------------------------------------
// one wavefront per frame: every tile's map entry (its frame, nothing indexed
// yet), then the sub-block index of the frame's FoR / Delta payload
__global__ void __launch_bounds__(256) k_dec_index(DecArgs a) {
    const int lane = threadIdx.x & 63;
    uint32_t f = a.f0 + blockIdx.x * 4 + qa_wave_id();
    if (f >= a.f1) return;
    const DFrame &fr = a.frames[f];
    for (uint32_t t = lane; t < fr.ntiles; t += WAVE) {
        DTileD d;
        d.off = 0;
        d.tail = 0;
        d.lpos = 0;
        d.frame = f;
        d.w = 0;
        a.tiles[fr.tile_base + t] = d;
    }
    __threadfence_block();  // the walk below rewrites entries other lanes wrote
    if (*a.err || !frame_used(a, f)) return;
    const DDecCol &col = a.cols[fr.col];
    const uint8_t *s = col.src;
    uint64_t pos = fr.payload_off, end = fr.payload_off + fr.plen;
    uint32_t err = 0;
    if (!fr.gpu_ok) {
        if (lane == 0) atomicOr(a.err, QA_ERR_UNSUPPORTED);
        return;
    }
    switch (fr.mode) {
    case QA_M_RAW:
        if (fr.plen != fr.n * lane_bytes_d(col.dtype)) err = QA_ERR_CORRUPT;
        break;
    case QA_M_CONST:
        if (fr.plen != 8) err = QA_ERR_CORRUPT;
        break;
    case QA_M_STRIDE:
        if (fr.plen != 16) err = QA_ERR_CORRUPT;
        break;
    case QA_M_FOR:
    case QA_M_DELTA:
        if (fr.mode == QA_M_DELTA) {
            if (fr.plen < 8) {
                err = QA_ERR_CORRUPT;
                break;
            }
            pos += 8;
        }
        if (for_walk_wave(s, pos, end, fr.n, a.tiles + fr.tile_base, 0, lane) == ~0ull)
            err = QA_ERR_CORRUPT;
        break;
    default:
        if (!fast_mode(fr.mode)) err = QA_ERR_UNSUPPORTED;  // fast modes: k_dec_index_fast
    }
    if (err && lane == 0) atomicOr(a.err, err);
}

// ---------------------------------------------------------------------------
// stage 4: logical start of each tile of a stream with nulls
// ---------------------------------------------------------------------------
__global__ void k_dec_lpos(DecArgs a) {
    uint64_t t = a.t0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.t1 || *a.err) return;
    DTileD &d = a.tiles[t];
    const uint32_t f = d.frame;
    if (!frame_used(a, f)) return;
    const DFrame &fr = a.frames[f];
    const DDecCol &col = a.cols[fr.col];
    if (!col.nvwords || a.dst[fr.col].nv == col.n) return;
    uint64_t c = fr.out_off + (t - fr.tile_base) * QA_TILE;  // compact index of the first value
    if (c == 0) {
        d.lpos = 0;  // the first tile also takes the leading nulls
        return;
    }
    // last bitmap word whose valid-count prefix is <= c; c < nv, so the word
    // holds the c-th valid value
    const uint64_t *vp = a.vpre + col.vword_base;
    const uint64_t p0 = vp[0];
    uint64_t lo = 0, hi = col.nvwords - 1;
    while (lo < hi) {
        uint64_t mid = (lo + hi + 1) >> 1;
        if (vp[mid] - p0 <= c) lo = mid; else hi = mid - 1;
    }
    uint64_t r = c - (vp[lo] - p0), w = col.bitmap[lo];
    if (r >= (uint64_t)__popcll(w)) {
        atomicOr(a.err, QA_ERR_CORRUPT);
        return;
    }
    d.lpos = lo * 64 + select64(w, (uint32_t)r);
}

// ---------------------------------------------------------------------------
// stage 5: DeltaBitpack carries
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wsum(T v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += (T)__shfl_xor(v, d, 64);
    return v;
}

// per DeltaBitpack tile: the wrapping sum of its unzigzagged deltas, into the
// tile descriptor's tail (k_dec_delta_scan turns the sums into carries)
__global__ void __launch_bounds__(256) k_dec_delta_sums(DecArgs a) {
    // staged words of a w <= 32 sub-block (32 w, plus the row-word read past the
    // last one); wider sub-blocks (rare) are read directly
    __shared__ __align__(16) uint32_t wbuf[4][1088];
    const int lane = threadIdx.x & 63, wv = qa_wave_id();
    uint64_t t = a.t0 + (uint64_t)blockIdx.x * 4 + wv;
    if (t >= a.t1) return;
    TileDec &td = a.tdec[t];
    if (td.mode != QA_M_DELTA) return;  // also every tile once an error is flagged
    const uint32_t cnt = td.cnt;
    const uint8_t *h = td.src + td.off;
    const uint32_t w = td.w;
    uint64_t mn = qa_ld64u(h + 1);
    uint64_t s = 0;
    if (w <= 32) {
        uint32_t *wd = wbuf[wv];
        stage_words(h + 9, 128 * w, wd, lane);
        wave_lds_sync();
        uint32_t v[16];
        unpack16_u32(wd, w, lane, v);
#pragma unroll
        for (int i = 0; i < 16; i++)
            s += 16u * (uint32_t)lane + i < cnt ? qa_unzigzag(mn + v[i]) : 0;
    } else {
        for (uint32_t k = lane; k < cnt; k += WAVE) s += qa_unzigzag(for_value_direct(h, k));
    }
    s = wsum(s);
    if (lane == 0) td.tail = s;
}

// one wavefront per frame: exclusive scan of the tile sums, seeded with the base
__global__ void __launch_bounds__(256) k_dec_delta_scan(DecArgs a) {
    const int lane = threadIdx.x & 63;
    uint32_t f = a.f0 + blockIdx.x * 4 + qa_wave_id();
    if (f >= a.f1 || *a.err) return;
    const DFrame &fr = a.frames[f];
    if (fr.mode != QA_M_DELTA || !frame_used(a, f)) return;
    uint64_t carry = qa_ld64u(a.cols[fr.col].src + fr.payload_off);  // base
    for (uint32_t t0 = 0; t0 < fr.ntiles; t0 += WAVE) {
        uint32_t t = t0 + lane;
        uint64_t s = t < fr.ntiles ? a.tdec[fr.tile_base + t].tail : 0;
        const uint64_t inc = qa_wave_scan_add64(s);
        if (t < fr.ntiles) a.tdec[fr.tile_base + t].tail = carry + inc - s;
        carry += __shfl(inc, 63, 64);
    }
}

// ---------------------------------------------------------------------------
// stage 7: all-null streams (no tile writes them), validity_out without bitmap
// ---------------------------------------------------------------------------
__global__ void k_dec_nullfill(DecArgs a) {
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.nvwords || *a.err) return;
    const int ci = dcol_by_vword(a, j);
    const DDecCol &c = a.cols[ci];
    if (a.dst[ci].nv != 0) return;
    uint64_t i0 = (j - c.vword_base) * 64;
    for (uint64_t i = i0; i < i0 + 64 && i < c.n; i++) store_typed(c, i, 0);
}

// validity_out of streams without a bitmap: all valid, trailing bits cleared
__global__ void k_dec_allvalid(DecArgs a) {
    uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.navbytes || *a.err) return;
    const DDecCol &col = a.cols[dcol_by_avbyte(a, g)];
    uint64_t b = g - col.avbyte_base, rem = col.n - b * 8;
    col.validity_out[b] = rem >= 8 ? 0xFF : (uint8_t)((1u << rem) - 1);
}

static inline unsigned nbk(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }
#define QA_LAUNCH(k, grid, ...)                                   \
    do {                                                          \
        dim3 g_ = (grid);                                         \
        if (g_.x) qa_launch(k, g_, __VA_ARGS__);         \
    } while (0)

// column-level: validity blob -> bitmap (+ validity_out), valid counts, the
// frames each stream's decoder reads
void qa_decode_prep_launch(const DecArgs &a, hipStream_t s) {
    if (a.nvbytes) {
        QA_LAUNCH(k_vb_ends, dim3(nbk(a.nvbytes, 256)), dim3(256), 0, s, a);
        qa_exclusive_scan2(a.vbpre, a.vbpre, a.vbval, a.vbval, a.nvbytes, a.scan_tmp, s);
        QA_LAUNCH(k_vb_runs, dim3(nbk(a.nvbytes, 256)), dim3(256), 0, s, a);
    }
    if (a.nvwords) {
        QA_LAUNCH(k_vb_words, dim3(nbk(a.nvwords, 4 * WAVE)), dim3(256), 0, s, a);
        qa_exclusive_scan(a.vpre, a.nvwords, a.vpre, a.scan_tmp, s);
    }
    QA_LAUNCH(k_dec_colfin, dim3(nbk(a.ncols, 64)), dim3(64), 0, s, a);
}

// per frame [f0, f1): tile map, sub-block index, Fast-pool frame stages
void qa_decode_index_launch(const DecArgs &a, hipStream_t s) {
    if (a.f1 <= a.f0) return;
    QA_LAUNCH(k_dec_index, dim3(nbk(a.f1 - a.f0, 4)), dim3(256), 0, s, a);
    qa_fast_dec_index_launch(a, s);
}

// tiles [t0, t1) of streams with nulls: logical start of each tile
void qa_decode_lpos_launch(const DecArgs &a, hipStream_t s) {
    if (a.nvwords && a.t1 > a.t0)
        QA_LAUNCH(k_dec_lpos, dim3(nbk(a.t1 - a.t0, 256)), dim3(256), 0, s, a);
}

// per frame [f0, f1) / tile [t0, t1): Delta carries, tile decode
void qa_decode_values_launch(const DecArgs &a, hipStream_t s) {
    if (a.t1 <= a.t0) return;
    qa_dec_desc_launch(a, s);
    QA_LAUNCH(k_dec_delta_sums, dim3(nbk(a.t1 - a.t0, 4)), dim3(256), 0, s, a);
    QA_LAUNCH(k_dec_delta_scan, dim3(nbk(a.f1 - a.f0, 4)), dim3(256), 0, s, a);
    qa_dec_tiles_launch(a, s);
}

// all-null streams, validity_out of streams without a bitmap
void qa_decode_final_launch(const DecArgs &a, hipStream_t s) {
    if (a.nvwords) QA_LAUNCH(k_dec_nullfill, dim3(nbk(a.nvwords, 256)), dim3(256), 0, s, a);
    QA_LAUNCH(k_dec_allvalid, dim3(nbk(a.navbytes, 256)), dim3(256), 0, s, a);
}

// the batched form: every frame of every stream in one launch per stage
void qa_decode_launch(const DecArgs &a, hipStream_t s) {
    qa_decode_prep_launch(a, s);
    qa_decode_index_launch(a, s);
    qa_decode_lpos_launch(a, s);
    qa_decode_values_launch(a, s);
    qa_decode_final_launch(a, s);
}
