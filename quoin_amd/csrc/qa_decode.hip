// qa_decode.hip — gfx950 kernels decoding a batch of quoin streams.
//
// The host reads only the 16-byte header and the 9-byte frame headers
// (src/format.rs:65-94, src/decoder.rs:65-94); every data-sized step runs here:
//   1 validity blob -> bitmap + validity_out (src/validity.rs:73-119): one
//     pass over the blob's chunks chained by a decoupled look-back gives every
//     run its index, its start and the valid values before it (a byte ends a
//     varint iff its top bit is clear); each bitmap word ORs the valid runs
//     that meet it
//   2 valid count and the frames that cover it (the reference scans frames
//     until it has counted the valid values), in the same pass
//   3 tile map and sub-block index: each frame's sub-block headers (widths,
//     offsets), bounds checked; one wavefront per frame confirms runs of
//     equal-width sub-blocks per round trip
//   4 streams with nulls: the logical position of each tile's first value
//     (select over the bitmap, with the tile's descriptor)
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

__device__ __forceinline__ int dcol_by_vchunk(const DecArgs &a, uint32_t t) {
    QA_SEARCH(a.ncols, a.cols[mid].vchunk_base, t)
}

// A column's valid count is known: keep it, and the frames the reference
// decoder reads for it (src/decoder.rs:65-94: frames are read while fewer
// than nv values are counted, so the first frame starting at or past nv ends
// the read; out_off is non-decreasing). The frames read must hold exactly nv.
// Called by a whole wavefront: a 64-ary search over the column's frames.
__device__ void col_finish_wave(const DecArgs &a, int ci, uint64_t nv) {
    const int lane = threadIdx.x & 63;
    const DDecCol &c = a.cols[ci];
    // lo: frames before lo start below nv; hi: frames from hi start at or past it
    uint32_t lo = c.fbegin, hi = c.fend;
    while (hi > lo) {
        const uint32_t span = hi - lo;
        // probe every frame of a short span, else 64 spread over it
        const uint32_t pi = lo + (span <= 64 ? (uint32_t)lane : (uint32_t)(((uint64_t)span * lane) / 64));
        const bool below = lane < (int)span && a.frames[pi].out_off < nv;
        const uint64_t m = __ballot(below);
        const uint32_t cnt = (uint32_t)__popcll(m);
        if (span <= 64) {
            lo += cnt;
            break;
        }
        // probes 0..cnt-1 are below nv: the boundary lies after probe cnt - 1
        // and at or before probe cnt
        const uint32_t nlo = cnt ? lo + (uint32_t)(((uint64_t)span * (cnt - 1)) / 64) + 1 : lo;
        const uint32_t nhi = cnt < 64 ? lo + (uint32_t)(((uint64_t)span * cnt) / 64) : hi;
        lo = nlo;
        hi = nhi;
    }
    if (lane == 0) {
        DDecState &s = a.dst[ci];
        s.nv = nv;
        s.fused_end = lo;
        const uint64_t got = lo > c.fbegin ? a.frames[lo - 1].out_off + a.frames[lo - 1].n : 0;
        if (got != nv) atomicOr(a.err, QA_ERR_CORRUPT);
    }
}

// x then y: runs and bits add; y's valid bits are those of its parity, which
// x's run count decides. Bits and valid bits saturate at VB_CAP: once past n
// their exact value no longer matters, and bit 63 stays free for the marks.
constexpr uint64_t VB_CAP = 1ull << 62;
constexpr uint64_t VB_RDY = 1ull << 63;
__device__ __forceinline__ uint64_t vb_add(uint64_t x, uint64_t y) {
    return min(x + y, VB_CAP);
}
struct VbAgg {
    uint64_t c, s, v0, v1;
};
__device__ __forceinline__ VbAgg vb_cat(const VbAgg &x, const VbAgg &y) {
    const bool ev = !(x.c & 1);
    return VbAgg{x.c + y.c, vb_add(x.s, y.s), vb_add(x.v0, ev ? y.v0 : y.v1),
                 vb_add(x.v1, ev ? y.v1 : y.v0)};
}
__device__ __forceinline__ void vb_put(uint64_t *w, uint64_t v) {
    __hip_atomic_store(w, v | VB_RDY, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t vb_get(const uint64_t *w) {
    return __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// exclusive prefix sums over the workgroup (256 threads) of up to four
// values, with the totals; every thread calls it
template <int K>
__device__ __forceinline__ void vb_block_scan(uint64_t (&v)[K], uint64_t (&tot)[K],
                                              uint64_t (*red)[4]) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint64_t inc[K];
#pragma unroll
    for (int i = 0; i < K; i++) {
        inc[i] = qa_wave_scan_add64(v[i]);
        if (lane == 63) red[wv][i] = inc[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < K; i++) {
        uint64_t before = 0, all = 0;
        for (int w = 0; w < 4; w++) {
            const uint64_t r = red[w][i];
            before += w < wv ? r : 0;
            all += r;
        }
        v[i] = before + inc[i] - v[i];
        tot[i] = all;
    }
    __syncthreads();
}

// Decoupled look-back over the column's chunks (wavefront 0): publish this
// chunk's aggregate, combine the predecessors' until one with an inclusive
// prefix, publish this chunk's inclusive prefix, and return the exclusive one.
// A chunk publishes its aggregate before it waits on anyone, and chunks take
// tickets in launch order, so every predecessor waited on is running or done.
__device__ VbAgg vb_lookback(VbChunk *ch, uint32_t j, const VbAgg &mine, VbAgg *s_items,
                             uint32_t *s_fl, uint32_t *err) {
    const int lane = threadIdx.x & 63;
    uint32_t spins = 0;
    VbAgg pre{0, 0, 0, 0};
    if (j > 0) {
        if (lane < 4) vb_put(&ch[j].a[lane], lane == 0 ? mine.c : lane == 1 ? mine.s
                                             : lane == 2 ? mine.v0 : mine.v1);
        VbAgg acc{0, 0, 0, 0};
        int64_t base = j;  // predecessors base - 1, base - 2, ... (lane 0 nearest)
        for (;;) {
            const int64_t idx = base - 1 - lane;
            uint32_t fl = 2;
            VbAgg it{0, 0, 0, 0};
            if (idx >= 0) {
                const VbChunk &q = ch[idx];
                for (;;) {
                    // one round trip: the inclusive and the aggregate words
                    const uint64_t p0 = vb_get(&q.p[0]), p1 = vb_get(&q.p[1]), p2 = vb_get(&q.p[2]);
                    const uint64_t a0 = vb_get(&q.a[0]), a1 = vb_get(&q.a[1]), a2 = vb_get(&q.a[2]),
                                   a3 = vb_get(&q.a[3]);
                    if (p0 & p1 & p2 & VB_RDY) {
                        it = VbAgg{p0 & ~VB_RDY, p1 & ~VB_RDY, p2 & ~VB_RDY, 0};
                        break;
                    }
                    if (a0 & a1 & a2 & a3 & VB_RDY) {
                        it = VbAgg{a0 & ~VB_RDY, a1 & ~VB_RDY, a2 & ~VB_RDY, a3 & ~VB_RDY};
                        fl = 1;
                        break;
                    }
                    if (++spins > QA_SPIN_MAX) {
                        atomicOr(err, QA_ERR_STALLED);
                        break;  // fl stays 2: the walk ends
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            s_items[lane] = it;
            s_fl[lane] = fl;
            const uint64_t incl = __ballot(fl == 2);
            wave_lds_sync();
            // the nearest inclusive prefix (a lane past chunk 0 counts as one)
            const int lp = incl ? __builtin_ctzll(incl) : 63;
            if (lane == 0) {
                VbAgg w = s_items[lp];
                for (int l = lp - 1; l >= 0; l--) w = vb_cat(w, s_items[l]);
                acc = vb_cat(w, acc);
            }
            wave_lds_sync();
            if (incl) break;
            base -= 64;
        }
        pre = acc;  // lane 0's
        pre.c = (uint64_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)pre.c) |
                ((uint64_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(pre.c >> 32)) << 32);
        pre.s = qa_readlane64(pre.s, 0);
        pre.v0 = qa_readlane64(pre.v0, 0);
        pre.v1 = qa_readlane64(pre.v1, 0);
    }
    const VbAgg inc = vb_cat(pre, mine);
    if (lane < 3) vb_put(&ch[j].p[lane], lane == 0 ? inc.c : lane == 1 ? inc.s : inc.v0);
    return pre;
}

// Validity blobs (src/validity.rs:73-119), one chunk per workgroup, in one
// pass. RLE blob: the run varints ending in the chunk's 4096 bytes (a byte
// ends one iff its top bit is clear; src/varint.rs:24-39: at most 10 bytes,
// bits past 64 dropped) give each run its start, its index (runs alternate,
// the first one null) and the valid values before it; the run that reaches n
// sets the runs the reference reads and must end at n. Raw blob: the bitmap
// words with the bits past n cleared, validity_out, and each word's valid
// prefix. Either way the column's valid count then fixes its frames (the work
// of col_finish); workgroup 0 also does that for columns without a bitmap.
__global__ void __launch_bounds__(256) k_vb_scan(DecArgs a) {
    __shared__ uint32_t s_t;
    __shared__ uint64_t red[4][4];
    __shared__ VbAgg s_items[64];
    __shared__ uint32_t s_fl[64];
    __shared__ VbAgg s_pre;
    __shared__ __align__(16) uint8_t s_blob[QA_VB_CHUNK];
    __shared__ uint64_t s_reach;  // the valid count, once the run reaching n is met
    const int tid = threadIdx.x, wv = tid >> 6;
    if (blockIdx.x == 0)
        for (uint32_t c = wv; c < a.ncols; c += 4)
            if (!a.cols[c].nvwords) col_finish_wave(a, (int)c, a.cols[c].n);
    if (!a.nvchunks) return;
    if (tid == 0) {
        s_t = atomicAdd(a.vticket, 1u);
        s_reach = ~0ull;
    }
    __syncthreads();
    const uint32_t t = s_t;
    const int ci = dcol_by_vchunk(a, t);
    const DDecCol &c = a.cols[ci];
    const uint32_t j = t - c.vchunk_base;
    const uint8_t *p = c.src + c.vblob_off + 1;
    const uint64_t n = c.n;
    VbChunk *ch = a.vchunks + c.vchunk_base;

    if (c.src[c.vblob_off] == 0) {
        // raw bitmap (its length was checked on the host): two words a thread
        const uint64_t nbytes = (n + 7) / 8;
        uint64_t wd[2], pc[2] = {0, 0};
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const uint64_t w = (uint64_t)j * QA_VB_WORDS + 2 * tid + u;
            uint64_t out = 0;
            if (w < c.nvwords) {
                for (int k = 0; k < 8; k++)
                    if (w * 8 + k < nbytes) out |= (uint64_t)p[w * 8 + k] << (8 * k);
                const uint64_t bits = n - w * 64;
                if (bits < 64) out &= (1ull << bits) - 1;
            }
            wd[u] = out;
            pc[u] = (uint64_t)__popcll(out);
        }
        uint64_t v[1] = {pc[0] + pc[1]}, tot[1];
        vb_block_scan<1>(v, tot, red);
        if (wv == 0) {
            const VbAgg pre = vb_lookback(ch, j, VbAgg{0, 0, tot[0], tot[0]}, s_items, s_fl, a.err);
            if (tid == 0) s_pre = pre;
        }
        __syncthreads();
        const uint64_t nch = (c.nvwords + QA_VB_WORDS - 1) / QA_VB_WORDS;
        if (wv == 0 && j + 1 == nch) col_finish_wave(a, ci, s_pre.v0 + tot[0]);
        uint64_t rank = s_pre.v0 + v[0];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const uint64_t w = (uint64_t)j * QA_VB_WORDS + 2 * tid + u;
            if (w >= c.nvwords) break;
            c.bitmap[w] = wd[u];
            a.vpre[c.vword_base + w] = rank;
            tile_start_in_word(a, c, w, rank, wd[u]);
            rank += pc[u];
            if (c.validity_out) {
                uint8_t *vo = c.validity_out + w * 8;
                if (w * 8 + 8 <= nbytes && ((uintptr_t)vo & 7) == 0) {
                    *(uint64_t *)vo = wd[u];
                } else {
                    for (int k = 0; k < 8 && w * 8 + k < nbytes; k++) vo[k] = (uint8_t)(wd[u] >> (8 * k));
                }
            }
        }
        return;
    }

    // RLE: 16 blob bytes a thread, staged in LDS so a varint continued from
    // the thread before is read there (the chunk's first one from HBM)
    const uint64_t nb = c.vblob_len - 1;
    const uint64_t cb0 = (uint64_t)j * QA_VB_CHUNK;
    const uint64_t b0 = cb0 + 16 * (uint64_t)tid;
    {
        const uint8_t *src = p + b0;
        uint32_t wds[4];
        if (b0 + 16 <= nb) {
            // five aligned dwords hold the 16 bytes (none past the blob's last)
            const uintptr_t ad = (uintptr_t)src;
            const uint32_t *al = (const uint32_t *)(ad & ~(uintptr_t)3);
            const uint32_t m = (uint32_t)(ad & 3);
            uint32_t d[5];
#pragma unroll
            for (int u = 0; u < 4; u++) d[u] = al[u];
            d[4] = m ? al[4] : 0;
#pragma unroll
            for (int u = 0; u < 4; u++) wds[u] = __builtin_amdgcn_alignbyte(d[u + 1], d[u], m);
        } else {
#pragma unroll
            for (int u = 0; u < 4; u++) {
                uint32_t x = 0;
                for (int k = 0; k < 4; k++)
                    if (b0 + 4 * u + k < nb) x |= (uint32_t)src[4 * u + k] << (8 * k);
                wds[u] = x;
            }
        }
        *(uint4 *)&s_blob[16 * tid] = make_uint4(wds[0], wds[1], wds[2], wds[3]);
    }
    // bytes past the blob read as varint ends of value 0 and are not counted
    const uint32_t have = b0 >= nb ? 0u : (uint32_t)min((uint64_t)16, nb - b0);
    __syncthreads();
    uint32_t endm = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) endm |= (uint32_t)(!(s_blob[16 * tid + k] & 0x80)) << k;
    endm &= have == 16 ? 0xFFFFu : ((1u << have) - 1);
    // the varint open at b0: its bytes before b0 (at most 10 are looked at)
    uint64_t acc0 = 0;
    uint32_t len0 = 0;
    if (have) {
        uint64_t q = b0;
        auto byte_at = [&](uint64_t g) -> uint32_t {
            return g >= cb0 ? s_blob[g - cb0] : p[g];
        };
        while (q > 0 && (byte_at(q - 1) & 0x80) && b0 - q < 10) q--;
        len0 = (uint32_t)(b0 - q);
        for (uint32_t k = 0; k < len0; k++) acc0 |= (uint64_t)(byte_at(q + k) & 0x7f) << (7 * k);
    }
    // run k of this thread (in byte order): its length (an overlong varint or a
    // run past n reads as n + 1, an error if it is read)
    auto runs = [&](auto &&fn) {
        uint64_t acc = acc0;
        uint32_t len = len0, r = 0;
        for (int k = 0; k < 16; k++) {
            if (k >= (int)have) break;
            const uint32_t b = s_blob[16 * tid + k];
            if (len < 10) acc |= (uint64_t)(b & 0x7f) << (7 * len);
            len++;
            if (!(b & 0x80)) {
                const uint64_t x = len > 10 || acc > n ? n + 1 : acc;
                fn(r, x);
                r++;
                acc = 0;
                len = 0;
            }
        }
    };
    // this thread's first run index within the chunk, then the bits and the
    // valid bits (by the chunk-local parity) before it
    uint64_t cv[1] = {(uint64_t)__popc(endm)}, ctot[1];
    vb_block_scan<1>(cv, ctot, red);
    uint64_t sv[3] = {0, 0, 0}, stot[3];
    const uint32_t odd0 = (uint32_t)(cv[0] & 1);
    runs([&](uint32_t r, uint64_t x) {
        sv[0] = vb_add(sv[0], x);
        const bool odd = ((r ^ odd0) & 1) != 0;
        sv[1] = vb_add(sv[1], odd ? x : 0);
        sv[2] = vb_add(sv[2], odd ? 0 : x);
    });
    vb_block_scan<3>(sv, stot, red);
    if (wv == 0) {
        const VbAgg pre = vb_lookback(ch, j, VbAgg{ctot[0], stot[0], stot[1], stot[2]}, s_items, s_fl, a.err);
        if (tid == 0) s_pre = pre;
    }
    __syncthreads();
    const VbAgg pre = s_pre;
    const uint64_t nch = (nb + QA_VB_CHUNK - 1) / QA_VB_CHUNK;
    // runs that never reach n: truncated (src/validity.rs:96-113)
    if (tid == 0 && j + 1 == nch && vb_add(pre.s, stot[0]) < n) atomicOr(a.err, QA_ERR_CORRUPT);
    const uint64_t k0 = pre.c + cv[0];
    uint64_t st = vb_add(pre.s, sv[0]);
    uint64_t vb = vb_add(pre.v0, (pre.c & 1) ? sv[2] : sv[1]);
    uint64_t *rs = a.vruns + c.vbyte_base + ci, *rv = a.vbval + c.vbyte_base + ci;
    uint64_t reach = ~0ull, nv = 0;
    runs([&](uint32_t r, uint64_t x) {
        const uint64_t k = k0 + r;
        rs[k] = st;
        rv[k] = vb;
        if (st < n && st + x >= n) {
            reach = k;
            nv = vb + ((k & 1) ? n - st : 0);
            rs[k + 1] = st + x;
            if (st + x != n) atomicOr(a.err, QA_ERR_CORRUPT);  // a run overruns
        }
        vb = vb_add(vb, (k & 1) ? x : 0);
        st = vb_add(st, x);
    });
    if (reach != ~0ull) {
        // the run that reaches n: the reference reads runs 0..reach
        a.dst[ci].nruns = reach + 1;
        s_reach = nv;
    }
    __syncthreads();
    if (wv == 0 && s_reach != ~0ull) col_finish_wave(a, ci, s_reach);
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

// RLE blobs: the bitmap words, the decoded validity_out bytes (trailing bits
// cleared, src/lib.rs:503-506) and each word's valid prefix (the valid values
// before the run holding its bit 0, plus that run's bits before it when the
// run is valid). One wavefront per 64 consecutive words. A wavefront inside
// one stream finds the last run starting at or before its first bit by a
// 64-ary search; each later run starting inside its 4096 bits marks the first
// word whose bit 0 it reaches, and a prefix maximum over the 64 words gives
// every lane the last run starting at or before its word, so long runs cost
// nothing extra. A wavefront spanning two streams searches per word. Raw
// blobs were written by k_vb_scan.
__device__ __forceinline__ void vb_words_wave(const DecArgs &a, uint64_t wi, uint32_t *mk) {
    const int lane = threadIdx.x & 63;
    const uint64_t j0 = wi * WAVE;
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
    uint64_t kw = 0;  // the run holding the word's bit 0
    if (j0 + WAVE <= cwc.vword_base + cwc.nvwords && vtag(cwc) == 1) {
        const uint64_t nr = a.dst[cw].nruns;
        const uint64_t *rs = a.vruns + cwc.vbyte_base + cw;
        const uint64_t base = 0;
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
        kw = lo + m;
        out = vb_word_from(rs, base, nr, kw, j - cwc.vword_base);
    } else {
        if (j >= a.nvwords) return;
        ci = dcol_by_vword(a, j);
        const DDecCol &c = a.cols[ci];
        if (vtag(c) != 1) return;  // raw blob: k_vb_scan
        const uint64_t w = j - c.vword_base;
        const uint64_t nr = a.dst[ci].nruns;
        const uint64_t *rs = a.vruns + c.vbyte_base + ci;
        const uint64_t bit0 = w * 64;
        // last run k with start <= bit0
        uint64_t lo = 0, hi = nr ? nr - 1 : 0;
        while (lo < hi) {
            uint64_t mid = (lo + hi + 1) >> 1;
            if (rs[mid] <= bit0) lo = mid; else hi = mid - 1;
        }
        kw = lo;
        out = vb_word_from(rs, 0, nr, kw, w);
    }
    const DDecCol &c = a.cols[ci];
    const uint64_t w = j - c.vword_base, bits = c.n - w * 64;
    if (bits < 64) out &= (1ull << bits) - 1;
    c.bitmap[w] = out;
    {
        const uint64_t *rs = a.vruns + c.vbyte_base + ci, *rv = a.vbval + c.vbyte_base + ci;
        const uint64_t P = rv[kw] + ((kw & 1) ? w * 64 - rs[kw] : 0);
        a.vpre[j] = P;
        tile_start_in_word(a, c, w, P, out);
    }
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
        d.run0 = 0;
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
// stage 5: DeltaBitpack carries
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wsum(T v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += (T)__shfl_xor(v, d, 64);
    return v;
}

// per DeltaBitpack tile: the wrapping sum of its unzigzagged deltas, into the
// tile map's tail (k_dec_delta_scan turns the sums into carries). Reads only
// what k_dec_index wrote: the tile's frame, sub-block offset and width. wd
// stages the words of a w <= 32 sub-block (32 w, plus the row word read past
// the last one); wider sub-blocks (rare) are read directly.
__device__ __forceinline__ void delta_sum_tile(const DecArgs &a, uint64_t t, uint32_t *wd) {
    const int lane = threadIdx.x & 63;
    DTileD &ts = a.tiles[t];
    const uint32_t f = ts.frame;
    const DFrame &fr = a.frames[f];
    if (fr.mode != QA_M_DELTA || *a.err || !frame_used(a, f)) return;
    const uint32_t lt = (uint32_t)(t - fr.tile_base);
    const uint32_t cnt = min((uint32_t)QA_TILE, fr.n - lt * QA_TILE);
    const uint8_t *h = a.cols[fr.col].src + ts.off;
    const uint32_t w = ts.w;
    uint64_t mn = qa_ld64u(h + 1);
    uint64_t s = 0;
    if (w <= 32) {
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
    if (lane == 0) ts.tail = s;
}

// the Delta tiles [t0, t1) of the launch range, one wavefront per tile
__device__ __forceinline__ bool delta_span(const DecArgs &a, uint64_t *t0, uint64_t *t1) {
    *t0 = a.t0 > a.dt0 ? a.t0 : a.dt0;
    *t1 = a.t1 < a.dt1 ? a.t1 : a.dt1;
    return *t1 > *t0;
}

__global__ void __launch_bounds__(256) k_dec_delta_sums(DecArgs a) {
    __shared__ __align__(16) uint32_t wbuf[4][1088];
    uint64_t t0, t1;
    if (!delta_span(a, &t0, &t1)) return;
    const uint64_t t = t0 + (uint64_t)blockIdx.x * 4 + qa_wave_id();
    if (t < t1) delta_sum_tile(a, t, wbuf[qa_wave_id()]);
}

// The bitmap words and the Delta tile sums in one launch: workgroups
// [0, nbw) take 256 bitmap words each (k_vb_words), the rest four Delta tiles
// each. The two read disjoint inputs and write disjoint outputs, and both
// need only k_vb_scan and k_dec_index before them.
__global__ void __launch_bounds__(256) k_vb_words_dsums(DecArgs a, uint32_t nbw) {
    __shared__ __align__(16) uint32_t lds[4][1088];
    const int wv = qa_wave_id();
    if (blockIdx.x < nbw) {
        vb_words_wave(a, (uint64_t)blockIdx.x * 4 + wv, lds[wv]);
        return;
    }
    uint64_t t0, t1;
    if (!delta_span(a, &t0, &t1)) return;
    const uint64_t t = t0 + (uint64_t)(blockIdx.x - nbw) * 4 + wv;
    if (t < t1) delta_sum_tile(a, t, lds[wv]);
}

__global__ void __launch_bounds__(256) k_vb_words(DecArgs a) {
    __shared__ uint32_t marks[4][WAVE];
    vb_words_wave(a, (uint64_t)blockIdx.x * 4 + qa_wave_id(), marks[qa_wave_id()]);
}

// one wavefront per Delta frame: exclusive scan of its tile sums, seeded with
// the base, into the tile map (k_dec_desc copies each carry into its tile's
// descriptor)
__global__ void __launch_bounds__(256) k_dec_delta_scan(DecArgs a) {
    const int lane = threadIdx.x & 63;
    const uint32_t lo = a.f0 > a.df0 ? a.f0 : a.df0;
    uint32_t f = lo + blockIdx.x * 4 + qa_wave_id();
    if (f >= a.f1 || f >= a.df1 || *a.err) return;
    const DFrame &fr = a.frames[f];
    if (fr.mode != QA_M_DELTA || !frame_used(a, f)) return;
    uint64_t carry = qa_ld64u(a.cols[fr.col].src + fr.payload_off);  // base
    for (uint32_t t0 = 0; t0 < fr.ntiles; t0 += WAVE) {
        uint32_t t = t0 + lane;
        uint64_t s = t < fr.ntiles ? a.tiles[fr.tile_base + t].tail : 0;
        const uint64_t inc = qa_wave_scan_add64(s);
        if (t < fr.ntiles) a.tiles[fr.tile_base + t].tail = carry + inc - s;
        carry += __shfl(inc, 63, 64);
    }
}

// ---------------------------------------------------------------------------
// stage 7: all-null streams (no tile writes them), validity_out without bitmap
// ---------------------------------------------------------------------------
__global__ void k_dec_nullfill(DecArgs a) {
    dec_nullfill_one(a, (uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// validity_out of streams without a bitmap: all valid, trailing bits cleared
__global__ void k_dec_allvalid(DecArgs a) {
    dec_allvalid_one(a, (uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
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
    QA_LAUNCH(k_vb_scan, dim3(a.nvchunks ? a.nvchunks : 1), dim3(256), 0, s, a);
    if (a.nvbytes) QA_LAUNCH(k_vb_words, dim3(nbk(a.nvwords, 4 * WAVE)), dim3(256), 0, s, a);
}

// per frame [f0, f1): tile map, sub-block index, Fast-pool frame stages
void qa_decode_index_launch(const DecArgs &a, hipStream_t s) {
    if (a.f1 <= a.f0) return;
    QA_LAUNCH(k_dec_index, dim3(nbk(a.f1 - a.f0, 4)), dim3(256), 0, s, a);
    qa_fast_dec_index_launch(a, s);
}

// per frame [f0, f1) / tile [t0, t1): Delta sums and carries, tile decode
void qa_decode_values_launch(const DecArgs &a, hipStream_t s) {
    if (a.t1 <= a.t0) return;
    const uint64_t t0 = a.t0 > a.dt0 ? a.t0 : a.dt0, t1 = a.t1 < a.dt1 ? a.t1 : a.dt1;
    if (t1 > t0) {
        QA_LAUNCH(k_dec_delta_sums, dim3(nbk(t1 - t0, 4)), dim3(256), 0, s, a);
        const uint32_t f0 = a.f0 > a.df0 ? a.f0 : a.df0, f1 = a.f1 < a.df1 ? a.f1 : a.df1;
        if (f1 > f0) QA_LAUNCH(k_dec_delta_scan, dim3(nbk(f1 - f0, 4)), dim3(256), 0, s, a);
    }
    qa_dec_desc_launch(a, s);
    qa_dec_tiles_launch(a, s);
}

// all-null streams, validity_out of streams without a bitmap
void qa_decode_final_launch(const DecArgs &a, hipStream_t s) {
    if (a.nvwords) QA_LAUNCH(k_dec_nullfill, dim3(nbk(a.nvwords, 256)), dim3(256), 0, s, a);
    QA_LAUNCH(k_dec_allvalid, dim3(nbk(a.navbytes, 256)), dim3(256), 0, s, a);
}

// the batched form: every frame of every stream in one launch per stage; the
// bitmap words and the Delta tile sums share one launch after the index
void qa_decode_launch(const DecArgs &a, hipStream_t s) {
    QA_LAUNCH(k_vb_scan, dim3(a.nvchunks ? a.nvchunks : 1), dim3(256), 0, s, a);
    qa_decode_index_launch(a, s);
    const uint32_t nbw = a.nvbytes ? nbk(a.nvwords, 4 * WAVE) : 0;
    const uint32_t nbs = a.dt1 > a.dt0 ? nbk(a.dt1 - a.dt0, 4) : 0;
    QA_LAUNCH(k_vb_words_dsums, dim3(nbw + nbs), dim3(256), 0, s, a, nbw);
    if (a.df1 > a.df0) QA_LAUNCH(k_dec_delta_scan, dim3(nbk(a.df1 - a.df0, 4)), dim3(256), 0, s, a);
    // the tile descriptors, with the zeros of all-null streams and the
    // validity_out of streams without a bitmap (nothing after reads them)
    qa_dec_desc_fill_launch(a, s);
    qa_dec_tiles_launch(a, s);
}
