// qa_rans.hip — 4-way interleaved order-0 rANS (src/entropy/rans.rs) on gfx950.
//
// Format (rans.rs:52-108): varint(n) varint(nsym) (sym:u8 varint(f))^nsym
// x0 x1 x2 x3 (u32 LE) then the renormalization bytes in decode order. M = 4096,
// L = 2^23, symbol i uses state i & 3, and the four states share the byte
// stream: the bytes of symbol i come after those of every earlier symbol.
//
// Decode: one stream per 4 lanes (lane = state), four streams per wavefront.
// A step decodes 4 symbols: each lane looks its slot up in the stream's LDS
// table (sym | f | slot - cum, one 32-bit read) and reads the 12 payload bytes
// at the stream position from an LDS window (issued together, so one LDS
// latency per step); the byte counts follow from the new state alone (0, 1 or
// 2 bytes for every state a valid stream reaches, from its leading zeros), so
// a 4-lane DPP scan places each lane's bytes. The window is a 1 KiB ring per
// stream, refilled from registers loaded a round (64 steps) ahead. A round
// runs branch-free; a state that needs 3 or more bytes (only a corrupt stream
// reaches one) is found at the round's end, and the round is replayed on an
// exact sequential path. The chain of one step (table read, state update,
// byte count, quad scan, byte extraction) is the kernel's bound: every stream
// of the batch is resident at once, so the time is the steps of the longest
// stream times that chain's latency.
//
// Encode: a workgroup per stream builds the histogram, normalizes it
// (rans.rs:25-50), writes the header and the per-symbol encode table; then one
// stream per 4 lanes (16 per wavefront) runs the backward state recurrence,
// dividing by f with a per-symbol reciprocal, and writes the renormalization
// bytes backward into scratch; a copy places them after the header.
#include "qa_kernels.h"

namespace {

constexpr uint32_t RM = 4096;          // total frequency (rans.rs:17-18)
constexpr uint32_t RL = 1u << 23;      // lower renormalization bound (rans.rs:20)
constexpr int DG = 4;                  // decode: streams per wavefront
constexpr int RING = 1024;             // decode: payload window bytes per stream
constexpr int STAGE = 1024;            // decode: header staging bytes (in the table)
constexpr int ROUND = 64;              // decode: steps between window refills
constexpr int EG = 16;                 // encode: streams per wavefront

struct DecLds {
    uint32_t tab[DG][RM];              // slot -> sym | (f & 4095) << 8 | (slot - cum) << 20
    uint32_t ring[DG][RING / 4 + 4];   // payload dwords by aligned position; the
                                       // first 4 repeated after the end
    uint32_t cum[DG][260];             // freq during the header parse, then cum
    uint32_t outs[DG + 1][64];         // 256 staged output bytes per stream; the
                                       // last row takes the idle lanes' writes
};

// 4-lane (quad) helpers on DPP; every lane of the wavefront must be active
template <int SRC>
__device__ __forceinline__ uint32_t quad_bcast(uint32_t v) {
    constexpr int ctrl = SRC | (SRC << 2) | (SRC << 4) | (SRC << 6);
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, ctrl, 0xF, 0xF, false);
}
__device__ __forceinline__ uint32_t quad_bcast_dyn(uint32_t v, int j) {
    switch (j) {
    case 0: return quad_bcast<0>(v);
    case 1: return quad_bcast<1>(v);
    case 2: return quad_bcast<2>(v);
    default: return quad_bcast<3>(v);
    }
}
__device__ __forceinline__ uint64_t quad_bcast64_dyn(uint64_t v, int j) {
    return (uint64_t)quad_bcast_dyn((uint32_t)v, j) |
           ((uint64_t)quad_bcast_dyn((uint32_t)(v >> 32), j) << 32);
}
// inclusive prefix sum inside each quad (idx = lane & 3)
__device__ __forceinline__ uint32_t quad_scan(uint32_t v, int idx) {
    // quad_perm [0,0,1,2]: the left neighbour
    uint32_t t = (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x90, 0xF, 0xF, false);
    v += idx >= 1 ? t : 0u;
    // quad_perm [0,0,0,1]: lanes 2, 3 add lanes 0, 1
    t = (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x40, 0xF, 0xF, false);
    return v + (idx >= 2 ? t : 0u);
}
__device__ __forceinline__ uint32_t quad_or(uint32_t v) {
    v |= (uint32_t)__builtin_amdgcn_mov_dpp((int)v, QA_DPP_QP_1032, 0xF, 0xF, false);
    v |= (uint32_t)__builtin_amdgcn_mov_dpp((int)v, QA_DPP_QP_2301, 0xF, 0xF, false);
    return v;
}
// value of lane `src` (any lane, per-lane source) — round-level bookkeeping only
__device__ __forceinline__ uint32_t lane_pull(uint32_t v, int src) {
    return (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)v);
}
__device__ __forceinline__ uint64_t lane_pull64(uint64_t v, int src) {
    return (uint64_t)lane_pull((uint32_t)v, src) | ((uint64_t)lane_pull((uint32_t)(v >> 32), src) << 32);
}

// header byte at aligned position q (q = byte offset + (src & 15)): staged in
// LDS for q < STAGE, else read from HBM; false past the payload end
__device__ __forceinline__ bool hdr_byte(const DecLds &L, int g, const uint8_t *base, uint64_t qe,
                                         uint64_t q, uint32_t *b) {
    if (q >= qe) return false;
    if (q < STAGE) *b = ((const uint8_t *)L.tab[g])[q];
    else *b = base[q];
    return true;
}
// LEB128 (src/varint.rs:24-39): 0 ok, 1 truncated, 2 overflow
__device__ __forceinline__ int hdr_varint(const DecLds &L, int g, const uint8_t *base, uint64_t qe,
                                          uint64_t *q, uint64_t *out) {
    uint64_t r = 0;
    uint32_t shift = 0;
    for (;;) {
        uint32_t b;
        if (!hdr_byte(L, g, base, qe, *q, &b)) return 1;
        (*q)++;
        r |= (uint64_t)(b & 0x7f) << shift;
        if (!(b & 0x80)) {
            *out = r;
            return 0;
        }
        shift += 7;
        if (shift >= 64) return 2;
    }
}

// 16-byte aligned chunk at aligned position q of a payload whose bytes occupy
// [la, qe); zero when the chunk holds no payload byte (a chunk that holds one
// lies in the same page as that byte)
__device__ __forceinline__ uint4 chunk16(const uint8_t *base, uint32_t la, uint64_t qe, uint64_t q) {
    if (q >= qe || q + 16 <= la) return make_uint4(0, 0, 0, 0);
    return *(const uint4 *)(base + q);
}

// flush lanes: bytes [c0, c0 + 16) of an output of ln bytes from staged words
__device__ __forceinline__ void flush16(uint8_t *out, bool al16, uint64_t c0, uint64_t ln, uint4 v) {
    if (c0 >= ln) return;
    if (al16 && c0 + 16 <= ln) {
        *(uint4 *)(out + c0) = v;
        return;
    }
    const uint32_t vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int b = 0; b < 16; b++)
        if (c0 + b < ln) out[c0 + b] = (uint8_t)(vv[b >> 2] >> (8 * (b & 3)));
}

// Decode kernel: wavefront w decodes streams 4w .. 4w+3.
__global__ void __launch_bounds__(64) k_rans_dec(const RansDecD *__restrict__ descs,
                                                 RansDecR *__restrict__ res, uint32_t ns) {
    __shared__ DecLds L;
    const int lane = threadIdx.x;
    const uint32_t s0 = blockIdx.x * DG;
    // decode-side mapping: group g = lane >> 2 (lanes 0..15), state idx = lane & 3
    const int g = (lane >> 2) & (DG - 1);
    const int idx = lane & 3;
    const bool dlane = lane < 4 * DG && s0 + (lane >> 2) < ns;
    // load-side mapping: stream gl = lane >> 4, sub-lane t
    const int gl = lane >> 4;
    const int t = lane & 15;
    const bool llane = s0 + gl < ns;

    RansDecD dl = {};
    if (llane) dl = descs[s0 + gl];
    // payload base aligned down to 16 bytes, a = offset of the first byte
    uint32_t tagged = dl.tagged;
    const uint8_t *psrc = dl.src + (tagged ? 1 : 0);
    uint64_t plen = dl.src_len > 0 && tagged ? dl.src_len - 1 : dl.src_len;
    const uint32_t la = (uint32_t)((uintptr_t)psrc & 15);
    const uint8_t *lbase = psrc - la;
    const uint64_t lqe = la + plen;

    // stage the first STAGE bytes (aligned) of each payload into its table area
    for (int c = 0; c < STAGE / (16 * 16); c++) {
        uint64_t q = (uint64_t)(c * 16 + t) * 16;
        uint4 v = llane ? chunk16(lbase, la, lqe, q) : make_uint4(0, 0, 0, 0);
        *(uint4 *)((uint8_t *)L.tab[gl] + q) = v;
    }
    for (int i = lane; i < DG * 260; i += 64) (&L.cum[0][0])[i] = 0;
    wave_lds_sync();

    // header parse: the group's lane 0 (rans.rs:110-154)
    uint32_t st = 0;            // 0 ok, 1 corrupt, 2 unsupported (tagged, other coder)
    uint64_t n = 0, q = 0;
    uint32_t x[4] = {0, 0, 0, 0};
    // the group's stream descriptor, from the load-side lanes
    const uint8_t *gsrc = (const uint8_t *)lane_pull64((uint64_t)(uintptr_t)lbase, (lane >> 2) * 16);
    const uint64_t gqe = lane_pull64(lqe, (lane >> 2) * 16);
    const uint32_t ga = lane_pull(la, (lane >> 2) * 16);
    const uint64_t gmax = lane_pull64(dl.max_len, (lane >> 2) * 16);
    const uint32_t gtag = lane_pull(tagged, (lane >> 2) * 16);
    const uint64_t gslen = lane_pull64(dl.src_len, (lane >> 2) * 16);
    const uint8_t gtagb = (uint8_t)lane_pull(llane && tagged && dl.src_len ? dl.src[0] : 0,
                                             (lane >> 2) * 16);
    if (dlane && idx == 0) {
        q = ga;
        if (gtag) {
            // decode_entropy_only (src/entropy/mod.rs:153-161): TAG_RANS = 2 here;
            // range coder (0), tANS (1) and the LZ cascade (3) are host coders
            if (gslen == 0) st = 1;
            else if (gtagb == 2) st = 0;
            else if (gtagb <= 3) st = 2;
            else st = 1;
        }
        uint64_t nsym = 0;
        if (st == 0 && hdr_varint(L, g, gsrc, gqe, &q, &n)) st = 1;
        if (st == 0 && n > gmax) st = 1;
        if (st == 0 && n >= (1ull << 34)) st = 2;  // step counter is 32-bit
        if (st == 0 && hdr_varint(L, g, gsrc, gqe, &q, &nsym)) st = 1;
        if (st == 0 && (nsym == 0 || nsym > 256)) st = 1;
        uint32_t sum = 0;
        for (uint64_t k = 0; st == 0 && k < nsym; k++) {
            uint32_t s;
            uint64_t f64v;
            if (!hdr_byte(L, g, gsrc, gqe, q, &s)) {
                st = 1;
                break;
            }
            q++;
            if (hdr_varint(L, g, gsrc, gqe, &q, &f64v)) {
                st = 1;
                break;
            }
            uint32_t f = (uint32_t)f64v;  // `as u32` before the check (rans.rs:125-128)
            if (f == 0 || f > RM) {
                st = 1;
                break;
            }
            L.cum[g][s] = f;
            sum += f;
        }
        if (st == 0 && sum != RM) st = 1;
        for (int i = 0; st == 0 && i < 4; i++) {
            uint32_t v = 0;
            for (int b = 0; b < 4; b++) {
                uint32_t by;
                if (!hdr_byte(L, g, gsrc, gqe, q + b, &by)) {
                    st = 1;
                    break;
                }
                v |= by << (8 * b);
            }
            x[i] = v;
            q += 4;
        }
    }
    wave_lds_sync();
    // broadcast the leader's results inside each quad
    st = quad_bcast<0>(st);
    n = (uint64_t)quad_bcast<0>((uint32_t)n) | ((uint64_t)quad_bcast<0>((uint32_t)(n >> 32)) << 32);
    q = (uint64_t)quad_bcast<0>((uint32_t)q) | ((uint64_t)quad_bcast<0>((uint32_t)(q >> 32)) << 32);
    uint32_t x0 = quad_bcast<0>(x[0]), x1 = quad_bcast<0>(x[1]), x2 = quad_bcast<0>(x[2]),
             x3 = quad_bcast<0>(x[3]);
    uint32_t xs = idx == 0 ? x0 : idx == 1 ? x1 : idx == 2 ? x2 : x3;
    if (!dlane) st = 1;

    // slot tables (rans.rs:136-146): cum by a wavefront scan, then 64 slots per
    // lane. Entry: sym | (f - 1) << 8 | (slot - cum) << 20. A table whose
    // frequencies sum below M (a repeated symbol) leaves slots past cum[256] to
    // symbol 0; when freq[0] = 0 those slots have f = 0, which the entry cannot
    // hold: such a stream ("irregular") decodes on the exact path.
    uint32_t irr = 0, ctot = RM;
    for (int gg = 0; gg < DG; gg++) {
        const uint32_t sg = (uint32_t)__builtin_amdgcn_readlane((int)st, gg * 4);
        if (s0 + gg >= ns || sg != 0) continue;
        uint32_t f0 = L.cum[gg][4 * lane], f1 = L.cum[gg][4 * lane + 1], f2 = L.cum[gg][4 * lane + 2],
                 f3 = L.cum[gg][4 * lane + 3];
        uint32_t loc = f0 + f1 + f2 + f3;
        uint32_t incl = (uint32_t)qa_wave_scan_add64(loc);
        uint32_t ex = incl - loc;
        wave_lds_sync();
        L.cum[gg][4 * lane] = ex;
        L.cum[gg][4 * lane + 1] = ex + f0;
        L.cum[gg][4 * lane + 2] = ex + f0 + f1;
        L.cum[gg][4 * lane + 3] = ex + f0 + f1 + f2;
        if (lane == 63) L.cum[gg][256] = incl;
        wave_lds_sync();
        const uint32_t tot = L.cum[gg][256];
        const uint32_t freq0 = L.cum[gg][1] - L.cum[gg][0];
        if (g == gg) {
            irr = tot < RM && freq0 == 0 ? 1u : 0u;
            ctot = tot;
        }
        const uint32_t slot0 = (uint32_t)lane * 64;
        // largest s in [0, 255] with cum[s] <= slot0
        uint32_t lo = 0, hi = 255;
        while (lo < hi) {
            uint32_t mid = (lo + hi + 1) >> 1;
            if (L.cum[gg][mid] <= slot0) lo = mid;
            else hi = mid - 1;
        }
        uint32_t s = lo;
        for (int j = 0; j < 64; j++) {
            const uint32_t slot = slot0 + j;
            uint32_t e;
            if (slot >= tot) {
                // slots past cum[256] keep symbol 0 (rans.rs:141 zero-fills)
                e = (((freq0 - 1) & 4095) << 8) | (slot << 20);
            } else {
                while (L.cum[gg][s + 1] <= slot) s++;
                const uint32_t c = L.cum[gg][s];
                e = s | (((L.cum[gg][s + 1] - c - 1) & 4095) << 8) | ((slot - c) << 20);
            }
            L.tab[gg][slot] = e;
        }
        wave_lds_sync();
    }

    // ---- decode loop -------------------------------------------------------
    // lane idx decodes symbols 4k + idx < n: steps k < klim
    uint32_t klim = st == 0 ? (uint32_t)((n + 3 - idx) >> 2) : 0u;
    uint32_t err = st != 0 ? 1u : 0u;
    uint32_t pos = (uint32_t)q;              // aligned position of the next byte
    const uint32_t qe32 = (uint32_t)gqe;
    uint32_t rb = 0;                         // the ring holds [rb, rb + RING)
    uint4 pf0 = make_uint4(0, 0, 0, 0), pf1 = pf0;  // prefetched [rb + RING, rb + RING + 512)
    auto mx = [](uint64_t a, uint64_t b) { return a > b ? a : b; };
    const uint32_t kmax = (uint32_t)qa_wave_reduce64(klim, mx, 0);
    const uint32_t sh8 = 8u * (uint32_t)idx, below = (1u << sh8) - 1u;
    // a lane's shift placed in its byte of the quad word, and the byte mask
    // of the lower lanes (for the dot product that sums their shifts)
    const uint32_t q8 = 8u << sh8, q16 = 16u << sh8, lowb = 0x01010101u & below;
    const int og = lane < 4 * DG ? g : DG;
    // slow-path threshold: states below 128 need 3+ bytes; irregular tables
    // always take the exact path
    const uint32_t thr = irr ? 0xFFFFFFFFu : 128u;
    // load-side view of each stream (no bytes for a failed header)
    const uint64_t ln = lane_pull64(st == 0 ? n : 0, gl * 4);
    uint8_t *const lout = dl.out;
    const bool lout16 = ((uintptr_t)lout & 15) == 0;
    auto ring_put = [&](uint32_t h, uint4 v) {
        const uint32_t di = (h >> 2) & (RING / 4 - 1);
        *(uint4 *)&L.ring[gl][di] = v;
        if (di == 0) *(uint4 *)&L.ring[gl][RING / 4] = v;
    };

    // a wavefront with an irregular table runs every round on the exact loop
    // (the fast rounds keep positions in bits, so payloads of 2^29 bytes or
    // more take the exact loop)
    const bool any_irr = __builtin_amdgcn_ballot_w64(irr != 0 || gqe >= (1ull << 29)) != 0;
    bool first = true;
    uint32_t k = 0;
    while (k < kmax) {
        // window maintenance for the coming round (<= ROUND steps of <= 8 bytes,
        // each reading 12 bytes at pos & ~3): 0 keep, 1 shift a half, 2 reload
        const bool act_g = err == 0 && k < klim;
        // every stream done or failed: stop (a corrupt length must not keep
        // the wavefront stepping)
        if (!__builtin_amdgcn_ballot_w64(act_g)) break;
        uint32_t mode = 0;
        if (act_g) {
            // a round reads below pos + 516: keep while pos <= rb + 508; the
            // shift drops [rb, rb + 512), so it needs pos >= rb + 512
            if (first) mode = 2;
            else if (pos > rb + (RING - 516))
                mode = pos >= rb + 512 && pos <= rb + 512 + (RING - 516) ? 1 : 2;
        }
        mode = quad_or(mode);  // a group whose lanes 1..3 finished still moves
        const uint32_t lmode = lane_pull(mode, gl * 4);
        const uint32_t lpos = lane_pull(pos, gl * 4);
        uint32_t lrb = lane_pull(rb, gl * 4);
        if (llane && lmode != 0) {
            if (lmode == 1) {
                // the prefetched [rb + RING, rb + RING + 512) replaces [rb, rb + 512)
                const uint32_t h = lrb + RING + (uint32_t)t * 32;
                ring_put(h, pf0);
                ring_put(h + 16, pf1);
                lrb += 512;
            } else {
                lrb = lpos & ~31u;
                for (int c = 0; c < RING / 512; c++) {
                    const uint32_t h = lrb + (uint32_t)c * 512 + (uint32_t)t * 32;
                    const uint4 a0 = chunk16(lbase, la, lqe, h), a1 = chunk16(lbase, la, lqe, h + 16);
                    ring_put(h, a0);
                    ring_put(h + 16, a1);
                }
            }
            const uint32_t h = lrb + RING + (uint32_t)t * 32;
            pf0 = chunk16(lbase, la, lqe, h);
            pf1 = chunk16(lbase, la, lqe, h + 16);
        }
        rb = lane_pull(lrb, ((lane >> 2) & (DG - 1)) * 16);
        first = false;
        wave_lds_sync();

        // steps to the end of the current 64-step output chunk
        const uint32_t rend = min(kmax, (k | 63u) + 1u);
        uint32_t sp = (uint32_t)og * 256 + ((k & 63) << 2) + (uint32_t)idx;  // staging byte
        bool slow = false;
        if (!any_irr) {
            // Fast round: no per-step branch. A state below 128 (3+ bytes: only a
            // corrupt stream reaches one) and a read past the payload end are
            // found at the round's end; the first replays the round on the exact
            // loop below from its start, the second fails the stream as that
            // loop would. Every step consumes at most 2 bytes per state (the
            // shift is clamped), so the round stays inside the window even on
            // a corrupt stream.
            const uint32_t k0 = k, xs0 = xs, pos0 = pos, sp0 = sp;
            uint32_t minx = 0xFFFFFFFFu;
            // One step of every lane. The payload position is kept in bits
            // (posb = 8 pos: a funnel shift by posb takes the byte offset pos & 3).
            uint32_t posb = pos << 3;
            // the payload words at posb, read one step ahead of their use
            uint32_t w0, w1, w2;
            auto window = [&]() {
                const uint32_t *wp = &L.ring[g][__builtin_amdgcn_ubfe(posb, 5, 8)];
                w0 = wp[0];
                w1 = wp[1];
                w2 = wp[2];
            };
            window();
            uint8_t *const outs = (uint8_t *)L.outs;
            auto step = [&](uint32_t thr1, uint32_t thr2, uint32_t dead, uint32_t spo) {
                const uint32_t e = L.tab[g][xs & (RM - 1)];
                const uint32_t xhi = xs >> 12;
                // x' = f (x >> 12) + slot - cum = (f - 1) xhi + (xhi + slot - cum)
                const uint32_t xn = __umul24((e >> 8) & 4095, xhi) + (xhi + (e >> 20));
                // renormalization bytes: 1 below 2^23, 2 below 2^15 (3 or more,
                // below 2^7, only on a corrupt stream: replayed). The shift in
                // bits, then the same in this lane's byte of the quad word.
                const bool b1 = xn < thr1, b2 = xn < thr2;
                const uint32_t sh = b2 ? 16u : (b1 ? 8u : 0u);
                uint32_t w = sh << sh8;
                w |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)w, QA_DPP_QP_1032, 0xF, 0xF, true);
                w |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)w, QA_DPP_QP_2301, 0xF, 0xF, true);
                // bit offset of this lane's bytes: the lower lanes' shifts
                const uint32_t exb = __builtin_amdgcn_udot4(w, lowb, 0u, false);
                // (the state's chain above is issued first; what follows fills its
                // waits)
                __builtin_amdgcn_sched_barrier(0);
                // the 8 payload bytes from pos, big-endian (byte pos in bits 56-63)
                const uint64_t win =
                    ((uint64_t)__builtin_amdgcn_perm(0u, __builtin_amdgcn_alignbit(w1, w0, posb), 0x00010203u)
                     << 32) |
                    __builtin_amdgcn_perm(0u, __builtin_amdgcn_alignbit(w2, w1, posb), 0x00010203u);
                // the quad's bits, summed onto the position
                posb = __builtin_amdgcn_udot4(w, 0x01010101u, posb, false);
                minx = min(minx, xn | dead);
                outs[spo] = (uint8_t)e;
                // this lane's bytes lead the shifted window; the new state is
                // x' followed by them, shifted left by sh (x' < 2^(32 - sh) when
                // sh > 0)
                const uint64_t xw = ((uint64_t)xn << 32) | (uint32_t)((win << exb) >> 32);
                xs = (uint32_t)((xw << sh) >> 32);
                window();
            };
            if (__builtin_amdgcn_ballot_w64(klim > k && klim < rend)) {
                for (; k < rend; k++, sp += 4) {
                    const bool act = k < klim;
                    uint32_t thr1 = act ? RL : 0u, thr2 = act ? (1u << 15) : 0u;
                    // kept as per-lane values, so each compare feeds its select
                    // directly (no scalar mask arithmetic on the state's chain)
                    asm volatile("" : "+v"(thr1), "+v"(thr2));
                    step(thr1, thr2, act ? 0u : 0xFFFFFFFFu, sp);
                }
            } else {
                const bool act = k < klim;
                uint32_t thr1 = act ? RL : 0u, thr2 = act ? (1u << 15) : 0u;
                const uint32_t dead = act ? 0u : 0xFFFFFFFFu;
                asm volatile("" : "+v"(thr1), "+v"(thr2));
                // four steps a trip (a round's steps end at a multiple of 64, or
                // at kmax)
                for (; k + 4 <= rend; k += 4, sp += 16) {
                    step(thr1, thr2, dead, sp);
                    step(thr1, thr2, dead, sp + 4);
                    step(thr1, thr2, dead, sp + 8);
                    step(thr1, thr2, dead, sp + 12);
                }
                for (; k < rend; k++, sp += 4) step(thr1, thr2, dead, sp);
            }
            pos = posb >> 3;
            if (__builtin_amdgcn_ballot_w64(minx < 128u)) {
                k = k0;
                xs = xs0;
                pos = pos0;
                sp = sp0;
            } else {
                const bool over = pos > qe32;
                err |= over ? 1u : 0u;
                klim = over ? 0u : klim;
            }
        }
        for (; k < rend; k++, sp += 4) {
            const bool act = k < klim;
            const uint32_t e = L.tab[g][xs & (RM - 1)];
            const uint32_t *wp = &L.ring[g][(pos >> 2) & (RING / 4 - 1)];
            const uint32_t w0 = wp[0], w1 = wp[1], w2 = wp[2];
            // keep the window reads next to the table read (one LDS wait per
            // step) instead of sunk into the branch that uses them
            asm volatile("" ::"v"(w0), "v"(w1), "v"(w2), "v"(e));
            const uint32_t xhi = xs >> 12;
            // x' = f (x >> 12) + slot - cum with f - 1 from the table
            const uint32_t xn = __umul24(((e >> 8) & 4095) + 1, xhi) + (e >> 20);
            // lanes past their end (or of a failed stream) need no bytes
            const uint32_t xe = act ? xn : 0xFFFFFFFFu;
            if (__builtin_amdgcn_ballot_w64(xe < thr)) {
                // 3+ bytes or an irregular table: the exact step below
                slow = true;
                break;
            }
            // staged symbol (lanes past their end write bytes the flush skips;
            // lanes without a stream write the spare row)
            ((uint8_t *)L.outs)[sp] = (uint8_t)e;
            // bytes per state: 0, 1 or 2 (xn >= 128); the quad's counts packed in
            // one word give each lane's offset and the total by byte sums
            const uint32_t cnt = (xe < RL ? 1u : 0u) + (xe < (1u << 15) ? 1u : 0u);
            uint32_t w = cnt << sh8;
            w |= (uint32_t)__builtin_amdgcn_mov_dpp((int)w, QA_DPP_QP_1032, 0xF, 0xF, false);
            w |= (uint32_t)__builtin_amdgcn_mov_dpp((int)w, QA_DPP_QP_2301, 0xF, 0xF, false);
            const uint32_t ex = __builtin_amdgcn_sad_u8(w & below, 0u, 0u);
            const uint32_t tot = __builtin_amdgcn_sad_u8(w, 0u, 0u);
            const uint32_t off = (pos & 3) + ex;
            const uint32_t lo = off < 4 ? w0 : (off < 8 ? w1 : w2);
            const uint32_t hi = off < 4 ? w1 : w2;
            const uint32_t d = __builtin_amdgcn_alignbyte(hi, lo, off);
            const uint32_t bs = __builtin_amdgcn_perm(0u, d, 0x0c0c0001u);  // b0 << 8 | b1
            const uint32_t sh = cnt << 3;
            // (a lane past its end keeps stepping on a state nothing reads)
            xs = (xn << sh) | (bs >> (16 - sh));
            pos += tot;
            // a byte at or past the payload end: truncated (rans.rs:167)
            const bool over = pos > qe32;
            err |= over ? 1u : 0u;
            klim = over ? 0u : klim;
        }
        if (slow) {
            // exact renormalization of step k, state by state in symbol order
            // (rans.rs:157-173): states needing 3+ bytes, irregular tables. The
            // window may not cover pos afterwards, so this step ends the round.
            const bool act = k < klim;
            const uint32_t slot = xs & (RM - 1);
            const uint32_t e = L.tab[g][slot];
            ((uint8_t *)L.outs)[sp] = (uint8_t)e;
            const uint32_t fx = irr && slot >= ctot ? 0u : ((e >> 8) & 4095) + 1;
            uint32_t xv = fx * (xs >> 12) + (e >> 20);
            uint64_t p = pos;
            uint32_t serr = 0;
            for (int jj = 0; jj < 4; jj++) {
                if (act && idx == jj) {
                    while (xv < RL) {
                        if (p >= gqe) {
                            serr = 1;
                            break;
                        }
                        xv = (xv << 8) | gsrc[p];
                        p++;
                    }
                }
                p = quad_bcast64_dyn(p, jj);
            }
            xs = xv;
            pos = (uint32_t)p;
            const uint32_t bad = quad_or(serr);
            err |= bad;
            klim = bad ? 0u : klim;
            k++;
        }
        if ((k & 63) == 0) {
            // the staged 256 bytes of chunk (k - 1) >> 6 are complete
            wave_lds_sync();
            if (llane)
                flush16(lout, lout16, ((uint64_t)((k - 1) >> 6) << 8) + (uint64_t)t * 16, ln,
                        *(const uint4 *)&L.outs[gl][t * 4]);
            wave_lds_sync();
        }
    }
    // the last, partial chunk (k = kmax, or the step where every stream ended)
    if (k & 63) {
        wave_lds_sync();
        if (llane)
            flush16(lout, lout16, ((uint64_t)((k - 1) >> 6) << 8) + (uint64_t)t * 16, ln,
                    *(const uint4 *)&L.outs[gl][t * 4]);
    }
    if (dlane && idx == 0) {
        RansDecR r;
        r.n = n;
        r.status = st == 2 ? 2 : (err ? 1 : 0);
        r.pad = 0;
        res[s0 + (lane >> 2)] = r;
    }
}


// ---- encode ----------------------------------------------------------------

// histogram, normalize (rans.rs:25-50), header and encode table of one stream
// per workgroup. Table entry: f | cum << 13 | l << 25 | m << 32 with
// l = ceil(log2 f) and m = ceil(2^(31+l) / f), so that for every x < 2^31
// floor(x / f) = (x * m) >> (31 + l) (Granlund and Montgomery 1994, thm 4.2:
// 2^(31+l) <= m f <= 2^(31+l) + 2^l).
__global__ void __launch_bounds__(256) k_rans_hist(const RansEncD *__restrict__ descs,
                                                   RansEncR *__restrict__ res,
                                                   uint64_t *__restrict__ tabs) {
    __shared__ uint32_t h[4][8][256];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const RansEncD D = descs[blockIdx.x];
    for (int i = tid; i < 4 * 8 * 256; i += 256) (&h[0][0][0])[i] = 0;
    __syncthreads();
    const uint32_t la = (uint32_t)((uintptr_t)D.src & 15);
    const uint8_t *base = D.src - la;
    const uint64_t qe = la + D.n;
    for (uint64_t q = (uint64_t)tid * 16; q < qe; q += 256 * 16) {
        const uint4 v = chunk16(base, la, qe, q);
        const uint32_t vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int b = 0; b < 16; b++) {
            const uint64_t p = q + b;
            if (p >= la && p < qe) atomicAdd(&h[w][lane & 7][(vv[b >> 2] >> (8 * (b & 3))) & 255], 1u);
        }
    }
    __syncthreads();
    if (w != 0) return;
    // lane owns symbols 4 lane + j
    uint32_t f[4];
    uint64_t loc = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int s = 4 * lane + j;
        uint32_t c = 0;
        for (int ww = 0; ww < 4; ww++)
            for (int cc = 0; cc < 8; cc++) c += h[ww][cc][s];
        uint32_t fj = 0;
        if (c) {
            fj = (uint32_t)(((uint64_t)c * RM) / D.n);
            fj = fj > 1 ? fj : 1;
        }
        f[j] = fj;
        loc += fj;
    }
    auto add = [](uint64_t a, uint64_t b) { return a + b; };
    auto mx = [](uint64_t a, uint64_t b) { return a > b ? a : b; };
    uint64_t sum = qa_wave_reduce64(loc, add, 0);
    // while sum > M: the largest f > 1 loses one, the last such symbol on ties
    while (sum > RM) {
        uint64_t key = 0;
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (f[j] > 1) key = mx(key, ((uint64_t)f[j] << 8) | (uint64_t)(4 * lane + j));
        key = qa_wave_reduce64(key, mx, 0);
        const uint32_t sb = (uint32_t)(key & 255);
#pragma unroll
        for (int j = 0; j < 4; j++)
            if ((uint32_t)(4 * lane + j) == sb) f[j]--;
        sum--;
    }
    if (sum < RM) {
        uint64_t key = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) key = mx(key, ((uint64_t)f[j] << 8) | (uint64_t)(4 * lane + j));
        key = qa_wave_reduce64(key, mx, 0);
        const uint32_t sb = (uint32_t)(key & 255);
#pragma unroll
        for (int j = 0; j < 4; j++)
            if ((uint32_t)(4 * lane + j) == sb) f[j] += RM - (uint32_t)sum;
    }
    // cum, present symbols and header entry bytes (sym, varint f)
    uint64_t fl = 0, el = 0, np = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        fl += f[j];
        if (f[j]) {
            el += 1 + qa_varint_len(f[j]);
            np++;
        }
    }
    const uint64_t cinc = qa_wave_scan_add64(fl), einc = qa_wave_scan_add64(el);
    const uint64_t nsym = qa_wave_reduce64(np, add, 0);
    uint64_t cx = cinc - fl, ex = einc - el;
    const uint32_t h0 = qa_varint_len(D.n) + qa_varint_len(nsym);
    const uint64_t hdr = h0 + qa_wave_reduce64(el, add, 0);
    if (lane == 0) {
        uint32_t k = qa_put_varint(D.dst, D.n);
        qa_put_varint(D.dst + k, nsym);
        res[blockIdx.x].hdr_len = hdr;
    }
    uint64_t *tab = tabs + (uint64_t)blockIdx.x * 256;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t s = 4 * lane + j, fj = f[j];
        uint64_t e = 0;
        if (fj) {
            D.dst[h0 + ex] = (uint8_t)s;
            ex += 1 + qa_put_varint(D.dst + h0 + ex + 1, fj);
            const uint32_t l = fj > 1 ? 32u - (uint32_t)__clz((int)(fj - 1)) : 0u;
            const uint64_t m = ((1ull << (31 + l)) + fj - 1) / fj;
            e = (uint64_t)fj | (cx << 13) | ((uint64_t)l << 25) | (m << 32);
        }
        tab[s] = e;
        cx += fj;
    }
}

// backward encode (rans.rs:73-97): one stream per 4 lanes, state = lane & 3;
// the bytes go backward from the stream's scratch end, then the 4 states
// trash: 64 bytes past every stream's scratch region. Steps that emit fewer
// than two bytes (and steps past a stream's end) store the unused byte there,
// so the byte stores and the symbol loads need no branch.
__global__ void __launch_bounds__(64) k_rans_enc(const RansEncD *__restrict__ descs,
                                                 RansEncR *__restrict__ res,
                                                 const uint64_t *__restrict__ tabs,
                                                 uint8_t *__restrict__ scr, uint32_t ns,
                                                 uint8_t *__restrict__ trash) {
    __shared__ uint64_t T[EG][256];
    const int lane = threadIdx.x;
    const int g = lane >> 2, idx = lane & 3;
    const uint32_t s0 = blockIdx.x * EG;
    for (int i = lane; i < EG * 256; i += 64) {
        const uint32_t sg = s0 + (i >> 8);
        T[i >> 8][i & 255] = sg < ns ? tabs[(uint64_t)sg * 256 + (i & 255)] : 0;
    }
    const bool valid = s0 + g < ns;
    RansEncD D = {};
    if (valid) D = descs[s0 + g];
    __syncthreads();
    const uint64_t n = D.n;
    const uint64_t nk = (n + 3) >> 2;
    auto mx = [](uint64_t a, uint64_t b) { return a > b ? a : b; };
    const uint64_t tmax = qa_wave_reduce64(nk, mx, 0);
    uint32_t x = RL;
    uint64_t P = D.scr_end;
    constexpr int U = 16;
    uint32_t cur[U], nxt[U];
    // symbol loads at a clamped index (a stream with no symbols reads trash)
    typedef const __attribute__((address_space(1))) uint8_t *gsrc_t;
    const gsrc_t src = (gsrc_t)(n ? D.src : trash);
    auto fetch = [&](uint64_t tb, uint32_t *buf) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint64_t t = tb + u;
            const uint64_t i = ((nk - 1 - t) << 2) + idx;
            const bool ok = t < nk && i < n;
            const uint32_t v = src[ok ? i : 0];
            buf[u] = ok ? v : 0u;
        }
    };
    fetch(0, cur);
    for (uint64_t tb = 0; tb < tmax; tb += U) {
        fetch(tb + U, nxt);
#pragma unroll
        for (int u = 0; u < U; u++) {
            // steps past tmax are inactive in every group
            const uint64_t t = tb + u;
            const uint64_t i = ((nk - 1 - t) << 2) + idx;
            const bool act = t < nk && i < n;
            const uint64_t e = T[g][cur[u]];
            const uint32_t f = (uint32_t)e & 0x1FFF, c = (uint32_t)(e >> 13) & 0xFFF;
            const uint32_t l = (uint32_t)(e >> 25) & 0xF, m = (uint32_t)(e >> 32);
            const uint32_t xmax = f << 19;  // ((L >> 12) << 8) * f
            const uint32_t cnt = act ? (uint32_t)(x >= xmax) + (uint32_t)((x >> 8) >= xmax) : 0u;
            const uint32_t xsh = x >> (8 * cnt);
            const uint32_t qd = (uint32_t)(((uint64_t)xsh * m) >> (31 + l));
            const uint32_t xn = (qd << 12) + (xsh - qd * f) + c;
            const uint32_t incl = quad_scan(cnt, idx);
            const uint32_t tot = quad_bcast<3>(incl);
            const uint64_t at = P - (tot - incl);  // one past this lane's bytes
            uint8_t *const b1 = cnt >= 1 ? scr + (at - 1) : trash + lane;
            uint8_t *const b2 = cnt == 2 ? scr + (at - 2) : trash + lane;
            *b1 = (uint8_t)x;
            *b2 = (uint8_t)(x >> 8);
            P -= tot;
            if (act) x = xn;
        }
#pragma unroll
        for (int u = 0; u < U; u++) cur[u] = nxt[u];
    }
    // states x0..x3 in front of the bytes (rans.rs:99-103)
    P -= 16;
    if (valid && n) {
#pragma unroll
        for (int b = 0; b < 4; b++) scr[P + 4 * idx + b] = (uint8_t)(x >> (8 * b));
    }
    if (valid && idx == 0) {
        res[s0 + g].start = P;
        res[s0 + g].out_len = n ? res[s0 + g].hdr_len + (D.scr_end - P) : 0;
    }
}

// place each coded stream after its header
__global__ void __launch_bounds__(256) k_rans_copy(const RansEncD *__restrict__ descs,
                                                   const RansEncR *__restrict__ res,
                                                   const uint8_t *__restrict__ scr) {
    const RansEncD D = descs[blockIdx.x];
    const RansEncR R = res[blockIdx.x];
    const uint64_t len = D.scr_end - R.start;
    const uint8_t *src = scr + R.start;
    uint8_t *dst = D.dst + R.hdr_len;
    const uint32_t head = (uint32_t)((4 - ((uintptr_t)dst & 3)) & 3);
    const uint64_t h = head < len ? head : len;
    if (threadIdx.x < h) dst[threadIdx.x] = src[threadIdx.x];
    const uint64_t body = (len - h) >> 2;
    uint32_t *dw = (uint32_t *)(dst + h);
    for (uint64_t j = threadIdx.x; j < body; j += 256) dw[j] = qa_ld32u(src + h + 4 * j);
    const uint64_t tail0 = h + 4 * body;
    if (tail0 + threadIdx.x < len) dst[tail0 + threadIdx.x] = src[tail0 + threadIdx.x];
}

}  // namespace

__global__ void k_put_bytes(BytePuts b) {
    const uint32_t i = threadIdx.x;
    if (i < b.n) b.p[i][0] = b.v[i];
}
void qa_put_bytes_launch(const BytePuts &b, hipStream_t s) {
    if (b.n) qa_launch(k_put_bytes, dim3(1), dim3(128), 0, s, b);
}

void qa_rans_decode_launch(const RansDecD *d, RansDecR *r, uint32_t ns, hipStream_t s) {
    if (!ns) return;
    qa_launch(k_rans_dec, dim3((ns + DG - 1) / DG), dim3(64), 0, s, d, r, ns);
}

void qa_rans_encode_launch(const RansEncD *d, RansEncR *r, uint64_t *tabs, uint8_t *scr, uint32_t ns,
                           uint8_t *trash, hipStream_t s) {
    if (!ns) return;
    qa_launch(k_rans_hist, dim3(ns), dim3(256), 0, s, d, r, tabs);
    qa_launch(k_rans_enc, dim3((ns + EG - 1) / EG), dim3(64), 0, s, d, r, tabs, scr, ns, trash);
    qa_launch(k_rans_copy, dim3(ns), dim3(256), 0, s, d, r, scr);
}
