// qa_hybrid.cpp — the engine's frames that pass through the host coders.
//
// Decode (src/decoder.rs:145-190): a stream holding frames the device tile
// decoders cannot take (the predictor and linear residual modes, LZ,
// ByteTranspose, pco, and the entropy-coded payloads of FloatMult, ALP-RD and
// Dict) is rewritten, frame by frame, into a stream of the same values that
// they can: the entropy layer runs through quoin_gpu_decode_residuals (rANS on
// the GPU; the range coder, tANS and LZ cascade on host cores), the
// recurrences of the residual modes on the device, LZ / ByteTranspose / pco
// on host cores. Value frames of those modes become RAW frames; FloatMult,
// ALP-RD and Dict keep their mode with their bit-packed sub-streams, so the
// device still reconstructs their values. The validity blob and every other
// frame are copied unchanged.
#include "qa_hybrid.h"

#include <algorithm>
#include <cstring>

#include "qa_hostcoders.h"
#include "qa_pco.h"

namespace qahyb {

namespace {
enum {
    M_PRED = 0, M_LZ = 4, M_ORDERED_DELTA = 7, M_BYTE_TRANSPOSE = 11, M_DELTA2 = 20, M_PRED2 = 25,
    M_PRED_RC = 29, M_DELTA_DP = 45, M_PCO = 54
};

uint64_t rd64(const uint8_t *p) {
    uint64_t v;
    memcpy(&v, p, 8);
    return v;
}
uint32_t rd32(const uint8_t *p) {
    uint32_t v;
    memcpy(&v, p, 4);
    return v;
}
void put64(Bytes &o, uint64_t v) {
    for (int k = 0; k < 8; k++) o.push_back((uint8_t)(v >> (8 * k)));
}
void put32(Bytes &o, uint32_t v) {
    for (int k = 0; k < 4; k++) o.push_back((uint8_t)(v >> (8 * k)));
}
uint64_t zigzag(uint64_t x) { return (x << 1) ^ (uint64_t)((int64_t)x >> 63); }
uint64_t unzigzag(uint64_t z) { return (z >> 1) ^ (0ull - (z & 1)); }
size_t lane_bytes(uint32_t dt) { return (dt == QA_I32 || dt == QA_U32 || dt == QA_F32) ? 4 : 8; }
int pco_type(uint32_t dt) {
    switch (dt) {
    case QA_F64: return QA_PCO_F64;
    case QA_F32: return QA_PCO_F32;
    case QA_I64: return QA_PCO_I64;
    case QA_U64: return QA_PCO_U64;
    case QA_I32: return QA_PCO_I32;
    default: return QA_PCO_U32;
    }
}
// decoder.rs:18-20
uint64_t resid_bound(uint64_t n) { return n > (UINT64_MAX - 16) / 10 ? UINT64_MAX : n * 10 + 16; }

// FastLanes pack of 1024 residuals at width w (bitpack.rs:47-133)
void pack_sub(const uint64_t *res, uint32_t w, Bytes &out) {
    if (w <= 32) {
        uint32_t words[32 * 32];
        memset(words, 0, sizeof(uint32_t) * 32 * w);
        const uint32_t mask = w == 32 ? 0xFFFFFFFFu : ((1u << w) - 1);
        for (uint32_t r = 0; r < 32; r++) {
            const uint32_t bit = r * w, word = bit / 32, shift = bit % 32;
            for (uint32_t l = 0; l < 32; l++) {
                const uint32_t v = (uint32_t)res[r * 32 + l] & mask;
                words[word * 32 + l] |= v << shift;
                if (shift + w > 32) words[(word + 1) * 32 + l] |= v >> (32 - shift);
            }
        }
        for (uint32_t i = 0; i < 32 * w; i++) put32(out, words[i]);
    } else {
        uint64_t words[16 * 64];
        memset(words, 0, sizeof(uint64_t) * 16 * w);
        const uint64_t mask = w == 64 ? ~0ull : ((1ull << w) - 1);
        for (uint32_t r = 0; r < 64; r++) {
            const uint32_t bit = r * w, word = bit / 64, shift = bit % 64;
            for (uint32_t l = 0; l < 16; l++) {
                const uint64_t v = res[r * 16 + l] & mask;
                words[word * 16 + l] |= v << shift;
                if (shift + w > 64) words[(word + 1) * 16 + l] |= v >> (64 - shift);
            }
        }
        for (uint32_t i = 0; i < 16 * w; i++) put64(out, words[i]);
    }
}

// ---- frames -----------------------------------------------------------------
struct Frame {
    int mode;
    uint32_t n;
    size_t poff;  // payload offset in the stream
    uint32_t plen;
    bool host;
};
struct Stream {
    const uint8_t *h = nullptr, *d = nullptr;
    size_t len = 0;
    uint32_t dtype = 0, plog2 = 16;
    size_t body = 0;  // first frame
    size_t tail = 0;  // bytes past the last parsed frame
    std::vector<Frame> frames;
    bool any_host = false;
};

bool scan(Stream &st) {
    const uint8_t *p = st.h;
    const size_t len = st.len;
    if (len < 16 || memcmp(p, "FCR1", 4) != 0 || p[4] != 2) return false;
    if (p[5] & 0x02) return false;  // decimal container: its limbs are streams of their own
    st.plog2 = p[6];
    st.dtype = p[7];
    if (st.dtype > 5) return false;
    size_t pos = 16;
    if (p[5] & 1) {
        uint64_t vlen;
        if (!get_varint(p, len, pos, vlen) || vlen > len - pos) return false;
        pos += vlen;
    }
    st.body = pos;
    while (len - pos >= 9) {
        Frame f;
        f.mode = p[pos];
        f.n = rd32(p + pos + 1);
        f.plen = rd32(p + pos + 5);
        if (f.n > QA_MAX_Q || f.plen > len - pos - 9) break;
        f.poff = pos + 9;
        f.host = frame_needs_host(f.mode, p + f.poff, f.plen);
        st.any_host |= f.host;
        st.frames.push_back(f);
        pos += 9 + (size_t)f.plen;
    }
    st.tail = pos;
    return true;
}

// one entropy-coded sub-stream of a frame, decoded by quoin_gpu_decode_residuals
struct ResJob {
    size_t stream, frame;
    size_t off;      // in the stream
    uint64_t len, max_len;
    uint64_t out_off;  // in the residual scratch
    uint64_t out_len = 0;
    int32_t status = QUOIN_OK;
};
// one recurrence (predictor / linear decode) producing n lanes
struct RecJob {
    size_t stream, frame;
    int kind;                // 0 FCM, 25 DFCM, or a linear kind (7, 20, 45)
    uint32_t plog2;
    const uint8_t *src_dev;  // device residual bytes
    uint64_t src_len;
    uint64_t lane_off;       // in the lane scratch
    uint64_t n;
    int32_t status = QUOIN_OK;
};

// where a frame's sub-streams live (offsets relative to the payload)
struct Plan {
    int res[3] = {-1, -1, -1};  // ResJob indices
    int rec = -1;
    // ALP-RD: codes tag, codes blob [coff, coff+clen), rights tag, rights at roff
    uint8_t ctag = 0, rtag = 0;
    size_t coff = 0, clen = 0, roff = 0, head = 0;
    // Dict
    uint64_t card = 0;
    uint8_t vtag = 0, ctag_d = 0, nplanes = 0;
    size_t voff = 0, vlen = 0, cpos = 0;
};

bool alprd_layout(const uint8_t *p, size_t len, Plan &pl) {
    size_t pos = 0;
    if (len < 2) return false;
    const uint32_t left = p[0], dl = p[1];
    if (left < 1 || left > 63 || dl > 8) return false;
    pos = 2 + 8 * (size_t)dl;
    if (pos > len) return false;
    uint64_t nexc;
    if (!get_varint(p, len, pos, nexc)) return false;
    for (uint64_t k = 0; k < nexc; k++) {
        uint64_t x;
        if (!get_varint(p, len, pos, x) || len - pos < 8) return false;
        pos += 8;
    }
    pl.head = pos;
    if (pos >= len) return false;
    pl.ctag = p[pos++];
    uint64_t cl;
    if (!get_varint(p, len, pos, cl) || cl > len - pos) return false;
    pl.coff = pos;
    pl.clen = (size_t)cl;
    pos += cl;
    if (pos >= len) return false;
    pl.rtag = p[pos++];
    pl.roff = pos;
    return true;
}

bool dict_layout(const uint8_t *p, size_t len, Plan &pl) {
    size_t pos = 0;
    if (!get_varint(p, len, pos, pl.card)) return false;
    if (pos >= len) return false;
    pl.vtag = p[pos++];
    uint64_t vl;
    if (!get_varint(p, len, pos, vl) || vl > len - pos) return false;
    pl.voff = pos;
    pl.vlen = (size_t)vl;
    pos += vl;
    if (pos >= len) return false;
    pl.ctag_d = p[pos++];
    pl.cpos = pos;
    return true;
}
}  // namespace

void put_varint(Bytes &o, uint64_t v) {
    while (v >= 0x80) {
        o.push_back((uint8_t)(v | 0x80));
        v >>= 7;
    }
    o.push_back((uint8_t)v);
}
bool get_varint(const uint8_t *p, size_t len, size_t &pos, uint64_t &v) {
    uint64_t r = 0;
    for (uint32_t sh = 0; sh < 70; sh += 7) {
        if (pos >= len) return false;
        const uint8_t b = p[pos++];
        if (sh == 63 && (b & 0x7e)) return false;  // varint.rs:25-35: overflow
        r |= (uint64_t)(b & 0x7f) << sh;
        if (!(b & 0x80)) {
            v = r;
            return true;
        }
    }
    return false;
}

void for_encode(const uint64_t *v, size_t n, bool sgn, Bytes &out) {
    put_varint(out, n);
    uint64_t res[QA_TILE];
    for (size_t i = 0; i < n; i += QA_TILE) {
        const size_t m = std::min<size_t>(QA_TILE, n - i);
        uint64_t mn, range;
        if (sgn) {
            int64_t a = INT64_MAX, b = INT64_MIN;
            for (size_t k = 0; k < m; k++) {
                a = std::min(a, (int64_t)v[i + k]);
                b = std::max(b, (int64_t)v[i + k]);
            }
            mn = (uint64_t)a;
            range = (uint64_t)b - (uint64_t)a;
        } else {
            uint64_t a = UINT64_MAX, b = 0;
            for (size_t k = 0; k < m; k++) {
                a = std::min(a, v[i + k]);
                b = std::max(b, v[i + k]);
            }
            mn = a;
            range = b - a;
        }
        const uint32_t w = range ? 64 - (uint32_t)__builtin_clzll(range) : 0;
        out.push_back((uint8_t)w);
        put64(out, mn);
        if (!w) continue;
        for (size_t k = 0; k < QA_TILE; k++) res[k] = k < m ? v[i + k] - mn : 0;
        pack_sub(res, w, out);
    }
}

void delta_encode(const uint64_t *v, size_t n, Bytes &out) {
    const uint64_t base = n ? v[0] : 0;
    put64(out, base);
    std::vector<uint64_t> d(n);
    uint64_t prev = base;
    for (size_t i = 0; i < n; i++) {
        d[i] = zigzag(v[i] - prev);
        prev = v[i];
    }
    for_encode(d.data(), n, false, out);
}

bool frame_needs_host(int mode, const uint8_t *p, size_t len) {
    switch (mode) {
    case M_PRED: case M_LZ: case M_ORDERED_DELTA: case M_BYTE_TRANSPOSE: case M_DELTA2: case M_PRED2:
    case M_PRED_RC: case M_DELTA_DP: case M_PCO:
        return true;
    case QA_M_FLOAT_MULT: return len >= 1 && p[0] == 1;
    case QA_M_ALP_RD: {
        Plan pl;
        return alprd_layout(p, len, pl) && (pl.ctag == 1 || pl.rtag == 1);
    }
    case QA_M_DICT: {
        Plan pl;
        return dict_layout(p, len, pl) && (pl.vtag == 2 || pl.ctag_d == 1);
    }
    default: return false;
    }
}

int rewrite_streams(QuoinGpu *g, size_t ns, const uint8_t *const *src_host, const uint8_t *const *src_dev,
                    const uint64_t *lens, std::vector<Bytes> &out, hipStream_t s) {
    out.assign(ns, Bytes());
    std::vector<Stream> st(ns);
    bool any = false;
    for (size_t i = 0; i < ns; i++) {
        st[i].h = src_host[i];
        st[i].d = src_dev[i];
        st[i].len = (size_t)lens[i];
        if (scan(st[i]) && st[i].any_host) any = true;
        else st[i].any_host = false;
    }
    if (!any) return QUOIN_OK;

    // ---- plan the sub-streams -------------------------------------------------
    std::vector<std::vector<Plan>> plans(ns);
    std::vector<ResJob> rj;
    std::vector<RecJob> cj;
    uint64_t res_total = 0, lane_total = 0;
    auto add_res = [&](size_t si, size_t fi, size_t off, uint64_t len, uint64_t max_len) {
        ResJob j{};
        j.stream = si;
        j.frame = fi;
        j.off = off;
        j.len = len;
        j.max_len = max_len;
        j.out_off = res_total;
        res_total += (max_len + 255) & ~255ull;
        rj.push_back(j);
        return (int)rj.size() - 1;
    };
    for (size_t si = 0; si < ns; si++) {
        Stream &S = st[si];
        if (!S.any_host) continue;
        plans[si].resize(S.frames.size());
        for (size_t fi = 0; fi < S.frames.size(); fi++) {
            const Frame &f = S.frames[fi];
            if (!f.host) continue;
            Plan &pl = plans[si][fi];
            const uint8_t *p = S.h + f.poff;
            const uint64_t n = f.n;
            switch (f.mode) {
            case M_PRED_RC: case M_PRED2: case M_DELTA2: case M_DELTA_DP: case M_ORDERED_DELTA:
                pl.res[0] = add_res(si, fi, f.poff, f.plen, resid_bound(n));
                break;
            case M_LZ: pl.res[0] = add_res(si, fi, f.poff, f.plen, n * 16 + 1024); break;
            case M_BYTE_TRANSPOSE: pl.res[0] = add_res(si, fi, f.poff, f.plen, n * 8 + 16); break;
            case QA_M_FLOAT_MULT:
                if (f.plen < 2) return QUOIN_ERR_CORRUPT;
                pl.res[0] = add_res(si, fi, f.poff + 2, f.plen - 2, n * 10 + 16);
                break;
            case QA_M_ALP_RD:
                alprd_layout(p, f.plen, pl);
                if (pl.ctag == 1) pl.res[0] = add_res(si, fi, f.poff + pl.coff, pl.clen, n);
                if (pl.rtag == 1) pl.res[1] = add_res(si, fi, f.poff + pl.roff, f.plen - pl.roff, n * 8 + 64);
                break;
            case QA_M_DICT: {
                dict_layout(p, f.plen, pl);
                if (pl.card > n) return QUOIN_ERR_CORRUPT;
                if (pl.vtag == 2) pl.res[0] = add_res(si, fi, f.poff + pl.voff, pl.vlen, pl.card * 8);
                if (pl.ctag_d == 1) {
                    size_t pos = pl.cpos;
                    if (pos >= f.plen) return QUOIN_ERR_CORRUPT;
                    pl.nplanes = p[pos++];
                    if (pl.nplanes < 1 || pl.nplanes > 2) return QUOIN_ERR_CORRUPT;
                    for (int k = 0; k < pl.nplanes; k++) {
                        uint64_t bl;
                        if (!get_varint(p, f.plen, pos, bl) || bl > f.plen - pos) return QUOIN_ERR_CORRUPT;
                        pl.res[1 + k] = add_res(si, fi, f.poff + pos, bl, n);
                        pos += bl;
                    }
                }
                break;
            }
            default: break;
            }
        }
    }

    // ---- entropy layer: rANS on the GPU, the other tags on host cores ----------
    uint8_t *dres = nullptr;
    uint64_t *dlane = nullptr;
    auto cleanup = [&]() {
        if (dres) hipFree(dres);
        if (dlane) hipFree(dlane);
    };
    if (res_total && hipMalloc(&dres, res_total) != hipSuccess) return QUOIN_ERR_DEVICE;
    if (!rj.empty()) {
        std::vector<QuoinRansStream> in(rj.size());
        std::vector<uint64_t> olen(rj.size(), 0);
        std::vector<int32_t> ost(rj.size(), QUOIN_OK);
        for (size_t k = 0; k < rj.size(); k++) {
            in[k].src = st[rj[k].stream].d + rj[k].off;
            in[k].src_len = rj[k].len;
            in[k].max_len = rj[k].max_len;
            in[k].out = dres + rj[k].out_off;
        }
        const int rc = quoin_gpu_decode_residuals(g, in.data(), in.size(), olen.data(), ost.data(), s);
        if (rc == QUOIN_ERR_DEVICE || rc == QUOIN_ERR_NULL) {
            cleanup();
            return QUOIN_ERR_DEVICE;
        }
        for (size_t k = 0; k < rj.size(); k++) {
            rj[k].out_len = olen[k];
            rj[k].status = ost[k];
            if (ost[k] != QUOIN_OK) {
                cleanup();
                return QUOIN_ERR_CORRUPT;
            }
        }
    }

    // ---- recurrences on the device ------------------------------------------------
    for (size_t si = 0; si < ns; si++) {
        Stream &S = st[si];
        if (!S.any_host) continue;
        for (size_t fi = 0; fi < S.frames.size(); fi++) {
            const Frame &f = S.frames[fi];
            if (!f.host) continue;
            Plan &pl = plans[si][fi];
            RecJob j{};
            j.stream = si;
            j.frame = fi;
            j.plog2 = S.plog2;
            j.n = f.n;
            switch (f.mode) {
            case M_PRED:
                j.kind = 0;
                j.src_dev = S.d + f.poff;
                j.src_len = f.plen;
                break;
            case M_PRED_RC: case M_PRED2: case M_DELTA2: case M_DELTA_DP: case M_ORDERED_DELTA:
                j.kind = f.mode == M_PRED_RC ? 0 : f.mode == M_PRED2 ? 25 : f.mode;
                j.src_dev = dres + rj[pl.res[0]].out_off;
                j.src_len = rj[pl.res[0]].out_len;
                break;
            default: continue;
            }
            j.lane_off = lane_total;
            lane_total += f.n;
            pl.rec = (int)cj.size();
            cj.push_back(j);
        }
    }
    if (lane_total && hipMalloc(&dlane, lane_total * 8) != hipSuccess) {
        cleanup();
        return QUOIN_ERR_DEVICE;
    }
    {
        // one batched call per (kind, table size)
        std::vector<char> done(cj.size(), 0);
        for (size_t a = 0; a < cj.size(); a++) {
            if (done[a]) continue;
            std::vector<size_t> grp;
            for (size_t b = a; b < cj.size(); b++)
                if (!done[b] && cj[b].kind == cj[a].kind && (cj[a].kind > 25 || cj[a].kind == 7 ||
                                                             cj[b].plog2 == cj[a].plog2)) {
                    grp.push_back(b);
                    done[b] = 1;
                }
            std::vector<QuoinLinearStream> ls(grp.size());
            std::vector<int32_t> sts(grp.size(), QUOIN_OK);
            for (size_t k = 0; k < grp.size(); k++) {
                const RecJob &j = cj[grp[k]];
                ls[k].src = j.src_dev;
                ls[k].src_len = j.src_len;
                ls[k].n = j.n;
                ls[k].out = dlane + j.lane_off;
            }
            const int kind = cj[a].kind;
            int rc = (kind == 0 || kind == 25)
                         ? quoin_gpu_pred_decode(g, kind, cj[a].plog2, ls.data(), ls.size(), sts.data(), s)
                         : quoin_gpu_linear_decode(g, kind, ls.data(), ls.size(), sts.data(), s);
            if (rc == QUOIN_ERR_DEVICE || rc == QUOIN_ERR_NULL || rc == QUOIN_ERR_BUFFER_TOO_SMALL) {
                cleanup();
                return rc == QUOIN_ERR_BUFFER_TOO_SMALL ? QUOIN_ERR_DEVICE : rc;
            }
            for (size_t k = 0; k < grp.size(); k++)
                if (sts[k] != QUOIN_OK) {
                    cleanup();
                    return QUOIN_ERR_CORRUPT;
                }
        }
    }

    // ---- back to the host ------------------------------------------------------------
    Bytes hres(res_total);
    std::vector<uint64_t> hlane(lane_total);
    if ((res_total && hipMemcpyAsync(hres.data(), dres, res_total, hipMemcpyDeviceToHost, s) != hipSuccess) ||
        (lane_total &&
         hipMemcpyAsync(hlane.data(), dlane, lane_total * 8, hipMemcpyDeviceToHost, s) != hipSuccess) ||
        hipStreamSynchronize(s) != hipSuccess) {
        cleanup();
        return QUOIN_ERR_DEVICE;
    }
    cleanup();

    // ---- per frame: the device-decodable payload ------------------------------------
    struct Out {
        int mode;
        Bytes payload;
    };
    std::vector<std::vector<Out>> fo(ns);
    std::vector<std::pair<size_t, size_t>> work;
    for (size_t si = 0; si < ns; si++) {
        if (!st[si].any_host) continue;
        fo[si].resize(st[si].frames.size());
        for (size_t fi = 0; fi < st[si].frames.size(); fi++)
            if (st[si].frames[fi].host) work.push_back({si, fi});
    }
    auto res_of = [&](int k, const uint8_t **p, size_t *len) {
        *p = hres.data() + rj[k].out_off;
        *len = (size_t)rj[k].out_len;
    };
    auto raw_of = [&](const uint64_t *lanes, size_t n, uint32_t dt, Bytes &o) {
        const size_t lb = lane_bytes(dt);
        o.resize(n * lb);
        for (size_t i = 0; i < n; i++) {
            uint64_t v = lanes[i];
            if (dt == QA_F32) {
                double d;
                memcpy(&d, &v, 8);
                const float f = (float)d;
                uint32_t b;
                memcpy(&b, &f, 4);
                v = b;
            }
            memcpy(o.data() + i * lb, &v, lb);
        }
    };
    std::vector<int> err(work.size(), QUOIN_OK);
    qahost::parallel_for(work.size(), err, [&](size_t w) {
        const size_t si = work[w].first, fi = work[w].second;
        const Stream &S = st[si];
        const Frame &f = S.frames[fi];
        const Plan &pl = plans[si][fi];
        const uint8_t *p = S.h + f.poff;
        const size_t n = f.n;
        Out &o = fo[si][fi];
        o.mode = QA_M_RAW;
        switch (f.mode) {
        case M_PRED: case M_PRED_RC: case M_PRED2: case M_DELTA2: case M_DELTA_DP: case M_ORDERED_DELTA:
            raw_of(hlane.data() + cj[pl.rec].lane_off, n, S.dtype, o.payload);
            return;
        case M_LZ: {  // lz.rs:164-171
            const uint8_t *r;
            size_t rl;
            res_of(pl.res[0], &r, &rl);
            Bytes b;
            if (!qahost::lz_decompress(r, rl, 8 * (uint64_t)n, b)) {
                err[w] = QUOIN_ERR_CORRUPT;
                return;
            }
            std::vector<uint64_t> lanes(n);
            memcpy(lanes.data(), b.data(), 8 * n);
            raw_of(lanes.data(), n, S.dtype, o.payload);
            return;
        }
        case M_BYTE_TRANSPOSE: {  // transpose.rs:21-32
            const uint8_t *r;
            size_t rl;
            res_of(pl.res[0], &r, &rl);
            if (rl != 8 * n) {
                err[w] = QUOIN_ERR_CORRUPT;
                return;
            }
            std::vector<uint64_t> lanes(n, 0);
            for (int pl8 = 0; pl8 < 8; pl8++)
                for (size_t i = 0; i < n; i++) lanes[i] |= (uint64_t)r[(size_t)pl8 * n + i] << (8 * pl8);
            raw_of(lanes.data(), n, S.dtype, o.payload);
            return;
        }
        case M_PCO: {  // pcodec.rs:66-76
            std::vector<uint64_t> lanes(n);
            if (qapco::decode_lane(p, f.plen, n, pco_type(S.dtype), lanes.data()) != 0) {
                err[w] = QUOIN_ERR_CORRUPT;
                return;
            }
            raw_of(lanes.data(), n, S.dtype, o.payload);
            return;
        }
        case QA_M_FLOAT_MULT: {  // float_mult.rs:100-133, the FM_ENTROPY arm
            const uint8_t *r;
            size_t rl;
            res_of(pl.res[0], &r, &rl);
            std::vector<uint64_t> ks(n);
            size_t pos = 0;
            uint64_t prev = 0;
            for (size_t i = 0; i < n; i++) {
                uint64_t z;
                if (!get_varint(r, rl, pos, z)) {
                    err[w] = QUOIN_ERR_CORRUPT;
                    return;
                }
                prev += unzigzag(z);
                ks[i] = prev;
            }
            if (pos != rl) {
                err[w] = QUOIN_ERR_CORRUPT;
                return;
            }
            o.mode = QA_M_FLOAT_MULT;
            o.payload.push_back(0);  // FM_BITPACK
            o.payload.push_back(p[1]);
            for_encode(ks.data(), n, true, o.payload);
            return;
        }
        case QA_M_ALP_RD: {  // alp_rd.rs:208-285
            if (pl.ctag > 1 || pl.rtag > 1) {
                err[w] = QUOIN_ERR_CORRUPT;
                return;
            }
            o.mode = QA_M_ALP_RD;
            o.payload.assign(p, p + pl.head);
            Bytes codes;
            if (pl.ctag == 1) {
                const uint8_t *r;
                size_t rl;
                res_of(pl.res[0], &r, &rl);
                if (rl != n) {
                    err[w] = QUOIN_ERR_CORRUPT;
                    return;
                }
                std::vector<uint64_t> c(r, r + n);
                for_encode(c.data(), n, false, codes);
            } else {
                codes.assign(p + pl.coff, p + pl.coff + pl.clen);
            }
            o.payload.push_back(0);
            put_varint(o.payload, codes.size());
            o.payload.insert(o.payload.end(), codes.begin(), codes.end());
            o.payload.push_back(0);
            if (pl.rtag == 1) {
                const uint8_t *r;
                size_t rl;
                res_of(pl.res[1], &r, &rl);
                o.payload.insert(o.payload.end(), r, r + rl);
            } else {
                o.payload.insert(o.payload.end(), p + pl.roff, p + f.plen);
            }
            return;
        }
        case QA_M_DICT: {  // dict.rs:154-207
            if (pl.vtag > 2 || pl.ctag_d > 1) {
                err[w] = QUOIN_ERR_CORRUPT;
                return;
            }
            o.mode = QA_M_DICT;
            put_varint(o.payload, pl.card);
            if (pl.vtag == 2) {
                const uint8_t *r;
                size_t rl;
                res_of(pl.res[0], &r, &rl);
                if (rl != pl.card * 8) {
                    err[w] = QUOIN_ERR_CORRUPT;
                    return;
                }
                o.payload.push_back(0);  // VAL_RAW
                put_varint(o.payload, pl.card * 8);
                for (uint64_t c = 0; c < pl.card; c++) {
                    uint64_t x = 0;
                    for (int b = 0; b < 8; b++) x |= (uint64_t)r[(size_t)b * pl.card + c] << (8 * b);
                    put64(o.payload, x);
                }
            } else {
                o.payload.push_back(pl.vtag);
                put_varint(o.payload, pl.vlen);
                o.payload.insert(o.payload.end(), p + pl.voff, p + pl.voff + pl.vlen);
            }
            o.payload.push_back(0);  // CODES_BITPACK
            if (pl.ctag_d == 1) {
                std::vector<uint64_t> codes(n, 0);
                for (int k = 0; k < pl.nplanes; k++) {
                    const uint8_t *r;
                    size_t rl;
                    res_of(pl.res[1 + k], &r, &rl);
                    if (rl != n) {
                        err[w] = QUOIN_ERR_CORRUPT;
                        return;
                    }
                    for (size_t i = 0; i < n; i++) codes[i] |= (uint64_t)r[i] << (8 * k);
                }
                for_encode(codes.data(), n, false, o.payload);
            } else {
                o.payload.insert(o.payload.end(), p + pl.cpos, p + f.plen);
            }
            return;
        }
        default: err[w] = QUOIN_ERR_CORRUPT;
        }
    });
    for (int e : err)
        if (e != QUOIN_OK) return QUOIN_ERR_CORRUPT;

    // ---- the rewritten streams -------------------------------------------------------
    for (size_t si = 0; si < ns; si++) {
        const Stream &S = st[si];
        if (!S.any_host) continue;
        Bytes &o = out[si];
        o.assign(S.h, S.h + S.body);
        for (size_t fi = 0; fi < S.frames.size(); fi++) {
            const Frame &f = S.frames[fi];
            if (!f.host) {
                o.insert(o.end(), S.h + f.poff - 9, S.h + f.poff + f.plen);
                continue;
            }
            const Out &x = fo[si][fi];
            o.push_back((uint8_t)x.mode);
            put32(o, f.n);
            put32(o, (uint32_t)x.payload.size());
            o.insert(o.end(), x.payload.begin(), x.payload.end());
        }
        o.insert(o.end(), S.h + S.tail, S.h + S.len);
    }
    return QUOIN_OK;
}

}  // namespace qahyb

// ===========================================================================
// Encode: the competition at Balanced / High / Max
// ===========================================================================
#include <cmath>
#include <unordered_map>

namespace qahyb {

namespace {
const double FM_SCALES[7] = {10.0, 100.0, 1000.0, 10000.0, 100000.0, 1000000.0, 10000000.0};

// encoder.rs:79-89
uint64_t decode_weight(int m) {
    switch (m) {
    case QA_M_RAW: case QA_M_CONST: case QA_M_STRIDE: return 0;
    case QA_M_XORZ: case QA_M_FOR: case QA_M_DELTA: case QA_M_RLE: return 1;
    case QA_M_ALP: case QA_M_ALP_RD: case QA_M_DICT: return 2;
    case M_PRED: return 3;
    case M_BYTE_TRANSPOSE: case QA_M_FLOAT_MULT: return 6;
    case M_ORDERED_DELTA: case M_DELTA2: case M_DELTA_DP: case M_LZ: case M_PCO: return 7;
    default: return 10;  // PredRc, Pred2
    }
}
uint64_t sat_mul(uint64_t a, uint64_t b) {
    unsigned __int128 p = (unsigned __int128)a * b;
    return p > UINT64_MAX ? UINT64_MAX : (uint64_t)p;
}
uint64_t penalty(int m, uint64_t lambda, uint64_t dec) {
    return sat_mul(sat_mul(lambda, decode_weight(m)), dec) >> 8;  // encoder.rs:94-97
}
bool looks_compressible(uint64_t res, uint64_t raw) { return sat_mul(res, 20) < sat_mul(raw, 19); }
bool cascades_lz(int m) {  // encoder.rs:564-578
    return m == M_PRED_RC || m == M_PRED2 || m == M_DELTA2 || m == M_DELTA_DP || m == M_ORDERED_DELTA ||
           m == QA_M_FLOAT_MULT || m == M_LZ || m == M_BYTE_TRANSPOSE || m == QA_M_DICT;
}

// entropy/mod.rs:90-125: order-1 size estimate, glibc log2 in the reference's order
size_t estimate_order1(const uint8_t *b, size_t n) {
    if (n < 64) return n;
    std::vector<uint32_t> joint(65536, 0);
    uint32_t marg[256] = {0};
    size_t prev = 0;
    for (size_t i = 0; i < n; i++) {
        joint[prev * 256 + b[i]]++;
        marg[prev]++;
        prev = b[i];
    }
    double bits = 0.0;
    for (int p = 0; p < 256; p++) {
        if (!marg[p]) continue;
        const double mf = (double)marg[p];
        for (int c = 0; c < 256; c++) {
            const uint32_t k = joint[(size_t)p * 256 + c];
            if (k) bits += (double)k * std::log2(mf / (double)k);
        }
    }
    return (size_t)std::ceil(bits / 8.0);
}

struct Best {
    int mode = QA_M_RAW;
    uint64_t len = 0, score = 0;
    bool host = false;
    Bytes payload;
    int runner = -1;
    uint64_t runner_score = UINT64_MAX;
    uint64_t lambda = 0, dec = 0;
    void consider(int m, uint64_t l, bool h, Bytes *p) {  // encoder.rs:126-139
        const uint64_t sc = l + penalty(m, lambda, dec);
        if (sc < score) {
            runner = mode;
            runner_score = score;
            mode = m;
            len = l;
            score = sc;
            host = h;
            if (h) payload.swap(*p);
        } else if (sc < runner_score && m != mode) {
            runner = m;
            runner_score = sc;
        }
    }
};

void le_bytes(const uint64_t *v, size_t n, Bytes &o) {
    o.resize(8 * n);
    memcpy(o.data(), v, 8 * n);
}
void transpose(const uint64_t *v, size_t n, Bytes &o) {  // transpose.rs:10-19
    o.resize(8 * n);
    for (int p = 0; p < 8; p++)
        for (size_t i = 0; i < n; i++) o[(size_t)p * n + i] = (uint8_t)(v[i] >> (8 * p));
}

// float_mult.rs:66-98 at an entropy level, with the scale the device found
bool fm_payload(const EncBlock &e, uint64_t lambda, bool allow_lz, bool entropy, Bytes &out) {
    if (e.fm_scale >= 7 || !e.n) return false;
    const double s = FM_SCALES[e.fm_scale];
    std::vector<uint64_t> ks(e.n);
    for (uint32_t i = 0; i < e.n; i++) {
        double x;
        memcpy(&x, &e.v[i], 8);
        const double k = std::round(x * s);
        ks[i] = (uint64_t)(int64_t)k;
    }
    Bytes bp;
    for_encode(ks.data(), e.n, true, bp);
    Bytes d, coded;
    bool ent = false;
    if (entropy) {
        uint64_t prev = 0;
        for (uint32_t i = 0; i < e.n; i++) {
            put_varint(d, zigzag(ks[i] - prev));
            prev = ks[i];
        }
        qahost::code_residuals(d.data(), d.size(), lambda, allow_lz, coded);
        ent = coded.size() < bp.size();
    }
    out.clear();
    out.push_back(ent ? 1 : 0);
    out.push_back((uint8_t)e.fm_scale);
    const Bytes &b = ent ? coded : bp;
    out.insert(out.end(), b.begin(), b.end());
    return true;
}

// alp_rd.rs:132-206 at an entropy level, with the cut and dictionary the device chose
bool alprd_payload(const EncBlock &e, uint64_t lambda, bool allow_lz, bool entropy, Bytes &out) {
    if (e.rd_bail || !e.n) return false;
    const uint32_t rbw = 64 - e.rd_left;
    const uint64_t mask = rbw >= 64 ? ~0ull : ((1ull << rbw) - 1);
    std::vector<uint64_t> codes(e.n), rights(e.n);
    std::vector<std::pair<uint64_t, uint64_t>> exc;
    for (uint32_t i = 0; i < e.n; i++) {
        const uint64_t left = e.v[i] >> rbw;
        rights[i] = e.v[i] & mask;
        uint32_t c = 0;
        while (c < e.rd_dl && e.rd_dict[c] != left) c++;
        if (c < e.rd_dl) {
            codes[i] = c;
        } else {
            codes[i] = 0;
            exc.push_back({i, left});
        }
    }
    if (exc.size() * 4 > e.n) return false;
    out.clear();
    out.push_back((uint8_t)e.rd_left);
    out.push_back((uint8_t)e.rd_dl);
    for (uint32_t c = 0; c < e.rd_dl; c++) put64(out, e.rd_dict[c]);
    put_varint(out, exc.size());
    for (const auto &x : exc) {
        put_varint(out, x.first);
        put64(out, x.second);
    }
    Bytes cbp, cent;
    for_encode(codes.data(), e.n, false, cbp);
    bool ce = false;
    if (entropy) {
        Bytes cb(e.n);
        for (uint32_t i = 0; i < e.n; i++) cb[i] = (uint8_t)codes[i];
        qahost::code_residuals(cb.data(), cb.size(), lambda, allow_lz, cent);
        ce = cent.size() < cbp.size();
    }
    out.push_back(ce ? 1 : 0);
    put_varint(out, ce ? cent.size() : cbp.size());
    const Bytes &c = ce ? cent : cbp;
    out.insert(out.end(), c.begin(), c.end());
    Bytes rbp, rent;
    for_encode(rights.data(), e.n, false, rbp);
    bool re = false;
    if (entropy) {
        qahost::code_residuals(rbp.data(), rbp.size(), lambda, allow_lz, rent);
        re = rent.size() < rbp.size();
    }
    out.push_back(re ? 1 : 0);
    const Bytes &r = re ? rent : rbp;
    out.insert(out.end(), r.begin(), r.end());
    return true;
}

// dict.rs:87-152 at an entropy level
bool dict_payload(const EncBlock &e, uint64_t lambda, bool allow_lz, bool entropy, Bytes &out) {
    if (!e.n) return false;
    const size_t max_card = e.n / 2 + 1;
    std::unordered_map<uint64_t, uint32_t> map;
    map.reserve(std::min<size_t>(max_card + 2, 1 << 16) * 2);
    std::vector<uint64_t> dict, codes(e.n);
    for (uint32_t i = 0; i < e.n; i++) {
        auto it = map.find(e.v[i]);
        uint32_t c;
        if (it == map.end()) {
            c = (uint32_t)dict.size();
            map.emplace(e.v[i], c);
            dict.push_back(e.v[i]);
        } else {
            c = it->second;
        }
        codes[i] = c;
        if (dict.size() > max_card) return false;
    }
    std::vector<uint32_t> order(dict.size());
    for (uint32_t i = 0; i < order.size(); i++) order[i] = i;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return dict[a] < dict[b]; });
    std::vector<uint32_t> remap(dict.size());
    std::vector<uint64_t> sorted(dict.size());
    for (uint32_t k = 0; k < order.size(); k++) {
        remap[order[k]] = k;
        sorted[k] = dict[order[k]];
    }
    for (auto &c : codes) c = remap[c];
    const size_t card = sorted.size();
    // encode_values (dict.rs:40-61)
    uint8_t vtag = 0;
    Bytes vb, alt;
    le_bytes(sorted.data(), card, vb);
    delta_encode(sorted.data(), card, alt);
    if (alt.size() < vb.size()) {
        vtag = 1;
        vb.swap(alt);
    }
    if (entropy) {
        Bytes tr;
        transpose(sorted.data(), card, tr);
        qahost::code_residuals(tr.data(), tr.size(), lambda, allow_lz, alt);
        if (alt.size() < vb.size()) {
            vtag = 2;
            vb.swap(alt);
        }
    }
    uint8_t ctag = 0;
    Bytes cb;
    for_encode(codes.data(), e.n, false, cb);
    if (entropy && card <= (1u << 16)) {
        const int np = card <= 256 ? 1 : 2;
        Bytes blob, plane(e.n), coded;
        blob.push_back((uint8_t)np);
        for (int p = 0; p < np; p++) {
            for (uint32_t i = 0; i < e.n; i++) plane[i] = (uint8_t)(codes[i] >> (8 * p));
            qahost::code_residuals(plane.data(), plane.size(), lambda, allow_lz, coded);
            put_varint(blob, coded.size());
            blob.insert(blob.end(), coded.begin(), coded.end());
        }
        if (blob.size() < cb.size()) {
            ctag = 1;
            cb.swap(blob);
        }
    }
    out.clear();
    put_varint(out, card);
    out.push_back(vtag);
    put_varint(out, vb.size());
    out.insert(out.end(), vb.begin(), vb.end());
    out.push_back(ctag);
    out.insert(out.end(), cb.begin(), cb.end());
    return true;
}

// encoder.rs:263-273
bool coded_if_competitive(const uint8_t *r, size_t n, uint64_t lambda, bool allow_lz, uint64_t best_score,
                          Bytes &out) {
    if (lambda == 0 && sat_mul(estimate_order1(r, n), 9) / 10 >= best_score) return false;
    qahost::code_residuals(r, n, lambda, allow_lz, out);
    return true;
}

// encoder.rs:543-581: one host-coded mode's payload at lambda, with the LZ
// cascade inside code_residuals when allow_lz
bool encode_host_mode(const EncBlock &e, int m, uint64_t lambda, bool allow_lz, unsigned pco_level,
                      bool entropy, Bytes &out) {
    const int k = m == M_PRED_RC ? 0 : m == M_PRED2 ? 1 : m == M_DELTA2 ? 2 : m == M_DELTA_DP ? 3 :
                  m == M_ORDERED_DELTA ? 4 : -1;
    if (k >= 0) {
        if (!e.has_res[k]) return false;
        qahost::code_residuals(e.res[k], e.rlen[k], lambda, allow_lz, out);
        return true;
    }
    Bytes tmp;
    switch (m) {
    case M_PRED:
        if (!e.has_res[0]) return false;
        out.assign(e.res[0], e.res[0] + e.rlen[0]);
        return true;
    case QA_M_FLOAT_MULT: return fm_payload(e, lambda, allow_lz, entropy, out);
    case QA_M_ALP_RD: return alprd_payload(e, lambda, allow_lz, entropy, out);
    case QA_M_DICT: return dict_payload(e, lambda, allow_lz, entropy, out);
    case M_LZ:
        le_bytes(e.v, e.n, tmp);
        {
            Bytes z;
            qahost::lz_compress(tmp.data(), tmp.size(), z);
            qahost::code_residuals(z.data(), z.size(), lambda, allow_lz, out);
        }
        return true;
    case M_BYTE_TRANSPOSE:
        transpose(e.v, e.n, tmp);
        qahost::code_residuals(tmp.data(), tmp.size(), lambda, allow_lz, out);
        return true;
    case M_PCO:
        out.clear();
        return qapco::encode_lane(e.v, e.n, pco_type(e.dt), pco_level, out);
    default: return false;
    }
}
}  // namespace

void compete(EncBlock &e, int level) {
    const bool isf = e.dt == QA_F64 || e.dt == QA_F32;
    const bool predictors = level >= 3, cascade = level == 4;
    const uint64_t lambda = level == 2 ? 2 : 0;
    const uint64_t n = e.n, raw_bytes = n * 8, dec = n * lane_bytes(e.dt);
    Best b;
    b.lambda = lambda;
    b.dec = dec;
    b.mode = QA_M_RAW;
    b.len = e.sz[C_RAW];
    b.score = b.len;  // RAW's decode weight is 0
    auto emit = [&]() {
        e.mode = b.mode;
        e.host = b.host;
        e.len = b.len;
        if (b.host) e.payload.swap(b.payload);
    };
    if (!isf && e.incompressible) {  // encoder.rs:291-298
        emit();
        return;
    }
    auto gpu = [&](int m, int c) {
        if (e.sz[c] != UINT64_MAX) b.consider(m, e.sz[c], false, nullptr);
    };
    Bytes p;
    gpu(QA_M_CONST, C_CONST);
    gpu(QA_M_STRIDE, C_STRIDE);
    gpu(QA_M_XORZ, C_XORZ);
    gpu(QA_M_FOR, C_FOR);
    gpu(QA_M_DELTA, C_DELTA);
    if (isf) {
        if (fm_payload(e, lambda, false, true, p)) b.consider(QA_M_FLOAT_MULT, p.size(), true, &p);
        gpu(QA_M_ALP, C_ALP);
        if (alprd_payload(e, lambda, false, true, p)) b.consider(QA_M_ALP_RD, p.size(), true, &p);
    }
    if (predictors) {
        if (e.has_res[0]) {
            if (looks_compressible(e.rlen[0], raw_bytes) &&
                coded_if_competitive(e.res[0], e.rlen[0], lambda, false, b.score, p))
                b.consider(M_PRED_RC, p.size(), true, &p);
            p.assign(e.res[0], e.res[0] + e.rlen[0]);
            b.consider(M_PRED, p.size(), true, &p);
        }
        if (e.has_res[1] && looks_compressible(e.rlen[1], raw_bytes) &&
            coded_if_competitive(e.res[1], e.rlen[1], lambda, false, b.score, p))
            b.consider(M_PRED2, p.size(), true, &p);
    }
    if (isf && predictors && e.has_res[2] && looks_compressible(e.rlen[2], raw_bytes)) {
        if (coded_if_competitive(e.res[2], e.rlen[2], lambda, false, b.score, p))
            b.consider(M_DELTA2, p.size(), true, &p);
        if (e.has_res[3] && coded_if_competitive(e.res[3], e.rlen[3], lambda, false, b.score, p))
            b.consider(M_DELTA_DP, p.size(), true, &p);
    }
    if (predictors && e.has_res[4] && looks_compressible(e.rlen[4], raw_bytes) &&
        coded_if_competitive(e.res[4], e.rlen[4], lambda, false, b.score, p))
        b.consider(M_ORDERED_DELTA, p.size(), true, &p);
    const bool compressible = looks_compressible(b.len, raw_bytes);
    const bool lz_worth = lambda == 0 || e.distinct_low || e.repeats;
    if (compressible && lz_worth) {
        Bytes bytes, z;
        le_bytes(e.v, n, bytes);
        qahost::lz_compress(bytes.data(), bytes.size(), z);
        qahost::code_residuals(z.data(), z.size(), lambda, false, p);
        b.consider(M_LZ, p.size(), true, &p);
    }
    if (compressible) {
        uint32_t emin = UINT32_MAX, emax = 0;  // probe_block_features (encoder.rs:705-750)
        for (uint64_t i = 0; i < n; i++) {
            const uint32_t x = (uint32_t)((e.v[i] >> 52) & 0x7FF);
            emin = std::min(emin, x);
            emax = std::max(emax, x);
        }
        const uint32_t exp_range = n ? emax - emin : 0;
        if (exp_range <= 64 || e.repeats) {
            Bytes t;
            transpose(e.v, n, t);
            if (coded_if_competitive(t.data(), t.size(), lambda, false, b.score, p))
                b.consider(M_BYTE_TRANSPOSE, p.size(), true, &p);
        }
        if (dict_payload(e, lambda, false, true, p)) b.consider(QA_M_DICT, p.size(), true, &p);
        gpu(QA_M_RLE, C_RLE);
    }
    {
        std::vector<uint8_t> pc;
        if (qapco::encode_lane(e.v, n, pco_type(e.dt), level == 4 ? 12 : 8, pc))
            b.consider(M_PCO, pc.size(), true, &pc);
    }
    if (cascade) {
        const int ms[2] = {b.mode, b.runner};
        for (int m : ms) {
            if (m < 0 || !cascades_lz(m)) continue;
            if (encode_host_mode(e, m, lambda, true, 12, true, p)) b.consider(m, p.size(), true, &p);
        }
    }
    emit();
}


namespace {
// encoder.rs:24-73: which modes may compete for a family at a level
bool is_entropy_mode(int m) {
    return m == M_PRED_RC || m == M_PRED2 || m == M_DELTA2 || m == M_DELTA_DP || m == M_ORDERED_DELTA ||
           m == QA_M_FLOAT_MULT || m == M_LZ || m == M_BYTE_TRANSPOSE;
}
bool is_predictor_mode(int m) {
    return m == M_PRED || m == M_PRED_RC || m == M_PRED2 || m == M_DELTA2 || m == M_DELTA_DP ||
           m == M_ORDERED_DELTA;
}
bool mode_runs(int m, bool isf, int level) {
    const bool fam_float = m == M_DELTA2 || m == M_DELTA_DP || m == QA_M_FLOAT_MULT || m == QA_M_ALP ||
                           m == QA_M_ALP_RD;
    if (fam_float && !isf) return false;
    if (is_predictor_mode(m)) return level >= 3;
    return level >= 2 || !is_entropy_mode(m);
}
uint64_t level_lambda(int level) {  // src/lib.rs:124-139
    static const uint64_t kLambda[5] = {16, 4, 2, 0, 0};
    return kLambda[level];
}
// the device candidate slot of a mode the device sizes and emits, else -1
int device_slot(int m) {
    switch (m) {
    case QA_M_CONST: return C_CONST;
    case QA_M_STRIDE: return C_STRIDE;
    case QA_M_XORZ: return C_XORZ;
    case QA_M_FOR: return C_FOR;
    case QA_M_DELTA: return C_DELTA;
    case QA_M_ALP: return C_ALP;
    case QA_M_RLE: return C_RLE;
    default: return -1;
    }
}
// encode_mode at a level (encoder.rs:543-581): a device mode's size, or a
// host-coded mode's payload (the level's lambda, its LZ cascade at Max)
bool mode_at_level(const EncBlock &e, int m, int level, uint64_t &len, bool &host, Bytes &out) {
    const int slot = device_slot(m);
    if (slot >= 0) {
        if (e.sz[slot] == UINT64_MAX) return false;
        len = e.sz[slot];
        host = false;
        return true;
    }
    if (!encode_host_mode(e, m, level_lambda(level), level == 4, level == 4 ? 12 : 8, level >= 2, out))
        return false;
    len = out.size();
    host = true;
    return true;
}
// encoder.rs:483-641, the sampled modes in the reference's order
const int kSampleModes[15] = {QA_M_XORZ, M_PRED, M_PRED_RC, M_PRED2, M_DELTA2, M_DELTA_DP, M_ORDERED_DELTA,
                              QA_M_FLOAT_MULT, M_BYTE_TRANSPOSE, QA_M_FOR, QA_M_DELTA, QA_M_ALP,
                              QA_M_ALP_RD, QA_M_DICT, QA_M_RLE};
}  // namespace

void compete_sampled(EncBlock &e, const EncBlock &smp, int level) {
    const bool isf = e.dt == QA_F64 || e.dt == QA_F32;
    const uint64_t n = e.n;
    Best b;
    b.lambda = level_lambda(level);
    b.dec = n * lane_bytes(e.dt);
    b.mode = QA_M_RAW;
    b.len = e.sz[C_RAW];
    b.score = b.len;  // RAW's decode weight is 0
    uint64_t len;
    bool host;
    Bytes p;
    for (int m : {QA_M_CONST, QA_M_STRIDE})
        if (mode_at_level(e, m, level, len, host, p)) b.consider(m, len, host, &p);
    if (mode_runs(M_LZ, isf, level) && (e.distinct_low || e.repeats) && mode_at_level(e, M_LZ, level, len, host, p))
        b.consider(M_LZ, len, host, &p);
    if (level >= 2 && encode_host_mode(e, M_PCO, 0, false, level == 4 ? 12 : 8, true, p))
        b.consider(M_PCO, p.size(), true, &p);
    // the winner on the sample (smallest payload, the earlier on ties), then
    // that mode on the whole block
    int win = -1;
    uint64_t win_len = UINT64_MAX;
    for (int m : kSampleModes) {
        if (!mode_runs(m, isf, level)) continue;
        if (mode_at_level(smp, m, level, len, host, p) && len < win_len) {
            win_len = len;
            win = m;
        }
    }
    if (win >= 0 && mode_at_level(e, win, level, len, host, p)) b.consider(win, len, host, &p);
    e.mode = b.mode;
    e.host = b.host;
    e.len = b.len;
    if (b.host) e.payload.swap(b.payload);
}

}  // namespace qahyb

// Test hook: one block's host competition with the device inputs supplied by
// the caller. sz[11] are the device candidate sizes (UINT64_MAX = none),
// feats = {incompressible, distinct_low, repeats}, rd = {left, dict_len,
// bail}, res/rlen the five residual streams (null = none). Writes the winner's
// mode, whether its payload is host-written, and that payload (len bytes; for
// a device-emitted mode len is its size and nothing is written). Returns 0,
// or -2 when cap is below the payload.
extern "C" int qa_hybrid_compete(int level, uint32_t dt, const uint64_t *v, uint32_t n, const uint64_t *sz,
                                 const int32_t *feats, uint32_t fm_scale, const uint32_t *rd,
                                 const uint64_t *rd_dict, const uint8_t *const *res, const uint64_t *rlen,
                                 int32_t *mode, int32_t *host, uint8_t *payload, size_t cap, size_t *len) {
    qahyb::EncBlock e;
    e.v = v;
    e.n = n;
    e.dt = dt;
    for (int j = 0; j < C_N; j++) e.sz[j] = sz[j];
    e.incompressible = feats[0] != 0;
    e.distinct_low = feats[1] != 0;
    e.repeats = feats[2] != 0;
    e.fm_scale = fm_scale;
    e.rd_left = rd[0];
    e.rd_dl = rd[1];
    e.rd_bail = rd[2];
    for (int q = 0; q < 8; q++) e.rd_dict[q] = rd_dict[q];
    for (int j = 0; j < 5; j++) {
        e.has_res[j] = res && res[j] != nullptr;
        e.res[j] = e.has_res[j] ? res[j] : nullptr;
        e.rlen[j] = e.has_res[j] ? rlen[j] : 0;
    }
    qahyb::compete(e, level);
    *mode = e.mode;
    *host = e.host ? 1 : 0;
    *len = e.len;
    if (!e.host) return 0;
    if (cap < e.payload.size()) return -2;
    if (!e.payload.empty()) memcpy(payload, e.payload.data(), e.payload.size());
    return 0;
}

// Test hook: one host-coded mode's payload (encoder.rs:543-581) at lambda
// with allow_lz, from the same device inputs as qa_hybrid_compete. Returns 1
// (payload written, *len bytes), 0 when the mode declines, -2 when cap is
// below the payload.
extern "C" int qa_hybrid_encode_mode(int mode, uint64_t lambda, int allow_lz, unsigned pco_level, uint32_t dt,
                                     const uint64_t *v, uint32_t n, uint32_t fm_scale, const uint32_t *rd,
                                     const uint64_t *rd_dict, const uint8_t *const *res, const uint64_t *rlen,
                                     uint8_t *payload, size_t cap, size_t *len) {
    qahyb::EncBlock e;
    e.v = v;
    e.n = n;
    e.dt = dt;
    e.fm_scale = fm_scale;
    e.rd_left = rd[0];
    e.rd_dl = rd[1];
    e.rd_bail = rd[2];
    for (int q = 0; q < 8; q++) e.rd_dict[q] = rd_dict[q];
    for (int j = 0; j < 5; j++) {
        e.has_res[j] = res && res[j] != nullptr;
        e.res[j] = e.has_res[j] ? res[j] : nullptr;
        e.rlen[j] = e.has_res[j] ? rlen[j] : 0;
    }
    qahyb::Bytes out;
    if (!qahyb::encode_host_mode(e, mode, lambda, allow_lz != 0, pco_level, true, out)) return 0;
    *len = out.size();
    if (cap < out.size()) return -2;
    if (!out.empty()) memcpy(payload, out.data(), out.size());
    return 1;
}
