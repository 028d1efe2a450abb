// qa_entropy.cpp — the residual entropy layer (src/entropy/mod.rs) over device
// buffers. rANS runs on the GPU (qa_rans.hip). The adaptive order-1 range coder
// (src/entropy/rc.rs) and the LZ cascade (src/codecs/lz.rs) are sequential
// byte coders, so they run here on host cores, one stream per worker thread,
// and only the streams whose level asks for them leave the device.
//
// Tags (mod.rs:24-30): 0 range coder, 1 tANS (decode only: no encoder writes it
// any more), 2 rANS, 3 LZ cascade.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>
#include <vector>

#include "../../include/quoin.h"
#include "qa_kernels.h"
#include "qa_pool.h"

namespace {

using Bytes = std::vector<uint8_t>;

constexpr uint8_t kTagRc = 0, kTagTans = 1, kTagRans = 2, kTagLz = 3;
constexpr uint64_t kWeightRc = 30, kWeightRans = 3;  // mod.rs:32-35
constexpr uint64_t kRcLambdaCutoff = 2;               // mod.rs:42

// ---- LEB128 (src/varint.rs) ------------------------------------------------
void put_varint(Bytes &o, uint64_t v) {
    while (v >= 0x80) {
        o.push_back((uint8_t)(v | 0x80));
        v >>= 7;
    }
    o.push_back((uint8_t)v);
}
bool get_varint(const uint8_t *p, size_t len, size_t &pos, uint64_t &v) {
    v = 0;
    for (uint32_t shift = 0;; shift += 7) {
        if (shift >= 64 || pos >= len) return false;
        const uint8_t b = p[pos++];
        v |= (uint64_t)(b & 0x7F) << shift;
        if (!(b & 0x80)) return true;
    }
}

uint64_t sat_mul(uint64_t a, uint64_t b) {
    return a && b > UINT64_MAX / a ? UINT64_MAX : a * b;
}
// decode-cost penalty of a coder (mod.rs:63-64)
uint64_t coder_penalty(uint64_t lambda, uint64_t w, uint64_t n) {
    return sat_mul(sat_mul(lambda, w), n) >> 8;
}

// ---- range coder (src/entropy/rc.rs) ---------------------------------------
// LZMA-style binary coder: 11-bit probabilities of a 0 bit, adapted by 1/32,
// 32-bit range renormalized a byte at a time below 2^24, carries resolved
// through a cached byte and a count of pending 0xFF bytes. The byte model
// walks a 256-node bit tree (MSB first) chosen by the previous byte.
constexpr uint32_t kProbBits = 11, kMoveBits = 5, kTop = 1u << 24;

struct ByteModel {
    std::unique_ptr<uint16_t[]> p{new uint16_t[256 * 256]};
    ByteModel() { std::fill(p.get(), p.get() + 256 * 256, (uint16_t)(1u << (kProbBits - 1))); }
};

class RcWriter {
  public:
    explicit RcWriter(Bytes &out) : out_(out) {}
    void bit(uint16_t &prob, uint32_t b) {
        const uint32_t bound = (range_ >> kProbBits) * prob;
        if (b) {
            low_ += bound;
            range_ -= bound;
            prob = (uint16_t)(prob - (prob >> kMoveBits));
        } else {
            range_ = bound;
            prob = (uint16_t)(prob + (((1u << kProbBits) - prob) >> kMoveBits));
        }
        while (range_ < kTop) {
            range_ <<= 8;
            shift_low();
        }
    }
    void finish() {
        for (int i = 0; i < 5; i++) shift_low();
    }

  private:
    void shift_low() {
        if (low_ < 0xFF000000ull || (low_ >> 32)) {
            const uint8_t carry = (uint8_t)(low_ >> 32);
            out_.push_back((uint8_t)(cache_ + carry));
            for (; pending_ > 1; pending_--) out_.push_back((uint8_t)(0xFF + carry));
            pending_ = 0;
            cache_ = (uint8_t)(low_ >> 24);
        }
        pending_++;
        low_ = (low_ & 0x00FFFFFFull) << 8;
    }
    Bytes &out_;
    uint64_t low_ = 0, pending_ = 1;
    uint32_t range_ = 0xFFFFFFFFu;
    uint8_t cache_ = 0;
};

// varint(len) ++ coded bytes (rc.rs:206-219)
void rc_encode(const uint8_t *src, size_t n, Bytes &out) {
    put_varint(out, n);
    ByteModel m;
    RcWriter w(out);
    uint32_t ctx = 0;
    for (size_t i = 0; i < n; i++) {
        uint16_t *tree = m.p.get() + ctx * 256;
        uint32_t node = 1;
        for (int k = 7; k >= 0; k--) {
            const uint32_t b = (src[i] >> k) & 1;
            w.bit(tree[node], b);
            node = node << 1 | b;
        }
        ctx = src[i];
    }
    w.finish();
}

// rc.rs:221-243 (decoder rc.rs:94-150): a skipped first byte, four code bytes,
// bytes past the end read as 0. Fails only on the length varint or bound.
bool rc_decode(const uint8_t *src, size_t len, uint64_t max_len, Bytes &out) {
    size_t pos = 0;
    uint64_t n;
    if (!get_varint(src, len, pos, n) || n > max_len) return false;
    auto next = [&]() -> uint32_t { return pos < len ? src[pos++] : (pos++, 0u); };
    next();
    uint32_t code = 0, range = 0xFFFFFFFFu;
    for (int i = 0; i < 4; i++) code = code << 8 | next();
    ByteModel m;
    out.resize(n);
    uint32_t ctx = 0;
    for (uint64_t i = 0; i < n; i++) {
        uint16_t *tree = m.p.get() + ctx * 256;
        uint32_t node = 1;
        for (int k = 0; k < 8; k++) {
            uint16_t &prob = tree[node];
            const uint32_t bound = (range >> kProbBits) * prob;
            const uint32_t b = code >= bound;
            if (b) {
                code -= bound;
                range -= bound;
                prob = (uint16_t)(prob - (prob >> kMoveBits));
            } else {
                range = bound;
                prob = (uint16_t)(prob + (((1u << kProbBits) - prob) >> kMoveBits));
            }
            while (range < kTop) {
                range <<= 8;
                code = code << 8 | next();
            }
            node = node << 1 | b;
        }
        out[i] = (uint8_t)node;
        ctx = node & 0xFF;
    }
    return true;
}

// ---- LZSS over bytes (src/codecs/lz.rs) ------------------------------------
// Greedy parse: at each position the longest match among the last 64 earlier
// positions with the same 4-byte hash (17 bits, multiplicative), the nearest
// on ties; a candidate is scanned only when it can beat the current best at
// the best length. Matches of 4 or more are taken. Framing: varint(n), then
// per 8 tokens a control byte (MSB first, 1 = match) and the tokens: a literal
// byte, or varint(offset) and varint(length - 4).
constexpr size_t kMinMatch = 4, kMaxChain = 64;
constexpr uint32_t kHashBits = 17;

uint32_t hash4(const uint8_t *w) {
    uint32_t v;
    std::memcpy(&v, w, 4);  // little-endian host
    return (v * 2654435761u) >> (32 - kHashBits);
}

void lz_encode(const uint8_t *in, size_t n, Bytes &out) {
    std::vector<int64_t> head((size_t)1 << kHashBits, -1), chain(n ? n : 1, -1);
    auto insert = [&](size_t i) {
        if (i + kMinMatch <= n) {
            const uint32_t h = hash4(in + i);
            chain[i] = head[h];
            head[h] = (int64_t)i;
        }
    };
    put_varint(out, n);
    size_t ctrl_at = 0, ntok = 0;
    for (size_t i = 0; i < n;) {
        size_t best = 0, best_off = 0;
        if (i + kMinMatch <= n) {
            const size_t lim = n - i;
            int64_t c = head[hash4(in + i)];
            for (size_t k = 0; c >= 0 && k < kMaxChain; k++, c = chain[(size_t)c]) {
                const size_t cp = (size_t)c;
                if (best && !(cp + best < n && in[cp + best] == in[i + best])) continue;
                size_t l = 0;
                while (l < lim && in[cp + l] == in[i + l]) l++;
                if (l > best) {
                    best = l;
                    best_off = i - cp;
                    if (l == lim) break;
                }
            }
        }
        if (ntok % 8 == 0) {
            ctrl_at = out.size();
            out.push_back(0);
        }
        if (best >= kMinMatch) {
            out[ctrl_at] |= (uint8_t)(0x80u >> (ntok % 8));
            put_varint(out, best_off);
            put_varint(out, best - kMinMatch);
            for (size_t e = i + best; i < e; i++) insert(i);
        } else {
            out.push_back(in[i]);
            insert(i);
            i++;
        }
        ntok++;
    }
}

// lz.rs:115-148; a match that would overrun the output is corrupt
bool lz_decode(const uint8_t *s, size_t len, uint64_t expected, Bytes &out) {
    size_t pos = 0;
    uint64_t n;
    if (!get_varint(s, len, pos, n) || n != expected) return false;
    out.resize(n);
    uint64_t o = 0;
    while (o < n) {
        if (pos >= len) return false;
        const uint8_t ctrl = s[pos++];
        for (int k = 0; k < 8 && o < n; k++) {
            if (ctrl & (0x80u >> k)) {
                uint64_t off, l;
                if (!get_varint(s, len, pos, off) || !get_varint(s, len, pos, l)) return false;
                if (off == 0 || off > o || n - o < kMinMatch || l > n - o - kMinMatch) return false;
                for (uint64_t e = o + l + kMinMatch; o < e; o++) out[o] = out[o - off];
            } else {
                if (pos >= len) return false;
                out[o++] = s[pos++];
            }
        }
    }
    return true;
}

// ---- rANS on the host (src/entropy/rans.rs) --------------------------------
// Only the LZ cascade's inner stream, which this host path produces and reads,
// is rANS-coded here; every other rANS stream is coded on the GPU.
constexpr uint32_t kM = 4096, kL = 1u << 23;

void rans_freqs(const uint8_t *d, size_t n, uint32_t f[256]) {
    uint64_t cnt[256] = {0};
    for (size_t i = 0; i < n; i++) cnt[d[i]]++;
    uint32_t sum = 0;
    for (int s = 0; s < 256; s++) {
        f[s] = cnt[s] ? std::max<uint32_t>(1, (uint32_t)(cnt[s] * kM / n)) : 0;
        sum += f[s];
    }
    auto last_max = [&](bool above_one) {  // the last index on ties (rans.rs:25-50)
        int b = -1;
        for (int s = 0; s < 256; s++)
            if ((!above_one || f[s] > 1) && (b < 0 || f[s] >= f[b])) b = s;
        return b;
    };
    for (; sum > kM; sum--) f[last_max(true)]--;
    if (sum < kM) f[last_max(false)] += kM - sum;
}

void rans_encode(const uint8_t *d, size_t n, Bytes &out) {  // n > 0
    uint32_t f[256], c[257];
    rans_freqs(d, n, f);
    c[0] = 0;
    for (int s = 0; s < 256; s++) c[s + 1] = c[s] + f[s];
    Bytes back(2 * n + 64);
    size_t p = back.size();
    uint32_t x[4] = {kL, kL, kL, kL};
    for (size_t i = n; i-- > 0;) {
        const uint32_t s = d[i], fs = f[s];
        uint32_t &xi = x[i & 3];
        for (const uint32_t lim = ((kL >> 12) << 8) * fs; xi >= lim; xi >>= 8) back[--p] = (uint8_t)xi;
        xi = ((xi / fs) << 12) + xi % fs + c[s];
    }
    for (int k = 3; k >= 0; k--)
        for (int b = 3; b >= 0; b--) back[--p] = (uint8_t)(x[k] >> (8 * b));
    put_varint(out, n);
    int present = 0;
    for (int s = 0; s < 256; s++) present += f[s] > 0;
    put_varint(out, (uint64_t)present);
    for (int s = 0; s < 256; s++)
        if (f[s]) {
            out.push_back((uint8_t)s);
            put_varint(out, f[s]);
        }
    out.insert(out.end(), back.begin() + (ptrdiff_t)p, back.end());
}

// rans.rs:110-175 (u32 arithmetic wraps as in the reference's release build)
bool rans_decode(const uint8_t *s, size_t len, uint64_t max_len, Bytes &out) {
    size_t pos = 0;
    uint64_t n, ns;
    if (!get_varint(s, len, pos, n) || n > max_len) return false;
    if (!get_varint(s, len, pos, ns) || ns == 0 || ns > 256) return false;
    uint32_t f[256] = {0}, sum = 0;
    for (uint64_t k = 0; k < ns; k++) {
        if (pos >= len) return false;
        const uint8_t sym = s[pos++];
        uint64_t fv;
        if (!get_varint(s, len, pos, fv)) return false;
        const uint32_t fq = (uint32_t)fv;
        if (fq == 0 || fq > kM) return false;
        f[sym] = fq;
        sum += fq;
    }
    if (sum != kM) return false;
    uint32_t c[257];
    c[0] = 0;
    for (int k = 0; k < 256; k++) c[k + 1] = c[k] + f[k];
    std::vector<uint8_t> sym_of(kM, 0);
    for (int k = 0; k < 256; k++)
        for (uint32_t sl = c[k]; sl < c[k + 1] && sl < kM; sl++) sym_of[sl] = (uint8_t)k;
    uint32_t x[4];
    for (auto &xi : x) {
        if (len - pos < 4) return false;
        xi = (uint32_t)s[pos] | (uint32_t)s[pos + 1] << 8 | (uint32_t)s[pos + 2] << 16 |
             (uint32_t)s[pos + 3] << 24;
        pos += 4;
    }
    out.resize(n);
    for (uint64_t i = 0; i < n; i++) {
        uint32_t &xi = x[i & 3];
        const uint32_t slot = xi & (kM - 1);
        const uint8_t sym = sym_of[slot];
        out[i] = sym;
        uint32_t xn = f[sym] * (xi >> 12) + slot - c[sym];
        while (xn < kL) {
            if (pos >= len) return false;
            xn = xn << 8 | s[pos++];
        }
        xi = xn;
    }
    return true;
}

// ---- tANS decode (src/entropy/tans.rs:197-303) ------------------------------
// The legacy TAG_TANS coder: no encoder writes it any more, the reference still
// decodes it. 1024-state table; symbols are spread with the odd stride
// 3 * 1024 / 5 (+1) over the weights in header order, and each table entry's
// node gives the symbol, the renorm bit count and the next state's base.
// Renorm bits are read MSB first (src/bitio.rs), 0 past the end.
constexpr uint32_t kTansLog = 10, kTansSize = 1u << kTansLog;

bool tans_decode(const uint8_t *s, size_t len, uint64_t max_len, Bytes &out) {
    size_t pos = 0;
    uint64_t n;
    if (!get_varint(s, len, pos, n) || n > max_len || pos >= len) return false;
    const uint32_t num = s[pos++] + 1u;
    if (len - pos < num) return false;
    const uint8_t *which = s + pos;
    pos += num;
    if (len - pos < 2ull * num + 4) return false;
    uint32_t weight[256], xs[256] = {0}, wsum = 0;
    bool seen[256] = {false};
    for (uint32_t k = 0; k < num; k++, pos += 2) {
        weight[k] = (uint32_t)s[pos] | (uint32_t)s[pos + 1] << 8;
        if (!weight[k] || seen[which[k]]) return false;
        seen[which[k]] = true;
        wsum += weight[k];
        xs[which[k]] = weight[k];
    }
    uint32_t state = (uint32_t)s[pos] | (uint32_t)s[pos + 1] << 8 | (uint32_t)s[pos + 2] << 16 |
                     (uint32_t)s[pos + 3] << 24;
    pos += 4;
    // the reference checks the state before the model
    if (state < kTansSize || state >= 2 * kTansSize || wsum != kTansSize) return false;
    uint8_t sym[kTansSize], nbits[kTansSize];
    uint16_t nbase[kTansSize];
    uint32_t stride = 3 * kTansSize / 5 | 1u, step = 0;
    for (uint32_t k = 0; k < num; k++)
        for (uint32_t r = 0; r < weight[k]; r++, step++) sym[(stride * step) & (kTansSize - 1)] = which[k];
    for (uint32_t i = 0; i < kTansSize; i++) {
        const uint32_t base = xs[sym[i]]++;
        const uint32_t bits = (uint32_t)__builtin_clz(base) - (uint32_t)__builtin_clz(kTansSize);
        nbits[i] = (uint8_t)bits;
        nbase[i] = (uint16_t)((base << bits) - kTansSize);
    }
    out.resize(n);
    uint64_t acc = 0;
    uint32_t have = 0;
    for (uint64_t k = 0; k < n; k++) {
        const uint32_t idx = state - kTansSize, bits = nbits[idx];
        uint32_t extra = 0;
        if (bits) {
            for (; have < bits; have += 8) acc = acc << 8 | (pos < len ? s[pos++] : (pos++, 0u));
            have -= bits;
            extra = (uint32_t)(acc >> have) & ((1u << bits) - 1);
            acc &= (1ull << have) - 1;
        }
        out[k] = sym[idx];
        state = kTansSize + nbase[idx] + extra;
    }
    return true;
}

// ---- the tag layer ---------------------------------------------------------
// entropy_pick (mod.rs:45-78) with every coder on the host
void pick_host(const uint8_t *d, size_t n, uint64_t lambda, Bytes &out) {
    Bytes r, c;
    if (n) rans_encode(d, n, r);
    if (lambda >= kRcLambdaCutoff && n) {
        out.push_back(kTagRans);
        out.insert(out.end(), r.begin(), r.end());
        return;
    }
    rc_encode(d, n, c);
    const bool rans = n && r.size() + coder_penalty(lambda, kWeightRans, n) <
                               c.size() + coder_penalty(lambda, kWeightRc, n);
    out.push_back(rans ? kTagRans : kTagRc);
    const Bytes &w = rans ? r : c;
    out.insert(out.end(), w.begin(), w.end());
}

// decode_entropy_only (mod.rs:151-161); QUOIN_ERR_* or QUOIN_OK
int entropy_only_host(const uint8_t *p, size_t len, uint64_t max_len, Bytes &out) {
    if (!len) return QUOIN_ERR_CORRUPT;
    switch (p[0]) {
    case kTagRc: return rc_decode(p + 1, len - 1, max_len, out) ? QUOIN_OK : QUOIN_ERR_CORRUPT;
    case kTagRans: return rans_decode(p + 1, len - 1, max_len, out) ? QUOIN_OK : QUOIN_ERR_CORRUPT;
    case kTagTans: return tans_decode(p + 1, len - 1, max_len, out) ? QUOIN_OK : QUOIN_ERR_CORRUPT;
    default: return QUOIN_ERR_CORRUPT;
    }
}
// decode_residuals (mod.rs:163-178)
int residuals_host(const uint8_t *p, size_t len, uint64_t max_len, Bytes &out) {
    if (len && p[0] == kTagLz) {
        size_t pos = 1;
        uint64_t orig;
        if (!get_varint(p, len, pos, orig) || orig > max_len) return QUOIN_ERR_CORRUPT;
        const uint64_t bound = orig > (UINT64_MAX - 64) / 2 ? UINT64_MAX : 2 * orig + 64;
        Bytes z;
        const int rc = entropy_only_host(p + pos, len - pos, bound, z);
        if (rc != QUOIN_OK) return rc;
        return lz_decode(z.data(), z.size(), orig, out) ? QUOIN_OK : QUOIN_ERR_CORRUPT;
    }
    return entropy_only_host(p, len, max_len, out);
}

// f(i) for i < n over up to 16 host threads; a thrown exception turns into
// QUOIN_ERR_PANIC for its item
template <class F>
void host_for(size_t n, std::vector<int> &err, F f) {
    qapool::run(n, [&](size_t i) {
        try {
            f(i);
        } catch (...) {
            err[i] = QUOIN_ERR_PANIC;
        }
    });
}

// one byte at each of the given device addresses, in kernel-argument chunks
void put_bytes(const std::vector<std::pair<uint8_t *, uint8_t>> &puts, hipStream_t s) {
    for (size_t i = 0; i < puts.size(); i += QA_PUT_MAX) {
        BytePuts b{};
        b.n = (uint32_t)std::min<size_t>(QA_PUT_MAX, puts.size() - i);
        for (uint32_t k = 0; k < b.n; k++) {
            b.p[k] = puts[i + k].first;
            b.v[k] = puts[i + k].second;
        }
        qa_put_bytes_launch(b, s);
    }
}

}  // namespace

extern "C" size_t quoin_residual_bound(size_t n) { return n ? 1 + quoin_rans_bound(n) : 7; }

extern "C" int quoin_gpu_code_residuals(QuoinGpu *g, uint64_t lambda, int allow_lz,
                                        const QuoinRansInput *in, size_t ns, uint64_t *out_lens,
                                        void *stream) {
    if (!g || (ns && (!in || !out_lens))) return QUOIN_ERR_NULL;
    for (size_t i = 0; i < ns; i++) {
        out_lens[i] = 0;
        if (!in[i].dst || (in[i].n && !in[i].src)) return QUOIN_ERR_NULL;
        if (in[i].dst_cap < quoin_residual_bound(in[i].n)) {
            out_lens[i] = quoin_residual_bound(in[i].n);
            return QUOIN_ERR_BUFFER_TOO_SMALL;
        }
    }
    if (!ns) return QUOIN_OK;
    // rANS of every non-empty stream on the GPU, behind the tag byte
    std::vector<QuoinRansInput> r(in, in + ns);
    for (auto &x : r) {
        x.dst += 1;
        x.dst_cap -= 1;
    }
    std::vector<uint64_t> rlen(ns, 0);
    int rc = quoin_gpu_rans_compress(g, r.data(), ns, rlen.data(), stream);
    if (rc != QUOIN_OK) return rc;
    if (hipSetDevice(qa_gpu_device(g)) != hipSuccess) return QUOIN_ERR_DEVICE;
    hipStream_t s = qa_gpu_stream(g, stream);
    // the streams that need the host: the range coder competes below the
    // cutoff (and codes the empty stream), the cascade runs with allow_lz
    std::vector<size_t> host;
    for (size_t i = 0; i < ns; i++)
        if (lambda < kRcLambdaCutoff || in[i].n == 0 || (allow_lz && in[i].n > 64)) host.push_back(i);
    std::vector<Bytes> src(host.size()), res(host.size());
    for (size_t k = 0; k < host.size(); k++) {
        src[k].resize(in[host[k]].n);
        if (in[host[k]].n &&
            hipMemcpyAsync(src[k].data(), in[host[k]].src, in[host[k]].n, hipMemcpyDeviceToHost, s) !=
                hipSuccess)
            return QUOIN_ERR_DEVICE;
    }
    if (hipStreamSynchronize(s) != hipSuccess) return QUOIN_ERR_DEVICE;
    std::vector<int> err(host.size(), QUOIN_OK);
    host_for(host.size(), err, [&](size_t k) {
        const size_t i = host[k], n = in[i].n;
        const uint8_t *d = src[k].data();
        // entropy_pick with the GPU's rANS (res stays empty when it wins)
        uint64_t best = 1 + rlen[i];
        bool gpu_rans = n > 0;
        if (lambda < kRcLambdaCutoff || n == 0) {
            Bytes c{kTagRc};
            rc_encode(d, n, c);
            if (!(n && rlen[i] + coder_penalty(lambda, kWeightRans, n) <
                           (c.size() - 1) + coder_penalty(lambda, kWeightRc, n))) {
                gpu_rans = false;
                best = c.size();
                res[k] = std::move(c);
            }
        }
        if (allow_lz && n > 64) {  // mod.rs:121-149
            Bytes z;
            lz_encode(d, n, z);
            if (z.size() < n) {
                Bytes cand{kTagLz};
                put_varint(cand, n);
                pick_host(z.data(), z.size(), lambda, cand);
                if (cand.size() < best) {
                    res[k] = std::move(cand);
                    gpu_rans = false;
                }
            }
        }
        if (gpu_rans) res[k].clear();
    });
    for (int e : err)
        if (e != QUOIN_OK) return e;
    std::vector<char> from_host(ns, 0);
    for (size_t k = 0; k < host.size(); k++) {
        if (res[k].empty()) continue;
        const size_t i = host[k];
        from_host[i] = 1;
        out_lens[i] = res[k].size();
        if (hipMemcpyAsync(in[i].dst, res[k].data(), res[k].size(), hipMemcpyHostToDevice, s) !=
            hipSuccess)
            return QUOIN_ERR_DEVICE;
    }
    std::vector<std::pair<uint8_t *, uint8_t>> tags;
    for (size_t i = 0; i < ns; i++)
        if (!from_host[i]) {
            tags.emplace_back(in[i].dst, kTagRans);
            out_lens[i] = 1 + rlen[i];
        }
    put_bytes(tags, s);
    if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess)
        return QUOIN_ERR_DEVICE;
    return QUOIN_OK;
}

extern "C" int quoin_gpu_decode_residuals(QuoinGpu *g, const QuoinRansStream *in, size_t ns,
                                          uint64_t *out_lens, int32_t *status, void *stream) {
    if (!g || (ns && (!in || !out_lens))) return QUOIN_ERR_NULL;
    if (!ns) return QUOIN_OK;
    // TAG_RANS streams decode on the GPU; the others come back UNSUPPORTED
    std::vector<int32_t> st(ns, QUOIN_OK);
    int rc = quoin_gpu_rans_decompress(g, in, ns, 1, out_lens, st.data(), stream);
    if (rc == QUOIN_ERR_NULL || rc == QUOIN_ERR_DEVICE) return rc;
    if (hipSetDevice(qa_gpu_device(g)) != hipSuccess) return QUOIN_ERR_DEVICE;
    hipStream_t s = qa_gpu_stream(g, stream);
    std::vector<size_t> host;
    for (size_t i = 0; i < ns; i++)
        if (st[i] == QUOIN_ERR_UNSUPPORTED) host.push_back(i);
    std::vector<Bytes> src(host.size()), res(host.size());
    for (size_t k = 0; k < host.size(); k++) {
        src[k].resize(in[host[k]].src_len);
        if (hipMemcpyAsync(src[k].data(), in[host[k]].src, src[k].size(), hipMemcpyDeviceToHost,
                           s) != hipSuccess)
            return QUOIN_ERR_DEVICE;
    }
    if (hipStreamSynchronize(s) != hipSuccess) return QUOIN_ERR_DEVICE;
    std::vector<int> err(host.size(), QUOIN_OK);
    host_for(host.size(), err, [&](size_t k) {
        err[k] = residuals_host(src[k].data(), src[k].size(), in[host[k]].max_len, res[k]);
    });
    for (size_t k = 0; k < host.size(); k++) {
        const size_t i = host[k];
        st[i] = err[k];
        out_lens[i] = err[k] == QUOIN_OK ? res[k].size() : 0;
        if (err[k] == QUOIN_OK && !res[k].empty() &&
            hipMemcpyAsync(in[i].out, res[k].data(), res[k].size(), hipMemcpyHostToDevice, s) !=
                hipSuccess)
            return QUOIN_ERR_DEVICE;
    }
    if (hipStreamSynchronize(s) != hipSuccess) return QUOIN_ERR_DEVICE;
    rc = QUOIN_OK;
    for (size_t i = 0; i < ns; i++) {
        if (status) status[i] = st[i];
        if (rc == QUOIN_OK) rc = st[i];
    }
    return rc;
}

// ---------------------------------------------------------------------------
// the host coders for the engine's entropy levels (qa_hybrid.cpp)
// ---------------------------------------------------------------------------
#include "qa_hostcoders.h"

namespace qahost {

// code_residuals (mod.rs:121-149) with every coder on host cores
void code_residuals(const uint8_t *d, size_t n, uint64_t lambda, bool allow_lz, std::vector<uint8_t> &out) {
    out.clear();
    pick_host(d, n, lambda, out);
    if (allow_lz && n > 64) {
        Bytes z;
        lz_encode(d, n, z);
        if (z.size() < n) {
            Bytes cand;
            cand.push_back(kTagLz);
            put_varint(cand, n);
            pick_host(z.data(), z.size(), lambda, cand);
            if (cand.size() < out.size()) out.swap(cand);
        }
    }
}
int decode_residuals(const uint8_t *p, size_t len, uint64_t max_len, std::vector<uint8_t> &out) {
    out.clear();
    return residuals_host(p, len, max_len, out);
}
void lz_compress(const uint8_t *d, size_t n, std::vector<uint8_t> &out) {
    out.clear();
    lz_encode(d, n, out);
}
bool lz_decompress(const uint8_t *s, size_t len, uint64_t expected, std::vector<uint8_t> &out) {
    out.clear();
    return lz_decode(s, len, expected, out);
}
void parallel_for(size_t n, std::vector<int> &err, const std::function<void(size_t)> &f) {
    host_for(n, err, f);
}

}  // namespace qahost

// Test hooks of the host coders (no device work): code_residuals with every
// coder on host cores, and decode_residuals. 0 = ok (*len bytes written),
// -2 = cap below the result (*len = its size), QUOIN_ERR_CORRUPT for a stream
// decode_residuals refuses.
extern "C" int qa_host_code_residuals(const uint8_t *d, size_t n, uint64_t lambda, int allow_lz,
                                      uint8_t *out, size_t cap, size_t *len) {
    std::vector<uint8_t> o;
    qahost::code_residuals(d, n, lambda, allow_lz != 0, o);
    *len = o.size();
    if (o.size() > cap) return -2;
    if (!o.empty()) memcpy(out, o.data(), o.size());
    return 0;
}
extern "C" int qa_host_decode_residuals(const uint8_t *p, size_t plen, uint64_t max_len, uint8_t *out,
                                        size_t cap, size_t *len) {
    std::vector<uint8_t> o;
    const int rc = qahost::decode_residuals(p, plen, max_len, o);
    if (rc != QUOIN_OK) return rc;
    *len = o.size();
    if (o.size() > cap) return -2;
    if (!o.empty()) memcpy(out, o.data(), o.size());
    return 0;
}
