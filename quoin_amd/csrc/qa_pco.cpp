// qa_pco.cpp — pco (the vendored pcodec 1.0.2 fork) on host cores: the
// standalone compressor at ChunkConfig::default() with a compression level
// (the only configuration quoin uses, src/codecs/pcodec.rs:33-49) and the
// standalone decompressor.
//
// Layout of the bytes (vendor/quoin-pco/src/standalone/compressor.rs,
// wrapped/chunk_compressor.rs, metadata/*.rs): a standalone header ("pco!",
// version 3, uniform-type byte 0, a bit varint of n), the wrapped header
// (format 4.1), then per chunk a type byte, n-1 in 24 bits, the chunk
// metadata (mode, delta encoding, bins of every latent variable) and one page
// (delta state and final ANS states per latent variable, then batches of 256
// numbers: each variable's ANS bits, then its offset bits), and the
// termination byte 0. Bits are written least-significant first.
//
// Every floating-point step is restated in the reference's operation order:
// f32 in the bin optimizer and the weight quantizer, f64 with glibc log2,
// log10, pow and cbrt elsewhere, no contraction. Two sums in the reference
// run over a std HashMap, whose iteration order Rust randomizes per process
// (sampling.rs:86-120, int_mult.rs:141-158); they are taken here in
// first-insertion order. The float-mult savings are integers, so their sum
// is exact in any order; the others can differ from a given reference run
// only in the last bit of a sum compared against a threshold.
#include "qa_pco.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <type_traits>
#include <unordered_map>
#include <utility>

namespace {

using Bitlen = uint32_t;
constexpr size_t BATCH = 256;        // FULL_BATCH_N (constants.rs:66)
constexpr int INTERLEAVE = 4;        // ANS_INTERLEAVING
constexpr Bitlen MAX_ANS_BITS = 14;
constexpr Bitlen MAX_LEVEL = 12;
constexpr size_t MAX_ENTRIES = 1u << 24;
constexpr size_t MAX_PAGE_N = 1u << 18;  // DEFAULT_MAX_PAGE_N
constexpr Bitlen LIMITED_UBL = 6;        // LIMITED_UNOPTIMIZED_BINS_LOG
constexpr size_t DELTA_MAX_BITS = 4 + 5 + 5 + 64 + 32 * 32;  // DeltaEncoding::MAX_BIT_SIZE

// ---------------------------------------------------------------------------
// bit helpers (bits.rs)
// ---------------------------------------------------------------------------
template <class L> constexpr Bitlen BITS = (Bitlen)(sizeof(L) * 8);
template <class L> constexpr L MID = (L)((L)1 << (BITS<L> - 1));
template <class L> constexpr L LMAX = (L)~(L)0;
template <class L> inline Bitlen lzc(L x) {
    if (!x) return BITS<L>;
    return sizeof(L) == 8 ? (Bitlen)__builtin_clzll((unsigned long long)x)
                          : (Bitlen)__builtin_clz((unsigned)x);
}
template <class L> inline Bitlen tzc(L x) {
    if (!x) return BITS<L>;
    return sizeof(L) == 8 ? (Bitlen)__builtin_ctzll((unsigned long long)x)
                          : (Bitlen)__builtin_ctz((unsigned)x);
}
template <class L> inline Bitlen bte_offset(L max_offset) { return BITS<L> - lzc(max_offset); }
// bits_to_encode_offset_bits: 6 for 32-bit latents, 7 for 64-bit
template <class L> constexpr Bitlen OBB = BITS<L> == 32 ? 6 : 7;
inline Bitlen ilog2_64(uint64_t x) { return 63 - (Bitlen)__builtin_clzll(x); }
inline uint64_t low_bits(uint64_t x, Bitlen n) { return n >= 64 ? x : x & ((1ull << n) - 1); }

// ---------------------------------------------------------------------------
// Xoroshiro128++ seeded from a u64 through SplitMix64 (rand_xoshiro 0.6.0
// seed_from_u64, Blackman and Vigna's published generators)
// ---------------------------------------------------------------------------
struct Xoro {
    uint64_t s0, s1;
    explicit Xoro(uint64_t seed) {
        uint64_t x = seed;
        auto sm = [&x]() {
            x += 0x9e3779b97f4a7c15ull;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            return z ^ (z >> 31);
        };
        s0 = sm();
        s1 = sm();
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next() {
        const uint64_t a = s0;
        uint64_t b = s1;
        const uint64_t r = rotl(a + b, 17) + a;
        b ^= a;
        s0 = rotl(a, 49) ^ b ^ (b << 21);
        s1 = rotl(b, 28);
        return r;
    }
};

// ---------------------------------------------------------------------------
// bit writer / reader (bit_writer.rs, bit_reader.rs): LSB-first
// ---------------------------------------------------------------------------
struct BitWriter {
    std::vector<uint8_t> &out;
    uint64_t acc = 0;
    Bitlen nacc = 0;
    explicit BitWriter(std::vector<uint8_t> &o) : out(o) {}
    void put(uint64_t x, Bitlen n) {
        if (n == 0) return;
        if (n > 56) {
            put(x & 0xFFFFFFFFull, 32);
            put(x >> 32, n - 32);
            return;
        }
        acc |= (x & ((1ull << n) - 1)) << nacc;
        nacc += n;
        while (nacc >= 8) {
            out.push_back((uint8_t)acc);
            acc >>= 8;
            nacc -= 8;
        }
    }
    void finish_byte() {
        if (nacc) {
            out.push_back((uint8_t)acc);
            acc = 0;
            nacc = 0;
        }
    }
    void byte(uint8_t b) { out.push_back(b); }  // only when aligned
};

struct BitReader {
    const uint8_t *p;  // padded with >= 16 zero bytes past len
    size_t len;
    uint64_t bit = 0;
    uint64_t u64_at(size_t byte) const {
        uint64_t v;
        memcpy(&v, p + byte, 8);
        return v;
    }
    uint64_t get(Bitlen n) {
        if (n > 56) {
            const uint64_t lo = get(32);
            const uint64_t hi = get(n - 32);
            return lo | (hi << 32);
        }
        if (n == 0) return 0;
        const uint64_t b = bit;
        bit += n;
        if (bit > (uint64_t)len * 8) return 0;  // overrun: ok() reports it
        return (u64_at((size_t)(b >> 3)) >> (b & 7)) & ((1ull << n) - 1);
    }
    bool ok() const { return bit <= (uint64_t)len * 8; }
    bool drain_empty_byte() {
        if (!ok()) return false;
        const Bitlen off = (Bitlen)(bit & 7);
        if (off) {
            if ((p[bit >> 3] >> off) != 0) return false;
            bit += 8 - off;
        }
        return true;
    }
};

// ---------------------------------------------------------------------------
// number types (data_types/*.rs): latent type, ordered latent map
// ---------------------------------------------------------------------------
template <class F> struct FT;
template <> struct FT<float> {
    using L = uint32_t;
    static constexpr Bitlen PREC = 23;
    static constexpr int OFF = 127;
    static constexpr Bitlen MANT = 24;
    static float rnd(float x) { return std::roundf(x); }
};
template <> struct FT<double> {
    using L = uint64_t;
    static constexpr Bitlen PREC = 52;
    static constexpr int OFF = 1023;
    static constexpr Bitlen MANT = 53;
    static double rnd(double x) { return std::round(x); }
};
template <class F> inline typename FT<F>::L fbits(F x) {
    typename FT<F>::L l;
    memcpy(&l, &x, sizeof x);
    return l;
}
template <class F> inline F ffrom(typename FT<F>::L l) {
    F x;
    memcpy(&x, &l, sizeof x);
    return x;
}
// Float::exp2 (float.rs:93-96): the exponent field written directly
template <class F> inline F fexp2(int power) {
    using L = typename FT<F>::L;
    const int32_t e = (int32_t)((int64_t)FT<F>::OFF + power);
    const L l = sizeof(L) == 8 ? (L)(uint64_t)(int64_t)e : (L)(uint32_t)e;
    return ffrom<F>((L)(l << FT<F>::PREC));
}
template <class F> inline int fexponent(F x) {
    return (int)(fbits<F>(std::fabs(x)) >> FT<F>::PREC) - FT<F>::OFF;
}
template <class F> inline bool fnormal(F x) { return std::fpclassify(x) == FP_NORMAL; }

template <class T> struct NT;
template <> struct NT<uint32_t> {
    using L = uint32_t;
    static constexpr uint8_t BYTE = 1;
    static L to_lat(uint32_t x) { return x; }
    static uint32_t from_lat(L l) { return l; }
};
template <> struct NT<uint64_t> {
    using L = uint64_t;
    static constexpr uint8_t BYTE = 2;
    static L to_lat(uint64_t x) { return x; }
    static uint64_t from_lat(L l) { return l; }
};
template <> struct NT<int32_t> {
    using L = uint32_t;
    static constexpr uint8_t BYTE = 3;
    static L to_lat(int32_t x) { return (uint32_t)x ^ 0x80000000u; }
    static int32_t from_lat(L l) { return (int32_t)(l ^ 0x80000000u); }
};
template <> struct NT<int64_t> {
    using L = uint64_t;
    static constexpr uint8_t BYTE = 4;
    static L to_lat(int64_t x) { return (uint64_t)x ^ 0x8000000000000000ull; }
    static int64_t from_lat(L l) { return (int64_t)(l ^ 0x8000000000000000ull); }
};
template <class F> struct NTF {
    using L = typename FT<F>::L;
    static L to_lat(F x) {
        const L m = fbits<F>(x);
        return (m & MID<L>) ? (L)~m : (L)(m ^ MID<L>);
    }
    static F from_lat(L l) { return (l & MID<L>) ? ffrom<F>(l ^ MID<L>) : ffrom<F>((L)~l); }
};
template <> struct NT<float> : NTF<float> {
    static constexpr uint8_t BYTE = 5;
};
template <> struct NT<double> : NTF<double> {
    static constexpr uint8_t BYTE = 6;
};

// Float::int_float_to_latent / int_float_from_latent (float.rs:137-178)
template <class F> typename FT<F>::L int_float_to_latent(F x) {
    using L = typename FT<F>::L;
    const F a = std::fabs(x);
    const L gpi = (L)1 << FT<F>::MANT;
    const F gpif = (F)gpi;
    const L ai = a < gpif ? (L)a : (L)(gpi + (fbits<F>(a) - fbits<F>(gpif)));
    return !std::signbit(x) ? (L)(MID<L> + ai) : (L)(MID<L> - 1 - ai);
}
template <class F> F int_float_from_latent(typename FT<F>::L l) {
    using L = typename FT<F>::L;
    bool neg;
    L ai;
    if (l >= MID<L>) {
        neg = false;
        ai = l - MID<L>;
    } else {
        neg = true;
        ai = MID<L> - 1 - l;
    }
    const L gpi = (L)1 << FT<F>::MANT;
    const F af = ai < gpi ? (F)ai : ffrom<F>((L)(fbits<F>((F)gpi) + (ai - gpi)));
    return neg ? -af : af;
}

// f64::total_cmp key
inline int64_t total_key(double x) {
    int64_t b;
    memcpy(&b, &x, 8);
    b ^= (int64_t)(((uint64_t)(b >> 63)) >> 1);
    return b;
}

// ---------------------------------------------------------------------------
// sampling.rs
// ---------------------------------------------------------------------------
constexpr size_t MIN_SAMPLE = 10;

template <class S, class Filter>
bool choose_sample(size_t n, Filter filter, std::vector<S> &res) {
    if (n < MIN_SAMPLE) return false;
    const size_t target = MIN_SAMPLE + (n - MIN_SAMPLE) / 40;
    Xoro rng(0);
    std::vector<uint8_t> visited((n + 7) / 8, 0);
    res.clear();
    res.reserve(target);
    size_t iters = 0;
    while (res.size() < target && iters < 4 * target) {
        const size_t idx = (size_t)(rng.next() % (uint64_t)n);
        const uint8_t mask = (uint8_t)(1u << (idx % 8));
        if (!(visited[idx / 8] & mask)) {
            S v;
            if (filter(idx, v)) res.push_back(v);
            visited[idx / 8] |= mask;
        }
        iters++;
    }
    return res.size() >= MIN_SAMPLE;
}

// est_bits_saved_per_num (sampling.rs:86-120); fn(x) -> (primary, bits_saved)
template <class L, class S, class Fn>
double est_bits_saved_per_num(const std::vector<S> &sample, Fn fn) {
    std::unordered_map<L, size_t> idx;
    idx.reserve(sample.size() * 2 + 1);
    std::vector<std::pair<size_t, double>> acc;
    for (const S &x : sample) {
        const std::pair<L, double> pr = fn(x);
        auto it = idx.find(pr.first);
        size_t k;
        if (it == idx.end()) {
            k = acc.size();
            idx.emplace(pr.first, k);
            acc.push_back({0, 0.0});
        } else {
            k = it->second;
        }
        acc[k].first += 1;
        acc[k].second += pr.second;
    }
    const size_t cutoff = std::max<size_t>(1, (size_t)((double)sample.size() / 256.0));
    double s = 0.0;
    for (const auto &a : acc)
        if (a.first <= cutoff) s += a.second;
    return s / (double)sample.size();
}

// mode/mod.rs
inline double single_category_entropy(double p) {
    return (p == 0.0 || p == 1.0) ? 0.0 : -p * std::log2(p);
}
inline double worst_case_categorical_entropy(double p, double m1) {
    return single_category_entropy(p) + m1 * single_category_entropy((1.0 - p) / m1);
}

// ---------------------------------------------------------------------------
// mode/int_mult.rs
// ---------------------------------------------------------------------------
const double ZETA_OF_2 = M_PI * M_PI / 6.0;

template <class L> L calc_gcd(L x, L y) {
    if (x == 0) return y;
    for (;;) {
        if (y == 0) return x;
        x %= y;
        std::swap(x, y);
    }
}

template <class L> L calc_triple_gcd(const L *t) {
    L a = t[0], b = t[1], c = t[2];
    if (a > b) std::swap(a, b);
    if (b > c) std::swap(b, c);
    if (a > b) std::swap(a, b);
    return calc_gcd<L>(b - a, c - a);
}

template <class Fn> bool solve_root_by_false_position(Fn f, double lb, double ub, double *out) {
    const double X_TOLERANCE = 1E-4;
    double flb = f(lb), fub = f(ub);
    if (flb > 0.0 || fub < 0.0) return false;
    while (ub - lb > X_TOLERANCE && fub - flb > 0.0) {
        const double lb_prop = 0.001 + 0.998 * fub / (fub - flb);
        const double mid = lb_prop * lb + (1.0 - lb_prop) * ub;
        const double fmid = f(mid);
        if (fmid < 0.0) {
            lb = mid;
            flb = fmid;
        } else {
            ub = mid;
            fub = fmid;
        }
    }
    *out = (lb + ub) / 2.0;
    return true;
}

bool filter_score_triple_gcd(double gcd, size_t triples_w_gcd_n, size_t total_triples_n,
                             double *score) {
    const double triples_w_gcd = (double)triples_w_gcd_n;
    const double total_triples = (double)total_triples_n;
    const double prob_per_triple = triples_w_gcd / total_triples;
    const double natural = 1.0 / (ZETA_OF_2 * gcd * gcd);
    const double stdev = std::sqrt(natural * (1.0 - natural) / total_triples);
    const double z = (prob_per_triple - natural) / stdev;
    if (z < 3.0) return false;
    const double lcb = triples_w_gcd - 1.0 * std::sqrt(triples_w_gcd);
    if (lcb <= 0.0) return false;
    const double cong = std::min(ZETA_OF_2 * lcb / total_triples, 1.0);
    const double gcd_m1 = gcd - 1.0;
    const double inv_sq = 1.0 / (gcd_m1 * gcd_m1);
    auto f = [&](double p) {
        const double q = 1.0 - p;
        return p * (p * p) + q * (q * q) * inv_sq - cong;
    };
    const double lb = 1.0 / gcd;
    const double ub = std::cbrt(cong) + 2.220446049250313e-16;
    double p;
    if (!solve_root_by_false_position(f, lb, ub, &p)) return false;
    const double ent = worst_case_categorical_entropy(p, gcd_m1);
    const double bits = std::log2(gcd) - ent;
    if (bits < 0.5) return false;
    *score = bits;
    return true;
}

template <class L> bool choose_candidate_base(const std::vector<L> &sample, L *base, double *bits) {
    std::vector<L> gcds;
    for (size_t i = 0; i + 3 <= sample.size(); i += 3) {
        const L g = calc_triple_gcd<L>(&sample[i]);
        if (g > 1) gcds.push_back(g);
    }
    const size_t total = sample.size() / 3;
    std::unordered_map<L, size_t> idx;
    std::vector<std::pair<L, size_t>> counts;
    for (L g : gcds) {
        auto it = idx.find(g);
        if (it == idx.end()) {
            idx.emplace(g, counts.size());
            counts.push_back({g, 1});
        } else {
            counts[it->second].second++;
        }
    }
    bool found = false;
    uint64_t best_key = 0;
    for (const auto &c : counts) {
        double score;
        if (!filter_score_triple_gcd((double)(uint64_t)c.first, c.second, total, &score)) continue;
        const uint64_t key = NT<double>::to_lat(score);
        if (!found || key >= best_key) {  // max_by_key keeps the last maximum
            found = true;
            best_key = key;
            *base = c.first;
            *bits = score;
        }
    }
    return found;
}

template <class T> bool int_choose_base(const T *nums, size_t n, typename NT<T>::L *out) {
    using L = typename NT<T>::L;
    std::vector<L> sample;
    if (!choose_sample<L>(n, [&](size_t i, L &v) { v = NT<T>::to_lat(nums[i]); return true; }, sample))
        return false;
    L cand;
    double bits;
    if (!choose_candidate_base<L>(sample, &cand, &bits)) return false;
    const double saved = est_bits_saved_per_num<L>(
        sample, [&](L x) { return std::pair<L, double>(x / cand, bits); });
    if (saved > 0.5) {
        *out = cand;
        return true;
    }
    return false;
}

// ---------------------------------------------------------------------------
// mode/float_mult.rs
// ---------------------------------------------------------------------------
template <class F> struct FMConfig {
    F base, inv_base;
    static FMConfig from_base(F b) { return {b, (F)1.0 / b}; }
    static FMConfig from_inv_base(F ib) { return {(F)1.0 / ib, ib}; }
};

template <class F> F insignificant_float_to(F x) {
    const Bitlen spare = FT<F>::PREC > 6 ? FT<F>::PREC - 6 : 0;
    return x * fexp2<F>(-(int)spare);
}
template <class F> bool is_approx_zero(F small, F big) { return small <= insignificant_float_to(big); }
template <class F> bool is_small_remainder(F rem, F orig) { return rem <= orig * fexp2<F>(-16); }
template <class F> bool is_imprecise(F value, F err) { return value <= err * fexp2<F>(6); }

template <class F> bool approx_pair_gcd(F greater, F lesser, F *out) {
    if (is_approx_zero(lesser, greater) || lesser == greater) return false;
    struct PM {
        F value, err;
    };
    const F eps = fexp2<F>(-(int)FT<F>::PREC);
    auto rem_assign = [&](PM &lhs, const PM &rhs) {
        const F ratio = FT<F>::rnd(lhs.value / rhs.value);
        lhs.err += ratio * rhs.err + lhs.value * eps;
        lhs.value = std::fabs(lhs.value - ratio * rhs.value);
    };
    PM pg{greater, (F)0}, pl{lesser, (F)0};
    for (;;) {
        const F prev = pg.value;
        rem_assign(pg, pl);
        if (is_small_remainder(pg.value, prev) || pg.value <= pg.err) {
            *out = pl.value;
            return true;
        }
        if (is_approx_zero(pg.value, greater) || is_imprecise(pg.value, pg.err)) return false;
        std::swap(pg, pl);
    }
}

template <class F> bool choose_config_by_trailing_zeros(const std::vector<F> &sample, FMConfig<F> *cfg) {
    using L = typename FT<F>::L;
    const Bitlen prec = FT<F>::PREC;
    auto p2d = [&](int exponent, Bitlen tz) { return exponent - (int)(prec > tz ? prec - tz : 0); };
    int k = 0x7FFFFFFF;
    size_t count = 0;
    for (F x : sample) {
        const Bitlen tz = tzc<L>(fbits<F>(x));
        if (x != (F)0 && tz >= 5) {
            const int kp = p2d(fexponent(x), tz);
            count++;
            k = std::min(k, kp);
        }
    }
    const size_t required =
        std::max((size_t)std::ceil((double)sample.size() * 0.5), MIN_SAMPLE);
    if (count < required) return false;
    std::vector<L> ints;
    const Bitlen lshift = BITS<L> - prec - 1;
    for (F x : sample) {
        const int e = fexponent(x);
        const int kp = p2d(e, tzc<L>(fbits<F>(x)));
        if (kp >= k && e < k + (int)BITS<L>) {
            const Bitlen rshift = BITS<L> - 1 - (Bitlen)(e - k);
            const L l = (L)((L)(fbits<F>(x) << lshift) | MID<L>);
            ints.push_back((L)(l >> rshift));
        }
    }
    if (ints.size() < required) return false;
    L ib;
    double bits;
    if (!choose_candidate_base<L>(ints, &ib, &bits)) ib = 1;
    *cfg = FMConfig<F>::from_base((F)ib * fexp2<F>(k));
    return true;
}

template <class F> bool approx_sample_gcd_euclidean(const std::vector<F> &sample, F *out) {
    std::vector<F> gcds;
    for (size_t i = 0; i + 1 < sample.size(); i += 2) {
        const F a = sample[i], b = sample[i + 1];
        F g;
        if (approx_pair_gcd<F>(std::fmax(a, b), std::fmin(a, b), &g)) gcds.push_back(g);
    }
    const size_t required = 1 + (size_t)std::ceil((double)sample.size() * 0.001);
    if (gcds.size() < required) return false;
    std::sort(gcds.begin(), gcds.end());
    for (double pct : {0.1, 0.3, 0.5}) {
        const F cand = gcds[(size_t)(pct * (double)gcds.size())];
        size_t similar = 0;
        for (F g : gcds)
            if (std::fabs(g - cand) < (F)0.01 * cand) similar++;
        if (similar >= required) {
            *out = cand;
            return true;
        }
    }
    return false;
}

template <class F> F center_sample_base(F base, const std::vector<F> &sample) {
    const F inv = (F)1.0 / base;
    F tsum = 0, tweight = 0;
    for (F x : sample) {
        const F mult = FT<F>::rnd(x * inv);
        const Bitlen me = (Bitlen)fexponent(mult);
        if (me < FT<F>::PREC && mult != (F)0) {
            const F overshoot = (mult * base) - x;
            const F weight = (F)(double)(FT<F>::PREC - me);
            tsum += weight * (overshoot / mult);
            tweight += weight;
        }
    }
    return base - tsum / tweight;
}

template <class F> FMConfig<F> snap_to_int_reciprocal(F base) {
    const F inv = (F)1.0 / base;
    const F round_inv = FT<F>::rnd(inv);
    const F dec_inv = (F)std::pow(10.0, std::round(std::log10((double)inv)));
    if (std::fabs(inv - round_inv) < (F)0.02) return FMConfig<F>::from_inv_base(round_inv);
    if (std::fabs(inv - dec_inv) / inv < (F)0.01) return FMConfig<F>::from_inv_base(dec_inv);
    return FMConfig<F>::from_base(base);
}

template <class F> bool fm_choose_config(const std::vector<F> &sample, FMConfig<F> *cfg) {
    if (choose_config_by_trailing_zeros(sample, cfg)) return true;
    F base;
    if (!approx_sample_gcd_euclidean(sample, &base)) return false;
    base = center_sample_base(base, sample);
    *cfg = snap_to_int_reciprocal(base);
    return true;
}

template <class F> bool fm_bits_saved(const FMConfig<F> &c, const std::vector<F> &sample, double *out) {
    using L = typename FT<F>::L;
    const double saved = est_bits_saved_per_num<L>(sample, [&](F x) {
        const F mult = FT<F>::rnd(x * c.inv_base);
        const L primary = int_float_to_latent<F>(mult);
        const Bitlen me = (Bitlen)fexponent(mult);
        const Bitlen inter = FT<F>::PREC > me ? FT<F>::PREC - me : 0;
        const L approx = NT<F>::to_lat(mult * c.base);
        const L xu = NT<F>::to_lat(x);
        const L abs_adj = std::max(xu, approx) - std::min(xu, approx);
        const Bitlen adj_bits = 1 + 2 * (BITS<L> - lzc<L>(abs_adj));
        return std::pair<L, double>(primary, (double)inter - (double)adj_bits);
    });
    if (saved >= 0.5) {
        *out = saved;
        return true;
    }
    return false;
}

// ---------------------------------------------------------------------------
// mode/float_quant.rs
// ---------------------------------------------------------------------------
template <class F> Bitlen fq_best_k(const std::vector<F> &sample, double *bits_saved) {
    using L = typename FT<F>::L;
    std::vector<uint32_t> hist(FT<F>::PREC + 1, 0);
    for (F x : sample) hist[std::min(FT<F>::PREC, tzc<L>(fbits<F>(x)))] += 1;
    uint32_t rc = 0;
    for (size_t i = hist.size(); i-- > 0;) {
        rc += hist[i];
        hist[i] = rc;
    }
    const double len = (double)sample.size();
    Bitlen best_k = 0;
    double best = 0.0;
    for (size_t k = 1; k < hist.size(); k++) {
        if (hist[k] == 0) continue;
        const double freq = (double)hist[k] / len;
        const double ncat = (double)((1ull << k) - 1);
        const double worst = worst_case_categorical_entropy(freq, ncat);
        const double saved = (double)k - worst;
        if (saved > best) {
            best_k = (Bitlen)k;
            best = saved;
        } else {
            break;
        }
    }
    *bits_saved = best;
    return best_k;
}

// ---------------------------------------------------------------------------
// modes and latent splits
// ---------------------------------------------------------------------------
enum { M_CLASSIC = 0, M_INT_MULT = 1, M_FLOAT_MULT = 2, M_FLOAT_QUANT = 3, M_DICT = 4 };
struct Mode {
    int kind = M_CLASSIC;
    uint64_t base = 0;  // IntMult base / FloatMult base latent (ordered)
    Bitlen k = 0;       // FloatQuant
    std::vector<uint64_t> dict;
};

enum { D_NOOP = 0, D_CONSEC = 1, D_LOOKBACK = 2, D_CONV1 = 3 };
struct DeltaEnc {
    int kind = D_NOOP;
    Bitlen order = 0;
    bool sec_uses = false;
    Bitlen window_n_log = 0, state_n_log = 0;
};
// per latent variable (LatentVarDeltaEncoding)
struct VarEnc {
    int kind = D_NOOP;
    Bitlen order = 0;
    Bitlen window_n_log = 0, state_n_log = 0;
    size_t n_latents_per_state() const {
        switch (kind) {
        case D_CONSEC: return order;
        case D_LOOKBACK: return (size_t)1 << state_n_log;
        default: return 0;
        }
    }
};
enum { K_DELTA = 0, K_PRIMARY = 1, K_SECONDARY = 2 };
// DeltaEncoding::for_latent_var (metadata/delta_encoding.rs:255-297)
VarEnc for_var(const DeltaEnc &d, int key) {
    VarEnc v;
    if (d.kind == D_NOOP || key == K_DELTA) return v;
    if (key == K_SECONDARY && !d.sec_uses) return v;
    if (d.kind == D_CONV1 && key == K_SECONDARY) return v;
    v.kind = d.kind;
    v.order = d.order;
    v.window_n_log = d.window_n_log;
    v.state_n_log = d.state_n_log;
    return v;
}
DeltaEnc new_lookback(size_t n) {  // delta/mod.rs:29-41
    DeltaEnc d;
    d.kind = D_LOOKBACK;
    d.window_n_log = std::min<Bitlen>(std::max<Bitlen>(bte_offset<uint32_t>((uint32_t)n - 1), 4), 15);
    d.state_n_log = 0;
    return d;
}

// ---------------------------------------------------------------------------
// delta/consecutive.rs, delta/lookback.rs
// ---------------------------------------------------------------------------
template <class L> void toggle_center(L *p, size_t n) {
    for (size_t i = 0; i < n; i++) p[i] = (L)(p[i] + MID<L>);
}

template <class L> std::vector<L> consec_encode(Bitlen order, L *lat, size_t n) {
    std::vector<L> moments;
    for (Bitlen o = 0; o < order; o++) {
        moments.push_back(n ? lat[0] : 0);
        for (size_t i = n; i-- > 1;) lat[i] = (L)(lat[i] - lat[i - 1]);
        const size_t t = std::min<size_t>(n, 1);
        lat += t;
        n -= t;
    }
    toggle_center(lat, n);
    return moments;
}

template <class L> void consec_decode(std::vector<L> &moments, L *lat, size_t n) {
    toggle_center(lat, n);
    for (size_t m = moments.size(); m-- > 0;) {
        L mom = moments[m];
        for (size_t i = 0; i < n; i++) {
            const L tmp = lat[i];
            lat[i] = mom;
            mom = (L)(mom + tmp);
        }
        moments[m] = mom;
    }
}

template <class L>
std::vector<uint32_t> choose_lookbacks(Bitlen window_n_log, Bitlen state_n_log, const L *lat, size_t n) {
    const size_t PROPOSED = 16, BRUTE = 6, REPEATING = 4;
    const size_t state_n = (size_t)1 << state_n_log;
    if (n <= state_n) return {};
    const size_t ht_n = (size_t)1 << (window_n_log + 1);
    const size_t window_n = (size_t)1 << window_n_log;
    std::vector<uint32_t> counts(std::min(window_n, n), 1);
    std::vector<uint32_t> out(n - state_n);
    std::vector<size_t> table(2 * ht_n, 0);
    size_t prop[16];
    for (size_t i = 0; i < PROPOSED; i++) prop[i] = std::min(i + 1, state_n);
    size_t best = 1, rep = 0;
    const size_t hmask = ht_n - 1;
    auto hash_fn = [&](uint64_t x) {
        x = (x ^ (x >> 32)) * 11400714819323197441ull;
        x = x ^ (x >> 32);
        return (size_t)x & hmask;
    };
    for (size_t i = state_n; i < n; i++) {
        const L l = lat[i];
        const size_t nb = std::min(i, PROPOSED);
        prop[nb - 1] = nb;
        // hash_lookup
        size_t pidx = BRUTE + REPEATING, off = 0;
        for (Bitlen coarse : {0u, 8u}) {
            const uint64_t bucket = (uint64_t)l >> coarse;
            const uint64_t buckets[3] = {bucket - 1, bucket, bucket + 1};
            size_t hs[3];
            for (int q = 0; q < 3; q++) hs[q] = hash_fn(buckets[q]);
            for (int q = 0; q < 3; q++) {
                const size_t last = i - table[off + hs[q]];
                prop[pidx] = last <= window_n ? last : std::min(pidx, i);
                pidx++;
            }
            table[off + hs[1]] = i;
            off += ht_n;
        }
        // find_best_lookback
        Bitlen best_good = 0;
        size_t nbest = 0;
        for (size_t q = 0; q < PROPOSED; q++) {
            const size_t lb = prop[q];
            const uint32_t cnt = counts[lb - 1];
            const L other = lat[i - lb];
            const Bitlen lg = 32 - lzc<uint32_t>(cnt);
            const L d = std::min<L>((L)(l - other), (L)(other - l));
            const Bitlen good = lg + lzc<L>(d);
            if (good > best_good) {
                best_good = good;
                nbest = lb;
            }
        }
        if (nbest != best) rep++;
        prop[BRUTE + rep % REPEATING] = nbest;
        best = nbest;
        out[i - state_n] = (uint32_t)best;
        counts[best - 1] += 1;
    }
    return out;
}

template <class L>
std::vector<L> lookback_encode(Bitlen state_n_log, const std::vector<uint32_t> &lbs, L *lat, size_t n) {
    const size_t state_n = (size_t)1 << state_n_log;
    const size_t real = std::min(n, state_n);
    for (size_t i = n; i-- > real;) lat[i] = (L)(lat[i] - lat[i - lbs[i - state_n]]);
    std::vector<L> state(state_n, 0);
    for (size_t i = 0; i < real; i++) state[state_n - real + i] = lat[i];
    toggle_center(lat, n);
    return state;
}

// ---------------------------------------------------------------------------
// sort_utils.rs and histograms.rs
// ---------------------------------------------------------------------------
template <class L> L choose_pivot(const L *lat, size_t len) {
    size_t a = len / 4, b = len / 2, c = (len * 3) / 4;
    auto sort2 = [&](size_t &x, size_t &y) {
        if (lat[y] < lat[x]) std::swap(x, y);
    };
    auto sort3 = [&](size_t &x, size_t &y, size_t &z) {
        sort2(x, y);
        sort2(y, z);
        sort2(x, y);
    };
    if (len >= 8) {
        if (len >= 50) {
            auto adj = [&](size_t &x) {
                size_t lo = x - 1, hi = x + 1;
                sort3(lo, x, hi);
            };
            adj(a);
            adj(b);
            adj(c);
        }
        sort3(a, b, c);
    }
    return lat[b];
}

template <class L> void break_patterns(L *v, size_t len) {
    if (len < 8) return;
    uint64_t seed = len;
    auto gen = [&]() {
        uint64_t r = seed;
        r ^= r << 13;
        r ^= r >> 7;
        r ^= r << 17;
        seed = r;
        return (size_t)seed;
    };
    size_t modulus = 1;
    while (modulus < len) modulus <<= 1;
    const size_t pos = len / 4 * 2;
    for (size_t i = 0; i < 3; i++) {
        size_t other = gen() & (modulus - 1);
        if (other >= len) other -= len;
        std::swap(v[pos - 1 + i], v[other]);
    }
}

template <class L> size_t partition(L *lat, size_t len, L pivot, bool *bad) {
    size_t left = 0;
    for (size_t pos = 0; pos < len; pos++) {
        const L value = lat[pos];
        const bool lt = value < pivot;
        lat[pos] = lat[left];
        lat[left] = value;
        left += lt;
    }
    *bad = 1 + std::min(left, len - left) < len / 8;
    return left;
}

template <class L> void heapsort(L *x, size_t len) {
    auto sift = [&](size_t n, size_t node) {
        for (;;) {
            size_t child = 2 * node + 1;
            if (child >= n) break;
            if (child + 1 < n) child += (x[child] < x[child + 1]);
            if (x[node] >= x[child]) break;
            std::swap(x[node], x[child]);
            node = child;
        }
    };
    for (size_t i = len / 2; i-- > 0;) sift(len, i);
    for (size_t i = len; i-- > 1;) {
        std::swap(x[0], x[i]);
        sift(i, 0);
    }
}

template <class L> struct HBin {
    size_t count;
    L lower, upper;
};
template <class L> struct Bound {
    bool tight;
    L v;
};

template <class L> struct Hist {
    uint64_t n, n_bins;
    Bitlen log;
    size_t applied = 0, next_avail = 0;
    bool has_inc = false;
    HBin<L> inc{0, 0, 0};
    std::vector<HBin<L>> dst;

    size_t bin_idx(size_t c) const { return (size_t)(((uint64_t)c << log) / n); }
    size_t c_count(size_t b) const { return (size_t)(((uint64_t)(b + 1) * n + n_bins - 1) >> log); }
    static L smin(const L *p, size_t m) {
        L r = LMAX<L>;
        for (size_t i = 0; i < m; i++) r = std::min(r, p[i]);
        return r;
    }
    static L smax(const L *p, size_t m) {
        L r = 0;
        for (size_t i = 0; i < m; i++) r = std::max(r, p[i]);
        return r;
    }
    void apply_incomplete(const L *p, size_t m, Bound<L> lo, Bound<L> up) {
        if (!m) return;
        if (has_inc) {
            inc.upper = up.tight ? up.v : smax(p, m);
            inc.count += m;
        } else {
            const L tl = lo.tight ? lo.v : smin(p, m);
            const L tu = up.tight ? up.v : smax(p, m);
            inc = {m, tl, tu};
            has_inc = true;
        }
        applied += m;
    }
    bool complete_bin(size_t b) {
        if (!has_inc) return false;
        next_avail = b + 1;
        dst.push_back(inc);
        has_inc = false;
        return true;
    }
    void apply_constant_run(const L *p, size_t m) {
        const size_t start = applied, mid = start + m / 2, end = start + m;
        size_t b = bin_idx(mid);
        if (b > next_avail) {
            const size_t spare = b - 1;
            if (!complete_bin(spare)) b = spare;
        }
        apply_incomplete(p, m, Bound<L>{true, p[0]}, Bound<L>{true, p[0]});
        if (end >= c_count(b)) complete_bin(b);
    }
    void apply_sorted(const L *p, size_t m) {
        while (m) {
            const size_t tb = bin_idx(applied);
            const size_t tc = c_count(tb);
            const size_t ti = tc - applied;
            if (ti >= m) {
                apply_incomplete(p, m, Bound<L>{true, p[0]}, Bound<L>{true, p[m - 1]});
                if (ti == m) complete_bin(tb);
                break;
            }
            size_t l = ti - 1, r = ti;
            const L x = p[l];
            while (l > 0 && p[l - 1] == x) l--;
            while (r < m && p[r] == x) r++;
            if (l > 0) apply_incomplete(p, l, Bound<L>{true, p[0]}, Bound<L>{true, p[l - 1]});
            apply_constant_run(p + l, r - l);
            p += r;
            m -= r;
        }
    }
    void recurse(L *p, size_t m, Bound<L> lb, Bound<L> ub, uint32_t bad_limit) {
        if (!m) return;
        const size_t tb = bin_idx(applied);
        const size_t tc = c_count(tb);
        const size_t end = applied + m;
        if (end <= tc) {
            apply_incomplete(p, m, lb, ub);
            if (end == tc) complete_bin(tb);
            return;
        }
        const L loose_lb = lb.v;
        if (loose_lb == ub.v || m == 1) {
            apply_constant_run(p, m);
            return;
        }
        const L tp = choose_pivot(p, m);
        L pivot;
        Bound<L> lhs_ub, rhs_lb;
        if (tp > loose_lb) {
            pivot = tp;
            lhs_ub = {false, (L)(tp - 1)};
            rhs_lb = {true, tp};
        } else {
            pivot = (L)(tp + 1);
            lhs_ub = {true, tp};
            rhs_lb = {false, (L)(tp + 1)};
        }
        bool bad;
        const size_t lc = partition(p, m, pivot, &bad);
        if (bad) {
            bad_limit--;
            if (bad_limit == 0) {
                heapsort(p, lc);
                heapsort(p + lc, m - lc);
                apply_sorted(p, m);
                return;
            }
            break_patterns(p, lc);
            break_patterns(p + lc, m - lc);
        }
        recurse(p, lc, lb, lhs_ub, bad_limit);
        recurse(p + lc, m - lc, rhs_lb, ub, bad_limit);
    }
};

template <class L> std::vector<HBin<L>> histogram(L *lat, size_t n, Bitlen n_bins_log) {
    Hist<L> h;
    h.n = n;
    h.log = n_bins_log;
    h.n_bins = 1ull << n_bins_log;
    h.recurse(lat, n, Bound<L>{false, 0}, Bound<L>{false, LMAX<L>}, 1 + ilog2_64((uint64_t)n + 1));
    return h.dst;
}

// ---------------------------------------------------------------------------
// bin_optimization.rs
// ---------------------------------------------------------------------------
inline float log2_approx(float x) {
    const float Z = 0.674f;
    const uint32_t SIGNIF_MASK = 0x7FFFFF;
    uint32_t zb;
    memcpy(&zb, &Z, 4);
    const uint32_t Z_SIGNIF = zb & SIGNIF_MASK;
    const float B = 2.0f / Z;
    const float C = -B / (6.0f * Z);
    const float A = -B - C;
    uint32_t bits;
    memcpy(&bits, &x, 4);
    const uint32_t e = bits >> 23;
    const uint32_t signif = bits & SIGNIF_MASK;
    const uint32_t high = signif > Z_SIGNIF ? 1u : 0u;
    const uint32_t log_int = e + high - 127;
    const uint32_t ne = 0x7F ^ high;
    const uint32_t nb = (ne << 23) | signif;
    float normalized;
    memcpy(&normalized, &nb, 4);
    const float t = C * normalized;
    const float u = B + t;
    const float v = normalized * u;
    const float w = (float)log_int + A;
    return w + v;
}

template <class L> constexpr Bitlen bin_exact_bits(Bitlen asl) { return asl + BITS<L> + OBB<L>; }

template <class L>
float bin_cost(float meta_cost, L lower, L upper, uint32_t count, float total_log2) {
    const float c = (float)count;
    const float ans_cost = total_log2 - log2_approx(c);
    const float offset_cost = (float)bte_offset<L>((L)(upper - lower));
    const float s = ans_cost + offset_cost;
    const float p = s * c;
    return meta_cost + p;
}

template <class L> struct Info {
    uint32_t weight;
    L lower, upper;
    Bitlen offset_bits;
    uint32_t symbol;
};

// One row of the bin optimizer's dynamic program: over j = i .. 0, the cost
// best_cost[j] + bin_cost(bins j..i) with bin_cost's f32 operations in their
// order, and the first minimum in that order (the largest j among equal
// costs). ob[j] holds the offset bits of bins j..i as a float.
struct DpRow {
    float cost;
    size_t j;
};
DpRow dp_row_scalar(const float *best_cost, const uint32_t *cc, const float *ob, size_t i, uint32_t ci,
                    float total_log2, float meta) {
    float bc = 3.40282347e+38f;
    size_t bj = (size_t)-1;
    for (size_t j = i + 1; j-- > 0;) {
        const float c = (float)(ci - cc[j]);
        const float ans_cost = total_log2 - log2_approx(c);
        const float sv = ans_cost + ob[j];
        const float cost = best_cost[j] + (meta + sv * c);
        if (cost < bc) {
            bc = cost;
            bj = j;
        }
    }
    return DpRow{bc, bj};
}

#if defined(__x86_64__)
typedef float v8f __attribute__((vector_size(32)));
typedef int32_t v8i __attribute__((vector_size(32)));
typedef uint32_t v8u __attribute__((vector_size(32)));

// The same row eight j at a time (AVX2): every lane runs log2_approx's and
// bin_cost's f32 operations in the same order, so each cost is bit-identical
// to the scalar one; the minimum keeps the largest j among equal costs.
__attribute__((target("avx2"))) DpRow dp_row_avx2(const float *best_cost, const uint32_t *cc, const float *ob,
                                                  size_t i, uint32_t ci, float total_log2, float meta) {
    const float Z = 0.674f;
    uint32_t zb;
    memcpy(&zb, &Z, 4);
    const uint32_t ZS = zb & 0x7FFFFF;
    const float B = 2.0f / Z, C = -B / (6.0f * Z), A = -B - C;
    const size_t n = i + 1;
    // an empty bin i (count 0 somewhere in the row) converts differently as a
    // signed lane: the scalar row takes it
    if (ci == cc[i]) return dp_row_scalar(best_cost, cc, ob, i, ci, total_log2, meta);
    v8f best;
    v8i bestj;
    for (int l = 0; l < 8; l++) {
        best[l] = 3.40282347e+38f;
        bestj[l] = -1;
    }
    const v8i lanes = {0, 1, 2, 3, 4, 5, 6, 7};
    size_t j0 = 0;
    for (; j0 + 8 <= n; j0 += 8) {
        v8u cj, cnt;
        v8f bcj, obj;
        memcpy(&cj, cc + j0, 32);
        memcpy(&bcj, best_cost + j0, 32);
        memcpy(&obj, ob + j0, 32);
        cnt = (v8u){ci, ci, ci, ci, ci, ci, ci, ci} - cj;
        const v8f c = __builtin_convertvector((v8i)cnt, v8f);
        const v8u bits = (v8u)c;
        const v8u e = bits >> 23, signif = bits & 0x7FFFFF;
        const v8u high = (v8u)(signif > ZS) & 1u;
        const v8u log_int = e + high - 127u;
        const v8u nbits = ((0x7Fu ^ high) << 23) | signif;
        const v8f normalized = (v8f)nbits;
        const v8f t = C * normalized;
        const v8f u = B + t;
        const v8f v = normalized * u;
        const v8f w = __builtin_convertvector((v8i)log_int, v8f) + A;
        const v8f lg = w + v;
        const v8f ans_cost = total_log2 - lg;
        const v8f sv = ans_cost + obj;
        const v8f cost = bcj + (meta + sv * c);
        const v8i jj = (v8i){(int)j0, (int)j0, (int)j0, (int)j0, (int)j0, (int)j0, (int)j0, (int)j0} + lanes;
        // later chunks hold larger j: they win ties
        const v8i take = (cost < best) | (cost == best);
        best = take ? cost : best;
        bestj = take ? jj : bestj;
    }
    float bc = 3.40282347e+38f;
    long bj = -1;
    for (int l = 0; l < 8; l++) {
        if (bestj[l] < 0) continue;
        if (best[l] < bc || (best[l] == bc && bestj[l] > bj)) {
            bc = best[l];
            bj = bestj[l];
        }
    }
    // the j past the last full chunk, scalar (each larger than every j above)
    for (size_t j = j0; j < n; j++) {
        const float c = (float)(ci - cc[j]);
        const float ans_cost = total_log2 - log2_approx(c);
        const float sv = ans_cost + ob[j];
        const float cost = best_cost[j] + (meta + sv * c);
        if (cost < bc || (cost == bc && (long)j > bj)) {
            bc = cost;
            bj = (long)j;
        }
    }
    return DpRow{bc, (size_t)bj};
}
#endif

DpRow (*pick_dp_row())(const float *, const uint32_t *, const float *, size_t, uint32_t, float, float) {
#if defined(__x86_64__)
    if (__builtin_cpu_supports("avx2")) return dp_row_avx2;
#endif
    return dp_row_scalar;
}

template <class L> std::vector<Info<L>> optimize_bins(const std::vector<HBin<L>> &bins, Bitlen asl) {
    static DpRow (*const dp_row)(const float *, const uint32_t *, const float *, size_t, uint32_t, float,
                                 float) = pick_dp_row();
    const size_t nb = bins.size();
    std::vector<uint32_t> cc(nb + 1, 0);
    std::vector<float> best_cost(nb + 1, 0.0f), ob(nb + 8, 0.0f);
    uint32_t c = 0;
    for (size_t i = 0; i < nb; i++) {
        c += (uint32_t)bins[i].count;
        cc[i + 1] = c;
    }
    const uint32_t total = c;
    const float total_log2 = log2_approx((float)c);
    std::vector<size_t> best_js(nb);
    const float meta = (float)bin_exact_bits<L>(asl);
    std::vector<L> lowers(nb);
    for (size_t i = 0; i < nb; i++) lowers[i] = bins[i].lower;
    // pk[k]: the first j with upper_i - lower_j < 2^k. The uppers ascend with
    // i, so each pointer only moves forward over the rows.
    size_t pk[BITS<L> + 1];
    for (Bitlen k = 0; k <= BITS<L>; k++) pk[k] = 0;
    for (size_t i = 0; i < nb; i++) {
        const L upper = bins[i].upper;
        const uint32_t ci = cc[i + 1];
        // offset bits of bins j..i, bte_offset(upper - lower_j): the lowers
        // ascend with j, so the bits are k on the run [pk[k], pk[k - 1]) of j
        size_t hi = i + 1;
        for (Bitlen k = bte_offset<L>((L)(upper - lowers[i])); hi > 0; k++) {
            size_t p = 0;
            if (k < BITS<L> && (upper >> k) != 0) {
                const L t = (L)(upper - ((L)1 << k));  // lower_j > t <=> upper - lower_j < 2^k
                size_t q = pk[k];
                while (q < hi && lowers[q] <= t) q++;
                pk[k] = q;
                p = q;
            }
            std::fill(ob.begin() + p, ob.begin() + hi, (float)k);
            hi = p;
        }
        const DpRow r = dp_row(best_cost.data(), cc.data(), ob.data(), i, ci, total_log2, meta);
        best_cost[i + 1] = r.cost;
        best_js[i] = r.j;
    }
    const float bcost = best_cost[nb];
    std::vector<std::pair<size_t, size_t>> part;
    const float single = bin_cost<L>(meta, bins[0].lower, bins[nb - 1].upper, total, total_log2);
    bool done = false;
    if (single < bcost + 0.1f * (float)total) {
        part.push_back({0, nb - 1});
        done = true;
    }
    if (!done) {
        bool trivial = true;
        for (const auto &b : bins)
            if (b.lower != b.upper) trivial = false;
        if (trivial) {
            float tc = 0.0f;
            for (const auto &b : bins) tc += bin_cost<L>(meta, b.lower, b.upper, (uint32_t)b.count, total_log2);
            if (tc < bcost + 0.1f * (float)total) {
                for (size_t i = 0; i < nb; i++) part.push_back({i, i});
                done = true;
            }
        }
    }
    if (!done) {
        size_t i = nb - 1;
        for (;;) {
            const size_t j = best_js[i];
            part.push_back({j, i});
            if (j > 0) i = j - 1;
            else break;
        }
        std::reverse(part.begin(), part.end());
    }
    std::vector<Info<L>> res;
    res.reserve(part.size());
    for (size_t s = 0; s < part.size(); s++) {
        const size_t j = part[s].first, i = part[s].second;
        size_t count = 0;
        for (size_t q = j; q <= i; q++) count += bins[q].count;
        res.push_back(Info<L>{(uint32_t)count, bins[j].lower, bins[i].upper,
                              bte_offset<L>((L)(bins[i].upper - bins[j].lower)), (uint32_t)s});
    }
    return res;
}

// ---------------------------------------------------------------------------
// ans/encoding.rs, ans/spec.rs, ans/decoding.rs
// ---------------------------------------------------------------------------
std::vector<uint32_t> quantize_weights_to(const std::vector<uint32_t> &counts, size_t total, Bitlen size_log) {
    if (size_log == 0) return {1};
    const uint32_t required = 1u << size_log;
    const float mult = (float)required / (float)total;
    std::vector<float> desired(counts.size());
    float dsum = 0.0f;
    for (size_t i = 0; i < counts.size(); i++) {
        const float d = (float)counts[i] * mult - 1.0f;
        desired[i] = d > 0.0f ? d : 0.0f;
    }
    for (float d : desired) dsum += d;
    const uint32_t required_surplus = required - (uint32_t)counts.size();
    const float smult = dsum == 0.0f ? 0.0f : (float)required_surplus / dsum;
    std::vector<float> fw(counts.size());
    std::vector<uint32_t> w(counts.size());
    uint64_t sum = 0;
    for (size_t i = 0; i < counts.size(); i++) {
        fw[i] = 1.0f + desired[i] * smult;
        const float r = std::roundf(fw[i]);
        w[i] = r <= 0.0f ? 0u : (r >= 4294967295.0f ? 0xFFFFFFFFu : (uint32_t)r);
        sum += w[i];
    }
    size_t i = 0;
    while (sum > required && i < w.size()) {
        if (w[i] > 1 && (float)w[i] > fw[i]) {
            w[i] -= 1;
            sum -= 1;
        }
        i++;
    }
    i = 0;
    while (sum < required && i < w.size()) {
        if ((float)w[i] < fw[i]) {
            w[i] += 1;
            sum += 1;
        }
        i++;
    }
    return w;
}

Bitlen quantize_weights(std::vector<uint32_t> &counts, size_t total, Bitlen max_size_log) {
    if (counts.size() == 1) {
        counts[0] = 1;
        return 0;
    }
    const Bitlen min_size_log = 64 - (Bitlen)__builtin_clzll((unsigned long long)(counts.size() - 1));
    Bitlen size_log = std::max(min_size_log, max_size_log);
    std::vector<uint32_t> w = quantize_weights_to(counts, total, size_log);
    Bitlen p2 = 32;
    for (uint32_t x : w) p2 = std::min(p2, tzc<uint32_t>(x));
    size_log -= p2;
    for (auto &x : w) x >>= p2;
    counts = w;
    return size_log;
}

// Spec::from_weights: the state -> symbol table; false when the weights do
// not sum to the table size
bool spread_state_symbols(Bitlen size_log, std::vector<uint32_t> &weights, std::vector<uint32_t> &states) {
    if (weights.empty()) weights.push_back(1);
    uint64_t ts = 0;
    for (uint32_t w : weights) ts += w;
    if (size_log > 31 || ts != (1ull << size_log)) return false;
    const uint32_t table = (uint32_t)ts;
    uint32_t stride = (3 * table) / 5;
    if (stride % 2 == 0) stride += 1;
    const uint32_t mod = table - 1;
    states.assign(table, 0);
    uint32_t step = 0;
    for (size_t s = 0; s < weights.size(); s++)
        for (uint32_t k = 0; k < weights[s]; k++) {
            states[(stride * step) & mod] = (uint32_t)s;
            step++;
        }
    return true;
}

struct AnsEncoder {
    Bitlen size_log = 0;
    std::vector<uint32_t> cutoff, minbits, base, weight, next;
    void build(Bitlen sl, const std::vector<uint32_t> &weights, const std::vector<uint32_t> &states) {
        size_log = sl;
        const size_t ns = weights.size();
        cutoff.resize(ns);
        minbits.resize(ns);
        base.resize(ns);
        weight = weights;
        uint32_t acc = 0;
        for (size_t s = 0; s < ns; s++) {
            const uint32_t w = weights[s];
            const uint32_t max_x_s = 2 * w - 1;
            const Bitlen mb = sl - (31 - (Bitlen)__builtin_clz(max_x_s));
            minbits[s] = mb;
            cutoff[s] = 2 * w * (1u << mb);
            base[s] = acc;
            acc += w;
        }
        next.assign(acc, 0);
        std::vector<uint32_t> fill(ns, 0);
        const uint32_t table = 1u << sl;
        for (size_t i = 0; i < states.size(); i++) {
            const uint32_t s = states[i];
            next[base[s] + fill[s]++] = table + (uint32_t)i;
        }
    }
    uint32_t default_state() const { return 1u << size_log; }
    std::pair<uint32_t, Bitlen> encode(uint32_t state, uint32_t sym) const {
        const Bitlen rb = state >= cutoff[sym] ? minbits[sym] + 1 : minbits[sym];
        return {next[base[sym] + ((state >> rb) - weight[sym])], rb};
    }
};

template <class L> struct Bin {
    uint32_t weight;
    L lower;
    Bitlen offset_bits;
};

// ---------------------------------------------------------------------------
// chunk latent compressor (chunk_latent_compressor.rs) and the chunk
// compressor (wrapped/chunk_compressor.rs)
// ---------------------------------------------------------------------------
template <class L> struct VarC {
    bool present = false;
    std::vector<L> latents;
    size_t start = 0, end = 0;  // stored range
    std::vector<L> delta_state;
    VarEnc enc;
    std::vector<Info<L>> infos;
    std::vector<Bin<L>> bins;
    std::vector<uint32_t> counts;
    Bitlen asl = 0;
    AnsEncoder ans;
    double avg_bits = 0.0;
    bool trivial = true, needs_ans = false;
    std::vector<L> search_lowers;
    size_t ssl = 0;

    void train(Bitlen ubl) {
        infos.clear();
        counts.clear();
        asl = 0;
        const size_t n = end - start;
        if (n) {
            std::vector<L> lat(latents.begin() + start, latents.begin() + end);
            std::vector<HBin<L>> h = histogram(lat.data(), n, ubl);
            const Bitlen n_log_ceil = n <= 1 ? 0 : ilog2_64((uint64_t)n - 1) + 1;
            const Bitlen est = std::min(std::min<Bitlen>(ubl + 2, MAX_LEVEL), n_log_ceil);
            infos = optimize_bins(h, est);
            std::vector<uint32_t> w;
            for (const auto &i : infos) w.push_back(i.weight);
            counts = w;
            asl = quantize_weights(w, n, est);
            for (size_t i = 0; i < infos.size(); i++) infos[i].weight = w[i];
        }
        bins.clear();
        for (const auto &i : infos) bins.push_back(Bin<L>{i.weight, i.lower, i.offset_bits});
        needs_ans = bins.size() != 1;
        std::sort(infos.begin(), infos.end(), [](const Info<L> &a, const Info<L> &b) { return a.lower < b.lower; });
        ssl = infos.size() <= 1 ? 0 : 1 + ilog2_64(infos.size() - 1);
        search_lowers.clear();
        for (const auto &i : infos) search_lowers.push_back(i.lower);
        while (search_lowers.size() < ((size_t)1 << ssl)) search_lowers.push_back(LMAX<L>);
        std::vector<uint32_t> weights;
        for (const auto &b : bins) weights.push_back(b.weight);
        std::vector<uint32_t> states;
        spread_state_symbols(asl, weights, states);
        ans.build(asl, weights, states);
        const double tw = (double)(1ull << asl);
        double s = 0.0;
        for (const auto &b : bins) {
            const double ab = (double)asl - std::log2((double)b.weight);
            s += (ab + (double)b.offset_bits) * (double)b.weight / tw;
        }
        avg_bits = s;
        trivial = bins.empty() || (bins.size() == 1 && bins[0].offset_bits == 0);
    }
    size_t meta_bits() const { return 4 + 15 + bins.size() * (size_t)bin_exact_bits<L>(asl); }
    size_t page_meta_bits() const { return (size_t)asl * INTERLEAVE + (size_t)BITS<L> * enc.n_latents_per_state(); }

    // dissected page (compression_intermediates.rs PageDissectedVar)
    std::vector<uint32_t> ans_vals;
    std::vector<Bitlen> ans_bits, off_bits;
    std::vector<L> offsets;
    uint32_t finals[INTERLEAVE];

    void dissect() {
        for (int j = 0; j < INTERLEAVE; j++) finals[j] = ans.default_state();
        const size_t n = trivial ? 0 : end - start;
        ans_vals.assign(n, 0);
        ans_bits.assign(n, 0);
        off_bits.assign(n, 0);
        offsets.assign(n, 0);
        if (!n) return;
        const bool table_trivial = infos.size() <= 1;
        const L default_lower = table_trivial && !infos.empty() ? infos[0].lower : 0;
        L lowers[BATCH];
        uint32_t symbols[BATCH];
        for (size_t i = 0; i < BATCH; i++) {
            lowers[i] = default_lower;
            symbols[i] = 0;
        }
        const size_t nb = (n + BATCH - 1) / BATCH;
        for (size_t b = nb; b-- > 0;) {
            const size_t r0 = b * BATCH, r1 = std::min(n, r0 + BATCH), m = r1 - r0;
            const L *lat = latents.data() + start + r0;
            if (table_trivial) {
                const Bitlen ob = infos.empty() ? 0 : infos[0].offset_bits;
                for (size_t i = 0; i < m; i++) off_bits[r0 + i] = ob;
            } else {
                for (size_t i = 0; i < m; i++) {
                    size_t idx = 0;
                    for (size_t d = 0; d < ssl; d++) {
                        const size_t bis = (size_t)1 << (ssl - 1 - d);
                        idx += (lat[i] >= search_lowers[idx + bis]) ? bis : 0;
                    }
                    idx = std::min(idx, infos.size() - 1);
                    const Info<L> &inf = infos[idx];
                    lowers[i] = inf.lower;
                    symbols[i] = inf.symbol;
                    off_bits[r0 + i] = inf.offset_bits;
                }
            }
            for (size_t i = 0; i < m; i++) offsets[r0 + i] = (L)(lat[i] - lowers[i]);
            // encode_ans_in_reverse
            if (ans.size_log == 0) continue;  // ans_bits stay 0
            const size_t fbase = (m / INTERLEAVE) * INTERLEAVE, fj = m % INTERLEAVE;
            auto step = [&](size_t i, int j) {
                const auto r = ans.encode(finals[j], symbols[i]);
                ans_vals[r0 + i] = finals[j] & ((1u << r.second) - 1);
                ans_bits[r0 + i] = r.second;
                finals[j] = r.first;
            };
            for (size_t j = fj; j-- > 0;) step(fbase + j, (int)j);
            for (size_t bi = fbase; bi > 0;) {
                bi -= INTERLEAVE;
                for (int j = INTERLEAVE; j-- > 0;) step(bi + j, j);
            }
        }
    }
    void write_batch(size_t bs, BitWriter &w) const {
        if (bs >= offsets.size()) return;
        const size_t be = std::min(offsets.size(), bs + BATCH);
        if (needs_ans)
            for (size_t i = bs; i < be; i++) w.put(ans_vals[i], ans_bits[i]);
        for (size_t i = bs; i < be; i++) w.put((uint64_t)offsets[i], off_bits[i]);
    }
    void write_meta(BitWriter &w) const {
        w.put(asl, 4);
        w.put(bins.size(), 15);
        for (const auto &b : bins) {
            w.put(b.weight - 1, asl);
            w.put((uint64_t)b.lower, BITS<L>);
            w.put(b.offset_bits, OBB<L>);
        }
    }
    void write_page_meta(BitWriter &w) const {
        for (L x : delta_state) w.put((uint64_t)x, BITS<L>);
        for (int j = 0; j < INTERLEAVE; j++) w.put(finals[j] - ans.default_state(), asl);
    }
};

template <class L> struct Chunk {
    Mode mode;
    DeltaEnc de;
    size_t n = 0;
    VarC<uint32_t> dv;  // lookbacks
    VarC<L> pv, sv;

    size_t mode_max_bits() const {
        switch (mode.kind) {
        case M_INT_MULT:
        case M_FLOAT_MULT: return 4 + BITS<L>;
        case M_FLOAT_QUANT: return 4 + 8;
        default: return 4;
        }
    }
    size_t meta_max_size() const {
        size_t b = mode_max_bits() + DELTA_MAX_BITS + pv.meta_bits();
        if (dv.present) b += dv.meta_bits();
        if (sv.present) b += sv.meta_bits();
        return (b + 7) / 8;
    }
    size_t page_meta_size() const {
        size_t b = pv.page_meta_bits();
        if (dv.present) b += dv.page_meta_bits();
        if (sv.present) b += sv.page_meta_bits();
        return (b + 7) / 8;
    }
    size_t page_size_hint(double over) const {
        size_t body = 0;
        auto add = [&](size_t stored, double avg) {
            const double nb = (double)stored * avg;
            body += (size_t)std::ceil(nb * over);
        };
        if (dv.present) add(dv.end - dv.start, dv.avg_bits);
        add(pv.end - pv.start, pv.avg_bits);
        if (sv.present) add(sv.end - sv.start, sv.avg_bits);
        return page_meta_size() + (body + 7) / 8;
    }
};

template <class L>
void new_candidate(Chunk<L> &c, std::vector<L> &&primary, std::vector<L> *secondary, const Mode &mode,
                   const DeltaEnc &de, Bitlen ubl) {
    c.mode = mode;
    c.de = de;
    c.n = primary.size();
    const size_t n = c.n;
    std::vector<uint32_t> lbs;
    if (de.kind == D_LOOKBACK) lbs = choose_lookbacks<L>(de.window_n_log, de.state_n_log, primary.data(), n);
    c.pv = VarC<L>();
    c.pv.present = true;
    c.pv.enc = for_var(de, K_PRIMARY);
    c.pv.latents = std::move(primary);
    auto encode = [&](VarC<L> &v) {
        switch (v.enc.kind) {
        case D_CONSEC: v.delta_state = consec_encode<L>(v.enc.order, v.latents.data(), n); break;
        case D_LOOKBACK: v.delta_state = lookback_encode<L>(v.enc.state_n_log, lbs, v.latents.data(), n); break;
        default: break;
        }
        v.start = std::min(v.enc.n_latents_per_state(), n);
        v.end = n;
    };
    encode(c.pv);
    c.sv = VarC<L>();
    if (secondary) {
        c.sv.present = true;
        c.sv.enc = for_var(de, K_SECONDARY);
        c.sv.latents = std::move(*secondary);
        encode(c.sv);
    }
    c.dv = VarC<uint32_t>();
    if (de.kind == D_LOOKBACK) {
        c.dv.present = true;
        c.dv.latents = std::move(lbs);
        c.dv.start = 0;
        c.dv.end = c.dv.latents.size();
    }
    if (c.dv.present) c.dv.train(ubl);
    c.pv.train(ubl);
    if (c.sv.present) c.sv.train(std::min(ubl, LIMITED_UBL));
}

template <class L> float sample_size(const std::vector<L> &sample, Bitlen ubl, const DeltaEnc &de) {
    Chunk<L> c;
    std::vector<L> s(sample);
    new_candidate<L>(c, std::move(s), nullptr, Mode(), de, ubl);
    return (float)(c.meta_max_size() + c.page_size_hint(1.0));
}

template <class L> DeltaEnc choose_auto_delta(const std::vector<L> &primary, Bitlen ubl) {
    const size_t n = primary.size();
    const size_t group = 200, n_extra = 1 + n / 10000;
    const size_t nominal = (n_extra + 1) * group;
    const size_t padding = n_extra == 0 ? 0 : (n > nominal ? n - nominal : 0) / n_extra;
    std::vector<L> sample;
    for (size_t i = 0; i < std::min(group, n); i++) sample.push_back(primary[i]);
    size_t i = group;
    for (size_t g = 0; g < n_extra; g++) {
        i += padding;
        for (size_t q = i; q < std::min(n, i + group); q++) sample.push_back(primary[q]);
        i += group;
    }
    const size_t sn = sample.size();
    DeltaEnc best;
    float best_cost = sample_size<L>(sample, ubl, best);
    const float penalty = 0.25f * (float)sn;
    if (best_cost > penalty) {
        const float lc = sample_size<L>(sample, ubl, new_lookback(sn)) + penalty;
        if (lc < best_cost) {
            best = new_lookback(n);
            best_cost = lc;
        }
    }
    for (Bitlen order = 1; order <= 7; order++) {
        DeltaEnc d;
        d.kind = D_CONSEC;
        d.order = order;
        const float cost = sample_size<L>(sample, ubl, d);
        if (cost < best_cost) {
            best = d;
            best_cost = cost;
        } else {
            break;
        }
    }
    return best;
}

Bitlen choose_ubl(Bitlen level, size_t n) {
    const Bitlen log_n = (Bitlen)std::floor(std::log2((double)n));
    const Bitlen fast = log_n > 4 ? log_n - 4 : 0;
    return level <= fast ? level : fast + (level - fast) / 2;
}

template <class L> bool should_fallback(const Chunk<L> &c) {
    if (c.de.kind == D_NOOP && c.mode.kind == M_CLASSIC) return false;
    size_t worst = 7;
    auto add = [&](const auto &v) {
        for (size_t i = 0; i < v.bins.size(); i++)
            worst += (size_t)v.counts[i] *
                     (size_t)(v.bins[i].offset_bits + v.asl - (31 - (Bitlen)__builtin_clz(v.bins[i].weight)));
    };
    if (c.dv.present) add(c.dv);
    add(c.pv);
    if (c.sv.present) add(c.sv);
    const size_t wc = c.meta_max_size() + c.page_meta_size() + (worst + 7) / 8;
    const size_t baseline_meta = (4 + DELTA_MAX_BITS + 4 + 15 + bin_exact_bits<L>(0) + 7) / 8;
    const size_t baseline = baseline_meta + (c.n * BITS<L> + 7) / 8;
    return wc > baseline;
}

template <class L> void fallback_chunk(Chunk<L> &c, std::vector<L> &&classic) {
    c.mode = Mode();
    c.de = DeltaEnc();
    c.n = classic.size();
    c.dv = VarC<uint32_t>();
    c.sv = VarC<L>();
    c.pv = VarC<L>();
    VarC<L> &v = c.pv;
    v.present = true;
    v.latents = std::move(classic);
    v.start = 0;
    v.end = c.n;
    v.infos = {Info<L>{1, 0, LMAX<L>, BITS<L>, 0}};
    v.bins = {Bin<L>{1, 0, BITS<L>}};
    v.counts = {(uint32_t)c.n};
    v.asl = 0;
    v.needs_ans = false;
    v.trivial = false;
    v.ssl = 0;
    v.search_lowers = {0};
    std::vector<uint32_t> w{1}, st;
    spread_state_symbols(0, w, st);
    v.ans.build(0, w, st);
    v.avg_bits = (double)BITS<L>;
}

// mode selection of each number type, then the chunk compressor
template <class T> struct Split {
    Mode mode;
    std::vector<typename NT<T>::L> primary, secondary;
    bool has_secondary = false;
};

template <class T> void classic_split(const T *nums, size_t n, Split<T> &s) {
    s.mode = Mode();
    s.primary.resize(n);
    for (size_t i = 0; i < n; i++) s.primary[i] = NT<T>::to_lat(nums[i]);
    s.has_secondary = false;
}

template <class T> bool choose_mode_int(const T *nums, size_t n, Split<T> &s) {
    using L = typename NT<T>::L;
    L base;
    if (int_choose_base<T>(nums, n, &base)) {
        s.mode.kind = M_INT_MULT;
        s.mode.base = base;
        s.primary.resize(n);
        s.secondary.resize(n);
        for (size_t i = 0; i < n; i++) {
            const L u = NT<T>::to_lat(nums[i]);
            s.primary[i] = u / base;
            s.secondary[i] = u % base;
        }
        s.has_secondary = true;
    } else {
        classic_split(nums, n, s);
    }
    return true;
}

template <class F> bool choose_mode_float(const F *nums, size_t n, Split<F> &s) {
    using L = typename FT<F>::L;
    std::vector<F> sample;
    const F maxs = std::numeric_limits<F>::max() * (F)0.5;
    const bool have = choose_sample<F>(n, [&](size_t i, F &v) {
        const F x = nums[i];
        if (!fnormal(x)) return false;
        const F a = std::fabs(x);
        if (a <= maxs) {
            v = a;
            return true;
        }
        return false;
    }, sample);
    // bids: classic (0), float mult, float quant; the last maximum wins
    int win = M_CLASSIC;
    double win_bits = 0.0;
    FMConfig<F> fmc{};
    Bitlen qk = 0;
    if (have) {
        FMConfig<F> c;
        double b;
        if (fm_choose_config(sample, &c) && fm_bits_saved(c, sample, &b)) {
            if (total_key(b) >= total_key(win_bits)) {
                win = M_FLOAT_MULT;
                win_bits = b;
                fmc = c;
            }
        }
        double qb;
        const Bitlen k = fq_best_k(sample, &qb);
        const double saved = est_bits_saved_per_num<L>(sample, [&](F x) {
            return std::pair<L, double>((L)(fbits<F>(x) >> k), qb);
        });
        if (saved > 1.5 && total_key(saved) >= total_key(win_bits)) {
            win = M_FLOAT_QUANT;
            win_bits = saved;
            qk = k;
        }
    }
    if (win == M_CLASSIC) {
        classic_split(nums, n, s);
        return true;
    }
    s.primary.resize(n);
    s.secondary.resize(n);
    s.has_secondary = true;
    if (win == M_FLOAT_MULT) {
        s.mode.kind = M_FLOAT_MULT;
        s.mode.base = NT<F>::to_lat(fmc.base);
        // mode_is_valid (float.rs:213-218)
        if (!(std::isfinite(fmc.base) && std::fabs(fmc.base) > (F)0)) return false;
        for (size_t i = 0; i < n; i++) {
            const F x = nums[i];
            const F mult = FT<F>::rnd(x * fmc.inv_base);
            s.primary[i] = int_float_to_latent<F>(mult);
            s.secondary[i] = (L)((L)(NT<F>::to_lat(x) - NT<F>::to_lat(mult * fmc.base)) + MID<L>);
        }
    } else {
        s.mode.kind = M_FLOAT_QUANT;
        s.mode.k = qk;
        if (!(qk > 0 && qk <= FT<F>::PREC)) return false;
        const L lmax = (L)(((L)1 << qk) - 1);
        for (size_t i = 0; i < n; i++) {
            const L u = NT<F>::to_lat(nums[i]);
            s.primary[i] = (L)(u >> qk);
            const L low = u & lmax;
            s.secondary[i] = !std::signbit(nums[i]) ? low : (L)(lmax - low);
        }
    }
    return true;
}

template <class T> bool choose_mode(const T *nums, size_t n, Split<T> &s) { return choose_mode_int<T>(nums, n, s); }
template <> bool choose_mode<float>(const float *nums, size_t n, Split<float> &s) {
    return choose_mode_float<float>(nums, n, s);
}
template <> bool choose_mode<double>(const double *nums, size_t n, Split<double> &s) {
    return choose_mode_float<double>(nums, n, s);
}

template <class L> void write_chunk_body(Chunk<L> &c, BitWriter &w) {
    // chunk metadata (metadata/chunk.rs write_to)
    w.put((uint64_t)c.mode.kind, 4);
    if (c.mode.kind == M_INT_MULT || c.mode.kind == M_FLOAT_MULT) w.put(c.mode.base, BITS<L>);
    else if (c.mode.kind == M_FLOAT_QUANT) w.put(c.mode.k, 8);
    w.put((uint64_t)c.de.kind, 4);
    if (c.de.kind == D_CONSEC) {
        w.put(c.de.order, 3);
        w.put(c.de.sec_uses ? 1 : 0, 1);
    } else if (c.de.kind == D_LOOKBACK) {
        w.put(c.de.window_n_log - 1, 5);
        w.put(c.de.state_n_log, 4);
        w.put(c.de.sec_uses ? 1 : 0, 1);
    }
    if (c.dv.present) c.dv.write_meta(w);
    c.pv.write_meta(w);
    if (c.sv.present) c.sv.write_meta(w);
    w.finish_byte();
    // page (wrapped/chunk_compressor.rs write_page)
    if (c.dv.present) c.dv.dissect();
    c.pv.dissect();
    if (c.sv.present) c.sv.dissect();
    if (c.dv.present) c.dv.write_page_meta(w);
    c.pv.write_page_meta(w);
    if (c.sv.present) c.sv.write_page_meta(w);
    w.finish_byte();
    for (size_t bs = 0; bs < c.n; bs += BATCH) {
        if (c.dv.present) c.dv.write_batch(bs, w);
        c.pv.write_batch(bs, w);
        if (c.sv.present) c.sv.write_batch(bs, w);
    }
    w.finish_byte();
}

template <class T> bool compress_chunk(const T *nums, size_t n, Bitlen level, BitWriter &w) {
    using L = typename NT<T>::L;
    if (n == 0 || n > MAX_ENTRIES || level > MAX_LEVEL) return false;
    Split<T> s;
    if (!choose_mode<T>(nums, n, s)) return false;
    const Bitlen ubl = choose_ubl(level, n);
    const DeltaEnc de = choose_auto_delta<L>(s.primary, ubl);
    Chunk<L> c;
    new_candidate<L>(c, std::move(s.primary), s.has_secondary ? &s.secondary : nullptr, s.mode, de, ubl);
    if (should_fallback(c)) {
        std::vector<L> cl(n);
        for (size_t i = 0; i < n; i++) cl[i] = NT<T>::to_lat(nums[i]);
        fallback_chunk<L>(c, std::move(cl));
    }
    w.byte(NT<T>::BYTE);
    w.put(n - 1, 24);
    write_chunk_body(c, w);
    return true;
}

template <class T> bool simple_compress(const T *nums, size_t n, Bitlen level, std::vector<uint8_t> &out) {
    out.clear();
    BitWriter w(out);
    const uint8_t magic[4] = {112, 99, 111, 33};
    for (uint8_t b : magic) w.byte(b);
    w.put(3, 8);  // standalone version
    w.byte(0);    // no uniform number type
    const Bitlen power = n == 0 ? 1 : ilog2_64(n) + 1;
    w.put(power - 1, 6);
    w.put(low_bits(n, power), power);
    w.finish_byte();
    w.byte(4);  // format version 4.1
    w.byte(1);
    if (n) {
        // EqualPagesUpTo(2^18): one chunk per page
        const size_t np = (n + MAX_PAGE_N - 1) / MAX_PAGE_N;
        const size_t lo = n / np, r = n % np;
        size_t start = 0;
        for (size_t p = 0; p < np; p++) {
            const size_t pn = lo + (p < r ? 1 : 0);
            if (!compress_chunk<T>(nums + start, pn, level, w)) return false;
            start += pn;
        }
    }
    w.byte(0);  // termination
    return true;
}

// ---------------------------------------------------------------------------
// decompressor
// ---------------------------------------------------------------------------
template <class L> struct VarD {
    bool present = false;
    Bitlen asl = 0;
    std::vector<Bin<L>> bins;
    VarEnc enc;
    // ans decoder (ans/decoding.rs)
    std::vector<uint32_t> nbase;
    std::vector<uint8_t> nread, noff;
    std::vector<L> lowers;
    size_t bytes_per_offset = 0;
    // page state
    uint32_t st[INTERLEAVE];
    std::vector<L> dstate;  // moments, or the lookback window buffer
    size_t dpos = 0;
    L scratch[BATCH];
    Bitlen sob[BATCH];

    bool read_meta(BitReader &r) {
        asl = (Bitlen)r.get(4);
        const size_t nb = (size_t)r.get(15);
        if (!r.ok()) return false;
        if ((1ull << asl) < nb) return false;
        if (nb == 1 && asl > 0) return false;
        if (asl > MAX_ANS_BITS) return false;
        bins.resize(nb);
        for (size_t i = 0; i < nb; i++) {
            bins[i].weight = (uint32_t)r.get(asl) + 1;
            bins[i].lower = (L)r.get(BITS<L>);
            bins[i].offset_bits = (Bitlen)r.get(OBB<L>);
            if (bins[i].offset_bits > BITS<L>) return false;
        }
        return r.ok();
    }
    bool build() {
        Bitlen mob = 0;
        for (const auto &b : bins) mob = std::max(mob, b.offset_bits);
        bytes_per_offset = mob == 0 ? 0 : (mob + 14) / 8;
        std::vector<uint32_t> w;
        for (const auto &b : bins) w.push_back(b.weight);
        std::vector<uint32_t> states;
        if (!spread_state_symbols(asl, w, states)) return false;
        const uint32_t table = 1u << asl;
        std::vector<uint32_t> xs = w;
        nbase.resize(table);
        nread.resize(table);
        noff.resize(table);
        lowers.resize(table);
        for (uint32_t i = 0; i < table; i++) {
            const uint32_t s = states[i];
            const uint32_t nsb = xs[s];
            const Bitlen br = lzc<uint32_t>(nsb) - lzc<uint32_t>(table);
            nbase[i] = (nsb << br) - table;
            nread[i] = (uint8_t)br;
            noff[i] = (uint8_t)(s < bins.size() ? bins[s].offset_bits : 0);
            lowers[i] = s < bins.size() ? bins[s].lower : 0;
            xs[s] += 1;
        }
        for (size_t i = 0; i < BATCH; i++) {
            scratch[i] = bins.size() == 1 ? bins[0].lower : 0;
            sob[i] = bins.size() == 1 ? bins[0].offset_bits : 0;
        }
        return true;
    }
    bool read_page_meta(BitReader &r) {
        const size_t ns = enc.n_latents_per_state();
        std::vector<L> stored(ns);
        for (size_t i = 0; i < ns; i++) stored[i] = (L)r.get(BITS<L>);
        for (int j = 0; j < INTERLEAVE; j++) st[j] = (uint32_t)r.get(asl);
        if (enc.kind == D_LOOKBACK) {
            const size_t wn = (size_t)1 << enc.window_n_log;
            dstate.assign(std::max(wn, BATCH) * 2, 0);
            for (size_t i = 0; i < ns; i++) dstate[wn - ns + i] = stored[i];
            dpos = wn;
        } else {
            dstate = stored;
            dpos = 0;
        }
        return r.ok();
    }
    void read_pre_delta(BitReader &r, size_t m) {
        if (m == 0) return;
        if (bins.size() > 1) {
            for (size_t i = 0; i < m; i++) {
                const int j = (int)(i % INTERLEAVE);
                const uint32_t s = st[j];
                const Bitlen br = nread[s];
                const uint32_t v = (uint32_t)r.get(br);
                scratch[i] = lowers[s];
                sob[i] = noff[s];
                st[j] = nbase[s] + v;
            }
        } else {
            for (size_t i = 0; i < m; i++) scratch[i] = lowers[0];
        }
        if (bytes_per_offset)
            for (size_t i = 0; i < m; i++) scratch[i] = (L)(scratch[i] + (L)r.get(sob[i]));
    }
    // read_batch (page_latent_decompressor.rs:236-258); false on a lookback
    // past the window
    bool read_batch(BitReader &r, const uint32_t *lookbacks, size_t n_remaining) {
        const size_t nlps = enc.n_latents_per_state();
        const size_t pre = std::min(BATCH, n_remaining > nlps ? n_remaining - nlps : 0);
        read_pre_delta(r, pre);
        const size_t m = std::min(n_remaining, BATCH);
        if (enc.kind == D_CONSEC) {
            consec_decode<L>(dstate, scratch, m);
        } else if (enc.kind == D_LOOKBACK) {
            toggle_center(scratch, m);
            const size_t wn = (size_t)1 << enc.window_n_log, sn = (size_t)1 << enc.state_n_log;
            size_t sp = dpos;
            if (sp + m > dstate.size()) {
                std::memmove(dstate.data(), dstate.data() + sp - wn, wn * sizeof(L));
                sp = wn;
            }
            bool oob = false;
            for (size_t i = 0; i < m; i++) {
                size_t lb = lookbacks[i];
                if (lb > wn) {
                    oob = true;
                    lb = 1;
                }
                dstate[sp + i] = (L)(scratch[i] + dstate[sp + i - lb]);
            }
            for (size_t i = 0; i < m; i++) scratch[i] = dstate[sp - sn + i];
            dpos = sp + m;
            if (oob) return false;
        }
        return true;
    }
};

template <class T> int decode_chunks(const uint8_t *src, size_t len, size_t n_expected, std::vector<T> &out) {
    using L = typename NT<T>::L;
    if (len > ((size_t)1 << 40)) return -1;
    std::vector<uint8_t> buf(len + 32, 0);
    if (len) memcpy(buf.data(), src, len);
    BitReader r{buf.data(), len};
    if (len < 4 || memcmp(buf.data(), "pco!", 4) != 0) return -1;
    r.bit = 32;
    const uint64_t sver = r.get(8);
    bool has_uniform = false;
    uint8_t uniform = 0;
    if (sver < 2) {
        r.bit -= 8;
    } else {
        if (sver >= 3) {
            const uint8_t b = (uint8_t)r.get(8);
            if (b != 0) {
                if (!(b >= 1 && b <= 11 && b != 9)) return -1;
                has_uniform = true;
                uniform = b;
            }
        }
        const Bitlen power = 1 + (Bitlen)r.get(6);
        r.get(power);
        if (!r.drain_empty_byte()) return -1;
    }
    if (sver > 3) return -1;
    // wrapped header: format version
    const uint8_t major = (uint8_t)r.get(8);
    if (major >= 4) r.get(8);
    if (!r.ok() || major > 4) return -1;
    out.clear();
    for (;;) {
        const uint8_t tb = (uint8_t)r.get(8);
        if (!r.ok()) return -1;
        if (tb == 0) break;
        if (has_uniform && uniform != tb) return -1;
        if (tb != NT<T>::BYTE) return -1;
        const size_t n = (size_t)r.get(24) + 1;
        // chunk metadata
        Mode mode;
        mode.kind = (int)r.get(4);
        switch (mode.kind) {
        case M_CLASSIC: break;
        case M_INT_MULT:
            if (major == 0) return -1;
            mode.base = r.get(BITS<L>);
            break;
        case M_FLOAT_MULT: mode.base = r.get(BITS<L>); break;
        case M_FLOAT_QUANT: mode.k = (Bitlen)r.get(8); break;
        case M_DICT: {
            const size_t nu = (size_t)r.get(25);
            if (!r.drain_empty_byte()) return -1;
            if (nu > len * 8) return -1;
            mode.dict.resize(nu);
            for (size_t i = 0; i < nu; i++) mode.dict[i] = r.get(BITS<L>);
            break;
        }
        default: return -1;
        }
        DeltaEnc de;
        if (major < 3) {
            const Bitlen order = (Bitlen)r.get(3);
            if (order) {
                de.kind = D_CONSEC;
                de.order = order;
            }
        } else {
            de.kind = (int)r.get(4);
            switch (de.kind) {
            case D_NOOP: break;
            case D_CONSEC:
                de.order = (Bitlen)r.get(3);
                if (de.order == 0) return -1;
                de.sec_uses = r.get(1) != 0;
                break;
            case D_LOOKBACK:
                de.window_n_log = 1 + (Bitlen)r.get(5);
                de.state_n_log = (Bitlen)r.get(4);
                if (de.state_n_log > de.window_n_log) return -1;
                de.sec_uses = r.get(1) != 0;
                break;
            default: return -1;  // Conv1: never produced by the Auto delta spec
            }
        }
        VarD<uint32_t> dv;
        VarD<uint32_t> pv32;  // Dict primary (u32 indices)
        VarD<L> pv, sv;
        const bool dict = mode.kind == M_DICT;
        dv.present = de.kind == D_LOOKBACK;
        if (dv.present && !dv.read_meta(r)) return -1;
        if (dict) {
            pv32.present = true;
            if (!pv32.read_meta(r)) return -1;
        } else {
            pv.present = true;
            if (!pv.read_meta(r)) return -1;
        }
        sv.present = mode.kind == M_INT_MULT || mode.kind == M_FLOAT_MULT || mode.kind == M_FLOAT_QUANT;
        if (sv.present && !sv.read_meta(r)) return -1;
        if (!r.drain_empty_byte()) return -1;
        if (dv.present) {
            const uint32_t wn = 1u << de.window_n_log;
            for (const auto &b : dv.bins)
                if (b.lower < 1 || b.lower > wn) return -1;
        }
        // mode_is_valid
        if (mode.kind == M_INT_MULT && (L)mode.base == 0) return -1;
        if (mode.kind == M_FLOAT_MULT || mode.kind == M_FLOAT_QUANT) {
            if (!std::is_floating_point<T>::value) return -1;
        }
        if (mode.kind == M_INT_MULT && std::is_floating_point<T>::value) return -1;
        dv.enc = for_var(de, K_DELTA);
        pv.enc = pv32.enc = for_var(de, K_PRIMARY);
        sv.enc = for_var(de, K_SECONDARY);
        if (dv.present && !dv.build()) return -1;
        if (pv.present && !pv.build()) return -1;
        if (pv32.present && !pv32.build()) return -1;
        if (sv.present && !sv.build()) return -1;
        // page
        if (dv.present && !dv.read_page_meta(r)) return -1;
        if (pv.present && !pv.read_page_meta(r)) return -1;
        if (pv32.present && !pv32.read_page_meta(r)) return -1;
        if (sv.present && !sv.read_page_meta(r)) return -1;
        if (!r.drain_empty_byte()) return -1;
        const size_t nlps = pv.present ? pv.enc.n_latents_per_state() : pv32.enc.n_latents_per_state();
        const size_t n_in_body = n > nlps ? n - nlps : 0;
        if (n_in_body > 0) {
            if (dv.present && dv.bins.empty()) return -1;
            if (pv.present && pv.bins.empty()) return -1;
            if (pv32.present && pv32.bins.empty()) return -1;
            if (sv.present && sv.bins.empty()) return -1;
        }
        if (out.size() + n > n_expected) return -1;  // more numbers than the frame declares
        const size_t base = out.size();
        out.resize(base + n);
        size_t rem = n;
        while (rem) {
            const size_t bn = std::min(rem, BATCH);
            if (dv.present) dv.read_pre_delta(r, std::min(rem > nlps ? rem - nlps : 0, bn));
            bool okb = true;
            if (pv.present) okb = pv.read_batch(r, dv.scratch, rem);
            else okb = pv32.read_batch(r, dv.scratch, rem);
            if (sv.present && okb) okb = sv.read_batch(r, dv.scratch, rem);
            if (!okb || !r.ok()) return -1;
            T *dst = out.data() + base + (n - rem);
            switch (mode.kind) {
            case M_CLASSIC:
                for (size_t i = 0; i < bn; i++) dst[i] = NT<T>::from_lat(pv.scratch[i]);
                break;
            case M_INT_MULT:
                for (size_t i = 0; i < bn; i++)
                    dst[i] = NT<T>::from_lat((L)((L)(pv.scratch[i] * (L)mode.base) + sv.scratch[i]));
                break;
            case M_DICT:
                for (size_t i = 0; i < BATCH; i++)
                    if (pv32.scratch[i] >= mode.dict.size()) return -1;
                for (size_t i = 0; i < bn; i++) dst[i] = NT<T>::from_lat((L)mode.dict[pv32.scratch[i]]);
                break;
            default:
                if (!join_float(mode, pv.scratch, sv.scratch, dst, bn)) return -1;
                break;
            }
            rem -= bn;
        }
        if (!r.drain_empty_byte()) return -1;
    }
    if (out.size() != n_expected) return -1;
    return 0;
}

template <class T, class L>
bool join_float(const Mode &, const L *, const L *, T *, size_t) { return false; }

template <class F>
bool join_float_impl(const Mode &mode, const typename FT<F>::L *p, const typename FT<F>::L *s, F *dst, size_t bn) {
    using L = typename FT<F>::L;
    if (mode.kind == M_FLOAT_MULT) {
        const F base = NT<F>::from_lat((L)mode.base);
        if (!(std::isfinite(base) && std::fabs(base) > (F)0)) return false;
        for (size_t i = 0; i < bn; i++) {
            const F un = int_float_from_latent<F>(p[i]) * base;
            dst[i] = NT<F>::from_lat((L)((L)(NT<F>::to_lat(un) + s[i]) + MID<L>));
        }
        return true;
    }
    const Bitlen k = mode.k;
    if (!(k > 0 && k <= FT<F>::PREC)) return false;
    const L cut = MID<L> >> k;
    const L lmax = (L)(((L)1 << k) - 1);
    for (size_t i = 0; i < bn; i++) {
        const L y = p[i], m = s[i];
        const L low = y >= cut ? m : (L)(lmax - m);
        dst[i] = NT<F>::from_lat((L)((L)(y << k) + low));
    }
    return true;
}
template <> bool join_float<float, uint32_t>(const Mode &m, const uint32_t *p, const uint32_t *s, float *d, size_t n) {
    return join_float_impl<float>(m, p, s, d, n);
}
template <> bool join_float<double, uint64_t>(const Mode &m, const uint64_t *p, const uint64_t *s, double *d, size_t n) {
    return join_float_impl<double>(m, p, s, d, n);
}

// quoin lane <-> pco numbers (pcodec.rs:33-76)
template <class T> T from_lane(uint64_t v);
template <> uint32_t from_lane<uint32_t>(uint64_t v) { return (uint32_t)v; }
template <> uint64_t from_lane<uint64_t>(uint64_t v) { return v; }
template <> int32_t from_lane<int32_t>(uint64_t v) { return (int32_t)(uint32_t)v; }
template <> int64_t from_lane<int64_t>(uint64_t v) { return (int64_t)v; }
template <> double from_lane<double>(uint64_t v) { return ffrom<double>(v); }
template <> float from_lane<float>(uint64_t v) { return (float)ffrom<double>(v); }
inline uint64_t to_lane(uint32_t x) { return x; }
inline uint64_t to_lane(uint64_t x) { return x; }
inline uint64_t to_lane(int32_t x) { return (uint64_t)(int64_t)x; }
inline uint64_t to_lane(int64_t x) { return (uint64_t)x; }
inline uint64_t to_lane(double x) { return fbits<double>(x); }
inline uint64_t to_lane(float x) { return fbits<double>((double)x); }

template <class T> bool enc_t(const uint64_t *lane, size_t n, unsigned level, std::vector<uint8_t> &out) {
    std::vector<T> nums(n);
    for (size_t i = 0; i < n; i++) nums[i] = from_lane<T>(lane[i]);
    return simple_compress<T>(nums.data(), n, level, out);
}
template <class T> int dec_t(const uint8_t *src, size_t len, size_t n, uint64_t *lane) {
    std::vector<T> nums;
    if (decode_chunks<T>(src, len, n, nums) != 0) return -1;
    for (size_t i = 0; i < n; i++) lane[i] = to_lane(nums[i]);
    return 0;
}

}  // namespace

namespace qapco {

bool encode_lane(const uint64_t *lane, size_t n, int ntype, unsigned level, std::vector<uint8_t> &out) {
    if (n == 0) return false;
    switch (ntype) {
    case QA_PCO_U32: return enc_t<uint32_t>(lane, n, level, out);
    case QA_PCO_U64: return enc_t<uint64_t>(lane, n, level, out);
    case QA_PCO_I32: return enc_t<int32_t>(lane, n, level, out);
    case QA_PCO_I64: return enc_t<int64_t>(lane, n, level, out);
    case QA_PCO_F32: return enc_t<float>(lane, n, level, out);
    case QA_PCO_F64: return enc_t<double>(lane, n, level, out);
    default: return false;
    }
}

int decode_lane(const uint8_t *src, size_t len, size_t n, int ntype, uint64_t *lane) {
    switch (ntype) {
    case QA_PCO_U32: return dec_t<uint32_t>(src, len, n, lane);
    case QA_PCO_U64: return dec_t<uint64_t>(src, len, n, lane);
    case QA_PCO_I32: return dec_t<int32_t>(src, len, n, lane);
    case QA_PCO_I64: return dec_t<int64_t>(src, len, n, lane);
    case QA_PCO_F32: return dec_t<float>(src, len, n, lane);
    case QA_PCO_F64: return dec_t<double>(src, len, n, lane);
    default: return -1;
    }
}

}  // namespace qapco

// ---------------------------------------------------------------------------
// C entry points
// ---------------------------------------------------------------------------
extern "C" int qa_pco_compress(int ntype, unsigned level, const uint64_t *lane, size_t n, uint8_t *dst,
                               size_t cap, size_t *out_len) {
    if (!out_len || (n && !lane)) return -1;
    std::vector<uint8_t> o;
    if (!qapco::encode_lane(lane, n, ntype, level, o)) return 1;
    *out_len = o.size();
    if (o.size() > cap || !dst) return -2;
    memcpy(dst, o.data(), o.size());
    return 0;
}

extern "C" int qa_pco_decompress(int ntype, const uint8_t *src, size_t len, size_t n, uint64_t *lane) {
    if ((len && !src) || (n && !lane)) return -1;
    return qapco::decode_lane(src, len, n, ntype, lane);
}

extern "C" void qa_pco_xoroshiro(uint64_t seed, uint64_t *out, size_t count) {
    Xoro r(seed);
    for (size_t i = 0; i < count; i++) out[i] = r.next();
}

extern "C" uint32_t qa_pco_quantize_weights(uint32_t *counts, size_t n, uint64_t total, uint32_t max_size_log) {
    std::vector<uint32_t> c(counts, counts + n);
    const Bitlen sl = quantize_weights(c, (size_t)total, max_size_log);
    for (size_t i = 0; i < n && i < c.size(); i++) counts[i] = c[i];
    return sl;
}

extern "C" int qa_pco_spread(uint32_t size_log, const uint32_t *weights, size_t n, uint32_t *out) {
    std::vector<uint32_t> w(weights, weights + n), st;
    if (!spread_state_symbols(size_log, w, st)) return -1;
    memcpy(out, st.data(), st.size() * 4);
    return 0;
}

extern "C" size_t qa_pco_sample_f32(const float *x, size_t n, float *out, size_t cap) {
    std::vector<float> s;
    const float maxs = std::numeric_limits<float>::max() * 0.5f;
    if (!choose_sample<float>(n, [&](size_t i, float &v) {
            if (!fnormal(x[i])) return false;
            const float a = std::fabs(x[i]);
            if (a <= maxs) {
                v = a;
                return true;
            }
            return false;
        }, s))
        return 0;
    for (size_t i = 0; i < s.size() && i < cap; i++) out[i] = s[i];
    return s.size();
}

extern "C" int qa_pco_pair_gcd_f64(double g, double l, double *out) { return approx_pair_gcd<double>(g, l, out) ? 1 : 0; }
extern "C" int qa_pco_pair_gcd_f32(float g, float l, float *out) { return approx_pair_gcd<float>(g, l, out) ? 1 : 0; }
extern "C" void qa_pco_snap_f32(float base, float *ob, float *oi) {
    const FMConfig<float> c = snap_to_int_reciprocal<float>(base);
    *ob = c.base;
    *oi = c.inv_base;
}
extern "C" int qa_pco_float_mult_config_f32(const float *sample, size_t n, float *base, float *inv) {
    std::vector<float> s(sample, sample + n);
    FMConfig<float> c;
    if (!fm_choose_config(s, &c)) return 0;
    *base = c.base;
    *inv = c.inv_base;
    return 1;
}
extern "C" int qa_pco_int_candidate_base(const uint32_t *sample, size_t n, uint32_t *base) {
    std::vector<uint32_t> s(sample, sample + n);
    double b;
    return choose_candidate_base<uint32_t>(s, base, &b) ? 1 : 0;
}
extern "C" uint32_t qa_pco_quant_k_f32(const float *sample, size_t n, double *bits) {
    std::vector<float> s(sample, sample + n);
    return fq_best_k<float>(s, bits);
}
extern "C" uint32_t qa_pco_quant_k_f64(const double *sample, size_t n, double *bits) {
    std::vector<double> s(sample, sample + n);
    return fq_best_k<double>(s, bits);
}

namespace {
template <class T> int describe_t(const uint64_t *lane, size_t n, unsigned level, uint64_t info[4]) {
    using L = typename NT<T>::L;
    if (n == 0) return -1;
    std::vector<T> nums(n);
    for (size_t i = 0; i < n; i++) nums[i] = from_lane<T>(lane[i]);
    Split<T> s;
    if (!choose_mode<T>(nums.data(), n, s)) return -1;
    const DeltaEnc de = choose_auto_delta<L>(s.primary, choose_ubl(level, n));
    info[0] = (uint64_t)s.mode.kind;
    info[1] = s.mode.kind == M_FLOAT_QUANT ? s.mode.k : s.mode.base;
    info[2] = (uint64_t)de.kind;
    info[3] = de.kind == D_CONSEC ? de.order : de.window_n_log;
    return 0;
}
}  // namespace

extern "C" int qa_pco_describe(int ntype, unsigned level, const uint64_t *lane, size_t n, uint64_t info[4]) {
    switch (ntype) {
    case QA_PCO_U32: return describe_t<uint32_t>(lane, n, level, info);
    case QA_PCO_U64: return describe_t<uint64_t>(lane, n, level, info);
    case QA_PCO_I32: return describe_t<int32_t>(lane, n, level, info);
    case QA_PCO_I64: return describe_t<int64_t>(lane, n, level, info);
    case QA_PCO_F32: return describe_t<float>(lane, n, level, info);
    case QA_PCO_F64: return describe_t<double>(lane, n, level, info);
    default: return -1;
    }
}

extern "C" size_t qa_pco_histogram_u32(const uint32_t *latents, size_t n, uint32_t n_bins_log, uint64_t *out,
                                       size_t cap) {
    std::vector<uint32_t> l(latents, latents + n);
    std::vector<HBin<uint32_t>> h = histogram(l.data(), n, n_bins_log);
    for (size_t i = 0; i < h.size() && 3 * i + 2 < cap; i++) {
        out[3 * i] = h[i].count;
        out[3 * i + 1] = h[i].lower;
        out[3 * i + 2] = h[i].upper;
    }
    return h.size();
}

extern "C" size_t qa_pco_lookbacks_u32(uint32_t window_n_log, uint32_t state_n_log, const uint32_t *latents,
                                       size_t n, uint32_t *out) {
    std::vector<uint32_t> lb = choose_lookbacks<uint32_t>(window_n_log, state_n_log, latents, n);
    if (out) memcpy(out, lb.data(), lb.size() * 4);
    return lb.size();
}

extern "C" float qa_pco_log2_approx(float x) { return log2_approx(x); }

// optimize_bins (bin_optimization.rs:178-197) over u32 histogram bins given as
// (count, lower, upper) triples; writes (weight, lower, upper, offset_bits,
// symbol) quintuples, returns the bin count
extern "C" size_t qa_pco_optimize_bins_u32(const uint64_t *hist, size_t nbins, uint32_t asl, uint64_t *out,
                                           size_t cap) {
    std::vector<HBin<uint32_t>> h(nbins);
    for (size_t i = 0; i < nbins; i++) h[i] = {(size_t)hist[3 * i], (uint32_t)hist[3 * i + 1], (uint32_t)hist[3 * i + 2]};
    std::vector<Info<uint32_t>> r = optimize_bins<uint32_t>(h, asl);
    for (size_t i = 0; i < r.size() && 5 * i + 4 < cap; i++) {
        out[5 * i] = r[i].weight;
        out[5 * i + 1] = r[i].lower;
        out[5 * i + 2] = r[i].upper;
        out[5 * i + 3] = r[i].offset_bits;
        out[5 * i + 4] = r[i].symbol;
    }
    return r.size();
}

// snap_to_int_reciprocal on f64 (float_mult.rs:216-231)
extern "C" void qa_pco_snap_f64(double base, double *ob, double *oi) {
    const FMConfig<double> c = snap_to_int_reciprocal<double>(base);
    *ob = c.base;
    *oi = c.inv_base;
}

// float_mult::compute_bid on an f32 sample: 1 with the base and the bits
// saved per number when the bid exists
extern "C" int qa_pco_float_mult_bid_f32(const float *sample, size_t n, float *base, double *bits) {
    std::vector<float> s(sample, sample + n);
    FMConfig<float> c;
    if (!fm_choose_config(s, &c)) return 0;
    double b;
    if (!fm_bits_saved(c, s, &b)) return 0;
    *base = c.base;
    *bits = b;
    return 1;
}
