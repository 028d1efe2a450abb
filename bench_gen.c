/* bench_gen.c — synthetic columns of the benchmark workloads.
 *
 * One named generator for every workload: xoshiro256** (Blackman and Vigna,
 * "Scrambled linear pseudorandom number generators", 2018), its state filled by
 * SplitMix64 from a 64-bit seed. Uniform doubles take the top 53 bits of a
 * draw; normals come from the Box–Muller transform on two such uniforms; a
 * bounded integer is Lemire's multiply-shift with rejection (2019). Each
 * column function fills a caller buffer from its own seed, so a column's bytes
 * depend only on (function, seed, n), and qg_checksum() names them.
 *
 * Built by __graft_entry__.build() as bench_gen.so; bench.py loads it with
 * ctypes. Host code only: it is not part of the library.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define QG_PI 3.14159265358979323846

typedef struct {
    uint64_t s[4];
} qg_rng;

static uint64_t splitmix64(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void rng_seed(qg_rng *r, uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t rng_next(qg_rng *r) {
    uint64_t *s = r->s;
    const uint64_t out = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return out;
}

/* [0, 1) and (0, 1] from the top 53 bits */
static inline double unif(qg_rng *r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }
static inline double unif_pos(qg_rng *r) { return (double)((rng_next(r) >> 11) + 1) * 0x1.0p-53; }

/* uniform in [0, m), m > 0 */
static inline uint64_t below(qg_rng *r, uint64_t m) {
    unsigned __int128 p = (unsigned __int128)rng_next(r) * m;
    uint64_t lo = (uint64_t)p;
    if (lo < m) {
        const uint64_t t = (0 - m) % m;
        while (lo < t) {
            p = (unsigned __int128)rng_next(r) * m;
            lo = (uint64_t)p;
        }
    }
    return (uint64_t)(p >> 64);
}

/* standard normals, two per Box–Muller draw */
typedef struct {
    qg_rng r;
    int have;
    double spare;
} qg_normal;

static double normal(qg_normal *g) {
    if (g->have) {
        g->have = 0;
        return g->spare;
    }
    const double u1 = unif_pos(&g->r), u2 = unif(&g->r);
    const double rad = sqrt(-2.0 * log(u1)), th = 2.0 * QG_PI * u2;
    g->spare = rad * sin(th);
    g->have = 1;
    return rad * cos(th);
}

/* the first `count` outputs of the generator seeded with `seed` */
void qg_raw(uint64_t seed, uint64_t count, uint64_t *out) {
    qg_rng r;
    rng_seed(&r, seed);
    for (uint64_t i = 0; i < count; i++) out[i] = rng_next(&r);
}

/* 2-decimal random walk: v_i = round(100 (1000 + S_i)) / 100, S_i the sum of
 * i + 1 standard normals (C round: half away from zero; IEEE division) */
void qg_walk_f64(uint64_t seed, uint64_t n, double *out) {
    qg_normal g;
    rng_seed(&g.r, seed);
    g.have = 0;
    double S = 0.0;
    for (uint64_t i = 0; i < n; i++) {
        S += normal(&g);
        out[i] = round(100.0 * (1000.0 + S)) / 100.0;
    }
}

/* monotone timestamps: t_0 = 1.7e18, t_i = t_(i-1) + U{1..1000} */
void qg_timestamps_i64(uint64_t seed, uint64_t n, int64_t *out) {
    qg_rng r;
    rng_seed(&r, seed);
    int64_t t = 1700000000000000000ll;
    for (uint64_t i = 0; i < n; i++) {
        if (i) t += 1 + (int64_t)below(&r, 1000);
        out[i] = t;
    }
}

/* k-bit values: the draw's top 32 bits shifted right by 32 - k */
void qg_kbits_u32(uint64_t seed, uint64_t n, uint32_t k, uint32_t *out) {
    qg_rng r;
    rng_seed(&r, seed);
    for (uint64_t i = 0; i < n; i++) out[i] = (uint32_t)(rng_next(&r) >> 32) >> (32 - k);
}

/* Bernoulli(p_valid) validity, bit i of the bitmap (LSB first) for value i;
 * values at null slots, when given, are set to 0 (4-byte elements) */
void qg_validity(uint64_t seed, uint64_t n, double p_valid, uint8_t *bitmap, uint32_t *values) {
    qg_rng r;
    rng_seed(&r, seed);
    memset(bitmap, 0, (n + 7) / 8);
    for (uint64_t i = 0; i < n; i++) {
        if (unif(&r) < p_valid) bitmap[i >> 3] |= (uint8_t)(1u << (i & 7));
        else if (values) values[i] = 0;
    }
}

/* decimals at scale 10^3: k / 1000.0 with k uniform in [0, 10^9) */
void qg_decimal_f64(uint64_t seed, uint64_t n, double *out) {
    qg_rng r;
    rng_seed(&r, seed);
    for (uint64_t i = 0; i < n; i++) out[i] = (double)below(&r, 1000000000ull) / 1000.0;
}

/* full-precision uniform [0, 1) doubles: (u64 >> 11) 2^-53 */
void qg_full_f64(uint64_t seed, uint64_t n, double *out) {
    qg_rng r;
    rng_seed(&r, seed);
    for (uint64_t i = 0; i < n; i++) out[i] = unif(&r);
}

/* a dictionary of 256 decimals (as qg_decimal_f64), written in runs of
 * Geometric(1/8) length (trials to the first success, mean 8), each run's
 * value uniform over the 256 */
void qg_dict_runs_f64(uint64_t seed, uint64_t n, double *out) {
    qg_rng r;
    rng_seed(&r, seed);
    double dict[256];
    for (int j = 0; j < 256; j++) dict[j] = (double)below(&r, 1000000000ull) / 1000.0;
    const double lq = log(1.0 - 1.0 / 8.0);
    uint64_t i = 0;
    while (i < n) {
        uint64_t len = 1 + (uint64_t)floor(log(unif_pos(&r)) / lq);
        const double v = dict[below(&r, 256)];
        for (; len && i < n; len--) out[i++] = v;
    }
}

/* sin(2 pi i / 1000) 1000 + N(0, 1e-3) */
void qg_sine_f64(uint64_t seed, uint64_t n, double *out) {
    qg_normal g;
    rng_seed(&g.r, seed);
    g.have = 0;
    for (uint64_t i = 0; i < n; i++)
        out[i] = sin(2.0 * QG_PI * (double)i / 1000.0) * 1000.0 + 1e-3 * normal(&g);
}

/* f32 uniform [-1, 1): ((u32 >> 8) 2^-23) - 1 */
void qg_unif_f32(uint64_t seed, uint64_t n, float *out) {
    qg_rng r;
    rng_seed(&r, seed);
    for (uint64_t i = 0; i < n; i++)
        out[i] = (float)((uint32_t)(rng_next(&r) >> 32) >> 8) * 0x1.0p-23f - 1.0f;
}

/* exp(U[ln lo, ln hi]) */
void qg_loguniform(uint64_t seed, uint64_t count, double lo, double hi, double *out) {
    qg_rng r;
    rng_seed(&r, seed);
    const double a = log(lo), b = log(hi);
    for (uint64_t i = 0; i < count; i++) out[i] = exp(a + (b - a) * unif(&r));
}

/* FNV-1a over the bytes, 64-bit (offset basis 0xcbf29ce484222325, prime
 * 0x100000001b3), taken eight bytes at a time as little-endian words */
uint64_t qg_checksum(const void *p, uint64_t bytes) {
    const uint8_t *b = (const uint8_t *)p;
    uint64_t h = 0xcbf29ce484222325ull;
    uint64_t i = 0;
    for (; i + 8 <= bytes; i += 8) {
        uint64_t w;
        memcpy(&w, b + i, 8);
        h = (h ^ w) * 0x100000001b3ull;
    }
    for (; i < bytes; i++) h = (h ^ b[i]) * 0x100000001b3ull;
    return h;
}
