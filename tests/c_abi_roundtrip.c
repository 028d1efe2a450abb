/* c_abi_roundtrip.c — the reference C ABI driven from C, built against
 * include/quoin.h and linked with libquoin_amd.so. It covers the checks of the
 * reference's capi/test/roundtrip.c:102-222 with inputs of its own:
 *   - f64 columns (ramp, cents, sine, specials, empty) through quoin_compress /
 *     quoin_decompress, and through a 4-thread QuoinCtx, which must give the
 *     same bytes;
 *   - junk input refused as QUOIN_ERR_CORRUPT by both decoders;
 *   - QUOIN_ERR_BUFFER_TOO_SMALL with the needed size, on both directions;
 *   - the typed API: i64, f32, i32 with nulls, the Arrow widenings (i8, i16,
 *     u8, u16, bool), an empty column, unaligned input and output buffers;
 *   - the decimal container at 64, 128 and 256 bits.
 * Usage: c_abi_roundtrip OUTDIR. Each compressed case leaves OUTDIR/<case>.in
 * (the bytes handed to the compressor) and OUTDIR/<case>.qs (the stream), and
 * prints "CASE <case> <api> <dtype> <n>" so a caller can compare the streams
 * with another encoder. Prints one line per check; exits with the number of
 * failed checks. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "quoin.h"

static const char *g_dir;
static int g_fail;

static void check(int ok, const char *what, const char *name) {
    printf("%-5s %-10s %s\n", ok ? "ok" : "FAIL", what, name);
    if (!ok) g_fail++;
}

static void save(const char *name, const char *ext, const void *p, size_t len) {
    char path[4096];
    snprintf(path, sizeof path, "%s/%s.%s", g_dir, name, ext);
    FILE *f = fopen(path, "wb");
    if (!f || (len && fwrite(p, 1, len, f) != len)) {
        check(0, "write", path);
    }
    if (f) fclose(f);
}

/* f64 round trip through the plain or the context entry points; returns the
 * stream (caller frees) and its length */
static uint8_t *f64_case(const char *name, const double *x, size_t n, const QuoinCtx *ctx, size_t *slen) {
    size_t cap = quoin_compress_bound(n), len = 0;
    uint8_t *s = malloc(cap ? cap : 1);
    int rc = ctx ? quoin_compress_ctx(ctx, x, n, s, cap, &len) : quoin_compress(x, n, s, cap, &len);
    check(rc == QUOIN_OK, "compress", name);
    check(quoin_decompressed_value_count(s, len) == (ptrdiff_t)n, "count", name);
    double *y = malloc((n + 1) * sizeof(double));
    size_t got = 0;
    rc = ctx ? quoin_decompress_ctx(ctx, s, len, y, n, &got) : quoin_decompress(s, len, y, n, &got);
    /* bit patterns: NaN payloads and -0.0 must survive */
    check(rc == QUOIN_OK && got == n && (n == 0 || memcmp(x, y, n * sizeof(double)) == 0), "roundtrip",
          name);
    free(y);
    *slen = len;
    return s;
}

static void typed_case(const char *name, int dtype, size_t n, const void *x, const uint8_t *valid) {
    size_t w = quoin_dtype_width(dtype), cap = quoin_typed_compress_bound(dtype, n);
    uint8_t *s = malloc(cap ? cap : 1);
    size_t len = quoin_typed_compress(dtype, n, x, valid, s, cap);
    check(len > 0, "compress", name);
    save(name, "in", x, n * w);
    save(name, "qs", s, len);
    if (valid) save(name, "valid", valid, (n + 7) / 8);
    printf("CASE %s typed %d %zu\n", name, dtype, n);
    uint8_t *y = malloc(n * w + 8), *vy = malloc((n + 7) / 8 + 1);
    int rc = quoin_typed_decompress(dtype, n, s, len, y, valid ? vy : NULL);
    int ok = rc == QUOIN_OK && (n == 0 || memcmp(x, y, n * w) == 0);
    if (ok && valid && n) {
        /* the decoder writes canonical bitmaps: bits past n are zero */
        size_t nb = (n + 7) / 8;
        uint8_t last = valid[nb - 1];
        if (n % 8) last &= (uint8_t)((1u << (n % 8)) - 1);
        ok = memcmp(valid, vy, nb - 1) == 0 && vy[nb - 1] == last;
    }
    check(ok, "roundtrip", name);
    free(s);
    free(y);
    free(vy);
}

static void decimal_case(const char *name, int dtype, size_t n, const void *x, int8_t scale, uint8_t prec) {
    size_t w = quoin_dtype_width(dtype), cap = quoin_typed_compress_bound(dtype, n);
    uint8_t *s = malloc(cap);
    size_t len = quoin_typed_compress_decimal(dtype, n, x, NULL, scale, prec, s, cap);
    check(len > 0, "compress", name);
    save(name, "in", x, n * w);
    save(name, "qs", s, len);
    printf("CASE %s decimal %d %zu %d %u\n", name, dtype, n, (int)scale, (unsigned)prec);
    uint8_t *y = malloc(n * w + 8);
    int rc = quoin_typed_decompress(dtype, n, s, len, y, NULL);
    check(rc == QUOIN_OK && memcmp(x, y, n * w) == 0, "roundtrip", name);
    free(s);
    free(y);
}

int main(int argc, char **argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s OUTDIR\n", argv[0]);
        return 2;
    }
    g_dir = argv[1];
    printf("library: %s\n", quoin_version());

    /* ---- f64 columns ---------------------------------------------------- */
    const size_t N = 80000;
    double *ramp = malloc(N * sizeof(double)), *cents = malloc(N * sizeof(double)),
           *sine = malloc(N * sizeof(double));
    for (size_t i = 0; i < N; i++) {
        ramp[i] = 0.25 * (double)i - 3.0;
        cents[i] = (double)((i * 37) % 2000) / 100.0;
        sine[i] = sin((double)i / 700.0) * 50.0;
    }
    double special[9] = {0.0, -0.0, NAN, -NAN, INFINITY, -INFINITY, 1.7e308, 4.9e-324, -2.5};
    struct { const char *name; const double *x; size_t n; } cols[] = {
        {"ramp", ramp, N}, {"cents", cents, N}, {"sine", sine, N}, {"special", special, 9},
        {"empty", ramp, 0}};
    QuoinCtx *ctx = quoin_ctx_create(4);
    check(ctx != NULL, "ctx", "create(4)");
    for (size_t c = 0; c < sizeof cols / sizeof cols[0]; c++) {
        size_t la = 0, lb = 0;
        uint8_t *a = f64_case(cols[c].name, cols[c].x, cols[c].n, NULL, &la);
        uint8_t *b = f64_case(cols[c].name, cols[c].x, cols[c].n, ctx, &lb);
        check(la == lb && memcmp(a, b, la) == 0, "ctx-same", cols[c].name);
        save(cols[c].name, "in", cols[c].x, cols[c].n * sizeof(double));
        save(cols[c].name, "qs", a, la);
        printf("CASE %s f64 %d %zu\n", cols[c].name, QUOIN_DTYPE_F64, cols[c].n);
        free(a);
        free(b);
    }
    quoin_ctx_free(ctx);

    /* ---- error paths ---------------------------------------------------- */
    uint8_t junk[40];
    for (size_t i = 0; i < sizeof junk; i++) junk[i] = (uint8_t)(0x5A ^ (i * 29));
    double tmp[4];
    size_t got = 12345;
    check(quoin_decompress(junk, sizeof junk, tmp, 4, &got) == QUOIN_ERR_CORRUPT && got == 12345,
          "corrupt", "f64 junk");
    int64_t ttmp[4];
    check(quoin_typed_decompress(QUOIN_DTYPE_I64, 4, junk, sizeof junk, ttmp, NULL) == QUOIN_ERR_CORRUPT,
          "corrupt", "typed junk");
    {
        size_t cap = quoin_compress_bound(N), len = 0, need = 0;
        uint8_t *s = malloc(cap);
        check(quoin_compress(cents, N, s, cap, &len) == QUOIN_OK, "compress", "cents again");
        check(quoin_compress(cents, N, s, 10, &need) == QUOIN_ERR_BUFFER_TOO_SMALL && need == len,
              "too-small", "compress reports the size");
        double *y = malloc(N * sizeof(double));
        got = 0;
        check(quoin_decompress(s, len, y, N - 1, &got) == QUOIN_ERR_BUFFER_TOO_SMALL && got == N,
              "too-small", "decompress reports the count");
        free(y);
        free(s);
    }

    /* ---- typed columns -------------------------------------------------- */
    const size_t M = 40000;
    int64_t *i64 = malloc(M * 8);
    float *f32 = malloc(M * 4);
    int16_t *i16 = malloc(M * 2);
    int8_t *i8 = malloc(M);
    uint16_t *u16 = malloc(M * 2);
    uint8_t *u8 = malloc(M), *bools = malloc(M);
    for (size_t i = 0; i < M; i++) {
        i64[i] = 5000000 + (int64_t)((i * 7919) % 3001);
        f32[i] = cosf((float)i * 2e-4f) * 80.0f;
        i16[i] = (int16_t)((int)(i % 400) - 200);
        i8[i] = (int8_t)((int)(i % 9) - 4);
        u16[i] = (uint16_t)(60000 + i % 5000);
        u8[i] = (uint8_t)(i % 11);
        bools[i] = (uint8_t)((i / 3) % 2);
    }
    typed_case("i64", QUOIN_DTYPE_I64, M, i64, NULL);
    typed_case("f32", QUOIN_DTYPE_F32, M, f32, NULL);
    typed_case("i16_widen", QUOIN_DTYPE_I16, M, i16, NULL);
    typed_case("i8_widen", QUOIN_DTYPE_I8, M, i8, NULL);
    typed_case("u16_widen", QUOIN_DTYPE_U16, M, u16, NULL);
    typed_case("u8_widen", QUOIN_DTYPE_U8, M, u8, NULL);
    typed_case("bool", QUOIN_DTYPE_BOOL, M, bools, NULL);
    typed_case("i64_empty", QUOIN_DTYPE_I64, 0, i64, NULL);
    {
        /* nulls: every fifth slot null, its value 0 (the decoder restores 0) */
        const size_t K = 1500;
        int32_t *v = calloc(K, 4);
        uint8_t *valid = calloc((K + 7) / 8, 1);
        for (size_t i = 0; i < K; i++)
            if (i % 5) {
                v[i] = -300 + (int32_t)(i % 77);
                valid[i / 8] |= (uint8_t)(1u << (i % 8));
            }
        typed_case("i32_nulls", QUOIN_DTYPE_I32, K, v, valid);
        free(v);
        free(valid);
    }
    {
        /* unaligned: values at an odd address, decoded to an odd address */
        uint8_t *in = malloc(M * 8 + 3), *out = malloc(M * 8 + 3);
        for (size_t i = 0; i < M; i++) {
            int64_t v = -40000 + (int64_t)(i % 211) * 3;
            memcpy(in + 3 + i * 8, &v, 8);
        }
        size_t cap = quoin_typed_compress_bound(QUOIN_DTYPE_I64, M);
        uint8_t *s = malloc(cap);
        size_t len = quoin_typed_compress(QUOIN_DTYPE_I64, M, in + 3, NULL, s, cap);
        int rc = len ? quoin_typed_decompress(QUOIN_DTYPE_I64, M, s, len, out + 3, NULL) : -99;
        check(len > 0 && rc == QUOIN_OK && memcmp(in + 3, out + 3, M * 8) == 0, "roundtrip",
              "i64_unaligned");
        save("i64_unaligned", "in", in + 3, M * 8);
        save("i64_unaligned", "qs", s, len);
        printf("CASE i64_unaligned typed %d %zu\n", QUOIN_DTYPE_I64, M);
        free(in);
        free(out);
        free(s);
    }

    /* ---- decimals ------------------------------------------------------- */
    int64_t *d64 = malloc(M * 8);
    uint8_t *d128 = calloc(M, 16), *d256 = calloc(M, 32);
    for (size_t i = 0; i < M; i++) {
        d64[i] = 250000 + (int64_t)(i % 4000);
        int64_t lo = -7000000000LL + (int64_t)(i % 9000);
        int64_t hi = lo < 0 ? -1 : 0; /* sign extension into the upper limb */
        memcpy(d128 + 16 * i, &lo, 8);
        memcpy(d128 + 16 * i + 8, &hi, 8);
        uint64_t w = 123456789ULL + i % 3000;
        memcpy(d256 + 32 * i, &w, 8);
    }
    decimal_case("dec64", QUOIN_DTYPE_DECIMAL64, M, d64, 3, 18);
    decimal_case("dec128", QUOIN_DTYPE_DECIMAL128, M, d128, 6, 38);
    decimal_case("dec256", QUOIN_DTYPE_DECIMAL256, M, d256, 10, 60);

    free(ramp); free(cents); free(sine); free(i64); free(f32); free(i16); free(i8); free(u16);
    free(u8); free(bools); free(d64); free(d128); free(d256);
    printf(g_fail ? "FAILED %d\n" : "all passed\n", g_fail);
    return g_fail;
}
