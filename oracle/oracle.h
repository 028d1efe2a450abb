/*
 * oracle.h — CPU restatement of quoin's block path, used ONLY as a test checker.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library. The product (quoin_amd/) never links,
 * loads or calls it.
 *
 * Scope: the whole Level::Fastest / Level::Fast competition (encoder.rs:275-469
 * with entropy disabled), every decoder those levels can emit, the validity
 * container (validity.rs), the frame format (format.rs), and the 4-way rANS
 * coder (entropy/rans.rs), the range coder (entropy/rc.rs), the LZ cascade
 * (codecs/lz.rs) and the entropy tag layer (entropy/mod.rs). Whole streams at
 * Levels Balanced/High/Max return ORC_UNSUPPORTED (their competition needs pco).
 */
#ifndef QUOIN_ORACLE_H
#define QUOIN_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes (negative) — mirror src/error.rs variants coarsely. */
#define ORC_OK 0
#define ORC_TRUNCATED (-10)
#define ORC_BAD_MAGIC (-11)
#define ORC_BAD_VERSION (-12)
#define ORC_UNKNOWN_MODE (-13)
#define ORC_CORRUPT (-14)
#define ORC_LENGTH_MISMATCH (-15)
#define ORC_BAD_DTYPE (-16)
#define ORC_UNSUPPORTED (-17)
#define ORC_NOMEM (-18)

/* DType wire ids (dtype.rs:70-80). */
enum { ORC_F64 = 0, ORC_I64 = 1, ORC_U64 = 2, ORC_I32 = 3, ORC_U32 = 4, ORC_F32 = 5 };
/* Levels (lib.rs:97-118). */
enum { ORC_FASTEST = 0, ORC_FAST = 1, ORC_BALANCED = 2, ORC_HIGH = 3, ORC_MAX = 4 };

/* Compress a typed column (values in the dtype's native width, native endian),
 * optional Arrow validity bitmap. block_size 0 = adaptive. On success *out is a
 * malloc'ed stream of *out_len bytes (free with orc_free). */
int orc_compress_column(int dtype, size_t n, const void *values, const uint8_t *validity,
                        int level, size_t block_size, uint8_t **out, size_t *out_len);

/* Decompress: values written in the dtype's native width into a malloc'ed
 * buffer (*values, n * width bytes); *validity is NULL when the stream has no
 * bitmap, else a malloc'ed ceil(n/8) bitmap. */
int orc_decompress_column(const uint8_t *src, size_t len, int *dtype, size_t *n,
                          void **values, uint8_t **validity);

void orc_free(void *p);

/* Per-mode win counters (diag.rs). */
void orc_mode_wins(uint64_t out[64]);
void orc_reset_mode_wins(void);

/* Kernel-level entry points for unit tests. */
void orc_pack32(const uint32_t *values1024, uint32_t width, uint32_t *out);
void orc_unpack32(const uint32_t *packed, uint32_t width, uint32_t *out1024);
void orc_pack64(const uint64_t *values1024, uint32_t width, uint64_t *out);
void orc_unpack64(const uint64_t *packed, uint32_t width, uint64_t *out1024);

/* Codec payload encoders over a u64 lane; return malloc'ed payload (NULL when
 * the mode bails; *len = 0 then). */
uint8_t *orc_encode_mode(int mode, const uint64_t *vals, size_t n, int dtype, size_t *len);
/* Decode a single payload of `mode` into n lane words (out must hold n). */
int orc_decode_mode(int mode, const uint8_t *payload, size_t len, size_t n, int dtype,
                    uint64_t *out);

/* rANS (entropy/rans.rs). compress returns NULL for empty input. */
uint8_t *orc_rans_compress(const uint8_t *data, size_t n, size_t *len);
int orc_rans_decompress(const uint8_t *payload, size_t len, size_t max_len, uint8_t **out,
                        size_t *out_len);

/* CRC32C step of hash.rs (8 little-endian bytes, no inversion) and the FCM
 * (kind 0) / DFCM (kind 25) predictors of codecs/pred.rs. */
uint32_t orc_crc32c_u64(uint32_t crc, uint64_t v);
uint8_t *orc_pred_encode(int kind, const uint64_t *vals, size_t n, int log2, size_t *len);
int orc_pred_decode(int kind, const uint8_t *payload, size_t len, size_t n, int log2,
                    uint64_t *out);
/* Range coder with the order-1 byte model (entropy/rc.rs): varint(len) ++
 * coded bytes. */
uint8_t *orc_rc_compress(const uint8_t *src, size_t n, size_t *len);
int orc_rc_decompress(const uint8_t *src, size_t len, size_t max_len, uint8_t **out,
                      size_t *out_len);
/* LZSS over bytes (codecs/lz.rs). decompress writes `expected` bytes. */
uint8_t *orc_lz_compress(const uint8_t *in, size_t n, size_t *len);
int orc_lz_decompress(const uint8_t *s, size_t len, size_t expected, uint8_t **out);
/* tANS (entropy/tans.rs): the legacy coder behind TAG_TANS. compress returns
 * NULL for empty input or a model that does not quantize. */
uint8_t *orc_tans_compress(const uint8_t *src, size_t n, size_t *len);
int orc_tans_decompress(const uint8_t *src, size_t len, size_t max_len, uint8_t **out,
                        size_t *out_len);
/* Entropy tag layer (entropy/mod.rs): code_residuals picks rANS / range coder
 * by lambda and, with allow_lz, tries the LZ cascade; decode_residuals takes
 * every tag (TAG_RC, TAG_TANS, TAG_RANS, TAG_LZ). */
uint8_t *orc_code_residuals(const uint8_t *r, size_t n, uint64_t lambda, int allow_lz,
                            size_t *len);
int orc_decode_residuals(const uint8_t *p, size_t len, size_t max_len, uint8_t **out,
                         size_t *out_len);

/* Linear residual coders (codecs/linear.rs). kind = the frame mode: 7
 * OrderedDelta, 20 Delta2, 45 DeltaDp. encode: order 1..4 (ignored by
 * OrderedDelta); returns NULL with *len = 0 when DeltaDp bails. decode: n values
 * into out. */
int orc_linear_select_order(const uint64_t *vals, size_t n);
uint8_t *orc_linear_encode(int kind, const uint64_t *vals, size_t n, int order, size_t *len);
int orc_linear_decode(int kind, const uint8_t *payload, size_t len, size_t n, uint64_t *out);

/* Decimal container (decimal.rs). wire 6 = Decimal128 (2 limbs), 7 =
 * Decimal256 (4 limbs); values are n significands of limbs * 8 little-endian
 * two's-complement bytes. decompress returns malloc'ed values (null slots 0)
 * and the bitmap (or NULL). */
int orc_compress_decimal(int wire, size_t n, const uint8_t *values, const uint8_t *validity,
                         int8_t scale, uint8_t precision, int level, uint8_t **out, size_t *out_len);
int orc_decompress_decimal(const uint8_t *src, size_t len, int *wire, size_t *n, int8_t *scale,
                           uint8_t *precision, uint8_t **values, uint8_t **validity);

/* Feature probe + planning (encoder.rs:154-178, 705-750). ranges gets
 * (start,end) pairs; returns the number of blocks. ranges must hold n/1+1 pairs
 * in the worst case of block_size 1; callers size it as n+1 pairs. */
size_t orc_plan_blocks(const uint64_t *vals, size_t n, size_t block_size, uint64_t *ranges);

#ifdef __cplusplus
}
#endif
#endif
