/*
 * quoin.h — C ABI of the MI355X-native quoin block path (libquoin_amd.so).
 *
 * The first block of declarations keeps the reference C ABI symbol for symbol
 * (reference: capi/include/quoin.h:28-142): same names, signatures, error codes
 * and buffer ownership. The second block adds the device-resident batched and
 * served entry points the GPU build exists for.
 *
 * Levels. The reference-named compress entry points run at Config::default,
 * Level::Max (capi/src/lib.rs:29-34, capi/src/typed.rs:71-73), as the reference
 * does; the *_cfg and quoin_gpu_compress entry points take any Level. Fastest
 * and Fast run wholly on the GPU. Balanced, High and Max plan blocks, size the
 * device-emitted candidates and write the predictor residual streams on the
 * GPU, and run the sequential coders of the competition (range coder, rANS,
 * LZ, pco, the entropy forms and the LZ cascade) on host cores; the device
 * then lays out and emits every frame. Every stream is bit-identical to the
 * reference encoder at the same Config. Decoding accepts every frame mode:
 * host-coded frames are first rewritten on host cores (and the device's
 * recurrence kernels) into device-decodable frames.
 *
 * Threading: the library owns one HIP device context per process (device 0
 * unless quoin_gpu_create picks another). Calls are serialized per handle.
 */
#ifndef QUOIN_H
#define QUOIN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QUOIN_OK                    0
#define QUOIN_ERR_NULL             (-1) /* a required pointer argument was NULL   */
#define QUOIN_ERR_BUFFER_TOO_SMALL (-2) /* dst too small; needed size in out_*    */
#define QUOIN_ERR_CORRUPT          (-3) /* malformed / truncated input stream     */
#define QUOIN_ERR_PANIC            (-4) /* internal error caught at the boundary  */
/* Additions of this build (no reference counterpart). */
#define QUOIN_ERR_UNSUPPORTED      (-5) /* level / frame mode not built (see top) */
#define QUOIN_ERR_DEVICE           (-6) /* HIP runtime error or no usable GPU     */

/* ---- reference C ABI (capi/include/quoin.h) ---------------------------- */

/* capi/include/quoin.h:37 — static NUL-terminated version string. */
const char *quoin_version(void);

/* capi/include/quoin.h:40 — 9*n + 1024. */
size_t quoin_compress_bound(size_t n_values);

/* capi/include/quoin.h:48-49 — compress doubles (at QUOIN_DEFAULT_LEVEL). */
int quoin_compress(const double *src, size_t n_values, uint8_t *dst, size_t dst_cap,
                   size_t *out_len);

/* capi/include/quoin.h:57-58 */
int quoin_decompress(const uint8_t *src, size_t src_len, double *dst, size_t dst_cap_values,
                     size_t *out_values);

/* capi/include/quoin.h:65 */
ptrdiff_t quoin_decompressed_value_count(const uint8_t *src, size_t src_len);

/* capi/include/quoin.h:68-76 — a context; `threads` sizes the host worker
 * pool used for host-side frame scans (0 = one per core). */
typedef struct QuoinCtx QuoinCtx;
QuoinCtx *quoin_ctx_create(size_t threads);
void quoin_ctx_free(QuoinCtx *ctx);

/* capi/include/quoin.h:80-83 */
int quoin_compress_ctx(const QuoinCtx *ctx, const double *src, size_t n_values, uint8_t *dst,
                       size_t dst_cap, size_t *out_len);
int quoin_decompress_ctx(const QuoinCtx *ctx, const uint8_t *src, size_t src_len, double *dst,
                         size_t dst_cap_values, size_t *out_values);

/* capi/include/quoin.h:97-115 — Arrow-native dtype tags. */
typedef enum {
    QUOIN_DTYPE_BOOL = 0,
    QUOIN_DTYPE_I8 = 1,
    QUOIN_DTYPE_I16 = 2,
    QUOIN_DTYPE_I32 = 3,
    QUOIN_DTYPE_I64 = 4,
    QUOIN_DTYPE_U8 = 5,
    QUOIN_DTYPE_U16 = 6,
    QUOIN_DTYPE_U32 = 7,
    QUOIN_DTYPE_U64 = 8,
    QUOIN_DTYPE_F32 = 9,
    QUOIN_DTYPE_F64 = 10,
    QUOIN_DTYPE_DECIMAL32 = 11,
    QUOIN_DTYPE_DECIMAL64 = 12,
    QUOIN_DTYPE_DECIMAL128 = 13,
    QUOIN_DTYPE_DECIMAL256 = 14
} QuoinDType;

/* capi/include/quoin.h:119 */
size_t quoin_dtype_width(int32_t dtype);
/* capi/include/quoin.h:122 */
size_t quoin_typed_compress_bound(int32_t dtype, size_t n);
/* capi/include/quoin.h:128-129 — returns bytes written, 0 on any failure. */
size_t quoin_typed_compress(int32_t dtype, size_t n, const void *values, const uint8_t *validity,
                            uint8_t *dst, size_t dst_cap);
/* capi/include/quoin.h:136-138 (capi/src/typed.rs:176-222, src/decimal.rs):
 * the decimal container, byte-identical to the reference at the same level.
 * DECIMAL32 / 64 widen to the 128-bit container. vmin and the 64-bit limbs of
 * v - vmin are computed on the device, each limb is compressed by the engine
 * as a U64 column, and the host lays out the container. Returns bytes written,
 * 0 on any failure. quoin_typed_decompress decodes the container. */
size_t quoin_typed_compress_decimal(int32_t dtype, size_t n, const void *values,
                                    const uint8_t *validity, int8_t scale, uint8_t precision,
                                    uint8_t *dst, size_t dst_cap);
/* capi/include/quoin.h:144-145 */
int quoin_typed_decompress(int32_t dtype, size_t n, const uint8_t *src, size_t src_len,
                           void *values_out, uint8_t *validity_out);

/* ---- additions: explicit configuration (src/lib.rs:97-245) -------------- */

/* Level (src/lib.rs:97-118). */
#define QUOIN_LEVEL_FASTEST  0
#define QUOIN_LEVEL_FAST     1
#define QUOIN_LEVEL_BALANCED 2
#define QUOIN_LEVEL_HIGH     3
#define QUOIN_LEVEL_MAX      4
#define QUOIN_DEFAULT_LEVEL  QUOIN_LEVEL_MAX

/* Selection (src/lib.rs:150-161). */
#define QUOIN_SELECTION_FULL   0
#define QUOIN_SELECTION_SAMPLE 1

/* Config (src/lib.rs:194-245): level, fixed block size in values (0 =
 * adaptive; clamped to [1, 131072]), mode selection, and the FCM / DFCM table
 * size predictor_log2 (0 = 16; clamped to [10, 16]). */
typedef struct {
    int32_t level;
    uint32_t block_size;
    int32_t selection;
    uint32_t predictor_log2;
} QuoinConfig;

/* As quoin_typed_compress with an explicit config; NULL cfg = defaults. */
size_t quoin_typed_compress_cfg(const QuoinConfig *cfg, int32_t dtype, size_t n,
                                const void *values, const uint8_t *validity, uint8_t *dst,
                                size_t dst_cap);
/* As quoin_compress with an explicit config. */
int quoin_compress_cfg(const QuoinConfig *cfg, const double *src, size_t n_values, uint8_t *dst,
                       size_t dst_cap, size_t *out_len);

/* Error code behind the last 0 returned by quoin_typed_compress* on this
 * thread (QUOIN_OK after a success). */
int quoin_last_error(void);

/* Per-mode win counters since the last reset (src/lib.rs:540-547), indexed by
 * mode id. */
void quoin_mode_win_counts(uint64_t out[64]);
void quoin_reset_mode_win_counts(void);

/* ---- additions: device-resident batched path --------------------------- */

/* A GPU handle: device, streams, scratch arenas sized by an HBM budget. */
typedef struct QuoinGpu QuoinGpu;

/* device < 0 picks device 0. hbm_budget_bytes bounds the scratch the handle may
 * allocate (0 = 64 GiB). Returns NULL when no usable GPU is present. */
QuoinGpu *quoin_gpu_create(int device, uint64_t hbm_budget_bytes);
void quoin_gpu_destroy(QuoinGpu *g);

/* One column in device memory. dtype is a QuoinDType among
 * I32, I64, U32, U64, F32, F64. */
typedef struct {
    int32_t dtype;
    uint32_t reserved;
    uint64_t n;              /* logical value count (nulls included)        */
    const void *values;      /* device, n * width bytes                     */
    const uint8_t *validity; /* device Arrow bitmap (ceil(n/8) B) or NULL   */
    uint8_t *dst;            /* device output stream                        */
    uint64_t dst_cap;        /* bytes; quoin_typed_compress_bound() is safe */
} QuoinDeviceColumn;

/* Compress a batch of device columns with one launch sequence covering every
 * block of every column. out_lens (host, ncols entries) receives each stream's
 * length. `stream` is a hipStream_t (NULL = the handle's stream). Synchronous:
 * returns once out_lens is valid. */
int quoin_gpu_compress(QuoinGpu *g, const QuoinConfig *cfg, const QuoinDeviceColumn *cols,
                       size_t ncols, uint64_t *out_lens, void *stream);

/* One stream to decode. The frame index is built on the host from host_src
 * (the same bytes as src), the way the reference decoder's sequential frame
 * scan reads the 9-byte frame headers (src/decoder.rs:65-94). */
typedef struct {
    int32_t dtype;            /* expected QuoinDType                          */
    uint32_t reserved;
    uint64_t n;               /* expected logical value count                 */
    const uint8_t *src;       /* device copy of the stream                    */
    const uint8_t *host_src;  /* host copy of the stream (frame scan)          */
    uint64_t src_len;
    void *values_out;         /* device, n * width bytes                      */
    uint8_t *validity_out;    /* device, ceil(n/8) bytes, or NULL             */
} QuoinDeviceStream;

/* Decode a batch of device streams with one launch sequence. Returns QUOIN_OK
 * or the first error. */
int quoin_gpu_decompress(QuoinGpu *g, const QuoinDeviceStream *streams, size_t nstreams,
                         void *stream);

/* Launch form of the two calls above: 0 = batched (default: every block of every
 * column in one launch per kernel stage; a batch shape met on two calls in a
 * row is captured into a hipGraph on the second and replayed while it repeats,
 * at Level::Fastest and Fast and for streams of device-coded frames),
 * 1 = plain (one launch sequence per block, the per-block-launch baseline),
 * 2 = batched, always launched kernel by kernel. Streams and decodes do not
 * depend on the form. */
int quoin_gpu_set_launch_form(QuoinGpu *g, int form);

/* No reference counterpart: the autotuner of the tile kernels' launch shape
 * (wavefronts per workgroup of the FoR / Delta tile decoders and the FoR / Delta
 * / FloatMult tile emitter). With on = 1 (the default) a handle times the
 * shapes 1, 2, 4 and 8 on its first batch with at least 4096 tiles of a
 * stage, re-running that idempotent stage, and keeps the fastest; 0 uses the
 * fixed shape 4. Results do not depend on the shape. */
int quoin_gpu_set_autotune(QuoinGpu *g, int on);
/* The tuned shapes: out[0] k_dec_fl<4>, out[1] k_dec_fl<8>, out[2] k_emit_fl
 * (0 = not tuned yet). */
int quoin_gpu_tuned_shapes(const QuoinGpu *g, uint32_t out[3]);
/* No reference counterpart: launch form 0's graphs so far, out[0] compress
 * sequences captured, out[1] compress graph launches, out[2] / out[3] the
 * same for decompress. */
int quoin_gpu_graph_counts(const QuoinGpu *g, uint64_t out[4]);

/* Device time (ms, HIP events on the handle's stream) of the last compress /
 * decompress call, and of its dominant kernel with its name. */
int quoin_gpu_last_timing(const QuoinGpu *g, double *total_ms, double *kernel_ms,
                          const char **kernel_name);

/* ---- additions: serving loop (continuous batching) --------------------- */

/* No reference counterpart: the reference has no request loop. Jobs name
 * device buffers (as quoin_gpu_compress / quoin_gpu_decompress take them) that
 * stay valid until quoin_server_wait returns for the job. */
#define QUOIN_JOB_COMPRESS   0
#define QUOIN_JOB_DECOMPRESS 1
typedef struct {
    int32_t kind;              /* QUOIN_JOB_*                              */
    uint32_t reserved;
    QuoinConfig cfg;           /* compress: level / block size             */
    QuoinDeviceColumn column;  /* compress: input column, output stream    */
    QuoinDeviceStream stream;  /* decompress: stream, typed output         */
} QuoinJob;

typedef struct QuoinServer QuoinServer;

/* A request loop on one device. `slots` worker slots (each with its own HIP
 * stream and scratch) take queued jobs as they free up: a slot admits the
 * oldest waiting job and every other waiting job of the same kind (and, for
 * compress, the same config) while the bytes of all admitted, unfinished jobs
 * stay within hbm_budget_bytes (0 = 64 GiB), up to max_batch jobs (0 = 256),
 * and runs them as one batched launch sequence. Jobs that arrive while others
 * are on the device join the next batch; batches of different slots overlap
 * on the device. A job's result does not depend on the jobs it is batched
 * with. Returns NULL when no usable GPU is present. */
QuoinServer *quoin_server_create(int device, uint64_t hbm_budget_bytes, uint32_t slots,
                                 uint32_t max_batch);
/* Queue a job; *ticket identifies it. QUOIN_ERR_BUFFER_TOO_SMALL when the job
 * alone exceeds the admission budget. Thread-safe. */
int quoin_server_submit(QuoinServer *s, const QuoinJob *job, uint64_t *ticket);
/* Block until the job is done: its return code (as the quoin_gpu_* call
 * would give for it alone), the stream length for compress, and the latency
 * from submit to completion in milliseconds. Each ticket is waited once. */
int quoin_server_wait(QuoinServer *s, uint64_t ticket, uint64_t *out_len, double *latency_ms);
/* Jobs completed and batched launches run since creation. */
void quoin_server_stats(const QuoinServer *s, uint64_t *jobs, uint64_t *batches);
/* Finishes every queued job, then frees the server. */
void quoin_server_destroy(QuoinServer *s);

/* ---- additions: rANS entropy coder on device buffers -------------------- */

/* One byte stream for quoin_gpu_rans_compress. */
typedef struct {
    const uint8_t *src;        /* device, n bytes                          */
    uint64_t n;
    uint8_t *dst;              /* device, dst_cap bytes                    */
    uint64_t dst_cap;          /* >= quoin_rans_bound(n)                   */
} QuoinRansInput;

/* One coded stream for quoin_gpu_rans_decompress. */
typedef struct {
    const uint8_t *src;        /* device, src_len bytes                    */
    uint64_t src_len;
    uint64_t max_len;          /* decode bound: a stream declaring more is corrupt */
    uint8_t *out;              /* device, max_len bytes                    */
} QuoinRansStream;

/* Capacity a coded stream of n bytes needs (0 for n = 0). */
size_t quoin_rans_bound(size_t n);

/* Replaces rans::compress_bytes (src/entropy/rans.rs:52-108) for a batch:
 * out_lens[i] = coded length of stream i, byte-identical to the reference
 * (0 for an empty input, where the reference returns None). Every stream is
 * coded by one launch sequence on the device. QUOIN_ERR_BUFFER_TOO_SMALL when
 * a dst_cap is below quoin_rans_bound(n); n must be below 2^32. */
int quoin_gpu_rans_compress(QuoinGpu *g, const QuoinRansInput *in, size_t n_streams,
                            uint64_t *out_lens, void *stream);

/* Replaces rans::decompress_bytes (src/entropy/rans.rs:110-175) for a batch,
 * and with tagged = 1 the TAG_RANS arm of decode_entropy_only
 * (src/entropy/mod.rs:153-161): src starts with the entropy tag byte; tags 0,
 * 1 and 3 (range coder, tANS, LZ cascade) are host coders and give
 * QUOIN_ERR_UNSUPPORTED, other tags QUOIN_ERR_CORRUPT. out_lens[i] = decoded
 * bytes; status[i] (optional) = QUOIN_OK or the stream's error, where the
 * reference returns Err (the bytes of such a stream's out are unspecified).
 * Returns QUOIN_OK or the first stream's error. */
int quoin_gpu_rans_decompress(QuoinGpu *g, const QuoinRansStream *streams, size_t n_streams,
                              int tagged, uint64_t *out_lens, int32_t *status, void *stream);

/* ---- additions: the residual entropy layer (src/entropy/mod.rs) ---------- */

/* Capacity the coded form of n residual bytes needs: 1 + quoin_rans_bound(n),
 * and 7 for n = 0 (the tag and an empty range-coded stream). */
size_t quoin_residual_bound(size_t n);

/* Replaces entropy::code_residuals (src/entropy/mod.rs:121-149, with
 * entropy_pick at mod.rs:45-78) for a batch of device byte streams: dst gets
 * [tag] ++ coded, byte-identical to the reference at the same lambda and
 * allow_lz. rANS is coded on the GPU; the range coder (src/entropy/rc.rs)
 * competes on host cores when lambda < 2 (and codes the empty stream), and
 * with allow_lz the LZ cascade (src/codecs/lz.rs) runs on host cores for
 * streams over 64 bytes. lambda is Level::lambda (src/lib.rs:124-139: Max and
 * High 0, Balanced 2, Fast 4, Fastest 16); allow_lz is Level::Max's
 * (lib.rs:145-147). out_lens[i] = coded bytes. QUOIN_ERR_BUFFER_TOO_SMALL when a
 * dst_cap is below quoin_residual_bound(n) (out_lens[i] = the bound). */
int quoin_gpu_code_residuals(QuoinGpu *g, uint64_t lambda, int allow_lz, const QuoinRansInput *in,
                             size_t n_streams, uint64_t *out_lens, void *stream);

/* Replaces entropy::decode_residuals (src/entropy/mod.rs:151-178) for a batch:
 * TAG_RANS streams decode on the GPU; TAG_RC, TAG_LZ and the legacy TAG_TANS
 * (src/entropy/tans.rs, still decoded, no longer written) on host cores.
 * out_lens[i] = decoded bytes; status[i]
 * (optional) = QUOIN_OK or the stream's error (QUOIN_ERR_CORRUPT where the
 * reference returns Err). Returns QUOIN_OK or the first stream's error. */
int quoin_gpu_decode_residuals(QuoinGpu *g, const QuoinRansStream *streams, size_t n_streams,
                               uint64_t *out_lens, int32_t *status, void *stream);

/* ------------------------------------------------------------------------ */
/* Linear residual coders on the device (src/codecs/linear.rs)               */
/* ------------------------------------------------------------------------ */
/* kind = the frame mode the residuals belong to (src/mode.rs:27-34). */
#define QUOIN_LINEAR_ORDERED_DELTA 7  /* varint(zigzag(v - pred)), integer pred */
#define QUOIN_LINEAR_DELTA2        20 /* [order] ++ varint(bits(v) ^ bits(pred)) */
#define QUOIN_LINEAR_DELTA_DP      45 /* [order] ++ varint(bits(r) ^ bits(r_prev)) */

/* One block of lane words for quoin_gpu_linear_encode. */
typedef struct {
    const uint64_t *vals;      /* device, n u64 lane words (f64 bit patterns) */
    uint64_t n;
    uint8_t *dst;              /* device, dst_cap bytes                    */
    uint64_t dst_cap;          /* >= quoin_linear_bound(n)                 */
} QuoinLinearInput;

/* One residual stream for quoin_gpu_linear_decode. */
typedef struct {
    const uint8_t *src;        /* device, src_len bytes                    */
    uint64_t src_len;
    uint64_t n;                /* values the block holds                   */
    uint64_t *out;             /* device, n u64 lane words                 */
} QuoinLinearStream;

/* Largest residual stream of n values: 10 n + 1. */
uint64_t quoin_linear_bound(uint64_t n);

/* Replaces, for a batch of blocks coded by one launch sequence:
 *   kind 7:  linear::idelta2_encode (src/codecs/linear.rs:31-45);
 *   kind 20: linear::encode (linear.rs:149-159) with order from
 *            linear::select_order (linear.rs:125-143) when order = 0;
 *   kind 45: linear::dp_encode (linear.rs:191-212), same order rule.
 * order: 0 = select_order per block, 1..4 = that order for every block
 * (ignored by kind 7). out_lens[i] = stream bytes; orders[i] (optional) = the
 * order used; status[i] (optional) = 1 where dp_encode returns None (an empty
 * block, or a value whose residual does not add back exactly), 0 otherwise.
 * f64 arithmetic follows x86-64 SSE2 (NaN results keep the first NaN operand,
 * quieted; invalid operations give 0xFFF8000000000000). QUOIN_ERR_UNSUPPORTED
 * for another kind or order, QUOIN_ERR_BUFFER_TOO_SMALL when a dst_cap is
 * below quoin_linear_bound(n). */
int quoin_gpu_linear_encode(QuoinGpu *g, int32_t kind, int32_t order, const QuoinLinearInput *in,
                            size_t n_blocks, uint64_t *out_lens, int32_t *orders, int32_t *status,
                            void *stream);

/* Replaces linear::idelta2_decode (linear.rs:47-66), linear::decode
 * (linear.rs:161-180) and linear::dp_decode (linear.rs:214-236) for a batch.
 * status[i] (optional) = QUOIN_OK, or QUOIN_ERR_CORRUPT where the reference
 * returns Err (a bad order byte, too few varints, trailing bytes, a varint
 * over 10 bytes; the block's out is then unspecified). Returns QUOIN_OK or
 * the first block's error. */
int quoin_gpu_linear_decode(QuoinGpu *g, int32_t kind, const QuoinLinearStream *streams,
                            size_t n_blocks, int32_t *status, void *stream);

/* ---- additions: FCM / DFCM predictors (src/codecs/pred.rs) --------------- */
#define QUOIN_PRED_FCM  0   /* Mode::Pred residuals: v ^ table[ctx] (pred.rs:23-37)          */
#define QUOIN_PRED_DFCM 25  /* Mode::Pred2 residuals: v ^ (last + table[dctx]) (pred.rs:59-75) */

/* Replaces pred::encode (kind 0) and pred::dfcm_encode (kind 25) for a batch
 * of blocks, byte-identical: out_lens[i] = residual stream bytes (varints,
 * written to in[i].dst, dst_cap >= quoin_linear_bound(n)). The stream is
 * Mode::Pred's payload as it is; Mode::PredRc and Mode::Pred2 code it with
 * quoin_gpu_code_residuals. predictor_log2 1..16 (Config::predictor_log2
 * clamps to 10..16, src/lib.rs:236-238); fewer than 65535 blocks per call. */
int quoin_gpu_pred_encode(QuoinGpu *g, int32_t kind, uint32_t predictor_log2,
                          const QuoinLinearInput *in, size_t n_blocks, uint64_t *out_lens,
                          void *stream);

/* Replaces pred::decode / pred::dfcm_decode (pred.rs:39-57, 77-97) for a
 * batch: the recurrence runs on the device, one block per lane, each with its
 * 8 << predictor_log2 byte table in HBM (refused as
 * QUOIN_ERR_BUFFER_TOO_SMALL past the handle's budget). status[i] (optional) =
 * QUOIN_OK or QUOIN_ERR_CORRUPT where the reference returns Err (too few
 * varints, trailing bytes, a varint over 10 bytes). */
int quoin_gpu_pred_decode(QuoinGpu *g, int32_t kind, uint32_t predictor_log2,
                          const QuoinLinearStream *streams, size_t n_blocks, int32_t *status,
                          void *stream);

/* ---- addition with no counterpart in the reference: no device allocation - */
/* quoin_gpu_linear_encode for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): same arguments, results and errors, but the call makes no
 * hipMalloc / hipFree. Its scratch is one fixed 2 MiB device array per GPU,
 * held until the process ends, which limits a call to 1024 blocks and 65536
 * tiles of 2048 values (134217728 values); larger calls return
 * QUOIN_ERR_BUFFER_TOO_SMALL before any device work. Calls share their GPU's
 * array: calls that overlap on one GPU race. */
int quoin_gpu_linear_encode_noalloc(QuoinGpu *g, int32_t kind, int32_t order,
                                    const QuoinLinearInput *in, size_t n_blocks,
                                    uint64_t *out_lens, int32_t *orders, int32_t *status,
                                    void *stream);


#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* QUOIN_H */
