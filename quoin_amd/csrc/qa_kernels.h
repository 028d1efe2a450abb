// qa_kernels.h — launch-side declarations shared by the kernel files and the
// host engine.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <tuple>
#include <utility>
#include <vector>

#include "qa_common.h"

// One step of a recorded launch sequence: a kernel launch with its argument
// values, or (ev set) a timing event recorded after the steps before it.
struct QaRecStep {
    const void *func;
    dim3 grid, block;
    uint32_t shm;
    std::shared_ptr<void> hold;  // the argument values
    std::vector<void *> params;  // a pointer to each of them
    void *ev;
};
// The calling thread's recorder: while it is set, qa_launch and qa_mark append
// steps to it instead of enqueueing work (the steps become a hipGraph's nodes).
std::vector<QaRecStep> *&qa_recorder();

// Every kernel launch of the engine: enqueued on s, or recorded.
template <typename... P, typename... A>
inline void qa_launch(void (*k)(P...), dim3 grid, dim3 block, uint32_t shm, hipStream_t s,
                      A &&...args) {
    auto vals = std::make_shared<std::tuple<P...>>(std::forward<A>(args)...);
    std::vector<void *> ps =
        std::apply([](auto &...x) { return std::vector<void *>{(void *)&x...}; }, *vals);
    if (std::vector<QaRecStep> *r = qa_recorder()) {
        r->push_back(QaRecStep{(const void *)k, grid, block, shm, vals, std::move(ps), nullptr});
        return;
    }
    (void)hipLaunchKernel((const void *)k, grid, block, ps.data(), shm, s);
}

size_t qa_scan_tmp_words(uint64_t n);
void qa_exclusive_scan(const uint64_t *in, uint64_t n, uint64_t *out, uint64_t *tmp,
                       hipStream_t s);
void qa_exclusive_scan_dn(const uint64_t *in, uint64_t nmax, const uint64_t *ndev, uint64_t *out,
                          uint64_t *tmp, hipStream_t s);
void qa_exclusive_scan2(const uint64_t *in0, uint64_t *out0, const uint64_t *in1, uint64_t *out1,
                        uint64_t n, uint64_t *tmp, hipStream_t s);

// Everything a wavefront needs to write one tile's payload (RAW, FoR or
// DeltaBitpack), resolved by k_tile_desc so the emitting wavefront starts
// with one load instead of a chain of lookups.
struct TileEmit {
    const void *lp;   // lane storage of the column
    uint8_t *dst;     // RAW: the tile's first byte; FoR/Delta: its sub-block header
    uint64_t k0;      // lane index of the tile's first value
    uint64_t mn;      // FoR minimum / Delta zig-zag minimum
    uint32_t cnt;     // values in the tile
    uint8_t w;        // FastLanes width
    uint8_t mode;     // QA_M_RAW / FOR / DELTA / FLOAT_MULT, 0 = written elsewhere
    uint8_t lk;       // lane storage kind
    uint8_t first;    // the first tile of its block (Delta: own predecessor)
    uint32_t dtype;
    uint32_t bmode;   // mode of the block owning this tile slot (0: none, or failed column)
    uint32_t fast;    // FoR / Delta / FloatMult with w <= 32: written by k_emit_fl
    uint32_t scale;   // FloatMult scale index
};

// Launch shapes picked by the autotuner: wavefronts per workgroup of the tile
// kernels, per handle, 0 until tuned. A stage is tuned on the first batch with
// at least QA_TUNE_MIN_TILES tiles of it, by timing each candidate (1, 2, 4, 8)
// on that batch (the tile stages are idempotent) and keeping the fastest.
#define QA_TUNE_DEC4 0  // k_dec_fl<4>
#define QA_TUNE_DEC8 1  // k_dec_fl<8>
#define QA_TUNE_EMIT 2  // k_emit_fl
#define QA_TUNE_N 3
#define QA_TUNE_MIN_TILES 4096
#include <functional>
uint32_t qa_tune_pick(const std::function<void(uint32_t)> &launch, hipStream_t s);
// a timing event (null: none) recorded on s, as an event node when s is capturing
void qa_mark(void *ev, hipStream_t s);

// Device pointers of one encode batch (all arrays sized by the host from upper
// bounds; counts that depend on the data live in DColState).
struct EncArgs {
    const DCol *cols;
    DColState *cst;
    uint32_t ncols;
    uint32_t nregions;     // total region slots
    uint64_t ntiles;       // total tile slots
    uint64_t nvals;        // total logical values
    uint64_t nvwords;      // total validity words
    uint64_t *lane;        // compacted copies of columns with nulls
    uint64_t *vpre;        // nvwords+1: popcount, then exclusive prefix
    uint64_t *tpre;        // nvwords+1: transition counts, then prefix
    uint64_t *tpos;        // transition positions (<= total bits)
    uint64_t *rsz;         // run varint sizes (+ prefix), ntrans_max + ncols + 1
    uint64_t *colbase;     // per column: transition base (ncols+1), then [ncols+1] = run count
    uint64_t *scan_tmp;
    uint32_t *rfeat;       // per region slot features
    uint32_t *owner;       // per region slot: global slot of the owning block
    DBlock *blocks;        // per region slot
    DTile *tiles;          // per tile slot
    uint64_t ntrans_max;
    unsigned long long *wins;  // 64 per-mode win counters
    // Level::Fast codecs
    uint32_t fast;         // some column runs at Level::Fast
    // Levels Balanced / High / Max and Selection::Sample: the device sizes
    // its own candidates (1: for the blocks the Full competition's gates let
    // through; 2: for every block), leaves Dict to the host, counts no wins,
    // and with `force` set lays out the modes the host chose
    uint32_t hybrid;
    DTileF *tilesf;        // per tile slot
    DBlockX *blocksx;      // per region slot
    uint64_t *dsorted;     // per lane value slot: sorted dictionaries
    uint32_t *rstart;      // per lane value slot (+1 per block): RLE run starts
    uint64_t *hkeys;       // 4 per lane value slot: global hash fallback keys
    uint32_t *hcnt;        // 4 per lane value slot: their counts
    TileEmit *temit;       // per tile slot
    uint32_t *tlist;       // [0]: count, [16 + i]: the tiles k_emit_tiles writes
    uint32_t *dpart;       // per region slot: Dict part set size (Level::Fast)
    uint64_t *dparts;      // per region slot: DH_LIMIT keys of its part set
    DRegionG *rgrown;      // per region slot
    uint64_t *gsets;       // 256 words per region slot
    const struct HForce *force;  // per region slot, or null
    // launch range of the per-block stages: region slots [r0, r1), tile slots
    // [t0, t1) (the whole batch, or one block in the plain form)
    uint32_t r0, r1;
    uint64_t t0, t1;
    // host side: HIP events recorded around the tile emit kernel, or null
    void *kev0, *kev1;
    // host side: the handle's tuned launch shapes (QA_TUNE_*), or null
    uint32_t *tune;
};

// The winner of a region slot's block at Levels Balanced / High / Max: a mode
// the device emits itself, or (host = 1) a payload of `len` bytes written by
// k_host_frames.
struct HForce {
    uint32_t mode, host;
    uint64_t len;
};
// One host-written frame: the mode byte and the payload bytes at src_off of
// the uploaded payload buffer go to the frame of region slot `slot`.
struct HScatter {
    uint32_t slot, mode;
    uint64_t src_off, len;
};
// the compacted lane words of every column, at lane_off + i (i < nv)
void qa_lane_words_launch(const EncArgs &a, uint64_t total, uint64_t *out, hipStream_t s);
// after the emission: the host-written frames
void qa_host_frames_launch(const EncArgs &a, const HScatter *sc, uint32_t ns, const uint8_t *src,
                           hipStream_t s);

void qa_fast_sizes_launch(const EncArgs &a, hipStream_t s);
void qa_fast_force_launch(const EncArgs &a, hipStream_t s);
void qa_fast_final_launch(const EncArgs &a, hipStream_t s);
void qa_fast_emit_launch(const EncArgs &a, hipStream_t s);

void qa_encode_launch(const EncArgs &a, hipStream_t s, int level_max, hipStream_t side = nullptr,
                      hipEvent_t fork = nullptr, hipEvent_t join = nullptr);
void qa_encode_plan_launch(const EncArgs &a, hipStream_t s);
void qa_encode_blocks_launch(const EncArgs &a, hipStream_t s);
void qa_encode_layout_launch(const EncArgs &a, hipStream_t s);
void qa_encode_emit_launch(const EncArgs &a, hipStream_t s);

struct DFrame {
    uint64_t payload_off;  // byte offset of the payload in the stream
    uint64_t out_off;      // first compact value of this frame (within column)
    uint64_t tile_base;    // first global decode tile slot
    uint32_t n, plen;
    uint32_t mode, col;
    uint32_t ntiles, gpu_ok;  // gpu_ok: the mode has a GPU decoder
    uint64_t scr;             // base of this frame's n slots in the decode scratch
};

// device-parsed frame state of the Level::Fast modes
struct DFrameX {
    uint64_t dict[8];      // ALP-RD dictionary
    uint64_t exc_off;      // ALP-RD: stream offset of the exceptions list
    uint64_t nexc;         // ALP-RD exceptions
    uint64_t aux_off;      // Dict: values blob offset; RLE: run values offset
    uint64_t card;         // Dict cardinality / RLE runs
    uint32_t left, dl;     // ALP-RD cut and dictionary size
    uint32_t sorted;       // ALP-RD exception positions strictly increasing
    uint32_t vtag;         // Dict values tag
    uint32_t scale, pad;   // FloatMult scale index
};

// per 1024-value decode tile
struct DTileD {
    uint64_t off;    // stream offset of the tile's sub-block header (FoR-family modes, ALP)
    uint64_t tail;   // DeltaBitpack: tile sum, then carry; ALP-RD rights / Dict values /
                     // RLE lengths: sub-block header of the second for_bitpack walk
    uint64_t run0;   // RLE: the run holding the tile's first position (k_dec_frame_wg)
    uint32_t frame;  // owning frame
    uint32_t w;      // FastLanes width of the sub-block at off
};

// Everything a wavefront needs to decode one tile, resolved by k_dec_desc
// (frame, stream, output, logical range) so the decoding wavefront starts with
// one load instead of a chain of lookups.
struct TileDec {
    const uint8_t *src;       // the stream
    void *out;                // typed output
    const uint64_t *bitmap;   // decoded validity words (streams with nulls)
    uint64_t off, tail;       // DTileD.off / .tail (FloatMult: the scale, as f64 bits)
    uint64_t payload;         // stream offset of the frame's payload
    uint64_t base;            // compact index of the tile's first value
    uint64_t p0;              // index of the tile's first value within its frame
    uint64_t L0, L1;          // streams with nulls: logical range of the tile
    uint32_t frame, cnt, w, mode;  // mode 0: nothing to do
    uint32_t dtype, oform, nulls;
    uint32_t fl;  // FoR / Delta, w <= 32, into 4 or 8-byte integers: k_dec_fl<fl>
};

struct DDecCol {
    const uint8_t *src;     // device stream
    void *out;              // typed output
    uint8_t *validity_out;  // device bitmap out or null
    uint64_t *bitmap;       // library-owned decoded bitmap (words), or null
    uint64_t src_len;
    uint64_t n;             // logical values
    uint64_t vword_base;    // rank prefix slots
    uint64_t val_base;
    uint64_t vblob_off;     // stream offset of the validity blob (its tag byte)
    uint64_t vblob_len;     // blob bytes including the tag
    uint64_t vbyte_base;    // global slot of the blob's first byte after the tag
    uint64_t avbyte_base;   // all-valid validity_out bytes before this column
    uint32_t dtype, nvwords;
    uint32_t out_kind;      // QA_IN_* narrow output element type
    uint32_t fbegin, fend;  // this column's frames
    uint32_t vchunk_base;   // first validity chunk of this column (k_vb_scan)
    uint32_t tl_aligned;    // every frame starts at a multiple of 1024 values: the
                            // column's tile m starts at valid value 1024 m
    uint64_t tile0, tile1;  // the column's decode tiles [tile0, tile1)
};

// Validity blob decode: a column's blob is cut into chunks (QA_VB_CHUNK run
// bytes, or QA_VB_WORDS bitmap words of a raw blob), one workgroup each,
// chained by a decoupled look-back. A chunk's aggregate: runs (c), bits (s),
// valid bits when its first run has an even index (v0) or an odd one (v1);
// its inclusive prefix from the column's start: runs, bits, valid bits.
#define QA_VB_CHUNK 4096u
#define QA_VB_WORDS 512u
// Every word carries bit 63 once written (values saturate below it), so a
// reader takes a set only when each of its words is marked: no ordering
// between the words is needed, and none is asked of the memory system.
struct VbChunk {
    uint64_t a[4];          // aggregate: c, s, v0, v1
    uint64_t p[3];          // inclusive prefix: c, s, v
    uint64_t pad;
};

// device-computed per-column decode state
struct DDecState {
    uint64_t nv;            // valid values
    uint64_t nruns;         // RLE runs used
    uint32_t fused_end;     // frames [fbegin, fused_end) are decoded
    uint32_t err;
};

struct DecArgs {
    const DDecCol *cols;
    DDecState *dst;
    uint32_t ncols;
    const DFrame *frames;
    uint32_t nframes;
    uint64_t ntiles;
    uint64_t nvals;
    uint64_t nvwords;
    uint64_t nvbytes;      // total validity blob bytes (after tags)
    uint64_t navbytes;     // validity_out bytes of streams without a bitmap
    uint64_t *vpre;        // per bitmap word: the column's valid values before it (nvwords + 1)
    uint64_t *vbval;       // per run: the column's valid values before it (nvbytes + ncols + 1)
    uint64_t *tlpos;       // per tile of an aligned stream with nulls: the logical
                           // position of its first value (written with the bitmap)
    uint64_t *vruns;       // per run: its start (nvbytes + ncols + 1); a column's runs
                           // start at slot vbyte_base + column index
    DTileD *tiles;         // per decode tile
    TileDec *tdec;         // per decode tile
    uint32_t *err;         // error flags
    uint64_t *scan_tmp;
    DFrameX *fx;           // per frame
    uint64_t *scr, *scr2;  // per frame value slots (Dict values, RLE starts, XORZ, ALP-RD exceptions)
    uint32_t fast;         // some frame needs the Level::Fast decoders
    uint32_t nvchunks;     // validity chunks (k_vb_scan workgroups)
    VbChunk *vchunks;      // per chunk look-back state, zeroed per call
    uint32_t *vticket;     // chunk ticket counter, zeroed per call
    // launch range of the per-frame stages: frames [f0, f1), tiles [t0, t1)
    // (the whole batch, or one frame in the plain form)
    uint32_t f0, f1;
    uint64_t t0, t1;
    // host side: HIP events recorded around the tile decode kernel, or null
    void *kev0, *kev1;
    // host side: the handle's tuned launch shapes (QA_TUNE_*), or null
    uint32_t *tune;
    // host side: runs of consecutive tiles of one tile_kind, or null
    const struct TileRun *runs;
    uint32_t nruns, pad2;
    // device table of every tile kind's runs (kinds 4, 8, 0 at kr_off[0..2]),
    // or null; a tile decoder's launch takes one kind's slice as krun
    const struct KindRun *krun_all;
    uint32_t kr_off[3], kr_n[3];
    uint64_t kr_tot[3];
    const struct KindRun *krun;
    uint32_t nkrun, pad3;
    uint64_t ktotal;
    // the DeltaBitpack frames [df0, df1) and their tiles [dt0, dt1) (spans
    // that hold every Delta frame; empty when there is none)
    uint32_t df0, df1;
    uint64_t dt0, dt1;
};

// Which tile decoder takes a tile (from its frame's mode and the output
// form): 4 = k_dec_fl<4> (FoR into 4-byte integers), 8 = k_dec_fl<8> (FoR and
// Delta into 8-byte integers, Delta into 4-byte ones, FloatMult and ALP into
// 8-byte lanes), 0 = k_dec_tiles.
__host__ __device__ inline uint32_t tile_kind(uint32_t mode, uint32_t oform) {
    if (mode == QA_M_FOR && oform == QA_OUT_U32) return 4;
    if ((mode == QA_M_FOR || mode == QA_M_DELTA) && oform == QA_OUT_U64) return 8;
    if (mode == QA_M_DELTA && oform == QA_OUT_U32) return 8;
    if ((mode == QA_M_FLOAT_MULT || mode == QA_M_ALP) && oform == QA_OUT_U64) return 8;
    return 0;
}
struct TileRun {
    uint64_t t0, t1;
    uint32_t kind, pad;
};
// a run of one kind's tiles in its launch: launch index `pre` is tile t0
struct KindRun {
    uint64_t pre, t0;
};

void qa_decode_launch(const DecArgs &a, hipStream_t s);
void qa_decode_prep_launch(const DecArgs &a, hipStream_t s);
void qa_decode_index_launch(const DecArgs &a, hipStream_t s);
void qa_decode_values_launch(const DecArgs &a, hipStream_t s);
void qa_decode_final_launch(const DecArgs &a, hipStream_t s);
void qa_fast_dec_index_launch(const DecArgs &a, hipStream_t s);
void qa_dec_desc_launch(const DecArgs &a, hipStream_t s);
void qa_dec_desc_fill_launch(const DecArgs &a, hipStream_t s);
void qa_dec_tiles_launch(const DecArgs &a, hipStream_t s);

// rANS (src/entropy/rans.rs): per-stream descriptors of a batch
struct RansDecD {
    const uint8_t *src;   // payload, or [tag] + payload when tagged
    uint8_t *out;         // max_len bytes
    uint64_t src_len;
    uint64_t max_len;
    uint32_t tagged;      // 1: entropy tag byte first (src/entropy/mod.rs:153-161)
    uint32_t pad;
};
struct RansDecR {
    uint64_t n;           // decoded bytes
    int32_t status;       // 0 ok, 1 corrupt, 2 other coder (tagged)
    uint32_t pad;
};
struct RansEncD {
    const uint8_t *src;   // bytes to code (n >= 1)
    uint8_t *dst;         // header + states + bytes
    uint64_t n;
    uint64_t scr_end;     // end of the stream's backward scratch region
};
struct RansEncR {
    uint64_t hdr_len;
    uint64_t start;       // scratch offset of the states
    uint64_t out_len;
};
// one byte written at each of up to QA_PUT_MAX device addresses (kernel
// arguments only: no device descriptor array)
#define QA_PUT_MAX 120
struct BytePuts {
    uint8_t *p[QA_PUT_MAX];
    uint8_t v[QA_PUT_MAX];
    uint32_t n;
};
void qa_put_bytes_launch(const BytePuts &b, hipStream_t s);

// the handle's device, and the stream a call runs on (the caller's or the
// handle's own); the handle itself is defined in qa_host.cpp
struct QuoinGpu;
int qa_gpu_device(const QuoinGpu *g);
hipStream_t qa_gpu_stream(const QuoinGpu *g, void *stream);
// the lock that serializes a handle's calls
std::recursive_mutex &qa_gpu_mutex(QuoinGpu *g);

void qa_rans_decode_launch(const RansDecD *d, RansDecR *r, uint32_t ns, hipStream_t s);
// trash: 64 device bytes outside every stream's scratch region
void qa_rans_encode_launch(const RansEncD *d, RansEncR *r, uint64_t *tabs, uint8_t *scr, uint32_t ns,
                           uint8_t *trash, hipStream_t s);

// linear residual coders (src/codecs/linear.rs): one block of a batch
struct LinD {
    const uint64_t *vals;  // encode: the block's lane words
    uint64_t *out;         // decode: n lane words
    const uint8_t *src;    // decode: the residual stream
    uint8_t *dst;          // encode: the residual stream (10 n + 1 bytes)
    uint64_t n;            // values
    uint64_t len;          // decode: stream bytes
    uint64_t tile0;        // first tile (encode: 2048 values, decode: 4096 stream bytes)
    uint32_t order;        // encode: 0 = select_order, else 1..4
    uint32_t pad;
};
struct LinR {
    uint64_t len;          // encode: stream bytes
    int32_t status;        // encode: 1 = DeltaDp declined; decode: 1 = corrupt
    uint32_t order;
};
// decimal container (src/decimal.rs): significands -> limbs of v - vmin, and back
#define QA_DECM_PARTS 1024
struct DecmArgs {
    const uint8_t *values;  // encode: in_w-byte significands
    const uint8_t *bitmap;  // Arrow validity bitmap (bytes) or null
    uint64_t n;             // logical values
    uint32_t in_w, nl;      // bytes per input value; limbs (2: Decimal128, 4: Decimal256)
    uint64_t *part;         // encode: QA_DECM_PARTS x 5 words of partial minima
    uint64_t *vmin;         // nl words
    uint64_t *limbs;        // nl x n words at logical positions
    uint8_t *out;           // decode: out_w bytes per value
    uint32_t out_w, pad;
};
void qa_decm_limbs_launch(const DecmArgs &a, hipStream_t s);
void qa_decm_join_launch(const DecmArgs &a, hipStream_t s);

// FCM / DFCM predictors (kind 0 / 25), on the linear coders' descriptors
size_t qa_pred_sort_bytes(uint64_t total);
void qa_pred_encode_launch(uint32_t kind, const LinD *d, LinR *r, uint32_t nblk, uint64_t ntiles,
                           uint32_t log2, uint32_t *keys, uint32_t *idx, uint32_t *keys2,
                           uint32_t *idx2, uint64_t *res, void *sort_tmp, size_t sort_bytes,
                           uint64_t *tsz, uint64_t *tpre, uint64_t *tmp, hipStream_t s);
void qa_pred_decode_launch(uint32_t kind, const LinD *d, LinR *r, uint32_t nblk, uint64_t ntiles,
                           uint32_t log2, uint64_t *tables, uint64_t *cnt, uint64_t *pre,
                           uint64_t *tmp, hipStream_t s, void *kev0, void *kev1);
uint64_t qa_lin_enc_tiles(uint64_t n);
uint64_t qa_lin_dec_tiles(uint64_t bytes);
void qa_lin_encode_launch(uint32_t kind, const LinD *d, LinR *r, uint32_t nblk, uint64_t ntiles,
                          uint64_t *tsz, uint64_t *tpre, uint64_t *tmp, hipStream_t s);
void qa_lin_decode_launch(uint32_t kind, const LinD *d, LinR *r, uint32_t nblk, uint64_t ntiles,
                          uint64_t *cnt, uint64_t *pre, uint64_t *tmp, hipStream_t s, void *kev0,
                          void *kev1);
