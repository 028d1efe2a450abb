// qa_hybrid.h — the engine at Levels Balanced / High / Max: the frames whose
// payloads come from the sequential host coders (range coder, LZ, pco) or the
// entropy tag layer, next to the device tile path.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/quoin.h"
#include "qa_common.h"

namespace qahyb {

using Bytes = std::vector<uint8_t>;

// LEB128 (src/varint.rs)
void put_varint(Bytes &o, uint64_t v);
bool get_varint(const uint8_t *p, size_t len, size_t &pos, uint64_t &v);

// for_bitpack::encode / delta_bitpack::encode on host cores (the byte form the
// device tile decoders read back): FastLanes sub-blocks of 1024 values
void for_encode(const uint64_t *v, size_t n, bool sgn, Bytes &out);
void delta_encode(const uint64_t *v, size_t n, Bytes &out);

// Whether a frame needs the host stage before the device decoder takes it:
// a host-coded mode (predictors, linear coders, LZ, ByteTranspose, pco) or an
// entropy-coded payload form of FloatMult, ALP-RD or Dict.
bool frame_needs_host(int mode, const uint8_t *payload, size_t len);

// The decode side: every stream whose frames need the host stage is rewritten
// into a stream of device-decodable frames of the same values. Entropy-coded
// sub-streams decode through quoin_gpu_decode_residuals (rANS on the GPU, the
// range coder / LZ / tANS on host cores), predictor and linear recurrences
// through the device recurrence kernels, LZ / ByteTranspose / pco on host
// cores; FloatMult / ALP-RD / Dict keep their frame mode in the bit-packed
// form. out[i] empty = stream i is unchanged. src_dev[i] is the device copy of
// stream i. QUOIN_OK, QUOIN_ERR_CORRUPT or QUOIN_ERR_DEVICE.
int rewrite_streams(QuoinGpu *g, size_t ns, const uint8_t *const *src_host, const uint8_t *const *src_dev,
                    const uint64_t *lens, std::vector<Bytes> &out, hipStream_t s);

}  // namespace qahyb

namespace qahyb {

// One block of the competition at Levels Balanced / High / Max
// (src/encoder.rs:275-469). The device supplies the planning, the features and
// the sizes of the candidates it emits itself (RAW, CONST, STRIDE, XORZ,
// ForBitpack, DeltaBitpack, ALP, RLE), the FloatMult scale and the ALP-RD cut
// and dictionary, and (High / Max) the FCM / DFCM / Delta2 / DeltaDp /
// OrderedDelta residual streams. The host cores run the rest of the
// competition in the reference's order: the entropy forms of FloatMult,
// ALP-RD and Dict, the order-1 estimate gate, the range coder and rANS, LZ,
// ByteTranspose, pco and the LZ cascade of the top two.
struct EncBlock {
    const uint64_t *v = nullptr;  // the block's lane words (nulls compacted out)
    uint32_t n = 0, dt = 0;
    uint64_t sz[C_N];             // device candidate sizes, UINT64_MAX = none
    bool incompressible = false, distinct_low = false, repeats = false;
    uint32_t fm_scale = 7;        // first passing FloatMult scale, 7 = none
    uint32_t rd_left = 0, rd_dl = 0, rd_bail = 1;
    uint64_t rd_dict[8] = {};
    const uint8_t *res[5] = {};   // FCM, DFCM, Delta2, DeltaDp, OrderedDelta
    size_t rlen[5] = {};
    bool has_res[5] = {};
    // result
    int mode = 5;                 // RAW
    bool host = false;            // payload below; else the device emits mode
    Bytes payload;
    uint64_t len = 0;
};
void compete(EncBlock &e, int level);
// Selection::Sample (src/encoder.rs:483-641) at any level: RAW, CONST, STRIDE,
// LZ (gated) and pco on the whole block, and the one mode that wins on the
// block's sample `smp` (8 runs of 64 values, predictor tables of 2^10), again
// on the whole block.
void compete_sampled(EncBlock &e, const EncBlock &smp, int level);

}  // namespace qahyb
