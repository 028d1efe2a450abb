// qa_pco.h — the pco (pcodec 1.0.2 fork) block codec on host cores.
//
// quoin's Mode::Pco frames carry a standalone pco file of the block's values
// (src/codecs/pcodec.rs:33-76, vendor/quoin-pco/src/standalone/simple.rs:58-100).
// pco's encoder is a sequential search (mode and delta selection on samples,
// a quicksort histogram, a dynamic program over bins, ANS), so it runs on host
// cores; a block whose winner is pco is also decoded there.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

// pco number-type bytes (data_types/{unsigned,signed,float}.rs impl macros)
enum {
    QA_PCO_U32 = 1,
    QA_PCO_U64 = 2,
    QA_PCO_I32 = 3,
    QA_PCO_I64 = 4,
    QA_PCO_F32 = 5,
    QA_PCO_F64 = 6,
};

namespace qapco {

// Compress n quoin lane words as pco numbers of type `ntype` at compression
// level `level` (0..12), the way pcodec::encode maps a lane (pcodec.rs:33-49):
// f64 bits, f32 narrowed from the exact f64, i32/u32 from the low 32 bits.
// Returns false (no candidate) for n == 0, as the reference does.
bool encode_lane(const uint64_t *lane, size_t n, int ntype, unsigned level,
                 std::vector<uint8_t> &out);

// Decode a pco payload that must hold exactly n numbers of `ntype` back to
// quoin lane words (pcodec.rs:66-76). Returns 0, or -1 when the payload is
// corrupt, truncated, of another number type or of another count.
int decode_lane(const uint8_t *src, size_t len, size_t n, int ntype, uint64_t *lane);

}  // namespace qapco

// C entry points of the same codec (test and tooling hooks; the engine calls
// the C++ functions above).
extern "C" {
// 0 = ok (*out_len bytes written), 1 = declined (n == 0), -2 = dst too small
// (*out_len = needed), -1 = bad argument.
int qa_pco_compress(int ntype, unsigned level, const uint64_t *lane, size_t n, uint8_t *dst,
                    size_t cap, size_t *out_len);
int qa_pco_decompress(int ntype, const uint8_t *src, size_t len, size_t n, uint64_t *lane);
// Known-answer hooks of pco's own helpers (the functions its unit tests pin):
// the first `count` outputs of Xoroshiro128++ seeded through SplitMix64.
void qa_pco_xoroshiro(uint64_t seed, uint64_t *out, size_t count);
// quantize_weights (ans/encoding.rs:139-158): returns the table size log,
// weights written over counts.
uint32_t qa_pco_quantize_weights(uint32_t *counts, size_t n, uint64_t total, uint32_t max_size_log);
// spread_state_symbols (ans/spec.rs:22-45): table_size entries into out.
int qa_pco_spread(uint32_t size_log, const uint32_t *weights, size_t n, uint32_t *out);
// choose_sample over n floats with the float-mode filter (float.rs:61-71):
// writes the sampled |x| values, returns their count (0 = no sample).
size_t qa_pco_sample_f32(const float *x, size_t n, float *out, size_t cap);
// approx_pair_gcd / snap_to_int_reciprocal / choose_config (float_mult.rs).
int qa_pco_pair_gcd_f64(double greater, double lesser, double *out);
int qa_pco_pair_gcd_f32(float greater, float lesser, float *out);
void qa_pco_snap_f32(float base, float *out_base, float *out_inv);
int qa_pco_float_mult_config_f32(const float *sample, size_t n, float *base, float *inv);
// int_mult::choose_candidate_base on a u32 sample (int_mult.rs:160-168).
int qa_pco_int_candidate_base(const uint32_t *sample, size_t n, uint32_t *base);
// float_quant::estimate_best_k_and_bits_saved on f32 / f64 samples.
uint32_t qa_pco_quant_k_f32(const float *sample, size_t n, double *bits_saved);
uint32_t qa_pco_quant_k_f64(const double *sample, size_t n, double *bits_saved);
// The chosen pco mode and delta encoding of a lane block:
// info[0] mode (0 classic, 1 int mult, 2 float mult, 3 float quant),
// info[1] mode parameter (base latent or k), info[2] delta (0 none,
// 1 consecutive, 2 lookback), info[3] order / window log.
int qa_pco_describe(int ntype, unsigned level, const uint64_t *lane, size_t n, uint64_t info[4]);
// histogram (histograms.rs) of u32 latents into n_bins_log bins:
// writes (count, lower, upper) triples, returns the bin count.
size_t qa_pco_histogram_u32(const uint32_t *latents, size_t n, uint32_t n_bins_log, uint64_t *out,
                            size_t cap);
// choose_lookbacks (delta/lookback.rs:86-138) on u32 latents.
size_t qa_pco_lookbacks_u32(uint32_t window_n_log, uint32_t state_n_log, const uint32_t *latents,
                            size_t n, uint32_t *out);
}
