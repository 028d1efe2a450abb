// qa_hostcoders.h — the sequential byte coders that run on host cores
// (src/entropy/rc.rs, src/codecs/lz.rs, the tag layer of src/entropy/mod.rs),
// for the engine's Balanced / High / Max path.
#pragma once
#include <cstddef>
#include <cstdint>
#include <functional>
#include <vector>

namespace qahost {
// entropy::code_residuals: [tag] ++ coded, the range coder and rANS chosen by
// lambda, with the LZ cascade when allow_lz
void code_residuals(const uint8_t *d, size_t n, uint64_t lambda, bool allow_lz, std::vector<uint8_t> &out);
// entropy::decode_residuals; QUOIN_OK or QUOIN_ERR_CORRUPT
int decode_residuals(const uint8_t *p, size_t len, uint64_t max_len, std::vector<uint8_t> &out);
// lz::lz_compress / lz_decompress over bytes
void lz_compress(const uint8_t *d, size_t n, std::vector<uint8_t> &out);
bool lz_decompress(const uint8_t *s, size_t len, uint64_t expected, std::vector<uint8_t> &out);
// f(i) for i < n over up to 16 host threads; err[i] = QUOIN_ERR_PANIC where f throws
void parallel_for(size_t n, std::vector<int> &err, const std::function<void(size_t)> &f);
}  // namespace qahost
