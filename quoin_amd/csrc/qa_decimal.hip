// qa_decimal.hip — the decimal container's device work (src/decimal.rs):
// significands of 4, 8, 16 or 32 bytes (DECIMAL32 / 64 sign-extended to the
// 128-bit container, as capi/src/typed.rs:209-218 widens them), the signed
// minimum over the valid values, the 64-bit limbs of v - vmin at their logical
// positions (compressed afterwards as U64 columns by the engine), and on
// decode the recombination vmin + limbs into the caller's width, null slots 0.
#include "qa_kernels.h"

namespace {

// significand i as nl little-endian 64-bit limbs, sign-extended from in_w bytes
__device__ __forceinline__ void load_sig(const uint8_t *v, uint32_t in_w, uint64_t i, uint32_t nl,
                                         uint64_t (&x)[4]) {
    const uint8_t *p = v + i * in_w;
    if (in_w == 4) {
        x[0] = (uint64_t)(int64_t)(int32_t)((uint32_t)p[0] | (uint32_t)p[1] << 8 |
                                            (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24);
    } else {
        x[0] = qa_ld64u(p);
    }
    const uint32_t given = in_w <= 8 ? 1 : in_w / 8;
    for (uint32_t l = 1; l < nl; l++)
        x[l] = l < given ? qa_ld64u(p + 8 * l) : (uint64_t)((int64_t)x[given - 1] >> 63);
}

// signed a < b over nl limbs (decimal.rs:286-302)
__device__ __forceinline__ bool lt_signed(const uint64_t (&a)[4], const uint64_t (&b)[4], uint32_t nl) {
    const uint64_t sa = a[nl - 1] >> 63, sb = b[nl - 1] >> 63;
    if (sa != sb) return sa == 1;
    for (int l = (int)nl - 1; l >= 0; l--)
        if (a[l] != b[l]) return a[l] < b[l];
    return false;
}

__device__ __forceinline__ bool is_valid(const uint8_t *bm, uint64_t i) {
    return !bm || ((bm[i >> 3] >> (i & 7)) & 1);
}

#define DM_WG 256

// per workgroup: (any valid, the minimum) into part[blockIdx]
__global__ void __launch_bounds__(DM_WG) k_decm_min(DecmArgs a) {
    __shared__ uint64_t sm[DM_WG][5];
    uint64_t best[4] = {0, 0, 0, 0};
    bool has = false;
    for (uint64_t i = (uint64_t)blockIdx.x * DM_WG + threadIdx.x; i < a.n;
         i += (uint64_t)gridDim.x * DM_WG) {
        if (!is_valid(a.bitmap, i)) continue;
        uint64_t x[4] = {0, 0, 0, 0};
        load_sig(a.values, a.in_w, i, a.nl, x);
        if (!has || lt_signed(x, best, a.nl)) {
            for (int l = 0; l < 4; l++) best[l] = x[l];
            has = true;
        }
    }
    for (int l = 0; l < 4; l++) sm[threadIdx.x][l] = best[l];
    sm[threadIdx.x][4] = has;
    __syncthreads();
    if (threadIdx.x == 0) {
        bool h = false;
        uint64_t m[4] = {0, 0, 0, 0};
        for (int t = 0; t < DM_WG; t++) {
            if (!sm[t][4]) continue;
            uint64_t x[4] = {sm[t][0], sm[t][1], sm[t][2], sm[t][3]};
            if (!h || lt_signed(x, m, a.nl)) {
                for (int l = 0; l < 4; l++) m[l] = x[l];
                h = true;
            }
        }
        for (int l = 0; l < 4; l++) a.part[blockIdx.x * 5 + l] = m[l];
        a.part[blockIdx.x * 5 + 4] = h;
    }
}

// the column minimum from the parts (0 when no value is valid: decimal.rs:196)
__global__ void k_decm_min_final(DecmArgs a, uint32_t nparts) {
    if (threadIdx.x) return;
    bool h = false;
    uint64_t m[4] = {0, 0, 0, 0};
    for (uint32_t t = 0; t < nparts; t++) {
        if (!a.part[t * 5 + 4]) continue;
        uint64_t x[4] = {a.part[t * 5], a.part[t * 5 + 1], a.part[t * 5 + 2], a.part[t * 5 + 3]};
        if (!h || lt_signed(x, m, a.nl)) {
            for (int l = 0; l < 4; l++) m[l] = x[l];
            h = true;
        }
    }
    for (int l = 0; l < 4; l++) a.vmin[l] = m[l];
}

// limb l of v - vmin (wrapping, decimal.rs:304-318) at every logical position
__global__ void __launch_bounds__(DM_WG) k_decm_limbs(DecmArgs a) {
    const uint64_t i = (uint64_t)blockIdx.x * DM_WG + threadIdx.x;
    if (i >= a.n) return;
    uint64_t x[4] = {0, 0, 0, 0};
    if (is_valid(a.bitmap, i)) load_sig(a.values, a.in_w, i, a.nl, x);
    uint64_t borrow = 0;
    for (uint32_t l = 0; l < a.nl; l++) {
        const uint64_t m = a.vmin[l];
        const uint64_t d1 = x[l] - m, b1 = x[l] < m;
        const uint64_t d2 = d1 - borrow, b2 = d1 < borrow;
        a.limbs[l * a.n + i] = d2;
        borrow = b1 | b2;
    }
}

// decode: vmin + limbs (wrapping, decimal.rs:320-335) into out_w bytes per
// value (the low bytes: DECIMAL64 / 32 narrow as `as i64` / `as i32`), null 0
__global__ void __launch_bounds__(DM_WG) k_decm_join(DecmArgs a) {
    const uint64_t i = (uint64_t)blockIdx.x * DM_WG + threadIdx.x;
    if (i >= a.n) return;
    uint64_t v[4] = {0, 0, 0, 0};
    if (is_valid(a.bitmap, i)) {
        uint64_t carry = 0;
        for (uint32_t l = 0; l < a.nl; l++) {
            const uint64_t m = a.vmin[l], o = a.limbs[l * a.n + i];
            const uint64_t s1 = m + o, c1 = s1 < m;
            const uint64_t s2 = s1 + carry, c2 = s2 < s1;
            v[l] = s2;
            carry = c1 | c2;
        }
    }
    uint8_t *o = a.out + i * a.out_w;
    for (uint32_t b = 0; b < a.out_w; b++) o[b] = (uint8_t)(v[b >> 3] >> (8 * (b & 7)));
}

}  // namespace

void qa_decm_limbs_launch(const DecmArgs &a, hipStream_t s) {
    if (!a.n) {
        qa_launch(k_decm_min_final, dim3(1), dim3(64), 0, s, a, 0u);
        return;
    }
    uint64_t g = (a.n + DM_WG - 1) / DM_WG;
    const uint32_t parts = (uint32_t)(g < QA_DECM_PARTS ? g : QA_DECM_PARTS);
    qa_launch(k_decm_min, dim3(parts), dim3(DM_WG), 0, s, a);
    qa_launch(k_decm_min_final, dim3(1), dim3(64), 0, s, a, parts);
    qa_launch(k_decm_limbs, dim3((unsigned)g), dim3(DM_WG), 0, s, a);
}

void qa_decm_join_launch(const DecmArgs &a, hipStream_t s) {
    if (!a.n) return;
    qa_launch(k_decm_join, dim3((unsigned)((a.n + DM_WG - 1) / DM_WG)), dim3(DM_WG), 0, s, a);
}
