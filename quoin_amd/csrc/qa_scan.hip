// qa_scan.hip — device exclusive prefix sums over u64 counts, used for
// variable-length layout (validity compaction, run lists, varint offsets).
// Three phases: per-chunk reduce, one-workgroup scan of the chunk sums, per-chunk
// scan with the chunk's carry. out has n+1 entries: out[n] is the total.
#include "qa_kernels.h"

#define SCAN_WG 256
#define SCAN_PER 8
#define SCAN_CHUNK (SCAN_WG * SCAN_PER)

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v) { return qa_wave_scan_add64(v); }

// block-wide exclusive scan of one value per thread; returns exclusive prefix,
// *total receives the block sum.
__device__ uint64_t block_excl_scan(uint64_t v, uint64_t *total) {
    __shared__ uint64_t wsum[17];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    uint64_t inc = wave_incl_scan(v);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        uint64_t s = lane < nw ? wsum[lane] : 0;
        uint64_t si = wave_incl_scan(s);
        if (lane < nw) wsum[lane] = si - s;
        if (lane == nw - 1) wsum[16] = si;
    }
    __syncthreads();
    uint64_t r = inc - v + wsum[wid];
    *total = wsum[16];
    __syncthreads();
    return r;
}

__device__ __forceinline__ uint64_t scan_n(uint64_t n, const uint64_t *ndev) {
    return ndev ? *ndev : n;
}

__global__ void __launch_bounds__(SCAN_WG) k_scan_reduce(const uint64_t *in, uint64_t nmax,
                                                         const uint64_t *ndev, uint64_t *chunk) {
    uint64_t n = scan_n(nmax, ndev);
    uint64_t base = (uint64_t)blockIdx.x * SCAN_CHUNK;
    if (base >= n) {
        if (threadIdx.x == 0) chunk[blockIdx.x] = 0;
        return;
    }
    uint64_t s = 0;
    for (int j = 0; j < SCAN_PER; j++) {
        uint64_t i = base + (uint64_t)j * SCAN_WG + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint64_t tot;
    block_excl_scan(s, &tot);
    if (threadIdx.x == 0) chunk[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(1024) k_scan_chunks(uint64_t *chunk, uint64_t nmax,
                                                      const uint64_t *ndev) {
    uint64_t n = scan_n(nmax, ndev);
    uint64_t nchunks = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
    uint64_t carry = 0;
    for (uint64_t b = 0; b < nchunks; b += blockDim.x) {
        uint64_t i = b + threadIdx.x;
        uint64_t v = i < nchunks ? chunk[i] : 0;
        uint64_t tot;
        uint64_t ex = block_excl_scan(v, &tot);
        if (i < nchunks) chunk[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) chunk[nchunks] = carry;
}

__global__ void __launch_bounds__(SCAN_WG) k_scan_apply(const uint64_t *in, uint64_t nmax,
                                                        const uint64_t *ndev, const uint64_t *chunk,
                                                        uint64_t *out) {
    __shared__ uint64_t vals[SCAN_CHUNK];
    uint64_t n = scan_n(nmax, ndev);
    uint64_t base = (uint64_t)blockIdx.x * SCAN_CHUNK;
    if (base > n) return;
    uint64_t nch = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
    if (blockIdx.x == n / SCAN_CHUNK && threadIdx.x == 0) out[n] = chunk[nch];
    if (base == n) return;
    for (int j = 0; j < SCAN_PER; j++) {
        uint64_t i = base + (uint64_t)j * SCAN_WG + threadIdx.x;
        vals[j * SCAN_WG + threadIdx.x] = i < n ? in[i] : 0;
    }
    __syncthreads();
    uint64_t loc[SCAN_PER];
    uint64_t s = 0;
    for (int j = 0; j < SCAN_PER; j++) {
        loc[j] = s;
        s += vals[threadIdx.x * SCAN_PER + j];
    }
    uint64_t tot;
    uint64_t ex = block_excl_scan(s, &tot);
    uint64_t carry = chunk[blockIdx.x] + ex;
    for (int j = 0; j < SCAN_PER; j++) {
        uint64_t i = base + (uint64_t)threadIdx.x * SCAN_PER + j;
        if (i < n) out[i] = carry + loc[j];
    }
}

// Exclusive scan of in[0..n) into out[0..n], out[n] = total. When ndev is
// given, n is read from device memory (n <= nmax) so a data-dependent length
// needs no host round trip. tmp needs qa_scan_tmp_words(nmax) words; in and
// out may alias.
size_t qa_scan_tmp_words(uint64_t n) { return (n + SCAN_CHUNK - 1) / SCAN_CHUNK + 2; }

void qa_exclusive_scan_dn(const uint64_t *in, uint64_t nmax, const uint64_t *ndev, uint64_t *out,
                          uint64_t *tmp, hipStream_t s) {
    unsigned grid = (unsigned)(nmax / SCAN_CHUNK + 1);
    hipLaunchKernelGGL(k_scan_reduce, dim3(grid), dim3(SCAN_WG), 0, s, in, nmax, ndev, tmp);
    hipLaunchKernelGGL(k_scan_chunks, dim3(1), dim3(1024), 0, s, tmp, nmax, ndev);
    hipLaunchKernelGGL(k_scan_apply, dim3(grid), dim3(SCAN_WG), 0, s, in, nmax, ndev, tmp, out);
}

void qa_exclusive_scan(const uint64_t *in, uint64_t n, uint64_t *out, uint64_t *tmp,
                       hipStream_t s) {
    qa_exclusive_scan_dn(in, n, nullptr, out, tmp, s);
}
