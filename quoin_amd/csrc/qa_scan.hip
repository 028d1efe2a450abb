// qa_scan.hip — device exclusive prefix sums over u64 counts, used for
// variable-length layout (validity compaction, run lists, varint offsets).
// Two kernels: per-chunk sums, the last workgroup to finish (an atomic ticket,
// no waiting) scanning the chunk sums; then per-chunk scans seeded with the
// chunk's carry. Each thread holds 16 consecutive elements (four 32-byte
// loads), so both passes stream at full width. out has n + 1 entries: out[n]
// is the total. tmp[y] is scan y's ticket (y = 0, or 0 and 1 for a pair),
// tmp[2 + y (nb + 1) + b] its chunk sums (then their exclusive prefix) and
// tmp[2 + y (nb + 1) + nb] its total; every scan leaves its ticket at zero
// (the apply kernel resets it) and the host zeroes both tickets when the
// workspace is allocated.
#include "qa_kernels.h"

#define SCAN_WG 256
#define SCAN_PER 16
#define SCAN_CHUNK (SCAN_WG * SCAN_PER)

// block-wide exclusive scan of one value per thread; returns the exclusive
// prefix, *total receives the block sum.
__device__ uint64_t block_excl_scan(uint64_t v, uint64_t *total) {
    __shared__ uint64_t wsum[17];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    const uint64_t inc = qa_wave_scan_add64(v);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        const uint64_t s = lane < nw ? wsum[lane] : 0;
        const uint64_t si = qa_wave_scan_add64(s);
        if (lane < nw) wsum[lane] = si - s;
        if (lane == nw - 1) wsum[16] = si;
    }
    __syncthreads();
    const uint64_t r = inc - v + wsum[wid];
    *total = wsum[16];
    __syncthreads();
    return r;
}

__device__ __forceinline__ uint64_t scan_n(uint64_t n, const uint64_t *ndev) {
    return ndev ? *ndev : n;
}

// the 16 elements of thread threadIdx.x in chunk blockIdx.x (0 past n)
__device__ __forceinline__ void scan_load16(const uint64_t *in, uint64_t n, uint64_t i0,
                                            uint64_t (&v)[SCAN_PER]) {
    if (i0 + SCAN_PER <= n && ((uintptr_t)(in + i0) & 15) == 0) {
        const uint4 *p = (const uint4 *)(in + i0);
#pragma unroll
        for (int c = 0; c < SCAN_PER / 2; c++) {
            const uint4 q = p[c];
            v[2 * c] = ((uint64_t)q.y << 32) | q.x;
            v[2 * c + 1] = ((uint64_t)q.w << 32) | q.z;
        }
    } else {
#pragma unroll
        for (int j = 0; j < SCAN_PER; j++) v[j] = i0 + j < n ? in[i0 + j] : 0;
    }
}

// up to two independent scans of the same length per launch pair: blockIdx.y
// picks the array (its own input, output and workspace)
struct ScanJobs {
    const uint64_t *in[2];
    uint64_t *out[2];
};

__global__ void __launch_bounds__(SCAN_WG) k_scan_reduce(ScanJobs j, uint64_t nmax,
                                                         const uint64_t *ndev, uint64_t *tmp) {
    __shared__ bool last;
    const uint64_t *in = j.in[blockIdx.y];
    const uint64_t n = scan_n(nmax, ndev);
    // chunks holding data (at least one): the others take no ticket, so a grid
    // sized for nmax costs nothing past n
    const uint64_t nu = n ? (n + SCAN_CHUNK - 1) / SCAN_CHUNK : 1;
    if (blockIdx.x >= nu) return;
    const uint64_t i0 = (uint64_t)blockIdx.x * SCAN_CHUNK + (uint64_t)threadIdx.x * SCAN_PER;
    uint64_t v[SCAN_PER];
    scan_load16(in, n, i0, v);
    uint64_t s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_PER; j++) s += v[j];
    uint64_t tot;
    block_excl_scan(s, &tot);
    const uint32_t nb = gridDim.x;
    unsigned long long *ticket = (unsigned long long *)tmp + blockIdx.y;
    uint64_t *cs = tmp + 2 + (uint64_t)blockIdx.y * (nb + 1);
    if (threadIdx.x == 0) {
        cs[blockIdx.x] = tot;
        __threadfence();
        last = atomicAdd(ticket, 1ull) == nu - 1;
    }
    __syncthreads();
    if (!last) return;
    // the last workgroup: every chunk sum is visible; scan them in place
    __threadfence();
    uint64_t carry = 0;
    for (uint64_t b = 0; b < nu; b += SCAN_WG) {
        const uint64_t i = b + threadIdx.x;
        const uint64_t c =
            i < nu ? __hip_atomic_load(&cs[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
        uint64_t t2;
        const uint64_t ex = block_excl_scan(c, &t2);
        if (i < nu) cs[i] = carry + ex;
        carry += t2;
    }
    if (threadIdx.x == 0) cs[nb] = carry;
}

__global__ void __launch_bounds__(SCAN_WG) k_scan_apply(ScanJobs j, uint64_t nmax,
                                                        const uint64_t *ndev, uint64_t *tmp) {
    const uint64_t *in = j.in[blockIdx.y];
    uint64_t *out = j.out[blockIdx.y];
    const uint64_t n = scan_n(nmax, ndev);
    const uint32_t nb = gridDim.x;
    const uint64_t *cs = tmp + 2 + (uint64_t)blockIdx.y * (nb + 1);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        out[n] = cs[nb];
        tmp[blockIdx.y] = 0;  // the ticket, for the next scan
    }
    if (blockIdx.x * (uint64_t)SCAN_CHUNK >= n) return;
    const uint64_t i0 = (uint64_t)blockIdx.x * SCAN_CHUNK + (uint64_t)threadIdx.x * SCAN_PER;
    uint64_t v[SCAN_PER];
    scan_load16(in, n, i0, v);
    uint64_t s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_PER; j++) {
        const uint64_t x = v[j];
        v[j] = s;
        s += x;
    }
    uint64_t tot;
    const uint64_t carry = cs[blockIdx.x] + block_excl_scan(s, &tot);
    if (i0 + SCAN_PER <= n && ((uintptr_t)(out + i0) & 15) == 0) {
        uint4 *p = (uint4 *)(out + i0);
#pragma unroll
        for (int c = 0; c < SCAN_PER / 2; c++) {
            const uint64_t a = carry + v[2 * c], b = carry + v[2 * c + 1];
            p[c] = make_uint4((uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32));
        }
    } else {
#pragma unroll
        for (int j = 0; j < SCAN_PER; j++)
            if (i0 + j < n) out[i0 + j] = carry + v[j];
    }
}

// Exclusive scan of in[0..n) into out[0..n], out[n] = total. When ndev is
// given, n is read from device memory (n <= nmax) so a data-dependent length
// needs no host round trip. tmp needs qa_scan_tmp_words(nmax) words, the
// first of them zero before the first scan; in and out may alias.
size_t qa_scan_tmp_words(uint64_t n) { return n / SCAN_CHUNK + 1 + 3; }

void qa_exclusive_scan_dn(const uint64_t *in, uint64_t nmax, const uint64_t *ndev, uint64_t *out,
                          uint64_t *tmp, hipStream_t s) {
    const unsigned grid = (unsigned)(nmax / SCAN_CHUNK + 1);
    const ScanJobs j{{in, nullptr}, {out, nullptr}};
    qa_launch(k_scan_reduce, dim3(grid), dim3(SCAN_WG), 0, s, j, nmax, ndev, tmp);
    qa_launch(k_scan_apply, dim3(grid), dim3(SCAN_WG), 0, s, j, nmax, ndev, tmp);
}

// Two exclusive scans of n elements each in one launch pair (twice the
// workgroups of one scan, half the launches). tmp needs twice
// qa_scan_tmp_words(n) words, its first two (the tickets) zero.
void qa_exclusive_scan2(const uint64_t *in0, uint64_t *out0, const uint64_t *in1, uint64_t *out1,
                        uint64_t n, uint64_t *tmp, hipStream_t s) {
    const unsigned grid = (unsigned)(n / SCAN_CHUNK + 1);
    const ScanJobs j{{in0, in1}, {out0, out1}};
    qa_launch(k_scan_reduce, dim3(grid, 2), dim3(SCAN_WG), 0, s, j, n, nullptr, tmp);
    qa_launch(k_scan_apply, dim3(grid, 2), dim3(SCAN_WG), 0, s, j, n, nullptr, tmp);
}

void qa_exclusive_scan(const uint64_t *in, uint64_t n, uint64_t *out, uint64_t *tmp,
                       hipStream_t s) {
    qa_exclusive_scan_dn(in, n, nullptr, out, tmp, s);
}
