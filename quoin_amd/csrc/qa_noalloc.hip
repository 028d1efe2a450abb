// qa_noalloc.hip — quoin_gpu_linear_encode_noalloc: the linear residual
// encoder (src/codecs/linear.rs) for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc). Its scratch (block descriptors, per-block results, tile
// sizes and their prefix, the scan workspace) lives in one fixed device array
// of QA_NA_BYTES per GPU, held until the process ends; the call allocates
// nothing on the device.
//
// Every call on a GPU shares that GPU's array, which keeps the last call's
// contents: calls that overlap on one GPU (two host threads, or two streams,
// at once) race on it. The scan tickets are cleared by a fill kernel at the
// start of every call.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "../../include/quoin.h"
#include "qa_kernels.h"

#define QA_NA_BLOCKS 1024           // blocks per call
#define QA_NA_TILES 65536           // 2048-value tiles per call (134217728 values)
#define QA_NA_BYTES (2u << 20)      // the scratch array

static __device__ __align__(256) uint8_t scratch[QA_NA_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex na_mu;
void *na_ptr[kMaxDevices] = {};  // the array's device address, per device

size_t na_align(size_t x) { return (x + 255) & ~(size_t)255; }

// the array's address on the current device, looked up on the first call on
// that device (the name `scratch` is a host-side stand-in, never passed on)
int na_scratch(int dev, uint8_t **out) {
    if (dev < 0 || dev >= kMaxDevices) return QUOIN_ERR_DEVICE;
    std::lock_guard<std::mutex> lk(na_mu);
    if (!na_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return QUOIN_ERR_DEVICE;
        na_ptr[dev] = p;
    }
    *out = (uint8_t *)na_ptr[dev];
    return QUOIN_OK;
}

__global__ void k_na_zero(uint64_t *p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace

extern "C" int quoin_gpu_linear_encode_noalloc(QuoinGpu *g, int32_t kind, int32_t order,
                                               const QuoinLinearInput *in, size_t nb,
                                               uint64_t *out_lens, int32_t *orders,
                                               int32_t *status, void *stream) {
    if (!g || (nb && (!in || !out_lens))) return QUOIN_ERR_NULL;
    if ((kind != QUOIN_LINEAR_ORDERED_DELTA && kind != QUOIN_LINEAR_DELTA2 &&
         kind != QUOIN_LINEAR_DELTA_DP) ||
        order < 0 || order > 4)
        return QUOIN_ERR_UNSUPPORTED;
    if (nb > QA_NA_BLOCKS) return QUOIN_ERR_BUFFER_TOO_SMALL;  // beyond the fixed scratch
    LinD d[QA_NA_BLOCKS];
    LinR r[QA_NA_BLOCKS];
    uint64_t tiles = 0;
    for (size_t i = 0; i < nb; i++) {
        const uint64_t n = in[i].n;
        const bool writes = n > 0 || kind == QUOIN_LINEAR_DELTA2;  // Delta2 of nothing is [order]
        if ((n && !in[i].vals) || (writes && !in[i].dst)) return QUOIN_ERR_NULL;
        if (writes && in[i].dst_cap < quoin_linear_bound(n)) return QUOIN_ERR_BUFFER_TOO_SMALL;
        memset(&d[i], 0, sizeof d[i]);
        d[i].vals = in[i].vals;
        d[i].dst = in[i].dst;
        d[i].n = n;
        d[i].tile0 = tiles;
        d[i].order = (uint32_t)order;
        tiles += qa_lin_enc_tiles(n);
        if (tiles > QA_NA_TILES) return QUOIN_ERR_BUFFER_TOO_SMALL;  // beyond the fixed scratch
    }
    std::lock_guard<std::recursive_mutex> lk(qa_gpu_mutex(g));
    const int dev = qa_gpu_device(g);
    if (hipSetDevice(dev) != hipSuccess) return QUOIN_ERR_DEVICE;
    hipStream_t s = qa_gpu_stream(g, stream);
    if (nb == 0) return QUOIN_OK;
    uint8_t *base = nullptr;
    int rc = na_scratch(dev, &base);
    if (rc != QUOIN_OK) return rc;
    // layout: descriptors, results, tile sizes and prefix, scan workspace
    const size_t o_r = na_align(QA_NA_BLOCKS * sizeof(LinD));
    const size_t o_t = o_r + na_align(QA_NA_BLOCKS * sizeof(LinR));
    const size_t o_s = o_t + na_align((2 * (size_t)QA_NA_TILES + 2) * 8);
    const size_t scan_words = qa_scan_tmp_words(QA_NA_TILES);
    if (o_s + scan_words * 8 > QA_NA_BYTES) return QUOIN_ERR_PANIC;
    LinD *dd = (LinD *)base;
    LinR *dr = (LinR *)(base + o_r);
    uint64_t *tsz = (uint64_t *)(base + o_t), *tpre = tsz + tiles + 1;
    uint64_t *tmp = (uint64_t *)(base + o_s);
    if (hipMemcpyAsync(dd, d, nb * sizeof(LinD), hipMemcpyHostToDevice, s) != hipSuccess)
        return QUOIN_ERR_DEVICE;
    qa_launch(k_na_zero, dim3(1), dim3(64), 0, s, tmp, 2u);  // the scan tickets
    qa_lin_encode_launch((uint32_t)kind, dd, dr, (uint32_t)nb, tiles, tsz, tpre, tmp, s);
    if (hipMemcpyAsync(r, dr, nb * sizeof(LinR), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess)
        return QUOIN_ERR_DEVICE;
    for (size_t i = 0; i < nb; i++) {
        out_lens[i] = r[i].status ? 0 : r[i].len;
        if (orders) orders[i] = (int32_t)r[i].order;
        if (status) status[i] = r[i].status ? 1 : 0;
    }
    return QUOIN_OK;
}
