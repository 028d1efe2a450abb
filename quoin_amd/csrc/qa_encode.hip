// qa_encode.hip — gfx950 kernels of the batched block competition (encode).
//
// Stages (one launch each, every block of every column of the batch at once):
//   1 validity: popcount words -> scan -> compact+lower to the u64 lane
//   2 probe every 32 Ki region (src/encoder.rs:705-750)
//   3 plan blocks per column (src/encoder.rs:154-178)
//   4 full-block features + looks_incompressible (src/encoder.rs:655-683)
//   5 per-tile statistics (one wavefront per 1024-value FastLanes sub-block)
//   6 per-block candidate sizes + the ordered Best replay (src/encoder.rs:99-140,
//     275-469)
//   7 validity blob (src/validity.rs:40-70) and frame layout per column
//   8 emit headers and payloads
#include "qa_kernels.h"

#define WAVE 64
#define NONE32 0xFFFFFFFFu

// ---------------------------------------------------------------------------
// column lookups (columns are sorted by every *_base field)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int col_by_region(const DCol *c, int nc, uint32_t r) {
    int lo = 0, hi = nc - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (c[mid].region_base <= r) lo = mid; else hi = mid - 1;
    }
    return lo;
}
__device__ __forceinline__ int col_by_tile(const DCol *c, int nc, uint64_t t) {
    int lo = 0, hi = nc - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (c[mid].tile_base <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}
__device__ __forceinline__ int col_by_val(const DCol *c, int nc, uint64_t v) {
    int lo = 0, hi = nc - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (c[mid].val_base <= v) lo = mid; else hi = mid - 1;
    }
    return lo;
}
// among columns WITH a bitmap (nvwords > 0): column owning validity word j
__device__ __forceinline__ int col_by_vword(const DCol *c, int nc, uint64_t j) {
    int lo = 0, hi = nc - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (c[mid].vword_base <= j) lo = mid; else hi = mid - 1;
    }
    return lo;  // zero-word columns share the next column's base and lose ties
}
__device__ __forceinline__ uint64_t lane_bytes_of(uint32_t dt) {
    return (dt == QA_I32 || dt == QA_U32 || dt == QA_F32) ? 4 : 8;
}
__device__ __forceinline__ bool is_float(uint32_t dt) { return dt == QA_F64 || dt == QA_F32; }
__device__ __forceinline__ bool is_signed(uint32_t dt) { return dt == QA_I32 || dt == QA_I64; }

// bitmap word j of a column (bits >= n cleared); byte loads: the caller's
// bitmap is ceil(n/8) bytes with no padding or alignment promise
__device__ __forceinline__ uint64_t bm_word(const uint8_t *bm, uint64_t n, uint64_t j) {
    uint64_t nbytes = (n + 7) >> 3, b0 = j * 8;
    uint64_t w = 0;
    for (int k = 0; k < 8; k++)
        if (b0 + k < nbytes) w |= (uint64_t)bm[b0 + k] << (8 * k);
    uint64_t bits = n - j * 64;
    if (bits < 64) w &= (1ull << bits) - 1;
    return w;
}

// ---------------------------------------------------------------------------
// stage 1: validity popcounts, compaction and lowering (src/lib.rs:429-455)
// ---------------------------------------------------------------------------
__global__ void k_vcount(EncArgs a) {
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.nvwords) return;
    int c = col_by_vword(a.cols, a.ncols, j);
    const DCol &col = a.cols[c];
    uint64_t lj = j - col.vword_base;
    uint64_t w = bm_word(col.validity, col.n, lj);
    a.vpre[j] = (uint64_t)__popcll(w);
    // transitions for the validity RLE (src/validity.rs:40-60): runs alternate
    // starting from a 0-run, so bit -1 counts as 0
    uint64_t prev_top = lj ? (bm_word(col.validity, col.n, lj - 1) >> 63) : 0;
    uint64_t t = w ^ ((w << 1) | prev_top);
    uint64_t bits = col.n - lj * 64;
    if (bits < 64) t &= (1ull << bits) - 1;
    a.tpre[j] = (uint64_t)__popcll(t);
}

// per-column counts from the global scans
__global__ void k_col_counts(EncArgs a) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (int)a.ncols) return;
    const DCol &col = a.cols[c];
    DColState &st = a.cst[c];
    if (col.nvwords) {
        uint64_t nv = a.vpre[col.vword_base + col.nvwords] - a.vpre[col.vword_base];
        st.nv = nv;
        st.has_nulls = nv < col.n;
        a.colbase[c] = a.tpre[col.vword_base];
    } else {
        st.nv = col.n;
        st.has_nulls = 0;
        a.colbase[c] = a.nvwords ? a.tpre[col.vword_base] : 0;
    }
    if (c == (int)a.ncols - 1) {
        uint64_t T = a.nvwords ? a.tpre[a.nvwords] : 0;
        a.colbase[a.ncols] = T;
        a.colbase[a.ncols + 1] = T + a.ncols;  // run slots in use
    }
}

__device__ __forceinline__ uint64_t load_lane(const DCol &col, uint64_t i) {
    switch (col.in_kind) {
    case QA_IN_I8: return (uint64_t)(int64_t)((const int8_t *)col.values)[i];
    case QA_IN_I16: return (uint64_t)(int64_t)((const int16_t *)col.values)[i];
    case QA_IN_U8: return (uint64_t)((const uint8_t *)col.values)[i];
    case QA_IN_U16: return (uint64_t)((const uint16_t *)col.values)[i];
    default: break;
    }
    switch (col.dtype) {
    case QA_I32: return (uint64_t)(int64_t)((const int32_t *)col.values)[i];
    case QA_U32: return (uint64_t)((const uint32_t *)col.values)[i];
    case QA_F32: return qa_f32_to_lane(((const uint32_t *)col.values)[i]);
    default: return ((const uint64_t *)col.values)[i];
    }
}

// one thread per logical value: lower to the lane, compacting out nulls
__global__ void k_to_lane(EncArgs a) {
    uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.nvals) return;
    int c = col_by_val(a.cols, a.ncols, g);
    const DCol &col = a.cols[c];
    uint64_t i = g - col.val_base;
    if (i >= col.n) return;
    uint64_t v = load_lane(col, i);
    if (col.nvwords && a.cst[c].has_nulls) {
        uint64_t j = i >> 6;
        uint64_t w = bm_word(col.validity, col.n, j);
        uint32_t b = (uint32_t)(i & 63);
        if (!((w >> b) & 1)) return;
        uint64_t rank = a.vpre[col.vword_base + j] - a.vpre[col.vword_base] +
                        (uint64_t)__popcll(w & ((1ull << b) - 1));
        a.lane[col.lane_off + rank] = v;
    } else {
        a.lane[col.lane_off + i] = v;
    }
}

// ---------------------------------------------------------------------------
// stage 2: region probe (src/encoder.rs:705-750)
// ---------------------------------------------------------------------------
#define RF_DLOW 1u
#define RF_REP 2u

// workgroup-wide features of lane[base .. base+len)
__device__ void probe_range(const uint64_t *v, uint64_t len, uint32_t *exp_range, bool *dlow,
                            bool *rep) {
    __shared__ uint64_t seen[256];
    __shared__ uint32_t red[3];
    for (int k = threadIdx.x; k < 256; k += blockDim.x) seen[k] = 0;
    if (threadIdx.x == 0) {
        red[0] = 0xFFFFFFFFu;
        red[1] = 0;
        red[2] = 0;
    }
    __syncthreads();
    uint32_t mn = 0xFFFFFFFFu, mx = 0, eq = 0;
    for (uint64_t i = threadIdx.x; i < len; i += blockDim.x) {
        uint64_t x = v[i];
        uint32_t e = (uint32_t)((x >> 52) & 0x7FF);
        mn = min(mn, e);
        mx = max(mx, e);
        if (i == 0 || v[i - 1] == x) eq++;
    }
    uint64_t stride = len / 4096 > 1 ? len / 4096 : 1;
    for (uint64_t k = (uint64_t)threadIdx.x * stride; k < len; k += (uint64_t)blockDim.x * stride) {
        uint64_t h = v[k] * 0x9E3779B97F4A7C15ull;
        uint32_t bucket = (uint32_t)((h >> 50) & 0x3FFF);
        atomicOr((unsigned long long *)&seen[bucket >> 6], 1ull << (bucket & 63));
    }
    atomicMin(&red[0], mn);
    atomicMax(&red[1], mx);
    atomicAdd(&red[2], eq);
    __syncthreads();
    uint32_t d = 0;
    for (int k = threadIdx.x; k < 256; k += blockDim.x) d += (uint32_t)__popcll(seen[k]);
    __shared__ uint32_t dsum;
    if (threadIdx.x == 0) dsum = 0;
    __syncthreads();
    atomicAdd(&dsum, d);
    __syncthreads();
    if (len == 0) {
        *exp_range = 0;
        *dlow = true;
        *rep = false;
    } else {
        *exp_range = red[1] - red[0];
        *dlow = dsum < 2048;
        uint64_t c2 = (uint64_t)red[2] * 2;
        if (c2 > 0xFFFFFFFFull) c2 = 0xFFFFFFFFull;
        *rep = c2 > (uint32_t)len;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(256) k_probe_regions(EncArgs a) {
    uint32_t r = blockIdx.x;
    int c = col_by_region(a.cols, a.ncols, r);
    const DCol &col = a.cols[c];
    uint64_t nv = a.cst[c].nv;
    uint64_t start = (uint64_t)(r - col.region_base) * col.region_size;
    if (!col.adaptive || start >= nv) return;
    uint64_t len = nv - start < col.region_size ? nv - start : col.region_size;
    uint32_t er;
    bool dl, rp;
    probe_range(a.lane + col.lane_off + start, len, &er, &dl, &rp);
    if (threadIdx.x == 0) a.rfeat[r] = (dl ? RF_DLOW : 0) | (rp ? RF_REP : 0);
}

// ---------------------------------------------------------------------------
// stage 3: block planning, one thread per column (src/encoder.rs:154-178)
// ---------------------------------------------------------------------------
__global__ void k_plan(EncArgs a) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (int)a.ncols) return;
    const DCol &col = a.cols[c];
    uint64_t nv = a.cst[c].nv;
    uint32_t nb = 0;
    uint64_t start = 0;
    uint32_t R = col.region_size;
    for (uint32_t lr = 0; lr < col.nregions; lr++) {
        uint32_t slot = col.region_base + lr;
        DBlock &b = a.blocks[slot];
        uint64_t rs = (uint64_t)lr * R;
        if (rs >= nv) {
            b.len = 0;
            a.owner[slot] = NONE32;
            continue;
        }
        if (rs < start) {  // covered by a grown block
            b.len = 0;
            continue;
        }
        uint64_t end;
        if (col.adaptive) {
            uint32_t f = a.rfeat[slot];
            uint64_t be = rs + R < nv ? rs + R : nv;
            end = (f & (RF_DLOW | RF_REP)) ? (rs + QA_MAX_Q < nv ? rs + QA_MAX_Q : nv) : be;
        } else {
            end = rs + R < nv ? rs + R : nv;
        }
        b.start = rs;
        b.len = (uint32_t)(end - rs);
        b.col = (uint32_t)c;
        b.ntiles = (uint32_t)((end - rs + QA_TILE - 1) / QA_TILE);
        for (uint32_t k = lr; (uint64_t)k * R < end && k < col.nregions; k++)
            a.owner[col.region_base + k] = slot;
        start = end;
        nb++;
    }
    a.cst[c].nblocks = nb;
}

// ---------------------------------------------------------------------------
// stage 4: features of the whole block + incompressibility probe
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_block_feats(EncArgs a) {
    uint32_t slot = blockIdx.x;
    DBlock &b = a.blocks[slot];
    if (b.len == 0) return;
    const DCol &col = a.cols[b.col];
    const uint64_t *v = a.lane + col.lane_off + b.start;
    uint64_t n = b.len;
    uint32_t er;
    bool dl, rp;
    probe_range(v, n, &er, &dl, &rp);
    // looks_incompressible (src/encoder.rs:655-683): sampled value range and
    // zig-zag delta range both within 4 bits of the lane width
    __shared__ uint64_t smin, smax, sdz;
    if (threadIdx.x == 0) {
        smin = v[0];
        smax = v[0];
        sdz = 0;
    }
    __syncthreads();
    uint64_t step = n / 2048 > 1 ? n / 2048 : 1;
    uint64_t mn = v[0], mx = v[0], dz = 0;
    for (uint64_t i = step * (1 + threadIdx.x); i < n; i += step * blockDim.x) {
        uint64_t x = v[i];
        mn = x < mn ? x : mn;
        mx = x > mx ? x : mx;
        uint64_t z = qa_zigzag(x - v[i - 1]);
        dz = z > dz ? z : dz;
    }
    atomicMin((unsigned long long *)&smin, mn);
    atomicMax((unsigned long long *)&smax, mx);
    atomicMax((unsigned long long *)&sdz, dz);
    __syncthreads();
    if (threadIdx.x == 0) {
        b.exp_range = er;
        b.distinct_low = dl;
        b.repeats = rp;
        bool inc = false;
        if (!is_float(col.dtype) && !dl && !rp && n >= 256) {
            uint32_t thr = (uint32_t)(lane_bytes_of(col.dtype) * 8) - 4;
            uint32_t vw = qa_width(smax - smin), dw = qa_width(sdz);
            inc = vw >= thr && dw >= thr;
        }
        b.incompressible = inc;
        b.v0 = v[0];
        b.v1 = n > 1 ? v[1] : 0;
    }
}

// ---------------------------------------------------------------------------
// stage 5: per-tile statistics, one wavefront per 1024-value sub-block
// ---------------------------------------------------------------------------
template <typename T, typename Op>
__device__ __forceinline__ T wave_reduce(T v, Op op) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = op(v, (T)__shfl_xor(v, d, 64));
    return v;
}
struct OpMinU { __device__ uint64_t operator()(uint64_t a, uint64_t b) const { return a < b ? a : b; } };
struct OpMaxU { __device__ uint64_t operator()(uint64_t a, uint64_t b) const { return a > b ? a : b; } };
struct OpMinS { __device__ int64_t operator()(int64_t a, int64_t b) const { return a < b ? a : b; } };
struct OpMaxS { __device__ int64_t operator()(int64_t a, int64_t b) const { return a > b ? a : b; } };
struct OpAdd { __device__ uint64_t operator()(uint64_t a, uint64_t b) const { return a + b; } };

// resolve a global tile slot to (column, block slot, tile index within block);
// returns false for an inactive slot
__device__ __forceinline__ bool tile_locate(const EncArgs &a, uint64_t t, int *cp, uint32_t *slotp,
                                            uint32_t *btp) {
    int c = col_by_tile(a.cols, a.ncols, t);
    const DCol &col = a.cols[c];
    uint64_t lt = t - col.tile_base;
    uint32_t lr = (uint32_t)(lt / col.tpr), tin = (uint32_t)(lt % col.tpr);
    if (lr >= col.nregions) return false;
    uint32_t o = a.owner[col.region_base + lr];
    if (o == NONE32) return false;
    const DBlock &b = a.blocks[o];
    uint32_t bt = (col.region_base + lr - o) * col.tpr + tin;
    if ((uint64_t)bt * QA_TILE >= b.len) return false;
    *cp = c;
    *slotp = o;
    *btp = bt;
    return true;
}

__global__ void __launch_bounds__(256) k_tile_stats(EncArgs a) {
    const int lane = threadIdx.x & 63;
    uint64_t t = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= a.ntiles) return;
    int c;
    uint32_t slot, bt;
    DTile &ts = a.tiles[t];
    if (!tile_locate(a, t, &c, &slot, &bt)) {
        if (lane == 0) ts.active = 0;
        return;
    }
    const DCol &col = a.cols[c];
    const DBlock &b = a.blocks[slot];
    const uint64_t *v = a.lane + col.lane_off + b.start;
    uint64_t k0 = (uint64_t)bt * QA_TILE;
    uint32_t cnt = (uint32_t)min((uint64_t)QA_TILE, (uint64_t)b.len - k0);
    uint64_t umin = ~0ull, umax = 0, dzmin = ~0ull, dzmax = 0, xb = 0;
    int64_t smin = INT64_MAX, smax = INT64_MIN;
    uint32_t cbad = 0, sbad = 0, chg = 0;
    uint64_t stride = b.v1 - b.v0;
    for (uint32_t k = lane; k < cnt; k += WAVE) {
        uint64_t i = k0 + k;
        uint64_t x = v[i];
        uint64_t p = i ? v[i - 1] : b.v0;   // delta predecessor (base = v0)
        uint64_t px = i ? v[i - 1] : 0;     // xorz predecessor
        umin = x < umin ? x : umin;
        umax = x > umax ? x : umax;
        smin = (int64_t)x < smin ? (int64_t)x : smin;
        smax = (int64_t)x > smax ? (int64_t)x : smax;
        uint64_t z = qa_zigzag(x - p);
        dzmin = z < dzmin ? z : dzmin;
        dzmax = z > dzmax ? z : dzmax;
        xb += qa_varint_len(x ^ px);
        cbad += x != b.v0;
        sbad += x != b.v0 + i * stride;
        chg += (i > 0 && x != p);
    }
    umin = wave_reduce(umin, OpMinU());
    umax = wave_reduce(umax, OpMaxU());
    smin = wave_reduce(smin, OpMinS());
    smax = wave_reduce(smax, OpMaxS());
    dzmin = wave_reduce(dzmin, OpMinU());
    dzmax = wave_reduce(dzmax, OpMaxU());
    xb = wave_reduce(xb, OpAdd());
    uint64_t cb = wave_reduce((uint64_t)cbad, OpAdd());
    uint64_t sb = wave_reduce((uint64_t)sbad, OpAdd());
    uint64_t ch = wave_reduce((uint64_t)chg, OpAdd());
    if (lane == 0) {
        ts.umin = umin;
        ts.umax = umax;
        ts.smin = smin;
        ts.smax = smax;
        ts.dzmin = dzmin;
        ts.dzmax = dzmax;
        ts.xorz_bytes = xb;
        ts.const_bad = (uint32_t)cb;
        ts.stride_bad = (uint32_t)sb;
        ts.changes = (uint32_t)ch;
        ts.cnt = cnt;
        ts.active = 1;
    }
}

// ---------------------------------------------------------------------------
// stage 6: candidate sizes and the ordered replay of Best::consider
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t weight_of(uint32_t mode) {
    switch (mode) {  // src/encoder.rs:79-89
    case QA_M_RAW: case QA_M_CONST: case QA_M_STRIDE: return 0;
    case QA_M_XORZ: case QA_M_FOR: case QA_M_DELTA: case QA_M_RLE: return 1;
    case QA_M_ALP: case QA_M_ALP_RD: case QA_M_DICT: return 2;
    default: return 6;  // FLOAT_MULT
    }
}
__device__ __forceinline__ uint64_t penalty_of(uint32_t mode, uint64_t lambda, uint64_t dec) {
    return (lambda * weight_of(mode) * dec) >> 8;  // src/encoder.rs:94-97
}

__global__ void k_block_decide(EncArgs a) {
    uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= a.nregions) return;
    DBlock &b = a.blocks[slot];
    if (b.len == 0) return;
    const DCol &col = a.cols[b.col];
    const DTile *tl = a.tiles + col.tile_base + (uint64_t)(slot - col.region_base) * col.tpr;
    uint64_t n = b.len;
    bool fl = is_float(col.dtype), sg = is_signed(col.dtype);
    bool reduced = col.level == QA_FASTEST;
    uint64_t lambda = reduced ? 16 : 4;
    uint64_t dec = n * lane_bytes_of(col.dtype);
    uint64_t raw = n * lane_bytes_of(col.dtype);
    for (int k = 0; k < C_N; k++) b.sz[k] = ~0ull;
    b.sz[C_RAW] = raw;
    uint32_t mode = QA_M_RAW;
    uint64_t best_len = raw, best_score = raw;
    if (!fl && b.incompressible) {
        b.mode = mode;
        b.payload_len = raw;
        atomicAdd(&a.wins[mode & 63], 1ull);
        return;
    }
    bool cst = true, str = n >= 2;
    uint64_t xz = 0, fsz = qa_varint_len(n), dsz = 8 + qa_varint_len(n);
    for (uint32_t t = 0; t < b.ntiles; t++) {
        const DTile &s = tl[t];
        cst &= s.const_bad == 0;
        str &= s.stride_bad == 0;
        xz += s.xorz_bytes;
        uint64_t range = sg ? (uint64_t)s.smax - (uint64_t)s.smin : s.umax - s.umin;
        fsz += 9 + 128ull * qa_width(range);
        dsz += 9 + 128ull * qa_width(s.dzmax - s.dzmin);
    }
    if (cst) b.sz[C_CONST] = 8;
    if (str) b.sz[C_STRIDE] = 16;
    if (!reduced) b.sz[C_XORZ] = xz;
    b.sz[C_FOR] = fsz;
    b.sz[C_DELTA] = dsz;
    const uint32_t order[6] = {QA_M_RAW, QA_M_CONST, QA_M_STRIDE, QA_M_XORZ, QA_M_FOR, QA_M_DELTA};
    for (int k = 1; k < 6; k++) {
        uint64_t sz = b.sz[k];
        if (sz == ~0ull) continue;
        uint64_t sc = sz + penalty_of(order[k], lambda, dec);
        if (sc < best_score) {
            best_score = sc;
            best_len = sz;
            mode = order[k];
        }
    }
    b.compressible = best_len * 20 < (n * 8) * 19;
    b.mode = mode;
    atomicAdd(&a.wins[mode & 63], 1ull);
    b.payload_len = best_len;
    // sub-block offsets inside the winner payload
    if (mode == QA_M_FOR || mode == QA_M_DELTA) {
        uint64_t off = mode == QA_M_FOR ? qa_varint_len(n) : 8 + qa_varint_len(n);
        for (uint32_t t = 0; t < b.ntiles; t++) {
            DTile &s = a.tiles[col.tile_base + (uint64_t)(slot - col.region_base) * col.tpr + t];
            uint64_t range = mode == QA_M_DELTA ? s.dzmax - s.dzmin
                             : sg ? (uint64_t)s.smax - (uint64_t)s.smin : s.umax - s.umin;
            s.w = qa_width(range);
            s.off = off;
            off += 9 + 128ull * s.w;
        }
    }
}

// ---------------------------------------------------------------------------
// stage 7: validity blob + frame layout
// ---------------------------------------------------------------------------
// transition positions (column-local bit index) into tpos[global transition]
__global__ void k_vtrans_pos(EncArgs a) {
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.nvwords) return;
    int c = col_by_vword(a.cols, a.ncols, j);
    const DCol &col = a.cols[c];
    uint64_t lj = j - col.vword_base;
    uint64_t w = bm_word(col.validity, col.n, lj);
    uint64_t prev_top = lj ? (bm_word(col.validity, col.n, lj - 1) >> 63) : 0;
    uint64_t t = w ^ ((w << 1) | prev_top);
    uint64_t bits = col.n - lj * 64;
    if (bits < 64) t &= (1ull << bits) - 1;
    uint64_t k = a.tpre[j];
    while (t) {
        uint32_t b = (uint32_t)__ffsll((long long)t) - 1;
        a.tpos[k++] = lj * 64 + b;
        t &= t - 1;
    }
}

// run slot s = global transition index + column index; run k of column c ends
// at transition k (or at n for the last run)
__device__ __forceinline__ int col_by_run(const EncArgs &a, uint64_t s) {
    int lo = 0, hi = (int)a.ncols - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (a.colbase[mid] + (uint64_t)mid <= s) lo = mid; else hi = mid - 1;
    }
    return lo;
}
__global__ void k_vrun_sizes(EncArgs a) {
    uint64_t total = a.colbase[a.ncols + 1];
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < total;
         s += (uint64_t)gridDim.x * blockDim.x) {
        int c = col_by_run(a, s);
        const DCol &col = a.cols[c];
        uint64_t first = a.colbase[c] + c, T = a.colbase[c + 1] - a.colbase[c];
        uint64_t k = s - first;
        if (!col.nvwords || !a.cst[c].has_nulls) {
            a.rsz[s] = 0;
            continue;
        }
        uint64_t st = k == 0 ? 0 : a.tpos[a.colbase[c] + k - 1];
        uint64_t en = k == T ? col.n : a.tpos[a.colbase[c] + k];
        a.rsz[s] = qa_varint_len(en - st);
    }
}

__global__ void __launch_bounds__(256) k_col_layout(EncArgs a) {
    // one workgroup per column: validity blob choice, then frame offsets
    int c = blockIdx.x;
    const DCol &col = a.cols[c];
    DColState &st = a.cst[c];
    __shared__ uint64_t vlen_s, frames_off_s;
    if (threadIdx.x == 0) {
        uint64_t vlen = 0;
        if (st.has_nulls) {
            uint64_t first = a.colbase[c] + c, T = a.colbase[c + 1] - a.colbase[c];
            uint64_t rle = a.rsz[first + T + 1] - a.rsz[first];
            uint64_t raw = (col.n + 7) / 8;
            st.vblob_rle = rle;
            vlen = 1 + (rle < raw ? rle : raw);
        }
        st.vblob_len = vlen;
        frames_off_s = 16 + (st.has_nulls ? qa_varint_len(vlen) + vlen : 0);
        st.frames_off = frames_off_s;
    }
    __syncthreads();
    // frame offsets: exclusive scan of (9 + payload) over the column's slots
    uint64_t carry = frames_off_s;
    for (uint32_t base = 0; base < col.nregions; base += blockDim.x) {
        uint32_t lr = base + threadIdx.x;
        uint64_t v = 0;
        if (lr < col.nregions) {
            const DBlock &b = a.blocks[col.region_base + lr];
            if (b.len) v = 9 + b.payload_len;
        }
        // simple workgroup scan through LDS
        __shared__ uint64_t buf[256];
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int d = 1; d < 256; d <<= 1) {
            uint64_t o = threadIdx.x >= (unsigned)d ? buf[threadIdx.x - d] : 0;
            __syncthreads();
            buf[threadIdx.x] += o;
            __syncthreads();
        }
        uint64_t incl = buf[threadIdx.x];
        if (lr < col.nregions) a.blocks[col.region_base + lr].frame_off = carry + incl - v;
        uint64_t tot = buf[255];
        __syncthreads();
        carry += tot;
    }
    if (threadIdx.x == 0) {
        st.stream_len = carry;
        if (carry > col.dst_cap) st.err |= QA_ERR_OVERFLOW;
    }
}

// ---------------------------------------------------------------------------
// stage 8: emission
// ---------------------------------------------------------------------------
__global__ void k_emit_header(EncArgs a) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (int)a.ncols) return;
    const DCol &col = a.cols[c];
    const DColState &st = a.cst[c];
    if (st.err) return;
    uint8_t *d = col.dst;
    // src/format.rs:56-63
    d[0] = 'F'; d[1] = 'C'; d[2] = 'R'; d[3] = '1';
    d[4] = 2;
    d[5] = st.has_nulls ? 1 : 0;
    d[6] = 16;
    d[7] = (uint8_t)col.dtype;
    qa_st64u(d + 8, col.n);
    if (st.has_nulls) {
        uint32_t k = qa_put_varint(d + 16, st.vblob_len);
        uint64_t raw = (col.n + 7) / 8;
        d[16 + k] = st.vblob_rle < raw ? 1 : 0;
    }
}

// RLE validity blob: one thread per run
__global__ void k_emit_vruns(EncArgs a) {
    uint64_t total = a.colbase[a.ncols + 1];
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < total;
         s += (uint64_t)gridDim.x * blockDim.x) {
        int c = col_by_run(a, s);
        const DCol &col = a.cols[c];
        const DColState &st = a.cst[c];
        if (!col.nvwords || !st.has_nulls || st.err) continue;
        uint64_t raw = (col.n + 7) / 8;
        if (!(st.vblob_rle < raw)) continue;
        uint64_t first = a.colbase[c] + c, T = a.colbase[c + 1] - a.colbase[c];
        uint64_t k = s - first;
        uint64_t stp = k == 0 ? 0 : a.tpos[a.colbase[c] + k - 1];
        uint64_t en = k == T ? col.n : a.tpos[a.colbase[c] + k];
        uint64_t off = 16 + qa_varint_len(st.vblob_len) + 1 + (a.rsz[s] - a.rsz[first]);
        qa_put_varint(col.dst + off, en - stp);
    }
}

// raw validity blob: copy the bitmap bytes verbatim (src/validity.rs:63-68)
__global__ void k_emit_vraw(EncArgs a) {
    uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.nvwords * 8) return;
    int c = col_by_vword(a.cols, a.ncols, g / 8);
    const DCol &col = a.cols[c];
    const DColState &st = a.cst[c];
    if (!col.nvwords || !st.has_nulls || st.err) return;
    uint64_t raw = (col.n + 7) / 8;
    if (st.vblob_rle < raw) return;
    uint64_t bi = g - col.vword_base * 8;
    if (bi >= raw) return;
    uint64_t off = 16 + qa_varint_len(st.vblob_len) + 1;
    col.dst[off + bi] = col.validity[bi];
}

__global__ void k_emit_frames(EncArgs a) {
    uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= a.nregions) return;
    const DBlock &b = a.blocks[slot];
    if (b.len == 0) return;
    const DCol &col = a.cols[b.col];
    if (a.cst[b.col].err) return;
    uint8_t *d = col.dst + b.frame_off;
    d[0] = (uint8_t)b.mode;  // src/encoder.rs:752-759
    qa_st32u(d + 1, b.len);
    qa_st32u(d + 5, (uint32_t)b.payload_len);
    uint8_t *p = d + 9;
    switch (b.mode) {
    case QA_M_CONST: qa_st64u(p, b.v0); break;
    case QA_M_STRIDE:
        qa_st64u(p, b.v0);
        qa_st64u(p + 8, b.v1 - b.v0);
        break;
    case QA_M_FOR: qa_put_varint(p, b.len); break;
    case QA_M_DELTA:
        qa_st64u(p, b.v0);
        qa_put_varint(p + 8, b.len);
        break;
    default: break;
    }
}

// FastLanes pack of 1024 residuals held in LDS into the stream (bitpack.rs:47-133):
// output word (k, l) gathers the rows whose bit range meets word k of lane l.
__device__ __forceinline__ void pack_store(const uint64_t *res, uint32_t w, uint8_t *dst, int lane) {
    if (w == 0) return;
    if (w <= 32) {
        for (uint32_t idx = lane; idx < 32 * w; idx += WAVE) {
            uint32_t k = idx >> 5, l = idx & 31;
            uint32_t lo_bit = 32 * k;
            uint32_t r0 = lo_bit / w, r1 = min(31u, (lo_bit + 31) / w);
            uint32_t word = 0;
            for (uint32_t r = r0; r <= r1; r++) {
                uint32_t v = (uint32_t)res[r * 32 + l];
                int off = (int)(r * w) - (int)lo_bit;
                word |= off >= 0 ? (off < 32 ? v << off : 0u) : v >> (-off);
            }
            qa_st32u(dst + 4 * (uint64_t)idx, word);
        }
    } else {
        for (uint32_t idx = lane; idx < 16 * w; idx += WAVE) {
            uint32_t k = idx >> 4, l = idx & 15;
            uint32_t lo_bit = 64 * k;
            uint32_t r0 = lo_bit / w, r1 = min(63u, (lo_bit + 63) / w);
            uint64_t word = 0;
            for (uint32_t r = r0; r <= r1; r++) {
                uint64_t v = res[r * 16 + l];
                int off = (int)(r * w) - (int)lo_bit;
                word |= off >= 0 ? (off < 64 ? v << off : 0ull) : v >> (-off);
            }
            qa_st64u(dst + 8 * (uint64_t)idx, word);
        }
    }
}

__global__ void __launch_bounds__(64) k_emit_tiles(EncArgs a) {
    // one wavefront per workgroup: every early exit below is wave-uniform, so
    // the barrier before packing is reached by the whole group
    __shared__ uint64_t res[QA_TILE];
    const int lane = threadIdx.x;
    uint64_t t = blockIdx.x;
    if (t >= a.ntiles) return;
    const DTile &ts = a.tiles[t];
    if (!ts.active) return;
    int c;
    uint32_t slot, bt;
    if (!tile_locate(a, t, &c, &slot, &bt)) return;
    const DCol &col = a.cols[c];
    if (a.cst[c].err) return;
    const DBlock &b = a.blocks[slot];
    const uint64_t *v = a.lane + col.lane_off + b.start;
    uint64_t k0 = (uint64_t)bt * QA_TILE;
    uint32_t cnt = ts.cnt;
    uint8_t *p = col.dst + b.frame_off + 9;
    if (b.mode == QA_M_RAW) {
        uint32_t lb = (uint32_t)lane_bytes_of(col.dtype);
        for (uint32_t k = lane; k < cnt; k += WAVE) {
            uint64_t x = v[k0 + k];
            uint8_t *q = p + (k0 + k) * lb;
            if (col.dtype == QA_F32) qa_st32u(q, qa_lane_to_f32(x));
            else if (lb == 4) qa_st32u(q, (uint32_t)x);
            else qa_st64u(q, x);
        }
        return;
    }
    if (b.mode != QA_M_FOR && b.mode != QA_M_DELTA) return;
    bool sg = is_signed(col.dtype);
    uint64_t mn = b.mode == QA_M_DELTA ? ts.dzmin : sg ? (uint64_t)ts.smin : ts.umin;
    for (uint32_t k = lane; k < QA_TILE; k += WAVE) {
        uint64_t r = 0;
        if (k < cnt) {
            uint64_t i = k0 + k, x = v[i];
            if (b.mode == QA_M_DELTA) {
                uint64_t pv = i ? v[i - 1] : b.v0;
                r = qa_zigzag(x - pv) - mn;
            } else {
                r = x - mn;
            }
        }
        res[k] = r;
    }
    __syncthreads();
    uint8_t *q = p + ts.off;
    if (lane == 0) {
        q[0] = (uint8_t)ts.w;
        qa_st64u(q + 1, mn);
    }
    pack_store(res, ts.w, q + 9, lane);
}

// ---------------------------------------------------------------------------
// host-side launcher
// ---------------------------------------------------------------------------
static inline unsigned nb(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }
// launch unless the grid is empty (a zero-size grid is a launch error)
#define QA_LAUNCH(k, grid, ...)                                   \
    do {                                                          \
        dim3 g_ = (grid);                                         \
        if (g_.x) hipLaunchKernelGGL(k, g_, __VA_ARGS__);         \
    } while (0)

void qa_encode_launch(const EncArgs &a, hipStream_t s, int) {
    if (a.nvwords) {
        QA_LAUNCH(k_vcount, dim3(nb(a.nvwords, 256)), dim3(256), 0, s, a);
        qa_exclusive_scan(a.vpre, a.nvwords, a.vpre, a.scan_tmp, s);
        qa_exclusive_scan(a.tpre, a.nvwords, a.tpre, a.scan_tmp, s);
    }
    QA_LAUNCH(k_col_counts, dim3(nb(a.ncols, 64)), dim3(64), 0, s, a);
    QA_LAUNCH(k_to_lane, dim3(nb(a.nvals, 256)), dim3(256), 0, s, a);
    QA_LAUNCH(k_probe_regions, dim3(a.nregions), dim3(256), 0, s, a);
    QA_LAUNCH(k_plan, dim3(nb(a.ncols, 64)), dim3(64), 0, s, a);
    QA_LAUNCH(k_block_feats, dim3(a.nregions), dim3(256), 0, s, a);
    QA_LAUNCH(k_tile_stats, dim3(nb(a.ntiles, 4)), dim3(256), 0, s, a);
    QA_LAUNCH(k_block_decide, dim3(nb(a.nregions, 64)), dim3(64), 0, s, a);
    uint64_t nruns_max = a.ntrans_max + a.ncols;
    if (a.nvwords) {
        QA_LAUNCH(k_vtrans_pos, dim3(nb(a.nvwords, 256)), dim3(256), 0, s, a);
        unsigned g = nb(nruns_max, 256) < 4096 ? nb(nruns_max, 256) : 4096;
        QA_LAUNCH(k_vrun_sizes, dim3(g), dim3(256), 0, s, a);
        qa_exclusive_scan_dn(a.rsz, nruns_max, a.colbase + a.ncols + 1, a.rsz, a.scan_tmp, s);
    }
    QA_LAUNCH(k_col_layout, dim3(a.ncols), dim3(256), 0, s, a);
    QA_LAUNCH(k_emit_header, dim3(nb(a.ncols, 64)), dim3(64), 0, s, a);
    if (a.nvwords) {
        unsigned g = nb(nruns_max, 256) < 4096 ? nb(nruns_max, 256) : 4096;
        QA_LAUNCH(k_emit_vruns, dim3(g), dim3(256), 0, s, a);
        QA_LAUNCH(k_emit_vraw, dim3(nb(a.nvwords * 8, 256)), dim3(256), 0, s, a);
    }
    QA_LAUNCH(k_emit_frames, dim3(nb(a.nregions, 64)), dim3(64), 0, s, a);
    QA_LAUNCH(k_emit_tiles, dim3((unsigned)a.ntiles), dim3(64), 0, s, a);
}
