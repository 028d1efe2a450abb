// qa_encode.hip — gfx950 kernels of the batched block competition (encode).
//
// Two passes over the values:
//   pass 1 (k_region_stats): one workgroup per region slot computes every
//     block-independent statistic of its 1024-value tiles and the region's
//     probe features (src/encoder.rs:696-750) and incompressibility samples
//     (src/encoder.rs:655-683).
//   pass 2 (k_emit_tiles): one wavefront per tile writes the winner's payload.
// Between them: block planning (src/encoder.rs:154-178), features of grown
// blocks, the per-block candidate sizes and the ordered Best::consider replay
// (src/encoder.rs:99-140, 275-469), the validity blob (src/validity.rs:40-70)
// and the frame layout (src/encoder.rs:184-214, 752-759).
#include "qa_dev.h"


// ---------------------------------------------------------------------------
// stage 1: validity popcounts / transitions, lane setup, compaction
// ---------------------------------------------------------------------------
__global__ void k_vcount(EncArgs a) {
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.nvwords) return;
    int c = col_by_vword(a.cols, a.ncols, j);
    const DCol &col = a.cols[c];
    uint64_t lj = j - col.vword_base;
    uint64_t w = bm_word(col.validity, col.n, lj);
    a.vpre[j] = (uint64_t)__popcll(w);
    // transitions for the validity RLE: runs alternate starting from a 0-run,
    // so bit -1 counts as 0 (src/validity.rs:40-60)
    uint64_t prev_top = lj ? (bm_word(col.validity, col.n, lj - 1) >> 63) : 0;
    uint64_t t = w ^ ((w << 1) | prev_top);
    uint64_t bits = col.n - lj * 64;
    if (bits < 64) t &= (1ull << bits) - 1;
    a.tpre[j] = (uint64_t)__popcll(t);
}

__device__ __forceinline__ uint32_t lkind_of(const DCol &col) {
    switch (col.in_kind) {
    case QA_IN_I8: return QA_LK_I8;
    case QA_IN_I16: return QA_LK_I16;
    case QA_IN_U8: return QA_LK_U8;
    case QA_IN_U16: return QA_LK_U16;
    default: break;
    }
    switch (col.dtype) {
    case QA_I32: return QA_LK_I32;
    case QA_U32: return QA_LK_U32;
    case QA_F32: return QA_LK_F32;
    default: return QA_LK_U64;
    }
}

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
    st.lkind = lkind_of(col);
    st.lwidth = col.in_width;
    // without nulls the kernels read the caller's values in place
    st.lanep = st.has_nulls ? (const void *)((const uint8_t *)a.lane + col.lane_off * 8)
                            : col.values;
    if (c == (int)a.ncols - 1) {
        uint64_t T = a.nvwords ? a.tpre[a.nvwords] : 0;
        a.colbase[a.ncols] = T;
        a.colbase[a.ncols + 1] = T + a.ncols;  // run slots in use
    }
}

// compaction of the valid values, in the input's own element width
// (src/validity.rs:122-127); one workgroup per 16 validity words
__global__ void __launch_bounds__(256) k_compact(EncArgs a) {
    __shared__ uint64_t words[16], wpre[16];
    __shared__ int wcol[16];
    uint64_t w0 = (uint64_t)blockIdx.x * 16;
    if (threadIdx.x < 16) {
        uint64_t j = w0 + threadIdx.x;
        if (j < a.nvwords) {
            int c = col_by_vword(a.cols, a.ncols, j);
            const DCol &col = a.cols[c];
            wcol[threadIdx.x] = c;
            words[threadIdx.x] = bm_word(col.validity, col.n, j - col.vword_base);
            wpre[threadIdx.x] = a.vpre[j] - a.vpre[col.vword_base];
        } else {
            wcol[threadIdx.x] = -1;
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 16 * 64; k += 256) {
        int wi = k >> 6;
        int c = wcol[wi];
        if (c < 0) continue;
        const DCol &col = a.cols[c];
        if (!a.cst[c].has_nulls) continue;
        uint64_t lj = w0 + wi - col.vword_base;
        uint64_t i = lj * 64 + (k & 63);
        if (i >= col.n) continue;
        uint64_t w = words[wi];
        uint32_t b = k & 63;
        if (!((w >> b) & 1)) continue;
        uint64_t rank = wpre[wi] + (uint64_t)__popcll(w & ((1ull << b) - 1));
        uint8_t *dst = (uint8_t *)a.lane + col.lane_off * 8;
        const uint8_t *src = (const uint8_t *)col.values;
        switch (col.in_width) {
        case 1: dst[rank] = src[i]; break;
        case 2: ((uint16_t *)dst)[rank] = ((const uint16_t *)src)[i]; break;
        case 4: ((uint32_t *)dst)[rank] = ((const uint32_t *)src)[i]; break;
        default: ((uint64_t *)dst)[rank] = ((const uint64_t *)src)[i];
        }
    }
}

// ---------------------------------------------------------------------------
// stage 2: fused tile statistics + region probe, one workgroup per region
// ---------------------------------------------------------------------------
struct RegionAcc {
    uint32_t exp_min, exp_max, eq, dsum;
    uint64_t smin, smax, sdz;
};

__global__ void __launch_bounds__(1024) k_region_stats(EncArgs a) {
    __shared__ uint64_t seen[256];
    __shared__ RegionAcc acc;
    const uint32_t r = blockIdx.x;
    const int c = col_by_region(a.cols, a.ncols, r);
    const DCol &col = a.cols[c];
    const DColState &st = a.cst[c];
    const uint64_t nv = st.nv;
    const uint64_t rs = (uint64_t)(r - col.region_base) * col.region_size;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    DTile *tl = a.tiles + col.tile_base + (uint64_t)(r - col.region_base) * col.tpr;
    if (rs >= nv) {
        for (uint32_t t = threadIdx.x; t < col.tpr; t += blockDim.x) tl[t].active = 0;
        return;
    }
    const uint64_t len = nv - rs < col.region_size ? nv - rs : col.region_size;
    for (int k = threadIdx.x; k < 256; k += blockDim.x) seen[k] = 0;
    if (threadIdx.x == 0) {
        acc.exp_min = 0xFFFFFFFFu;
        acc.exp_max = 0;
        acc.eq = 0;
        acc.dsum = 0;
        acc.smin = ~0ull;
        acc.smax = 0;
        acc.sdz = 0;
    }
    __syncthreads();
    const uint32_t ntl = (uint32_t)((len + QA_TILE - 1) / QA_TILE);
    uint32_t eq_region = 0;  // consecutive-equal pairs inside the region (lane 0 sums)
    for (uint32_t t = wv; t < col.tpr; t += nwv) {
        DTile &ts = tl[t];
        if (t >= ntl) {
            if (lane == 0) ts.active = 0;
            continue;
        }
        const uint64_t k0 = rs + (uint64_t)t * QA_TILE;  // column index of the tile start
        const uint32_t cnt = (uint32_t)min((uint64_t)QA_TILE, rs + len - k0);
        uint64_t umin = ~0ull, umax = 0, dzmin = ~0ull, dzmax = 0, xb = 0;
        int64_t smin = INT64_MAX, smax = INT64_MIN;
        uint32_t emin = 0xFFFFFFFFu, emax = 0, chg = 0, eqp = 0, arith = 1, same = 1;
        const uint64_t vf = lane_at(st, k0);
        const uint64_t step = cnt > 1 ? lane_at(st, k0 + 1) - vf : 0;
        for (uint32_t k = lane; k < cnt; k += WAVE) {
            uint64_t x = lane_at(st, k0 + k);
            umin = x < umin ? x : umin;
            umax = x > umax ? x : umax;
            smin = (int64_t)x < smin ? (int64_t)x : smin;
            smax = (int64_t)x > smax ? (int64_t)x : smax;
            uint32_t e = (uint32_t)((x >> 52) & 0x7FF);
            emin = min(emin, e);
            emax = max(emax, e);
            same &= x == vf;
            if (k > 0) {
                uint64_t p = lane_at(st, k0 + k - 1);
                uint64_t z = qa_zigzag(x - p);
                dzmin = z < dzmin ? z : dzmin;
                dzmax = z > dzmax ? z : dzmax;
                xb += qa_varint_len(x ^ p);
                chg += x != p;
                arith &= (x - p) == step;
                eqp += x == p;
            } else if (k0 > rs) {
                eqp += x == lane_at(st, k0 - 1);  // pair across tiles of this region
            }
        }
        umin = wave_reduce(umin, OpMinU());
        umax = wave_reduce(umax, OpMaxU());
        smin = wave_reduce(smin, OpMinS());
        smax = wave_reduce(smax, OpMaxS());
        dzmin = wave_reduce(dzmin, OpMinU());
        dzmax = wave_reduce(dzmax, OpMaxU());
        xb = wave_reduce(xb, OpAdd());
        uint64_t chg64 = wave_reduce((uint64_t)chg, OpAdd());
        uint64_t eq64 = wave_reduce((uint64_t)eqp, OpAdd());
        emin = (uint32_t)wave_reduce((uint64_t)emin, OpMinU());
        emax = (uint32_t)wave_reduce((uint64_t)emax, OpMaxU());
        arith = wave_reduce(arith, OpAnd());
        same = wave_reduce(same, OpAnd());
        if (lane == 0) {
            ts.umin = umin;
            ts.umax = umax;
            ts.smin = smin;
            ts.smax = smax;
            ts.dzmin = dzmin;
            ts.dzmax = dzmax;
            ts.xorz_rest = xb;
            ts.chg_rest = (uint32_t)chg64;
            ts.exp_min = emin;
            ts.exp_max = emax;
            ts.v_first = vf;
            ts.v_last = lane_at(st, k0 + cnt - 1);
            ts.step = step;
            ts.dz_first = k0 ? qa_zigzag(vf - lane_at(st, k0 - 1)) : 0;
            ts.cnt = cnt;
            ts.arith = (uint8_t)arith;
            ts.allsame = (uint8_t)same;
            ts.active = 1;
            eq_region += (uint32_t)eq64;
            atomicMin(&acc.exp_min, emin);
            atomicMax(&acc.exp_max, emax);
        }
    }
    if (lane == 0) atomicAdd(&acc.eq, eq_region);
    // distinct estimate over a strided sample (src/encoder.rs:728-743)
    const uint64_t stride = len / 4096 > 1 ? len / 4096 : 1;
    for (uint64_t k = (uint64_t)threadIdx.x * stride; k < len; k += (uint64_t)blockDim.x * stride) {
        uint64_t h = lane_at(st, rs + k) * 0x9E3779B97F4A7C15ull;
        uint32_t bucket = (uint32_t)((h >> 50) & 0x3FFF);
        atomicOr((unsigned long long *)&seen[bucket >> 6], 1ull << (bucket & 63));
    }
    // incompressibility samples for the region taken as a block
    const uint64_t step = len / 2048 > 1 ? len / 2048 : 1;
    const uint64_t v0 = lane_at(st, rs);
    uint64_t mn = v0, mx = v0, dz = 0;
    for (uint64_t i = step * (1 + threadIdx.x); i < len; i += step * blockDim.x) {
        uint64_t x = lane_at(st, rs + i);
        mn = x < mn ? x : mn;
        mx = x > mx ? x : mx;
        uint64_t z = qa_zigzag(x - lane_at(st, rs + i - 1));
        dz = z > dz ? z : dz;
    }
    mn = wave_reduce(mn, OpMinU());
    mx = wave_reduce(mx, OpMaxU());
    dz = wave_reduce(dz, OpMaxU());
    if (lane == 0) {
        atomicMin((unsigned long long *)&acc.smin, mn);
        atomicMax((unsigned long long *)&acc.smax, mx);
        atomicMax((unsigned long long *)&acc.sdz, dz);
    }
    __syncthreads();
    uint32_t d = 0;
    for (int k = threadIdx.x; k < 256; k += blockDim.x) d += (uint32_t)__popcll(seen[k]);
    d = (uint32_t)wave_reduce((uint64_t)d, OpAdd());
    if (lane == 0) atomicAdd(&acc.dsum, d);
    __syncthreads();
    if (threadIdx.x == 0) {
        bool dl = acc.dsum < 2048;
        uint64_t c2 = (uint64_t)(acc.eq + 1) * 2;  // consec_eq counts element 0 against itself
        if (c2 > 0xFFFFFFFFull) c2 = 0xFFFFFFFFull;
        bool rp = c2 > (uint32_t)len;
        bool inc = false;
        if (!is_float(col.dtype) && !dl && !rp && len >= 256) {
            uint32_t thr = (uint32_t)(lane_bytes_of(col.dtype) * 8) - 4;
            inc = qa_width(acc.smax - acc.smin) >= thr && qa_width(acc.sdz) >= thr;
        }
        a.rfeat[r] = (dl ? RF_DLOW : 0) | (rp ? RF_REP : 0) | (inc ? RF_INC : 0) |
                     ((acc.exp_max - acc.exp_min) << 8);
    }
}

// ---------------------------------------------------------------------------
// stage 3: block planning (src/encoder.rs:154-178), one workgroup per column
// ---------------------------------------------------------------------------
#define PLAN_CHUNK 4096

__global__ void __launch_bounds__(256) k_plan(EncArgs a) {
    __shared__ uint8_t feat[PLAN_CHUNK];
    __shared__ uint32_t own[PLAN_CHUNK];
    __shared__ uint32_t next_s, nb_s, cur_owner;
    const int c = blockIdx.x;
    const DCol &col = a.cols[c];
    const uint64_t nv = a.cst[c].nv;
    const uint32_t R = col.region_size;
    const uint32_t nreg = (uint32_t)((nv + R - 1) / R);  // regions holding values
    if (threadIdx.x == 0) {
        next_s = 0;
        nb_s = 0;
        cur_owner = NONE32;
    }
    for (uint32_t base = 0; base < col.nregions; base += PLAN_CHUNK) {
        uint32_t cn = min((uint32_t)PLAN_CHUNK, col.nregions - base);
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < cn; k += blockDim.x)
            feat[k] = base + k < nreg ? (uint8_t)(a.rfeat[col.region_base + base + k] & 3) : 0;
        __syncthreads();
        if (threadIdx.x == 0) {
            // the sequential walk, in LDS: a region starts a block iff it is
            // the region after the previous block's end
            uint32_t ns = next_s, nb = nb_s, ow = cur_owner;
            for (uint32_t k = 0; k < cn; k++) {
                uint32_t lr = base + k;
                if (lr >= nreg) {
                    own[k] = NONE32;
                    continue;
                }
                if (lr == ns) {
                    uint64_t rs = (uint64_t)lr * R;
                    uint64_t end = (col.adaptive && feat[k]) ? min(rs + (uint64_t)QA_MAX_Q, nv)
                                                             : min(rs + (uint64_t)R, nv);
                    ns = (uint32_t)((end + R - 1) / R);
                    nb++;
                    ow = col.region_base + lr;
                }
                own[k] = ow;
            }
            next_s = ns;
            nb_s = nb;
            cur_owner = ow;
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < cn; k += blockDim.x) {
            uint32_t lr = base + k, slot = col.region_base + lr;
            DBlock &b = a.blocks[slot];
            uint32_t o = own[k];
            a.owner[slot] = o;
            if (o == slot) {
                uint64_t rs = (uint64_t)lr * R;
                uint64_t end = (col.adaptive && feat[k]) ? min(rs + (uint64_t)QA_MAX_Q, nv)
                                                         : min(rs + (uint64_t)R, nv);
                b.start = rs;
                b.len = (uint32_t)(end - rs);
                b.col = (uint32_t)c;
                b.ntiles = (uint32_t)((end - rs + QA_TILE - 1) / QA_TILE);
            } else {
                b.len = 0;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) a.cst[c].nblocks = nb_s;
}

// ---------------------------------------------------------------------------
// stage 4: features of blocks grown past one region (131072 values)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_grown_feats(EncArgs a) {
    __shared__ uint64_t seen[256];
    __shared__ uint64_t smin, smax, sdz;
    __shared__ uint32_t dsum, emin_s, emax_s, eq_s;
    const uint32_t slot = blockIdx.x;
    DBlock &b = a.blocks[slot];
    if (b.len == 0) return;
    const DCol &col = a.cols[b.col];
    const DColState &st = a.cst[b.col];
    const uint32_t f = a.rfeat[slot];
    if (b.len <= col.region_size) {  // a one-region block: its region's probe
        if (threadIdx.x == 0) {
            b.distinct_low = (f & RF_DLOW) != 0;
            b.repeats = (f & RF_REP) != 0;
            b.incompressible = (f & RF_INC) != 0;
            b.exp_range = f >> 8;
        }
        return;
    }
    const uint64_t n = b.len, s0 = b.start;
    for (int k = threadIdx.x; k < 256; k += blockDim.x) seen[k] = 0;
    if (threadIdx.x == 0) {
        uint64_t v0 = lane_at(st, s0);
        smin = v0;
        smax = v0;
        sdz = 0;
        dsum = 0;
        emin_s = 0xFFFFFFFFu;
        emax_s = 0;
        eq_s = 0;
    }
    __syncthreads();
    const uint64_t stride = n / 4096 > 1 ? n / 4096 : 1;
    for (uint64_t k = (uint64_t)threadIdx.x * stride; k < n; k += (uint64_t)blockDim.x * stride) {
        uint64_t h = lane_at(st, s0 + k) * 0x9E3779B97F4A7C15ull;
        uint32_t bucket = (uint32_t)((h >> 50) & 0x3FFF);
        atomicOr((unsigned long long *)&seen[bucket >> 6], 1ull << (bucket & 63));
    }
    const uint64_t step = n / 2048 > 1 ? n / 2048 : 1;
    uint64_t mn = ~0ull, mx = 0, dz = 0;
    for (uint64_t i = step * (1 + threadIdx.x); i < n; i += step * blockDim.x) {
        uint64_t x = lane_at(st, s0 + i);
        mn = x < mn ? x : mn;
        mx = x > mx ? x : mx;
        uint64_t z = qa_zigzag(x - lane_at(st, s0 + i - 1));
        dz = z > dz ? z : dz;
    }
    atomicMin((unsigned long long *)&smin, mn);
    atomicMax((unsigned long long *)&smax, mx);
    atomicMax((unsigned long long *)&sdz, dz);
    // exp range and consecutive-equal count combine from the tiles
    const DTile *tl = a.tiles + col.tile_base + (uint64_t)(slot - col.region_base) * col.tpr;
    for (uint32_t t = threadIdx.x; t < b.ntiles; t += blockDim.x) {
        const DTile &ts = tl[t];
        atomicMin(&emin_s, ts.exp_min);
        atomicMax(&emax_s, ts.exp_max);
        uint32_t e = ts.cnt - 1 - ts.chg_rest;
        if (t > 0) e += ts.v_first == tl[t - 1].v_last;
        atomicAdd(&eq_s, e);
    }
    __syncthreads();
    uint32_t d = 0;
    for (int k = threadIdx.x; k < 256; k += blockDim.x) d += (uint32_t)__popcll(seen[k]);
    atomicAdd(&dsum, d);
    __syncthreads();
    if (threadIdx.x == 0) {
        bool dl = dsum < 2048;
        uint64_t c2 = (uint64_t)(eq_s + 1) * 2;
        if (c2 > 0xFFFFFFFFull) c2 = 0xFFFFFFFFull;
        bool rp = c2 > (uint32_t)n;
        bool inc = false;
        if (!is_float(col.dtype) && !dl && !rp && n >= 256) {
            uint32_t thr = (uint32_t)(lane_bytes_of(col.dtype) * 8) - 4;
            inc = qa_width(smax - smin) >= thr && qa_width(sdz) >= thr;
        }
        b.distinct_low = dl;
        b.repeats = rp;
        b.incompressible = inc;
        b.exp_range = emax_s - emin_s;
    }
}

// ---------------------------------------------------------------------------
// stage 5: candidate sizes and the ordered replay of Best::consider, one
// wavefront per block
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// stage 6: validity blob + frame layout
// ---------------------------------------------------------------------------
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

// run slot s = global transition index + column index; run k of a column ends
// at its transition k (or at n for the last run)
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
    const int c = blockIdx.x;
    const DCol &col = a.cols[c];
    DColState &st = a.cst[c];
    __shared__ uint64_t frames_off_s;
    __shared__ uint64_t buf[256];
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
    uint64_t carry = frames_off_s;
    for (uint32_t base = 0; base < col.nregions; base += blockDim.x) {
        uint32_t lr = base + threadIdx.x;
        uint64_t v = 0;
        if (lr < col.nregions) {
            const DBlock &b = a.blocks[col.region_base + lr];
            if (b.len) v = 9 + b.payload_len;
        }
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
// stage 7: emission
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

// raw validity blob: the bitmap bytes verbatim (src/validity.rs:63-68)
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

__global__ void __launch_bounds__(64) k_emit_tiles(EncArgs a) {
    // one wavefront per workgroup: every early exit below is wave-uniform, so
    // the barriers are reached by the whole group
    __shared__ uint64_t res[QA_TILE];
    __shared__ uint32_t outw[2 * QA_TILE + 4];
    const int lane = threadIdx.x;
    const uint64_t t = blockIdx.x;
    const DTile &ts = a.tiles[t];
    if (!ts.active) return;
    const int c = col_by_tile(a.cols, a.ncols, t);
    const DCol &col = a.cols[c];
    const DColState &st = a.cst[c];
    if (st.err) return;
    const uint64_t lt = t - col.tile_base;
    const uint32_t lr = (uint32_t)(lt / col.tpr);
    const uint32_t o = a.owner[col.region_base + lr];
    const DBlock &b = a.blocks[o];
    const uint32_t bt = (col.region_base + lr - o) * col.tpr + (uint32_t)(lt % col.tpr);
    const uint64_t k0 = b.start + (uint64_t)bt * QA_TILE;
    const uint32_t cnt = ts.cnt;
    uint8_t *p = col.dst + b.frame_off + 9;
    if (b.mode == QA_M_RAW) {
        const uint32_t lb = (uint32_t)lane_bytes_of(col.dtype);
        for (uint32_t k = lane; k < cnt; k += WAVE) {
            uint64_t x = lane_at(st, k0 + k);
            if (col.dtype == QA_F32) outw[k] = qa_lane_to_f32(x);
            else if (lb == 4) outw[k] = (uint32_t)x;
            else {
                outw[2 * k] = (uint32_t)x;
                outw[2 * k + 1] = (uint32_t)(x >> 32);
            }
        }
        __syncthreads();
        wave_store_bytes(p + (uint64_t)bt * QA_TILE * lb, outw, cnt * lb, lane);
        return;
    }
    if (b.mode != QA_M_FOR && b.mode != QA_M_DELTA) return;
    const bool sg = is_signed(col.dtype);
    uint64_t mn;
    if (b.mode == QA_M_DELTA) {
        uint64_t dmx;
        tile_dz(ts, bt == 0, &mn, &dmx);
    } else {
        mn = sg ? (uint64_t)ts.smin : ts.umin;
    }
    for (uint32_t k = lane; k < QA_TILE; k += WAVE) {
        uint64_t r = 0;
        if (k < cnt) {
            uint64_t x = lane_at(st, k0 + k);
            if (b.mode == QA_M_DELTA) {
                uint64_t pv = (bt == 0 && k == 0) ? x : lane_at(st, k0 + k - 1);
                r = qa_zigzag(x - pv) - mn;
            } else {
                r = x - mn;
            }
        }
        res[k] = r;
    }
    __syncthreads();
    uint8_t *q = p + ts.off;
    if (lane == 0) q[0] = (uint8_t)ts.w;
    if (lane >= 1 && lane < 9) q[lane] = (uint8_t)(mn >> (8 * (lane - 1)));
    if (ts.w == 0) return;
    pack_lds(res, ts.w, outw, lane);
    __syncthreads();
    wave_store_bytes(q + 9, outw, 128 * ts.w, lane);
}

// ---------------------------------------------------------------------------
// launcher
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
    if (a.nvwords) QA_LAUNCH(k_compact, dim3(nb(a.nvwords, 16)), dim3(256), 0, s, a);
    QA_LAUNCH(k_region_stats, dim3(a.nregions), dim3(1024), 0, s, a);
    QA_LAUNCH(k_plan, dim3(a.ncols), dim3(256), 0, s, a);
    QA_LAUNCH(k_grown_feats, dim3(a.nregions), dim3(256), 0, s, a);
    qa_fast_sizes_launch(a, s);
    qa_fast_final_launch(a, s);
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
    qa_fast_emit_launch(a, s);
}
