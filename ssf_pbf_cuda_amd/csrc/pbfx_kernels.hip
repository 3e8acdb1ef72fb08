// pbfx_kernels.hip -- HIP kernels of the PBF solver step for gfx950 (CDNA4).
//
// One step = bookkeeping kernels (predict+hash+histogram, two-level scan,
// scatter, stable reorder), then:
//   k_ns          neighbour lists (frozen, capped), per work unit of cells
//                 whose 27-cell neighbourhood is staged in LDS
//   k_lambda      lambda over the lists, one thread per particle    (x I)
//   k_deltap      delta-p + tensile correction + boundary response  (x I)
//   k_cur_*       regrouping by final cell                (vort_on or xsph_on)
//   k_final_list  in-range pairs around each final cell, per work unit
//   k_vort        vorticity over the final lists                     (vort_on)
//   k_confine     vorticity confinement                              (vort_on)
//   k_xsph_out    XSPH + AoS write-back                              (xsph_on)
//   k_out         AoS write-back                                     (otherwise)
// All kernels are HBM/LDS-bound integer or FP32 vector work; nothing is a
// contraction, so MFMA is not used.  Arithmetic follows the reference's
// operation order with IEEE division/sqrt and explicit fmaf (the library is
// compiled with -ffp-contract=off), so results are reproducible bit-for-bit
// on the host.
#include "pbfx_device.hpp"

#include <climits>

namespace pbfx {
namespace {

// ------------------------------------------------- K1: predict + hash + histogram
// pbf_predict.cu:15-45 (v* = fmaf(g,dt,v), x* = fmaf(v*,dt,x)), calcHashCUDA
// (pbf_neighbor_search_util.cu:20-37), checkDeleteParticleCUDA
// (pbf_delete.cu:9-28, sentinel key CT) and the histogram of insertAndCountCUDA
// (pbf_counting_sort.cu:21-32) with wave-aggregated atomics; each particle
// marks its 4096-cell scan tile as occupied (a plain store, no contention).
// Input i belongs to the scene whose input range holds it.  A batch takes
// its input count from the device (in_off[nscenes], at most n, the count the
// grid was sized for), so that a recorded graph of steps stays valid while
// deletion shrinks the scenes.
template <bool kB>
__global__ __launch_bounds__(256) void k_predict_hash(const float* __restrict__ x, const float* __restrict__ v,
                                                      uint32_t n, Bufs b, Scenes<kB> SC, uint32_t* __restrict__ key,
                                                      uint32_t* __restrict__ rank, uint32_t* __restrict__ count,
                                                      uint32_t* __restrict__ tflag, uint32_t lgT) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = i < (kB ? min(n, b.in_off[b.nscenes]) : n);
    uint32_t k = 0;
    const uint32_t si = active ? scene_index<kB>(b.in_off, b.nscenes, i) : 0u;
    by_scene<kB>(si, active, [&](uint32_t sc) {
        const Scene& S = SC[sc];
        const float dt = S.P.dt;
        float q[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float vs = __builtin_fmaf(S.P.g[c], dt, v[3 * i + c]);
            q[c] = __builtin_fmaf(vs, dt, x[3 * i + c]);
        }
        const Grid G = S.G;
        int gp[3];
        grid_pos(mk3(q[0], q[1], q[2]), G.cw, gp);
        k = scene_key(gp[0], gp[1], gp[2], G);
        if (q[0] < S.lo[0] || S.hi[0] < q[0] || q[1] < S.lo[1] || S.hi[1] < q[1] || q[2] < S.lo[2] ||
            S.hi[2] < q[2])
            k = b.CT;
    });
    const uint32_t r = wave_rank(count, k, active, k);
    if (active) {
        mark_tile(tflag, k >> lgT);
        key[i] = k;
        rank[i] = r;
        if (kB) b.isid[i] = si;
    }
}

// Where the counter of key c is stored.  lgnc = 0: at c.  Otherwise keys are
// (run << lgnc | copy) and copy j of every run lives in region j of C >> lgnc
// words, so that the copies of one run sit on different cache lines: a run
// that gathers particles from the whole domain (a wall or corner pile) is
// counted through nc separate lines instead of one.
__device__ __forceinline__ uint32_t count_slot(uint32_t c, uint32_t C, uint32_t lgnc) {
    return (c & ((1u << lgnc) - 1u)) * (C >> lgnc) + (c >> lgnc);
}

// ------------------------------------------------- K2a: per-tile totals
// Totals of the cell counts of each occupied scan tile (unoccupied tiles are
// all zero); tile nT is the deleted bucket (cell C).
// Scan workgroups have kScanBlock threads; a tile of at most kScanTile cells
// gives each thread at most kScanPer of them.
constexpr uint32_t kScanBlock = 256;
constexpr uint32_t kScanPer = kScanTile / kScanBlock;

__global__ __launch_bounds__(256) void k_tile_reduce(const uint32_t* __restrict__ count,
                                                     const uint32_t* __restrict__ tflag, uint32_t* __restrict__ tsum,
                                                     uint32_t nT, uint32_t lgT, uint32_t C, uint32_t lgnc) {
    __shared__ uint32_t wsum[kWaves];
    const uint32_t t = blockIdx.x;
    if (t == nT) {
        if (threadIdx.x == 0) tsum[nT] = count[C];
        return;
    }
    if (tflag[t] == 0u) {
        if (threadIdx.x == 0) tsum[t] = 0u;
        return;
    }
    const uint32_t T = 1u << lgT, c0 = t << lgT;
    // all of a thread's loads are issued before the first add
    uint32_t v[kScanPer];
#pragma unroll
    for (uint32_t k = 0; k < kScanPer; ++k) {
        const uint32_t j = threadIdx.x + k * kScanBlock;
        v[k] = j < T ? count[count_slot(c0 + j, C, lgnc)] : 0u;
    }
    uint32_t s = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanPer; ++k) s += v[k];
    uint32_t total;
    block_excl_scan(s, wsum, total);
    if (threadIdx.x == 0) tsum[t] = total;
}

// ------------------------------------------------- K2: scan of the tile totals
// One workgroup: exclusive scan of tsum[0..nT] (nT real tiles + the deleted
// bucket), and reset of the step counters.
__global__ __launch_bounds__(1024) void k_scan_tiles(const uint32_t* __restrict__ tsum,
                                                     uint32_t* __restrict__ tprefix, uint32_t nT1,
                                                     uint32_t* __restrict__ ctr, uint32_t reset) {
    __shared__ uint32_t wsum[16];
    const uint32_t per = (nT1 + blockDim.x - 1) / blockDim.x;
    const uint32_t b0 = threadIdx.x * per;
    uint32_t s = 0;
    for (uint32_t j = 0; j < per; ++j)
        if (b0 + j < nT1) s += tsum[b0 + j];
    const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
    uint32_t xs = s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(xs, d, 64);
        if (lane >= (uint32_t)d) xs += y;
    }
    if (lane == 63) wsum[wid] = xs;
    __syncthreads();
    uint32_t off = 0;
    for (uint32_t w = 0; w < wid; ++w) off += wsum[w];
    uint32_t run = off + xs - s;
    for (uint32_t j = 0; j < per; ++j)
        if (b0 + j < nT1) {
            tprefix[b0 + j] = run;
            run += tsum[b0 + j];
        }
    if (reset)
        for (uint32_t t = threadIdx.x; t < kCtrWords; t += blockDim.x) ctr[t] = 0;
}

// ------------------------------------------------- K3: scan of the cell counts
// Workgroup t scans the counts of cell tile t (empty tiles only write their
// prefix), clears the counts it read and the tile's occupied flag; the last
// workgroup writes the deleted bucket: off[C] = live count, off[C+1] = n.
// (CUB DeviceScan::ExclusiveSum in pbf_counting_sort.cu:62-71.)  When
// `stat` is given, the occupied tiles are counted into its statistics
// copies (CTR_SCANTILES).
__global__ __launch_bounds__(256) void k_scan_cells(uint32_t* __restrict__ count, uint32_t* __restrict__ offset,
                                                    const uint32_t* __restrict__ tsum, uint32_t* __restrict__ tflag,
                                                    const uint32_t* __restrict__ tprefix, uint32_t nT,
                                                    uint32_t lgT, uint32_t C, uint32_t lgnc,
                                                    uint32_t* __restrict__ stat) {
    __shared__ uint32_t wsum[kWaves];
    const uint32_t t = blockIdx.x;
    if (t == nT) {
        if (threadIdx.x == 0) {
            offset[C] = tprefix[nT];
            offset[C + 1] = tprefix[nT] + tsum[nT];
            count[C] = 0;
            tflag[nT] = 0;
        }
        return;
    }
    const uint32_t T = 1u << lgT;
    const uint32_t c0 = t << lgT;
    const uint32_t tot = tsum[t];
    const uint32_t base = tprefix[t];
    if (tot == 0) {
        for (uint32_t i = threadIdx.x; i < T; i += blockDim.x) offset[c0 + i] = base;
        return;
    }
    if (stat && threadIdx.x == 0) atomicAdd(&stat[CTR_NUM * (1u + (t & (kStatRep - 1u))) + CTR_SCANTILES], 1u);
    // The tile passes through LDS so that global reads and writes are
    // coalesced; thread t scans cells t*per .. t*per+per-1 there.  Word w is
    // kept at w + w/16, so the threads' 16-word strides hit distinct banks.
    __shared__ uint32_t cell[kScanTile + kScanTile / 16];
    auto at = [](uint32_t w) { return w + (w >> 4); };
    {
        uint32_t v[kScanPer];
#pragma unroll
        for (uint32_t k = 0; k < kScanPer; ++k) {
            const uint32_t j = threadIdx.x + k * kScanBlock;
            v[k] = j < T ? count[count_slot(c0 + j, C, lgnc)] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < kScanPer; ++k) {
            const uint32_t j = threadIdx.x + k * kScanBlock;
            if (j < T) cell[at(j)] = v[k];
        }
    }
    __syncthreads();
    const uint32_t per = (T + blockDim.x - 1) / blockDim.x;
    const uint32_t my0 = threadIdx.x * per;
    const uint32_t myn = (my0 < T) ? min(per, T - my0) : 0u;
    uint32_t s = 0;
    for (uint32_t j = 0; j < myn; ++j) s += cell[at(my0 + j)];
    uint32_t total;
    uint32_t run = base + block_excl_scan(s, wsum, total);
    for (uint32_t j = 0; j < myn; ++j) {
        const uint32_t c = cell[at(my0 + j)];
        cell[at(my0 + j)] = run;
        run += c;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < T; i += blockDim.x) {
        offset[c0 + i] = cell[at(i)];
        count[count_slot(c0 + i, C, lgnc)] = 0;
    }
    if (threadIdx.x == 0) tflag[t] = 0;
}

// ------------------------------------------------- K4: scatter
// sortByCountingCUDA (pbf_counting_sort.cu:120-134): slot = off[key] + rank.
// The rank-0 particle of a cell above kSmallCell particles lists the cell
// for k_sort_cells.
// n_dev (a batch's device-resident input count, or NULL) bounds the inputs.
__global__ __launch_bounds__(256) void k_scatter(const uint32_t* __restrict__ key, const uint32_t* __restrict__ rank,
                                                 const uint32_t* __restrict__ offset, uint32_t* __restrict__ tmp,
                                                 uint32_t n, const uint32_t* __restrict__ n_dev,
                                                 uint32_t* __restrict__ big, uint32_t* __restrict__ nbig,
                                                 uint32_t cap, uint32_t* __restrict__ bslot) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (n_dev ? min(n, *n_dev) : n)) return;
    const uint32_t k = key[i], r = rank[i], o0 = offset[k];
    tmp[o0 + r] = i;
    if (r == 0u) list_big_cell(offset, k, o0, big, nbig, cap, bslot);
}

// ------------------------------------------------- K4b: large cells in stable order
// Ascending bitonic sort of src[0..len) (len <= kSortMax, distinct values)
// into dst[0..len) through LDS, by the whole workgroup (dst may be src: all
// of src is read before the first write).  Contains barriers: call uniformly.
__device__ void lds_sort_segment(const uint32_t* src, uint32_t* dst, uint32_t len, uint32_t* __restrict__ buf) {
    uint32_t P = 1u;
    while (P < len) P <<= 1;
    for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) buf[i] = i < len ? src[i] : 0xFFFFFFFFu;
    __syncthreads();
    for (uint32_t k = 2u; k <= P; k <<= 1)
        for (uint32_t j = k >> 1; j > 0u; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const uint32_t a = buf[i], bv = buf[l];
                    if ((a > bv) == ((i & k) == 0u)) {
                        buf[i] = bv;
                        buf[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) dst[i] = buf[i];
    __syncthreads();
}

// The segment of each listed cell (all indices distinct) in ascending order,
// so that the stable in-cell order is the segment's own order.  Cells of at
// most kSortMax (cells[0 .. ncells[0])) are sorted in place by one workgroup
// each; the segment of a larger cell (cells[huge_base(cap) ..] + ncells[1])
// is cut into chunks of kSortMax, each sorted by one workgroup into the same
// range of `scratch`, for k_merge_huge.
__global__ __launch_bounds__(1024) void k_sort_cells(const uint32_t* __restrict__ offset,
                                                     const uint32_t* __restrict__ cells,
                                                     const uint32_t* __restrict__ ncells, uint32_t* __restrict__ vals,
                                                     uint32_t* __restrict__ scratch, uint32_t cap) {
    __shared__ uint32_t buf[kSortMax];
    const uint32_t nc = ncells[0];
    for (uint32_t w = blockIdx.x; w < nc; w += gridDim.x) {
        const uint32_t c = cells[w];
        const uint32_t o0 = offset[c];
        lds_sort_segment(vals + o0, vals + o0, offset[c + 1] - o0, buf);
    }
    const uint32_t nh = ncells[1];
    const uint32_t* __restrict__ huge = cells + huge_base(cap);
    uint32_t w = blockIdx.x;  // chunk index over all huge cells, in list order
    for (uint32_t h = 0; h < nh; ++h) {
        const uint32_t c = huge[h];
        const uint32_t o0 = offset[c], len = offset[c + 1] - o0;
        const uint32_t nch = (len + kSortMax - 1u) / kSortMax;
        for (; w < nch; w += gridDim.x) {
            const uint32_t q0 = w * kSortMax;
            lds_sort_segment(vals + o0 + q0, scratch + o0 + q0, min(kSortMax, len - q0), buf);
        }
        w -= nch;
    }
}

// Merges the sorted chunks of every huge cell: the element at position p of
// chunk q lands at p + (the number of smaller elements in every other
// chunk), found by binary search.  Reads scratch, writes vals.
__global__ __launch_bounds__(256) void k_merge_huge(const uint32_t* __restrict__ offset,
                                                    const uint32_t* __restrict__ cells,
                                                    const uint32_t* __restrict__ ncells, uint32_t* __restrict__ vals,
                                                    const uint32_t* __restrict__ scratch, uint32_t cap) {
    const uint32_t nh = ncells[1];
    const uint32_t* __restrict__ huge = cells + huge_base(cap);
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t h = 0; h < nh; ++h) {
        const uint32_t c = huge[h];
        const uint32_t o0 = offset[c], len = offset[c + 1] - o0;
        const uint32_t* __restrict__ seg = scratch + o0;
        const uint32_t nch = (len + kSortMax - 1u) / kSortMax;
        for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < len; p += stride) {
            const uint32_t x = seg[p], q = p / kSortMax;
            uint32_t r = p - q * kSortMax;
            for (uint32_t o = 0; o < nch; ++o) {
                if (o == q) continue;
                uint32_t lo = o * kSortMax, hi = min(lo + kSortMax, len);  // first element >= x
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (seg[mid] < x) lo = mid + 1u;
                    else hi = mid;
                }
                r += lo - o * kSortMax;
            }
            vals[o0 + r] = x;
        }
    }
}

// Work items of one sort: the particle at the first slot f of a run of `uw`
// cells along x (cells kb .. kb+uw-1) appends the run's items.  A run whose
// estimated 27-cell neighbourhood (9 rows of W+2 cells at the run's mean
// density) exceeds `budget` particles is split into narrower units, and a
// unit with more than kChunk own particles into several items that share its
// neighbourhood (item = (first cell, W | chunk << 6)).
// Called by all 64 lanes of the wave (k = kNoKey for lanes with nothing to
// emit): the lanes' items are counted, scanned across the wave and placed
// with one atomic per wave on the item counter, which every run of every
// wave would otherwise hit in turn.
constexpr uint32_t kNoKey = 0xFFFFFFFFu;
__device__ __forceinline__ void emit_units(const uint32_t* __restrict__ offset, uint32_t k, uint32_t f, uint32_t uw,
                                           uint32_t budget, uint2* __restrict__ units,
                                           uint32_t* __restrict__ nunits) {
    const uint32_t kb = (k != kNoKey ? k : 0u) & ~(uw - 1u);
    const bool head = k != kNoKey && f == offset[kb];
    uint32_t W = uw, total = 0;
    if (head) {
        const uint32_t nseg = offset[kb + uw] - offset[kb];
        while (W > 1u && 9ull * (W + 2u) * nseg > (unsigned long long)budget * uw) W >>= 1;
        for (uint32_t q0 = 0; q0 < uw; q0 += W) {
            const uint32_t own = offset[kb + q0 + W] - offset[kb + q0];
            total += (own + kChunk - 1u) / kChunk;
        }
    }
    const uint32_t lane = lane_id();
    uint32_t incl = total;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)incl, d, 64);
        if (lane >= (uint32_t)d) incl += y;
    }
    const uint32_t wave_total = (uint32_t)__shfl((int)incl, 63, 64);
    uint32_t base = 0;
    if (lane == 0 && wave_total) base = atomicAdd(nunits, wave_total);
    base = (uint32_t)__shfl((int)base, 0, 64);
    if (head && total) {
        uint32_t u = base + incl - total;
        for (uint32_t q0 = 0; q0 < uw; q0 += W) {
            const uint32_t kq = kb + q0;
            const uint32_t nch = (offset[kq + W] - offset[kq] + kChunk - 1u) / kChunk;
            for (uint32_t c = 0; c < nch; ++c) units[u++] = make_uint2(kq, W | (c << 6));
        }
    }
}

// ------------------------------------------------- K5: stable reorder
// Makes the in-cell order ascending by original index (the stable order of
// the reference's radix-sort paths, pbf_neighbor_search.cpp:112-131), gathers
// x* (as float4) and x (SoA) (reorderDataCUDA, pbf_neighbor_search_util.cu:143-154)
// and emits the work units: each run of up to `uw` cells along x is split into
// narrower units where it is dense, so that a unit's 27-cell neighbourhood
// fits the on-chip tile.
template <bool kB>
__global__ __launch_bounds__(256) void k_reorder(const float* __restrict__ x, const float* __restrict__ v, Bufs b, Scenes<kB> SC,
                                                 const uint32_t* __restrict__ key, const uint32_t* __restrict__ tmp,
                                                 uint32_t n, uint32_t uw, uint32_t budget) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = s < (kB ? min(n, b.in_off[b.nscenes]) : n);
    const uint32_t i = active ? tmp[s] : 0u;
    const uint32_t si = (kB && active) ? b.isid[i] : 0u;
    uint32_t ek = kNoKey, ef = 0;
    by_scene<kB>(si, active, [&](uint32_t sc) {
        const uint32_t k = key[i];
        const uint32_t o0 = b.offset[k], o1 = b.offset[k + 1];
        uint32_t f = s;  // a sorted cell (k_sort_cells, k_merge_huge): the rank is the position
        if (!sorted_cell(o1 - o0)) {
            uint32_t r = 0;
            for (uint32_t t = o0; t < o1; ++t) r += (tmp[t] < i) ? 1u : 0u;
            f = o0 + r;
        }
        b.sidx[f] = i;
        b.rank[f] = 0u;  // the histogram ranks are dead: the slot's big-cell group mark (k_big_seg)
        if (kB) b.sid[f] = sc;
        const Phys& P = SC[sc].P;
        const float dt = P.dt;
        // one 12-byte load each for the gathered x and v
        const F3 xi = reinterpret_cast<const F3*>(x)[i], vi = reinterpret_cast<const F3*>(v)[i];
        const float xo[3] = {xi.x, xi.y, xi.z}, vv[3] = {vi.x, vi.y, vi.z};
        float q[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float vs = __builtin_fmaf(P.g[c], dt, vv[c]);
            q[c] = __builtin_fmaf(vs, dt, xo[c]);
        }
        b.Pf[f] = make_float4(q[0], q[1], q[2], 0.f);
        st3(b.X, b.cap, f, mk3(xo[0], xo[1], xo[2]));
        if (k < b.CT) {
            ek = k;
            ef = f;
        }
    });
    emit_units(b.offset, ek, ef, uw, budget, b.units, &b.ctr[CTR_UNITS]);  // all lanes
}

// ------------------------------------------------- K5b: groups of big cells
// Image of a position in the wrapped grid, as a group code:
// (ix+2) | (iy+2) << 2 | (iz+2) << 4 for the images ix = floor(cx / gx) of
// the unwrapped cell coordinate cx (the hash's, grid_pos) in [-2, 1] on each
// axis; kCodeOther for non-finite positions and farther images.  The code
// only groups particles; the boxes decide every skip.
__device__ __forceinline__ uint32_t image_code(F3 p, const Grid& G) {
    if (!(__builtin_isfinite(p.x) && __builtin_isfinite(p.y) && __builtin_isfinite(p.z))) return kCodeOther;
    int gp[3];
    grid_pos(p, G.cw, gp);
    const int ix = gp[0] >> (int)G.lgx, iy = gp[1] >> (int)G.lgy, iz = gp[2] >> (31 - __builtin_clz(G.gz));
    if (ix < -2 || ix > 1 || iy < -2 || iy > 1 || iz < -2 || iz > 1) return kCodeOther;
    return (uint32_t)(ix + 2) | ((uint32_t)(iy + 2) << 2) | ((uint32_t)(iz + 2) << 4);
}

// Running box of positions; `bad` once a non-finite one was seen (its group
// is then never skipped).
struct Box3 {
    F3 lo, hi;
    bool bad;
};
__device__ __forceinline__ Box3 box_empty() {
    const float inf = __builtin_huge_valf();
    return Box3{mk3(inf, inf, inf), mk3(-inf, -inf, -inf), false};
}
__device__ __forceinline__ void box_add(Box3& B, F3 p) {
    if (!(__builtin_isfinite(p.x) && __builtin_isfinite(p.y) && __builtin_isfinite(p.z))) B.bad = true;
    B.lo = mk3(fminf(B.lo.x, p.x), fminf(B.lo.y, p.y), fminf(B.lo.z, p.z));
    B.hi = mk3(fmaxf(B.hi.x, p.x), fmaxf(B.hi.y, p.y), fmaxf(B.hi.z, p.z));
}

// The box over the whole block (every thread calls it); valid in thread 0.
__device__ Box3 block_box(Box3 B, float (*red)[8]) {
    const uint32_t lane = lane_id(), wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        B.lo = mk3(fminf(B.lo.x, __shfl_xor(B.lo.x, d, 64)), fminf(B.lo.y, __shfl_xor(B.lo.y, d, 64)),
                   fminf(B.lo.z, __shfl_xor(B.lo.z, d, 64)));
        B.hi = mk3(fmaxf(B.hi.x, __shfl_xor(B.hi.x, d, 64)), fmaxf(B.hi.y, __shfl_xor(B.hi.y, d, 64)),
                   fmaxf(B.hi.z, __shfl_xor(B.hi.z, d, 64)));
    }
    B.bad = __any(B.bad);
    if (lane == 0) {
        red[wid][0] = B.lo.x;
        red[wid][1] = B.lo.y;
        red[wid][2] = B.lo.z;
        red[wid][3] = B.hi.x;
        red[wid][4] = B.hi.y;
        red[wid][5] = B.hi.z;
        red[wid][6] = B.bad ? 1.f : 0.f;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (uint32_t w = 1; w < nw; ++w) {
            B.lo = mk3(fminf(B.lo.x, red[w][0]), fminf(B.lo.y, red[w][1]), fminf(B.lo.z, red[w][2]));
            B.hi = mk3(fmaxf(B.hi.x, red[w][3]), fmaxf(B.hi.y, red[w][4]), fmaxf(B.hi.z, red[w][5]));
            B.bad = B.bad || red[w][6] != 0.f;
        }
    __syncthreads();
    return B;
}

// The box as stored: a group with a non-finite member spans everything.
__device__ __forceinline__ void box_store(const Box3& B, float4& lo, float4& hi) {
    const float inf = __builtin_huge_valf();
    lo = B.bad ? make_float4(-inf, -inf, -inf, 0.f) : make_float4(B.lo.x, B.lo.y, B.lo.z, 0.f);
    hi = B.bad ? make_float4(inf, inf, inf, 0.f) : make_float4(B.hi.x, B.hi.y, B.hi.z, 0.f);
}

// Order-preserving key of a float (a < b  =>  fkey(a) <= fkey(b), -0 below
// +0), so that boxes gather with integer atomics; funkey inverts it.
__device__ __forceinline__ uint32_t fkey(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float funkey(uint32_t k) { return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k); }

// Group mark of a big cell's member: descriptor d, group k (0 marks the
// slots of other cells).
__device__ __forceinline__ uint32_t big_mark(uint32_t d, uint32_t k) { return d * kMaxSeg + k + 1u; }

// Adds the final x* of the lanes' marked slots to their groups' final boxes.
// Lanes of one group merge over the wave first and one of them issues the
// atomics.  Called by every lane of the wave (bm = 0 where nothing is added).
__device__ void big_box_add(BigSeg* __restrict__ bseg, uint32_t bm, F3 p) {
    const uint32_t lane = lane_id();
    const bool bad = !(__builtin_isfinite(p.x) && __builtin_isfinite(p.y) && __builtin_isfinite(p.z));
    const uint32_t k[3] = {fkey(p.x), fkey(p.y), fkey(p.z)};
    uint64_t todo = __ballot(bm != 0u);
    while (todo) {
        const int l = __ffsll((unsigned long long)todo) - 1;
        const uint32_t key = (uint32_t)__shfl((int)bm, l, 64);
        const bool mine = bm == key;
        uint32_t lo[3], hi[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            lo[c] = mine ? k[c] : 0xFFFFFFFFu;
            hi[c] = mine ? k[c] : 0u;
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                lo[c] = min(lo[c], (uint32_t)__shfl_xor((int)lo[c], d, 64));
                hi[c] = max(hi[c], (uint32_t)__shfl_xor((int)hi[c], d, 64));
            }
        const bool anybad = __any(mine && bad);
        if (lane == (uint32_t)l) {
            BigSeg& D = bseg[(key - 1u) / kMaxSeg];
            const uint32_t g = (key - 1u) % kMaxSeg;
            atomicMin(&D.fmin[g].x, lo[0]);
            atomicMin(&D.fmin[g].y, lo[1]);
            atomicMin(&D.fmin[g].z, lo[2]);
            atomicMax(&D.fmax[g].x, hi[0]);
            atomicMax(&D.fmax[g].y, hi[1]);
            atomicMax(&D.fmax[g].z, hi[2]);
            if (anybad) atomicOr(&D.fmin[g].w, 1u);
        }
        todo &= ~__ballot(mine);
    }
}

// Big cell of work item `it` (huge cells first: they take longest): its
// descriptor index d and key c.
__device__ __forceinline__ uint32_t big_item(const Bufs& b, uint32_t it, uint32_t nh, uint32_t& c) {
    const uint32_t d = it < nh ? huge_base(b.cap) + it : it - nh;
    c = b.big[d];
    return d;
}

// Groups the slots of every big cell by image code: a stable partition into
// aux (ascending slots inside each group, the groups in code order), the
// groups' ends and the boxes of their x* (after k_reorder, before the
// search).  One workgroup per cell; a cell of one group keeps aux unused.
template <bool kB>
__global__ __launch_bounds__(1024) void k_big_seg(Bufs b, Scenes<kB> SC) {
    __shared__ uint32_t cnt[kCodes];
    __shared__ uint32_t segof[kCodes];       // group of each code (kMaxSeg: none)
    __shared__ uint32_t sbeg[kMaxSeg + 1];   // first aux position of each group, then the cell's size
    __shared__ uint32_t put[kMaxSeg];        // where each group's next chunk goes
    __shared__ uint32_t wcnt[16][kMaxSeg];   // members of each group per wave in a chunk
    __shared__ uint32_t nseg_s;
    __shared__ float red[16][8];
    const uint32_t lane = lane_id(), wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint32_t nb0 = b.ctr[CTR_BIGCELLS], nh = b.ctr[CTR_HUGECELLS];
    for (uint32_t it = blockIdx.x; it < nh + nb0; it += gridDim.x) {
        uint32_t c;
        const uint32_t d = big_item(b, it, nh, c);
        const uint32_t o0 = b.offset[c], n = b.offset[c + 1] - o0;
        const Grid G = SC[scene_index<kB>(b.kbase, b.nscenes, c)].G;
        for (uint32_t k = threadIdx.x; k < kCodes; k += blockDim.x) cnt[k] = 0u;
        __syncthreads();
        for (uint32_t p = threadIdx.x; p < n; p += blockDim.x) atomicAdd(&cnt[image_code(f3of(b.Pf[o0 + p]), G)], 1u);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t ns = 0, run = 0;
            for (uint32_t k = 0; k < kCodes; ++k) {
                segof[k] = kMaxSeg;
                if (cnt[k]) {
                    if (ns < kMaxSeg) {
                        segof[k] = ns;
                        sbeg[ns] = run;
                        put[ns] = run;
                    }
                    ++ns;
                    run += cnt[k];
                }
            }
            nseg_s = ns <= kMaxSeg ? ns : 0u;
            if (ns <= kMaxSeg) sbeg[ns] = n;
        }
        __syncthreads();
        const uint32_t ns = nseg_s;
        if (ns >= 2u) {
            for (uint32_t q0 = 0; q0 < n; q0 += blockDim.x) {
                const uint32_t p = q0 + threadIdx.x;
                const uint32_t sg = p < n ? segof[image_code(f3of(b.Pf[o0 + p]), G)] : kMaxSeg;
                uint32_t lrank = 0;
                for (uint32_t k = 0; k < ns; ++k) {
                    const uint64_t m = __ballot(sg == k);
                    if (sg == k) lrank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                    if (lane == 0) wcnt[wid][k] = (uint32_t)__popcll(m);
                }
                __syncthreads();
                if (sg < kMaxSeg) {
                    uint32_t o = put[sg] + lrank;
                    for (uint32_t w = 0; w < wid; ++w) o += wcnt[w][sg];
                    b.aux[o0 + o] = o0 + p;
                    b.rank[o0 + p] = big_mark(d, sg);
                }
                __syncthreads();
                if (threadIdx.x < ns) {
                    uint32_t t = 0;
                    for (uint32_t w = 0; w < nw; ++w) t += wcnt[w][threadIdx.x];
                    put[threadIdx.x] += t;
                }
                __syncthreads();
            }
        }
        BigSeg& D = b.bseg[d];
        if (ns == 1u)
            for (uint32_t p = threadIdx.x; p < n; p += blockDim.x) b.rank[o0 + p] = big_mark(d, 0u);
        for (uint32_t k = 0; k < ns; ++k) {
            const uint32_t p0 = sbeg[k], p1 = sbeg[k + 1];
            Box3 B = box_empty();
            for (uint32_t p = p0 + threadIdx.x; p < p1; p += blockDim.x)
                box_add(B, f3of(b.Pf[ns == 1u ? o0 + p : b.aux[o0 + p]]));
            B = block_box(B, red);
            if (threadIdx.x == 0) {
                box_store(B, D.slo[k], D.shi[k]);
                D.end[k] = p1;
                // the final box, gathered by k_cur_hash
                D.fmin[k] = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u);
                D.fmax[k] = make_uint4(0u, 0u, 0u, 0u);
            }
        }
        if (threadIdx.x == 0) D.nseg = ns;
        __syncthreads();
    }
}

// What a scan of big cell D (n particles) for a particle at `self` must
// visit.  kFinal: the final positions, whose pair test passes
// !(glm_dot(d, d) > thr); otherwise the search's, which passes
// fmaf-dot(d, d) < thr.  Every component of a member's difference to self
// lies, in magnitude, between the rounded gap from self to the group's box
// and the rounded distance to the box's far side (rounding is monotone), and
// both dot products are monotone in the components' magnitudes: a group
// whose gap fails the test holds no pair, and one whose far corner passes it
// holds only pairs.  A non-finite self skips nothing.
struct BigPart {
    uint32_t how;     // 0: no member passes; 1: visit the whole cell; 2: visit aux positions [p0, p1) only
    uint32_t p0, p1;  // the one group that may pass (how 2, or how 1 with a single group: the whole cell)
    bool all_in;      // every member of that group passes
    uint32_t inside;  // members of the cell's groups whose every member passes
};
__device__ __forceinline__ float box_gap(float x, float lo, float hi) {
    return (x < lo) ? (lo - x) : ((x > hi) ? (x - hi) : 0.f);
}
__device__ __forceinline__ float box_far(float x, float lo, float hi) { return fmaxf(x - lo, hi - x); }
template <bool kFinal>
__device__ __forceinline__ bool pair_test(F3 d, float thr) {
    return kFinal ? !(glm_dot(d, d) > thr) : (__builtin_fmaf(d.z, d.z, __builtin_fmaf(d.y, d.y, d.x * d.x)) < thr);
}
template <bool kFinal>
__device__ __forceinline__ BigPart big_part(const BigSeg* __restrict__ D, F3 self, float thr, uint32_t n) {
    BigPart r{1u, 0u, n, false, 0u};
    const uint32_t ns = D->nseg;
    if (ns == 0u) return r;
    uint32_t hits = 0, beg = 0, p0 = 0, p1 = 0;
    bool in0 = false;
    for (uint32_t k = 0; k < ns; ++k) {
        const uint32_t end = D->end[k];
        float4 lo, hi;
        if (kFinal) {
            const uint4 a = D->fmin[k], z = D->fmax[k];
            const float inf = __builtin_huge_valf();
            lo = a.w ? make_float4(-inf, -inf, -inf, 0.f) : make_float4(funkey(a.x), funkey(a.y), funkey(a.z), 0.f);
            hi = a.w ? make_float4(inf, inf, inf, 0.f) : make_float4(funkey(z.x), funkey(z.y), funkey(z.z), 0.f);
        } else {
            lo = D->slo[k];
            hi = D->shi[k];
        }
        const F3 g = mk3(box_gap(self.x, lo.x, hi.x), box_gap(self.y, lo.y, hi.y), box_gap(self.z, lo.z, hi.z));
        const F3 f = mk3(box_far(self.x, lo.x, hi.x), box_far(self.y, lo.y, hi.y), box_far(self.z, lo.z, hi.z));
        const bool in = pair_test<kFinal>(f, thr);
        if (pair_test<kFinal>(g, thr)) {
            if (hits == 0u) {
                p0 = beg;
                p1 = end;
                in0 = in;
            }
            ++hits;
        }
        if (in) r.inside += end - beg;
        beg = end;
    }
    if (hits == 0u) {
        r.how = 0u;
    } else if (hits == 1u) {
        r.how = ns == 1u ? 1u : 2u;
        r.p0 = p0;
        r.p1 = p1;
        r.all_in = in0;
    }
    return r;
}

// What the global-memory scans read of a step's cells: the cell offsets and
// the big cells' descriptors (passed by value, so that no kernel argument
// block is copied to scratch memory for its address).
struct CellView {
    const uint32_t* __restrict__ offset;
    const uint32_t* __restrict__ bslot;
    const uint32_t* __restrict__ aux;
    const BigSeg* __restrict__ bseg;
};
__device__ __forceinline__ CellView cell_view(const Bufs& b) { return CellView{b.offset, b.bslot, b.aux, b.bseg}; }

// Whether this step has cells of more than kSmallCell particles (a
// grid-uniform value: the scans without them keep their plain loops).
__device__ __forceinline__ bool big_cells(const Bufs& b) { return (b.ctr[CTR_BIGCELLS] | b.ctr[CTR_HUGECELLS]) != 0u; }

// a[i] by selects (i in 0..3, not known at compile time): no array in
// scratch memory
__device__ __forceinline__ uint32_t pick4(const uint32_t (&a)[4], uint32_t i) {
    const uint32_t lo = (i & 1u) ? a[1] : a[0], hi = (i & 1u) ? a[3] : a[2];
    return (i & 2u) ? hi : lo;
}

// Whether visiting one group through aux beats scanning the whole cell (the
// indirection costs a load per candidate): when it skips a quarter or more.
__device__ __forceinline__ bool aux_pays(const BigPart& bp, uint32_t n) {
    return bp.how == 2u && 4u * (bp.p1 - bp.p0) <= 3u * n;
}


// ------------------------------------------------- work units
struct UnitShared {
    uint32_t tstart[kMaxCells + 1];  // tile start of each neighbourhood cell (row-major 9 x (W+2))
    uint32_t gstart[kMaxCells];      // sorted slot of each neighbourhood cell's first particle
    uint32_t cnt[kMaxCells];
    uint32_t run_e[kMaxCells + 1];   // first cell of each run of cells contiguous in memory
    uint32_t wsum[kWaves];
};

struct UnitInfo {
    uint32_t W, ncell, xb, y, z, total, own_t0, own_g0, n_own, nruns;
    bool staged;  // the neighbourhood fits the LDS tile; otherwise it is read from global memory
};

// Per-thread event tallies, summed per wave and added once at kernel end.
struct Tally {
    uint32_t trunc = 0, slowpart = 0, slowunits = 0, maxtile = 0, sweepcap = 0, maxsweep = 0,
             nan = 0, listovf = 0;
};

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = max(v, (uint32_t)__shfl_xor(v, d, 64));
    return v;
}

// Called by every thread of the block at kernel end.
// Statistics go to one of kStatRep copies of the counters (by workgroup), so
// that waves of a large grid do not all update the same words; the host
// reduces the copies.
__device__ void flush_tally(uint32_t* __restrict__ ctr_base, const Tally& t) {
    uint32_t* ctr = ctr_base + CTR_NUM * (1u + (blockIdx.x & (kStatRep - 1u)));
    const uint32_t tr = wave_sum(t.trunc), sp = wave_sum(t.slowpart), su = wave_sum(t.slowunits),
                   mt = wave_max(t.maxtile), sc = wave_sum(t.sweepcap),
                   ms = wave_max(t.maxsweep), nn = wave_sum(t.nan), lo = wave_sum(t.listovf);
    if (lane_id() == 0) {
        if (tr) atomicAdd(&ctr[CTR_TRUNC], tr);
        if (sp) atomicAdd(&ctr[CTR_SLOWPART], sp);
        if (su) atomicAdd(&ctr[CTR_SLOWUNITS], su);
        if (mt && mt > ctr[CTR_MAXTILE]) atomicMax(&ctr[CTR_MAXTILE], mt);  // a stale read only costs an atomic
        if (sc) atomicAdd(&ctr[CTR_SWEEPCAP], sc);
        if (ms && ms > ctr[CTR_MAXSWEEP]) atomicMax(&ctr[CTR_MAXSWEEP], ms);
        if (nn) atomicAdd(&ctr[CTR_XSPHNAN], nn);
        if (lo) atomicAdd(&ctr[CTR_LISTOVF], lo);
    }
}

// XCD-aware split: workgroup b runs on XCD b % 8; each XCD takes a contiguous
// share of the (roughly key-ordered) unit list so neighbouring units share an L2.
__device__ __forceinline__ void unit_range(uint32_t nU, uint32_t& u0, uint32_t& u1, uint32_t& step) {
    const uint32_t nx = 8, b = blockIdx.x, Gd = gridDim.x;
    if (Gd < nx) {
        u0 = b;
        u1 = nU;
        step = Gd;
        return;
    }
    const uint32_t xcd = b % nx, lb = b / nx;
    const uint32_t nbx = (Gd - xcd + nx - 1) / nx;
    const uint32_t chunk = (nU + nx - 1) / nx;
    u0 = xcd * chunk + lb;
    u1 = min(nU, (xcd + 1) * chunk);
    step = nbx;
}

// Builds the neighbourhood table of a unit: 9 rows (dz outer, dy inner -- the
// reference's z,y order) of W+2 cells (x0-1 .. x0+W, wrapped), their tile
// starts and sorted slots, and the runs of cells whose particles are
// contiguous in memory (one per row unless the row wraps).  Thread t owns
// cells 2t and 2t+1 (up to 2 * kBlock cells).  Contains barriers: call
// uniformly.
static_assert(kMaxCells <= 2u * kBlock, "unit_setup gives each thread two neighbourhood cells");
__device__ UnitInfo unit_setup(const Grid& G, const uint32_t* __restrict__ offset, uint2 unit, uint32_t tile_cap,
                               UnitShared& S, Tally& tl) {
    UnitInfo I;
    I.W = unit.y & 63u;
    const uint32_t chunk = unit.y >> 6;
    I.ncell = 9u * (I.W + 2u);
    const uint32_t kb = unit.x - G.base;  // key within the scene's grid
    I.xb = kb & (G.gx - 1u);
    I.y = (kb >> G.lgx) & (G.gy - 1u);
    I.z = kb >> (G.lgx + G.lgy);
    const uint32_t e0 = 2u * threadIdx.x;
    uint32_t cnt[2] = {0u, 0u};
#pragma unroll
    for (uint32_t h = 0; h < 2u; ++h) {
        const uint32_t e = e0 + h;
        if (e < I.ncell) {
            const uint32_t r = e / (I.W + 2u), kk = e - r * (I.W + 2u);
            const uint32_t cx = (I.xb + kk - 1u) & (G.gx - 1u);
            const uint32_t cy = (I.y + (r % 3u) - 1u) & (G.gy - 1u);
            const uint32_t cz = (I.z + (r / 3u) - 1u) & (G.gz - 1u);
            const uint32_t key = G.base + (cz * G.gy + cy) * G.gx + cx;
            const uint32_t gs = offset[key];
            cnt[h] = offset[key + 1] - gs;
            S.gstart[e] = gs;
            S.cnt[e] = cnt[h];
        }
    }
    uint32_t total;
    const uint32_t ex = block_excl_scan(cnt[0] + cnt[1], S.wsum, total);
    if (e0 < I.ncell) S.tstart[e0] = ex;
    if (e0 + 1u < I.ncell) S.tstart[e0 + 1u] = ex + cnt[0];
    if (threadIdx.x == 0) S.tstart[I.ncell] = total;
    uint32_t flag[2] = {0u, 0u};
#pragma unroll
    for (uint32_t h = 0; h < 2u; ++h) {
        const uint32_t e = e0 + h;
        if (e < I.ncell) flag[h] = (e == 0 || S.gstart[e] != S.gstart[e - 1] + S.cnt[e - 1]) ? 1u : 0u;
    }
    uint32_t nruns;
    const uint32_t ri = block_excl_scan(flag[0] + flag[1], S.wsum, nruns);
    if (flag[0]) S.run_e[ri] = e0;
    if (flag[1]) S.run_e[ri + flag[0]] = e0 + 1u;
    if (threadIdx.x == 0) S.run_e[nruns] = I.ncell;
    __syncthreads();
    const uint32_t base = 4u * (I.W + 2u);
    I.total = total;
    I.nruns = nruns;
    const uint32_t skip = chunk * kChunk;
    I.own_t0 = S.tstart[base + 1] + skip;
    I.own_g0 = S.gstart[base + 1] + skip;
    I.n_own = min(kChunk, S.tstart[base + I.W + 1] - I.own_t0);
    I.staged = total <= tile_cap;
    if (threadIdx.x == 0) {
        tl.maxtile = max(tl.maxtile, total);
        if (!I.staged) tl.slowunits++;
    }
    return I;
}

// Copies the unit's neighbourhood into the tile, one run per wave at a time
// (coalesced): put(tile index, sorted slot).
template <class F>
__device__ __forceinline__ void load_runs(const UnitShared& S, const UnitInfo& I, F&& put) {
    const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
    for (uint32_t ri = wid; ri < I.nruns; ri += kWaves) {
        const uint32_t e0 = S.run_e[ri], e1 = S.run_e[ri + 1];
        const uint32_t t0 = S.tstart[e0], len = S.tstart[e1] - t0, g0 = S.gstart[e0];
#pragma unroll 4
        for (uint32_t i = lane; i < len; i += 64) put(t0 + i, g0 + i);
    }
}

// cell column (1..W) of an own particle's tile index
__device__ __forceinline__ uint32_t own_column(const UnitShared& S, uint32_t W, uint32_t t) {
    const uint32_t base = 4u * (W + 2u);
    uint32_t lo = base + 1u, hi = base + W;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (S.tstart[mid] <= t) lo = mid;
        else hi = mid - 1u;
    }
    return lo - base;
}

// ------------------------------------------------- K6: neighbour lists
// detectNeighborsCUDA (pbf_neighbor_search_util.cu:361-482): per particle the
// 27 cells in z,y,x order, slots ascending, keep r2 < h^2 while fewer than
// cap are kept; the list is frozen for all Jacobi iterations.  Stored as
// sorted slots in groups of four (list_at), with the count in
// ncnt[slot].  Candidates come from the unit's staged tile, four at a time.

// Appends v to the queue of long list builds counted by ctr[which], a
// wave-aggregated append to crank (dead from the final passes to the next
// regrouping: the search's queue is read before it, the final lists' queue
// after its placement).  A slot or position is queued at most once per step,
// so a queue never exceeds cap.
constexpr uint32_t kDeferred = 0xFFFFFFFFu;  // a filter's count when it left its list to a wave
__device__ __forceinline__ void queue_push(const Bufs& b, uint32_t which, uint32_t v) {
    const uint64_t m = __ballot(true);
    const uint32_t lane = lane_id(), leader = (uint32_t)(__ffsll((unsigned long long)m) - 1);
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&b.ctr[which], (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, (int)leader, 64);
    const uint32_t k = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (k < b.cap) b.crank[k] = v;
}

// The filter of one own particle (column kc of the unit, slot g): candidate
// tile index t is read at src[sb + t] -- the tile itself (kStaged) or the
// sorted positions, with sb the cell's slot base.  Returns the number kept
// (at most ncap); trunc is set when an (ncap+1)-th pair exists.  Each group
// of four candidates gives a 4-bit hit mask whose pairs are stored in order.
template <bool kStaged, class E>
__device__ __forceinline__ uint32_t ns_filter(const UnitShared& S, uint32_t W, uint32_t kc, F3 self, float h2,
                                              uint32_t ncap, const E* __restrict__ src,
                                              uint32_t* __restrict__ nl, uint32_t cap, uint32_t g, bool& trunc) {
    uint32_t* __restrict__ dst = nl + list_at(cap, g, 0u);  // where entry cnt goes
    const size_t jump = (size_t)4u * cap - 3u;              // from entry 4q+3 to 4q+4
    uint32_t cnt = 0;
    for (uint32_t c = 0; c < 27u; ++c) {
        const uint32_t e = (c / 3u) * (W + 2u) + kc - 1u + (c % 3u);
        const uint32_t lo = S.tstart[e], hi = S.tstart[e + 1];
        const uint32_t gb = S.gstart[e] - lo;  // slot = gb + tile index
        const uint32_t sb = kStaged ? 0u : gb;
        for (uint32_t t = lo; t < hi; t += 4u) {
            F3 pj[4];
#pragma unroll
            for (uint32_t i = 0; i < 4u; ++i)  // the tile is padded: no clamp
                pj[i] = f3of(kStaged ? src[t + i] : src[sb + min(t + i, hi - 1u)]);
            uint32_t m = 0;
#pragma unroll
            for (uint32_t i = 0; i < 4u; ++i) {
                const F3 d = self - pj[i];
                const float r2 = __builtin_fmaf(d.z, d.z, __builtin_fmaf(d.y, d.y, d.x * d.x));
                m |= (r2 < h2) ? (1u << i) : 0u;
            }
            m &= (hi - t >= 4u) ? 0xFu : ((1u << (hi - t)) - 1u);
            // keep the first ncap - cnt hits of the group; any beyond them
            // truncate the list
            const uint32_t room = ncap - cnt;
            bool full = false;
            while ((uint32_t)__builtin_popcount(m) > room) {
                m &= ~(0x80000000u >> __builtin_clz(m));
                full = true;
            }
            while (m) {
                const uint32_t i = (uint32_t)__builtin_ctz(m);
                m &= m - 1u;
                *dst = gb + t + i;
                ++cnt;
                dst += ((cnt & 3u) == 0u) ? jump : 1u;
            }
            if (full) {
                trunc = true;
                return cnt;
            }
        }
    }
    return cnt;
}

// The same filter from global memory, for units whose neighbourhood exceeds
// the tile: candidate x*_j = Pf[j] by slot.  In a big cell, groups out of
// range are skipped, a group wholly in range is listed without tests, and a
// single group left is scanned through aux when that skips a quarter of the
// cell or more.  Every scan is one loop whose slots come, per lane, from aux
// or from the cell's own run, the next group's slots in flight.  Returns
// kDeferred, leaving the list unfinished, before a range that would take the
// candidates tested past `budget` (only a neighbourhood holding a cell of
// more than kSmallCell particles can: 27 * kSmallCell < kHeavyScan).
__device__ __forceinline__ uint32_t ns_filter_global(const UnitShared& S, uint32_t W, uint32_t kc, F3 self,
                                                     float h2, uint32_t ncap, const float4* __restrict__ Pf,
                                                     CellView b, uint32_t* __restrict__ nl, uint32_t cap,
                                                     uint32_t g, uint32_t budget, bool& trunc) {
    uint32_t* __restrict__ dst = nl + list_at(cap, g, 0u);  // where entry cnt goes
    const size_t jump = (size_t)4u * cap - 3u;              // from entry 4q+3 to 4q+4
    uint32_t cnt = 0;
    for (uint32_t c = 0; c < 27u; ++c) {
        const uint32_t e = (c / 3u) * (W + 2u) + kc - 1u + (c % 3u);
        const uint32_t o0 = S.gstart[e], ncell = S.tstart[e + 1] - S.tstart[e];
        uint32_t p0 = 0, p1 = ncell;
        bool via = false;
        if (ncell > kSmallCell) {
            const BigPart bp = big_part<false>(b.bseg + b.bslot[o0], self, h2, ncell);
            if (bp.how == 0u) continue;
            if (bp.all_in) {
                // every member of the group is a pair: its first slots, ascending
                const uint32_t room = ncap - cnt, size = bp.p1 - bp.p0, take = min(room, size);
                for (uint32_t q = 0; q < take; ++q) {
                    *dst = bp.how == 2u ? b.aux[o0 + bp.p0 + q] : o0 + bp.p0 + q;
                    ++cnt;
                    dst += ((cnt & 3u) == 0u) ? jump : 1u;
                }
                if (size > room) {
                    trunc = true;
                    return cnt;
                }
                continue;
            }
            if (aux_pays(bp, ncell)) {
                p0 = bp.p0;
                p1 = bp.p1;
                via = true;
            }
        }
        if (p1 == p0) continue;
        if (p1 - p0 > budget) return kDeferred;
        budget -= p1 - p0;
        const uint32_t* __restrict__ ax = b.aux + o0;
        uint32_t jn[4];
#pragma unroll
        for (uint32_t i = 0; i < 4u; ++i) jn[i] = via ? ax[min(p0 + i, p1 - 1u)] : o0 + min(p0 + i, p1 - 1u);
        for (uint32_t p = p0; p < p1; p += 4u) {
            const uint32_t j[4] = {jn[0], jn[1], jn[2], jn[3]};
            F3 pj[4];
#pragma unroll
            for (uint32_t i = 0; i < 4u; ++i) pj[i] = f3of(Pf[j[i]]);
#pragma unroll
            for (uint32_t i = 0; i < 4u; ++i)
                jn[i] = via ? ax[min(p + 4u + i, p1 - 1u)] : o0 + min(p + 4u + i, p1 - 1u);
            uint32_t m = 0;
#pragma unroll
            for (uint32_t i = 0; i < 4u; ++i) m |= pair_test<false>(self - pj[i], h2) ? (1u << i) : 0u;
            m &= (p1 - p >= 4u) ? 0xFu : ((1u << (p1 - p)) - 1u);
            // keep the first ncap - cnt hits of the group; any beyond them
            // truncate the list
            const uint32_t room = ncap - cnt;
            bool full = false;
            while ((uint32_t)__builtin_popcount(m) > room) {
                m &= ~(0x80000000u >> __builtin_clz(m));
                full = true;
            }
            while (m) {
                const uint32_t i = (uint32_t)__builtin_ctz(m);
                m &= m - 1u;
                *dst = pick4(j, i);
                ++cnt;
                dst += ((cnt & 3u) == 0u) ? jump : 1u;
            }
            if (full) {
                trunc = true;
                return cnt;
            }
        }
    }
    return cnt;
}

template <bool kB>
__global__ __launch_bounds__(kBlock) void k_ns(Bufs b, Scenes<kB> SC, uint32_t tile_cap) {
    __shared__ UnitShared S;
    extern __shared__ float4 smem[];
    F3* tile = reinterpret_cast<F3*>(smem);  // x*, 12 bytes a candidate
    const float4* __restrict__ Pf = b.Pf;
    uint32_t* __restrict__ nl = b.nlist;
    const uint32_t cap = b.cap;
    const uint32_t nU = b.ctr[CTR_UNITS];
    uint32_t u0, u1, us;
    unit_range(nU, u0, u1, us);
    Tally tl;
    for (uint32_t u = u0; u < u1; u += us) {
        const uint2 unit = b.units[u];
        const uint32_t sc = scene_index<kB>(b.kbase, b.nscenes, unit.x);
        const Grid G = SC[sc].G;
        const float h2 = SC[sc].P.h2;
        const uint32_t ncap = (uint32_t)SC[sc].P.cap;
        const UnitInfo I = unit_setup(G, b.offset, unit, tile_cap, S, tl);
        // a neighbourhood larger than the tile (any size) is filtered from global memory
        if (I.staged) load_runs(S, I, [&](uint32_t t, uint32_t g) { tile[t] = f3of(Pf[g]); });
        __syncthreads();
        for (uint32_t q = threadIdx.x; q < I.n_own; q += kBlock) {
            const uint32_t ts = I.own_t0 + q, g = I.own_g0 + q;
            const uint32_t kc = own_column(S, I.W, ts);
            bool trunc = false;
            const uint32_t cnt =
                I.staged ? ns_filter<true>(S, I.W, kc, tile[ts], h2, ncap, tile, nl, cap, g, trunc)
                         : ns_filter_global(S, I.W, kc, f3of(Pf[g]), h2, ncap, Pf, cell_view(b), nl, cap, g,
                                            kHeavyScan, trunc);
            if (cnt == kDeferred) {  // a long scan: k_ns_heavy builds this list with a whole wave
                queue_push(b, CTR_NSHEAVY, g);
            } else {
                b.ncnt[g] = cnt;
                tl.trunc += trunc ? 1u : 0u;
            }
        }
        __syncthreads();
    }
    flush_tally(b.ctr, tl);
}

// The search list of one queued slot g (x* = self), built by a whole wave:
// the 27 cells around its cell in z,y,x order with the same ranges and group
// skips as ns_filter_global, 64 candidates at a time, lane l testing the
// l-th, the hits appended in slot order through a ballot; the first ncap
// are kept and trunc is set when there are more.  Called by all lanes of the
// wave; every value it depends on is the same on every lane.
template <bool kBig>
__device__ uint32_t ns_heavy_list(const Grid& G, CellView b, const float4* __restrict__ Pf, F3 self, float h2,
                                  uint32_t ncap, uint32_t* __restrict__ nl, uint32_t cap, uint32_t g, bool& trunc) {
    const uint32_t lane = lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    int gp[3];
    grid_pos(self, G.cw, gp);
    uint32_t n = 0;
    for (int c = 0; c < 27; ++c) {
        const int r = c / 3, dx = c % 3;
        const uint32_t key = scene_key(gp[0] + dx - 1, gp[1] + (r % 3) - 1, gp[2] + (r / 3) - 1, G);
        const uint32_t o0 = b.offset[key], ncell = b.offset[key + 1] - o0;
        uint32_t p0 = 0, p1 = ncell;
        bool via = false;
        if (kBig && ncell > kSmallCell) {
            const BigPart bp = big_part<false>(b.bseg + b.bslot[o0], self, h2, ncell);
            if (bp.how == 0u) continue;
            if (bp.all_in) {  // every member is a pair: the first ones, ascending
                const uint32_t size = bp.p1 - bp.p0, take = min(ncap - n, size);
                for (uint32_t t = lane; t < take; t += 64u)
                    nl[list_at(cap, g, n + t)] = bp.how == 2u ? b.aux[o0 + bp.p0 + t] : o0 + bp.p0 + t;
                n += take;
                if (size > take) {
                    trunc = true;
                    return n;
                }
                continue;
            }
            if (aux_pays(bp, ncell)) {
                p0 = bp.p0;
                p1 = bp.p1;
                via = true;
            }
        }
        for (uint32_t p = p0; p < p1; p += 64u) {
            const uint32_t t = p + lane;
            uint32_t j = 0;
            bool hit = false;
            if (t < p1) {
                j = via ? b.aux[o0 + t] : o0 + t;
                hit = pair_test<false>(self - f3of(Pf[j]), h2);
            }
            const uint64_t m = __ballot(hit);
            const uint32_t k = n + (uint32_t)__popcll(m & below);
            if (hit && k < ncap) nl[list_at(cap, g, k)] = j;
            n += (uint32_t)__popcll(m);
            if (n > ncap) {
                trunc = true;
                return ncap;
            }
        }
    }
    return n;
}

// The search lists k_ns queued (global-memory filters that would test more
// than kHeavyScan candidates): one wave per slot, then its count and
// truncation tally as k_ns would have written them.
template <bool kB>
__global__ __launch_bounds__(1024) void k_ns_heavy(Bufs b, Scenes<kB> SC) {
    const uint32_t nh = min(b.ctr[CTR_NSHEAVY], b.cap);
    if (nh == 0u) return;  // uniform over the grid
    const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t w0 = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const CellView cv = cell_view(b);
    const bool any_big = big_cells(b);
    Tally tl;
    for (uint32_t h = w0; h < nh; h += waves) {
        const uint32_t g = __builtin_amdgcn_readfirstlane(b.crank[h]);
        const uint32_t sc = kB ? __builtin_amdgcn_readfirstlane(b.sid[g]) : 0u;
        const Grid G = SC[sc].G;
        const float h2 = SC[sc].P.h2;
        const uint32_t ncap = (uint32_t)SC[sc].P.cap;
        const F3 self = f3of(b.Pf[g]);
        bool trunc = false;
        const uint32_t n = any_big ? ns_heavy_list<true>(G, cv, b.Pf, self, h2, ncap, b.nlist, b.cap, g, trunc)
                                   : ns_heavy_list<false>(G, cv, b.Pf, self, h2, ncap, b.nlist, b.cap, g, trunc);
        if (lane_id() == 0u) {
            b.ncnt[g] = n;
            tl.trunc += trunc ? 1u : 0u;
        }
    }
    flush_tally(b.ctr, tl);
}

// ------------------------------------------------- K7: lambda
// calcScalingFactorCUDA, memo form (pbf_constraint.cu:168-209): over the
// frozen list, s = C / (m^2 inv_rho0^2 G + eps) with the reference's fma
// order; writes (x*, s).  One thread per slot.
template <bool kB>
__global__ __launch_bounds__(256) void k_lambda(Bufs b, Scenes<kB> SC) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    const bool active = i < b.offset[b.CT];
    by_scene<kB>((kB && active) ? b.sid[i] : 0u, active, [&](uint32_t sc) {
        const Phys P = SC[sc].P;
        const float4* __restrict__ Pf = b.Pf;
        const uint32_t* __restrict__ nl = b.nlist;
        const uint32_t cap = b.cap;
        const F3 self = f3of(Pf[i]);
        const uint32_t n = b.ncnt[i];
        float ks = 0.f, kg2s = 0.f;
        F3 kgs = mk3(0.f, 0.f, 0.f);
        walk_list(nl, cap, i, n, Pf, [&](uint32_t, float4 pj) {
            float W;
            F3 kg;
            pair_terms(self, f3of(pj), P, W, kg);
            const float kg2 = fdot(kg, kg);
            ks += W;
            acc(kgs, kg);
            kg2s += kg2;
        });
        const float Cc = __builtin_fmaf(ks * P.m, P.inv_rho0, -1.f);
        const float Gs =
            __builtin_fmaf(kgs.x, kgs.x, __builtin_fmaf(kgs.z, kgs.z, __builtin_fmaf(kgs.y, kgs.y, kg2s)));
        b.Q[i] = make_float4(self.x, self.y, self.z, Cc / __builtin_fmaf(P.m2r2, Gs, P.relax));
    });
}

// ------------------------------------------------- K8: delta-p + collision + apply
// calcPositionUpdateCUDA (memo form, pbf_position_update.cu:131-188),
// responseCollisionCUDA (pbf_response_collision.cu:12-66) and
// updateInterimPositionCUDA (pbf_constraint.cu:335-345).  Reads (x*, s),
// writes the new x*.  One thread per slot.  kPow4: every scene's s_corr
// exponent is 4, so powi unrolls to its two squarings at compile time.
template <bool kB, bool kPow4>
__global__ __launch_bounds__(256) void k_deltap(Bufs b, Scenes<kB> SC) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    const bool active = i < b.offset[b.CT];
    Tally tl;
    by_scene<kB>((kB && active) ? b.sid[i] : 0u, active, [&](uint32_t sc) {
        const Phys P = SC[sc].P;
        const Boundary B = SC[sc].B;
        const float4* __restrict__ Q = b.Q;
        const uint32_t* __restrict__ nl = b.nlist;
        const uint32_t cap = b.cap;
        const float4 se = Q[i];
        const F3 self = f3of(se);
        const float si = se.w;
        const float self_cl = (si < 0.f) ? -si : 0.f;
        const uint32_t n = b.ncnt[i];
        F3 sum = mk3(0.f, 0.f, 0.f);
        walk_list(nl, cap, i, n, Q, [&](uint32_t, float4 pe) {
            float W;
            F3 kg;
            pair_terms(self, f3of(pe), P, W, kg);
            const float sj = pe.w;
            float corr;
            if (P.mode == 0) {
                const float kr2 = W * W;
                const float kr4 = kr2 * kr2;
                const float tc = 10.f * kr4 * P.inv_k;
                const float sup = (sj < 0.f) ? -sj : 0.f;
                const float cl = self_cl + sup;
                corr = cl * tc;
            } else {
                corr = P.sk * powi(W * P.inv_wdq, kPow4 ? 4 : P.sn);
            }
            const float a = -(si + sj + corr);
            acc(sum, mulsv(a, kg));
        });
        F3 dp = mk3(sum.x * P.m * P.inv_rho0, sum.y * P.m * P.inv_rho0, sum.z * P.m * P.inv_rho0);
        bool capped = false;
        const int sw = (B.n_planes == 6 && B.n_inner == 0 && B.n_outer == 0)
                           ? collide_planes<6>(P, B, dp, self, capped)
                           : collide(P, B, dp, self, capped);
        const F3 np = self + dp;
        b.Pf[i] = make_float4(np.x, np.y, np.z, 0.f);
        tl.sweepcap = capped ? 1u : 0u;
        tl.maxsweep = (uint32_t)sw;
    });
    flush_tally(b.ctr, tl);
}

// ------------------------------------------------- final-cell regrouping
// The vorticity, confinement and XSPH sums visit the 27 cells around each
// particle's FINAL position, bucketed by the search's cell ranges
// (pbf/interaction/cuda/detail/pbf_contribution.inl:7-57).  Particles are
// regrouped by run of uw final cells along x (the work unit of the final
// passes; the order inside a run is irrelevant, every sum is per particle):
// a counting sort over C/uw runs.  Each run has nc = min(8, uw) counters,
// one per group of workgroups (b mod nc, i.e. per XCD), each on its own cache
// line (count_slot), so that a run that collects particles from all over the
// domain -- a wall or corner pile of ~20k particles on the 1M dam -- is not
// counted through one contended line; key = run * nc + copy.  The same
// pass packs each slot's record R = (x*, 0), (v, 0) with
// v = inv_dt*(x* - x_sorted) (updateVelocityCUDA, pbf_update.cu:25-39), so
// that a particle visited in run order costs one cache line.  Final cells
// need not follow slot order (a violent step scatters them), so equal keys
// are merged over the whole wave, not only over neighbouring lanes.
template <bool kB>
__global__ __launch_bounds__(256) void k_cur_hash(Bufs b, Scenes<kB> SC, uint32_t lguw, uint32_t lgnc, uint32_t lgT,
                                                  uint32_t CR) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = s < b.offset[b.CT];
    uint32_t k = 0;
    by_scene<kB>((kB && active) ? b.sid[s] : 0u, active, [&](uint32_t sc) {
        const Scene& S = SC[sc];
        const F3 p = f3of(b.Pf[s]);
        const F3 vv = mulsv(S.P.inv_dt, p - ld3(b.X, b.cap, s));
        b.R[2 * s] = make_float4(p.x, p.y, p.z, 0.f);
        b.R[2 * s + 1] = make_float4(vv.x, vv.y, vv.z, 0.f);
        const Grid G = S.G;
        int gp[3];
        grid_pos(p, G.cw, gp);
        k = ((scene_key(gp[0], gp[1], gp[2], G) >> lguw) << lgnc) | (blockIdx.x & ((1u << lgnc) - 1u));
    });
    // the final boxes of the big cells' groups (the marks k_big_seg left)
    const uint32_t bm = active ? b.rank[s] : 0u;
    if (__any(bm != 0u)) big_box_add(b.bseg, bm, active ? f3of(b.Pf[s]) : mk3(0.f, 0.f, 0.f));
    const uint32_t r = wave_rank_any(b.ccount, k, active, count_slot(k, CR, lgnc));
    if (active) {
        mark_tile(b.ctflag, k >> lgT);
        b.ckey[s] = k;
        b.crank[s] = r;
    }
}

// Places each particle at its run's position and emits the work items of
// the runs holding at least kSparse particles (one per kChunk of them);
// clears the overflow slots of the final-order blocks of 64.
__global__ __launch_bounds__(256) void k_cur_scatter(const uint32_t* __restrict__ offset, uint32_t CT,
                                                     const uint32_t* __restrict__ ckey,
                                                     const uint32_t* __restrict__ crank,
                                                     const uint32_t* __restrict__ coffset,
                                                     uint32_t* __restrict__ cperm, uint32_t uw, uint32_t lguw,
                                                     uint32_t lgnc, uint2* __restrict__ cunits,
                                                     uint32_t* __restrict__ ctr, uint32_t* __restrict__ oslot) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nl = offset[CT];
    if (s >= nl) return;
    if (s <= (nl - 1u) / 64u) oslot[s] = kNoSlot;  // the final passes' blocks of 64 positions
    const uint32_t k = ckey[s];
    const uint32_t slot = coffset[k] + crank[s];
    cperm[slot] = s;
    const uint32_t run = k >> lgnc;
    const uint32_t r0 = coffset[run << lgnc];
    if (slot == r0) {
        const uint32_t own = coffset[(run + 1u) << lgnc] - r0;
        if (own >= kSparse) {
            const uint32_t nch = (own + kChunk - 1u) / kChunk;
            const uint32_t u = atomicAdd(&ctr[CTR_CUNITS], nch);
            for (uint32_t c = 0; c < nch; ++c) cunits[u + c] = make_uint2(run << lguw, uw | (c << 6));
        }
    }
}

// The 27-cell neighbourhood of a particle's final cell, restated from
// calcContribution (pbf/interaction/cuda/detail/pbf_contribution.inl:7-57):
// cells in z,y,x order, slots ascending in each, per-cell partial sums added
// in that order.  Pairs with dot(d, d) > cut2 lie outside the kernel support
// and add exactly +-0 for finite states; they are skipped.
//
// The final passes run in two stages.  k_final_list filters each particle's
// candidates once and stores its in-range pairs (bit 31 marks the first pair
// of each cell); the vorticity, confinement and XSPH passes then walk those
// lists with every lane busy.  Adding only the cells that have pairs gives
// the same sums: the skipped cells add +0.

// Keeps the first `room` set bits of a group's hit mask; returns whether any
// were dropped.
__device__ __forceinline__ bool clip_hits(uint32_t& m, uint32_t room) {
    bool full = false;
    while ((uint32_t)__builtin_popcount(m) > room) {
        m &= ~(0x80000000u >> __builtin_clz(m));
        full = true;
    }
    return full;
}

// Emits the in-range tile elements of the staged neighbourhood around own
// column kc, in order: emit(slot, first pair of its cell), which counts them
// in n; stops once n reaches lim.
template <class EmitF>
__device__ __forceinline__ void list_tile(const UnitInfo& I, const UnitShared& S, const F3* __restrict__ tp,
                                          uint32_t kc, F3 self, float cut2, const uint32_t& n, uint32_t lim,
                                          EmitF&& emit) {
    for (uint32_t q = 0; q < 27u; ++q) {
        const uint32_t e = (q / 3u) * (I.W + 2u) + kc - 1u + (q % 3u);
        const uint32_t lo = S.tstart[e], hi = S.tstart[e + 1];
        const uint32_t gb = S.gstart[e] - lo;  // slot = gb + tile index
        bool first = true;
        for (uint32_t t = lo; t < hi; t += 4u) {
            F3 el[4];
#pragma unroll
            for (uint32_t i = 0; i < 4u; ++i) el[i] = tp[t + i];  // the tile is padded: no clamp
            uint32_t m = 0;
#pragma unroll
            for (uint32_t i = 0; i < 4u; ++i) {
                const F3 d = el[i] - self;
                m |= !(glm_dot(d, d) > cut2) ? (1u << i) : 0u;
            }
            m &= (hi - t >= 4u) ? 0xFu : ((1u << (hi - t)) - 1u);
            const bool full = clip_hits(m, lim - n);
            while (m) {
                const uint32_t i = (uint32_t)__builtin_ctz(m);
                m &= m - 1u;
                emit(gb + t + i, first);
                first = false;
            }
            if (full || n == lim) return;
        }
    }
}

// The same from global memory, row by row (dz outer, dy inner), the next
// row's three cell ranges loaded while the current row is filtered; pos(j)
// gives x*_j.  kBig (the step has big cells): a big cell's groups out of
// range are skipped by their final boxes, one wholly in range is listed
// without tests, and the scan stops at once when such groups alone take the
// list to lim.  Stops once n reaches lim.  kBig: returns false, leaving the
// list unfinished, before a cell that would take the candidates tested past
// `budget` (at least 27 * kSmallCell).
template <bool kBig, class PosF, class EmitF>
__device__ __forceinline__ bool list_global(const Grid& G, CellView b, F3 self, float cut2, uint32_t& n,
                                            uint32_t lim, uint32_t budget, PosF&& pos, EmitF&& emit) {
    const uint32_t* __restrict__ offset = b.offset;
    int gp[3];
    grid_pos(self, G.cw, gp);
    auto row = [&](int r, uint32_t* st, uint32_t* en) {
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
            const uint32_t c = scene_key(gp[0] + dx - 1, gp[1] + (r % 3) - 1, gp[2] + (r / 3) - 1, G);
            st[dx] = offset[c];
            en[dx] = offset[c + 1];
        }
    };
    uint32_t st[3], en[3];
    row(0, st, en);
    for (int r = 0; r < 9; ++r) {
        uint32_t ns[3] = {0, 0, 0}, ne[3] = {0, 0, 0};
        if (r < 8) row(r + 1, ns, ne);
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
            bool first = true;
            if (!kBig) {
                // a step without big cells: the plain scan (27 cells of at
                // most kSmallCell particles stay within any budget)
                for (uint32_t j = st[dx]; j < en[dx]; j += 4u) {
                    F3 pj[4];
#pragma unroll
                    for (uint32_t i = 0; i < 4u; ++i) pj[i] = pos(min(j + i, en[dx] - 1u));
                    uint32_t m = 0;
#pragma unroll
                    for (uint32_t i = 0; i < 4u; ++i) m |= pair_test<true>(pj[i] - self, cut2) ? (1u << i) : 0u;
                    m &= (en[dx] - j >= 4u) ? 0xFu : ((1u << (en[dx] - j)) - 1u);
                    const bool full = clip_hits(m, lim - n);
                    while (m) {
                        const uint32_t i = (uint32_t)__builtin_ctz(m);
                        m &= m - 1u;
                        emit(j + i, first);
                        first = false;
                    }
                    if (full || n == lim) return true;
                }
                continue;
            }
            const uint32_t o0 = st[dx], ncell = en[dx] - st[dx];
            uint32_t p0 = 0, p1 = ncell;
            bool via = false;
            if (ncell > kSmallCell) {
                const BigPart bp = big_part<true>(b.bseg + b.bslot[o0], self, cut2, ncell);
                if (bp.how == 0u) continue;
                if (n + bp.inside >= lim) {  // the list reaches lim here: no need to store it
                    n = lim;
                    return true;
                }
                if (bp.all_in) {  // fewer than lim - n members, all pairs
                    for (uint32_t q = bp.p0; q < bp.p1; ++q) {
                        emit(bp.how == 2u ? b.aux[o0 + q] : o0 + q, first);
                        first = false;
                    }
                    continue;
                }
                if (aux_pays(bp, ncell)) {
                    p0 = bp.p0;
                    p1 = bp.p1;
                    via = true;
                }
            }
            if (p1 == p0) continue;
            if (p1 - p0 > budget) return false;
            budget -= p1 - p0;
            const uint32_t* __restrict__ ax = b.aux + o0;
            uint32_t jn[4];
#pragma unroll
            for (uint32_t i = 0; i < 4u; ++i) jn[i] = via ? ax[min(p0 + i, p1 - 1u)] : o0 + min(p0 + i, p1 - 1u);
            for (uint32_t p = p0; p < p1; p += 4u) {
                const uint32_t jj[4] = {jn[0], jn[1], jn[2], jn[3]};
                F3 pj[4];
#pragma unroll
                for (uint32_t i = 0; i < 4u; ++i) pj[i] = pos(jj[i]);
#pragma unroll
                for (uint32_t i = 0; i < 4u; ++i)
                    jn[i] = via ? ax[min(p + 4u + i, p1 - 1u)] : o0 + min(p + 4u + i, p1 - 1u);
                uint32_t m = 0;
#pragma unroll
                for (uint32_t i = 0; i < 4u; ++i) m |= pair_test<true>(pj[i] - self, cut2) ? (1u << i) : 0u;
                m &= (p1 - p >= 4u) ? 0xFu : ((1u << (p1 - p)) - 1u);
                const bool full = clip_hits(m, lim - n);
                while (m) {
                    const uint32_t i = (uint32_t)__builtin_ctz(m);
                    m &= m - 1u;
                    emit(pick4(jj, i), first);
                    first = false;
                }
                if (full || n == lim) return true;
            }
        }
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
            st[dx] = ns[dx];
            en[dx] = ne[dx];
        }
    }
    return true;
}

// A candidate's records for a pass's term: p = (x*_j, pass-specific w),
// q = the second record the pass needs (the velocity, or unused).  Loaded
// together by the sums, so no load waits on another.
struct El2 {
    float4 p, q;
};

// Key of cell c (0..26, in z,y,x order) of the 27 around grid position gp.
__device__ __forceinline__ uint32_t cell27(const int gp[3], uint32_t c, const Grid& G) {
    const int dz = (int)(c / 9u) - 1, q = (int)(c % 9u);
    return scene_key(gp[0] + (q % 3) - 1, gp[1] + (q / 3) - 1, gp[2] + dz, G);
}

// Partial sum of one cell (key) in slot order: the per-cell sum of
// calcContribution; elem(j) builds the pass's element of slot j.  A big
// cell's groups out of range (final boxes) add nothing and are skipped; a
// single group left is summed through aux, in slot order.
template <bool kBig, class ElemF, class TermF>
__device__ __forceinline__ F3 sum_cell(CellView b, uint32_t key, F3 self, float cut2, ElemF&& elem,
                                       TermF&& term) {
    const uint32_t st = b.offset[key], n = b.offset[key + 1] - st;
    F3 cs = mk3(0.f, 0.f, 0.f);
    if (!kBig) {
        // four candidates' records loaded together, then added in order
        for (uint32_t j = st; j < st + n; j += 4u) {
            El2 el[4];
#pragma unroll
            for (uint32_t i = 0; i < 4u; ++i) el[i] = elem(min(j + i, st + n - 1u));
#pragma unroll
            for (uint32_t i = 0; i < 4u; ++i) {
                const F3 d = f3of(el[i].p) - self;
                if (j + i < st + n && !(glm_dot(d, d) > cut2)) acc(cs, term(el[i]));
            }
        }
        return cs;
    }
    uint32_t p0 = 0, p1 = n;
    bool via = false;
    if (n > kSmallCell) {
        const BigPart bp = big_part<true>(b.bseg + b.bslot[st], self, cut2, n);
        if (bp.how == 0u) return cs;
        if (aux_pays(bp, n)) {
            p0 = bp.p0;
            p1 = bp.p1;
            via = true;
        }
    }
    if (p1 == p0) return cs;
    const uint32_t* __restrict__ ax = b.aux + st;
    uint32_t jn[4];
#pragma unroll
    for (uint32_t i = 0; i < 4u; ++i) jn[i] = via ? ax[min(p0 + i, p1 - 1u)] : st + min(p0 + i, p1 - 1u);
    // four candidates' records loaded together, then added in order
    for (uint32_t p = p0; p < p1; p += 4u) {
        El2 el[4];
#pragma unroll
        for (uint32_t i = 0; i < 4u; ++i) el[i] = elem(jn[i]);
#pragma unroll
        for (uint32_t i = 0; i < 4u; ++i) jn[i] = via ? ax[min(p + 4u + i, p1 - 1u)] : st + min(p + 4u + i, p1 - 1u);
#pragma unroll
        for (uint32_t i = 0; i < 4u; ++i) {
            const F3 d = f3of(el[i].p) - self;
            if (p + i < p1 && !(glm_dot(d, d) > cut2)) acc(cs, term(el[i]));
        }
    }
    return cs;
}

// Direct sum over the 27 cells from global memory, for particles whose list
// overflowed: the cells' partial sums added in z,y,x order.
template <bool kBig, class ElemF, class TermF>
__device__ __forceinline__ F3 sum27_global(const Grid& G, CellView b, F3 self, float cut2, ElemF&& elem,
                                           TermF&& term) {
    int gp[3];
    grid_pos(self, G.cw, gp);
    F3 sum = mk3(0.f, 0.f, 0.f);
    for (uint32_t c = 0; c < 27u; ++c) acc(sum, sum_cell<kBig>(b, cell27(gp, c, G), self, cut2, elem, term));
    return sum;
}

// Particles of final-cell runs too sparse for a unit (fewer than kSparse
// particles in their run of uw cells): one thread each, in final-cell order
// so that the lanes of a wave share their neighbourhoods' cache lines, with
// a contiguous share of that order per XCD.  Sparse runs are contiguous
// stretches of the order.  body(q, g): position q in that order, slot g.
template <class F>
__device__ __forceinline__ void for_sparse(const Bufs& b, uint32_t lgnc, F&& body) {
    const uint32_t nl = b.offset[b.CT];
    uint32_t c0, c1, cs;
    unit_range((nl + kBlock - 1) / kBlock, c0, c1, cs);
    for (uint32_t c = c0; c < c1; c += cs) {
        const uint32_t q = c * kBlock + threadIdx.x;
        if (q >= nl) continue;
        const uint32_t g = b.cperm[q];
        const uint32_t run = b.ckey[g] >> lgnc;
        if (b.coffset[(run + 1u) << lgnc] - b.coffset[run << lgnc] < kSparse) body(q, g);
    }
}

// Own particles of a final-run work item: positions coffset[run] + chunk*kChunk .. of the run order.
struct OwnCur {
    uint32_t c0, n;
};
__device__ __forceinline__ OwnCur own_cur(const Bufs& b, uint2 unit, uint32_t lguw, uint32_t lgnc) {
    OwnCur o;
    const uint32_t run = unit.x >> lguw, skip = (unit.y >> 6) * kChunk;
    o.c0 = b.coffset[run << lgnc] + skip;
    o.n = min(kChunk, b.coffset[(run + 1u) << lgnc] - o.c0);
    return o;
}

// Own column (1..W) of a final-cell unit's particle at position p.
__device__ __forceinline__ uint32_t final_column(const Grid& G, const UnitInfo& I, F3 p) {
    int gp[3];
    grid_pos(p, G.cw, gp);
    return (((uint32_t)gp[0] & (G.gx - 1u)) - I.xb) + 1u;
}

// ------------------------------------------------- overflow slots
// A final-order position whose list overflowed keeps the count kFinalK + 1;
// the first such position of each block of 64 takes the block's overflow
// slot (oslot, one atomic per block; ovf lists the slots' blocks), so that
// the passes add its per-cell partials, computed for all 27 cells at once by
// k_ovf_part, instead of summing the cells one after another (which they do
// when no slot was left).  Called by the lanes whose list overflowed (any
// subset of the wave); a wave's positions are consecutive, so its lanes of
// one block are contiguous and only the first of them competes for the slot.
__device__ __forceinline__ uint32_t ovf_mark(const Bufs& b, uint32_t q) {
    const uint64_t m = __ballot(true);
    const uint32_t lane = lane_id(), blk = q >> 6;
    const uint64_t prev = m & ((1ull << lane) - 1ull);
    const uint32_t pb = prev ? (uint32_t)__shfl((int)blk, 63 - __builtin_clzll(prev), 64) : kNoSlot;
    if (pb != blk && atomicCAS(&b.oslot[blk], kNoSlot, kPendingSlot) == kNoSlot) {
        const uint32_t k = atomicAdd(&b.ctr[CTR_OVF], 1u);
        if (k < b.ovf_max) b.ovf[k] = blk;
        b.oslot[blk] = k < b.ovf_max ? k : kFullSlot;
    }
    return kFinalK + 1u;
}

// ------------------------------------------------- K9: final pair lists
// For each particle (by position q of the final-cell order) the slots of its
// in-range pairs in groups of four (list_at(cap, q, k), k < kFinalK); fcnt[q] =
// their count, or kFinalK + 1 when there are more (the passes then fall back
// to the direct sum, so the filter stops at the (kFinalK+1)-th pair: a
// particle in a pile of thousands does not scan its whole neighbourhood).
template <bool kB>
__global__ __launch_bounds__(kBlock) void k_final_list(Bufs b, Scenes<kB> SC, uint32_t tile_cap, uint32_t lguw, uint32_t lgnc) {
    static_assert(kChunk <= kBlock, "one own particle per thread");
    __shared__ UnitShared S;
    __shared__ uint32_t colpos[kMaxW + 2];  // own particles per final column, then each column's first position
    __shared__ uint32_t corder[kChunk];     // the item's own slots in final-column order
    extern __shared__ float4 smem[];
    F3* tp = reinterpret_cast<F3*>(smem);  // x*, 12 bytes a candidate
    // the final x*, packed: the same values as the passes' records R (built
    // from Pf), at half the bytes per candidate
    const float4* __restrict__ Pf = b.Pf;
    const uint32_t cap = b.cap;
    uint32_t* __restrict__ fl = b.flist;
    const CellView cv = cell_view(b);
    const bool any_big = big_cells(b);
    Tally tl;
    UnitInfo I;
    // inlined at both call sites: a call would copy the kernel arguments to scratch memory
    auto build = [&](uint32_t q, uint32_t g, bool tiled, const Grid& G, float cut2) __attribute__((always_inline)) {
        const F3 self = f3of(Pf[g]);
        uint32_t n = 0;
        uint32_t* __restrict__ dst = fl + list_at(cap, q, 0u);  // where entry n goes
        const size_t jump = (size_t)4u * cap - 3u;              // from entry 4q+3 to 4q+4
        // the (kFinalK+1)-th pair is counted, not stored: it marks the overflow
        auto emit = [&](uint32_t j, bool first) {
            if (n < kFinalK) {
                *dst = j | (first ? 0x80000000u : 0u);
                dst += (((n + 1u) & 3u) == 0u) ? jump : 1u;
            }
            ++n;
        };
        if (tiled) {
            list_tile(I, S, tp, final_column(G, I, self), self, cut2, n, kFinalK + 1u, emit);
        } else if (any_big) {
            if (!list_global<true>(G, cv, self, cut2, n, kFinalK + 1u, kHeavyScan,
                                   [&](uint32_t j) { return f3of(Pf[j]); }, emit)) {
                queue_push(b, CTR_HEAVY, q);  // a long scan: k_final_heavy builds this list with a whole wave
                return;
            }
        } else {
            list_global<false>(G, cv, self, cut2, n, kFinalK + 1u, kHeavyScan, [&](uint32_t j) { return f3of(Pf[j]); },
                               emit);
        }
        if (n > kFinalK) n = ovf_mark(b, q);
        b.fcnt[q] = n;
        tl.listovf += n > kFinalK ? 1u : 0u;
    };
    auto cut2_of = [&](uint32_t sc) { return fmaxf(SC[sc].P.cut2_w, SC[sc].P.cut2_wd); };
    const uint32_t nU = b.ctr[CTR_CUNITS];
    uint32_t u0, u1, us;
    unit_range(nU, u0, u1, us);
    for (uint32_t u = u0; u < u1; u += us) {
        const uint2 unit = b.cunits[u];
        const uint32_t sc = scene_index<kB>(b.kbase, b.nscenes, unit.x);
        const Grid G = SC[sc].G;
        const float cut2 = cut2_of(sc);
        for (uint32_t c = threadIdx.x; c < kMaxW + 2u; c += kBlock) colpos[c] = 0u;
        I = unit_setup(G, b.offset, unit, tile_cap, S, tl);  // barriers order the clearing above
        // The item's own particles are put in final-column order (a
        // counting sort in LDS), written back to cperm, so that the lanes
        // of a wave share their candidate cells here and their pairs in
        // the passes; the order of the particles changes no sum.
        const OwnCur o = own_cur(b, unit, lguw, lgnc);
        const bool mine = threadIdx.x < o.n;
        uint32_t kc = 0, rk = 0, gm = 0;
        if (mine) {
            gm = b.cperm[o.c0 + threadIdx.x];
            kc = min(final_column(G, I, f3of(Pf[gm])), I.W + 1u);
            rk = atomicAdd(&colpos[kc], 1u);
        }
        if (I.staged)
            load_runs(S, I, [&](uint32_t t, uint32_t j) { tp[t] = f3of(Pf[j]); });
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t run = 0;
            for (uint32_t c = 0; c <= I.W + 1u; ++c) {
                const uint32_t k = colpos[c];
                colpos[c] = run;
                run += k;
            }
        }
        __syncthreads();
        if (mine) corder[colpos[kc] + rk] = gm;
        __syncthreads();
        if (mine) {
            const uint32_t g = corder[threadIdx.x];
            b.cperm[o.c0 + threadIdx.x] = g;
            if (!I.staged) tl.slowpart++;
            build(o.c0 + threadIdx.x, g, I.staged, G, cut2);
        }
        __syncthreads();
    }
    for_sparse(b, lgnc, [&](uint32_t q, uint32_t g) {
        tl.slowpart++;
        by_scene<kB>(kB ? b.sid[g] : 0u, true, [&](uint32_t sc) { build(q, g, false, SC[sc].G, cut2_of(sc)); });
    });
    flush_tally(b.ctr, tl);
}

// The final list of one queued particle (x* = self, final-order position q),
// built by a whole wave: each cell's candidates (the same ranges and group
// skips as list_global) 64 at a time, lane l testing the l-th; the hits are
// appended in slot order through a ballot.  Returns the pair count, or lim
// once it reaches lim = kFinalK + 1 (the list then overflows and is not
// read).  Called by all lanes of the wave; every value it depends on is the
// same on every lane.
template <bool kBig>
__device__ uint32_t heavy_list(const Grid& G, CellView b, const float4* __restrict__ Pf, F3 self, float cut2,
                               uint32_t* __restrict__ fl, uint32_t cap, uint32_t q) {
    const uint32_t lim = kFinalK + 1u;
    const uint32_t lane = lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    int gp[3];
    grid_pos(self, G.cw, gp);
    uint32_t n = 0;
    for (int c = 0; c < 27; ++c) {
        const int r = c / 3, dx = c % 3;
        const uint32_t key = scene_key(gp[0] + dx - 1, gp[1] + (r % 3) - 1, gp[2] + (r / 3) - 1, G);
        const uint32_t o0 = b.offset[key], ncell = b.offset[key + 1] - o0;
        uint32_t p0 = 0, p1 = ncell;
        bool via = false;
        if (kBig && ncell > kSmallCell) {
            const BigPart bp = big_part<true>(b.bseg + b.bslot[o0], self, cut2, ncell);
            if (bp.how == 0u) continue;
            if (n + bp.inside >= lim) return lim;
            if (bp.all_in) {  // fewer than lim - n members, all pairs, ascending
                const uint32_t size = bp.p1 - bp.p0;
                for (uint32_t t = lane; t < size; t += 64u) {
                    const uint32_t j = bp.how == 2u ? b.aux[o0 + bp.p0 + t] : o0 + bp.p0 + t;
                    if (n + t < kFinalK) fl[list_at(cap, q, n + t)] = j | (t == 0u ? 0x80000000u : 0u);
                }
                n += size;
                continue;
            }
            if (aux_pays(bp, ncell)) {
                p0 = bp.p0;
                p1 = bp.p1;
                via = true;
            }
        }
        bool any = false;  // a pair of this cell was appended
        for (uint32_t p = p0; p < p1; p += 64u) {
            const uint32_t t = p + lane;
            uint32_t j = 0;
            bool hit = false;
            if (t < p1) {
                j = via ? b.aux[o0 + t] : o0 + t;
                hit = pair_test<true>(f3of(Pf[j]) - self, cut2);
            }
            const uint64_t m = __ballot(hit);
            const uint32_t k = n + (uint32_t)__popcll(m & below);
            if (hit && k < kFinalK) fl[list_at(cap, q, k)] = j | ((!any && (m & below) == 0ull) ? 0x80000000u : 0u);
            n += (uint32_t)__popcll(m);
            any = any || m != 0ull;
            if (n >= lim) return lim;
        }
    }
    return n;
}

// The final lists k_final_list queued (builds of more than kHeavyScan
// candidates): one wave per particle, then its count or
// overflow entry as k_final_list would have written them.
template <bool kB>
__global__ __launch_bounds__(1024) void k_final_heavy(Bufs b, Scenes<kB> SC) {
    const uint32_t nh = min(b.ctr[CTR_HEAVY], b.cap);
    if (nh == 0u) return;  // uniform over the grid
    const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t w0 = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const CellView cv = cell_view(b);
    const bool any_big = big_cells(b);
    Tally tl;
    for (uint32_t h = w0; h < nh; h += waves) {
        const uint32_t q = __builtin_amdgcn_readfirstlane(b.crank[h]);
        const uint32_t g = __builtin_amdgcn_readfirstlane(b.cperm[q]);
        const uint32_t sc = kB ? b.sid[g] : 0u;
        const Grid G = SC[sc].G;
        const float cut2 = fmaxf(SC[sc].P.cut2_w, SC[sc].P.cut2_wd);
        const F3 self = f3of(b.Pf[g]);
        const uint32_t n = any_big ? heavy_list<true>(G, cv, b.Pf, self, cut2, b.flist, b.cap, q)
                                   : heavy_list<false>(G, cv, b.Pf, self, cut2, b.flist, b.cap, q);
        if (lane_id() == 0u) {
            b.fcnt[q] = n > kFinalK ? ovf_mark(b, q) : n;
            tl.listovf += n > kFinalK ? 1u : 0u;
        }
    }
    flush_tally(b.ctr, tl);
}

// Sum over a particle's stored pair list, per-cell partial sums in order;
// elem(j) builds the pass's element of slot j.
template <class ElemF, class TermF>
__device__ __forceinline__ F3 sum_list(const uint32_t* __restrict__ fl, uint32_t cap, uint32_t q, uint32_t n,
                                       ElemF&& elem, TermF&& term) {
    F3 sum = mk3(0.f, 0.f, 0.f), cs = mk3(0.f, 0.f, 0.f);
    uint4 w = n ? list_word(fl, cap, q, 0u) : make_uint4(0u, 0u, 0u, 0u);
    for (uint32_t k = 0; k < n; k += 4u) {
        uint32_t e[4];
        El2 el[4];
        list_group(w, k, n, e);
#pragma unroll
        for (uint32_t i = 0; i < 4u; ++i) el[i] = elem(e[i] & 0x7FFFFFFFu);
        w = list_word(fl, cap, q, min(k + 4u, (n - 1u) & ~3u));  // next group, in flight during this one
#pragma unroll
        for (uint32_t i = 0; i < 4u; ++i)
            if (k + i < n) {
                if (e[i] >> 31) {
                    acc(sum, cs);
                    cs = mk3(0.f, 0.f, 0.f);
                }
                acc(cs, term(el[i]));
            }
    }
    acc(sum, cs);
    return sum;
}

// sum_list for an elem that takes any index: the entries past n in the last
// word are loaded as they are and not summed.
template <class ElemF, class TermF>
__device__ __forceinline__ F3 sum_list_any(const uint32_t* __restrict__ fl, uint32_t cap, uint32_t q, uint32_t n,
                                           ElemF&& elem, TermF&& term) {
    F3 sum = mk3(0.f, 0.f, 0.f), cs = mk3(0.f, 0.f, 0.f);
    uint4 w = n ? list_word(fl, cap, q, 0u) : make_uint4(0u, 0u, 0u, 0u);
    for (uint32_t k = 0; k < n; k += 4u) {
        const uint32_t e[4] = {w.x, w.y, w.z, w.w};
        El2 el[4];
#pragma unroll
        for (uint32_t i = 0; i < 4u; ++i) el[i] = elem(e[i] & 0x7FFFFFFFu);
        w = list_word(fl, cap, q, min(k + 4u, (n - 1u) & ~3u));  // next group, in flight during this one
#pragma unroll
        for (uint32_t i = 0; i < 4u; ++i)
            if (k + i < n) {
                if (e[i] >> 31) {
                    acc(sum, cs);
                    cs = mk3(0.f, 0.f, 0.f);
                }
                acc(cs, term(el[i]));
            }
    }
    acc(sum, cs);
    return sum;
}

// Visits position q of the final-cell order: body(q, slot, stored pair
// count, scene), grouped by scene (no wave-wide operations inside).
template <bool kB, class F>
__device__ __forceinline__ void for_final(const Bufs& b, F&& body) {
    const uint32_t q = xcd_block() * blockDim.x + threadIdx.x;
    const bool active = q < b.offset[b.CT];
    const uint32_t g = active ? b.cperm[q] : 0u;
    by_scene<kB>((kB && active) ? b.sid[g] : 0u, active, [&](uint32_t sc) { body(q, g, b.fcnt[q], sc); });
}

// vorticity pair term (pbf_vorticity_confinement.cu:15-31)
__device__ __forceinline__ F3 vort_pair(F3 self, F3 selfv, F3 pj, F3 vj, float warg) {
    const F3 pd = pj - self;
    const float r = glm_length(pd);
    const F3 vd = vj - selfv;
    F3 w = mk3(0.f, 0.f, 0.f);
    if (r > 0.f) w = mulvs(glm_cross(vd, divvs(pd, r)), wspiky_deriv(r, warg));
    return w;
}

// confinement location-vector pair term (pbf_vorticity_confinement.cu:121-138)
__device__ __forceinline__ F3 loc_pair(F3 self, F3 pj, float wj, float warg) {
    const F3 pd = pj - self;
    const float r = glm_length(pd);
    F3 l = mk3(0.f, 0.f, 0.f);
    if (r > 0.f) l = mulvs(mulsv(wj, divvs(pd, r)), wspiky_deriv(r, warg));
    return l;
}

// XSPH pair term (pbf_xsph.cu:15-29)
__device__ __forceinline__ F3 xsph_pair(F3 self, F3 selfv, F3 pj, F3 vj, float inv_h) {
    const F3 pd = pj - self;
    const float r = glm_length(pd);
    const F3 vd = vj - selfv;
    return mulvs(vd, wpoly6(r, inv_h));
}

// The three final passes as (elements, pair term, support) for the particle
// at slot g: the list walk, the direct sum, the per-cell partials and their
// combination all evaluate a pass through these.
enum FinalPass : int { PASS_VORT = 0, PASS_CONF = 1, PASS_XSPH = 2 };

template <int kPass>
struct PassTerms {
    const float4* __restrict__ R;
    const float4* __restrict__ OM4;
    F3 self, selfv;
    float arg, cut2;
    __device__ __forceinline__ PassTerms(const Bufs& b, const Phys& P, uint32_t g)
        : R(b.R), OM4(b.OM4), self(f3of(b.R[2 * g])),
          selfv(kPass == PASS_CONF ? mk3(0.f, 0.f, 0.f) : f3of(b.R[2 * g + 1])),
          arg(kPass == PASS_XSPH ? P.inv_h : P.warg), cut2(kPass == PASS_XSPH ? P.cut2_w : P.cut2_wd) {}
    __device__ __forceinline__ El2 elem(uint32_t j) const {
        if (kPass == PASS_CONF) {
            const float4 r = R[2 * j];
            return El2{make_float4(r.x, r.y, r.z, OM4[j].w), make_float4(0.f, 0.f, 0.f, 0.f)};
        }
        return El2{R[2 * j], R[2 * j + 1]};
    }
    // the same through buffer resources (records of slot j at byte 32j of R, 16j of OM4)
    __device__ __forceinline__ El2 elem_buf(const Gather16& gr, const Gather16& go, uint32_t j) const {
        if (kPass == PASS_CONF) {
            const float4 r = gr(2u * j);
            return El2{make_float4(r.x, r.y, r.z, go(j).w), make_float4(0.f, 0.f, 0.f, 0.f)};
        }
        return El2{gr(2u * j), gr(2u * j + 1u)};
    }
    __device__ __forceinline__ F3 term(const El2& e) const {
        if (kPass == PASS_VORT) return vort_pair(self, selfv, f3of(e.p), f3of(e.q), arg);
        if (kPass == PASS_CONF) return loc_pair(self, f3of(e.p), e.p.w, arg);
        return xsph_pair(self, selfv, f3of(e.p), f3of(e.q), arg);
    }
};

// Partial sum of cell c (0..26) of the 27 around the pass particle's final cell.
template <class Pass>
__device__ __forceinline__ F3 pass_cell(const Bufs& b, const Grid& G, const Pass& ps, uint32_t c) {
    int gp[3];
    grid_pos(ps.self, G.cw, gp);
    auto elem = [&](uint32_t j) { return ps.elem(j); };
    auto term = [&](const El2& e) { return ps.term(e); };
    const uint32_t key = cell27(gp, c, G);
    return big_cells(b) ? sum_cell<true>(cell_view(b), key, ps.self, ps.cut2, elem, term)
                        : sum_cell<false>(cell_view(b), key, ps.self, ps.cut2, elem, term);
}

// The pass's sum for final-order position q with stored count n: the list
// (n <= kFinalK), or, for n = kFinalK + 1, the 27 per-cell partials that
// k_ovf_part left in its block's overflow slot, added in the same z,y,x order
// as the direct 27-cell sum adds its cells (which it does when the block has
// no slot).
template <class Pass>
__device__ __forceinline__ F3 pass_sum(const Bufs& b, const Grid& G, const Pass& ps, uint32_t q, uint32_t n) {
    auto elem = [&](uint32_t j) { return ps.elem(j); };
    auto term = [&](const El2& e) { return ps.term(e); };
    if (n <= kFinalK) {
        if (b.cap < kGatherMaxSlots / 2u) {
            const Gather16 gr(b.R, 2u * b.cap), go(b.OM4, b.cap);
            return sum_list_any(b.flist, b.cap, q, n, [&](uint32_t j) { return ps.elem_buf(gr, go, j); }, term);
        }
        return sum_list(b.flist, b.cap, q, n, elem, term);
    }
    const uint32_t k = b.oslot[q >> 6];
    if (k >= b.ovf_max)
        return big_cells(b) ? sum27_global<true>(G, cell_view(b), ps.self, ps.cut2, elem, term)
                            : sum27_global<false>(G, cell_view(b), ps.self, ps.cut2, elem, term);
    const float4* __restrict__ pk = b.part + (size_t)k * 27u * 64u + (q & 63u);
    F3 sum = mk3(0.f, 0.f, 0.f);
    for (uint32_t c = 0; c < 27u; ++c) acc(sum, f3of(pk[c * 64u]));
    return sum;
}


// Per-cell partials of pass kPass for the overflow slots.  A wave task is
// (slot k, cell c): lane l sums cell c of the 27 around final-order position
// 64 * ovf[k] + l, when that position overflowed, in slot order (sum_cell, as
// the direct sum does) into part[(k * 27 + c) * 64 + l].  The lanes are
// neighbours in final-cell order, so they mostly scan the same cell and load
// the same records, and the pass reads each row of 64 whole.
template <bool kB, int kPass>
__global__ __launch_bounds__(256) void k_ovf_part(Bufs b, Scenes<kB> SC) {
    const uint32_t ns = min(b.ctr[CTR_OVF], b.ovf_max);
    const uint32_t ntask = ns * 27u;
    const uint32_t nl = b.offset[b.CT];
    const uint32_t lane = lane_id();
    const uint32_t w0 = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nw = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t t = w0; t < ntask; t += nw) {
        const uint32_t k = t / 27u, c = t - k * 27u;
        const uint32_t q = b.ovf[k] * 64u + lane;
        const bool active = q < nl && b.fcnt[q] == kFinalK + 1u;
        const uint32_t g = active ? b.cperm[q] : 0u;
        by_scene<kB>((kB && active) ? b.sid[g] : 0u, active, [&](uint32_t sc) {
            const PassTerms<kPass> ps(b, SC[sc].P, g);
            const F3 cs = pass_cell(b, SC[sc].G, ps, c);
            b.part[((size_t)k * 27u + c) * 64u + lane] = make_float4(cs.x, cs.y, cs.z, 0.f);
        });
    }
}

// ------------------------------------------------- K10: vorticity
// calcVorticityCUDA (pbf_vorticity_confinement.cu:87-115) with inv_h in the
// spiky derivative (or h, when reproducing the reference's argument);
// writes (omega, |omega|).  Elements: (x*_j, slot j).
template <bool kB>
__global__ __launch_bounds__(kBlock) void k_vort(Bufs b, Scenes<kB> SC) {
    for_final<kB>(b, [&](uint32_t q, uint32_t g, uint32_t n, uint32_t sc) {
        const Phys P = SC[sc].P;
        if (!P.vort_on) {
            b.OM4[g] = make_float4(0.f, 0.f, 0.f, 0.f);
            return;
        }
        const float mr = P.m * P.inv_rho0;
        const F3 w = pass_sum(b, SC[sc].G, PassTerms<PASS_VORT>(b, P, g), q, n);
        const F3 om = mulsv(mr, w);
        b.OM4[g] = make_float4(om.x, om.y, om.z, glm_length(om));
    });
}

// ------------------------------------------------- K11: vorticity confinement
// addVorticityConfinementCUDA (pbf_vorticity_confinement.cu:192-222), in place
// on the record's velocity.  Elements: (x*_j, |omega_j|).
template <bool kB>
__global__ __launch_bounds__(kBlock) void k_confine(Bufs b, Scenes<kB> SC) {
    float4* __restrict__ R = b.R;
    const float4* __restrict__ OM4 = b.OM4;
    for_final<kB>(b, [&](uint32_t q, uint32_t g, uint32_t n, uint32_t sc) {
        const Phys P = SC[sc].P;
        if (!P.vort_on) return;
        F3 N = pass_sum(b, SC[sc].G, PassTerms<PASS_CONF>(b, P, g), q, n);
        if (glm_length(N) > 0.f) {
            N = mulvs(mulvs(N, P.m), P.inv_rho0);
            N = divvs(N, glm_length(N));
            const F3 add = mulvs(divvs(mulsv(P.vc, glm_cross(N, f3of(OM4[g]))), P.inv_rho0), P.dtc);
            const F3 nv = f3of(R[2 * g + 1]) + add;
            R[2 * g + 1] = make_float4(nv.x, nv.y, nv.z, 0.f);
        }
    });
}

// ------------------------------------------------- K12: XSPH + write-back
// calcXSPHCUDA (pbf_xsph.cu:84-113) and updatePosition (pbf_update.cu:16-23):
// x_out = x*, v_out = v + c*m*inv_rho0*sum unless the sum is NaN, with v the
// record's velocity (after confinement when that ran).
template <bool kB>
__global__ __launch_bounds__(kBlock) void k_xsph_out(Bufs b, Scenes<kB> SC, float* __restrict__ xo, float* __restrict__ vo) {
    const float4* __restrict__ R = b.R;
    Tally tl;
    for_final<kB>(b, [&](uint32_t q, uint32_t g, uint32_t n, uint32_t sc) {
        const Phys P = SC[sc].P;
        const F3 self = f3of(R[2 * g]);
        const F3 selfv = f3of(R[2 * g + 1]);
        if (!P.xsph_on) {
            reinterpret_cast<F3*>(xo)[g] = self;
            reinterpret_cast<F3*>(vo)[g] = selfv;
            return;
        }
        const F3 a = pass_sum(b, SC[sc].G, PassTerms<PASS_XSPH>(b, P, g), q, n);
        F3 out;
        if (isnan(a.x) || isnan(a.y) || isnan(a.z)) {
            out = selfv;
            ++tl.nan;
        } else {
            out = selfv + mulsv(P.xsph_c * P.m * P.inv_rho0, a);
        }
        reinterpret_cast<F3*>(xo)[g] = self;  // 12-byte stores
        reinterpret_cast<F3*>(vo)[g] = out;
    });
    flush_tally(b.ctr, tl);
}

// ------------------------------------------------- K13: write-back without XSPH
// pbf_solve.cpp:125-145 as shipped: v = inv_dt*(x* - x_sorted), x = x*; with
// R (vorticity ran without XSPH) the record's position and confined velocity.
template <bool kB>
__global__ __launch_bounds__(256) void k_out(Bufs b, Scenes<kB> SC, const float4* __restrict__ R, float* __restrict__ xo,
                                             float* __restrict__ vo) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nl = b.offset[b.CT];
    if (s >= nl) return;
    const uint32_t cap = b.cap;
    F3 p, vv;
    if (R) {
        p = f3of(R[2 * s]);
        vv = f3of(R[2 * s + 1]);
    } else {
        p = f3of(b.Pf[s]);
        vv = mulsv(SC[kB ? b.sid[s] : 0u].P.inv_dt, p - ld3(b.X, cap, s));
    }
    reinterpret_cast<F3*>(xo)[s] = p;  // 12-byte stores
    reinterpret_cast<F3*>(vo)[s] = vv;
}

// ------------------------------------------------- views
__global__ __launch_bounds__(256) void k_neighbors_view(Bufs b, uint32_t* __restrict__ out,
                                                        uint32_t* __restrict__ counts, uint32_t max_pair) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= b.offset[b.CT]) return;
    const uint32_t cnt = b.ncnt[g];
    counts[g] = cnt;
    for (uint32_t c = 0; c < max_pair; ++c)
        out[(size_t)g * max_pair + c] = c < cnt ? b.nlist[list_at(b.cap, g, c)] : 0xFFFFFFFFu;
}

__global__ void k_cell_ranges(const uint32_t* __restrict__ offset, uint32_t C, uint32_t* __restrict__ start,
                              uint32_t* __restrict__ end) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const uint32_t a = offset[c], z = offset[c + 1];
    start[c] = (z > a) ? a : 0xFFFFFFFFu;
    end[c] = (z > a) ? z : 0u;
}

__global__ void k_copy_lambda(const uint32_t* __restrict__ offset, uint32_t C, const float4* __restrict__ Q,
                              float* __restrict__ out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= offset[C]) return;
    out[s] = Q[s].w;
}

// calcDensityCUDA (pbf/solver/cuda/pbf_density.cu:11-71) over the last
// step's frozen lists: rho_i = sum_j m * W(|x_i - x_j|), W with inv_h (the
// reference passes h as the kernel's inv_h argument, pbf_density.cu:22).
// x: AoS float3 in the step's sorted order.
template <bool kB>
__global__ void k_density(Bufs b, Scenes<kB> SC, const float* __restrict__ x, float* __restrict__ rho) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.offset[b.CT]) return;
    const Phys& P = SC[kB ? b.sid[i] : 0u].P;
    const F3 self = mk3(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
    const uint32_t n = b.ncnt[i];
    float sum = 0.f;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t j = b.nlist[list_at(b.cap, i, k)];
        const F3 d = self - mk3(x[3 * j], x[3 * j + 1], x[3 * j + 2]);
        sum += wpoly6(glm_length(d), P.inv_h) * P.m;
    }
    rho[i] = sum;
}

// restoreOrderCUDA (pbf/interaction/cuda/pbf_neighbor_search_util.cu:156-170):
// out[sorted_index[s]] = sorted[s] for the live slots of the last step.
__global__ void k_restore(Bufs b, const float* __restrict__ xs, const float* __restrict__ vs,
                          float* __restrict__ xo, float* __restrict__ vo) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= b.offset[b.CT]) return;
    const uint32_t o = b.sidx[s];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        xo[3 * o + c] = xs[3 * s + c];
        vo[3 * o + c] = vs[3 * s + c];
    }
}

// First sorted slot of each scene after the step (the particles of scene s
// are slots [out[s], out[s+1]) of the output), out[nscenes] = live total.
__global__ void k_scene_offsets(Bufs b, uint32_t* __restrict__ out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s <= b.nscenes) out[s] = b.offset[b.kbase[s]];
}

// p[0..n) = value (grid-stride): clears counters and flags inside code that a
// caller may capture in a graph, where memset nodes are avoided.
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t value) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = value;
}

__global__ void k_selftest(const float* __restrict__ a, const float* __restrict__ bb, float* __restrict__ out,
                           uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = a[i], y = bb[i];
    out[4 * i + 0] = x / y;
    out[4 * i + 1] = __builtin_sqrtf(x);
    out[4 * i + 2] = __builtin_fmaf(x, y, x);
    out[4 * i + 3] = floorf(x / y);
}

// inv_sqrt against 1.f / sqrtf on every bit pattern in [lo, hi] (grid-stride):
// counts the mismatches and records the first one's input bits (any one, if
// several race).
__global__ __launch_bounds__(256) void k_selftest_inv_sqrt(uint32_t lo, uint32_t hi,
                                                           unsigned long long* __restrict__ mismatches,
                                                           uint32_t* __restrict__ first) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint32_t bad = 0, which = 0;
    for (uint64_t u = (uint64_t)lo + blockIdx.x * blockDim.x + threadIdx.x; u <= hi; u += stride) {
        const float r2 = __uint_as_float((uint32_t)u);
        if (__float_as_uint(inv_sqrt(r2)) != __float_as_uint(1.f / __builtin_sqrtf(r2))) {
            if (!bad) which = (uint32_t)u;
            ++bad;
        }
    }
    if (bad) {
        atomicAdd(mismatches, (unsigned long long)bad);
        atomicExch(first, which);
    }
}


inline uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// Dynamic LDS of each unit kernel: the tile of 12-byte positions, plus
// kTilePad entries so that the filters read candidates four at a time past a
// cell's last one without clamping (those reads are masked out).
constexpr uint32_t kTilePad = 4;
static_assert(sizeof(F3) == 12, "the tile holds packed float3 positions");
size_t smem_of(uint32_t tile_cap) { return (size_t)(tile_cap + kTilePad) * sizeof(F3); }

// The final-list units' tile: kFinalTileMax entries whatever the search's
// tile.  A 32-cell-wide final unit in the dam's lattice (8 particles a cell)
// has a 9 x 34-cell neighbourhood of 2,448 particles; 2,560 entries stage all
// of them at four workgroups per CU (1,536 left them to the global-memory
// filter: 237 against 216 us on the 1M step, 1.95 against 1.63 ms at 8M;
// 3,072 costs a workgroup per CU and is slower again).
uint32_t final_tile(const Tuning&) { return kFinalTileMax; }

}  // namespace

void size_unit_grids(Tuning& T, int num_cu) {
    struct K {
        int id;
        const void* f;
    } ks[] = {{K_NS, (const void*)k_ns<true>}, {K_FINAL_LIST, (const void*)k_final_list<true>}};
    const uint32_t ncu = num_cu > 0 ? (uint32_t)num_cu : 1u;
    uint32_t lo = 0xFFFFFFFFu;
    for (int i = 0; i < K_NUM; ++i) T.grid[i] = 0;
    const void* single[] = {(const void*)k_ns<false>, (const void*)k_final_list<false>};
    for (const void* f : single)
        (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem_of(kTileCapMax));
    for (auto& k : ks) {
        // allow the largest tile, then ask how many workgroups of the
        // current one fit a CU
        (void)hipFuncSetAttribute(k.f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem_of(kTileCapMax));
        int nb = 0;
        const uint32_t tc = k.id == K_FINAL_LIST ? final_tile(T) : T.tile_cap;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k.f, kBlock, smem_of(tc)) != hipSuccess ||
            nb < 1)
            nb = 1;
        T.grid[k.id] = ncu * (uint32_t)nb < 8u ? 8u : ncu * (uint32_t)nb;
        lo = T.grid[k.id] < lo ? T.grid[k.id] : lo;
    }
    T.grid_units = lo;
}

const char* kernel_name(int id) {
    static const char* names[K_NUM] = {"k_predict_hash", "k_scan_tiles", "k_scan_cells", "k_scatter",
                                       "k_reorder",      "k_ns",         "k_lambda",     "k_deltap",
                                       "k_vort",         "k_confine",    "k_xsph_out",   "k_out",
                                       "k_cur_sort",     "k_final_list",  "k_big_seg"};
    return (id >= 0 && id < K_NUM) ? names[id] : "?";
}

namespace {
struct Mark {
    Timer* t;
    hipStream_t s;
    int slot;
    Mark(Timer* tm, hipStream_t st, int id) : t(tm), s(st), slot(-1) {
        if (t && t->n < kTimerSlots) {
            slot = t->n++;
            t->id[slot] = id;
            (void)hipEventRecord(t->ev[2 * slot], s);
        }
    }
    ~Mark() {
        if (slot >= 0) (void)hipEventRecord(t->ev[2 * slot + 1], s);
    }
};
}  // namespace

namespace {
template <bool kB>
hipError_t step_seq(const Bufs& b, const Scenes<kB>& SC, const Tuning& T, float* x, float* v, uint32_t n, int iters,
                    int any_xsph, int any_vort, uint32_t uw, hipStream_t s, Timer* tm, float4* qtape) {
    const uint32_t nb = ceil_div(n > 0 ? n : 1, 256);
    const uint32_t CT = b.CT;
    const uint32_t budget = (T.tile_cap * 3u) / 4u > 0 ? (T.tile_cap * 3u) / 4u : 1u;
    auto grid = [&](int k) { return dim3(T.grid[k] ? T.grid[k] : T.grid_units); };
    const size_t smem = smem_of(T.tile_cap);
    if (n > 0) {
        Mark m(tm, s, K_PREDICT_HASH);
        hipLaunchKernelGGL(k_predict_hash<kB>, dim3(nb), dim3(256), 0, s, x, v, n, b, SC, b.key, b.rank, b.count, b.tflag,
                           b.lgT);
    }
    {
        Mark m(tm, s, K_SCAN_TILES);
        hipLaunchKernelGGL(k_tile_reduce, dim3(b.nT + 1), dim3(kScanBlock), 0, s, b.count, b.tflag, b.tsum, b.nT, b.lgT, CT, 0u);
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, s, b.tsum, b.tprefix, b.nT + 1, b.ctr, 1u);
    }
    {
        Mark m(tm, s, K_SCAN_CELLS);
        hipLaunchKernelGGL(k_scan_cells, dim3(b.nT + 1), dim3(kScanBlock), 0, s, b.count, b.offset, b.tsum, b.tflag,
                           b.tprefix, b.nT, b.lgT, CT, 0u, b.ctr);
    }
    if (n == 0) return hipGetLastError();
    {
        Mark m(tm, s, K_SCATTER);
        hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(256), 0, s, b.key, b.rank, b.offset, b.tmp, n,
                           kB ? b.in_off + b.nscenes : (const uint32_t*)nullptr, b.big, &b.ctr[CTR_BIGCELLS], b.cap, b.bslot);
        // the histogram ranks are dead after the scatter: their array holds the huge cells' chunks
        static_assert(CTR_HUGECELLS == CTR_BIGCELLS + 1, "the sort reads the two cell counts together");
        hipLaunchKernelGGL(k_sort_cells, dim3(256), dim3(1024), 0, s, b.offset, b.big, &b.ctr[CTR_BIGCELLS], b.tmp,
                           b.rank, b.cap);
        hipLaunchKernelGGL(k_merge_huge, dim3(256), dim3(256), 0, s, b.offset, b.big, &b.ctr[CTR_BIGCELLS], b.tmp,
                           (const uint32_t*)b.rank, b.cap);
    }
    {
        Mark m(tm, s, K_REORDER);
        hipLaunchKernelGGL(k_reorder<kB>, dim3(nb), dim3(256), 0, s, x, v, b, SC, b.key, b.tmp, n, uw, budget);
    }
    {
        Mark m(tm, s, K_BIG_SEG);
        hipLaunchKernelGGL(k_big_seg<kB>, dim3(kBigBlocks), dim3(1024), 0, s, b, SC);
    }
    {
        Mark m(tm, s, K_NS);
        hipLaunchKernelGGL(k_ns<kB>, grid(K_NS), dim3(kBlock), smem, s, b, SC, T.tile_cap);
        hipLaunchKernelGGL(k_ns_heavy<kB>, dim3(kHeavyBlocks), dim3(1024), 0, s, b, SC);
    }
    const bool pow4 = !kB && SC.s0.P.sn == 4;  // a batch's scenes may differ: general powi
    for (int it = 0; it < iters; ++it) {
        Bufs bi = b;
        if (qtape) bi.Q = qtape + (size_t)it * b.cap;
        {
            Mark m(tm, s, K_LAMBDA);
            hipLaunchKernelGGL(k_lambda<kB>, dim3(nb), dim3(256), 0, s, bi, SC);
        }
        Mark m(tm, s, K_DELTAP);
        if (pow4)
            hipLaunchKernelGGL((k_deltap<kB, true>), dim3(nb), dim3(256), 0, s, bi, SC);
        else
            hipLaunchKernelGGL((k_deltap<kB, false>), dim3(nb), dim3(256), 0, s, bi, SC);
    }
    if (!any_vort && !any_xsph) {
        Mark m(tm, s, K_OUT);
        hipLaunchKernelGGL(k_out<kB>, dim3(nb), dim3(256), 0, s, b, SC, (const float4*)nullptr, x, v);
        return hipGetLastError();
    }
    // runs of uw final cells, nc = min(8, uw) counters each: CR = CT / uw * nc
    // counters (never more than CT), scanned in tiles of up to 4096 that divide CR
    uint32_t lguw = 0;
    while ((1u << lguw) < uw) ++lguw;
    const uint32_t lgnc = lguw < 3u ? lguw : 3u;
    const uint32_t CR = (CT >> lguw) << lgnc;
    uint32_t lgTr = 0;
    while (lgTr < 12u && ((CR >> (lgTr + 1)) << (lgTr + 1)) == CR && (1u << (lgTr + 1)) <= CR) ++lgTr;
    const uint32_t nTr = CR >> lgTr;
    {
        Mark m(tm, s, K_CUR_SORT);
        hipLaunchKernelGGL(k_cur_hash<kB>, dim3(nb), dim3(256), 0, s, b, SC, lguw, lgnc, lgTr, CR);
        hipLaunchKernelGGL(k_tile_reduce, dim3(nTr + 1), dim3(kScanBlock), 0, s, b.ccount, b.ctflag, b.ctsum, nTr, lgTr, CR,
                           lgnc);
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, s, b.ctsum, b.ctprefix, nTr + 1, b.ctr, 0u);
        hipLaunchKernelGGL(k_scan_cells, dim3(nTr + 1), dim3(kScanBlock), 0, s, b.ccount, b.coffset, b.ctsum, b.ctflag,
                           b.ctprefix, nTr, lgTr, CR, lgnc, (uint32_t*)nullptr);
        hipLaunchKernelGGL(k_cur_scatter, dim3(nb), dim3(256), 0, s, b.offset, CT, b.ckey, b.crank, b.coffset,
                           b.cperm, uw, lguw, lgnc, b.cunits, b.ctr, b.oslot);
    }
    {
        Mark m(tm, s, K_FINAL_LIST);
        const uint32_t ft = final_tile(T);
        hipLaunchKernelGGL(k_final_list<kB>, grid(K_FINAL_LIST), dim3(kBlock), smem_of(ft), s, b, SC, ft, lguw, lgnc);
        hipLaunchKernelGGL(k_final_heavy<kB>, dim3(kHeavyBlocks), dim3(1024), 0, s, b, SC);
    }
    // the overflow positions' per-cell partials, before each pass that adds them
    const dim3 ovf_grid(kOvfBlocks);
    if (any_vort) {
        {
            Mark m(tm, s, K_VORT);
            hipLaunchKernelGGL((k_ovf_part<kB, PASS_VORT>), ovf_grid, dim3(256), 0, s, b, SC);
            hipLaunchKernelGGL(k_vort<kB>, dim3(nb), dim3(kBlock), 0, s, b, SC);
        }
        Mark m(tm, s, K_CONFINE);
        hipLaunchKernelGGL((k_ovf_part<kB, PASS_CONF>), ovf_grid, dim3(256), 0, s, b, SC);
        hipLaunchKernelGGL(k_confine<kB>, dim3(nb), dim3(kBlock), 0, s, b, SC);
    }
    if (any_xsph) {
        Mark m(tm, s, K_XSPH_OUT);
        hipLaunchKernelGGL((k_ovf_part<kB, PASS_XSPH>), ovf_grid, dim3(256), 0, s, b, SC);
        hipLaunchKernelGGL(k_xsph_out<kB>, dim3(nb), dim3(kBlock), 0, s, b, SC, x, v);
    } else {
        Mark m(tm, s, K_OUT);
        hipLaunchKernelGGL(k_out<kB>, dim3(nb), dim3(256), 0, s, b, SC, (const float4*)b.R, x, v);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_step(const Bufs& b, const Scene& s0, const Tuning& T, float* x, float* v, uint32_t n, int iters,
                       int any_xsph, int any_vort, uint32_t uw, hipStream_t s, Timer* tm, float4* qtape) {
    if (b.nscenes == 1) {
        const Scenes<false> SC{b.scenes, s0};
        return step_seq<false>(b, SC, T, x, v, n, iters, any_xsph, any_vort, uw, s, tm, qtape);
    }
    if (qtape) return hipErrorInvalidValue;
    const Scenes<true> SC{b.scenes, s0};
    return step_seq<true>(b, SC, T, x, v, n, iters, any_xsph, any_vort, uw, s, tm, nullptr);
}

hipError_t launch_count_scan(uint32_t* count, uint32_t* offset, uint32_t* tsum, uint32_t* tflag, uint32_t* tprefix,
                             uint32_t nT, uint32_t lgT, uint32_t C, hipStream_t s) {
    hipLaunchKernelGGL(k_tile_reduce, dim3(nT + 1), dim3(kScanBlock), 0, s, count, tflag, tsum, nT, lgT, C, 0u);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, s, tsum, tprefix, nT + 1, (uint32_t*)nullptr, 0u);
    hipLaunchKernelGGL(k_scan_cells, dim3(nT + 1), dim3(kScanBlock), 0, s, count, offset, tsum, tflag, tprefix, nT, lgT, C, 0u,
                       (uint32_t*)nullptr);
    return hipGetLastError();
}

hipError_t launch_scene_offsets(const Bufs& b, uint32_t* out, hipStream_t s) {
    hipLaunchKernelGGL(k_scene_offsets, dim3(ceil_div(b.nscenes + 1u, 256)), dim3(256), 0, s, b, out);
    return hipGetLastError();
}

hipError_t launch_neighbors_view(const Bufs& b, uint32_t* out, uint32_t* counts, uint32_t max_pair, hipStream_t s) {
    hipLaunchKernelGGL(k_neighbors_view, dim3(ceil_div(b.cap > 0 ? b.cap : 1, 256)), dim3(256), 0, s, b, out, counts,
                       max_pair);
    return hipGetLastError();
}

hipError_t launch_cell_ranges(const Bufs& b, uint32_t C, uint32_t* start, uint32_t* end, hipStream_t s) {
    hipLaunchKernelGGL(k_cell_ranges, dim3(ceil_div(C, 256)), dim3(256), 0, s, b.offset, C, start, end);
    return hipGetLastError();
}

hipError_t launch_copy_lambda(const Bufs& b, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_copy_lambda, dim3(ceil_div(b.cap > 0 ? b.cap : 1, 256)), dim3(256), 0, s, b.offset, b.CT,
                       (const float4*)b.Q, out);
    return hipGetLastError();
}

hipError_t launch_sort_cells(const uint32_t* offset, const uint32_t* cells, const uint32_t* ncells, uint32_t* vals,
                             uint32_t* scratch, uint32_t cap, hipStream_t s) {
    hipLaunchKernelGGL(k_sort_cells, dim3(256), dim3(1024), 0, s, offset, cells, ncells, vals, scratch, cap);
    hipLaunchKernelGGL(k_merge_huge, dim3(256), dim3(256), 0, s, offset, cells, ncells, vals,
                       (const uint32_t*)scratch, cap);
    return hipGetLastError();
}

hipError_t launch_density(const Bufs& b, const Scene& s0, const float* x, float* rho, hipStream_t s) {
    const dim3 g(ceil_div(b.cap > 0 ? b.cap : 1, 256));
    if (b.nscenes == 1) {
        const Scenes<false> SC{b.scenes, s0};
        hipLaunchKernelGGL(k_density<false>, g, dim3(256), 0, s, b, SC, x, rho);
    } else {
        const Scenes<true> SC{b.scenes, s0};
        hipLaunchKernelGGL(k_density<true>, g, dim3(256), 0, s, b, SC, x, rho);
    }
    return hipGetLastError();
}

hipError_t launch_restore(const Bufs& b, const float* xs, const float* vs, float* xo, float* vo, hipStream_t s) {
    hipLaunchKernelGGL(k_restore, dim3(ceil_div(b.cap > 0 ? b.cap : 1, 256)), dim3(256), 0, s, b, xs, vs, xo, vo);
    return hipGetLastError();
}

hipError_t launch_selftest_inv_sqrt(uint32_t lo, uint32_t hi, unsigned long long* mismatches, uint32_t* first,
                                    hipStream_t s) {
    hipLaunchKernelGGL(k_selftest_inv_sqrt, dim3(16384), dim3(256), 0, s, lo, hi, mismatches, first);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t value, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const size_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_fill_u32, dim3(nb < 4096 ? (uint32_t)nb : 4096u), dim3(256), 0, s, p, n, value);
    return hipGetLastError();
}

hipError_t launch_selftest(const float* a, const float* bb, float* out, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_selftest, dim3(ceil_div(n, 256)), dim3(256), 0, s, a, bb, out, n);
    return hipGetLastError();
}

}  // namespace pbfx
