// pbfx_internal.hpp -- shared definitions of the pbfx HIP library (gfx950).
//
// Layout in HBM (capacity `cap` particles, C cells):
//   key, rank, tmp, sorted_index   u32[cap]      hash / histogram rank / scatter / slot->index
//   Pf                             f32x4[cap]    predicted positions x* (w unused)
//   Q                              f32x4[cap]    (x*, s) of the current Jacobi iteration (s = -lambda)
//   X                              f32[3*cap]    SoA sorted old positions
//   R                              f32x4[2*cap]  per slot: (x*, 0), (v, 0) -- the final passes' own record
//   OM4                            f32x4[cap]    per slot: (omega, |omega|)
//   nlist, ncnt                    u32[cap*stride], u32[cap]  frozen neighbour lists as sorted
//                                                 slots in groups of four: entries 4q..4q+3 of slot
//                                                 i are the 16-byte word q*cap + i (list_at)
//   count                          u32[C+1]      cell histogram (cell C = deleted)
//   offset                         u32[C+2]      exclusive scan: cell c owns [off[c], off[c+1])
//   tsum, tflag, tprefix           u32[nT+1]     per-4096-cell-tile totals, occupied flags, scan
//   units                          uint2[cap+]   work items (first cell key, width | chunk << 5)
//   ckey, crank, cperm, ccount, coffset, ctsum, ctprefix, cunits
//                                                 the same sort keyed by each particle's final cell;
//                                                 once placed, crank holds the queue of final-order
//                                                 positions whose list build k_final_heavy does (and,
//                                                 between the search and the regrouping, the slots
//                                                 whose search list k_ns_heavy builds)
//   flist, fcnt                    u32[64*cap], u32[cap]  in-range pairs of the final passes by
//                                                 final-order position (groups of four, list_at), counts
//   bslot, aux                     u32[cap]      big cell -> descriptor (at its first slot); its
//                                                 slots grouped by image.  After the sort, `rank`
//                                                 holds each slot's group mark (0: not in a big cell)
//   bseg                           BigSeg[big_list_len(cap)]  groups and bounding boxes of big cells
//   ovf, oslot, part               u32[M], u32[cap/64+1], f32x4[M*27*64]  overflow slots:
//                                                 each slot's block of 64 final-order positions,
//                                                 each block's slot, and the slots' per-cell partial
//                                                 sums (M = cap/64 + 1 in a context)
//   ctr                            u32[16*65]    allocation counters, then 64 copies of the
//                                                 statistics counters of the last step
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pbfx {

constexpr int kBlock = 256;               // threads per work-unit workgroup
constexpr int kWaves = kBlock / 64;
constexpr uint32_t kMaxW = 32;            // widest work unit along x (cells)
constexpr uint32_t kMaxCells = 9 * (kMaxW + 2);
constexpr uint32_t kTileCapMax = 4096;
constexpr uint32_t kScanTile = 4096;      // cells per scan tile
constexpr uint32_t kChunk = kBlock;       // own particles per work item
constexpr uint32_t kFinalTileMax = 2560;  // LDS tile of the final-list units (particles)
constexpr uint32_t kSparse = 64;          // final-cell runs below this go one thread per particle
// Cells above kSmallCell particles get their index segment sorted, so that a
// particle's in-cell rank is its position rather than a count over the whole
// cell: up to kSortMax in one LDS sort, above that ("huge" cells) in chunks of
// kSortMax that are sorted in LDS and then merged by rank.
constexpr uint32_t kSmallCell = 64;
constexpr uint32_t kSortMax = 16384;
// Entries of a big-cell list: the cells above kSmallCell, then (from
// huge_base on) the cells above kSortMax.
__host__ __device__ constexpr uint32_t huge_base(uint32_t cap) { return cap / kSmallCell + 2u; }
__host__ __device__ constexpr uint32_t big_list_len(uint32_t cap) { return huge_base(cap) + cap / kSortMax + 2u; }
// Segments of a big cell (more than kSmallCell particles).  The hash wraps,
// so one cell can hold particles of several images of the grid (a pile on
// the far wall x = gx*cw hashes into column 0, beside the particles at
// x = 0) and dense piles far from a scanning particle.  Each big cell's
// particles are grouped by image (a stable partition into `aux`), and each
// group's bounding box is kept at the search's positions and at the final
// ones: a scan skips every group whose box lies out of range, and when one
// group remains it visits that group's slots only, in ascending order --
// the same candidates in range, in the same order, as the whole cell.
constexpr uint32_t kMaxSeg = 8;           // groups per big cell (more: the cell is scanned whole)
constexpr uint32_t kCodeOther = 64;       // image code of non-finite positions and images beyond [-2, 1]
constexpr uint32_t kCodes = kCodeOther + 1;
constexpr uint32_t kBigBlocks = 256;      // workgroups (of 1024) of the per-big-cell kernels
struct BigSeg {
    uint32_t nseg;            // groups (0: scan the cell whole)
    uint32_t end[kMaxSeg];    // end of group k in aux positions from the cell's first slot
    uint32_t pad[7];
    float4 slo[kMaxSeg], shi[kMaxSeg];  // box of the group's x* at the search
    uint4 fmin[kMaxSeg];      // box of the group's final x* as order-preserving keys (fkey), w: a
    uint4 fmax[kMaxSeg];      // member is non-finite; gathered by atomics
};
constexpr uint32_t kFinalK = 64;          // stored in-range pairs per particle for the final passes
// Final-pass particles with more than kFinalK pairs (piles) get their 27
// per-cell partial sums computed in parallel before each pass (kOvfBlocks
// workgroups), by block of 64 consecutive final-order positions: a block
// holding such a particle takes one of ovf_max slots (every block can have one
// in a context; the allocation-free step has kStaticOvfSlots), and its
// partials are stored slot by slot, cell by cell, 64 lanes a row.  A block
// that finds no slot left sums its cells one after another in the pass.
constexpr uint32_t kOvfBlocks = 2048;
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;       // oslot: no particle of the block overflowed (yet)
constexpr uint32_t kPendingSlot = 0xFFFFFFFEu;  // oslot: being allocated (never seen across kernels)
constexpr uint32_t kFullSlot = 0xFFFFFFFDu;     // oslot: no slot was left
// A final-list or search-list build from global memory whose next cell would
// take it past kHeavyScan candidates (a particle beside a pile) is handed to
// k_final_heavy / k_ns_heavy, which scan it with a whole wave (kHeavyBlocks
// workgroups of 16 waves).
constexpr uint32_t kHeavyScan = 2048;
static_assert(kHeavyScan >= 27u * kSmallCell, "a build without big cells never reaches the budget");
constexpr uint32_t kHeavyBlocks = 256;

enum Ctr : int {
    CTR_UNITS = 0,
    CTR_TRUNC,
    CTR_SWEEPCAP,
    CTR_MAXSWEEP,
    CTR_XSPHNAN,
    CTR_SLOWUNITS,
    CTR_SLOWPART,
    CTR_HEAVY,      // final-list builds handed to k_final_heavy (allocation counter of the queue in crank)
    CTR_MAXTILE,
    CTR_CUNITS,
    CTR_LISTOVF,
    CTR_BIGCELLS,   // cells with kSmallCell < n <= kSortMax particles (sorted by k_sort_cells)
    CTR_HUGECELLS,  // cells with more than kSortMax particles (chunk-sorted, then merged)
    CTR_SCANTILES,  // occupied tiles of the cell scan (statistics)
    CTR_OVF,        // final-pass positions whose list overflowed (allocation counter of ovf)
    CTR_NSHEAVY,    // search lists handed to waves (allocation counter of the queue in crank)
    CTR_NUM = 16
};
constexpr uint32_t kStatRep = 64;                           // copies of the statistics counters
constexpr uint32_t kCtrWords = CTR_NUM * (1 + kStatRep);    // ctr[0..CTR_NUM): allocation counters

// One scene's hash grid.  Scenes occupy consecutive ranges of one cell-key
// space: cell (x, y, z) of a scene has key base + (z*gy + y)*gx + x.
struct Grid {
    uint32_t gx, gy, gz;
    uint32_t lgx, lgy;
    uint32_t C;
    float cw;
    uint32_t base;
};

// Constants of one step, all computed on the host in IEEE float exactly as
// the reference's host code does (pbf_solve.cpp:147-216, pbf_position_update.cu:224-238).
struct Phys {
    float h, inv_h, h2;
    float a, ad;            // memo: poly6*inv_h*inv_h^2, -spiky*inv_h^2*inv_h^2
    float m, inv_rho0, relax, m2r2;
    float inv_k;            // REFERENCE s_corr: 1/(float)pow(W(0.4h),4)
    float inv_wdq, sk;      // PAPER s_corr: 1/W(dq), k
    int32_t mode, sn;
    float dt, inv_dt, dtc;  // dtc = 1/inv_dt, the confinement time step
    float g[3];
    float xsph_c, vc, warg; // warg: inv_h, or h to reproduce the vorticity argument mix-up
    int32_t cap;            // neighbours kept per particle (max_pair - 1)
    int32_t stride;         // list entries allocated per particle (cap rounded up to a multiple of 4)
    int32_t max_sweeps;
    int32_t outer_as_inner;
    float cut2_w, cut2_wd;  // squared support bounds of W (arg inv_h) and of W' (arg warg): pairs beyond add 0
    int32_t xsph_on, vort_on;
};

struct Boundary {
    const float4* inner;
    uint32_t n_inner;
    const float4* outer;
    uint32_t n_outer;
    const float* pts;
    const float* nrm;
    uint32_t n_planes;
};

// Everything a kernel needs to know about one scene of the batch.
struct Scene {
    Grid G;
    Phys P;
    Boundary B;
    float lo[3], hi[3];   // simulation area: predicted positions outside are deleted
};

struct Bufs {
    uint32_t cap;
    // scenes: table, count, total cell keys (padded; key CT = deleted)
    const Scene* scenes;
    uint32_t nscenes;
    uint32_t CT;
    const uint32_t* in_off;   // [nscenes+1] first input index of each scene in this step
    const uint32_t* kbase;    // [nscenes+1] first cell key of each scene, then CT
    uint32_t* isid;           // scene of each input index
    uint32_t* sid;            // scene of each sorted slot
    uint32_t* key;
    uint32_t* rank;
    uint32_t* tmp;
    uint32_t* big;            // [big_list_len(cap)] cells whose segment of tmp k_sort_cells sorts (huge
                              // cells from huge_base(cap) on)
    uint32_t* bslot;          // [cap] at a big cell's first slot: its index in big (and bseg)
    uint32_t* aux;            // [cap] a big cell's slots grouped by image, ascending in each group
    BigSeg* bseg;             // [big_list_len(cap)] groups and boxes of each big cell
    uint32_t* sidx;
    float4* Pf;
    float4* Q;
    float* X;
    float4* R;
    float4* OM4;
    uint32_t* nlist;
    uint32_t* ncnt;
    uint32_t* count;
    uint32_t* offset;
    uint32_t* tsum;
    uint32_t* tflag;
    uint32_t* tprefix;
    uint2* units;
    // regrouping by final cell for the vorticity / confinement / XSPH passes
    uint32_t* ckey;
    uint32_t* crank;
    uint32_t* cperm;
    uint32_t* ccount;
    uint32_t* coffset;
    uint32_t* ctsum;
    uint32_t* ctflag;
    uint32_t* ctprefix;
    uint2* cunits;
    uint32_t* flist;   // final pair lists, kFinalK entries per position in groups of four (list_at)
    uint32_t* fcnt;    // pairs per final-order position (kFinalK + 1: more, see the block's slot)
    uint32_t* ovf;     // [ovf_max] the block (final-order position / 64) of each overflow slot
    uint32_t* oslot;   // [cap / 64 + 1] each block's overflow slot (kNoSlot, kFullSlot)
    float4* part;      // [ovf_max * 27 * 64] the slots' per-cell partial sums: (slot, cell, lane)
    uint32_t ovf_max;
    uint32_t* ctr;
    uint32_t lgT;
    uint32_t nT;
};

// Kernels of one step, for the optional per-launch HIP-event timing.
enum KernelId : int {
    K_PREDICT_HASH = 0,
    K_SCAN_TILES,
    K_SCAN_CELLS,
    K_SCATTER,
    K_REORDER,
    K_NS,
    K_LAMBDA,
    K_DELTAP,
    K_VORT,
    K_CONFINE,
    K_XSPH_OUT,
    K_OUT,
    K_CUR_SORT,
    K_FINAL_LIST,
    K_BIG_SEG,
    K_NUM
};
constexpr int kTimerSlots = 64;

// Event pairs recorded on the step's stream around each launch.
struct Timer {
    hipEvent_t ev[2 * kTimerSlots];
    int id[kTimerSlots];
    int n;
};

struct Tuning {
    uint32_t tile_cap;      // particles staged on chip per unit
    uint32_t unit_width;    // max cells per unit along x
    uint32_t grid_units;    // persistent workgroups for the unit kernels (the smallest of grid[])
    uint32_t grid[K_NUM];   // persistent workgroups per unit kernel, from its own occupancy
};

// ---------------------------------------------------------- differentiable form
// What the backward pass needs of one recorded step (device arrays of a
// single-scene context, capacity cap).
struct StepTape {
    uint32_t n_in, n_live;
    uint32_t* sidx;     // [cap]   slot -> input index
    float* X;           // [3*cap] sorted input positions, SoA with stride cap
    uint32_t* nlist;    // [stride*cap] frozen lists, column-major
    uint32_t* ncnt;     // [cap]
    uint32_t* offset;   // [CT+2]  cell ranges of the search
    float4* Q;          // [iters*cap] (x*, s) entering each Jacobi iteration
    float4* PI;         // [cap]   final positions
    float4* R;          // [2*cap] (x*, 0), (velocity after confinement, 0) when the final passes ran
    float4* OM4;        // [cap]   (omega, |omega|) when vorticity ran
};

// Work arrays of the adjoint of one step, in FP64: the map's conditioning
// (up to ~1e7 in exploding flows) would turn FP32 rounding into O(1) error.
// The adjoint carries a common power-of-two scale kept by the host.
struct AdjBufs {
    uint32_t cap;
    double* ax;         // [3*cap] dL/d(step output x), output order
    double* av;         // [3*cap] dL/d(step output v)
    double* axin;       // [3*cap] dL/d(step input x), input order
    double* avin;       // [3*cap] dL/d(step input v)
    double4* aP;        // dL/d positions of the current stage
    double4* aQ;        // dL/d positions, partial (pre-collision path + first pass)
    double4* aV;        // dL/d velocity before confinement
    double4* aVc;       // dL/d velocity after confinement
    double4* aOm;       // dL/d vorticity
    double4* etab;      // dL/d confinement location vector
    double4* abar;      // dL/d pre-collision positions
    double4* coef;      // (dL/d sum W, 2 dL/dG * sum kg) per particle
    double* g2;         // 2 dL/dG per particle
    double4* aX;        // dL/d sorted input positions
    uint32_t* xnan;     // XSPH NaN guard per slot
    // particles grouped by final cell: the transposed 27-cell sets
    uint32_t *ckey, *crank, *ctmp, *cperm, *ccount, *coffset, *ctsum, *ctflag, *ctprefix;
    uint32_t *big, *nbig;  // final cells whose segment of ctmp is sorted (big_list_len(cap)), and the two
                           // counts (big, huge)
    // per-kernel copies in final-cell order (position q holds slot cperm[q]):
    // (final position, slot bits), a float4 and a double4 of the pass
    float4* gP;
    float4* gV;
    double4* gD;
    // particles with more than kAdjHeavy candidates run one wave each, per
    // set: bit 0 the final passes' 27-cell set, bit 1 its transpose, bit 2
    // the search cells (reverse-list construction), bit 3 the reverse list
    uint32_t* hflag;    // [cap]
    uint32_t* hlist;    // [4*cap] the heavy particles of each set
    uint32_t* nh;       // [4] list lengths
    // reverse lists: the particles that list each slot
    uint2* last;        // (27-cell index, slot) of a full list's last entry; (255, 0) if not full
    uint32_t *rcount, *roff, *rtsum, *rtflag, *rtprefix, *rlist;
    uint32_t RT, lgRT, nRT;   // padded slot space of the reverse-count scan
    double* pc;         // [cap] each particle's share of dL/dc
    double* pe;         // [cap] each particle's share of dL/d eps
    double* psum;       // [2]
    unsigned long long* amax;  // [1] largest |adjoint| of the step input, as the bits of a double
};
constexpr uint32_t kAdjMaxBlocks = 4096;   // workgroups of a per-particle adjoint launch
constexpr uint32_t kAdjHeavy = 256;        // candidates above which a particle gets a wave

// Launchers (pbfx_kernels.hip).  All enqueue on `s` and never synchronise.
// n_in: particles entering (all scenes); uw: cells per work unit along x,
// a power of two no wider than any scene's grid; any_xsph / any_vort: some
// scene runs XSPH / vorticity.
// s0: the scene (a single-scene context passes it by value to the kernels).
// qtape (may be NULL, single scene only): (x*, s) of Jacobi iteration k is
// written to qtape + k*cap instead of b.Q.
hipError_t launch_step(const Bufs& b, const Scene& s0, const Tuning& T, float* x, float* v, uint32_t n_in, int iters,
                       int any_xsph, int any_vort, uint32_t uw, hipStream_t s, Timer* timer,
                       float4* qtape = nullptr);
// Exclusive scan of count[0..C] (C a multiple of 2^lgT, nT = C >> lgT tiles,
// tflag marking the occupied tiles) into offset[0..C+1]; clears count and tflag.
hipError_t launch_count_scan(uint32_t* count, uint32_t* offset, uint32_t* tsum, uint32_t* tflag, uint32_t* tprefix,
                             uint32_t nT, uint32_t lgT, uint32_t C, hipStream_t s);
// out[s] = first output slot of scene s, out[nscenes] = live total (device).
hipError_t launch_scene_offsets(const Bufs& b, uint32_t* out, hipStream_t s);
const char* kernel_name(int id);
hipError_t launch_neighbors_view(const Bufs& b, uint32_t* out, uint32_t* counts, uint32_t max_pair, hipStream_t s);
hipError_t launch_cell_ranges(const Bufs& b, uint32_t C, uint32_t* start, uint32_t* end, hipStream_t s);
hipError_t launch_copy_lambda(const Bufs& b, float* out, hipStream_t s);
hipError_t launch_selftest(const float* a, const float* bb, float* out, uint32_t n, hipStream_t s);
hipError_t launch_selftest_inv_sqrt(uint32_t lo, uint32_t hi, unsigned long long* mismatches, uint32_t* first,
                                    hipStream_t s);
hipError_t launch_density(const Bufs& b, const Scene& s0, const float* x, float* rho, hipStream_t s);
// Sorts, ascending, the segment vals[offset[c] .. offset[c+1]) (distinct
// values) of each cell c listed in cells: cells[0 .. ncells[0]) hold at most
// kSortMax each, cells[huge_base(cap) .. + ncells[1]) more.  scratch (as long
// as vals) holds the huge cells' sorted chunks in between.
hipError_t launch_sort_cells(const uint32_t* offset, const uint32_t* cells, const uint32_t* ncells, uint32_t* vals,
                             uint32_t* scratch, uint32_t cap, hipStream_t s);
hipError_t launch_restore(const Bufs& b, const float* xs, const float* vs, float* xo, float* vo, hipStream_t s);
// p[0..n) = value by a kernel (no memset node when the caller captures a graph).
hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t value, hipStream_t s);
// Fills T.grid[] (and T.grid_units) for num_cu compute units.
void size_unit_grids(Tuning& T, int num_cu);

// Adjoint launchers (pbfx_adjoint.hip).  launch_adj_step: dL/d(step input)
// into a.axin / a.avin from a.ax / a.av, and the step's dL/dc, dL/d eps into
// a.psum, for the recorded step t of the single scene s0.  It waits on `s`
// twice, to read the heavy-list lengths into h_nh (pinned, 4 words).
hipError_t launch_adj_step(const Bufs& b, const Scene& s0, const StepTape& t, const AdjBufs& a, int iters,
                           hipStream_t s, uint32_t* h_nh);
// a.amax = max |a.axin|, |a.avin| over n rows (as double bits).
hipError_t launch_adj_absmax(const AdjBufs& a, uint32_t n, hipStream_t s);
// x[0..n) *= f for each of the arrays (may be NULL).
hipError_t launch_adj_scale(double* x, double* y, uint32_t n, double f, hipStream_t s);
// id_out[q] = id_in[sidx[q]] for q < n; id_in NULL: the identity.
hipError_t launch_adj_ident(const uint32_t* id_in, const uint32_t* sidx, uint32_t* id_out, uint32_t n, hipStream_t s);

}  // namespace pbfx
