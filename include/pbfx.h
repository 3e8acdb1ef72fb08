/*
 * pbfx.h -- C-ABI of the MI355X-native Position Based Fluids solver step.
 *
 * The drop-in boundary for the reference's PBF step (yodaplusplus/ssf_pbf_cuda,
 * paths below relative to cuda_pbf_with_ssf/src/).  Plain pointers and sizes
 * only; device pointers are HIP device memory (or any memory the device can
 * read and write); `stream` is a hipStream_t passed as void* (NULL = default
 * stream).  A context is re-entrant: all parameters travel as kernel
 * arguments, nothing lives in global device state, and each context owns its
 * scratch, allocated at create time -- no allocation happens inside a step.
 *
 * Semantics follow pbf::one_step (pbf/solver/pbf_solve.cpp:147-216):
 *  predict -> hashed-grid neighbour search (stable counting sort, frozen
 *  capped neighbour lists) -> I Jacobi iterations (lambda, delta-p with
 *  tensile correction, boundary response, apply) -> velocity update
 *  (+ vorticity confinement and XSPH when enabled).  Positions and velocities
 *  are written back in the new cell-sorted order, and particles whose
 *  predicted position leaves the simulation area are deleted.
 */
#ifndef PBFX_H
#define PBFX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PBFX_ABI_VERSION 1
#define PBFX_EMPTY 0xFFFFFFFFu

typedef enum pbfx_status {
    PBFX_OK = 0,
    PBFX_INVALID_ARG = 1,      /* null pointer, non-positive parameter, bad iteration count */
    PBFX_CAPACITY = 2,         /* num exceeds the context capacity (pbf_type.cpp:42-44 exits) */
    PBFX_GRID_TOO_LARGE = 3,   /* a grid dimension exceeds 1028 (pbf_type.cpp:155-158 exits) */
    PBFX_HIP_ERROR = 4,        /* a HIP runtime call failed; see pbfx_last_hip_error */
    PBFX_NOT_CONVERGED = 5,    /* information: a collision sweep loop hit its cap */
    PBFX_NEIGHBOR_TRUNCATED = 6,/* information: some particle had more in-range
                                  candidates than neighbour slots (the reference
                                  truncates silently, pbf_neighbor_search_util.cu:367) */
    PBFX_TILE_OVERFLOW = 7,    /* reserved: never returned (neighbourhoods of any size
                                  are searched) */
    PBFX_NOT_SUPPORTED = 8
} pbfx_status;

enum { PBFX_SCORR_REFERENCE = 0, PBFX_SCORR_PAPER = 1 };

/* Superset of pbf_parameter (pbf/pbf_type.h:49-63) and pbf_external
 * (pbf/pbf_type.h:69-72). */
typedef struct pbfx_params {
    float stable_density;      /* rho0 [kg/m^3] */
    float stable_distance;     /* lattice spacing [m] (informational) */
    float particle_mass;       /* m [kg] */
    float smoothing_length;    /* h [m]; also the hash cell width */
    float time_step;           /* dt [s] */
    float relaxation;          /* epsilon in the lambda denominator [m^-2] */
    float xsph_parameter;      /* XSPH c */
    float vc_parameter;        /* vorticity confinement epsilon */
    float body_force[3];       /* [m/s^2] */
    int32_t scorr_mode;        /* PBFX_SCORR_REFERENCE: 10*(max(-s_i,0)+max(-s_j,0))*(W/W(0.4h))^4
                                  (pbf/solver/cuda/pbf_position_update.cu:131-151);
                                  PBFX_SCORR_PAPER: scorr_k*(W/W(scorr_dq_over_h*h))^scorr_n */
    float scorr_k;
    int32_t scorr_n;           /* >= 1 */
    float scorr_dq_over_h;
    int32_t max_pair;          /* neighbour slots per particle, 2..64; the reference uses 30
                                  and keeps at most max_pair-1 neighbours */
    int32_t xsph_on;           /* 0 reproduces the shipped update (pbf_solve.cpp:125-145) */
    int32_t vort_on;
    int32_t vort_h_as_inv_h;   /* 1 reproduces the reference's h-for-inv_h argument in the
                                  vorticity kernels (pbf_vorticity_confinement.cu:28,134) */
    int32_t outer_as_inner;    /* 1 reproduces pbf_response_collision.cu:46-51 (outer spheres
                                  answered as inner spheres) */
    int32_t max_sweeps;        /* collision sweeps per Jacobi iteration (reference: unbounded) */
} pbfx_params;

/* Fills `p` with the reference's defaults for the drop-in path: REFERENCE
 * s_corr, 30 neighbour slots, update as shipped (no XSPH, no vorticity),
 * 32 collision sweeps; physical constants zero (the caller must set them). */
void pbfx_params_default(pbfx_params* p);

typedef struct pbfx_ctx pbfx_ctx;

/* Replaces pbf_particle::allocate(num, domain_end) + neighbor_search::
 * neighbor_search (pbf/pbf_type.cpp:165-175, pbf/interaction/pbf_neighbor_search.cpp:11-69)
 * and pbf_buffer::allocate (pbf/pbf_type.cpp:80-101): grid = calcGridSize(2h, domain_end)
 * (pbf/pbf_type.cpp:141-162), cell width h, scratch for `capacity` particles.
 * Allocates on the current HIP device. */
pbfx_status pbfx_create(const pbfx_params* params, uint32_t capacity, const float domain_end[3],
                        pbfx_ctx** out);

/* Replaces pbf_particle::free / pbf_buffer::free / ~neighbor_search. */
void pbfx_destroy(pbfx_ctx* ctx);

/* Updates the step parameters (the reference reads pbf_particle::parameter
 * at every one_step).  smoothing_length must keep the grid of pbfx_create. */
pbfx_status pbfx_set_params(pbfx_ctx* ctx, const pbfx_params* params);

/* Replaces pbf_boundary::allocate (pbf/pbf_boundary.cpp:4-25): inner and
 * outer spheres as float4 (center, radius), planes as point + normal float3.
 * Host pointers; copied to device memory owned by the context.  Counts may
 * be zero (no boundary). */
pbfx_status pbfx_set_boundary(pbfx_ctx* ctx, const float* inner_spheres4, uint32_t n_inner,
                              const float* outer_spheres4, uint32_t n_outer, const float* plane_points3,
                              const float* plane_normals3, uint32_t n_planes);

/* Replaces pbf::one_step(pbf_particle&, pbf_buffer&, const pbf_boundary&,
 * simulation_area, num_solver_iteration) (pbf/solver/pbf_solve.cpp:182-216;
 * pbf/solver/pbf_solve.h:64-70).
 *  x, v     device AoS float3 [num]; overwritten in the new sorted order.
 *  num_inout  live particles before; after the call, live particles after
 *             deletion (this call synchronises `stream` once, as the
 *             reference's D2H copy at pbf_neighbor_search.cpp:156 does).
 *  lo, hi   simulation area; particles whose predicted position is outside
 *           [lo, hi] (inclusive bounds) are deleted.
 *  iters    Jacobi iterations (>= 1).
 * Returns PBFX_OK, an error, or an information status
 * (PBFX_NEIGHBOR_TRUNCATED / PBFX_NOT_CONVERGED) when the corresponding
 * event occurred during the step and the context was created with
 * report_info enabled (pbfx_set_report_info). */
pbfx_status pbfx_step(pbfx_ctx* ctx, float* x, float* v, uint32_t* num_inout, const float lo[3],
                      const float hi[3], int iters, void* stream);

/* Same step without any host synchronisation: `num` is the live count
 * entering the step; the count after deletion stays on the device
 * (pbfx_get_num reads it), so a caller chaining steps whose simulation area
 * can delete particles reads it before passing the next count.  Use this
 * form for streams of steps and for graph capture. */
pbfx_status pbfx_step_async(pbfx_ctx* ctx, float* x, float* v, uint32_t num, const float lo[3],
                            const float hi[3], int iters, void* stream);

/* Appends add_n particles after the live ones: x[num..num+add_n) = add_x,
 * v[num..num+add_n) = add_v (device AoS float3, copied on `stream`) and
 * *num_inout += add_n.  Replaces pbf::cuda::addParticle
 * (pbf/util/cuda/pbf_add.cu:12-28), the emission of the reference's jet
 * (simulator/sphere/pbf_sphere_sim.cpp:84-86,386-387).  PBFX_CAPACITY if the
 * count would exceed the context's capacity (nothing is copied then). */
pbfx_status pbfx_add_particles(pbfx_ctx* ctx, float* x, float* v, uint32_t* num_inout, const float* add_x,
                               const float* add_v, uint32_t add_n, void* stream);

/* Live particle count after the last step on `stream` (synchronises it). */
pbfx_status pbfx_get_num(pbfx_ctx* ctx, void* stream, uint32_t* num);

/* Replaces the no-boundary pbf::one_step (pbf/solver/pbf_solve.cpp:147-180),
 * whose collision is the hard-coded inner sphere centre (3,3,3), radius 3
 * (pbf/solver/cuda/pbf_constraint.cu:262-315). */
pbfx_status pbfx_step_no_boundary(pbfx_ctx* ctx, float* x, float* v, uint32_t* num_inout,
                                  const float lo[3], const float hi[3], int iters, void* stream);

/* ---------------------------------------------------------------- batched form
 * Many independent scenes advanced by ONE sequence of launches per step: the
 * scenes occupy disjoint ranges of one cell-key space, are sorted together,
 * and every kernel reads its particle's scene parameters, grid and boundary
 * from a table.  Each scene evolves exactly as it would alone (pbfx_step):
 * same sort order within the scene, same neighbour lists, same arithmetic.
 * No counterpart in the reference, which runs one scene per process from
 * global constant memory (pbf/solver/cuda/pbf_position_update.cu:14-17,224-238);
 * the per-scene semantics are pbf::one_step's (pbf/solver/pbf_solve.cpp:182-216). */
typedef struct pbfx_scene_desc {
    pbfx_params params;          /* max_pair must be equal across the batch */
    uint32_t capacity;           /* particles this scene may hold */
    float domain_end[3];         /* grid = calcGridSize(2h, domain_end) */
    const float* inner_spheres4; /* host, float4 (center, radius) */
    uint32_t n_inner;
    const float* outer_spheres4;
    uint32_t n_outer;
    const float* plane_points3;  /* host, float3 */
    const float* plane_normals3;
    uint32_t n_planes;
    float area_lo[3], area_hi[3];/* simulation area (deletion bounds) */
} pbfx_scene_desc;

pbfx_status pbfx_batch_create(const pbfx_scene_desc* scenes, uint32_t n_scenes, pbfx_ctx** out);

/* x, v: device AoS float3, the scenes concatenated in order (scene s starts
 * at the sum of the counts before it).  counts: host [n_scenes], live
 * particles per scene in, and after the step (one host synchronisation)
 * live particles per scene out; the output is again concatenated in order. */
pbfx_status pbfx_batch_step(pbfx_ctx* ctx, float* x, float* v, uint32_t* counts, int iters, void* stream);

/* The same without host synchronisation, for streams of steps and graph
 * capture: in_offsets is device [n_scenes+1], the first input index of each
 * scene and then the total; pbfx_batch_offsets writes the output's offsets
 * in that form (the next step's in_offsets). */
pbfx_status pbfx_batch_step_async(pbfx_ctx* ctx, float* x, float* v, const uint32_t* in_offsets, uint32_t total,
                                  int iters, void* stream);
pbfx_status pbfx_batch_offsets(pbfx_ctx* ctx, uint32_t* out_offsets, void* stream);

/* Graph form of the batched step (the batched workload's "one recorded
 * launch per step"): records `steps` consecutive batched steps on x, v (as
 * pbfx_batch_step_async), each followed by the write of its output offsets
 * into `offsets` (as pbfx_batch_offsets), into one hipGraph kept by the
 * context; pbfx_batch_graph_launch replays it `launches` times on a stream.
 * offsets: device [n_scenes+1], the input offsets of the next step (set by
 * the caller before the first launch, updated in place by every step), so
 * replays chain, per-scene deletion included: the kernels read the live
 * input count from offsets[n_scenes] and size their grids for the capacity.
 * The buffers, iteration count, parameters, boundaries and tuning are those
 * at capture time (capture again after changing them).  `stream` must be a
 * created stream (not the null stream); nothing runs while recording.
 * No counterpart in the reference, which steps one scene per process. */
pbfx_status pbfx_batch_graph_capture(pbfx_ctx* ctx, float* x, float* v, uint32_t* offsets, int iters, uint32_t steps,
                                     void* stream);
pbfx_status pbfx_batch_graph_launch(pbfx_ctx* ctx, uint32_t launches, void* stream);

/* Replaces one scene's parameters (its grid must not change). */
pbfx_status pbfx_batch_set_params(pbfx_ctx* ctx, uint32_t scene, const pbfx_params* params);

/* Views of the last neighbour search, for bit-exact checks.  Replaces the
 * neighbor_search getters (pbf/interaction/pbf_neighbor_search.h:49-56).
 * Device pointers owned by the context, valid until the next step. */
typedef struct pbfx_ns_view {
    uint32_t grid[3];
    uint32_t num_cells;          /* C = grid[0]*grid[1]*grid[2] */
    float cell_width;
    uint32_t max_pair;
    const uint32_t* sorted_index;/* [num_in] slot -> index before the sort */
    const uint32_t* cell_offset; /* [C+2] exclusive scan of the cell counts; cell c owns
                                    slots [off[c], off[c+1]); off[C] = live count */
    const uint32_t* hash;        /* [num_in] key per index before the sort (C = deleted) */
} pbfx_ns_view;
pbfx_status pbfx_ns_view_get(pbfx_ctx* ctx, pbfx_ns_view* out);

/* Writes the last step's neighbour lists as sorted slot indices into
 * `out` (device, [num_live * max_pair], row per particle, PBFX_EMPTY
 * padded) and counts into `counts` (device, [num_live]); the reference
 * stores the same sets interleaved (pbf_neighbor_search_util.cu:471). */
pbfx_status pbfx_ns_neighbors(pbfx_ctx* ctx, uint32_t* out, uint32_t* counts, void* stream);

/* Writes cell_start / cell_end in the reference's layout
 * (pbf_neighbor_search_util.cu:39-81: empty cells get PBFX_EMPTY in start;
 * here end is 0 for empty cells).  Device buffers of C entries. */
pbfx_status pbfx_ns_cell_ranges(pbfx_ctx* ctx, uint32_t* cell_start, uint32_t* cell_end, void* stream);

/* Writes the scaling factor s (= -lambda of the paper) of the last Jacobi
 * iteration, in sorted order, into `out` (device float [num_live]). */
pbfx_status pbfx_scaling_factor(pbfx_ctx* ctx, float* out, void* stream);

/* Density of each live particle over the last step's neighbour lists,
 * rho_i = sum_j m * W(|x_i - x_j|) (calcDensity, pbf/solver/cuda/pbf_density.cu:11-99,
 * with inv_h in the kernel where the reference passes h, pbf_density.cu:22).
 * x: device AoS float3 in the last step's sorted order (its output);
 * rho: device float [num_live]. */
pbfx_status pbfx_density(pbfx_ctx* ctx, const float* x, float* rho, void* stream);

/* Writes the last step's sorted output back to the order before its sort:
 * x_out[sorted_index[s]] = x_sorted[s] (same for v) for every live slot s.
 * Replaces neighbor_search::restoreOrder (pbf/interaction/pbf_neighbor_search.cpp:194-200,
 * restoreOrderCUDA pbf_neighbor_search_util.cu:156-170).  Device AoS float3;
 * x_out, v_out hold num_in particles (entries of deleted particles are left as they were). */
pbfx_status pbfx_restore_order(pbfx_ctx* ctx, const float* x_sorted, const float* v_sorted, float* x_out,
                               float* v_out, void* stream);

typedef struct pbfx_stats {
    uint32_t num_in;           /* particles entering the last step */
    uint32_t num_live;         /* particles after deletion */
    uint32_t truncated;        /* particles that lost in-range neighbours to the slot cap */
    uint32_t sweep_cap_hits;   /* particle-iterations whose collision loop hit max_sweeps */
    uint32_t max_sweeps_used;
    uint32_t xsph_nan;         /* particles whose XSPH sum was NaN (velocity kept) */
    uint32_t units;            /* work units (cell blocks) of the last step */
    uint32_t slow_units;       /* units whose neighbourhood exceeded the on-chip tile */
    uint32_t slow_particles;   /* final-pass particles filtered from global memory (sparse runs, large tiles) */
    uint32_t tile_overflow;    /* reserved, always 0 (neighbourhoods of any size are searched) */
    uint32_t max_tile;         /* largest neighbourhood tile seen */
    uint32_t steps;            /* steps run by this context */
    uint32_t final_units;      /* work units of the final passes (final-cell runs) */
    uint32_t list_overflow;    /* particles with more than 64 in-range final pairs (summed cell by cell) */
    uint32_t scan_tiles;       /* occupied 4096-cell tiles of the cell scan (the others are not read) */
    uint32_t heavy_builds;     /* final-pass lists of more than 2048 candidates, built by a wave each */
    uint32_t heavy_search;     /* search lists of more than 2048 candidates, built by a wave each */
} pbfx_stats;

/* Statistics of the last step (synchronises `stream`). */
pbfx_status pbfx_get_stats(pbfx_ctx* ctx, void* stream, pbfx_stats* out);

/* Tuning knobs.  tile_cap: particles held on chip per work unit (0 forces
 * the global-memory path everywhere, for testing; default 2048, max 4096);
 * unit_width: max cells per work unit along x (power of two 1..32, default 16). */
pbfx_status pbfx_set_tuning(pbfx_ctx* ctx, uint32_t tile_cap, uint32_t unit_width);

/* Autotuner: times one step of the given state for every candidate
 * (tile_cap in {1024, 2048, 3072, 4096} x unit_width in {4, 8, 16, 32}, within
 * the grids' limits), best of four runs in two sweeps, on the scratch copies x_scratch /
 * v_scratch (x and v are only read), applies the fastest and returns it.  A single-scene context
 * steps n particles with area [lo, hi]; a batch steps the scenes described
 * by the device offsets in_offsets (lo, hi ignored).  No counterpart in the
 * reference. */
pbfx_status pbfx_autotune(pbfx_ctx* ctx, const float* x, const float* v, uint32_t n, const uint32_t* in_offsets,
                          const float lo[3], const float hi[3], int iters, float* x_scratch, float* v_scratch,
                          uint32_t* tile_cap_out, uint32_t* unit_width_out, void* stream);

/* When non-zero, pbfx_step returns PBFX_NEIGHBOR_TRUNCATED / PBFX_NOT_CONVERGED
 * after a step in which those events happened (default 0: they are only
 * counted in pbfx_stats, matching the reference's silent behaviour). */
void pbfx_set_report_info(pbfx_ctx* ctx, int enable);

/* Grid the context uses (calcGridSize(2h, domain_end)). */
pbfx_status pbfx_get_grid(pbfx_ctx* ctx, uint32_t grid[3]);

/* pbf/pbf_type.cpp:141-162 as a free function: 0 on success,
 * PBFX_GRID_TOO_LARGE past 1028 cells per axis. */
pbfx_status pbfx_calc_grid_size(double cell_size, const float domain_end[3], uint32_t grid[3]);

/* Last HIP error code recorded by the context (0 if none). */
int pbfx_last_hip_error(pbfx_ctx* ctx);

const char* pbfx_status_string(pbfx_status s);

/* Per-kernel timing with HIP events recorded on the step's stream around
 * every launch (off by default; at most 64 launches between reads, i.e. read
 * after each step).  No counterpart in the reference, which has no timers;
 * used by the benchmark to report each kernel's duration. */
typedef struct pbfx_kernel_time {
    char name[32];
    double total_ms;           /* accumulated since timing was enabled */
    uint32_t launches;
} pbfx_kernel_time;
pbfx_status pbfx_set_kernel_timing(pbfx_ctx* ctx, int enable);
/* Synchronises `stream`, folds the recorded launches into the totals and
 * writes up to `max` entries; *count = number of kernels with launches. */
pbfx_status pbfx_kernel_timing_read(pbfx_ctx* ctx, void* stream, pbfx_kernel_time* out, uint32_t max,
                                    uint32_t* count);

/* ---- Differentiable form.  No counterpart in the reference: its step has no
 * adjoint.  Gradients of a scalar loss L of the state after `steps` plain
 * steps of a single-scene context, with respect to the initial positions and
 * velocities, the XSPH coefficient c (xsph_parameter) and the vorticity
 * epsilon (vc_parameter).  The step is piecewise smooth; the gradient is that
 * of the smooth map left by the forward run's discrete decisions (sort and
 * frozen neighbour lists, boundary responses that fired, final-pass cells,
 * XSPH NaN guard).
 *
 * pbfx_adj_enable reserves, for up to max_steps steps, a checkpoint of the
 * state every `every` steps and the tape of one window of `every` steps
 * (about (200 + 16 iters) bytes per particle per step of the window).
 * Requires every grid axis to have at least 4 cells. */
pbfx_status pbfx_adj_enable(pbfx_ctx* ctx, uint32_t max_steps, uint32_t every);

/* Runs `steps` steps from (x, v, *num) as pbfx_step does (device AoS float3,
 * overwritten with the final state, *num updated), keeping the checkpoints
 * and the identity of every final slot. */
pbfx_status pbfx_adj_forward(pbfx_ctx* ctx, float* x, float* v, uint32_t* num, const float lo[3],
                             const float hi[3], int iters, uint32_t steps, void* stream);

/* id[s] = index in the initial state of the particle in final slot s
 * (device uint32 [num after the forward]). */
pbfx_status pbfx_adj_identity(pbfx_ctx* ctx, uint32_t* id, void* stream);

/* Backward pass: gx, gv = dL/d(final x, v) (device AoS double3, final
 * order); gx0, gv0 = dL/d(initial x, v) (device AoS double3, initial order);
 * gparam[0] = dL/dc, gparam[1] = dL/d(vorticity epsilon) (host).  Every
 * window of steps is recomputed from its checkpoint with its tape recorded,
 * then the adjoint kernels run through it in reverse, in FP64 at the FP32
 * forward's values.  The adjoint carries a power-of-two scale: the true
 * gradients are the returned values times 2^(*exponent) (0 while the
 * adjoint stays below 2^20).  gx and gv are not modified. */
pbfx_status pbfx_adj_backward(pbfx_ctx* ctx, const double* gx, const double* gv, double* gx0, double* gv0,
                              double gparam[2], int32_t* exponent, void* stream);

/* ---- Allocation-free step.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  One plain step of a
 * single scene, as pbfx_create + pbfx_set_boundary + pbfx_step_async would run
 * it (same kernels, same results bit for bit), with no context: every scratch
 * array lives in a static device array of PBFX_STATIC_SCRATCH_BYTES that the
 * library holds on each GPU it runs on until the process ends.  Nothing is
 * allocated, by the library or any routine it calls.
 *  params, domain_end  as for pbfx_create (grid = calcGridSize(2h, domain_end))
 *  inner_spheres4 .. plane_normals3  the boundary, as for pbfx_set_boundary but
 *               DEVICE pointers (counts may be zero)
 *  x, v, num, lo, hi, iters, stream  as for pbfx_step_async (the count after
 *               deletion is not reported: pass an area that deletes nothing,
 *               or use a context)
 * Scratch needed: about (417 + 4 * stride) bytes per particle, stride =
 * max_pair - 1 rounded up to a multiple of 4 (545 B at 30 slots), plus 16 bytes
 * per grid cell (the grid padded to a power of two or to 4096 cells) and
 * ~1.8 MB; PBFX_CAPACITY when that exceeds PBFX_STATIC_SCRATCH_BYTES (e.g. the
 * 32,768-particle dam on its 32 x 32 x 16 grid needs ~19.9 MB).  Every call on
 * a GPU shares that GPU's array: calls that overlap on one GPU (concurrent
 * streams or host threads) race and must be serialised by the caller.  The
 * array keeps the last call's contents; each call clears what must start at
 * zero with kernels (no memset), so a call can be captured in a graph. */
#define PBFX_STATIC_SCRATCH_BYTES (32u << 20)
pbfx_status pbfx_step_static(const pbfx_params* params, const float domain_end[3], const float* inner_spheres4,
                             uint32_t n_inner, const float* outer_spheres4, uint32_t n_outer,
                             const float* plane_points3, const float* plane_normals3, uint32_t n_planes, float* x,
                             float* v, uint32_t num, const float lo[3], const float hi[3], int iters, void* stream);

/* Arithmetic self-test on the device: for n inputs computes
 * out[4*i+0] = a/b, out[4*i+1] = sqrtf(a), out[4*i+2] = fmaf(a,b,a),
 * out[4*i+3] = floorf(a/b), bitwise comparable with IEEE host results.
 * Device pointers. */
pbfx_status pbfx_selftest_arith(const float* a, const float* b, float* out, uint32_t n, void* stream);

/* Self-test of the solver's 1/sqrt: over every float whose bits lie in
 * [lo_bits, hi_bits], compares the kernels' inv_sqrt (v_sqrt_f32 and v_rcp_f32
 * with correction steps) with 1.f / sqrtf in IEEE arithmetic, adds the number
 * of inputs whose results differ in any bit to *mismatches and, if there are
 * any, stores one such input's bits in *first_bits.  Device pointers; the
 * caller zeroes *mismatches.  No counterpart in the reference (diagnostic). */
pbfx_status pbfx_selftest_inv_sqrt(uint32_t lo_bits, uint32_t hi_bits, unsigned long long* mismatches,
                                   uint32_t* first_bits, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PBFX_H */
