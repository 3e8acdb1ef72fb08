/*
 * pbf_oracle.h -- CPU restatement of the reference PBF solver step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP solver
 * (libpbfx.so).  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it; the solver product never links or calls it.
 *
 * It restates, in host C++ with IEEE FP32 arithmetic and the reference's
 * operation order, the step driven by pbf::one_step
 * (cuda_pbf_with_ssf/src/pbf/solver/pbf_solve.cpp:182-216), with an FP64
 * twin that replays the same discrete decisions (sort, neighbour lists,
 * cell lookups) in double precision to state the FP32 rounding error.
 *
 * Deliberate deviations from the shipped reference (all documented in the
 * repository design notes):
 *   - stable counting sort (ties by original index) instead of the
 *     atomicAdd-order ranks of pbf/util/cuda/pbf_counting_sort.cu:21-32;
 *   - deleted particles use the sentinel key C = Gx*Gy*Gz, not the
 *     gz*gy*gz-1 cell of pbf/interaction/pbf_neighbor_search.cpp:146;
 *   - 1/sqrtf(x) (correctly rounded) where the reference calls rsqrtf;
 *   - the collision sweep loop is capped (max_sweeps) instead of unbounded;
 *   - vorticity uses inv_h in the spiky derivative (the reference passes h,
 *     pbf/solver/cuda/pbf_vorticity_confinement.cu:28,134) unless
 *     vort_h_as_inv_h is set;
 *   - outer spheres use the outer-sphere response (the reference calls the
 *     inner one, pbf/solver/cuda/pbf_response_collision.cu:46-51) unless
 *     outer_as_inner is set.
 */
#ifndef PBF_ORACLE_H
#define PBF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_EMPTY 0xFFFFFFFFu

enum { ORC_SCORR_REFERENCE = 0, ORC_SCORR_PAPER = 1 };

typedef struct orc_params {
    /* pbf_parameter (pbf/pbf_type.h:49-63) */
    float stable_density;
    float stable_distance;
    float particle_mass;
    float smoothing_length;   /* h; also the hash cell width */
    float time_step;
    float relaxation;         /* epsilon of the lambda denominator */
    float xsph_parameter;
    float vc_parameter;
    /* pbf_external (pbf/pbf_type.h:69-72) */
    float body_force[3];
    /* tensile correction */
    int32_t scorr_mode;       /* ORC_SCORR_REFERENCE or ORC_SCORR_PAPER */
    float scorr_k;
    int32_t scorr_n;
    float scorr_dq_over_h;
    /* neighbour slots per particle; the reference uses 30 (29 usable) */
    int32_t max_pair;
    /* update stage switches: both 0 reproduces the shipped update
       (pbf_solve.cpp:125-145); both 1 is the commented full order */
    int32_t xsph_on;
    int32_t vort_on;
    int32_t vort_h_as_inv_h;
    int32_t outer_as_inner;
    int32_t max_sweeps;       /* collision sweep cap per Jacobi iteration */
    uint32_t grid[3];         /* power-of-two grid from calc_grid_size */
} orc_params;

typedef struct orc_boundary {
    uint32_t n_inner;
    const float* inner;       /* 4 floats each: center xyz, radius */
    uint32_t n_outer;
    const float* outer;
    uint32_t n_planes;
    const float* plane_points;  /* 3 floats each */
    const float* plane_normals; /* 3 floats each */
} orc_boundary;

/* Optional views of the neighbour search; any pointer may be NULL. */
typedef struct orc_ns_out {
    uint32_t* key;           /* [n_in]  hash key per ORIGINAL particle (sentinel C if deleted) */
    uint32_t* sorted_index;  /* [n_in]  slot -> original particle */
    uint32_t* cell_offset;   /* [C+2]   exclusive scan of cell counts (sentinel cell C last) */
    uint32_t* neighbors;     /* [n_in*max_pair] sorted neighbour slots, row per particle, ORC_EMPTY padded */
    uint32_t* neighbor_count;/* [n_in] */
    float* pred0;            /* [3*n_in] sorted predicted positions (AoS) frozen for the list */
    float* lambda_last;      /* [n_in]  scaling factor s of the last Jacobi iteration */
} orc_ns_out;

typedef struct orc_stats {
    uint32_t num_in;
    uint32_t num_live;
    uint32_t num_deleted;
    uint32_t truncated;       /* particles that lost at least one neighbour to the cap */
    uint32_t sweep_cap_hits;  /* particle-iterations that hit max_sweeps */
    uint32_t max_sweeps_used;
    uint32_t xsph_nan;        /* particles whose XSPH sum was NaN (velocity kept) */
    uint32_t reserved;
} orc_stats;

/* pbf_type.cpp:141-162 (calcGridSize).  Returns 0, or 1 if a dimension
   exceeds 1028. */
int orc_calc_grid_size(double cell_size, const float domain_end[3], uint32_t out[3]);

/* pbf/kernel/cuda/detail/pbf_kernel.inl:34-63 */
float orc_wpoly6(float r, float inv_h);
float orc_wspiky_deriv(float r, float inv_h);
double orc_wpoly6_f64(double r, double inv_h);
double orc_wspiky_deriv_f64(double r, double inv_h);

/* pbf_grid.h:20-28,54-61: key per position (n positions, AoS). */
void orc_hash(const float* x, uint32_t n, float cell_width, const uint32_t grid[3], uint32_t* key);

/* Tensile-correction constants as the host computes them:
   ref_inv_k = 1/(float)pow(W(0.4h),4) (pbf_solve.cpp:92, pbf_position_update.cu:224-238),
   paper_inv_wdq = 1/W(dq). */
void orc_scorr_constants(const orc_params* p, float* ref_inv_k, float* paper_inv_wdq);

/* One full step, FP32.  x, v: AoS [num*3], overwritten in the new sorted
   order (F10).  *num is shrunk by deletion.  Returns 0 on success. */
int orc_step_f32(const orc_params* p, const orc_boundary* b,
                 float* x, float* v, uint32_t* num,
                 const float lo[3], const float hi[3], int iters,
                 orc_ns_out* ns, orc_stats* st);

/* FP64 twin: same inputs (float), same discrete decisions as orc_step_f32,
   continuous arithmetic in double.  Writes x_out, v_out (double AoS). */
int orc_step_f64(const orc_params* p, const orc_boundary* b,
                 const float* x, const float* v, uint32_t num,
                 const float lo[3], const float hi[3], int iters,
                 double* x_out, double* v_out, uint32_t* num_out);

/* Density-constraint and energy metrics on a state (sorted or not):
   C_i = rho_i/rho0 - 1 with an uncapped 27-cell poly6 sum.
   out[0]=mean C, out[1]=max C, out[2]=mean max(C,0), out[3]=kinetic energy. */
int orc_metrics(const orc_params* p, const float* x, const float* v, uint32_t num, double out[4]);

/* Density over given neighbour lists (calcDensity, pbf/solver/cuda/pbf_density.cu:11-71,
 * inv_h in the kernel): neighbors is n x max_pair, ORC_EMPTY padded, as the
 * lists orc_step_f32 returns; x in the same (sorted) order. */
int orc_density(const orc_params* p, const float* x, uint32_t n, const uint32_t* neighbors, float* rho);

/* Host IEEE reference of pbfx_selftest_arith: out[4i..4i+3] =
   a/b, sqrtf(a), fmaf(a,b,a), floorf(a/b). */
void orc_arith_f32(const float* a, const float* b, float* out, uint32_t n);

/* ---- The step as a map with frozen discrete decisions, and its FP64 adjoint
 * (pbf_adjoint.cpp).  A tape holds, per step, the decisions of an FP32 run
 * (sort, frozen lists, boundary responses fired, final-pass cells, XSPH NaN
 * guard) and the values computed under them. */
typedef struct orc_tape orc_tape;
orc_tape* orc_tape_new(void);
void orc_tape_free(orc_tape* t);
uint32_t orc_tape_steps(const orc_tape* t);
/* which = 0: particles entering step `step`; 1: live after it. */
uint32_t orc_tape_num(const orc_tape* t, uint32_t step, int which);

/* One FP32 step, bit-identical to orc_step_f32, appended to the tape. */
int orc_adj_step_f32(const orc_params* p, const orc_boundary* b, float* x, float* v, uint32_t* num,
                     const float lo[3], const float hi[3], int iters, orc_tape* t);

/* Every recorded step again in FP64 from (x0, v0) (initial order) under the
 * tape's decisions.  coef: {XSPH coefficient, vorticity epsilon} in double
 * (NULL: the float ones of p).  record != 0 replaces the tape's values by the
 * FP64 ones.  x_out, v_out: the final state (live count of the last step,
 * final order); may be NULL. */
int orc_adj_replay_f64(const orc_params* p, const orc_boundary* b, orc_tape* t, const double* x0, const double* v0,
                       const double* coef, double* x_out, double* v_out, int record);

/* Adjoint through every recorded step, at the tape's values.  gx, gv:
 * dL/d(final x, v) in the final order; gx0, gv0: dL/d(x0, v0) in the initial
 * order; gparam[0] = dL/d xsph_parameter, gparam[1] = dL/d vc_parameter;
 * coef as for the replay.  The adjoint is kept near 1 by a power-of-two
 * scale: the true gradients are the outputs times 2^(*exponent) (may be NULL). */
int orc_adj_backward_f64(const orc_params* p, const orc_boundary* b, const orc_tape* t, const double* coef,
                         const double* gx, const double* gv, double* gx0, double* gv0, double gparam[2],
                         int32_t* exponent);

/* id[s] = initial index of the particle in final slot s. */
int orc_adj_identity(const orc_tape* t, uint32_t* id);

/* Number of OpenMP threads the oracle uses (0 = runtime default). */
void orc_set_threads(int n);
int orc_get_threads(void);

/* Planted mistakes.  The closed-form tests of the oracle switch one of these
 * on and assert that their check then FAILS, which shows the check would
 * catch that mistake in the restatement.  ORC_PLANT_NONE (the default) is the
 * faithful restatement; nothing but those tests may set another value. */
enum {
    ORC_PLANT_NONE = 0,
    ORC_PLANT_XSPH_SIGN = 1,   /* (v_i - v_j) where pbf_xsph.cu:23 has v_j - v_i */
    ORC_PLANT_XSPH_NO_MR = 2,  /* XSPH without the m * inv_rho0 of pbf_xsph.cu:110 */
    ORC_PLANT_XSPH_H = 3,      /* poly6 evaluated with h where pbf_xsph.cu:25-26 passes 1/h */
    ORC_PLANT_VORT_CROSS = 4,  /* cross(dir, v_j - v_i) where pbf_vorticity_confinement.cu:28 has cross(v_j - v_i, dir) */
    ORC_PLANT_VORT_NO_MR = 5,  /* vorticity without its m * inv_rho0 factor */
    ORC_PLANT_CONF_CROSS = 6,  /* cross(omega, N) where pbf_vorticity_confinement.cu:220 has cross(N, omega) */
    ORC_PLANT_CONF_RHO = 7,    /* times inv_rho0 where pbf_vorticity_confinement.cu:220 divides by it */
    ORC_PLANT_LOC_SELF = 8,    /* |omega_i| where pbf_vorticity_confinement.cu:129 takes |omega_j| */
    ORC_PLANT_NUDGE = 9,       /* the sphere responses' nudges 1.0001 / 0.9999 swapped (pbf_sphere_boundary.cu:19,47) */
    ORC_PLANT_CROSS_FRAC = 10  /* |pb_n| / (|pa_n| + |pb_n|) where pbf_plane_boundary.cu:23 has |pa_n| / (...) */
};
void orc_set_plant(int code);

#ifdef __cplusplus
}
#endif

#endif
