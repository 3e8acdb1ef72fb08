// pbfx_api.hip -- C-ABI of the pbfx solver (see include/pbfx.h): context
// lifetime, host-side constants, step orchestration and views.
#include "pbfx.h"
#include "pbfx_internal.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>

using namespace pbfx;

struct pbfx_ctx {
    int device = 0;
    pbfx_params params{};
    uint32_t capacity = 0;
    float domain_end[3] = {0, 0, 0};
    Grid G{};
    Bufs b{};
    Tuning T{};
    int num_cu = 0;
    int blocks_per_cu = 1;
    float4* d_inner = nullptr;
    uint32_t n_inner = 0;
    float4* d_outer = nullptr;
    uint32_t n_outer = 0;
    float* d_pts = nullptr;
    float* d_nrm = nullptr;
    uint32_t n_planes = 0;
    float4* d_default_sphere = nullptr;  // pbf_constraint.cu:300: centre (3,3,3), radius 3
    uint32_t* h_pinned = nullptr;        // kCtrWords counters + live + in
    int last_err = 0;
    int report_info = 0;
    uint32_t steps = 0;
    uint32_t list_stride = 0;
    Timer* timer = nullptr;          // per-launch events when kernel timing is on
    double kt_ms[K_NUM] = {};
    uint32_t kt_n[K_NUM] = {};
};

namespace {

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

// Squared distance beyond which a kernel evaluated with argument `arg`
// (q = r * arg) is exactly zero: (1/arg)^2 with a 1e-5 relative margin, far
// above the rounding of r and q, so the pairs skipped add exactly +-0.
float support_bound2(float arg) {
    const double L = 1.0 / (double)arg;
    return (float)(L * L * (1.0 + 1e-5));
}

// pbf/kernel/cuda/detail/pbf_kernel.inl:34-48 on the host (IEEE float)
float wpoly6_host(float r, float inv_h) {
    float q = r * inv_h;
    float alpha = 1.56668147065f * inv_h * inv_h * inv_h;
    if (q <= 1.f) {
        float q2 = q * q;
        float t = 1.f - q2;
        return alpha * t * t * t;
    }
    return 0.f;
}

uint32_t ilog2(uint32_t v) {
    uint32_t r = 0;
    while ((1u << r) < v) ++r;
    return r;
}

bool is_pow2(uint32_t v) { return v && !(v & (v - 1)); }

pbfx_status check_params(const pbfx_params* p) {
    if (!p) return PBFX_INVALID_ARG;
    if (!(p->stable_density > 0.f) || !(p->particle_mass > 0.f) || !(p->smoothing_length > 0.f) ||
        !(p->time_step > 0.f))
        return PBFX_INVALID_ARG;  // pbf_type.cpp:167-170 exits on these
    if (p->max_pair < 2 || p->max_pair > 64) return PBFX_INVALID_ARG;
    if (p->scorr_mode != PBFX_SCORR_REFERENCE && p->scorr_mode != PBFX_SCORR_PAPER) return PBFX_INVALID_ARG;
    if (p->scorr_mode == PBFX_SCORR_PAPER && (p->scorr_n < 1 || !(p->scorr_dq_over_h > 0.f))) return PBFX_INVALID_ARG;
    if (p->max_sweeps < 1) return PBFX_INVALID_ARG;
    return PBFX_OK;
}

// Host constants of one_step / solveConstraint (pbf_solve.cpp:79-106,147-216)
// and setConstantMemory (pbf_position_update.cu:224-238), in IEEE float.
Phys make_phys(const pbfx_params& p) {
    Phys P{};
    P.h = p.smoothing_length;
    P.inv_h = 1.f / P.h;
    P.h2 = P.h * P.h;
    const float inv_h2 = P.inv_h * P.inv_h;
    P.a = 1.56668147065f * P.inv_h * inv_h2;
    P.ad = -14.3239448745f * inv_h2 * inv_h2;
    P.m = p.particle_mass;
    P.inv_rho0 = 1.f / p.stable_density;
    P.relax = p.relaxation;
    P.m2r2 = P.m * P.m * P.inv_rho0 * P.inv_rho0;
    const float w04 = wpoly6_host(P.h * 0.4f, P.inv_h);
    const float kf = (float)std::pow((double)w04, 4);
    P.inv_k = 1.f / kf;
    P.inv_wdq = (p.scorr_mode == PBFX_SCORR_PAPER) ? 1.f / wpoly6_host(p.scorr_dq_over_h * P.h, P.inv_h) : 0.f;
    P.sk = p.scorr_k;
    P.mode = p.scorr_mode;
    P.sn = p.scorr_n;
    P.dt = p.time_step;
    P.inv_dt = 1.f / P.dt;
    P.dtc = 1.f / P.inv_dt;
    P.g[0] = p.body_force[0];
    P.g[1] = p.body_force[1];
    P.g[2] = p.body_force[2];
    P.xsph_c = p.xsph_parameter;
    P.vc = p.vc_parameter;
    P.warg = p.vort_h_as_inv_h ? P.h : P.inv_h;
    P.cap = p.max_pair - 1;
    P.stride = P.cap;
    P.max_sweeps = p.max_sweeps;
    P.outer_as_inner = p.outer_as_inner;
    P.cut2_w = support_bound2(P.inv_h);
    P.cut2_wd = support_bound2(P.warg);
    return P;
}

template <class T>
hipError_t dalloc(T** p, size_t n) {
    return hipMalloc((void**)p, (n ? n : 1) * sizeof(T));
}

void free_timer(pbfx_ctx* c) {
    if (!c->timer) return;
    for (int i = 0; i < 2 * kTimerSlots; ++i) (void)hipEventDestroy(c->timer->ev[i]);
    delete c->timer;
    c->timer = nullptr;
}

void free_bufs(pbfx_ctx* c) {
    free_timer(c);
    Bufs& b = c->b;
    void* ptrs[] = {b.key,   b.rank,    b.tmp,     b.sidx,  b.Pf,     b.Q,     b.X,      b.R,
                    b.OM4,   b.ncnt,    b.nlist,   b.count, b.offset, b.tsum,  b.tprefix, b.units,
                    b.ckey,  b.crank,   b.cperm,   b.ccount, b.coffset, b.ctsum, b.ctprefix, b.cunits,
                    b.tflag, b.ctflag, b.flist, b.fcnt,
                    b.ctr,   c->d_inner, c->d_outer, c->d_pts, c->d_nrm, c->d_default_sphere};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (c->h_pinned) (void)hipHostFree(c->h_pinned);
    std::memset(&c->b, 0, sizeof(c->b));
    c->d_inner = c->d_outer = nullptr;
    c->d_pts = c->d_nrm = nullptr;
    c->d_default_sphere = nullptr;
    c->h_pinned = nullptr;
}

pbfx_status hip_fail(pbfx_ctx* c, hipError_t e) {
    if (c) c->last_err = (int)e;
    return PBFX_HIP_ERROR;
}

pbfx_status alloc_list(pbfx_ctx* c, uint32_t stride) {
    if (c->b.nlist && c->list_stride >= stride) return PBFX_OK;
    if (c->b.nlist) (void)hipFree(c->b.nlist);
    c->b.nlist = nullptr;
    hipError_t e = dalloc(&c->b.nlist, (size_t)c->capacity * stride);
    if (e != hipSuccess) return hip_fail(c, e);
    c->list_stride = stride;
    return PBFX_OK;
}

void set_grid_units(pbfx_ctx* c) {
    size_unit_grids(c->T, c->num_cu);
    c->blocks_per_cu = (int)(c->T.grid_units / (uint32_t)(c->num_cu > 0 ? c->num_cu : 1));
}

pbfx_status do_step(pbfx_ctx* c, float* x, float* v, uint32_t n, const float lo[3], const float hi[3], int iters,
                    const Boundary& B, hipStream_t s) {
    if (n > c->capacity) return PBFX_CAPACITY;
    if (iters < 1 || (n > 0 && (!x || !v)) || !lo || !hi) return PBFX_INVALID_ARG;
    Phys P = make_phys(c->params);
    if ((uint32_t)P.stride > c->list_stride) return PBFX_INVALID_ARG;
    hipError_t e = launch_step(c->G, P, B, c->b, c->T, x, v, n, lo, hi, iters, c->params.xsph_on,
                               c->params.vort_on, s, c->timer);
    if (e != hipSuccess) return hip_fail(c, e);
    c->steps++;
    return PBFX_OK;
}

Boundary user_boundary(const pbfx_ctx* c) {
    Boundary B{};
    B.inner = c->d_inner;
    B.n_inner = c->n_inner;
    B.outer = c->d_outer;
    B.n_outer = c->n_outer;
    B.pts = c->d_pts;
    B.nrm = c->d_nrm;
    B.n_planes = c->n_planes;
    return B;
}

pbfx_status read_counters(pbfx_ctx* c, hipStream_t s) {
    hipError_t e = hipMemcpyAsync(c->h_pinned, c->b.ctr, kCtrWords * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->h_pinned + kCtrWords, c->b.offset + c->G.C, 2 * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(c, e);
    // fold the statistics copies into h[0..CTR_NUM)
    uint32_t* h = c->h_pinned;
    for (int k = 0; k < CTR_NUM; ++k) {
        if (k == CTR_UNITS || k == CTR_CUNITS) continue;
        const bool is_max = (k == CTR_MAXTILE || k == CTR_MAXSWEEP);
        uint32_t v = 0;
        for (uint32_t r = 1; r <= kStatRep; ++r) {
            const uint32_t x = h[CTR_NUM * r + k];
            v = is_max ? (x > v ? x : v) : v + x;
        }
        h[k] = v;
    }
    return PBFX_OK;
}

pbfx_status sync_step_result(pbfx_ctx* c, uint32_t* num_inout, hipStream_t s) {
    pbfx_status st = read_counters(c, s);
    if (st != PBFX_OK) return st;
    *num_inout = c->h_pinned[kCtrWords];
    if (c->h_pinned[CTR_OVERFLOW]) return PBFX_TILE_OVERFLOW;
    if (c->report_info) {
        if (c->h_pinned[CTR_TRUNC]) return PBFX_NEIGHBOR_TRUNCATED;
        if (c->h_pinned[CTR_SWEEPCAP]) return PBFX_NOT_CONVERGED;
    }
    return PBFX_OK;
}

}  // namespace

extern "C" {

void pbfx_params_default(pbfx_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->scorr_mode = PBFX_SCORR_REFERENCE;
    p->scorr_k = 0.1f;
    p->scorr_n = 4;
    p->scorr_dq_over_h = 0.2f;
    p->max_pair = 30;
    p->xsph_on = 0;
    p->vort_on = 0;
    p->vort_h_as_inv_h = 0;
    p->outer_as_inner = 0;
    p->max_sweeps = 32;
}

pbfx_status pbfx_calc_grid_size(double cell_size, const float end[3], uint32_t grid[3]) {
    if (!end || !grid || !(cell_size > 0.0)) return PBFX_INVALID_ARG;
    uint32_t g[3] = {1, 1, 1};
    while (true) {
        if (g[0] * cell_size > end[0] && g[1] * cell_size > end[1] && g[2] * cell_size > end[2]) break;
        for (int a = 0; a < 3; ++a) g[a] = (g[a] * cell_size > end[a]) ? g[a] : g[a] * 2;
        if (g[0] > 1028 || g[1] > 1028 || g[2] > 1028) return PBFX_GRID_TOO_LARGE;
    }
    grid[0] = g[0];
    grid[1] = g[1];
    grid[2] = g[2];
    return PBFX_OK;
}

pbfx_status pbfx_create(const pbfx_params* params, uint32_t capacity, const float domain_end[3], pbfx_ctx** out) {
    if (!out || !domain_end) return PBFX_INVALID_ARG;
    *out = nullptr;
    pbfx_status st = check_params(params);
    if (st != PBFX_OK) return st;
    uint32_t grid[3];
    // pbf_type.cpp:172-174: cell width h, grid from calcGridSize(2h, domain_end)
    st = pbfx_calc_grid_size(params->smoothing_length * 2.0, domain_end, grid);
    if (st != PBFX_OK) return st;
    pbfx_ctx* c = new (std::nothrow) pbfx_ctx();
    if (!c) return PBFX_HIP_ERROR;
    (void)hipGetDevice(&c->device);
    c->params = *params;
    c->capacity = capacity;
    std::memcpy(c->domain_end, domain_end, sizeof(c->domain_end));
    Grid& G = c->G;
    G.gx = grid[0];
    G.gy = grid[1];
    G.gz = grid[2];
    G.lgx = ilog2(G.gx);
    G.lgy = ilog2(G.gy);
    G.C = G.gx * G.gy * G.gz;
    G.cw = params->smoothing_length;
    Bufs& b = c->b;
    b.cap = capacity > 0 ? capacity : 1;
    const uint32_t T = G.C < kScanTile ? G.C : kScanTile;
    b.lgT = ilog2(T);
    b.nT = G.C >> b.lgT;
    const size_t cap = b.cap;
    hipError_t e = hipSuccess;
#define PBFX_ALLOC(ptr, n)                           \
    if (e == hipSuccess) e = dalloc(&(ptr), (n));
    PBFX_ALLOC(b.key, cap);
    PBFX_ALLOC(b.rank, cap);
    PBFX_ALLOC(b.tmp, cap);
    PBFX_ALLOC(b.sidx, cap);
    PBFX_ALLOC(b.Pf, cap);
    PBFX_ALLOC(b.Q, cap);
    PBFX_ALLOC(b.X, 3 * cap);
    PBFX_ALLOC(b.R, 2 * cap);
    PBFX_ALLOC(b.OM4, cap);
    PBFX_ALLOC(b.ncnt, cap);
    PBFX_ALLOC(b.count, (size_t)G.C + 1);
    PBFX_ALLOC(b.offset, (size_t)G.C + 2);
    PBFX_ALLOC(b.tsum, (size_t)b.nT + 1);
    PBFX_ALLOC(b.tprefix, (size_t)b.nT + 1);
    const size_t items = cap + cap / kChunk + 64;  // a unit per own chunk of kChunk particles
    PBFX_ALLOC(b.units, items);
    PBFX_ALLOC(b.ckey, cap);
    PBFX_ALLOC(b.crank, cap);
    PBFX_ALLOC(b.cperm, cap);
    PBFX_ALLOC(b.ccount, (size_t)G.C + 1);
    PBFX_ALLOC(b.coffset, (size_t)G.C + 2);
    PBFX_ALLOC(b.ctsum, (size_t)b.nT + 1);
    PBFX_ALLOC(b.ctprefix, (size_t)b.nT + 1);
    PBFX_ALLOC(b.cunits, items);
    PBFX_ALLOC(b.flist, (size_t)kFinalK * cap);
    PBFX_ALLOC(b.fcnt, cap);
    PBFX_ALLOC(b.tflag, (size_t)b.nT + 1);
    PBFX_ALLOC(b.ctflag, (size_t)b.nT + 1);
    PBFX_ALLOC(b.ctr, kCtrWords);
    PBFX_ALLOC(c->d_default_sphere, 1);
#undef PBFX_ALLOC
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_pinned, (kCtrWords + 2) * sizeof(uint32_t), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMemset(b.count, 0, ((size_t)G.C + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b.offset, 0, ((size_t)G.C + 2) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b.tsum, 0, ((size_t)b.nT + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b.ccount, 0, ((size_t)G.C + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b.coffset, 0, ((size_t)G.C + 2) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b.ctsum, 0, ((size_t)b.nT + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b.tflag, 0, ((size_t)b.nT + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b.ctflag, 0, ((size_t)b.nT + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b.ctr, 0, kCtrWords * sizeof(uint32_t));
    if (e == hipSuccess) {
        const float4 sph = make_float4(3.f, 3.f, 3.f, 3.f);
        e = hipMemcpy(c->d_default_sphere, &sph, sizeof(sph), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, c->device);
    if (e != hipSuccess) {
        free_bufs(c);
        delete c;
        return PBFX_HIP_ERROR;
    }
    if (alloc_list(c, make_phys(*params).stride) != PBFX_OK) {
        free_bufs(c);
        delete c;
        return PBFX_HIP_ERROR;
    }
    c->T.tile_cap = 2048;
    c->T.unit_width = 16;
    set_grid_units(c);
    (void)hipDeviceSynchronize();
    *out = c;
    return PBFX_OK;
}

void pbfx_destroy(pbfx_ctx* c) {
    if (!c) return;
    DeviceGuard g(c->device);
    (void)hipDeviceSynchronize();
    free_bufs(c);
    delete c;
}

pbfx_status pbfx_set_params(pbfx_ctx* c, const pbfx_params* p) {
    if (!c) return PBFX_INVALID_ARG;
    pbfx_status st = check_params(p);
    if (st != PBFX_OK) return st;
    if (p->smoothing_length != c->params.smoothing_length) {
        uint32_t grid[3];
        st = pbfx_calc_grid_size(p->smoothing_length * 2.0, c->domain_end, grid);
        if (st != PBFX_OK) return st;
        if (grid[0] != c->G.gx || grid[1] != c->G.gy || grid[2] != c->G.gz) return PBFX_INVALID_ARG;
        c->G.cw = p->smoothing_length;
    }
    DeviceGuard g(c->device);
    st = alloc_list(c, make_phys(*p).stride);
    if (st != PBFX_OK) return st;
    c->params = *p;
    set_grid_units(c);
    return PBFX_OK;
}

pbfx_status pbfx_set_boundary(pbfx_ctx* c, const float* inner4, uint32_t n_inner, const float* outer4,
                              uint32_t n_outer, const float* pts3, const float* nrm3, uint32_t n_planes) {
    if (!c) return PBFX_INVALID_ARG;
    if ((n_inner && !inner4) || (n_outer && !outer4) || (n_planes && (!pts3 || !nrm3))) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    (void)hipDeviceSynchronize();
    if (c->d_inner) (void)hipFree(c->d_inner);
    if (c->d_outer) (void)hipFree(c->d_outer);
    if (c->d_pts) (void)hipFree(c->d_pts);
    if (c->d_nrm) (void)hipFree(c->d_nrm);
    c->d_inner = c->d_outer = nullptr;
    c->d_pts = c->d_nrm = nullptr;
    c->n_inner = c->n_outer = c->n_planes = 0;
    hipError_t e = hipSuccess;
    if (n_inner) {
        e = dalloc(&c->d_inner, n_inner);
        if (e == hipSuccess) e = hipMemcpy(c->d_inner, inner4, n_inner * sizeof(float4), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && n_outer) {
        e = dalloc(&c->d_outer, n_outer);
        if (e == hipSuccess) e = hipMemcpy(c->d_outer, outer4, n_outer * sizeof(float4), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && n_planes) {
        e = dalloc(&c->d_pts, 3 * (size_t)n_planes);
        if (e == hipSuccess) e = dalloc(&c->d_nrm, 3 * (size_t)n_planes);
        if (e == hipSuccess) e = hipMemcpy(c->d_pts, pts3, 3 * n_planes * sizeof(float), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(c->d_nrm, nrm3, 3 * n_planes * sizeof(float), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) return hip_fail(c, e);
    c->n_inner = n_inner;
    c->n_outer = n_outer;
    c->n_planes = n_planes;
    return PBFX_OK;
}

pbfx_status pbfx_step_async(pbfx_ctx* c, float* x, float* v, uint32_t num, const float lo[3], const float hi[3],
                            int iters, void* stream) {
    if (!c) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    return do_step(c, x, v, num, lo, hi, iters, user_boundary(c), (hipStream_t)stream);
}

pbfx_status pbfx_step(pbfx_ctx* c, float* x, float* v, uint32_t* num_inout, const float lo[3], const float hi[3],
                      int iters, void* stream) {
    if (!c || !num_inout) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    pbfx_status st = do_step(c, x, v, *num_inout, lo, hi, iters, user_boundary(c), (hipStream_t)stream);
    if (st != PBFX_OK) return st;
    return sync_step_result(c, num_inout, (hipStream_t)stream);
}

pbfx_status pbfx_step_no_boundary(pbfx_ctx* c, float* x, float* v, uint32_t* num_inout, const float lo[3],
                                  const float hi[3], int iters, void* stream) {
    if (!c || !num_inout) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    Boundary B{};
    B.inner = c->d_default_sphere;
    B.n_inner = 1;
    pbfx_status st = do_step(c, x, v, *num_inout, lo, hi, iters, B, (hipStream_t)stream);
    if (st != PBFX_OK) return st;
    return sync_step_result(c, num_inout, (hipStream_t)stream);
}

pbfx_status pbfx_get_num(pbfx_ctx* c, void* stream, uint32_t* num) {
    if (!c || !num) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    pbfx_status st = read_counters(c, (hipStream_t)stream);
    if (st != PBFX_OK) return st;
    *num = c->h_pinned[kCtrWords];
    return PBFX_OK;
}

pbfx_status pbfx_get_stats(pbfx_ctx* c, void* stream, pbfx_stats* out) {
    if (!c || !out) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    pbfx_status st = read_counters(c, (hipStream_t)stream);
    if (st != PBFX_OK) return st;
    const uint32_t* h = c->h_pinned;
    out->num_in = h[kCtrWords + 1];
    out->num_live = h[kCtrWords];
    out->truncated = h[CTR_TRUNC];
    out->sweep_cap_hits = h[CTR_SWEEPCAP];
    out->max_sweeps_used = h[CTR_MAXSWEEP];
    out->xsph_nan = h[CTR_XSPHNAN];
    out->units = h[CTR_UNITS];
    out->slow_units = h[CTR_SLOWUNITS];
    out->slow_particles = h[CTR_SLOWPART];
    out->tile_overflow = h[CTR_OVERFLOW];
    out->max_tile = h[CTR_MAXTILE];
    out->steps = c->steps;
    out->final_units = h[CTR_CUNITS];
    out->list_overflow = h[CTR_LISTOVF];
    return PBFX_OK;
}

pbfx_status pbfx_ns_view_get(pbfx_ctx* c, pbfx_ns_view* out) {
    if (!c || !out) return PBFX_INVALID_ARG;
    out->grid[0] = c->G.gx;
    out->grid[1] = c->G.gy;
    out->grid[2] = c->G.gz;
    out->num_cells = c->G.C;
    out->cell_width = c->G.cw;
    out->max_pair = (uint32_t)c->params.max_pair;
    out->sorted_index = c->b.sidx;
    out->cell_offset = c->b.offset;
    out->hash = c->b.key;
    return PBFX_OK;
}

pbfx_status pbfx_ns_neighbors(pbfx_ctx* c, uint32_t* out, uint32_t* counts, void* stream) {
    if (!c || !out || !counts) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    hipError_t e = launch_neighbors_view(c->G, make_phys(c->params), c->b, c->T, out, counts,
                                         (uint32_t)c->params.max_pair, (hipStream_t)stream);
    return e == hipSuccess ? PBFX_OK : hip_fail(c, e);
}

pbfx_status pbfx_ns_cell_ranges(pbfx_ctx* c, uint32_t* start, uint32_t* end, void* stream) {
    if (!c || !start || !end) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    hipError_t e = launch_cell_ranges(c->G, c->b, start, end, (hipStream_t)stream);
    return e == hipSuccess ? PBFX_OK : hip_fail(c, e);
}

pbfx_status pbfx_scaling_factor(pbfx_ctx* c, float* out, void* stream) {
    if (!c || !out) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    hipError_t e = launch_copy_lambda(c->G, c->b, out, (hipStream_t)stream);
    return e == hipSuccess ? PBFX_OK : hip_fail(c, e);
}

pbfx_status pbfx_density(pbfx_ctx* c, const float* x, float* rho, void* stream) {
    if (!c || !x || !rho) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    hipError_t e = launch_density(c->G, make_phys(c->params), c->b, x, rho, (hipStream_t)stream);
    return e == hipSuccess ? PBFX_OK : hip_fail(c, e);
}

pbfx_status pbfx_restore_order(pbfx_ctx* c, const float* x_sorted, const float* v_sorted, float* x_out,
                               float* v_out, void* stream) {
    if (!c || !x_sorted || !v_sorted || !x_out || !v_out) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    hipError_t e = launch_restore(c->G, c->b, x_sorted, v_sorted, x_out, v_out, (hipStream_t)stream);
    return e == hipSuccess ? PBFX_OK : hip_fail(c, e);
}

pbfx_status pbfx_set_tuning(pbfx_ctx* c, uint32_t tile_cap, uint32_t unit_width) {
    if (!c) return PBFX_INVALID_ARG;
    if (tile_cap > kTileCapMax || !is_pow2(unit_width) || unit_width > kMaxW) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    c->T.tile_cap = tile_cap;
    c->T.unit_width = unit_width;
    set_grid_units(c);
    return PBFX_OK;
}

void pbfx_set_report_info(pbfx_ctx* c, int enable) {
    if (c) c->report_info = enable;
}

pbfx_status pbfx_get_grid(pbfx_ctx* c, uint32_t grid[3]) {
    if (!c || !grid) return PBFX_INVALID_ARG;
    grid[0] = c->G.gx;
    grid[1] = c->G.gy;
    grid[2] = c->G.gz;
    return PBFX_OK;
}

int pbfx_last_hip_error(pbfx_ctx* c) { return c ? c->last_err : 0; }

const char* pbfx_status_string(pbfx_status s) {
    switch (s) {
        case PBFX_OK: return "ok";
        case PBFX_INVALID_ARG: return "invalid argument";
        case PBFX_CAPACITY: return "particle count exceeds capacity";
        case PBFX_GRID_TOO_LARGE: return "grid dimension exceeds 1028";
        case PBFX_HIP_ERROR: return "HIP runtime error";
        case PBFX_NOT_CONVERGED: return "collision sweep cap reached";
        case PBFX_NEIGHBOR_TRUNCATED: return "neighbour list truncated";
        case PBFX_TILE_OVERFLOW: return "neighbourhood tile overflow";
        case PBFX_NOT_SUPPORTED: return "not supported";
    }
    return "unknown status";
}

pbfx_status pbfx_set_kernel_timing(pbfx_ctx* c, int enable) {
    if (!c) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    (void)hipDeviceSynchronize();
    free_timer(c);
    std::memset(c->kt_ms, 0, sizeof(c->kt_ms));
    std::memset(c->kt_n, 0, sizeof(c->kt_n));
    if (!enable) return PBFX_OK;
    Timer* t = new (std::nothrow) Timer();
    if (!t) return PBFX_HIP_ERROR;
    for (int i = 0; i < 2 * kTimerSlots; ++i) {
        hipError_t e = hipEventCreate(&t->ev[i]);
        if (e != hipSuccess) {
            for (int j = 0; j < i; ++j) (void)hipEventDestroy(t->ev[j]);
            delete t;
            return hip_fail(c, e);
        }
    }
    t->n = 0;
    c->timer = t;
    return PBFX_OK;
}

pbfx_status pbfx_kernel_timing_read(pbfx_ctx* c, void* stream, pbfx_kernel_time* out, uint32_t max,
                                    uint32_t* count) {
    if (!c || !count || (max && !out)) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    if (c->timer) {
        hipError_t e = hipStreamSynchronize((hipStream_t)stream);
        if (e != hipSuccess) return hip_fail(c, e);
        for (int i = 0; i < c->timer->n; ++i) {
            float ms = 0.f;
            e = hipEventElapsedTime(&ms, c->timer->ev[2 * i], c->timer->ev[2 * i + 1]);
            if (e != hipSuccess) return hip_fail(c, e);
            c->kt_ms[c->timer->id[i]] += ms;
            c->kt_n[c->timer->id[i]] += 1;
        }
        c->timer->n = 0;
    }
    uint32_t k = 0;
    for (int id = 0; id < K_NUM; ++id) {
        if (!c->kt_n[id]) continue;
        if (k < max) {
            std::memset(out[k].name, 0, sizeof(out[k].name));
            std::strncpy(out[k].name, kernel_name(id), sizeof(out[k].name) - 1);
            out[k].total_ms = c->kt_ms[id];
            out[k].launches = c->kt_n[id];
        }
        ++k;
    }
    *count = k;
    return PBFX_OK;
}

pbfx_status pbfx_selftest_arith(const float* a, const float* b, float* out, uint32_t n, void* stream) {
    if ((n && (!a || !b || !out))) return PBFX_INVALID_ARG;
    hipError_t e = launch_selftest(a, b, out, n, (hipStream_t)stream);
    return e == hipSuccess ? PBFX_OK : PBFX_HIP_ERROR;
}

}  // extern "C"
