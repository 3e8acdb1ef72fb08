// pbfx_api.hip -- C-ABI of the pbfx solver (see include/pbfx.h): context
// lifetime, host-side constants, step orchestration and views.
#include "pbfx.h"
#include "pbfx_internal.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <type_traits>
#include <vector>

using namespace pbfx;

// One scene as the host keeps it: its parameters, grid, boundary arrays and area.
struct SceneHost {
    pbfx_params params{};
    uint32_t capacity = 0;
    float domain_end[3] = {0, 0, 0};
    Grid G{};
    float4* d_inner = nullptr;
    uint32_t n_inner = 0;
    float4* d_outer = nullptr;
    uint32_t n_outer = 0;
    float* d_pts = nullptr;
    float* d_nrm = nullptr;
    uint32_t n_planes = 0;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
};

// Differentiable form: checkpoints, the tape of one window of steps and the
// adjoint work arrays.
struct AdjState {
    uint32_t max_steps = 0, every = 0, iters = 0, steps = 0;
    uint32_t n0 = 0, n_final = 0;
    uint32_t stride = 0;              // list entries per particle the tapes hold
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    std::vector<float*> ck_x, ck_v;   // state at steps 0, every, 2*every, ... (device AoS)
    std::vector<uint32_t> ck_n;
    uint32_t* id[2] = {nullptr, nullptr};
    int id_cur = 0;
    std::vector<StepTape> tape;       // one window
    uint32_t tape_iters = 0;          // iterations the tapes' Q arrays hold
    float* xw = nullptr;              // recompute state
    float* vw = nullptr;
    AdjBufs a{};
    unsigned long long* h_amax = nullptr;  // pinned
    double* h_psum = nullptr;         // pinned [2]
    uint32_t* h_nh = nullptr;         // pinned [4]
};

struct pbfx_ctx {
    int device = 0;
    std::vector<SceneHost> sc;
    bool batch = false;                  // created by pbfx_batch_create
    uint32_t capacity = 0;               // all scenes
    uint32_t csum = 0;                   // cells of all scenes (kbase[n])
    uint32_t uw_max = kMaxW;             // narrowest grid row of any scene
    Bufs b{};
    Tuning T{};
    int num_cu = 0;
    int blocks_per_cu = 1;
    Scene* d_scenes = nullptr;           // scene table on the device
    Scene* h_scenes = nullptr;           // pinned staging of the table
    std::vector<Scene> last_table;       // what the device holds
    uint32_t* d_kbase = nullptr;
    uint32_t* d_in_off = nullptr;        // per-step input offsets (pbfx_batch_step)
    uint32_t* h_in_off = nullptr;        // pinned
    uint32_t* d_scene_off = nullptr;     // per-scene output offsets (pbfx_batch_step)
    uint32_t* h_scene_off = nullptr;     // pinned
    float4* d_default_sphere = nullptr;  // pbf_constraint.cu:300: centre (3,3,3), radius 3
    uint32_t* h_pinned = nullptr;        // kCtrWords counters + live + in
    int last_err = 0;
    int report_info = 0;
    uint32_t steps = 0;
    uint32_t list_stride = 0;
    Timer* timer = nullptr;              // per-launch events when kernel timing is on
    double kt_ms[K_NUM] = {};
    uint32_t kt_n[K_NUM] = {};
    AdjState* adj = nullptr;             // pbfx_adj_enable
    hipGraphExec_t graph = nullptr;      // pbfx_batch_graph_capture: one recorded step
    uint32_t graph_steps = 0;            // steps each launch of it advances
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
    P.stride = (P.cap + 3) & ~3;  // whole groups of four list entries
    P.max_sweeps = p.max_sweeps;
    P.outer_as_inner = p.outer_as_inner;
    P.cut2_w = support_bound2(P.inv_h);
    P.cut2_wd = support_bound2(P.warg);
    P.xsph_on = p.xsph_on;
    P.vort_on = p.vort_on;
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

void free_boundary(SceneHost& h) {
    void* ptrs[] = {h.d_inner, h.d_outer, h.d_pts, h.d_nrm};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    h.d_inner = h.d_outer = nullptr;
    h.d_pts = h.d_nrm = nullptr;
    h.n_inner = h.n_outer = h.n_planes = 0;
}

void free_tape_q(AdjState* A) {
    for (auto& t : A->tape)
        if (t.Q) {
            (void)hipFree(t.Q);
            t.Q = nullptr;
        }
    A->tape_iters = 0;
}

void free_adj(pbfx_ctx* c) {
    AdjState* A = c->adj;
    if (!A) return;
    free_tape_q(A);
    std::vector<void*> ptrs;
    for (float* p : A->ck_x) ptrs.push_back(p);
    for (float* p : A->ck_v) ptrs.push_back(p);
    for (auto& t : A->tape) {
        void* tp[] = {t.sidx, t.X, t.nlist, t.ncnt, t.offset, t.PI, t.R, t.OM4};
        for (void* p : tp) ptrs.push_back(p);
    }
    AdjBufs& a = A->a;
    void* ap[] = {A->id[0], A->id[1], A->xw, A->vw, a.ax, a.av, a.axin, a.avin, a.aP, a.aQ, a.aV, a.aVc, a.aOm,
                  a.etab, a.abar, a.coef, a.g2, a.aX, a.xnan, a.ckey, a.crank, a.ctmp, a.cperm, a.ccount,
                  a.coffset, a.ctsum, a.ctflag, a.ctprefix, a.gP, a.gV, a.gD, a.hflag, a.hlist, a.nh, a.last,
                  a.rcount, a.roff, a.rtsum, a.rtflag, a.rtprefix, a.rlist, a.pc, a.pe, a.psum, a.amax,
                  a.big, a.nbig};
    for (void* p : ap) ptrs.push_back(p);
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (A->h_amax) (void)hipHostFree(A->h_amax);
    if (A->h_psum) (void)hipHostFree(A->h_psum);
    if (A->h_nh) (void)hipHostFree(A->h_nh);
    delete A;
    c->adj = nullptr;
}

void free_bufs(pbfx_ctx* c) {
    free_adj(c);
    free_timer(c);
    Bufs& b = c->b;
    void* ptrs[] = {b.key,   b.rank,    b.tmp,     b.sidx,  b.Pf,     b.Q,     b.X,       b.R,
                    b.OM4,   b.ncnt,    b.nlist,   b.count, b.offset, b.tsum,  b.tprefix, b.units,
                    b.ckey,  b.crank,   b.cperm,   b.ccount, b.coffset, b.ctsum, b.ctprefix, b.cunits,
                    b.tflag, b.ctflag,  b.flist,   b.fcnt,  b.isid,   b.sid,   b.ctr,     b.big,
                    b.ovf,   b.oslot,   b.part,    b.bslot, b.aux,    b.bseg,
                    c->d_scenes, c->d_kbase, c->d_in_off, c->d_scene_off, c->d_default_sphere};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    void* pinned[] = {c->h_pinned, c->h_scenes, c->h_in_off, c->h_scene_off};
    for (void* p : pinned)
        if (p) (void)hipHostFree(p);
    for (auto& h : c->sc) free_boundary(h);
    std::memset(&c->b, 0, sizeof(c->b));
    c->d_scenes = nullptr;
    c->d_kbase = c->d_in_off = c->d_scene_off = nullptr;
    c->d_default_sphere = nullptr;
    c->h_pinned = c->h_in_off = c->h_scene_off = nullptr;
    c->h_scenes = nullptr;
}

pbfx_status hip_fail(pbfx_ctx* c, hipError_t e) {
    if (c) c->last_err = (int)e;
    return PBFX_HIP_ERROR;
}

pbfx_status alloc_list(pbfx_ctx* c, uint32_t stride) {
    if (c->b.nlist && c->list_stride >= stride) return PBFX_OK;
    if (c->b.nlist) (void)hipFree(c->b.nlist);
    c->b.nlist = nullptr;
    hipError_t e = dalloc(&c->b.nlist, (size_t)c->b.cap * stride);
    if (e != hipSuccess) return hip_fail(c, e);
    c->list_stride = stride;
    return PBFX_OK;
}

void set_grid_units(pbfx_ctx* c) {
    size_unit_grids(c->T, c->num_cu);
    c->blocks_per_cu = (int)(c->T.grid_units / (uint32_t)(c->num_cu > 0 ? c->num_cu : 1));
}

pbfx_status make_grid(const pbfx_params& p, const float end[3], Grid& G);

// The device table as the scenes stand now.
void build_table(pbfx_ctx* c, std::vector<Scene>& t, bool default_sphere) {
    t.resize(c->sc.size());
    for (size_t i = 0; i < c->sc.size(); ++i) {
        const SceneHost& h = c->sc[i];
        Scene S;
        std::memset(&S, 0, sizeof(S));
        S.G = h.G;
        S.P = make_phys(h.params);
        if (default_sphere) {
            S.B.inner = c->d_default_sphere;
            S.B.n_inner = 1;
        } else {
            S.B.inner = h.d_inner;
            S.B.n_inner = h.n_inner;
            S.B.outer = h.d_outer;
            S.B.n_outer = h.n_outer;
            S.B.pts = h.d_pts;
            S.B.nrm = h.d_nrm;
            S.B.n_planes = h.n_planes;
        }
        for (int a = 0; a < 3; ++a) {
            S.lo[a] = h.lo[a];
            S.hi[a] = h.hi[a];
        }
        t[i] = S;
    }
}

// Uploads the scene table when it differs from what the device holds (a
// parameter, boundary or area change); the copy completes before returning.
pbfx_status sync_table(pbfx_ctx* c, bool default_sphere, hipStream_t s) {
    std::vector<Scene> t;
    build_table(c, t, default_sphere);
    if (t.size() == c->last_table.size() &&
        std::memcmp(t.data(), c->last_table.data(), t.size() * sizeof(Scene)) == 0)
        return PBFX_OK;
    std::memcpy(c->h_scenes, t.data(), t.size() * sizeof(Scene));
    hipError_t e = hipMemcpyAsync(c->d_scenes, c->h_scenes, t.size() * sizeof(Scene), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(c, e);
    c->last_table = t;
    return PBFX_OK;
}

bool any_flag(const pbfx_ctx* c, int which) {
    for (const auto& h : c->sc)
        if (which ? h.params.vort_on : h.params.xsph_on) return true;
    return false;
}

uint32_t unit_width(const pbfx_ctx* c) {
    const uint32_t uw = c->T.unit_width;
    return uw < c->uw_max ? uw : c->uw_max;
}

pbfx_status do_step(pbfx_ctx* c, float* x, float* v, uint32_t n, const uint32_t* in_off, int iters,
                    bool default_sphere, hipStream_t s, float4* qtape = nullptr) {
    if (n > c->capacity) return PBFX_CAPACITY;
    if (iters < 1 || (n > 0 && (!x || !v))) return PBFX_INVALID_ARG;
    for (const auto& h : c->sc)
        if ((uint32_t)make_phys(h.params).stride > c->list_stride) return PBFX_INVALID_ARG;
    pbfx_status st = sync_table(c, default_sphere, s);
    if (st != PBFX_OK) return st;
    Bufs b = c->b;
    b.in_off = in_off;
    hipError_t e = launch_step(b, c->last_table[0], c->T, x, v, n, iters, any_flag(c, 0), any_flag(c, 1), unit_width(c),
                               s, c->timer, qtape);
    if (e != hipSuccess) return hip_fail(c, e);
    c->steps++;
    return PBFX_OK;
}

pbfx_status read_counters(pbfx_ctx* c, hipStream_t s) {
    hipError_t e = hipMemcpyAsync(c->h_pinned, c->b.ctr, kCtrWords * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->h_pinned + kCtrWords, c->b.offset + c->b.CT, 2 * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(c, e);
    // fold the statistics copies into h[0..CTR_NUM)
    uint32_t* h = c->h_pinned;
    for (int k = 0; k < CTR_NUM; ++k) {
        if (k == CTR_UNITS || k == CTR_CUNITS || k == CTR_BIGCELLS || k == CTR_HUGECELLS || k == CTR_OVF ||
            k == CTR_HEAVY || k == CTR_NSHEAVY)
            continue;
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

pbfx_status step_status(pbfx_ctx* c) {
    if (c->report_info) {
        if (c->h_pinned[CTR_TRUNC]) return PBFX_NEIGHBOR_TRUNCATED;
        if (c->h_pinned[CTR_SWEEPCAP]) return PBFX_NOT_CONVERGED;
    }
    return PBFX_OK;
}

pbfx_status sync_step_result(pbfx_ctx* c, uint32_t* num_inout, hipStream_t s) {
    pbfx_status st = read_counters(c, s);
    if (st != PBFX_OK) return st;
    *num_inout = c->h_pinned[kCtrWords];
    return step_status(c);
}

pbfx_status set_boundary(pbfx_ctx* c, SceneHost& h, const float* inner4, uint32_t n_inner, const float* outer4,
                         uint32_t n_outer, const float* pts3, const float* nrm3, uint32_t n_planes) {
    if ((n_inner && !inner4) || (n_outer && !outer4) || (n_planes && (!pts3 || !nrm3))) return PBFX_INVALID_ARG;
    (void)hipDeviceSynchronize();
    free_boundary(h);
    hipError_t e = hipSuccess;
    if (n_inner) {
        e = dalloc(&h.d_inner, n_inner);
        if (e == hipSuccess) e = hipMemcpy(h.d_inner, inner4, n_inner * sizeof(float4), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && n_outer) {
        e = dalloc(&h.d_outer, n_outer);
        if (e == hipSuccess) e = hipMemcpy(h.d_outer, outer4, n_outer * sizeof(float4), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && n_planes) {
        e = dalloc(&h.d_pts, 3 * (size_t)n_planes);
        if (e == hipSuccess) e = dalloc(&h.d_nrm, 3 * (size_t)n_planes);
        if (e == hipSuccess) e = hipMemcpy(h.d_pts, pts3, 3 * n_planes * sizeof(float), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(h.d_nrm, nrm3, 3 * n_planes * sizeof(float), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) return hip_fail(c, e);
    h.n_inner = n_inner;
    h.n_outer = n_outer;
    h.n_planes = n_planes;
    return PBFX_OK;
}

// Allocates every buffer for the scenes in c->sc (grids set) and the scene table.
pbfx_status create_common(pbfx_ctx* c) {
    (void)hipGetDevice(&c->device);
    const uint32_t n = (uint32_t)c->sc.size();
    uint32_t cap = 0, csum = 0, uwm = kMaxW;
    std::vector<uint32_t> kbase(n + 1);
    for (uint32_t i = 0; i < n; ++i) {
        SceneHost& h = c->sc[i];
        h.G.base = csum;
        kbase[i] = csum;
        if ((uint64_t)csum + h.G.C > 0x7FFFFFFFull) return PBFX_GRID_TOO_LARGE;
        csum += h.G.C;
        if ((uint64_t)cap + h.capacity > 0x7FFFFFFFull) return PBFX_CAPACITY;
        cap += h.capacity;
        if (h.G.gx < uwm) uwm = h.G.gx;
    }
    kbase[n] = csum;
    c->capacity = cap;
    c->csum = csum;
    c->uw_max = uwm;
    // key space padded so that scan tiles of T cells cover it exactly
    uint32_t CT = csum;
    if (CT <= kScanTile) {
        uint32_t p2 = 1;
        while (p2 < CT) p2 <<= 1;
        CT = p2;
    } else {
        CT = (CT + kScanTile - 1) / kScanTile * kScanTile;
    }
    Bufs& b = c->b;
    b.cap = cap > 0 ? cap : 1;
    b.CT = CT;
    b.nscenes = n;
    const uint32_t T = CT < kScanTile ? CT : kScanTile;
    b.lgT = ilog2(T);
    b.nT = CT >> b.lgT;
    const size_t bc = b.cap;
    const size_t C = CT;
    hipError_t e = hipSuccess;
#define PBFX_ALLOC(ptr, cnt)                           \
    if (e == hipSuccess) e = dalloc(&(ptr), (cnt));
    PBFX_ALLOC(b.key, bc);
    PBFX_ALLOC(b.rank, bc);
    PBFX_ALLOC(b.tmp, bc);
    PBFX_ALLOC(b.big, big_list_len(b.cap));
    PBFX_ALLOC(b.bslot, bc);
    PBFX_ALLOC(b.aux, bc);
    PBFX_ALLOC(b.bseg, big_list_len(b.cap));
    PBFX_ALLOC(b.sidx, bc);
    PBFX_ALLOC(b.isid, bc);
    PBFX_ALLOC(b.sid, bc);
    PBFX_ALLOC(b.Pf, bc);
    PBFX_ALLOC(b.Q, bc);
    PBFX_ALLOC(b.X, 3 * bc);
    PBFX_ALLOC(b.R, 2 * bc);
    PBFX_ALLOC(b.OM4, bc);
    PBFX_ALLOC(b.ncnt, bc);
    PBFX_ALLOC(b.count, C + 1);
    PBFX_ALLOC(b.offset, C + 2);
    PBFX_ALLOC(b.tsum, (size_t)b.nT + 1);
    PBFX_ALLOC(b.tprefix, (size_t)b.nT + 1);
    const size_t items = bc + bc / kChunk + 64 * (size_t)n;  // a unit per own chunk of kChunk particles
    PBFX_ALLOC(b.units, items);
    PBFX_ALLOC(b.ckey, bc);
    PBFX_ALLOC(b.crank, bc);
    PBFX_ALLOC(b.cperm, bc);
    PBFX_ALLOC(b.ccount, C + 1);
    PBFX_ALLOC(b.coffset, C + 2);
    PBFX_ALLOC(b.ctsum, (size_t)b.nT + 1);
    PBFX_ALLOC(b.ctprefix, (size_t)b.nT + 1);
    PBFX_ALLOC(b.cunits, items);
    PBFX_ALLOC(b.flist, (size_t)kFinalK * bc);
    PBFX_ALLOC(b.fcnt, bc);
    b.ovf_max = b.cap / 64u + 1u;  // every block of 64 positions can take a slot
    PBFX_ALLOC(b.ovf, b.ovf_max);
    PBFX_ALLOC(b.oslot, (size_t)b.cap / 64u + 1u);
    PBFX_ALLOC(b.part, (size_t)b.ovf_max * 27u * 64u);
    PBFX_ALLOC(b.tflag, (size_t)b.nT + 1);
    PBFX_ALLOC(b.ctflag, (size_t)b.nT + 1);
    PBFX_ALLOC(b.ctr, kCtrWords);
    PBFX_ALLOC(c->d_scenes, n);
    PBFX_ALLOC(c->d_kbase, n + 1);
    PBFX_ALLOC(c->d_in_off, n + 1);
    PBFX_ALLOC(c->d_scene_off, n + 1);
    PBFX_ALLOC(c->d_default_sphere, 1);
#undef PBFX_ALLOC
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_pinned, (kCtrWords + 2) * sizeof(uint32_t), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_scenes, n * sizeof(Scene), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_in_off, (n + 1) * sizeof(uint32_t), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_scene_off, (n + 1) * sizeof(uint32_t), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMemset(b.count, 0, (C + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b.offset, 0, (C + 2) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b.tsum, 0, ((size_t)b.nT + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b.ccount, 0, (C + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b.coffset, 0, (C + 2) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b.ctsum, 0, ((size_t)b.nT + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b.tflag, 0, ((size_t)b.nT + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b.ctflag, 0, ((size_t)b.nT + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b.ctr, 0, kCtrWords * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b.sid, 0, bc * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemcpy(c->d_kbase, kbase.data(), (n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const float4 sph = make_float4(3.f, 3.f, 3.f, 3.f);
        e = hipMemcpy(c->d_default_sphere, &sph, sizeof(sph), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, c->device);
    if (e != hipSuccess) return hip_fail(c, e);
    b.scenes = c->d_scenes;
    b.kbase = c->d_kbase;
    b.in_off = c->d_in_off;
    uint32_t stride = 0;
    for (const auto& h : c->sc) {
        const uint32_t st = (uint32_t)make_phys(h.params).stride;
        stride = st > stride ? st : stride;
    }
    pbfx_status st = alloc_list(c, stride);
    if (st != PBFX_OK) return st;
    c->T.tile_cap = 2048;
    c->T.unit_width = 16;
    set_grid_units(c);
    (void)hipDeviceSynchronize();
    return PBFX_OK;
}

pbfx_status make_grid(const pbfx_params& p, const float end[3], Grid& G) {
    uint32_t grid[3];
    // pbf_type.cpp:172-174: cell width h, grid from calcGridSize(2h, domain_end)
    pbfx_status st = pbfx_calc_grid_size(p.smoothing_length * 2.0, end, grid);
    if (st != PBFX_OK) return st;
    G.gx = grid[0];
    G.gy = grid[1];
    G.gz = grid[2];
    G.lgx = ilog2(G.gx);
    G.lgy = ilog2(G.gy);
    G.C = G.gx * G.gy * G.gz;
    G.cw = p.smoothing_length;
    G.base = 0;
    return PBFX_OK;
}

pbfx_status set_scene_params(pbfx_ctx* c, SceneHost& h, const pbfx_params* p) {
    pbfx_status st = check_params(p);
    if (st != PBFX_OK) return st;
    if (p->smoothing_length != h.params.smoothing_length) {
        Grid G{};
        st = make_grid(*p, h.domain_end, G);
        if (st != PBFX_OK) return st;
        if (G.gx != h.G.gx || G.gy != h.G.gy || G.gz != h.G.gz) return PBFX_INVALID_ARG;
        h.G.cw = p->smoothing_length;
    }
    st = alloc_list(c, (uint32_t)make_phys(*p).stride);
    if (st != PBFX_OK) return st;
    h.params = *p;
    set_grid_units(c);
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
    SceneHost h;
    h.params = *params;
    h.capacity = capacity;
    std::memcpy(h.domain_end, domain_end, sizeof(h.domain_end));
    st = make_grid(*params, domain_end, h.G);
    if (st != PBFX_OK) return st;
    pbfx_ctx* c = new (std::nothrow) pbfx_ctx();
    if (!c) return PBFX_HIP_ERROR;
    c->sc.push_back(h);
    st = create_common(c);
    if (st != PBFX_OK) {
        free_bufs(c);
        delete c;
        return st;
    }
    *out = c;
    return PBFX_OK;
}

pbfx_status pbfx_batch_create(const pbfx_scene_desc* scenes, uint32_t n_scenes, pbfx_ctx** out) {
    if (!out || !scenes || n_scenes == 0) return PBFX_INVALID_ARG;
    *out = nullptr;
    for (uint32_t i = 0; i < n_scenes; ++i) {
        pbfx_status st = check_params(&scenes[i].params);
        if (st != PBFX_OK) return st;
        if (scenes[i].params.max_pair != scenes[0].params.max_pair) return PBFX_INVALID_ARG;
    }
    pbfx_ctx* c = new (std::nothrow) pbfx_ctx();
    if (!c) return PBFX_HIP_ERROR;
    c->batch = true;
    pbfx_status st = PBFX_OK;
    for (uint32_t i = 0; i < n_scenes && st == PBFX_OK; ++i) {
        SceneHost h;
        h.params = scenes[i].params;
        h.capacity = scenes[i].capacity;
        std::memcpy(h.domain_end, scenes[i].domain_end, sizeof(h.domain_end));
        std::memcpy(h.lo, scenes[i].area_lo, sizeof(h.lo));
        std::memcpy(h.hi, scenes[i].area_hi, sizeof(h.hi));
        st = make_grid(h.params, h.domain_end, h.G);
        c->sc.push_back(h);
    }
    if (st == PBFX_OK) st = create_common(c);
    for (uint32_t i = 0; i < n_scenes && st == PBFX_OK; ++i)
        st = set_boundary(c, c->sc[i], scenes[i].inner_spheres4, scenes[i].n_inner, scenes[i].outer_spheres4,
                          scenes[i].n_outer, scenes[i].plane_points3, scenes[i].plane_normals3, scenes[i].n_planes);
    if (st != PBFX_OK) {
        free_bufs(c);
        delete c;
        return st;
    }
    *out = c;
    return PBFX_OK;
}

void pbfx_destroy(pbfx_ctx* c) {
    if (!c) return;
    DeviceGuard g(c->device);
    (void)hipDeviceSynchronize();
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    free_bufs(c);
    delete c;
}

pbfx_status pbfx_set_params(pbfx_ctx* c, const pbfx_params* p) {
    if (!c || c->batch) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    return set_scene_params(c, c->sc[0], p);
}

pbfx_status pbfx_batch_set_params(pbfx_ctx* c, uint32_t scene, const pbfx_params* p) {
    if (!c || scene >= c->sc.size() || !p) return PBFX_INVALID_ARG;
    if (p->max_pair != c->sc[0].params.max_pair) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    return set_scene_params(c, c->sc[scene], p);
}

pbfx_status pbfx_set_boundary(pbfx_ctx* c, const float* inner4, uint32_t n_inner, const float* outer4,
                              uint32_t n_outer, const float* pts3, const float* nrm3, uint32_t n_planes) {
    if (!c || c->batch) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    return set_boundary(c, c->sc[0], inner4, n_inner, outer4, n_outer, pts3, nrm3, n_planes);
}

// The plain step on a single-scene context; lo/hi become the scene's area.
static pbfx_status plain_step(pbfx_ctx* c, float* x, float* v, uint32_t num, const float lo[3], const float hi[3],
                              int iters, bool default_sphere, hipStream_t s, float4* qtape = nullptr) {
    if (c->batch || !lo || !hi) return PBFX_INVALID_ARG;
    for (int a = 0; a < 3; ++a) {
        c->sc[0].lo[a] = lo[a];
        c->sc[0].hi[a] = hi[a];
    }
    return do_step(c, x, v, num, c->d_in_off, iters, default_sphere, s, qtape);
}

pbfx_status pbfx_step_async(pbfx_ctx* c, float* x, float* v, uint32_t num, const float lo[3], const float hi[3],
                            int iters, void* stream) {
    if (!c) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    return plain_step(c, x, v, num, lo, hi, iters, false, (hipStream_t)stream);
}

pbfx_status pbfx_step(pbfx_ctx* c, float* x, float* v, uint32_t* num_inout, const float lo[3], const float hi[3],
                      int iters, void* stream) {
    if (!c || !num_inout) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    pbfx_status st = plain_step(c, x, v, *num_inout, lo, hi, iters, false, (hipStream_t)stream);
    if (st != PBFX_OK) return st;
    return sync_step_result(c, num_inout, (hipStream_t)stream);
}

pbfx_status pbfx_step_no_boundary(pbfx_ctx* c, float* x, float* v, uint32_t* num_inout, const float lo[3],
                                  const float hi[3], int iters, void* stream) {
    if (!c || !num_inout) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    pbfx_status st = plain_step(c, x, v, *num_inout, lo, hi, iters, true, (hipStream_t)stream);
    if (st != PBFX_OK) return st;
    return sync_step_result(c, num_inout, (hipStream_t)stream);
}

pbfx_status pbfx_batch_step_async(pbfx_ctx* c, float* x, float* v, const uint32_t* in_offsets, uint32_t total,
                                  int iters, void* stream) {
    if (!c || !c->batch || !in_offsets) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    return do_step(c, x, v, total, in_offsets, iters, false, (hipStream_t)stream);
}

pbfx_status pbfx_batch_offsets(pbfx_ctx* c, uint32_t* out, void* stream) {
    if (!c || !out) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    hipError_t e = launch_scene_offsets(c->b, out, (hipStream_t)stream);
    return e == hipSuccess ? PBFX_OK : hip_fail(c, e);
}

pbfx_status pbfx_batch_step(pbfx_ctx* c, float* x, float* v, uint32_t* counts, int iters, void* stream) {
    if (!c || !c->batch || !counts) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    const hipStream_t s = (hipStream_t)stream;
    const uint32_t n = (uint32_t)c->sc.size();
    uint64_t tot = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (counts[i] > c->sc[i].capacity) return PBFX_CAPACITY;
        c->h_in_off[i] = (uint32_t)tot;
        tot += counts[i];
    }
    c->h_in_off[n] = (uint32_t)tot;
    // the copy must complete before h_in_off is written again
    hipError_t e = hipMemcpyAsync(c->d_in_off, c->h_in_off, (n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(c, e);
    pbfx_status st = do_step(c, x, v, (uint32_t)tot, c->d_in_off, iters, false, s);
    if (st != PBFX_OK) {
        (void)hipStreamSynchronize(s);
        return st;
    }
    e = launch_scene_offsets(c->b, c->d_scene_off, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->h_scene_off, c->d_scene_off, (n + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return hip_fail(c, e);
    st = read_counters(c, s);  // synchronises the stream
    if (st != PBFX_OK) return st;
    for (uint32_t i = 0; i < n; ++i) counts[i] = c->h_scene_off[i + 1] - c->h_scene_off[i];
    return step_status(c);
}

pbfx_status pbfx_batch_graph_capture(pbfx_ctx* c, float* x, float* v, uint32_t* offsets, int iters, uint32_t steps,
                                     void* stream) {
    if (!c || !c->batch || !x || !v || !offsets || !stream || iters < 1 || steps < 1) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    const hipStream_t s = (hipStream_t)stream;
    for (const auto& h : c->sc)
        if ((uint32_t)make_phys(h.params).stride > c->list_stride) return PBFX_INVALID_ARG;
    // the scene table goes to the device before the recording (a copy and a
    // wait cannot be recorded)
    pbfx_status st = sync_table(c, false, s);
    if (st != PBFX_OK) return st;
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(c, e);
    if (c->graph) {
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
    Bufs b = c->b;
    b.in_off = offsets;
    hipGraph_t graph = nullptr;
    e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(c, e);
    hipError_t le = hipSuccess;
    for (uint32_t k = 0; k < steps && le == hipSuccess; ++k) {
        // grids sized for the capacity; the kernels read the live input count
        // from offsets[n_scenes]
        le = launch_step(b, c->last_table[0], c->T, x, v, c->capacity, iters, any_flag(c, 0), any_flag(c, 1),
                         unit_width(c), s, nullptr);
        // the output offsets become the next step's input offsets (every
        // reader of the input offsets has run by now)
        if (le == hipSuccess) le = launch_scene_offsets(b, offsets, s);
    }
    e = hipStreamEndCapture(s, &graph);
    if (le != hipSuccess || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return hip_fail(c, le != hipSuccess ? le : e);
    }
    e = hipGraphInstantiate(&c->graph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        c->graph = nullptr;
        return hip_fail(c, e);
    }
    c->graph_steps = steps;
    return PBFX_OK;
}

pbfx_status pbfx_batch_graph_launch(pbfx_ctx* c, uint32_t launches, void* stream) {
    if (!c || !c->batch || !c->graph) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    for (uint32_t k = 0; k < launches; ++k) {
        hipError_t e = hipGraphLaunch(c->graph, (hipStream_t)stream);
        if (e != hipSuccess) return hip_fail(c, e);
        c->steps += c->graph_steps;
    }
    return PBFX_OK;
}

pbfx_status pbfx_add_particles(pbfx_ctx* c, float* x, float* v, uint32_t* num_inout, const float* add_x,
                               const float* add_v, uint32_t add_n, void* stream) {
    if (!c || c->batch || !x || !v || !num_inout || (add_n && (!add_x || !add_v))) return PBFX_INVALID_ARG;
    if ((uint64_t)*num_inout + add_n > c->capacity) return PBFX_CAPACITY;
    if (add_n == 0) return PBFX_OK;
    DeviceGuard g(c->device);
    const hipStream_t s = (hipStream_t)stream;
    const size_t off = (size_t)*num_inout * 3, bytes = (size_t)add_n * 3 * sizeof(float);
    hipError_t e = hipMemcpyAsync(x + off, add_x, bytes, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(v + off, add_v, bytes, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return hip_fail(c, e);
    *num_inout += add_n;
    return PBFX_OK;
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
    out->tile_overflow = 0;
    out->max_tile = h[CTR_MAXTILE];
    out->steps = c->steps;
    out->final_units = h[CTR_CUNITS];
    out->list_overflow = h[CTR_LISTOVF];
    out->scan_tiles = h[CTR_SCANTILES];
    out->heavy_builds = h[CTR_HEAVY];
    out->heavy_search = h[CTR_NSHEAVY];
    return PBFX_OK;
}

pbfx_status pbfx_ns_view_get(pbfx_ctx* c, pbfx_ns_view* out) {
    if (!c || !out) return PBFX_INVALID_ARG;
    const Grid& G = c->sc[0].G;  // a batch: the first scene; offsets cover every scene's keys
    out->grid[0] = G.gx;
    out->grid[1] = G.gy;
    out->grid[2] = G.gz;
    out->num_cells = c->batch ? c->b.CT : G.C;
    out->cell_width = G.cw;
    out->max_pair = (uint32_t)c->sc[0].params.max_pair;
    out->sorted_index = c->b.sidx;
    out->cell_offset = c->b.offset;
    out->hash = c->b.key;
    return PBFX_OK;
}

pbfx_status pbfx_ns_neighbors(pbfx_ctx* c, uint32_t* out, uint32_t* counts, void* stream) {
    if (!c || !out || !counts) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    hipError_t e = launch_neighbors_view(c->b, out, counts, (uint32_t)c->sc[0].params.max_pair,
                                         (hipStream_t)stream);
    return e == hipSuccess ? PBFX_OK : hip_fail(c, e);
}

pbfx_status pbfx_ns_cell_ranges(pbfx_ctx* c, uint32_t* start, uint32_t* end, void* stream) {
    if (!c || !start || !end) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    hipError_t e = launch_cell_ranges(c->b, c->batch ? c->b.CT : c->sc[0].G.C, start, end, (hipStream_t)stream);
    return e == hipSuccess ? PBFX_OK : hip_fail(c, e);
}

pbfx_status pbfx_scaling_factor(pbfx_ctx* c, float* out, void* stream) {
    if (!c || !out) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    hipError_t e = launch_copy_lambda(c->b, out, (hipStream_t)stream);
    return e == hipSuccess ? PBFX_OK : hip_fail(c, e);
}

pbfx_status pbfx_density(pbfx_ctx* c, const float* x, float* rho, void* stream) {
    if (!c || !x || !rho) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    if (c->last_table.empty()) return PBFX_INVALID_ARG;  // no step yet
    hipError_t e = launch_density(c->b, c->last_table[0], x, rho, (hipStream_t)stream);
    return e == hipSuccess ? PBFX_OK : hip_fail(c, e);
}

pbfx_status pbfx_restore_order(pbfx_ctx* c, const float* x_sorted, const float* v_sorted, float* x_out,
                               float* v_out, void* stream) {
    if (!c || !x_sorted || !v_sorted || !x_out || !v_out) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    hipError_t e = launch_restore(c->b, x_sorted, v_sorted, x_out, v_out, (hipStream_t)stream);
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

pbfx_status pbfx_autotune(pbfx_ctx* c, const float* x, const float* v, uint32_t n, const uint32_t* in_offsets,
                          const float lo[3], const float hi[3], int iters, float* xs, float* vs, uint32_t* tile_out,
                          uint32_t* uw_out, void* stream) {
    if (!c || !x || !v || !xs || !vs || (c->batch && !in_offsets) || (!c->batch && (!lo || !hi)))
        return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    const hipStream_t s = (hipStream_t)stream;
    if (!c->batch)
        for (int a = 0; a < 3; ++a) {
            c->sc[0].lo[a] = lo[a];
            c->sc[0].hi[a] = hi[a];
        }
    Timer* keep = c->timer;
    c->timer = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    const Tuning saved = c->T;
    uint32_t best_tile = saved.tile_cap, best_uw = saved.unit_width;
    float best = 3.4e38f;
    const uint32_t tiles[] = {1024, 2048, 3072, 4096};
    const uint32_t widths[] = {4, 8, 16, 32};
    constexpr int kNT = 4, kNW = 4;
    float tbest[kNT][kNW];
    for (auto& r : tbest)
        for (float& t : r) t = 3.4e38f;
    pbfx_status st = PBFX_OK;
    // two sweeps over the candidates, best time of each kept: clock ramps and
    // the order of the candidates then bias no choice
    for (int sweep = 0; sweep < 2; ++sweep) {
        for (int ti = 0; ti < kNT; ++ti)
            for (int wi = 0; wi < kNW; ++wi) {
                const uint32_t tc = tiles[ti], uw = widths[wi];
                if (e != hipSuccess || st != PBFX_OK) break;
                if (uw > c->uw_max && uw != widths[0]) continue;
                c->T.tile_cap = tc;
                c->T.unit_width = uw;
                set_grid_units(c);
                float& t = tbest[ti][wi];
                for (int rep = 0; rep < 2 && e == hipSuccess && st == PBFX_OK; ++rep) {
                    e = hipMemcpyAsync(xs, x, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToDevice, s);
                    if (e == hipSuccess) e = hipMemcpyAsync(vs, v, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToDevice, s);
                    if (e == hipSuccess) e = hipEventRecord(e0, s);
                    if (e == hipSuccess)
                        st = do_step(c, xs, vs, n, c->batch ? in_offsets : c->d_in_off, iters, false, s);
                    if (e == hipSuccess && st == PBFX_OK) e = hipEventRecord(e1, s);
                    if (e == hipSuccess && st == PBFX_OK) e = hipEventSynchronize(e1);
                    float ms = 0.f;
                    if (e == hipSuccess && st == PBFX_OK) e = hipEventElapsedTime(&ms, e0, e1);
                    if (ms < t) t = ms;
                }
            }
    }
    for (int ti = 0; ti < kNT; ++ti)
        for (int wi = 0; wi < kNW; ++wi)
            if (tbest[ti][wi] < best) {
                best = tbest[ti][wi];
                best_tile = tiles[ti];
                best_uw = widths[wi];
            }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    c->timer = keep;
    c->T = saved;
    c->T.tile_cap = best_tile;
    c->T.unit_width = best_uw;
    set_grid_units(c);
    if (e != hipSuccess) return hip_fail(c, e);
    if (st != PBFX_OK) return st;
    if (tile_out) *tile_out = best_tile;
    if (uw_out) *uw_out = best_uw;
    return PBFX_OK;
}

void pbfx_set_report_info(pbfx_ctx* c, int enable) {
    if (c) c->report_info = enable;
}

pbfx_status pbfx_get_grid(pbfx_ctx* c, uint32_t grid[3]) {
    if (!c || !grid) return PBFX_INVALID_ARG;
    grid[0] = c->sc[0].G.gx;
    grid[1] = c->sc[0].G.gy;
    grid[2] = c->sc[0].G.gz;
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

// ---------------------------------------------------------- differentiable form

static uint32_t pad_tiles(uint32_t n, uint32_t& lgT, uint32_t& nT) {
    uint32_t N = n > 0 ? n : 1;
    if (N <= kScanTile) {
        uint32_t p2 = 1;
        while (p2 < N) p2 <<= 1;
        N = p2;
    } else {
        N = (N + kScanTile - 1) / kScanTile * kScanTile;
    }
    const uint32_t T = N < kScanTile ? N : kScanTile;
    lgT = ilog2(T);
    nT = N >> lgT;
    return N;
}

pbfx_status pbfx_adj_enable(pbfx_ctx* c, uint32_t max_steps, uint32_t every) {
    if (!c || c->batch || max_steps == 0 || every == 0) return PBFX_INVALID_ARG;
    const Grid& G = c->sc[0].G;
    if (G.gx < 4 || G.gy < 4 || G.gz < 4) return PBFX_NOT_SUPPORTED;
    if (c->b.cap > kAdjMaxBlocks * 256u) return PBFX_CAPACITY;
    DeviceGuard g(c->device);
    (void)hipDeviceSynchronize();
    free_adj(c);
    AdjState* A = new (std::nothrow) AdjState();
    if (!A) return PBFX_HIP_ERROR;
    c->adj = A;
    A->max_steps = max_steps;
    A->every = every < max_steps ? every : max_steps;
    A->stride = c->list_stride;
    const size_t cap = c->b.cap, CT = c->b.CT, nT = c->b.nT;
    const uint32_t nck = (max_steps + A->every - 1) / A->every;
    hipError_t e = hipSuccess;
#define PBFX_ALLOC(ptr, cnt)                           \
    if (e == hipSuccess) e = dalloc(&(ptr), (cnt));
    A->ck_x.assign(nck, nullptr);
    A->ck_v.assign(nck, nullptr);
    A->ck_n.assign(nck, 0u);
    for (uint32_t k = 0; k < nck; ++k) {
        PBFX_ALLOC(A->ck_x[k], 3 * cap);
        PBFX_ALLOC(A->ck_v[k], 3 * cap);
    }
    PBFX_ALLOC(A->id[0], cap);
    PBFX_ALLOC(A->id[1], cap);
    PBFX_ALLOC(A->xw, 3 * cap);
    PBFX_ALLOC(A->vw, 3 * cap);
    A->tape.assign(A->every, StepTape{});
    for (auto& t : A->tape) {
        PBFX_ALLOC(t.sidx, cap);
        PBFX_ALLOC(t.X, 3 * cap);
        PBFX_ALLOC(t.nlist, cap * c->list_stride);
        PBFX_ALLOC(t.ncnt, cap);
        PBFX_ALLOC(t.offset, CT + 2);
        PBFX_ALLOC(t.PI, cap);
        PBFX_ALLOC(t.R, 2 * cap);
        PBFX_ALLOC(t.OM4, cap);
    }
    AdjBufs& a = A->a;
    a.cap = (uint32_t)cap;
    PBFX_ALLOC(a.ax, 3 * cap);
    PBFX_ALLOC(a.av, 3 * cap);
    PBFX_ALLOC(a.axin, 3 * cap);
    PBFX_ALLOC(a.avin, 3 * cap);
    PBFX_ALLOC(a.aP, cap);
    PBFX_ALLOC(a.aQ, cap);
    PBFX_ALLOC(a.aV, cap);
    PBFX_ALLOC(a.aVc, cap);
    PBFX_ALLOC(a.aOm, cap);
    PBFX_ALLOC(a.etab, cap);
    PBFX_ALLOC(a.abar, cap);
    PBFX_ALLOC(a.coef, cap);
    PBFX_ALLOC(a.g2, cap);
    PBFX_ALLOC(a.aX, cap);
    PBFX_ALLOC(a.xnan, cap);
    PBFX_ALLOC(a.ckey, cap);
    PBFX_ALLOC(a.crank, cap);
    PBFX_ALLOC(a.ctmp, cap);
    PBFX_ALLOC(a.big, big_list_len(cap));
    PBFX_ALLOC(a.nbig, 2);
    PBFX_ALLOC(a.cperm, cap);
    PBFX_ALLOC(a.ccount, CT + 1);
    PBFX_ALLOC(a.coffset, CT + 2);
    PBFX_ALLOC(a.ctsum, nT + 1);
    PBFX_ALLOC(a.ctflag, nT + 1);
    PBFX_ALLOC(a.ctprefix, nT + 1);
    PBFX_ALLOC(a.gP, cap);
    PBFX_ALLOC(a.gV, cap);
    PBFX_ALLOC(a.gD, cap);
    PBFX_ALLOC(a.hflag, cap);
    PBFX_ALLOC(a.hlist, 4 * cap);
    PBFX_ALLOC(a.nh, 4);
    PBFX_ALLOC(a.last, cap);
    a.RT = pad_tiles((uint32_t)cap, a.lgRT, a.nRT);
    PBFX_ALLOC(a.rcount, (size_t)a.RT + 1);
    PBFX_ALLOC(a.roff, (size_t)a.RT + 2);
    PBFX_ALLOC(a.rtsum, (size_t)a.nRT + 1);
    PBFX_ALLOC(a.rtflag, (size_t)a.nRT + 1);
    PBFX_ALLOC(a.rtprefix, (size_t)a.nRT + 1);
    PBFX_ALLOC(a.rlist, cap * c->list_stride);
    PBFX_ALLOC(a.pc, cap);
    PBFX_ALLOC(a.pe, cap);
    PBFX_ALLOC(a.psum, 2);
    PBFX_ALLOC(a.amax, 1);
#undef PBFX_ALLOC
    if (e == hipSuccess) e = hipMemset(a.ccount, 0, (CT + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(a.ctflag, 0, (nT + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(a.rcount, 0, ((size_t)a.RT + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(a.rtflag, 0, ((size_t)a.nRT + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipHostMalloc((void**)&A->h_amax, sizeof(*A->h_amax), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&A->h_psum, 2 * sizeof(double), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&A->h_nh, 4 * sizeof(uint32_t), hipHostMallocDefault);
    if (e != hipSuccess) {
        free_adj(c);
        return hip_fail(c, e);
    }
    return PBFX_OK;
}

// Tapes hold (x*, s) of `iters` Jacobi iterations.
static pbfx_status adj_tape_iters(pbfx_ctx* c, uint32_t iters) {
    AdjState* A = c->adj;
    if (A->tape_iters >= iters) return PBFX_OK;
    free_tape_q(A);
    for (auto& t : A->tape) {
        hipError_t e = dalloc(&t.Q, (size_t)iters * c->b.cap);
        if (e != hipSuccess) {
            free_tape_q(A);
            return hip_fail(c, e);
        }
    }
    A->tape_iters = iters;
    return PBFX_OK;
}

pbfx_status pbfx_adj_forward(pbfx_ctx* c, float* x, float* v, uint32_t* num, const float lo[3], const float hi[3],
                             int iters, uint32_t steps, void* stream) {
    if (!c || !c->adj || c->batch || !x || !v || !num || !lo || !hi || iters < 1) return PBFX_INVALID_ARG;
    AdjState* A = c->adj;
    if (steps == 0 || steps > A->max_steps || *num > c->capacity) return PBFX_INVALID_ARG;
    if (c->list_stride != A->stride) return PBFX_INVALID_ARG;  // max_pair changed since pbfx_adj_enable
    DeviceGuard g(c->device);
    const hipStream_t s = (hipStream_t)stream;
    pbfx_status st = adj_tape_iters(c, (uint32_t)iters);
    if (st != PBFX_OK) return st;
    A->iters = (uint32_t)iters;
    A->steps = 0;
    for (int k = 0; k < 3; ++k) {
        A->lo[k] = lo[k];
        A->hi[k] = hi[k];
    }
    uint32_t n = *num;
    A->n0 = n;
    A->id_cur = 0;
    hipError_t e = launch_adj_ident(nullptr, nullptr, A->id[0], n, s);
    if (e != hipSuccess) return hip_fail(c, e);
    for (uint32_t t = 0; t < steps; ++t) {
        if (t % A->every == 0) {
            const uint32_t k = t / A->every;
            e = hipMemcpyAsync(A->ck_x[k], x, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToDevice, s);
            if (e == hipSuccess)
                e = hipMemcpyAsync(A->ck_v[k], v, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToDevice, s);
            if (e != hipSuccess) return hip_fail(c, e);
            A->ck_n[k] = n;
        }
        st = plain_step(c, x, v, n, lo, hi, iters, false, s);
        if (st == PBFX_OK) st = sync_step_result(c, &n, s);
        if (st != PBFX_OK) return st;
        e = launch_adj_ident(A->id[A->id_cur], c->b.sidx, A->id[1 - A->id_cur], n, s);
        if (e != hipSuccess) return hip_fail(c, e);
        A->id_cur ^= 1;
        A->steps = t + 1;
    }
    A->n_final = n;
    *num = n;
    return PBFX_OK;
}

pbfx_status pbfx_adj_identity(pbfx_ctx* c, uint32_t* id, void* stream) {
    if (!c || !c->adj || !id || c->adj->steps == 0) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    AdjState* A = c->adj;
    hipError_t e = hipMemcpyAsync(id, A->id[A->id_cur], (size_t)A->n_final * sizeof(uint32_t),
                                  hipMemcpyDeviceToDevice, (hipStream_t)stream);
    return e == hipSuccess ? PBFX_OK : hip_fail(c, e);
}

// Runs one step from (xw, vw, *n) recording its tape.
static pbfx_status adj_record_step(pbfx_ctx* c, StepTape& t, uint32_t* n, hipStream_t s) {
    AdjState* A = c->adj;
    const Bufs& b = c->b;
    t.n_in = *n;
    pbfx_status st = plain_step(c, A->xw, A->vw, *n, A->lo, A->hi, (int)A->iters, false, s, t.Q);
    if (st == PBFX_OK) st = sync_step_result(c, n, s);
    if (st != PBFX_OK) return st;
    t.n_live = *n;
    const size_t cap = b.cap;
    struct Cp {
        void* dst;
        const void* src;
        size_t bytes;
    } cps[] = {{t.sidx, b.sidx, cap * sizeof(uint32_t)},
               {t.X, b.X, 3 * cap * sizeof(float)},
               {t.nlist, b.nlist, cap * c->list_stride * sizeof(uint32_t)},
               {t.ncnt, b.ncnt, cap * sizeof(uint32_t)},
               {t.offset, b.offset, ((size_t)b.CT + 2) * sizeof(uint32_t)},
               {t.PI, b.Pf, cap * sizeof(float4)},
               {t.R, b.R, 2 * cap * sizeof(float4)},
               {t.OM4, b.OM4, cap * sizeof(float4)}};
    for (const Cp& q : cps) {
        hipError_t e = hipMemcpyAsync(q.dst, q.src, q.bytes, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return hip_fail(c, e);
    }
    return PBFX_OK;
}

pbfx_status pbfx_adj_backward(pbfx_ctx* c, const double* gx, const double* gv, double* gx0, double* gv0,
                              double gparam[2], int32_t* exponent, void* stream) {
    if (!c || !c->adj || !gx || !gv || !gx0 || !gv0 || !gparam || !exponent) return PBFX_INVALID_ARG;
    AdjState* A = c->adj;
    if (A->steps == 0 || c->list_stride != A->stride) return PBFX_INVALID_ARG;
    DeviceGuard g(c->device);
    const hipStream_t s = (hipStream_t)stream;
    AdjBufs a = A->a;
    hipError_t e = hipMemcpyAsync(a.ax, gx, (size_t)A->n_final * 3 * sizeof(double), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(a.av, gv, (size_t)A->n_final * 3 * sizeof(double), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return hip_fail(c, e);
    double gp[2] = {0.0, 0.0};
    int32_t E = 0;
    const uint32_t nwin = (A->steps + A->every - 1) / A->every;
    for (uint32_t w = nwin; w-- > 0;) {
        const uint32_t t0 = w * A->every;
        const uint32_t t1 = (t0 + A->every < A->steps) ? t0 + A->every : A->steps;
        uint32_t n = A->ck_n[w];
        e = hipMemcpyAsync(A->xw, A->ck_x[w], (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(A->vw, A->ck_v[w], (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return hip_fail(c, e);
        for (uint32_t t = t0; t < t1; ++t) {
            pbfx_status st = adj_record_step(c, A->tape[t - t0], &n, s);
            if (st != PBFX_OK) return st;
        }
        for (uint32_t t = t1; t-- > t0;) {
            const StepTape& T = A->tape[t - t0];
            e = launch_adj_step(c->b, c->last_table[0], T, a, (int)A->iters, s, A->h_nh);
            if (e == hipSuccess) e = launch_adj_absmax(a, T.n_in, s);
            if (e == hipSuccess) e = hipMemcpyAsync(A->h_psum, a.psum, 2 * sizeof(double), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipMemcpyAsync(A->h_amax, a.amax, sizeof(*a.amax), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) return hip_fail(c, e);
            gp[0] += A->h_psum[0];
            gp[1] += A->h_psum[1];
            double m;
            std::memcpy(&m, A->h_amax, sizeof(m));
            // keep the adjoint near 1 once it grows past 2^20: scale by 2^-k
            if (std::isfinite(m) && m > 1048576.0) {
                const int k = std::ilogb(m);
                const double f = std::ldexp(1.0, -k);
                e = launch_adj_scale(a.axin, a.avin, 3 * T.n_in, f, s);
                if (e != hipSuccess) return hip_fail(c, e);
                gp[0] = std::ldexp(gp[0], -k);
                gp[1] = std::ldexp(gp[1], -k);
                E += k;
            }
            std::swap(a.ax, a.axin);
            std::swap(a.av, a.avin);
        }
    }
    e = hipMemcpyAsync(gx0, a.ax, (size_t)A->n0 * 3 * sizeof(double), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(gv0, a.av, (size_t)A->n0 * 3 * sizeof(double), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(c, e);
    gparam[0] = gp[0];
    gparam[1] = gp[1];
    *exponent = E;
    return PBFX_OK;
}

}  // extern "C"

// ------------------------------------------------------------ allocation-free step
// pbfx_step_static keeps every scratch array of a single-scene step in this
// static device array (one copy per GPU, held until the process ends).  Host
// code never uses its name as an address: each device's address comes from
// hipGetSymbolAddress, taken once per device with that device current.
static __device__ __align__(256) unsigned char scratch[PBFX_STATIC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
// overflow slots of the allocation-free step (64 positions each, their
// partials 27 KiB a slot)
constexpr uint32_t kStaticOvfSlots = 64;

struct StaticDev {
    bool ready = false;
    unsigned char* ptr = nullptr;  // device address of `scratch` on this device
    Tuning T{};
};
StaticDev g_static[kMaxDevices];
std::mutex g_static_mu;

// The scratch address and launch sizing of the current device, set up on the
// first call on it.
pbfx_status static_device(int dev, StaticDev** out, int* hip_err) {
    if (dev < 0 || dev >= kMaxDevices) return PBFX_NOT_SUPPORTED;
    std::lock_guard<std::mutex> lock(g_static_mu);
    StaticDev& d = g_static[dev];
    if (!d.ready) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        int ncu = 0;
        if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) {
            *hip_err = (int)e;
            return PBFX_HIP_ERROR;
        }
        d.ptr = static_cast<unsigned char*>(p);
        d.T.tile_cap = 2048;
        d.T.unit_width = 16;
        size_unit_grids(d.T, ncu);
        d.ready = true;
    }
    *out = &d;
    return PBFX_OK;
}

// Places every array of a single-scene step for `cap` particles and CT cell
// keys (nT scan tiles) at 256-byte-aligned offsets from `base`; returns the
// bytes used.  base == 0 only measures.
size_t carve_static(Bufs& b, uintptr_t base, uint32_t cap, uint32_t CT, uint32_t nT, uint32_t stride) {
    size_t off = 0;
    auto take = [&](auto*& ptr, size_t count) {
        using T = std::remove_reference_t<decltype(*ptr)>;
        off = (off + 255) & ~(size_t)255;
        ptr = reinterpret_cast<T*>(base + off);
        off += (count ? count : 1) * sizeof(T);
    };
    const size_t bc = cap, C = CT, items = bc + bc / kChunk + 64;
    take(b.key, bc);
    take(b.rank, bc);
    take(b.tmp, bc);
    take(b.big, big_list_len(cap));
    take(b.bslot, bc);
    take(b.aux, bc);
    take(b.bseg, big_list_len(cap));
    take(b.sidx, bc);
    take(b.Pf, bc);
    take(b.Q, bc);
    take(b.X, 3 * bc);
    take(b.R, 2 * bc);
    take(b.OM4, bc);
    take(b.nlist, (size_t)stride * bc);
    take(b.ncnt, bc);
    take(b.count, C + 1);
    take(b.offset, C + 2);
    take(b.tsum, (size_t)nT + 1);
    take(b.tflag, (size_t)nT + 1);
    take(b.tprefix, (size_t)nT + 1);
    take(b.units, items);
    take(b.ckey, bc);
    take(b.crank, bc);
    take(b.cperm, bc);
    take(b.ccount, C + 1);
    take(b.coffset, C + 2);
    take(b.ctsum, (size_t)nT + 1);
    take(b.ctflag, (size_t)nT + 1);
    take(b.ctprefix, (size_t)nT + 1);
    take(b.cunits, items);
    take(b.flist, (size_t)kFinalK * bc);
    take(b.fcnt, bc);
    b.ovf_max = cap / 64u + 1u < kStaticOvfSlots ? cap / 64u + 1u : kStaticOvfSlots;
    take(b.ovf, b.ovf_max);
    take(b.oslot, (size_t)cap / 64u + 1u);
    take(b.part, (size_t)b.ovf_max * 27u * 64u);
    take(b.ctr, kCtrWords);
    return off;
}

}  // namespace

extern "C" {

pbfx_status pbfx_step_static(const pbfx_params* params, const float domain_end[3], const float* inner_spheres4,
                             uint32_t n_inner, const float* outer_spheres4, uint32_t n_outer,
                             const float* plane_points3, const float* plane_normals3, uint32_t n_planes, float* x,
                             float* v, uint32_t num, const float lo[3], const float hi[3], int iters, void* stream) {
    pbfx_status st = check_params(params);
    if (st != PBFX_OK) return st;
    if (!domain_end || !lo || !hi || iters < 1 || (num > 0 && (!x || !v))) return PBFX_INVALID_ARG;
    if ((n_inner && !inner_spheres4) || (n_outer && !outer_spheres4) || (n_planes && (!plane_points3 || !plane_normals3)))
        return PBFX_INVALID_ARG;
    Grid G{};
    st = make_grid(*params, domain_end, G);
    if (st != PBFX_OK) return st;
    // the key space padded as a context pads it (create_common)
    uint32_t CT = G.C;
    if (CT <= kScanTile) {
        uint32_t p2 = 1;
        while (p2 < CT) p2 <<= 1;
        CT = p2;
    } else {
        CT = (CT + kScanTile - 1) / kScanTile * kScanTile;
    }
    const uint32_t lgT = ilog2(CT < kScanTile ? CT : kScanTile), nT = CT >> lgT;
    const Phys P = make_phys(*params);
    const uint32_t cap = num > 0 ? num : 1u;
    Bufs b{};
    if (carve_static(b, 0, cap, CT, nT, (uint32_t)P.stride) > sizeof(scratch)) return PBFX_CAPACITY;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return PBFX_HIP_ERROR;
    StaticDev* d = nullptr;
    int herr = 0;
    st = static_device(dev, &d, &herr);
    if (st != PBFX_OK) return st;
    carve_static(b, reinterpret_cast<uintptr_t>(d->ptr), cap, CT, nT, (uint32_t)P.stride);
    b.cap = cap;
    b.CT = CT;
    b.lgT = lgT;
    b.nT = nT;
    b.nscenes = 1;
    Scene S;
    std::memset(&S, 0, sizeof(S));
    S.G = G;
    S.P = P;
    S.B.inner = reinterpret_cast<const float4*>(inner_spheres4);
    S.B.n_inner = n_inner;
    S.B.outer = reinterpret_cast<const float4*>(outer_spheres4);
    S.B.n_outer = n_outer;
    S.B.pts = plane_points3;
    S.B.nrm = plane_normals3;
    S.B.n_planes = n_planes;
    for (int a = 0; a < 3; ++a) {
        S.lo[a] = lo[a];
        S.hi[a] = hi[a];
    }
    hipStream_t s = (hipStream_t)stream;
    // the counters and flags the step accumulates into start at zero; the
    // array keeps the last call's contents, possibly of another layout
    hipError_t e = launch_fill_u32(b.count, (size_t)CT + 1, 0u, s);
    if (e == hipSuccess) e = launch_fill_u32(b.tflag, (size_t)nT + 1, 0u, s);
    if (e == hipSuccess) e = launch_fill_u32(b.ccount, (size_t)CT + 1, 0u, s);
    if (e == hipSuccess) e = launch_fill_u32(b.ctflag, (size_t)nT + 1, 0u, s);
    const uint32_t uw = d->T.unit_width < G.gx ? d->T.unit_width : G.gx;
    if (e == hipSuccess)
        e = launch_step(b, S, d->T, x, v, num, iters, params->xsph_on, params->vort_on, uw < kMaxW ? uw : kMaxW, s,
                        nullptr);
    return e == hipSuccess ? PBFX_OK : PBFX_HIP_ERROR;
}

pbfx_status pbfx_selftest_inv_sqrt(uint32_t lo_bits, uint32_t hi_bits, unsigned long long* mismatches,
                                   uint32_t* first_bits, void* stream) {
    if (!mismatches || !first_bits || lo_bits > hi_bits) return PBFX_INVALID_ARG;
    hipError_t e = launch_selftest_inv_sqrt(lo_bits, hi_bits, mismatches, first_bits, (hipStream_t)stream);
    return e == hipSuccess ? PBFX_OK : PBFX_HIP_ERROR;
}

pbfx_status pbfx_selftest_arith(const float* a, const float* b, float* out, uint32_t n, void* stream) {
    if ((n && (!a || !b || !out))) return PBFX_INVALID_ARG;
    hipError_t e = launch_selftest(a, b, out, n, (hipStream_t)stream);
    return e == hipSuccess ? PBFX_OK : PBFX_HIP_ERROR;
}

}  // extern "C"
