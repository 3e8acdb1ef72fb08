// pbf_oracle_core.hpp -- shared internals of the CPU oracle (test
// infrastructure only): vector helpers in GLM order, the SPH kernels, the
// hashed grid, the neighbour search, pair terms, boundary responses and the
// 27-cell sum.  Included by pbf_oracle.cpp (the step) and pbf_adjoint.cpp
// (the frozen-decision map and its adjoint).  Citations are relative to
// cuda_pbf_with_ssf/src/ of yodaplusplus/ssf_pbf_cuda.
#pragma once

#include "pbf_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

namespace orc_core {

// The planted mistake in force (ORC_PLANT_*, pbf_oracle.h); ORC_PLANT_NONE
// outside the tests that show their checks catch such mistakes.
inline int& plant_code() {
    static int code = ORC_PLANT_NONE;
    return code;
}

template <class R>
struct V3 {
    R x, y, z;
};

template <class R>
inline V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R>
inline V3<R> operator-(V3<R> a, V3<R> b) { return v3<R>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R>
inline V3<R> operator+(V3<R> a, V3<R> b) { return v3<R>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class R>
inline V3<R>& operator+=(V3<R>& a, V3<R> b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
// glm: vec * scalar and scalar * vec are componentwise products
template <class R>
inline V3<R> mulvs(V3<R> a, R s) { return v3<R>(a.x * s, a.y * s, a.z * s); }
template <class R>
inline V3<R> mulsv(R s, V3<R> a) { return v3<R>(s * a.x, s * a.y, s * a.z); }
template <class R>
inline V3<R> divvs(V3<R> a, R s) { return v3<R>(a.x / s, a.y / s, a.z / s); }
template <class R>
inline R glm_dot(V3<R> a, V3<R> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class R>
inline R glm_length(V3<R> a) { return std::sqrt(glm_dot(a, a)); }
template <class R>
inline V3<R> glm_cross(V3<R> x, V3<R> y) {
    return v3<R>(x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y);
}
// fmaf(a.z,b.z, fmaf(a.y,b.y, a.x*b.x)): the hand-fused dot used throughout the solver kernels
template <class R>
inline R fdot(V3<R> a, V3<R> b) { return std::fma(a.z, b.z, std::fma(a.y, b.y, a.x * b.x)); }

template <class R>
struct KC;
template <>
struct KC<float> {
    static constexpr float poly6 = 1.56668147065f;   // 315/(64 pi), pbf_kernel.inl:40
    static constexpr float spiky = -14.3239448745f;  // -45/pi,    pbf_kernel.inl:58
    static constexpr float ten = 10.f;
    static constexpr float nudge_in = 1.0001f;
    static constexpr float nudge_out = 0.9999f;
};
template <>
struct KC<double> {
    static constexpr double poly6 = 1.56668147065;
    static constexpr double spiky = -14.3239448745;
    static constexpr double ten = 10.0;
    static constexpr double nudge_in = 1.0001;
    static constexpr double nudge_out = 0.9999;
};

// pbf/kernel/cuda/detail/pbf_kernel.inl:34-48 (Wpoly6)
template <class R>
inline R wpoly6(R r, R inv_h) {
    R q = r * inv_h;
    R alpha = KC<R>::poly6 * inv_h * inv_h * inv_h;
    if (q <= R(1)) {
        R q2 = q * q;
        R t = R(1) - q2;
        return alpha * t * t * t;
    }
    return R(0);
}

// pbf/kernel/cuda/detail/pbf_kernel.inl:50-63 (WspikyDerivative)
template <class R>
inline R wspiky_deriv(R r, R inv_h) {
    R q = r * inv_h;
    R inv_h2 = inv_h * inv_h;
    if (q <= R(1)) {
        R t = R(1) - q;
        R inv_h4 = inv_h2 * inv_h2;
        return KC<R>::spiky * inv_h4 * t * t;
    }
    return R(0);
}

// float -> int with a defined result everywhere: NaN -> 0, saturating (the
// device conversion's behaviour; a plain cast is undefined out of range).
inline int f2i(float f) {
    if (!(f == f)) return 0;
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f <= -2147483648.0f) return INT32_MIN;
    return static_cast<int>(f);
}

// pbf/interaction/cuda/pbf_grid.h:20-28: cell = (int)floor(p / cell_width), IEEE division
inline void grid_pos(float px, float py, float pz, float cw, int g[3]) {
    g[0] = f2i(std::floor(px / cw));
    g[1] = f2i(std::floor(py / cw));
    g[2] = f2i(std::floor(pz / cw));
}

// pbf/interaction/cuda/pbf_grid.h:54-61: wrap each axis (power-of-two grid), z-major key
inline uint32_t grid_hash(int gx, int gy, int gz, const uint32_t G[3]) {
    uint32_t x = static_cast<uint32_t>(gx) & (G[0] - 1u);
    uint32_t y = static_cast<uint32_t>(gy) & (G[1] - 1u);
    uint32_t z = static_cast<uint32_t>(gz) & (G[2] - 1u);
    return z * G[0] * G[1] + y * G[0] + x;
}

// Integer power by binary exponentiation, the definition shared with the
// HIP solver for the PAPER s_corr (t^n): n=4 evaluates (t*t)*(t*t).
template <class R>
inline R powi(R t, int n) {
    R result = R(1);
    R base = t;
    bool first = true;
    while (n > 0) {
        if (n & 1) {
            result = first ? base : result * base;
            first = false;
        }
        n >>= 1;
        if (n) base = base * base;
    }
    return result;
}

// Discrete outcome of the neighbour search, computed once in FP32 and shared
// by the FP32 step and its FP64 twin.
struct NSData {
    uint32_t n_in = 0, n_live = 0, C = 0;
    uint32_t G[3] = {1, 1, 1};
    float cw = 0.f;
    std::vector<uint32_t> key;       // per original particle
    std::vector<uint32_t> idx;       // slot -> original
    std::vector<uint32_t> off;       // C+2 exclusive offsets
    std::vector<uint32_t> nb;        // n_live * max_pair
    std::vector<uint32_t> nbc;       // n_live
    std::vector<float> pred;         // sorted predicted positions (float), AoS
    uint32_t truncated = 0;
    int max_pair = 30;
};

// pbf/solver/cuda/pbf_predict.cu:15-45 via pbf_solve.cpp:16-25:
//   v* = fmaf(g, dt, v);  x* = fmaf(v*, dt, x)
template <class R>
inline void predict_one(const orc_params* p, const float* x, const float* v, uint32_t i, R out[3]) {
    R dt = R(p->time_step);
    for (int c = 0; c < 3; ++c) {
        R vs = std::fma(R(p->body_force[c]), dt, R(v[3 * i + c]));
        out[c] = std::fma(vs, dt, R(x[3 * i + c]));
    }
}

// pbf/interaction/pbf_neighbor_search.cpp:92-170 (counting-sort branch) with
// the stable order of its USE_CUB/USE_THRUST branches (:112-131):
//   hash (pbf_neighbor_search_util.cu:20-37) -> deletion flag
//   (pbf/util/cuda/pbf_delete.cu:9-28) -> counting sort
//   (pbf/util/cuda/pbf_counting_sort.cu:167-178) -> deleted count
//   (pbf_delete.cu:58-77) -> cell ranges (pbf_neighbor_search_util.cu:39-81)
//   -> reorder (:143-154) -> neighbour list (:361-482).
inline void neighbor_search(const orc_params* p, const float* x, const float* v, uint32_t n,
                     const float lo[3], const float hi[3], NSData& ns) {
    ns.n_in = n;
    ns.G[0] = p->grid[0];
    ns.G[1] = p->grid[1];
    ns.G[2] = p->grid[2];
    ns.C = ns.G[0] * ns.G[1] * ns.G[2];
    ns.cw = p->smoothing_length;
    ns.max_pair = p->max_pair;
    const uint32_t C = ns.C;

    std::vector<float> xs(3 * (size_t)n);
    ns.key.assign(n, 0);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        uint32_t i = (uint32_t)ii;
        float q[3];
        predict_one<float>(p, x, v, i, q);
        xs[3 * i + 0] = q[0];
        xs[3 * i + 1] = q[1];
        xs[3 * i + 2] = q[2];
        int g[3];
        grid_pos(q[0], q[1], q[2], ns.cw, g);
        uint32_t k = grid_hash(g[0], g[1], g[2], ns.G);
        if (q[0] < lo[0] || hi[0] < q[0] || q[1] < lo[1] || hi[1] < q[1] || q[2] < lo[2] || hi[2] < q[2])
            k = C;  // sentinel: sorts after every cell
        ns.key[i] = k;
    }

    // Stable counting sort: a sequential scatter in index order is stable.
    ns.off.assign((size_t)C + 2, 0);
    for (uint32_t i = 0; i < n; ++i) ns.off[ns.key[i] + 1]++;
    for (size_t c = 1; c < ns.off.size(); ++c) ns.off[c] += ns.off[c - 1];
    // off[c] = first slot of cell c; off[C] = live count; off[C+1] = n
    {
        std::vector<uint32_t> cur(ns.off.begin(), ns.off.end() - 1);
        ns.idx.assign(n, 0);
        for (uint32_t i = 0; i < n; ++i) ns.idx[cur[ns.key[i]]++] = i;
    }
    ns.n_live = ns.off[C];
    const uint32_t nl = ns.n_live;

    ns.pred.assign(3 * (size_t)n, 0.f);
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < (int64_t)n; ++s) {
        uint32_t i = ns.idx[s];
        ns.pred[3 * s + 0] = xs[3 * i + 0];
        ns.pred[3 * s + 1] = xs[3 * i + 1];
        ns.pred[3 * s + 2] = xs[3 * i + 2];
    }

    // detectNeighborsCUDA: 27 cells z,y,x in -1..1; ascending slots; keep
    // r2 < h*h while count < max_pair-1 (pbf_neighbor_search_util.cu:386-408,428-452)
    const int M = ns.max_pair;
    const int cap = M - 1;
    const float h = p->smoothing_length;
    const float h2 = h * h;
    ns.nb.assign((size_t)nl * M, ORC_EMPTY);
    ns.nbc.assign(nl, 0);
    uint32_t truncated = 0;
#pragma omp parallel for schedule(static) reduction(+ : truncated)
    for (int64_t ss = 0; ss < (int64_t)nl; ++ss) {
        uint32_t s = (uint32_t)ss;
        V3<float> self = v3<float>(ns.pred[3 * s], ns.pred[3 * s + 1], ns.pred[3 * s + 2]);
        int g[3];
        grid_pos(self.x, self.y, self.z, ns.cw, g);
        int cnt = 0;
        bool trunc = false;
        uint32_t* out = &ns.nb[(size_t)s * M];
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    uint32_t c = grid_hash(g[0] + dx, g[1] + dy, g[2] + dz, ns.G);
                    for (uint32_t j = ns.off[c]; j < ns.off[c + 1]; ++j) {
                        V3<float> pj = v3<float>(ns.pred[3 * j], ns.pred[3 * j + 1], ns.pred[3 * j + 2]);
                        V3<float> d = self - pj;
                        float r2 = std::fma(d.z, d.z, std::fma(d.y, d.y, d.x * d.x));
                        if (r2 < h2) {
                            if (cnt < cap)
                                out[cnt++] = j;
                            else
                                trunc = true;
                        }
                    }
                }
        ns.nbc[s] = (uint32_t)cnt;
        if (trunc) truncated++;
    }
    ns.truncated = truncated;
}

template <class R>
struct Consts {
    R h, inv_h, m, inv_rho0, relax, a, ad, inv_k, inv_wdq, sk, inv_dt, dtc, xsph, vc;
    int mode, n;
};

// Host-side constants of pbf_solve.cpp:79-106 / :147-216 and
// pbf_position_update.cu:224-238 (inv_k), plus the PAPER constants.
template <class R>
Consts<R> make_consts(const orc_params* p) {
    Consts<R> k;
    const float hf = p->smoothing_length;
    const float inv_rho0f = 1.f / p->stable_density;
    const float inv_tf = 1.f / p->time_step;
    k.h = R(hf);
    k.inv_h = R(1) / R(hf);
    k.m = R(p->particle_mass);
    k.inv_rho0 = R(inv_rho0f);
    k.relax = R(p->relaxation);
    R inv_h2 = k.inv_h * k.inv_h;
    // pbf_kernel_memorization.cu:96-101 / :110-114
    k.a = KC<R>::poly6 * k.inv_h * inv_h2;
    k.ad = KC<R>::spiky * inv_h2 * inv_h2;
    float ref_inv_k, inv_wdq;
    orc_scorr_constants(p, &ref_inv_k, &inv_wdq);
    k.inv_k = R(ref_inv_k);
    k.inv_wdq = R(inv_wdq);
    k.sk = R(p->scorr_k);
    k.mode = p->scorr_mode;
    k.n = p->scorr_n;
    k.inv_dt = R(inv_tf);
    k.dtc = R(1.f / inv_tf);  // pbf_solve.cpp:138 passes 1.f / inv_time_step
    k.xsph = R(p->xsph_parameter);
    k.vc = R(p->vc_parameter);
    if (!std::is_same<R, float>::value) {
        // FP64 twin: recompute the derived constants in double
        k.inv_rho0 = R(1) / R(p->stable_density);
        k.inv_dt = R(1) / R(p->time_step);
        k.dtc = R(p->time_step);
        R w04 = wpoly6<R>(R(hf) * R(0.4f), k.inv_h);
        k.inv_k = R(1) / (w04 * w04 * w04 * w04);
        k.inv_wdq = R(1) / wpoly6<R>(R(p->scorr_dq_over_h) * R(hf), k.inv_h);
    }
    return k;
}

// pbf_kernel_memorization.cu:15-56 (memorizeKernelPair): W and grad W of one
// listed pair; rsqrtf restated as 1/sqrt.
template <class R>
inline void pair_terms(V3<R> self, V3<R> pj, const Consts<R>& k, R& W, V3<R>& kg) {
    V3<R> d = self - pj;
    R r2 = std::fma(d.z, d.z, std::fma(d.y, d.y, d.x * d.x));
    R inv_r = R(1) / std::sqrt(r2);
    R q2 = r2 * k.inv_h * k.inv_h;
    W = R(0);
    if (q2 <= R(1)) {
        R t = R(1) - q2;
        W = k.a * t * t * t;
    }
    R r = r2 * inv_r;
    R u = std::fma(-r, k.inv_h, R(1));
    R kd = R(0);
    if (k.inv_h <= inv_r) kd = k.ad * u * u;
    kg = v3<R>(R(0), R(0), R(0));
    if (r2 > R(0)) {
        V3<R> dir = mulvs(d, inv_r);
        kg = mulsv(kd, dir);
    }
}

// pbf/solver/cuda/boundary/pbf_plane_boundary.cu:6-45.  Returns the case that
// fired: 0 none, 1 the step crossed the plane, 2 both ends behind it.  In
// both cases the intersection I lies on the plane, so the correction
// -((pred - I).n) n equals -((pred - P).n) n; a nonzero `force` applies that
// closed form without classifying (replay of a recorded decision).
template <class R>
inline int response_plane(bool& hit, V3<R>& dp, V3<R> old_p, V3<R> pred, V3<R> P, V3<R> n, int force = 0) {
    V3<R> pa = old_p - P;
    V3<R> pb = pred - P;
    R pa_n = fdot(pa, n);
    R pb_n = fdot(pb, n);
    V3<R> I;
    if (force) {
        // both cases move the point by -((pred - P).n) n: the closed form,
        // free of the crossing case's 0/0 when pred lies on the plane
        hit = true;
        dp += mulsv(-pb_n, n);
        return force;
    }
    int kase = 0;
    {
        if (pa_n == R(0) && pb_n == R(0))
            kase = 0;
        else if (pa_n >= R(0) && pb_n < R(0))
            kase = 1;
        else if (pa_n < R(0) && pb_n < R(0))
            kase = 2;
    }
    if (kase == 1) {
        hit = true;
        R frac = (plant_code() == ORC_PLANT_CROSS_FRAC ? std::fabs(pb_n) : std::fabs(pa_n)) /
                 (std::fabs(pa_n) + std::fabs(pb_n));
        I = old_p + mulvs(pred - old_p, frac);
    } else if (kase == 2) {
        hit = true;
        R d = -fdot(P, n);
        R t = -(fdot(pred, n) + d) / fdot(n, n);
        I = pred + mulsv(t, n);
    } else {
        return 0;
    }
    V3<R> di = pred - I;
    R c = fdot(di, n);
    dp += mulsv(-c, n);
    return kase;
}

// pbf/solver/cuda/boundary/pbf_sphere_boundary.cu:8-34 (inner) and :36-64
// (outer).  Returns 1 if the response fired (always, with `force`).
template <class R>
inline int response_sphere(bool& hit, V3<R>& dp, V3<R> pred, V3<R> center, R radius, bool inner, bool force = false) {
    V3<R> cd = center - pred;
    R d2 = fdot(cd, cd);
    R rinv = R(1) / std::sqrt(d2);
    bool coll = force || (inner ? (d2 > radius * radius) : (d2 < radius * radius));
    if (!coll) return 0;
    hit = true;
    R nudge = (inner != (plant_code() == ORC_PLANT_NUDGE)) ? KC<R>::nudge_in : KC<R>::nudge_out;
    dp.x = std::fma(cd.x, nudge, dp.x);
    dp.y = std::fma(cd.y, nudge, dp.y);
    dp.z = std::fma(cd.z, nudge, dp.z);
    V3<R> tc = mulvs(cd, rinv);
    dp.x = std::fma(-radius, tc.x, dp.x);
    dp.y = std::fma(-radius, tc.y, dp.y);
    dp.z = std::fma(-radius, tc.z, dp.z);
    return 1;
}

// Codes of the boundary responses logged by collide(): kind << 12 | index.
enum : uint16_t { EV_INNER = 1, EV_OUTER = 2, EV_PLANE_CROSS = 3, EV_PLANE_BEHIND = 4 };
inline uint16_t ev_code(int kind, uint32_t index) { return (uint16_t)((kind << 12) | (index & 0xFFFu)); }

// pbf/solver/cuda/pbf_response_collision.cu:12-66: sweep spheres and planes
// until a sweep sees no collision (capped at max_sweeps).
template <class R>
inline int collide(const orc_params* p, const orc_boundary* b, V3<R>& dp, V3<R> pk, bool& capped,
                   std::vector<uint16_t>* ev = nullptr) {
    V3<R> pred_x = pk, old_x = pk;
    int sweeps = 0;
    capped = false;
    while (true) {
        if (sweeps >= p->max_sweeps) {
            capped = true;
            break;
        }
        ++sweeps;
        V3<R> pp = pred_x + dp;
        bool hit = false;
        if (b) {
            for (uint32_t i = 0; i < b->n_inner; ++i) {
                const float* s = b->inner + 4 * i;
                if (response_sphere<R>(hit, dp, pp, v3<R>(R(s[0]), R(s[1]), R(s[2])), R(s[3]), true) && ev)
                    ev->push_back(ev_code(EV_INNER, i));
                pp = pred_x + dp;
            }
            for (uint32_t i = 0; i < b->n_outer; ++i) {
                const float* s = b->outer + 4 * i;
                if (response_sphere<R>(hit, dp, pp, v3<R>(R(s[0]), R(s[1]), R(s[2])), R(s[3]),
                                       p->outer_as_inner != 0) &&
                    ev)
                    ev->push_back(ev_code(EV_OUTER, i));
                pp = pred_x + dp;
            }
            for (uint32_t i = 0; i < b->n_planes; ++i) {
                const float* P = b->plane_points + 3 * i;
                const float* n = b->plane_normals + 3 * i;
                const int kase = response_plane<R>(hit, dp, old_x, pp, v3<R>(R(P[0]), R(P[1]), R(P[2])),
                                                   v3<R>(R(n[0]), R(n[1]), R(n[2])));
                if (kase && ev) ev->push_back(ev_code(kase == 1 ? EV_PLANE_CROSS : EV_PLANE_BEHIND, i));
                pp = pred_x + dp;
            }
        }
        if (!hit) break;
    }
    return sweeps;
}

template <class R>
inline V3<R> load3(const std::vector<R>& a, uint32_t s) {
    return v3<R>(a[3 * (size_t)s], a[3 * (size_t)s + 1], a[3 * (size_t)s + 2]);
}
template <class R>
inline void store3(std::vector<R>& a, uint32_t s, V3<R> v) {
    a[3 * (size_t)s] = v.x;
    a[3 * (size_t)s + 1] = v.y;
    a[3 * (size_t)s + 2] = v.z;
}

// Squared pair distance beyond which a kernel evaluated with argument `arg`
// (q = r * arg) is zero: 1/arg^2 widened by 1e-5, so that every skipped pair
// has q > 1 after rounding.
inline float support_bound2(float arg) {
    const double L = 1.0 / (double)arg;
    return (float)(L * L * (1.0 + 1e-5));
}

// calcContribution (pbf/interaction/cuda/detail/pbf_contribution.inl:7-57 and
// its copies in pbf_xsph.cu:51-82, pbf_vorticity_confinement.cu:51-82,156-186):
// the 27 cells around the cell of self_pos (its CURRENT position, hashed in
// float), cell ranges from the search, per-cell partial sums added in z,y,x
// order.  Pairs with dot(d, d) > cut2 lie outside the kernel support: their
// terms are exactly zero for finite states and are not evaluated (the sums
// are unchanged; only a non-finite velocity of a far particle, which the
// reference would add as NaN, is not propagated).  pair(j, d, r2) gets
// d = x_j - x_i and r2 = dot(d, d) (glm_length(d) = sqrt(r2)).
// The cell of a final-pass position: its coordinates hashed in float.
template <class R>
inline void final_cell(V3<R> self, float cw, int g[3]) {
    grid_pos((float)self.x, (float)self.y, (float)self.z, cw, g);
}

// The sum over the 27 cells around cell g.
template <class R, class F>
inline V3<R> sum27_at(const NSData& ns, const std::vector<R>& P, V3<R> self, const int g[3], R cut2, F&& pair) {
    V3<R> sum = v3<R>(R(0), R(0), R(0));
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                uint32_t c = grid_hash(g[0] + dx, g[1] + dy, g[2] + dz, ns.G);
                V3<R> cs = v3<R>(R(0), R(0), R(0));
                for (uint32_t j = ns.off[c]; j < ns.off[c + 1]; ++j) {
                    const V3<R> d = load3(P, j) - self;
                    const R r2 = glm_dot(d, d);
                    if (r2 > cut2) continue;
                    cs += pair(j, d, r2);
                }
                sum += cs;
            }
    return sum;
}

template <class R, class F>
inline V3<R> sum27(const NSData& ns, const std::vector<R>& P, V3<R> self, R cut2, F&& pair) {
    int g[3];
    final_cell(self, ns.cw, g);
    return sum27_at(ns, P, self, g, cut2, static_cast<F&&>(pair));
}

}  // namespace orc_core
