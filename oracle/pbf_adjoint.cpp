// pbf_adjoint.cpp -- the solver step as a map with frozen discrete decisions,
// and the FP64 adjoint of that map.
//
// TEST INFRASTRUCTURE ONLY: the checker of libpbfx.so's adjoint kernels,
// never part of the library.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this code.
//
// One step (pbf::one_step, pbf/solver/pbf_solve.cpp:182-216, with the full
// update order of the commented block :135-142) is piecewise smooth.  Its
// discrete decisions are the stable sort and the capped neighbour lists
// (frozen at the predicted positions, pbf_neighbor_search.cpp:92-170), the
// boundary responses that fire in each Jacobi iteration
// (pbf_response_collision.cu:12-66), the cell each vorticity / confinement /
// XSPH sum is centred on (pbf_contribution.inl:7-57) and the XSPH NaN guard
// (pbf_xsph.cu:84-113).  With those fixed, the step is a smooth function of
// the positions, the velocities, the XSPH coefficient c and the vorticity
// epsilon.
//
// A tape records, step by step, the decisions of an FP32 run (the run is
// bit-identical to orc_step_f32) and the values it computed.  An FP64 replay
// re-evaluates the smooth map under the recorded decisions -- the map whose
// central finite differences pin the adjoint -- and can overwrite the tape's
// values with its own.  The adjoint runs backwards through the tape's values
// in double precision.  Transposed (scatter) terms go to per-thread rows that
// are reduced in thread order.

#include "pbf_oracle.h"
#include "pbf_oracle_core.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using namespace orc_core;
using V3d = V3<double>;

inline int n_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
inline int thread_id() {
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}

struct StepRec {
    uint32_t n_in = 0, n_live = 0;
    NSData ns;                       // sort, cell ranges and frozen lists of the step
    std::vector<uint32_t> ev_off;    // [iters*nl + 1]: responses of (iteration, slot)
    std::vector<uint16_t> ev;        // boundary responses in the order they fired
    std::vector<int32_t> gcur;       // [3*nl] cell the final-pass sums are centred on
    std::vector<uint8_t> xnan;       // [nl] XSPH sum was NaN: velocity kept
    // values, in sorted order (an FP32 run widened to double, or an FP64 replay)
    std::vector<double> xin;         // [3*n_in] the step's input positions (input order)
    std::vector<double> P;           // [(iters+1)*3*nl] positions entering each iteration, then the final ones
    std::vector<double> S;           // [iters*nl] scaling factor of each iteration
    std::vector<double> V, Om, Vc;   // [3*nl] velocity, vorticity, velocity after confinement
    bool f32 = false;                // the values are the FP32 run's
};

template <class R>
inline V3<R> at(const std::vector<R>& a, size_t base, uint32_t i) {
    return v3<R>(a[base + 3 * (size_t)i], a[base + 3 * (size_t)i + 1], a[base + 3 * (size_t)i + 2]);
}
inline void add3(double* a, uint32_t i, V3d v) {
    a[3 * (size_t)i] += v.x;
    a[3 * (size_t)i + 1] += v.y;
    a[3 * (size_t)i + 2] += v.z;
}
inline void sub3(double* a, uint32_t i, V3d v) {
    a[3 * (size_t)i] -= v.x;
    a[3 * (size_t)i + 1] -= v.y;
    a[3 * (size_t)i + 2] -= v.z;
}
inline V3d vget(const double* a, uint32_t i) { return v3<double>(a[3 * (size_t)i], a[3 * (size_t)i + 1], a[3 * (size_t)i + 2]); }
inline double dot(V3d a, V3d b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3d cross(V3d a, V3d b) {
    return v3<double>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline V3d scale(V3d a, double s) { return v3<double>(a.x * s, a.y * s, a.z * s); }

// Replays recorded responses with the live responses' own arithmetic.
template <class R>
void apply_events(const orc_params* p, const orc_boundary* b, V3<R>& dp, V3<R> pk, const uint16_t* e, uint32_t ne) {
    const V3<R> pred_x = pk, old_x = pk;
    V3<R> pp = pred_x + dp;
    bool hit = false;
    for (uint32_t q = 0; q < ne; ++q) {
        const int kind = e[q] >> 12;
        const uint32_t i = e[q] & 0xFFFu;
        if (kind == EV_INNER || kind == EV_OUTER) {
            const float* s = (kind == EV_INNER ? b->inner : b->outer) + 4 * i;
            const bool inner = kind == EV_INNER || p->outer_as_inner != 0;
            response_sphere<R>(hit, dp, pp, v3<R>(R(s[0]), R(s[1]), R(s[2])), R(s[3]), inner, true);
        } else {
            const float* P = b->plane_points + 3 * i;
            const float* n = b->plane_normals + 3 * i;
            response_plane<R>(hit, dp, old_x, pp, v3<R>(R(P[0]), R(P[1]), R(P[2])), v3<R>(R(n[0]), R(n[1]), R(n[2])),
                              kind == EV_PLANE_CROSS ? 1 : 2);
        }
        pp = pred_x + dp;
    }
}

// Constants of the map: the FP32 step's own constants (make_consts<float>,
// and the float factors wpoly6 / wspiky_deriv form), widened when R is
// double, so that the FP64 replay and the adjoint differentiate the FP32
// map.  coef (may be NULL): the XSPH coefficient and the vorticity epsilon
// in double, in place of the float ones of p.
template <class R>
struct MapK {
    Consts<R> k;
    R m2r2, alpha_x, K0, warg;
};
template <class R>
MapK<R> map_consts(const orc_params* p, const double* coef) {
    const Consts<float> f = make_consts<float>(p);
    MapK<R> m;
    Consts<R>& k = m.k;
    k.h = R(f.h);
    k.inv_h = R(f.inv_h);
    k.m = R(f.m);
    k.inv_rho0 = R(f.inv_rho0);
    k.relax = R(f.relax);
    k.a = R(f.a);
    k.ad = R(f.ad);
    k.inv_k = R(f.inv_k);
    k.inv_wdq = R(f.inv_wdq);
    k.sk = R(f.sk);
    k.inv_dt = R(f.inv_dt);
    k.dtc = R(f.dtc);
    k.xsph = R(f.xsph);
    k.vc = R(f.vc);
    k.mode = f.mode;
    k.n = f.n;
    if (coef) {
        k.xsph = R(coef[0]);
        k.vc = R(coef[1]);
    }
    m.m2r2 = R(f.m * f.m * f.inv_rho0 * f.inv_rho0);
    m.alpha_x = R(KC<float>::poly6 * f.inv_h * f.inv_h * f.inv_h);
    const float wf = p->vort_h_as_inv_h ? f.h : f.inv_h;
    const float w2 = wf * wf;
    m.K0 = R(KC<float>::spiky * (w2 * w2));
    m.warg = R(wf);
    return m;
}

// wpoly6(r, inv_h) (pbf_kernel.inl:34-48) with the map's alpha: for R =
// float the same operations as the FP32 step.
template <class R>
inline R xsph_w(R r, const MapK<R>& m) {
    const R q = r * m.k.inv_h;
    if (q <= R(1)) {
        const R q2 = q * q;
        const R t = R(1) - q2;
        return m.alpha_x * t * t * t;
    }
    return R(0);
}

// wspiky_deriv(r, warg) (pbf_kernel.inl:50-63) with the map's factor.
template <class R>
inline R spiky_k(R r, const MapK<R>& m) {
    const R q = r * m.warg;
    if (q <= R(1)) {
        const R t = R(1) - q;
        return m.K0 * t * t;
    }
    return R(0);
}

// One step of the map.  live: the decisions are made here (R = float: the
// FP32 step, arithmetic identical to orc_step_f32) and recorded into st;
// otherwise they are read from st.  x, v: the step's input (st.n_in rows).
template <class R>
void step_map(const orc_params* p, const orc_boundary* b, int iters, StepRec& st, bool live, bool rec, const R* x,
              const R* v, uint32_t n_in, const float lo[3], const float hi[3], std::vector<R>& xo,
              std::vector<R>& vo, const double* coef = nullptr) {
    const MapK<R> mk = map_consts<R>(p, coef);
    const Consts<R>& k = mk.k;
    NSData& ns = st.ns;
    if constexpr (std::is_same<R, float>::value) {
        if (live) {
            neighbor_search(p, x, v, n_in, lo, hi, ns);
            st.n_in = n_in;
            st.n_live = ns.n_live;
        }
    }
    const uint32_t nl = st.n_live;
    const uint32_t n = st.n_in;
    const int M = ns.max_pair;
    std::vector<R> P(3 * (size_t)nl), X(3 * (size_t)nl);
    const R dt = R(p->time_step);
#pragma omp parallel for schedule(static)
    for (int64_t ss = 0; ss < (int64_t)nl; ++ss) {
        const uint32_t s = (uint32_t)ss, i = ns.idx[s];
        for (int c = 0; c < 3; ++c) {
            if (live) {
                P[3 * (size_t)s + c] = R(ns.pred[3 * (size_t)s + c]);
            } else {
                // pbf_predict.cu:15-45: v* = fma(g, dt, v), x* = fma(v*, dt, x)
                const R vs = std::fma(R(p->body_force[c]), dt, v[3 * (size_t)i + c]);
                P[3 * (size_t)s + c] = std::fma(vs, dt, x[3 * (size_t)i + c]);
            }
            X[3 * (size_t)s + c] = x[3 * (size_t)i + c];
        }
    }
    if (rec) {
        st.f32 = std::is_same<R, float>::value;
        st.xin.assign(x, x + 3 * (size_t)n);
        st.P.assign((size_t)(iters + 1) * 3 * nl, 0.0);
        st.S.assign((size_t)iters * nl, 0.0);
    }
    if (live) {
        st.ev_off.assign((size_t)iters * nl + 1, 0u);
        st.ev.clear();
    }

    // Jacobi loop (pbf_solve.cpp:79-106)
    std::vector<R> s(nl), Pn(3 * (size_t)nl);
    std::vector<std::vector<uint16_t>> evl(live ? nl : 0);
    for (int it = 0; it < iters; ++it) {
        if (rec) std::copy(P.begin(), P.end(), st.P.begin() + (size_t)it * 3 * nl);
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)nl; ++ii) {
            const uint32_t i = (uint32_t)ii;
            const V3<R> self = load3(P, i);
            R ks = R(0), kg2s = R(0);
            V3<R> kgs = v3<R>(R(0), R(0), R(0));
            const uint32_t* lst = &ns.nb[(size_t)i * M];
            for (uint32_t c = 0; c < ns.nbc[i]; ++c) {
                R W;
                V3<R> kg;
                pair_terms<R>(self, load3(P, lst[c]), k, W, kg);
                R kg2 = fdot(kg, kg);
                ks += W;
                kgs += kg;
                kg2s += kg2;
            }
            R C = std::fma(ks * k.m, k.inv_rho0, R(-1));
            R G = std::fma(kgs.x, kgs.x, std::fma(kgs.z, kgs.z, std::fma(kgs.y, kgs.y, kg2s)));
            s[i] = C / std::fma(mk.m2r2, G, k.relax);
        }
        if (rec)
            for (uint32_t i = 0; i < nl; ++i) st.S[(size_t)it * nl + i] = double(s[i]);
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)nl; ++ii) {
            const uint32_t i = (uint32_t)ii;
            const V3<R> self = load3(P, i);
            const R si = s[i];
            const R self_cl = (si < R(0)) ? -si : R(0);
            V3<R> sum = v3<R>(R(0), R(0), R(0));
            const uint32_t* lst = &ns.nb[(size_t)i * M];
            for (uint32_t c = 0; c < ns.nbc[i]; ++c) {
                const uint32_t j = lst[c];
                R W;
                V3<R> kg;
                pair_terms<R>(self, load3(P, j), k, W, kg);
                const R sj = s[j];
                R corr;
                if (k.mode == ORC_SCORR_REFERENCE) {
                    R kr2 = W * W;
                    R kr4 = kr2 * kr2;
                    R tc = KC<R>::ten * kr4 * k.inv_k;
                    R sup = (sj < R(0)) ? -sj : R(0);
                    corr = (self_cl + sup) * tc;
                } else {
                    corr = k.sk * powi<R>(W * k.inv_wdq, k.n);
                }
                sum += mulsv(-(si + sj + corr), kg);
            }
            V3<R> dp = v3<R>(sum.x * k.m * k.inv_rho0, sum.y * k.m * k.inv_rho0, sum.z * k.m * k.inv_rho0);
            if (live) {
                evl[i].clear();
                bool capped = false;
                collide<R>(p, b, dp, self, capped, &evl[i]);
            } else {
                const size_t e0 = st.ev_off[(size_t)it * nl + i], e1 = st.ev_off[(size_t)it * nl + i + 1];
                if (e1 > e0) apply_events<R>(p, b, dp, self, &st.ev[e0], (uint32_t)(e1 - e0));
            }
            store3(Pn, i, self + dp);
        }
        if (live)
            for (uint32_t i = 0; i < nl; ++i) {
                st.ev_off[(size_t)it * nl + i + 1] = st.ev_off[(size_t)it * nl + i] + (uint32_t)evl[i].size();
                st.ev.insert(st.ev.end(), evl[i].begin(), evl[i].end());
            }
        std::copy(Pn.begin(), Pn.end(), P.begin());
    }
    if (rec) std::copy(P.begin(), P.end(), st.P.begin() + (size_t)iters * 3 * nl);

    // update (pbf_update.cu:25-39), vorticity and confinement
    // (pbf_vorticity_confinement.cu:15-222), XSPH (pbf_xsph.cu:15-113)
    std::vector<R> V(3 * (size_t)nl);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)nl; ++ii) {
        const uint32_t i = (uint32_t)ii;
        store3(V, i, mulsv(k.inv_dt, load3(P, i) - load3(X, i)));
    }
    if (live) {
        st.gcur.assign(3 * (size_t)nl, 0);
        st.xnan.assign(nl, 0);
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)nl; ++ii) {
            int g[3];
            final_cell(load3(P, (uint32_t)ii), ns.cw, g);
            for (int c = 0; c < 3; ++c) st.gcur[3 * (size_t)ii + c] = g[c];
        }
    }
    const int32_t* gc = st.gcur.data();
    const R cut_w = R(support_bound2((float)k.inv_h));
    const R cut_wd = R(support_bound2(p->vort_h_as_inv_h ? p->smoothing_length : 1.f / p->smoothing_length));
    std::vector<R> Wv(3 * (size_t)nl, R(0));
    if (rec) st.V.assign(V.begin(), V.end());
    if (p->vort_on) {
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)nl; ++ii) {
            const uint32_t i = (uint32_t)ii;
            V3<R> self = load3(P, i), selfv = load3(V, i);
            V3<R> w = sum27_at<R>(ns, P, self, gc + 3 * (size_t)i, cut_wd, [&](uint32_t j, V3<R> pd, R r2) {
                R r = std::sqrt(r2);
                V3<R> vd = load3(V, j) - selfv;
                V3<R> vort = v3<R>(R(0), R(0), R(0));
                if (r > R(0)) vort = mulvs(glm_cross(vd, divvs(pd, r)), spiky_k<R>(r, mk));
                return vort;
            });
            store3(Wv, i, mulsv(k.m * k.inv_rho0, w));
        }
        std::vector<R> Vn(V);
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)nl; ++ii) {
            const uint32_t i = (uint32_t)ii;
            V3<R> self = load3(P, i), selfw = load3(Wv, i);
            V3<R> N = sum27_at<R>(ns, P, self, gc + 3 * (size_t)i, cut_wd, [&](uint32_t j, V3<R> pd, R r2) {
                R r = std::sqrt(r2);
                R w = glm_length(load3(Wv, j));
                V3<R> loc = v3<R>(R(0), R(0), R(0));
                if (r > R(0)) loc = mulvs(mulsv(w, divvs(pd, r)), spiky_k<R>(r, mk));
                return loc;
            });
            if (glm_length(N) > R(0)) {
                N = mulvs(mulvs(N, k.m), k.inv_rho0);
                N = divvs(N, glm_length(N));
                V3<R> add = mulvs(divvs(mulsv(k.vc, glm_cross(N, selfw)), k.inv_rho0), k.dtc);
                store3(Vn, i, load3(V, i) + add);
            }
        }
        V.swap(Vn);
    }
    if (rec) {
        st.Om.assign(Wv.begin(), Wv.end());
        st.Vc.assign(V.begin(), V.end());
    }
    xo.assign(3 * (size_t)nl, R(0));
    vo.assign(3 * (size_t)nl, R(0));
    if (p->xsph_on) {
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)nl; ++ii) {
            const uint32_t i = (uint32_t)ii;
            V3<R> self = load3(P, i), selfv = load3(V, i);
            V3<R> acc = sum27_at<R>(ns, P, self, gc + 3 * (size_t)i, cut_w, [&](uint32_t j, V3<R>, R r2) {
                R r = std::sqrt(r2);
                V3<R> vd = load3(V, j) - selfv;
                return mulvs(vd, xsph_w<R>(r, mk));
            });
            bool nan = std::isnan(acc.x) || std::isnan(acc.y) || std::isnan(acc.z);
            if (live)
                st.xnan[i] = nan ? 1 : 0;
            else
                nan = st.xnan[i] != 0;
            store3(vo, i, nan ? selfv : selfv + mulsv(k.xsph * k.m * k.inv_rho0, acc));
        }
    } else {
        std::copy(V.begin(), V.end(), vo.begin());
    }
    std::copy(P.begin(), P.end(), xo.begin());
}

// Per-thread rows for transposed terms, reduced in thread order.
struct Scatter {
    int T;
    size_t n;
    std::vector<double> buf;
    explicit Scatter(size_t n_) : T(n_threads()), n(n_), buf((size_t)T * n_, 0.0) {}
    double* mine() { return buf.data() + (size_t)thread_id() * n; }
    void add_to(std::vector<double>& out) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            double a = 0.0;
            for (int t = 0; t < T; ++t) a += buf[(size_t)t * n + (size_t)i];
            out[(size_t)i] += a;
        }
    }
};

// The slots of the 27 cells around cell g (the final passes' candidate set).
template <class F>
inline void for27(const NSData& ns, const int32_t* g, F&& f) {
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const uint32_t c = grid_hash(g[0] + dx, g[1] + dy, g[2] + dz, ns.G);
                for (uint32_t j = ns.off[c]; j < ns.off[c + 1]; ++j) f(j);
            }
}

// Memo kernels of a listed pair as smooth functions of d = p_i - p_j
// (pbf_kernel_memorization.cu:15-56): W = a t^3 (t = 1 - r^2/h^2),
// kg = kd(r) d/r with kd = ad (1 - r/h)^2.  grad_d W = gw d;
// d kg / d d = kdp e e^T + (kd/r)(I - e e^T).
struct PairD {
    double r, W, gw, kd, kdp;
    V3d e, kg;
    V3d H(V3d y) const {
        const double ey = dot(e, y);
        return scale(e, kdp * ey) + scale(y - scale(e, ey), kd / r);
    }
};
inline bool pair_d(V3d d, const Consts<double>& k, PairD& o) {
    const double r2 = dot(d, d);
    if (r2 <= 0.0) return false;
    o.r = std::sqrt(r2);
    o.e = scale(d, 1.0 / o.r);
    const double q2 = r2 * k.inv_h * k.inv_h;
    o.W = 0.0;
    o.gw = 0.0;
    if (q2 <= 1.0) {
        const double t = 1.0 - q2;
        o.W = k.a * t * t * t;
        o.gw = -6.0 * k.a * k.inv_h * k.inv_h * t * t;
    }
    o.kd = 0.0;
    o.kdp = 0.0;
    if (o.r * k.inv_h <= 1.0) {
        const double u = 1.0 - o.r * k.inv_h;
        o.kd = k.ad * u * u;
        o.kdp = -2.0 * k.ad * u * k.inv_h;
    }
    o.kg = scale(o.e, o.kd);
    return true;
}

// Spiky derivative with argument warg along pd = p_j - p_i
// (pbf/kernel/cuda/detail/pbf_kernel.inl:50-63): f = K(r) pd/r,
// df/dpd = Kp e e^T + (K/r)(I - e e^T).
struct SpikyD {
    double r, K, Kp;
    V3d e;
    V3d f() const { return scale(e, K); }
    V3d H(V3d y) const {
        const double ey = dot(e, y);
        return scale(e, Kp * ey) + scale(y - scale(e, ey), K / r);
    }
};
inline bool spiky_d(V3d pd, const MapK<double>& m, SpikyD& o) {
    const double r2 = dot(pd, pd);
    if (r2 <= 0.0) return false;
    o.r = std::sqrt(r2);
    const double q = o.r * m.warg;
    if (q > 1.0) return false;
    o.e = scale(pd, 1.0 / o.r);
    o.K = m.K0 * (1.0 - q) * (1.0 - q);
    o.Kp = -2.0 * m.K0 * m.warg * (1.0 - q);
    return true;
}

// Jacobian factors of the boundary responses, in firing order: kind 0 a
// sphere ((1 - nudge) I + (R/|cd|)(I - u u^T), u = cd/|cd|), kind 1 a plane
// (I - n n^T).
struct JacF {
    int kind;
    V3d n;
    double c0, c1;
};

inline JacF sphere_factor(V3d cd, double radius, bool inner) {
    const double nudge = inner ? (double)KC<float>::nudge_in : (double)KC<float>::nudge_out;
    const double L = std::sqrt(dot(cd, cd));
    return JacF{0, scale(cd, 1.0 / L), 1.0 - nudge, radius / L};
}

// FP32 tapes: collide() exactly as the FP32 step runs it (same decisions,
// same points), logging the factors; cd is formed in double from the float
// point.
void collide_factors_f32(const orc_params* p, const orc_boundary* b, V3<float>& dp, V3<float> pk,
                         std::vector<JacF>& fs) {
    const V3<float> pred_x = pk, old_x = pk;
    int sweeps = 0;
    while (sweeps < p->max_sweeps) {
        ++sweeps;
        V3<float> pp = pred_x + dp;
        bool hit = false;
        if (b) {
            for (uint32_t q = 0; q < b->n_inner + b->n_outer; ++q) {
                const bool is_inner = q < b->n_inner;
                const float* s = is_inner ? b->inner + 4 * q : b->outer + 4 * (q - b->n_inner);
                const bool inner = is_inner || p->outer_as_inner != 0;
                const V3d cd = v3<double>((double)s[0] - pp.x, (double)s[1] - pp.y, (double)s[2] - pp.z);
                if (response_sphere<float>(hit, dp, pp, v3<float>(s[0], s[1], s[2]), s[3], inner))
                    fs.push_back(sphere_factor(cd, s[3], inner));
                pp = pred_x + dp;
            }
            for (uint32_t q = 0; q < b->n_planes; ++q) {
                const float* P = b->plane_points + 3 * q;
                const float* n = b->plane_normals + 3 * q;
                if (response_plane<float>(hit, dp, old_x, pp, v3<float>(P[0], P[1], P[2]), v3<float>(n[0], n[1], n[2])))
                    fs.push_back(JacF{1, v3<double>(n[0], n[1], n[2]), 0.0, 0.0});
                pp = pred_x + dp;
            }
        }
        if (!hit) break;
    }
}

// Adjoint of one step.  gx, gv: dL/d(x_out, v_out) [3*nl]; out: dL/d(x_in,
// v_in) [3*n_in] in the step's input order; gc, gvc accumulate dL/dc and
// dL/d(vorticity epsilon).
void step_adjoint(const orc_params* p, const orc_boundary* b, int iters, const StepRec& st, const double* gx,
                  const double* gv, std::vector<double>& gxin, std::vector<double>& gvin, double& gc, double& gvc,
                  const double* coef) {
    const MapK<double> mk = map_consts<double>(p, coef);
    const Consts<double>& k = mk.k;
    const NSData& ns = st.ns;
    const uint32_t nl = st.n_live, n = st.n_in;
    const int M = ns.max_pair;
    const size_t n3 = 3 * (size_t)nl;
    const double beta = k.m * k.inv_rho0;
    const double gamma = k.xsph * beta;
    const double kappa = k.vc * k.dtc / k.inv_rho0;
    const double cut_w = (double)support_bound2((float)k.inv_h);
    const double alpha = mk.alpha_x;
    const size_t PI = (size_t)iters * n3;  // base of the final positions in st.P
    const std::vector<double>& Pv = st.P;
    const int32_t* g = st.gcur.data();

    std::vector<double> aP(n3, 0.0), aVc(n3, 0.0), aV(n3, 0.0);

    // XSPH: v_out = Vc + gamma * sum_j (Vc_j - Vc_i) W(|p_j - p_i|)
    if (p->xsph_on) {
        Scatter sV(n3), sP(n3);
        double gcs = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : gcs)
        for (int64_t ii = 0; ii < (int64_t)nl; ++ii) {
            const uint32_t i = (uint32_t)ii;
            const V3d gi = vget(gv, i);
            if (st.xnan[i]) {
                add3(aVc.data(), i, gi);
                continue;
            }
            double* tv = sV.mine();
            double* tp = sP.mine();
            const V3d self = at(Pv, PI, i), vi = at(st.Vc, 0, i);
            double sumW = 0.0;
            V3d acc = v3<double>(0, 0, 0), apos = v3<double>(0, 0, 0);
            for27(ns, g + 3 * (size_t)i, [&](uint32_t j) {
                if (j == i) return;
                const V3d pd = at(Pv, PI, j) - self;
                const double r2 = dot(pd, pd);
                if (r2 > cut_w) return;
                const double q2 = r2 * k.inv_h * k.inv_h;
                if (q2 > 1.0) return;
                const double t = 1.0 - q2;
                const double W = alpha * t * t * t;
                const V3d e = at(st.Vc, 0, j) - vi;
                acc += scale(e, W);
                sumW += W;
                add3(tv, j, scale(gi, gamma * W));
                const V3d gr = scale(pd, -6.0 * alpha * k.inv_h * k.inv_h * t * t * gamma * dot(gi, e));
                add3(tp, j, gr);
                apos = apos - gr;
            });
            gcs += beta * dot(gi, acc);
            add3(aVc.data(), i, scale(gi, 1.0 - gamma * sumW));
            add3(aP.data(), i, apos);
        }
        sV.add_to(aVc);
        sP.add_to(aP);
        gc += gcs;
    } else {
        std::copy(gv, gv + n3, aVc.begin());
    }

    std::copy(aVc.begin(), aVc.end(), aV.begin());
    if (p->vort_on) {
        // confinement: Vc = V + kappa * (N x omega_i), N = eta / |eta|,
        // eta_i = sum_j |omega_j| K(r) e_ij
        std::vector<double> aOm(n3, 0.0), etab(n3, 0.0);
        double gvs = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : gvs)
        for (int64_t ii = 0; ii < (int64_t)nl; ++ii) {
            const uint32_t i = (uint32_t)ii;
            const V3d self = at(Pv, PI, i);
            V3d eta = v3<double>(0, 0, 0);
            for27(ns, g + 3 * (size_t)i, [&](uint32_t j) {
                SpikyD sd;
                if (!spiky_d(at(Pv, PI, j) - self, mk, sd)) return;
                eta += scale(sd.f(), std::sqrt(dot(at(st.Om, 0, j), at(st.Om, 0, j))));
            });
            const double nr = std::sqrt(dot(eta, eta));
            if (!(nr > 0.0)) continue;
            const V3d Nh = scale(eta, 1.0 / nr), om = at(st.Om, 0, i), a = vget(aVc.data(), i);
            gvs += dot(a, scale(cross(Nh, om), k.dtc / k.inv_rho0));
            const V3d y = scale(a, kappa);
            add3(aOm.data(), i, cross(y, Nh));
            const V3d nb = cross(om, y);
            const V3d eb = scale(nb - scale(Nh, dot(Nh, nb)), 1.0 / nr);
            etab[3 * (size_t)i] = eb.x;
            etab[3 * (size_t)i + 1] = eb.y;
            etab[3 * (size_t)i + 2] = eb.z;
        }
        gvc += gvs;
        {
            Scatter sM(nl), sP(n3);
#pragma omp parallel for schedule(static)
            for (int64_t ii = 0; ii < (int64_t)nl; ++ii) {
                const uint32_t i = (uint32_t)ii;
                const V3d eb = vget(etab.data(), i);
                if (eb.x == 0.0 && eb.y == 0.0 && eb.z == 0.0) continue;
                double* tm = sM.mine();
                double* tp = sP.mine();
                const V3d self = at(Pv, PI, i);
                V3d apos = v3<double>(0, 0, 0);
                for27(ns, g + 3 * (size_t)i, [&](uint32_t j) {
                    SpikyD sd;
                    if (!spiky_d(at(Pv, PI, j) - self, mk, sd)) return;
                    tm[j] += dot(eb, sd.f());
                    const V3d pdb = scale(sd.H(eb), std::sqrt(dot(at(st.Om, 0, j), at(st.Om, 0, j))));
                    add3(tp, j, pdb);
                    apos = apos - pdb;
                });
                add3(aP.data(), i, apos);
            }
            std::vector<double> mb(nl, 0.0);
            sM.add_to(mb);
            sP.add_to(aP);
            for (uint32_t j = 0; j < nl; ++j) {
                const V3d om = at(st.Om, 0, j);
                const double nm = std::sqrt(dot(om, om));
                if (nm > 0.0) add3(aOm.data(), j, scale(om, mb[j] / nm));
            }
        }
        // vorticity: omega_i = beta * sum_j (V_j - V_i) x (K(r) e_ij)
        {
            Scatter sV(n3), sP(n3);
#pragma omp parallel for schedule(static)
            for (int64_t ii = 0; ii < (int64_t)nl; ++ii) {
                const uint32_t i = (uint32_t)ii;
                const V3d ab = vget(aOm.data(), i);
                if (ab.x == 0.0 && ab.y == 0.0 && ab.z == 0.0) continue;
                double* tv = sV.mine();
                double* tp = sP.mine();
                const V3d self = at(Pv, PI, i), vi = at(st.V, 0, i);
                V3d av = v3<double>(0, 0, 0), apos = v3<double>(0, 0, 0);
                for27(ns, g + 3 * (size_t)i, [&](uint32_t j) {
                    SpikyD sd;
                    if (!spiky_d(at(Pv, PI, j) - self, mk, sd)) return;
                    const V3d e = at(st.V, 0, j) - vi;
                    const V3d ebar = scale(cross(sd.f(), ab), beta);
                    add3(tv, j, ebar);
                    av = av - ebar;
                    const V3d pdb = sd.H(scale(cross(ab, e), beta));
                    add3(tp, j, pdb);
                    apos = apos - pdb;
                });
                add3(aV.data(), i, av);
                add3(aP.data(), i, apos);
            }
            sV.add_to(aV);
            sP.add_to(aP);
        }
    }

    // V = inv_dt (P - X); x_out = P
    std::vector<double> aX(n3, 0.0);
    for (size_t q = 0; q < n3; ++q) {
        aP[q] += k.inv_dt * aV[q] + gx[q];
        aX[q] = -k.inv_dt * aV[q];
    }

    // Jacobi iterations, last first
    const double m2r2 = mk.m2r2;
    std::vector<double> abar(n3), sbar(nl), aPn(n3);
    for (int it = iters - 1; it >= 0; --it) {
        const size_t PB = (size_t)it * n3;
        const double* s = st.S.data() + (size_t)it * nl;
        auto corr_of = [&](double W, double si, double sj) {
            if (k.mode == ORC_SCORR_REFERENCE) {
                const double cl = (si < 0 ? -si : 0.0) + (sj < 0 ? -sj : 0.0);
                return cl * 10.0 * W * W * W * W * k.inv_k;
            }
            return k.sk * std::pow(W * k.inv_wdq, k.n);
        };
        // collision: P' = events(P + dp); every response's Jacobian is
        // symmetric: planes I - n n^T, spheres (1 - nudge) I + (R/|cd|)(I - u u^T)
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)nl; ++ii) {
            const uint32_t i = (uint32_t)ii;
            const V3d self = at(Pv, PB, i);
            V3d a = vget(aP.data(), i);
            const size_t e0 = st.ev_off[(size_t)it * nl + i], e1 = st.ev_off[(size_t)it * nl + i + 1];
            if (e1 > e0) {
                std::vector<JacF> fs;
                if (st.f32) {
                    // the FP32 step's own delta p (pbf_position_update.cu:131-188)
                    const Consts<float> kf = make_consts<float>(p);
                    auto ldf = [&](uint32_t q) {
                        return v3<float>((float)Pv[PB + 3 * (size_t)q], (float)Pv[PB + 3 * (size_t)q + 1],
                                         (float)Pv[PB + 3 * (size_t)q + 2]);
                    };
                    const V3<float> sf = ldf(i);
                    const float si = (float)s[i];
                    const float self_cl = (si < 0.f) ? -si : 0.f;
                    V3<float> sum = v3<float>(0.f, 0.f, 0.f);
                    const uint32_t* lst = &ns.nb[(size_t)i * M];
                    for (uint32_t c = 0; c < ns.nbc[i]; ++c) {
                        const uint32_t j = lst[c];
                        float W;
                        V3<float> kg;
                        pair_terms<float>(sf, ldf(j), kf, W, kg);
                        const float sj = (float)s[j];
                        float corr;
                        if (kf.mode == ORC_SCORR_REFERENCE) {
                            const float kr2 = W * W;
                            const float kr4 = kr2 * kr2;
                            const float tc = KC<float>::ten * kr4 * kf.inv_k;
                            const float sup = (sj < 0.f) ? -sj : 0.f;
                            corr = (self_cl + sup) * tc;
                        } else {
                            corr = kf.sk * powi<float>(W * kf.inv_wdq, kf.n);
                        }
                        sum += mulsv(-(si + sj + corr), kg);
                    }
                    V3<float> dp = v3<float>(sum.x * kf.m * kf.inv_rho0, sum.y * kf.m * kf.inv_rho0,
                                             sum.z * kf.m * kf.inv_rho0);
                    collide_factors_f32(p, b, dp, sf, fs);
                } else {
                    // FP64 tape: the recorded responses at the FP64 points
                    V3d sum = v3<double>(0, 0, 0);
                    const uint32_t* lst = &ns.nb[(size_t)i * M];
                    for (uint32_t c = 0; c < ns.nbc[i]; ++c) {
                        const uint32_t j = lst[c];
                        PairD pd;
                        if (!pair_d(self - at(Pv, PB, j), k, pd)) continue;
                        sum += scale(pd.kg, -(s[i] + s[j] + corr_of(pd.W, s[i], s[j])));
                    }
                    V3d pp = self + scale(sum, beta);
                    for (size_t q = e0; q < e1; ++q) {
                        const int kind = st.ev[q] >> 12;
                        const uint32_t idx = st.ev[q] & 0xFFFu;
                        if (kind == EV_INNER || kind == EV_OUTER) {
                            const float* sp = (kind == EV_INNER ? b->inner : b->outer) + 4 * idx;
                            const bool inner = kind == EV_INNER || p->outer_as_inner != 0;
                            const double nudge = inner ? (double)KC<float>::nudge_in : (double)KC<float>::nudge_out;
                            const V3d cd = v3<double>(sp[0], sp[1], sp[2]) - pp;
                            const double L = std::sqrt(dot(cd, cd));
                            fs.push_back(sphere_factor(cd, sp[3], inner));
                            pp = pp + scale(cd, nudge) - scale(cd, (double)sp[3] / L);
                        } else {
                            const float* P0 = b->plane_points + 3 * idx;
                            const float* nn = b->plane_normals + 3 * idx;
                            const V3d nv = v3<double>(nn[0], nn[1], nn[2]);
                            fs.push_back(JacF{1, nv, 0, 0});
                            pp = pp - scale(nv, dot(pp - v3<double>(P0[0], P0[1], P0[2]), nv));
                        }
                    }
                }
                for (size_t q = fs.size(); q-- > 0;) {
                    const JacF& f = fs[q];
                    if (f.kind == 1)
                        a = a - scale(f.n, dot(f.n, a));
                    else
                        a = scale(a, f.c0) + scale(a - scale(f.n, dot(f.n, a)), f.c1);
                }
            }
            abar[3 * (size_t)i] = a.x;
            abar[3 * (size_t)i + 1] = a.y;
            abar[3 * (size_t)i + 2] = a.z;
        }
        // delta p = beta * sum_j -(s_i + s_j + corr(W_ij)) kg_ij
        std::fill(sbar.begin(), sbar.end(), 0.0);
        std::fill(aPn.begin(), aPn.end(), 0.0);
        {
            Scatter sS(nl), sP(n3);
#pragma omp parallel for schedule(static)
            for (int64_t ii = 0; ii < (int64_t)nl; ++ii) {
                const uint32_t i = (uint32_t)ii;
                double* ts = sS.mine();
                double* tp = sP.mine();
                const V3d self = at(Pv, PB, i), a = vget(abar.data(), i);
                double sb = 0.0;
                V3d apos = v3<double>(0, 0, 0);
                const uint32_t* lst = &ns.nb[(size_t)i * M];
                for (uint32_t c = 0; c < ns.nbc[i]; ++c) {
                    const uint32_t j = lst[c];
                    PairD pd;
                    if (!pair_d(self - at(Pv, PB, j), k, pd)) continue;
                    const double A = -(s[i] + s[j] + corr_of(pd.W, s[i], s[j]));
                    const double Ab = beta * dot(pd.kg, a);
                    sb -= Ab;
                    ts[j] -= Ab;
                    const double cb = -Ab;
                    double Wb;
                    if (k.mode == ORC_SCORR_REFERENCE) {
                        const double W3 = pd.W * pd.W * pd.W;
                        const double tc = 10.0 * W3 * pd.W * k.inv_k;
                        if (s[i] < 0) sb -= cb * tc;
                        if (s[j] < 0) ts[j] -= cb * tc;
                        const double cl = (s[i] < 0 ? -s[i] : 0.0) + (s[j] < 0 ? -s[j] : 0.0);
                        Wb = cb * cl * 40.0 * W3 * k.inv_k;
                    } else {
                        Wb = cb * k.sk * k.n * std::pow(pd.W * k.inv_wdq, k.n - 1) * k.inv_wdq;
                    }
                    const V3d db = scale(pd.H(a), beta * A) + scale(self - at(Pv, PB, j), Wb * pd.gw);
                    apos += db;
                    sub3(tp, j, db);
                }
                sbar[i] += sb;
                add3(aPn.data(), i, apos);
            }
            sS.add_to(sbar);
            sP.add_to(aPn);
        }
        // s = C / (m2r2 G + eps), C = beta sum W - 1, G = |sum kg|^2 + sum |kg|^2
        {
            Scatter sP(n3);
#pragma omp parallel for schedule(static)
            for (int64_t ii = 0; ii < (int64_t)nl; ++ii) {
                const uint32_t i = (uint32_t)ii;
                if (sbar[i] == 0.0) continue;
                double* tp = sP.mine();
                const V3d self = at(Pv, PB, i);
                const uint32_t* lst = &ns.nb[(size_t)i * M];
                double kg2s = 0.0;
                V3d kgs = v3<double>(0, 0, 0);
                for (uint32_t c = 0; c < ns.nbc[i]; ++c) {
                    PairD pd;
                    const V3d d = self - at(Pv, PB, lst[c]);
                    if (!pair_d(d, k, pd)) continue;  // r = 0: kg = 0
                    kgs += pd.kg;
                    kg2s += dot(pd.kg, pd.kg);
                }
                const double D = m2r2 * (dot(kgs, kgs) + kg2s) + k.relax;
                const double ksb = beta * sbar[i] / D;
                const double Gb = -m2r2 * sbar[i] * s[i] / D;
                V3d apos = v3<double>(0, 0, 0);
                for (uint32_t c = 0; c < ns.nbc[i]; ++c) {
                    const uint32_t j = lst[c];
                    PairD pd;
                    const V3d d = self - at(Pv, PB, j);
                    if (!pair_d(d, k, pd)) continue;
                    const V3d db = scale(d, ksb * pd.gw) + pd.H(scale(kgs + pd.kg, 2.0 * Gb));
                    apos += db;
                    sub3(tp, j, db);
                }
                add3(aPn.data(), i, apos);
            }
            sP.add_to(aPn);
        }
        for (size_t q = 0; q < n3; ++q) aP[q] = abar[q] + aPn[q];
    }

    // sort and predict: P0_s = x*_{idx[s]}, x* = x + dt (v + g dt); X_s = x_{idx[s]}
    const double dt = p->time_step;
    gxin.assign(3 * (size_t)n, 0.0);
    gvin.assign(3 * (size_t)n, 0.0);
    for (uint32_t s2 = 0; s2 < nl; ++s2) {
        const uint32_t i = ns.idx[s2];
        for (int c = 0; c < 3; ++c) {
            gxin[3 * (size_t)i + c] = aP[3 * (size_t)s2 + c] + aX[3 * (size_t)s2 + c];
            gvin[3 * (size_t)i + c] = dt * aP[3 * (size_t)s2 + c];
        }
    }
}

void set_threads_from_oracle() {
#ifdef _OPENMP
    const int t = orc_get_threads();
    if (t > 0) omp_set_num_threads(t);
#endif
}

}  // namespace

struct orc_tape {
    int iters = 0;
    std::vector<StepRec> steps;
};

extern "C" {

orc_tape* orc_tape_new(void) { return new orc_tape(); }
void orc_tape_free(orc_tape* t) { delete t; }
uint32_t orc_tape_steps(const orc_tape* t) { return t ? (uint32_t)t->steps.size() : 0u; }

uint32_t orc_tape_num(const orc_tape* t, uint32_t step, int which) {
    if (!t || step >= t->steps.size()) return 0;
    return which == 0 ? t->steps[step].n_in : t->steps[step].n_live;
}

int orc_adj_step_f32(const orc_params* p, const orc_boundary* b, float* x, float* v, uint32_t* num,
                     const float lo[3], const float hi[3], int iters, orc_tape* t) {
    if (!p || !x || !v || !num || !t || p->max_pair < 2) return 1;
    if (!t->steps.empty() && t->iters != iters) return 2;
    set_threads_from_oracle();
    t->iters = iters;
    t->steps.emplace_back();
    StepRec& st = t->steps.back();
    std::vector<float> xo, vo;
    step_map<float>(p, b, iters, st, true, true, x, v, *num, lo, hi, xo, vo);
    std::memcpy(x, xo.data(), sizeof(float) * xo.size());
    std::memcpy(v, vo.data(), sizeof(float) * vo.size());
    *num = st.n_live;
    return 0;
}

int orc_adj_replay_f64(const orc_params* p, const orc_boundary* b, orc_tape* t, const double* x0, const double* v0,
                       const double* coef, double* x_out, double* v_out, int record) {
    if (!p || !t || t->steps.empty() || !x0 || !v0) return 1;
    set_threads_from_oracle();
    const float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    std::vector<double> x(x0, x0 + 3 * (size_t)t->steps[0].n_in), v(v0, v0 + 3 * (size_t)t->steps[0].n_in);
    std::vector<double> xo, vo;
    for (StepRec& st : t->steps) {
        if (x.size() != 3 * (size_t)st.n_in) return 3;
        step_map<double>(p, b, t->iters, st, false, record != 0, x.data(), v.data(), st.n_in, lo, hi, xo, vo, coef);
        x.swap(xo);
        v.swap(vo);
    }
    if (x_out) std::memcpy(x_out, x.data(), sizeof(double) * x.size());
    if (v_out) std::memcpy(v_out, v.data(), sizeof(double) * v.size());
    return 0;
}

int orc_adj_backward_f64(const orc_params* p, const orc_boundary* b, const orc_tape* t, const double* coef,
                         const double* gx, const double* gv, double* gx0, double* gv0, double gparam[2],
                         int32_t* exponent) {
    if (!p || !t || t->steps.empty() || !gx || !gv) return 1;
    set_threads_from_oracle();
    const StepRec& last = t->steps.back();
    std::vector<double> ax(gx, gx + 3 * (size_t)last.n_live), av(gv, gv + 3 * (size_t)last.n_live);
    std::vector<double> nx, nv;
    double gc = 0.0, gvc = 0.0;
    int32_t E = 0;
    for (size_t q = t->steps.size(); q-- > 0;) {
        step_adjoint(p, b, t->iters, t->steps[q], ax.data(), av.data(), nx, nv, gc, gvc, coef);
        ax.swap(nx);
        av.swap(nv);
        // the adjoint grows by up to ~1e8 per step in exploding flows: keep it
        // near 1 with a power-of-two scale once it passes 2^20
        double m = 0.0;
        for (double z : ax) m = std::max(m, std::fabs(z));
        for (double z : av) m = std::max(m, std::fabs(z));
        if (std::isfinite(m) && m > 1048576.0) {
            const int k = std::ilogb(m);
            for (double& z : ax) z = std::ldexp(z, -k);
            for (double& z : av) z = std::ldexp(z, -k);
            gc = std::ldexp(gc, -k);
            gvc = std::ldexp(gvc, -k);
            E += k;
        }
    }
    if (exponent) *exponent = E;
    if (gx0) std::memcpy(gx0, ax.data(), sizeof(double) * ax.size());
    if (gv0) std::memcpy(gv0, av.data(), sizeof(double) * av.size());
    if (gparam) {
        gparam[0] = gc;
        gparam[1] = gvc;
    }
    return 0;
}

int orc_adj_identity(const orc_tape* t, uint32_t* id) {
    if (!t || t->steps.empty() || !id) return 1;
    std::vector<uint32_t> cur(t->steps[0].n_in);
    for (uint32_t i = 0; i < cur.size(); ++i) cur[i] = i;
    for (const StepRec& st : t->steps) {
        std::vector<uint32_t> nx(st.n_live);
        for (uint32_t s = 0; s < st.n_live; ++s) nx[s] = cur[st.ns.idx[s]];
        cur.swap(nx);
    }
    std::memcpy(id, cur.data(), sizeof(uint32_t) * cur.size());
    return 0;
}

}  // extern "C"
