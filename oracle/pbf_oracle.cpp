// pbf_oracle.cpp -- CPU restatement of the reference PBF solver step.
//
// TEST INFRASTRUCTURE ONLY: the checker for libpbfx.so, never part of it.
// Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
// may load this library.
//
// Every function cites the reference file:line it restates; paths are
// relative to cuda_pbf_with_ssf/src/ of yodaplusplus/ssf_pbf_cuda.
// Arithmetic is IEEE (built with -ffp-contract=off, no fast math); an
// explicit fma appears exactly where the reference calls fmaf, and GLM
// expressions keep GLM 0.9.7.1's evaluation order
// (external/include/glm/detail/func_geometric.inl:65-72,95-100,134-142):
//   dot(a,b)   = (a.x*b.x + a.y*b.y) + a.z*b.z
//   length(a)  = sqrt(dot(a,a))
//   cross(x,y) = (x.y*y.z - y.y*x.z, x.z*y.x - y.z*x.x, x.x*y.y - y.x*x.y)
//
// Pinning (tests/test_oracle_*.py): closed forms of the SPH kernels, the
// known-answer values computed from the reference's own header maths,
// brute-force neighbour sets, numpy stable argsort, finite differences of
// the spiky kernel and of the constraint, and conservation invariants.

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

struct SolveStats {
    uint32_t sweep_cap_hits = 0, max_sweeps_used = 0, xsph_nan = 0;
};

// Jacobi loop (pbf_solve.cpp:79-106: memo, scaling factor, position update,
// collision, interim update) and update (pbf_solve.cpp:125-145; the full
// order is the commented block :135-142: velocity, vorticity, confinement,
// XSPH).  P is the sorted predicted state (AoS, R), X the sorted old state.
template <class R>
void solve_and_update(const orc_params* p, const orc_boundary* b, const NSData& ns, std::vector<R>& P,
                      const std::vector<R>& X, int iters, std::vector<R>& x_out, std::vector<R>& v_out,
                      float* lambda_last, SolveStats& st) {
    const Consts<R> k = make_consts<R>(p);
    const uint32_t nl = ns.n_live;
    const int M = ns.max_pair;
    std::vector<R> s(nl), Pn(3 * (size_t)nl);
    uint32_t cap_hits = 0, max_sw = 0;

    for (int it = 0; it < iters; ++it) {
        // calcScalingFactorCUDA (memo) pbf/solver/cuda/pbf_constraint.cu:168-209
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)nl; ++ii) {
            uint32_t i = (uint32_t)ii;
            V3<R> self = load3(P, i);
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
            R m2r2 = k.m * k.m * k.inv_rho0 * k.inv_rho0;
            s[i] = C / std::fma(m2r2, G, k.relax);
        }
        // calcPositionUpdateCUDA (memo) pbf/solver/cuda/pbf_position_update.cu:131-188,
        // then responseCollision and updateInterimPosition (pbf_constraint.cu:335-345)
#pragma omp parallel for schedule(static) reduction(+ : cap_hits) reduction(max : max_sw)
        for (int64_t ii = 0; ii < (int64_t)nl; ++ii) {
            uint32_t i = (uint32_t)ii;
            V3<R> self = load3(P, i);
            R si = s[i];
            R self_cl = (si < R(0)) ? -si : R(0);
            V3<R> sum = v3<R>(R(0), R(0), R(0));
            const uint32_t* lst = &ns.nb[(size_t)i * M];
            for (uint32_t c = 0; c < ns.nbc[i]; ++c) {
                uint32_t j = lst[c];
                R W;
                V3<R> kg;
                pair_terms<R>(self, load3(P, j), k, W, kg);
                R sj = s[j];
                R corr;
                if (k.mode == ORC_SCORR_REFERENCE) {
                    R kr2 = W * W;
                    R kr4 = kr2 * kr2;
                    R tc = KC<R>::ten * kr4 * k.inv_k;
                    R sup = (sj < R(0)) ? -sj : R(0);
                    R cl = self_cl + sup;
                    corr = cl * tc;
                } else {
                    corr = k.sk * powi<R>(W * k.inv_wdq, k.n);
                }
                R a = -(si + sj + corr);
                sum += mulsv(a, kg);
            }
            V3<R> dp = v3<R>(sum.x * k.m * k.inv_rho0, sum.y * k.m * k.inv_rho0, sum.z * k.m * k.inv_rho0);
            bool capped = false;
            int sw = collide<R>(p, b, dp, self, capped);
            if (capped) cap_hits++;
            max_sw = std::max<uint32_t>(max_sw, (uint32_t)sw);
            store3(Pn, i, self + dp);
        }
        std::copy(Pn.begin(), Pn.end(), P.begin());
    }
    st.sweep_cap_hits = cap_hits;
    st.max_sweeps_used = max_sw;
    if (lambda_last)
        for (uint32_t i = 0; i < nl; ++i) lambda_last[i] = (float)s[i];

    // pbf/solver/cuda/pbf_update.cu:25-39: v = inv_dt * (x* - x_sorted)
    std::vector<R> V(3 * (size_t)nl);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)nl; ++ii) {
        uint32_t i = (uint32_t)ii;
        store3(V, i, mulsv(k.inv_dt, load3(P, i) - load3(X, i)));
    }

    const int plant = plant_code();
    const R warg = p->vort_h_as_inv_h ? k.h : k.inv_h;
    const R cut_w = R(support_bound2((float)k.inv_h));
    const R cut_wd = R(support_bound2(p->vort_h_as_inv_h ? p->smoothing_length : 1.f / p->smoothing_length));
    if (p->vort_on) {
        // calcVorticityCUDA pbf/solver/cuda/pbf_vorticity_confinement.cu:15-115
        std::vector<R> Wv(3 * (size_t)nl);
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)nl; ++ii) {
            uint32_t i = (uint32_t)ii;
            V3<R> self = load3(P, i), selfv = load3(V, i);
            V3<R> w = sum27<R>(ns, P, self, cut_wd, [&](uint32_t j, V3<R> pd, R r2) {
                R r = std::sqrt(r2);
                V3<R> vd = load3(V, j) - selfv;
                V3<R> vort = v3<R>(R(0), R(0), R(0));
                if (r > R(0)) {
                    V3<R> dir = divvs(pd, r);
                    vort = mulvs(plant == ORC_PLANT_VORT_CROSS ? glm_cross(dir, vd) : glm_cross(vd, dir),
                                 wspiky_deriv<R>(r, warg));
                }
                return vort;
            });
            store3(Wv, i, plant == ORC_PLANT_VORT_NO_MR ? w : mulsv(k.m * k.inv_rho0, w));
        }
        // addVorticityConfinementCUDA pbf_vorticity_confinement.cu:121-222
        std::vector<R> Vn(V);
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)nl; ++ii) {
            uint32_t i = (uint32_t)ii;
            V3<R> self = load3(P, i), selfw = load3(Wv, i);
            V3<R> N = sum27<R>(ns, P, self, cut_wd, [&](uint32_t j, V3<R> pd, R r2) {
                R r = std::sqrt(r2);
                R w = glm_length(load3(Wv, plant == ORC_PLANT_LOC_SELF ? i : j));
                V3<R> loc = v3<R>(R(0), R(0), R(0));
                if (r > R(0)) {
                    V3<R> dir = divvs(pd, r);
                    loc = mulvs(mulsv(w, dir), wspiky_deriv<R>(r, warg));
                }
                return loc;
            });
            if (glm_length(N) > R(0)) {
                N = mulvs(mulvs(N, k.m), k.inv_rho0);
                N = divvs(N, glm_length(N));
                const V3<R> nw = plant == ORC_PLANT_CONF_CROSS ? glm_cross(selfw, N) : glm_cross(N, selfw);
                V3<R> add = plant == ORC_PLANT_CONF_RHO ? mulvs(mulvs(mulsv(k.vc, nw), k.inv_rho0), k.dtc)
                                                        : mulvs(divvs(mulsv(k.vc, nw), k.inv_rho0), k.dtc);
                store3(Vn, i, load3(V, i) + add);
            }
        }
        V.swap(Vn);
    }

    x_out.assign(3 * (size_t)nl, R(0));
    v_out.assign(3 * (size_t)nl, R(0));
    uint32_t nan_count = 0;
    if (p->xsph_on) {
        // calcXSPHCUDA pbf/solver/cuda/pbf_xsph.cu:15-113
#pragma omp parallel for schedule(static) reduction(+ : nan_count)
        for (int64_t ii = 0; ii < (int64_t)nl; ++ii) {
            uint32_t i = (uint32_t)ii;
            V3<R> self = load3(P, i), selfv = load3(V, i);
            const R cut_x = plant == ORC_PLANT_XSPH_H ? R(support_bound2(p->smoothing_length)) : cut_w;
            V3<R> acc = sum27<R>(ns, P, self, cut_x, [&](uint32_t j, V3<R>, R r2) {
                R r = std::sqrt(r2);
                V3<R> vd = plant == ORC_PLANT_XSPH_SIGN ? selfv - load3(V, j) : load3(V, j) - selfv;
                R kr = wpoly6<R>(r, plant == ORC_PLANT_XSPH_H ? k.h : k.inv_h);
                return mulvs(vd, kr);
            });
            V3<R> out;
            if (std::isnan(acc.x) || std::isnan(acc.y) || std::isnan(acc.z)) {
                out = selfv;
                nan_count++;
            } else {
                out = selfv + mulsv(plant == ORC_PLANT_XSPH_NO_MR ? k.xsph : k.xsph * k.m * k.inv_rho0, acc);
            }
            store3(v_out, i, out);
        }
    } else {
        std::copy(V.begin(), V.end(), v_out.begin());
    }
    std::copy(P.begin(), P.begin() + 3 * (size_t)nl, x_out.begin());
    st.xsph_nan = nan_count;
}

int g_threads = 0;

void apply_threads() {
#ifdef _OPENMP
    if (g_threads > 0) omp_set_num_threads(g_threads);
#endif
}

}  // namespace

extern "C" {

void orc_set_threads(int n) { g_threads = n; }

void orc_set_plant(int code) { plant_code() = code; }

void orc_arith_f32(const float* a, const float* b, float* out, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        out[4 * i + 0] = a[i] / b[i];
        out[4 * i + 1] = std::sqrt(a[i]);
        out[4 * i + 2] = std::fma(a[i], b[i], a[i]);
        out[4 * i + 3] = std::floor(a[i] / b[i]);
    }
}

int orc_get_threads(void) {
#ifdef _OPENMP
    apply_threads();
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// pbf/pbf_type.cpp:141-162
int orc_calc_grid_size(double cell_size, const float end[3], uint32_t out[3]) {
    uint32_t g[3] = {1, 1, 1};
    while (true) {
        if (g[0] * cell_size > end[0] && g[1] * cell_size > end[1] && g[2] * cell_size > end[2]) break;
        for (int a = 0; a < 3; ++a) g[a] = (g[a] * cell_size > end[a]) ? g[a] : g[a] * 2;
        if (g[0] > 1028 || g[1] > 1028 || g[2] > 1028) return 1;
    }
    out[0] = g[0];
    out[1] = g[1];
    out[2] = g[2];
    return 0;
}

float orc_wpoly6(float r, float inv_h) { return wpoly6<float>(r, inv_h); }
float orc_wspiky_deriv(float r, float inv_h) { return wspiky_deriv<float>(r, inv_h); }
double orc_wpoly6_f64(double r, double inv_h) { return wpoly6<double>(r, inv_h); }
double orc_wspiky_deriv_f64(double r, double inv_h) { return wspiky_deriv<double>(r, inv_h); }

void orc_hash(const float* x, uint32_t n, float cw, const uint32_t grid[3], uint32_t* key) {
    for (uint32_t i = 0; i < n; ++i) {
        int g[3];
        grid_pos(x[3 * i], x[3 * i + 1], x[3 * i + 2], cw, g);
        key[i] = grid_hash(g[0], g[1], g[2], grid);
    }
}

// pbf_solve.cpp:90-93: static k = pow(weight(h*0.4f, inv_h), 4) in double,
// narrowed to float by setConstantMemory, then inv_k = 1.f / k
// (pbf_position_update.cu:224-238).  PAPER: inv_wdq = 1/W(dq).
void orc_scorr_constants(const orc_params* p, float* ref_inv_k, float* paper_inv_wdq) {
    const float h = p->smoothing_length;
    const float inv_h = 1.f / h;
    const float w04 = wpoly6<float>(h * 0.4f, inv_h);
    const float kf = (float)std::pow((double)w04, 4);
    if (ref_inv_k) *ref_inv_k = 1.f / kf;
    const float wdq = wpoly6<float>(p->scorr_dq_over_h * h, inv_h);
    if (paper_inv_wdq) *paper_inv_wdq = 1.f / wdq;
}

int orc_step_f32(const orc_params* p, const orc_boundary* b, float* x, float* v, uint32_t* num,
                 const float lo[3], const float hi[3], int iters, orc_ns_out* nso, orc_stats* st) {
    if (!p || !x || !v || !num || p->max_pair < 2) return 1;
    apply_threads();
    const uint32_t n = *num;
    NSData ns;
    neighbor_search(p, x, v, n, lo, hi, ns);
    const uint32_t nl = ns.n_live;

    std::vector<float> P(3 * (size_t)nl), X(3 * (size_t)nl);
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < (int64_t)nl; ++s) {
        uint32_t i = ns.idx[s];
        for (int c = 0; c < 3; ++c) {
            P[3 * s + c] = ns.pred[3 * s + c];
            X[3 * s + c] = x[3 * i + c];
        }
    }
    if (nso) {
        if (nso->key) std::memcpy(nso->key, ns.key.data(), sizeof(uint32_t) * n);
        if (nso->sorted_index) std::memcpy(nso->sorted_index, ns.idx.data(), sizeof(uint32_t) * n);
        if (nso->cell_offset) std::memcpy(nso->cell_offset, ns.off.data(), sizeof(uint32_t) * ns.off.size());
        if (nso->neighbors) {
            std::fill(nso->neighbors, nso->neighbors + (size_t)n * ns.max_pair, ORC_EMPTY);
            std::memcpy(nso->neighbors, ns.nb.data(), sizeof(uint32_t) * ns.nb.size());
        }
        if (nso->neighbor_count) {
            std::fill(nso->neighbor_count, nso->neighbor_count + n, 0u);
            std::memcpy(nso->neighbor_count, ns.nbc.data(), sizeof(uint32_t) * nl);
        }
        if (nso->pred0) std::memcpy(nso->pred0, ns.pred.data(), sizeof(float) * 3 * (size_t)n);
    }
    std::vector<float> xo, vo;
    SolveStats ss;
    solve_and_update<float>(p, b, ns, P, X, iters, xo, vo, nso ? nso->lambda_last : nullptr, ss);
    std::memcpy(x, xo.data(), sizeof(float) * xo.size());
    std::memcpy(v, vo.data(), sizeof(float) * vo.size());
    *num = nl;
    if (st) {
        st->num_in = n;
        st->num_live = nl;
        st->num_deleted = n - nl;
        st->truncated = ns.truncated;
        st->sweep_cap_hits = ss.sweep_cap_hits;
        st->max_sweeps_used = ss.max_sweeps_used;
        st->xsph_nan = ss.xsph_nan;
        st->reserved = 0;
    }
    return 0;
}

int orc_step_f64(const orc_params* p, const orc_boundary* b, const float* x, const float* v, uint32_t num,
                 const float lo[3], const float hi[3], int iters, double* x_out, double* v_out,
                 uint32_t* num_out) {
    if (!p || !x || !v || p->max_pair < 2) return 1;
    apply_threads();
    NSData ns;
    neighbor_search(p, x, v, num, lo, hi, ns);
    const uint32_t nl = ns.n_live;
    std::vector<double> P(3 * (size_t)nl), X(3 * (size_t)nl);
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < (int64_t)nl; ++s) {
        uint32_t i = ns.idx[s];
        double q[3];
        predict_one<double>(p, x, v, i, q);
        for (int c = 0; c < 3; ++c) {
            P[3 * s + c] = q[c];
            X[3 * s + c] = (double)x[3 * i + c];
        }
    }
    std::vector<double> xo, vo;
    SolveStats ss;
    solve_and_update<double>(p, b, ns, P, X, iters, xo, vo, nullptr, ss);
    std::memcpy(x_out, xo.data(), sizeof(double) * xo.size());
    std::memcpy(v_out, vo.data(), sizeof(double) * vo.size());
    if (num_out) *num_out = nl;
    return 0;
}

// calcDensity (pbf/solver/cuda/pbf_density.cu:11-71) over given neighbour
// lists: rho_i = sum over the listed j, in list order, of W(|x_i - x_j|) * m,
// with inv_h as the kernel argument (the reference passes h, pbf_density.cu:22).
int orc_density(const orc_params* p, const float* x, uint32_t n, const uint32_t* neighbors, float* rho) {
    const float inv_h = 1.f / p->smoothing_length;
    const uint32_t mp = (uint32_t)p->max_pair;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const V3<float> self = v3<float>(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
        float sum = 0.f;
        for (uint32_t c = 0; c < mp; ++c) {
            const uint32_t j = neighbors[(size_t)i * mp + c];
            if (j == ORC_EMPTY) break;
            const V3<float> d = self - v3<float>(x[3 * j], x[3 * j + 1], x[3 * j + 2]);
            sum += wpoly6<float>(glm_length(d), inv_h) * p->particle_mass;
        }
        rho[i] = sum;
    }
    return 0;
}

int orc_metrics(const orc_params* p, const float* x, const float* v, uint32_t n, double out[4]) {
    apply_threads();
    const uint32_t G[3] = {p->grid[0], p->grid[1], p->grid[2]};
    const uint32_t C = G[0] * G[1] * G[2];
    const float cw = p->smoothing_length;
    std::vector<uint32_t> key(n), off((size_t)C + 1, 0), idx(n);
    for (uint32_t i = 0; i < n; ++i) {
        int g[3];
        grid_pos(x[3 * i], x[3 * i + 1], x[3 * i + 2], cw, g);
        key[i] = grid_hash(g[0], g[1], g[2], G);
        off[key[i] + 1]++;
    }
    for (size_t c = 1; c < off.size(); ++c) off[c] += off[c - 1];
    {
        std::vector<uint32_t> cur(off.begin(), off.end() - 1);
        for (uint32_t i = 0; i < n; ++i) idx[cur[key[i]]++] = i;
    }
    const double inv_h = 1.0 / (double)p->smoothing_length;
    const double m = p->particle_mass, rho0 = p->stable_density;
    double sumC = 0, maxC = -1e300, sumPos = 0, ke = 0;
#pragma omp parallel for schedule(static) reduction(+ : sumC, sumPos, ke) reduction(max : maxC)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        uint32_t i = (uint32_t)ii;
        int g[3];
        grid_pos(x[3 * i], x[3 * i + 1], x[3 * i + 2], cw, g);
        double rho = 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    uint32_t c = grid_hash(g[0] + dx, g[1] + dy, g[2] + dz, G);
                    for (uint32_t t = off[c]; t < off[c + 1]; ++t) {
                        uint32_t j = idx[t];
                        double d0 = (double)x[3 * i] - x[3 * j], d1 = (double)x[3 * i + 1] - x[3 * j + 1],
                               d2 = (double)x[3 * i + 2] - x[3 * j + 2];
                        rho += m * wpoly6<double>(std::sqrt(d0 * d0 + d1 * d1 + d2 * d2), inv_h);
                    }
                }
        double Ci = rho / rho0 - 1.0;
        sumC += Ci;
        sumPos += Ci > 0 ? Ci : 0;
        maxC = std::max(maxC, Ci);
        double vv = (double)v[3 * i] * v[3 * i] + (double)v[3 * i + 1] * v[3 * i + 1] + (double)v[3 * i + 2] * v[3 * i + 2];
        ke += 0.5 * m * vv;
    }
    out[0] = n ? sumC / n : 0;
    out[1] = n ? maxC : 0;
    out[2] = n ? sumPos / n : 0;
    out[3] = ke;
    return 0;
}

}  // extern "C"
