// pbfx_adjoint.hip -- adjoint kernels of the solver step (differentiable form),
// gfx950.
//
// One step (pbf::one_step, pbf/solver/pbf_solve.cpp:182-216, in the full
// update order of :135-142) is piecewise smooth: the sort, the capped lists
// frozen at the predicted positions, the boundary responses that fire, the
// cell each final-pass sum is centred on and the XSPH NaN guard are discrete
// decisions.  The kernels here propagate dL/d(step output) back to dL/d(step
// input) and to dL/d(XSPH c), dL/d(vorticity epsilon) through the smooth map
// those decisions leave, at the values a recorded forward step computed (its
// tape).  Every transposed term is gathered, not scattered: a particle
// collects what the pairs it belongs to contribute, through
//   - reverse lists (the particles whose frozen list holds it), built by
//     re-running the list test from its own side: the relation r^2 < h^2 on
//     the frozen positions is symmetric, and a full list holds exactly the
//     candidates up to its last entry in the 27-cell scan order;
//   - the particles grouped by final cell (the transposed 27-cell sets of the
//     vorticity, confinement and XSPH sums), in slot order within a cell.
// Sums therefore run in a fixed order and results are reproducible run to run.
//
// Stages, last first (per step):
//   XSPH        v_out = Vc + gamma sum_j (Vc_j - Vc_i) W(|p_j - p_i|)
//   confinement Vc = V + kappa (N x omega),  N = eta / |eta|,
//               eta_i = sum_j |omega_j| K(r) e_ij
//   vorticity   omega_i = beta sum_j (V_j - V_i) x (K(r) e_ij)
//   update      V = (P - X) / dt,  x_out = P
//   Jacobi x I  P' = responses(P + beta sum_j -(s_i + s_j + corr) kg_ij),
//               s = C / (m^2/rho0^2 G + eps)
//   predict     P0 = x + dt (v + g dt), gathered by the sort
// with beta = m / rho0, gamma = c beta, kappa = epsilon dt / inv_rho0.
#include "pbfx_device.hpp"

namespace pbfx {
namespace {

constexpr float kSpiky = -14.3239448745f;

__device__ __forceinline__ float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 cross3(F3 a, F3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ F3 sc3(F3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ bool zero3(F3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
__device__ __forceinline__ float4 f4of(F3 a) { return make_float4(a.x, a.y, a.z, 0.f); }
__device__ __forceinline__ F3 ld3aos(const float* a, uint32_t i) { return mk3(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }

// Memo kernels of a listed pair as smooth functions of d = p_i - p_j
// (pbf_kernel_memorization.cu:15-56): W = a t^3 with t = 1 - r^2/h^2,
// kg = kd(r) e with kd = ad (1 - r/h)^2, e = d/r.  grad_d W = gw d and
// d kg/d d = kdp e e^T + (kd/r)(I - e e^T).
struct PairAdj {
    float W, gw, kd, kdp, r;
    F3 e, kg;
    __device__ __forceinline__ F3 H(F3 y) const {
        const float ey = dot3(e, y);
        return sc3(e, kdp * ey) + sc3(y - sc3(e, ey), kd / r);
    }
};
__device__ __forceinline__ bool pair_adj(F3 d, const Phys& P, PairAdj& o) {
    const float r2 = dot3(d, d);
    if (!(r2 > 0.f)) return false;
    o.r = __builtin_sqrtf(r2);
    o.e = sc3(d, 1.f / o.r);
    const float q2 = r2 * P.inv_h * P.inv_h;
    o.W = 0.f;
    o.gw = 0.f;
    if (q2 <= 1.f) {
        const float t = 1.f - q2;
        o.W = P.a * t * t * t;
        o.gw = -6.f * P.a * P.inv_h * P.inv_h * t * t;
    }
    o.kd = 0.f;
    o.kdp = 0.f;
    const float q = o.r * P.inv_h;
    if (q <= 1.f) {
        const float u = 1.f - q;
        o.kd = P.ad * u * u;
        o.kdp = -2.f * P.ad * u * P.inv_h;
    }
    o.kg = sc3(o.e, o.kd);
    return true;
}

// Spiky derivative with argument warg along pd = p_j - p_i
// (pbf/kernel/cuda/detail/pbf_kernel.inl:50-63): f = K(r) e,
// df/dpd = Kp e e^T + (K/r)(I - e e^T).
struct SpikyAdj {
    float K, Kp, r;
    F3 e;
    __device__ __forceinline__ F3 f() const { return sc3(e, K); }
    __device__ __forceinline__ F3 H(F3 y) const {
        const float ey = dot3(e, y);
        return sc3(e, Kp * ey) + sc3(y - sc3(e, ey), K / r);
    }
};
__device__ __forceinline__ bool spiky_adj(F3 pd, float warg, SpikyAdj& o) {
    const float r2 = dot3(pd, pd);
    if (!(r2 > 0.f)) return false;
    o.r = __builtin_sqrtf(r2);
    const float q = o.r * warg;
    if (q > 1.f) return false;
    o.e = sc3(pd, 1.f / o.r);
    const float w2 = warg * warg, w4 = w2 * w2;
    o.K = kSpiky * w4 * (1.f - q) * (1.f - q);
    o.Kp = -2.f * kSpiky * w4 * warg * (1.f - q);
    return true;
}

// poly6 with inv_h along pd (the XSPH weight, pbf_kernel.inl:34-48):
// W = a t^3, t = 1 - r^2/h^2; grad_pd W = gw pd.  False beyond the support.
__device__ __forceinline__ bool poly6_adj(F3 pd, const Phys& P, float& W, float& gw) {
    const float q2 = dot3(pd, pd) * P.inv_h * P.inv_h;
    if (q2 > 1.f) return false;
    const float t = 1.f - q2;
    W = P.a * t * t * t;
    gw = -6.f * P.a * P.inv_h * P.inv_h * t * t;
    return true;
}

// tensile correction of a pair and its derivatives (pbf_position_update.cu:131-188)
__device__ __forceinline__ float corr_of(const Phys& P, float W, float si, float sj) {
    if (P.mode == 0) {
        const float cl = (si < 0.f ? -si : 0.f) + (sj < 0.f ? -sj : 0.f);
        const float W2 = W * W;
        return cl * 10.f * W2 * W2 * P.inv_k;
    }
    return P.sk * powi(W * P.inv_wdq, P.sn);
}
// dcorr/dW
__device__ __forceinline__ float corr_dW(const Phys& P, float W, float si, float sj) {
    if (P.mode == 0) {
        const float cl = (si < 0.f ? -si : 0.f) + (sj < 0.f ? -sj : 0.f);
        return cl * 40.f * W * W * W * P.inv_k;
    }
    return P.sn > 0 ? P.sk * (float)P.sn * powi(W * P.inv_wdq, P.sn - 1) * P.inv_wdq : 0.f;
}
// dcorr/ds of one end whose s is `s` (REFERENCE form only)
__device__ __forceinline__ float corr_ds(const Phys& P, float W, float s) {
    if (P.mode != 0 || !(s < 0.f)) return 0.f;
    const float W2 = W * W;
    return -10.f * W2 * W2 * P.inv_k;
}

// the 27 search cells around the cell of p, in z, y, x order
template <class F>
__device__ __forceinline__ void for_cells(const Grid& G, F3 p, F&& f) {
    int g[3];
    grid_pos(p, G.cw, g);
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) f(grid_hash(g[0] + dx, g[1] + dy, g[2] + dz, G), dx, dy, dz);
}

// S(i): the slots of the 27 cells around the final cell of p (the forward
// sum's candidates, pbf_contribution.inl:7-57)
template <class F>
__device__ __forceinline__ void for_S(const StepTape& t, const Grid& G, F3 p, F&& f) {
    for_cells(G, p, [&](uint32_t c, int, int, int) {
        const uint32_t e = t.offset[c + 1];
        for (uint32_t j = t.offset[c]; j < e; ++j) f(j);
    });
}
// T(i) = {l : i in S(l)}: the particles whose final cell is one of the 27
// cells around the search cell of i (p0 = i's predicted position)
template <class F>
__device__ __forceinline__ void for_T(const AdjBufs& a, const Grid& G, F3 p0, F&& f) {
    for_cells(G, p0, [&](uint32_t c, int, int, int) {
        const uint32_t e = a.coffset[c + 1];
        for (uint32_t q = a.coffset[c]; q < e; ++q) f(a.cperm[q]);
    });
}

__device__ __forceinline__ uint32_t gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }

// sum of v over the workgroup (256 threads), valid in thread 0
__device__ __forceinline__ double block_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    return sh[0];
}

// ------------------------------------------------------------ grouping by final cell
__global__ __launch_bounds__(256) void k_adj_ckey(StepTape t, AdjBufs a, Grid G, uint32_t lgT) {
    const uint32_t s = gtid();
    const bool active = s < t.n_live;
    uint32_t k = 0;
    if (active) {
        int g[3];
        grid_pos(f3of(t.PI[s]), G.cw, g);
        k = grid_hash(g[0], g[1], g[2], G);
    }
    const uint32_t r = wave_rank(a.ccount, k, active);
    if (active) {
        mark_tile(a.ctflag, k >> lgT);
        a.ckey[s] = k;
        a.crank[s] = r;
    }
}

__global__ __launch_bounds__(256) void k_adj_cscatter(StepTape t, AdjBufs a) {
    const uint32_t s = gtid();
    if (s >= t.n_live) return;
    a.ctmp[a.coffset[a.ckey[s]] + a.crank[s]] = s;
}

// slot order inside each cell: position = number of smaller slots in the cell
__global__ __launch_bounds__(256) void k_adj_corder(StepTape t, AdjBufs a) {
    const uint32_t s = gtid();
    if (s >= t.n_live) return;
    const uint32_t k = a.ckey[s];
    const uint32_t o0 = a.coffset[k], o1 = a.coffset[k + 1];
    uint32_t r = 0;
    for (uint32_t q = o0; q < o1; ++q) r += a.ctmp[q] < s ? 1u : 0u;
    a.cperm[o0 + r] = s;
}

// ------------------------------------------------------------ reverse lists
// For a full list: the 27-cell index and slot of its last entry.
__global__ __launch_bounds__(256) void k_adj_last(StepTape t, AdjBufs a, Grid G, Phys P, uint32_t cap) {
    const uint32_t l = gtid();
    if (l >= t.n_live) return;
    const uint32_t n = t.ncnt[l];
    uint2 out = make_uint2(255u, 0u);
    if (n == (uint32_t)P.cap && n > 0) {
        const uint32_t jl = t.nlist[(size_t)(n - 1u) * cap + l];
        int gl[3], gj[3];
        grid_pos(f3of(t.Q[l]), G.cw, gl);
        grid_pos(f3of(t.Q[jl]), G.cw, gj);
        const uint32_t m[3] = {G.gx - 1u, G.gy - 1u, G.gz - 1u};
        int d[3];
        for (int c = 0; c < 3; ++c) {
            const uint32_t u = (uint32_t)(gj[c] - gl[c]) & m[c];
            d[c] = u == 0u ? 0 : (u == 1u ? 1 : -1);
        }
        out = make_uint2((uint32_t)((d[2] + 1) * 9 + (d[1] + 1) * 3 + (d[0] + 1)), jl);
    }
    a.last[l] = out;
}

// The slots l != j whose frozen list holds j, in j's 27-cell scan order
// (pbf_neighbor_search_util.cu:386-408,428-452: l lists the candidates with
// r^2 < h^2 in its own scan order until the list is full).  fill = false:
// count into rcount; fill = true: write rlist.
template <bool fill>
__global__ __launch_bounds__(256) void k_adj_rev(StepTape t, AdjBufs a, Grid G, Phys P) {
    const uint32_t j = gtid();
    if (j >= t.n_live) return;
    const F3 self = f3of(t.Q[j]);
    const uint32_t full = (uint32_t)P.cap;
    uint32_t cnt = 0;
    uint32_t* out = fill ? a.rlist + a.roff[j] : nullptr;
    for_cells(G, self, [&](uint32_t c, int dx, int dy, int dz) {
        const uint32_t e = t.offset[c + 1];
        const uint32_t pos = (uint32_t)((1 - dz) * 9 + (1 - dy) * 3 + (1 - dx));  // j seen from l
        for (uint32_t l = t.offset[c]; l < e; ++l) {
            if (l == j) continue;
            const F3 d = f3of(t.Q[l]) - self;  // the search's d = p_l - p_j
            const float r2 = __builtin_fmaf(d.z, d.z, __builtin_fmaf(d.y, d.y, d.x * d.x));
            if (!(r2 < P.h2)) continue;
            const uint32_t n = t.ncnt[l];
            bool listed = n > 0 && n < full;
            if (n == full) {
                const uint2 L = a.last[l];
                listed = pos < L.x || (pos == L.x && j <= L.y);
            }
            if (listed) {
                if (fill) out[cnt] = l;
                ++cnt;
            }
        }
    });
    if (!fill) {
        a.rcount[j] = cnt;
        if (cnt) mark_tile(a.rtflag, j >> a.lgRT);
    }
}

// ------------------------------------------------------------ final stage
// XSPH, own pairs: dL/dVc_i, the position terms of i's own sum, dL/dc.
__global__ __launch_bounds__(256) void k_adj_xsph_own(StepTape t, AdjBufs a, Grid G, Phys P) {
    __shared__ double sh[256];
    const uint32_t i = gtid();
    double pc = 0.0;
    if (i < t.n_live) {
        const float beta = P.m * P.inv_rho0;
        const float gamma = P.xsph_c * P.m * P.inv_rho0;
        const F3 self = f3of(t.PI[i]), vi = f3of(t.R[2 * i + 1]);
        const F3 gi = ld3aos(a.av, i);
        float sumW = 0.f;
        F3 acc = mk3(0.f, 0.f, 0.f), apos = mk3(0.f, 0.f, 0.f);
        for_S(t, G, self, [&](uint32_t j) {
            if (j == i) return;
            const F3 pd = f3of(t.PI[j]) - self;
            if (dot3(pd, pd) > P.cut2_w) return;
            float W, gw;
            if (!poly6_adj(pd, P, W, gw)) return;
            const F3 e = f3of(t.R[2 * j + 1]) - vi;
            acc = acc + sc3(e, W);
            sumW += W;
            apos = apos - sc3(pd, gw * gamma * dot3(gi, e));
        });
        const bool nan = isnan(acc.x) || isnan(acc.y) || isnan(acc.z);
        a.xnan[i] = nan ? 1u : 0u;
        F3 avc = gi;
        if (nan) {
            apos = mk3(0.f, 0.f, 0.f);
        } else {
            avc = sc3(gi, 1.f - gamma * sumW);
            pc = (double)beta * (double)dot3(gi, acc);
        }
        a.aVc[i] = f4of(avc);
        a.aP[i] = f4of(ld3aos(a.ax, i) + apos);
    }
    const double tot = block_sum(pc, sh);
    if (threadIdx.x == 0) a.part[blockIdx.x] = tot;
}

// XSPH, transposed pairs (l, i), i in S(l).
__global__ __launch_bounds__(256) void k_adj_xsph_rev(StepTape t, AdjBufs a, Grid G, Phys P) {
    const uint32_t i = gtid();
    if (i >= t.n_live) return;
    const float gamma = P.xsph_c * P.m * P.inv_rho0;
    const F3 pi = f3of(t.PI[i]), vi = f3of(t.R[2 * i + 1]);
    F3 avc = mk3(0.f, 0.f, 0.f), apos = mk3(0.f, 0.f, 0.f);
    for_T(a, G, f3of(t.Q[i]), [&](uint32_t l) {
        if (l == i || a.xnan[l]) return;
        const F3 pd = pi - f3of(t.PI[l]);
        if (dot3(pd, pd) > P.cut2_w) return;
        float W, gw;
        if (!poly6_adj(pd, P, W, gw)) return;
        const F3 gl = ld3aos(a.av, l);
        avc = avc + sc3(gl, gamma * W);
        apos = apos + sc3(pd, gw * gamma * dot3(gl, vi - f3of(t.R[2 * l + 1])));
    });
    a.aVc[i] = f4of(f3of(a.aVc[i]) + avc);
    a.aP[i] = f4of(f3of(a.aP[i]) + apos);
}

// Without XSPH: v_out = Vc.
__global__ __launch_bounds__(256) void k_adj_final_init(StepTape t, AdjBufs a) {
    const uint32_t i = gtid();
    if (i < t.n_live) {
        a.aVc[i] = f4of(ld3aos(a.av, i));
        a.aP[i] = f4of(ld3aos(a.ax, i));
    }
    if (i < kAdjMaxBlocks) {
        a.part[i] = 0.0;
        a.part[kAdjMaxBlocks + i] = 0.0;
    }
}

// Confinement, direct terms: dL/domega_i through N x omega_i, dL/deta_i,
// dL/d epsilon; dL/dV starts as dL/dVc.
__global__ __launch_bounds__(256) void k_adj_conf_own(StepTape t, AdjBufs a, Grid G, Phys P) {
    __shared__ double sh[256];
    const uint32_t i = gtid();
    double pv = 0.0;
    if (i < t.n_live) {
        const F3 self = f3of(t.PI[i]);
        F3 eta = mk3(0.f, 0.f, 0.f);
        for_S(t, G, self, [&](uint32_t j) {
            SpikyAdj sd;
            if (j == i || !spiky_adj(f3of(t.PI[j]) - self, P.warg, sd)) return;
            eta = eta + sc3(sd.f(), t.OM4[j].w);
        });
        const F3 avc = f3of(a.aVc[i]);
        a.aV[i] = f4of(avc);
        const float nr = __builtin_sqrtf(dot3(eta, eta));
        F3 aom = mk3(0.f, 0.f, 0.f), eb = mk3(0.f, 0.f, 0.f);
        if (nr > 0.f) {
            const F3 Nh = sc3(eta, 1.f / nr), om = f3of(t.OM4[i]);
            const float kappa = P.vc * P.dtc / P.inv_rho0;
            pv = (double)dot3(avc, cross3(Nh, om)) * (double)(P.dtc / P.inv_rho0);
            const F3 y = sc3(avc, kappa);
            aom = cross3(y, Nh);
            const F3 nb = cross3(om, y);
            eb = sc3(nb - sc3(Nh, dot3(Nh, nb)), 1.f / nr);
        }
        a.aOm[i] = f4of(aom);
        a.etab[i] = f4of(eb);
    }
    const double tot = block_sum(pv, sh);
    if (threadIdx.x == 0) a.part[kAdjMaxBlocks + blockIdx.x] = tot;
}

// Confinement through eta: the position terms and dL/d|omega_i| (own pairs
// and transposed pairs).
__global__ __launch_bounds__(256) void k_adj_conf_eta(StepTape t, AdjBufs a, Grid G, Phys P) {
    const uint32_t i = gtid();
    if (i >= t.n_live) return;
    const F3 self = f3of(t.PI[i]);
    const F3 eb = f3of(a.etab[i]);
    const float wi = t.OM4[i].w;
    F3 apos = mk3(0.f, 0.f, 0.f);
    if (!zero3(eb))
        for_S(t, G, self, [&](uint32_t j) {
            SpikyAdj sd;
            if (j == i || !spiky_adj(f3of(t.PI[j]) - self, P.warg, sd)) return;
            apos = apos - sc3(sd.H(eb), t.OM4[j].w);
        });
    float mb = 0.f;
    for_T(a, G, f3of(t.Q[i]), [&](uint32_t l) {
        if (l == i) return;
        const F3 el = f3of(a.etab[l]);
        if (zero3(el)) return;
        SpikyAdj sd;
        if (!spiky_adj(self - f3of(t.PI[l]), P.warg, sd)) return;
        mb += dot3(el, sd.f());
        apos = apos + sc3(sd.H(el), wi);
    });
    F3 aom = f3of(a.aOm[i]);
    if (wi > 0.f) aom = aom + sc3(f3of(t.OM4[i]), mb / wi);
    a.aOm[i] = f4of(aom);
    a.aP[i] = f4of(f3of(a.aP[i]) + apos);
}

// Vorticity: dL/dV and the position terms (own pairs and transposed pairs),
// then the update V = (P - X)/dt.
__global__ __launch_bounds__(256) void k_adj_vort(StepTape t, AdjBufs a, Grid G, Phys P, uint32_t cap) {
    const uint32_t i = gtid();
    if (i >= t.n_live) return;
    const float beta = P.m * P.inv_rho0;
    const F3 self = f3of(t.PI[i]);
    const F3 vi = mulsv(P.inv_dt, self - ld3(t.X, cap, i));
    const F3 ab = f3of(a.aOm[i]);
    F3 av = f3of(a.aV[i]), apos = mk3(0.f, 0.f, 0.f);
    if (!zero3(ab))
        for_S(t, G, self, [&](uint32_t j) {
            SpikyAdj sd;
            const F3 pj = f3of(t.PI[j]);
            if (j == i || !spiky_adj(pj - self, P.warg, sd)) return;
            const F3 e = mulsv(P.inv_dt, pj - ld3(t.X, cap, j)) - vi;
            av = av - sc3(cross3(sd.f(), ab), beta);
            apos = apos - sd.H(sc3(cross3(ab, e), beta));
        });
    for_T(a, G, f3of(t.Q[i]), [&](uint32_t l) {
        if (l == i) return;
        const F3 al = f3of(a.aOm[l]);
        if (zero3(al)) return;
        SpikyAdj sd;
        const F3 pl = f3of(t.PI[l]);
        if (!spiky_adj(self - pl, P.warg, sd)) return;
        const F3 e = vi - mulsv(P.inv_dt, pl - ld3(t.X, cap, l));
        av = av + sc3(cross3(sd.f(), al), beta);
        apos = apos + sd.H(sc3(cross3(al, e), beta));
    });
    a.aP[i] = f4of(f3of(a.aP[i]) + apos + sc3(av, P.inv_dt));
    a.aX[i] = f4of(sc3(av, -P.inv_dt));
}

// Without vorticity: the update only.
__global__ __launch_bounds__(256) void k_adj_update(StepTape t, AdjBufs a, Phys P) {
    const uint32_t i = gtid();
    if (i >= t.n_live) return;
    const F3 av = f3of(a.aVc[i]);
    a.aP[i] = f4of(f3of(a.aP[i]) + sc3(av, P.inv_dt));
    a.aX[i] = f4of(sc3(av, -P.inv_dt));
}

// ------------------------------------------------------------ Jacobi iteration
// Boundary responses of k_deltap's collide() with their Jacobian: every
// response that fires maps the point pp linearly around itself --
// planes: pp - ((pp - P).n) n, Jacobian I - n n^T (both cases of
// pbf_plane_boundary.cu:6-45 reduce to it); spheres (pbf_sphere_boundary.cu:8-64):
// pp + nudge cd - R cd/|cd| with cd = c - pp, Jacobian
// (1 - nudge) I + (R/|cd|)(I - u u^T), u = cd/|cd|.  The decisions follow
// collide() exactly (same arithmetic), J accumulates the product.
__device__ void collide_jac(const Phys& P, const Boundary& B, F3& dp, F3 pk, float J[9]) {
    auto apply = [&](const float M[9]) {
        float T[9];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) T[3 * r + c] = M[3 * r] * J[c] + M[3 * r + 1] * J[3 + c] + M[3 * r + 2] * J[6 + c];
        for (int q = 0; q < 9; ++q) J[q] = T[q];
    };
    const F3 pred_x = pk, old_x = pk;
    int sweeps = 0;
    while (sweeps < P.max_sweeps) {
        ++sweeps;
        F3 pp = pred_x + dp;
        bool hit = false;
        for (uint32_t k = 0; k < B.n_inner + B.n_outer; ++k) {
            const bool is_inner = k < B.n_inner;
            const float4 s = is_inner ? B.inner[k] : B.outer[k - B.n_inner];
            const bool inner = is_inner || P.outer_as_inner != 0;
            const F3 cd = mk3(s.x, s.y, s.z) - pp;
            const float d2 = fdot(cd, cd);
            if (inner ? (d2 > s.w * s.w) : (d2 < s.w * s.w)) {
                const float L = __builtin_sqrtf(dot3(cd, cd));
                const F3 u = sc3(cd, 1.f / L);
                const float nudge = inner ? 1.0001f : 0.9999f;
                const float c0 = 1.f - nudge, c1 = s.w / L;
                const float M[9] = {c0 + c1 * (1.f - u.x * u.x), -c1 * u.x * u.y, -c1 * u.x * u.z,
                                    -c1 * u.y * u.x, c0 + c1 * (1.f - u.y * u.y), -c1 * u.y * u.z,
                                    -c1 * u.z * u.x, -c1 * u.z * u.y, c0 + c1 * (1.f - u.z * u.z)};
                apply(M);
            }
            response_sphere(hit, dp, pp, mk3(s.x, s.y, s.z), s.w, inner);
            pp = pred_x + dp;
        }
        for (uint32_t k = 0; k < B.n_planes; ++k) {
            const F3 Pp = mk3(B.pts[3 * k], B.pts[3 * k + 1], B.pts[3 * k + 2]);
            const F3 n = mk3(B.nrm[3 * k], B.nrm[3 * k + 1], B.nrm[3 * k + 2]);
            bool h = false;
            response_plane(h, dp, old_x, pp, Pp, n);
            if (h) {
                hit = true;
                const float M[9] = {1.f - n.x * n.x, -n.x * n.y, -n.x * n.z, -n.y * n.x, 1.f - n.y * n.y,
                                    -n.y * n.z, -n.z * n.x, -n.z * n.y, 1.f - n.z * n.z};
                apply(M);
            }
            pp = pred_x + dp;
        }
        if (!hit) break;
    }
}

// dL/d(pre-collision position) = J^T dL/dP'.  dp is recomputed exactly as
// k_deltap computes it, so the responses replay the forward's decisions.
__global__ __launch_bounds__(256) void k_adj_coll(StepTape t, AdjBufs a, const float4* __restrict__ Q, Phys P,
                                                  Boundary B, uint32_t cap) {
    const uint32_t i = gtid();
    if (i >= t.n_live) return;
    F3 ag = f3of(a.aP[i]);
    if (B.n_planes + B.n_inner + B.n_outer > 0) {
        const float4 se = Q[i];
        const F3 self = f3of(se);
        const float si = se.w;
        const float self_cl = (si < 0.f) ? -si : 0.f;
        const uint32_t n = t.ncnt[i];
        F3 sum = mk3(0.f, 0.f, 0.f);
        for (uint32_t k = 0; k < n; ++k) {
            const float4 pe = Q[t.nlist[(size_t)k * cap + i]];
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
                corr = P.sk * powi(W * P.inv_wdq, P.sn);
            }
            const float aa = -(si + sj + corr);
            acc(sum, mulsv(aa, kg));
        }
        F3 dp = mk3(sum.x * P.m * P.inv_rho0, sum.y * P.m * P.inv_rho0, sum.z * P.m * P.inv_rho0);
        float J[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
        collide_jac(P, B, dp, self, J);
        ag = mk3(J[0] * ag.x + J[3] * ag.y + J[6] * ag.z, J[1] * ag.x + J[4] * ag.y + J[7] * ag.z,
                 J[2] * ag.x + J[5] * ag.y + J[8] * ag.z);
    }
    a.abar[i] = f4of(ag);
}

// delta p = beta sum_j A_ij kg_ij, A_ij = -(s_i + s_j + corr(W_ij)):
// dL/ds_i and the position terms through kg and corr, own pairs (i, j) and
// transposed pairs (l, i); then the coefficients of the lambda pass.
__global__ __launch_bounds__(256) void k_adj_lam(StepTape t, AdjBufs a, const float4* __restrict__ Q, Phys P,
                                                 uint32_t cap) {
    const uint32_t i = gtid();
    if (i >= t.n_live) return;
    const float beta = P.m * P.inv_rho0;
    const float4 se = Q[i];
    const F3 self = f3of(se);
    const float si = se.w;
    const F3 ai = f3of(a.abar[i]);
    float sb = 0.f, ks = 0.f, kg2s = 0.f;
    F3 kgs = mk3(0.f, 0.f, 0.f), apos = mk3(0.f, 0.f, 0.f);
    const uint32_t n = t.ncnt[i];
    for (uint32_t k = 0; k < n; ++k) {
        const float4 pe = Q[t.nlist[(size_t)k * cap + i]];
        const F3 d = self - f3of(pe);
        PairAdj pa;
        if (!pair_adj(d, P, pa)) {
            ks += P.a;  // r = 0: W = a, kg = 0
            continue;
        }
        ks += pa.W;
        kgs = kgs + pa.kg;
        kg2s += dot3(pa.kg, pa.kg);
        const float sj = pe.w;
        const float A = -(si + sj + corr_of(P, pa.W, si, sj));
        const float Ab = beta * dot3(pa.kg, ai);
        sb -= Ab;
        sb -= Ab * corr_ds(P, pa.W, si);
        const float Wb = -Ab * corr_dW(P, pa.W, si, sj);
        apos = apos + sc3(pa.H(ai), beta * A) + sc3(d, Wb * pa.gw);
    }
    const uint32_t r1 = a.roff[i + 1];
    for (uint32_t q = a.roff[i]; q < r1; ++q) {
        const uint32_t l = a.rlist[q];
        const float4 pe = Q[l];
        const F3 d = f3of(pe) - self;  // the pair (l, i): d = p_l - p_i
        PairAdj pa;
        if (!pair_adj(d, P, pa)) continue;
        const float sl = pe.w;
        const F3 al = f3of(a.abar[l]);
        const float A = -(sl + si + corr_of(P, pa.W, sl, si));
        const float Ab = beta * dot3(pa.kg, al);
        sb -= Ab;
        sb -= Ab * corr_ds(P, pa.W, si);
        const float Wb = -Ab * corr_dW(P, pa.W, sl, si);
        apos = apos - (sc3(pa.H(al), beta * A) + sc3(d, Wb * pa.gw));
    }
    // s = C / D, C = beta sum W - 1, D = m2r2 (|sum kg|^2 + sum |kg|^2) + eps
    const float D = P.m2r2 * (dot3(kgs, kgs) + kg2s) + P.relax;
    const float ksb = beta * sb / D;
    const float g2 = -2.f * P.m2r2 * sb * si / D;
    (void)ks;
    a.coef[i] = make_float4(ksb, g2 * kgs.x, g2 * kgs.y, g2 * kgs.z);
    a.g2[i] = g2;
    a.aQ[i] = f4of(ai + apos);
}

// lambda pass: the position terms through W and kg of s_i (own pairs) and
// s_l (transposed pairs); dL/dP of the iteration's input.
__global__ __launch_bounds__(256) void k_adj_pos(StepTape t, AdjBufs a, const float4* __restrict__ Q, Phys P,
                                                 uint32_t cap) {
    const uint32_t i = gtid();
    if (i >= t.n_live) return;
    const F3 self = f3of(Q[i]);
    const float4 ci = a.coef[i];
    const float g2i = a.g2[i];
    F3 apos = mk3(0.f, 0.f, 0.f);
    if (ci.x != 0.f || g2i != 0.f) {
        const uint32_t n = t.ncnt[i];
        for (uint32_t k = 0; k < n; ++k) {
            const F3 d = self - f3of(Q[t.nlist[(size_t)k * cap + i]]);
            PairAdj pa;
            if (!pair_adj(d, P, pa)) continue;
            apos = apos + sc3(d, ci.x * pa.gw) + pa.H(mk3(ci.y, ci.z, ci.w) + sc3(pa.kg, g2i));
        }
    }
    const uint32_t r1 = a.roff[i + 1];
    for (uint32_t q = a.roff[i]; q < r1; ++q) {
        const uint32_t l = a.rlist[q];
        const float4 cl = a.coef[l];
        const float g2l = a.g2[l];
        if (cl.x == 0.f && g2l == 0.f) continue;
        const F3 d = f3of(Q[l]) - self;
        PairAdj pa;
        if (!pair_adj(d, P, pa)) continue;
        apos = apos - (sc3(d, cl.x * pa.gw) + pa.H(mk3(cl.y, cl.z, cl.w) + sc3(pa.kg, g2l)));
    }
    a.aP[i] = f4of(f3of(a.aQ[i]) + apos);
}

// ------------------------------------------------------------ predict + sort
// P0_s = x*_{sidx[s]}, x* = x + dt (v + g dt); X_s = x_{sidx[s]}.
__global__ __launch_bounds__(256) void k_adj_in(StepTape t, AdjBufs a, Phys P) {
    const uint32_t s = gtid();
    if (s >= t.n_in) return;
    const uint32_t i = t.sidx[s];
    F3 gx = mk3(0.f, 0.f, 0.f), gv = mk3(0.f, 0.f, 0.f);
    if (s < t.n_live) {
        const F3 ap = f3of(a.aP[s]);
        gx = ap + f3of(a.aX[s]);
        gv = sc3(ap, P.dt);
    }
    a.axin[3 * i] = gx.x;
    a.axin[3 * i + 1] = gx.y;
    a.axin[3 * i + 2] = gx.z;
    a.avin[3 * i] = gv.x;
    a.avin[3 * i + 1] = gv.y;
    a.avin[3 * i + 2] = gv.z;
}

// psum[0] = sum of part[0..nb), psum[1] = sum of part[M..M+nb), in order
__global__ __launch_bounds__(256) void k_adj_reduce(AdjBufs a, uint32_t nb) {
    __shared__ double sh[256];
    for (int w = 0; w < 2; ++w) {
        double v = 0.0;
        for (uint32_t q = threadIdx.x; q < nb; q += blockDim.x) v += a.part[w * kAdjMaxBlocks + q];
        const double tot = block_sum(v, sh);
        if (threadIdx.x == 0) a.psum[w] = tot;
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_adj_absmax(AdjBufs a, uint32_t n3) {
    const uint32_t q = gtid();
    float m = 0.f;
    if (q < n3) m = fmaxf(fabsf(a.axin[q]), fabsf(a.avin[q]));
    if (q < n3 && (isnan(a.axin[q]) || isnan(a.avin[q]))) m = __builtin_inff();
    for (int d = 32; d > 0; d >>= 1) m = fmaxf(m, __shfl_xor(m, d, 64));
    if ((threadIdx.x & 63u) == 0 && m > 0.f) atomicMax(a.amax, __float_as_uint(m));
}

__global__ __launch_bounds__(256) void k_adj_scale(float* x, float* y, uint32_t n, float f) {
    const uint32_t q = gtid();
    if (q >= n) return;
    if (x) x[q] *= f;
    if (y) y[q] *= f;
}

__global__ __launch_bounds__(256) void k_adj_ident(const uint32_t* __restrict__ id_in,
                                                   const uint32_t* __restrict__ sidx, uint32_t* __restrict__ id_out,
                                                   uint32_t n) {
    const uint32_t q = gtid();
    if (q >= n) return;
    id_out[q] = id_in ? id_in[sidx[q]] : q;
}

inline uint32_t nblk(uint32_t n) { return (n + 255u) / 256u; }

}  // namespace

hipError_t launch_adj_step(const Bufs& b, const Scene& s0, const StepTape& t, const AdjBufs& a, int iters,
                           hipStream_t s) {
    const Grid& G = s0.G;
    const Phys& P = s0.P;
    const uint32_t nl = t.n_live;
    const dim3 blk(256);
    const dim3 gl(nblk(nl > 0 ? nl : 1));
    if (gl.x > kAdjMaxBlocks) return hipErrorInvalidValue;
    const uint32_t cap = b.cap;
    if (nl > 0) {
        // particles grouped by final cell
        hipLaunchKernelGGL(k_adj_ckey, gl, blk, 0, s, t, a, G, b.lgT);
        hipError_t e = launch_count_scan(a.ccount, a.coffset, a.ctsum, a.ctflag, a.ctprefix, b.nT, b.lgT, b.CT, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_adj_cscatter, gl, blk, 0, s, t, a);
        hipLaunchKernelGGL(k_adj_corder, gl, blk, 0, s, t, a);
        // reverse lists
        hipLaunchKernelGGL(k_adj_last, gl, blk, 0, s, t, a, G, P, cap);
        hipLaunchKernelGGL(k_adj_rev<false>, gl, blk, 0, s, t, a, G, P);
        e = launch_count_scan(a.rcount, a.roff, a.rtsum, a.rtflag, a.rtprefix, a.nRT, a.lgRT, a.RT, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_adj_rev<true>, gl, blk, 0, s, t, a, G, P);
        // final stage
        hipLaunchKernelGGL(k_adj_final_init, dim3(nblk(kAdjMaxBlocks > nl ? kAdjMaxBlocks : nl)), blk, 0, s, t, a);
        if (P.xsph_on) {
            hipLaunchKernelGGL(k_adj_xsph_own, gl, blk, 0, s, t, a, G, P);
            hipLaunchKernelGGL(k_adj_xsph_rev, gl, blk, 0, s, t, a, G, P);
        }
        if (P.vort_on) {
            hipLaunchKernelGGL(k_adj_conf_own, gl, blk, 0, s, t, a, G, P);
            hipLaunchKernelGGL(k_adj_conf_eta, gl, blk, 0, s, t, a, G, P);
            hipLaunchKernelGGL(k_adj_vort, gl, blk, 0, s, t, a, G, P, cap);
        } else {
            hipLaunchKernelGGL(k_adj_update, gl, blk, 0, s, t, a, P);
        }
        // Jacobi iterations, last first
        for (int it = iters - 1; it >= 0; --it) {
            const float4* Q = t.Q + (size_t)it * cap;
            hipLaunchKernelGGL(k_adj_coll, gl, blk, 0, s, t, a, Q, P, s0.B, cap);
            hipLaunchKernelGGL(k_adj_lam, gl, blk, 0, s, t, a, Q, P, cap);
            hipLaunchKernelGGL(k_adj_pos, gl, blk, 0, s, t, a, Q, P, cap);
        }
    } else {
        hipLaunchKernelGGL(k_adj_final_init, dim3(nblk(kAdjMaxBlocks)), blk, 0, s, t, a);
    }
    if (t.n_in > 0) hipLaunchKernelGGL(k_adj_in, dim3(nblk(t.n_in)), blk, 0, s, t, a, P);
    hipLaunchKernelGGL(k_adj_reduce, dim3(1), blk, 0, s, a, gl.x);
    return hipGetLastError();
}

hipError_t launch_adj_absmax(const AdjBufs& a, uint32_t n, hipStream_t s) {
    hipError_t e = hipMemsetAsync(a.amax, 0, sizeof(uint32_t), s);
    if (e != hipSuccess) return e;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adj_absmax, dim3(nblk(3 * n)), dim3(256), 0, s, a, 3 * n);
    return hipGetLastError();
}

hipError_t launch_adj_scale(float* x, float* y, uint32_t n, float f, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adj_scale, dim3(nblk(n)), dim3(256), 0, s, x, y, n, f);
    return hipGetLastError();
}

hipError_t launch_adj_ident(const uint32_t* id_in, const uint32_t* sidx, uint32_t* id_out, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adj_ident, dim3(nblk(n)), dim3(256), 0, s, id_in, sidx, id_out, n);
    return hipGetLastError();
}

}  // namespace pbfx
