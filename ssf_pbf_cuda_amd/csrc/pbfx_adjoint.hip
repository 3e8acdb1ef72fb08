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
// All adjoint arithmetic is FP64 (the tape's FP32 values are widened on
// load): in exploding flows the map is conditioned up to ~1e7, which turns
// FP32 rounding into O(1) error over tens of steps.
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

// ------------------------------------------------------------ FP64 vectors
struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 md3(double x, double y, double z) {
    D3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return md3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return md3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ D3 sc3(D3 a, double s) { return md3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ double dot3(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross3(D3 a, D3 b) {
    return md3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ bool zero3(D3 a) { return a.x == 0.0 && a.y == 0.0 && a.z == 0.0; }
__device__ __forceinline__ D3 dz3() { return md3(0.0, 0.0, 0.0); }
__device__ __forceinline__ D3 wid(float4 f) { return md3(f.x, f.y, f.z); }
__device__ __forceinline__ D3 wid3(F3 f) { return md3(f.x, f.y, f.z); }
__device__ __forceinline__ D3 ld(const double4* a, uint32_t i) {
    const double4 q = a[i];
    return md3(q.x, q.y, q.z);
}
__device__ __forceinline__ void st(double4* a, uint32_t i, D3 v) { a[i] = make_double4(v.x, v.y, v.z, 0.0); }
__device__ __forceinline__ D3 ld3d(const double* a, uint32_t i) { return md3(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }
// sorted input position i (SoA, stride cap)
__device__ __forceinline__ D3 ldX(const float* X, uint32_t cap, uint32_t i) {
    return md3(X[i], X[(size_t)cap + i], X[2 * (size_t)cap + i]);
}

__device__ __forceinline__ double powid(double t, int n) {
    double r = 1.0;
    for (int k = 0; k < n; ++k) r *= t;
    return r;
}

// Constants of the step in FP64: the FP32 step's own (float) constants,
// widened, so that the derivative is that of the FP32 map.
struct AdjConst {
    double beta, gamma, kappa, dpdvc, m2r2, relax, inv_dt, dt, warg, inv_k, inv_wdq, sk;
    double a, ad, ih, alpha_x, K0;  // memo W / kg, 1/h, XSPH poly6 and spiky-derivative factors
    int mode, sn;
};
__device__ __forceinline__ AdjConst adj_const(const Phys& P) {
    AdjConst k;
    k.beta = (double)P.m * (double)P.inv_rho0;
    k.gamma = (double)P.xsph_c * k.beta;
    k.dpdvc = (double)P.dtc / (double)P.inv_rho0;
    k.kappa = (double)P.vc * k.dpdvc;
    k.m2r2 = P.m2r2;
    k.relax = P.relax;
    k.inv_dt = P.inv_dt;
    k.dt = P.dt;
    k.warg = P.warg;
    k.inv_k = P.inv_k;
    k.inv_wdq = P.inv_wdq;
    k.sk = P.sk;
    k.mode = P.mode;
    k.sn = P.sn;
    k.a = P.a;
    k.ad = P.ad;
    k.ih = P.inv_h;
    // wpoly6 / wspiky_deriv form these in float (pbf_kernel.inl:34-63)
    k.alpha_x = 1.56668147065f * P.inv_h * P.inv_h * P.inv_h;
    const float w2 = P.warg * P.warg;
    k.K0 = -14.3239448745f * (w2 * w2);
    return k;
}

// Memo kernels of a listed pair as smooth functions of d = p_i - p_j
// (pbf_kernel_memorization.cu:15-56): W = a t^3 with t = 1 - r^2/h^2,
// kg = kd(r) e with kd = ad (1 - r/h)^2, e = d/r.  grad_d W = gw d and
// d kg/d d = kdp e e^T + (kd/r)(I - e e^T).
struct PairAdj {
    double W, gw, kd, kdp, r;
    D3 e, kg;
    __device__ __forceinline__ D3 H(D3 y) const {
        const double ey = dot3(e, y);
        return sc3(e, kdp * ey) + sc3(y - sc3(e, ey), kd / r);
    }
};
__device__ __forceinline__ bool pair_adj(D3 d, const AdjConst& k, PairAdj& o) {
    const double r2 = dot3(d, d);
    if (!(r2 > 0.0)) return false;
    o.r = sqrt(r2);
    o.e = sc3(d, 1.0 / o.r);
    const double q2 = r2 * k.ih * k.ih;
    o.W = 0.0;
    o.gw = 0.0;
    if (q2 <= 1.0) {
        const double t = 1.0 - q2;
        o.W = k.a * t * t * t;
        o.gw = -6.0 * k.a * k.ih * k.ih * t * t;
    }
    o.kd = 0.0;
    o.kdp = 0.0;
    const double q = o.r * k.ih;
    if (q <= 1.0) {
        const double u = 1.0 - q;
        o.kd = k.ad * u * u;
        o.kdp = -2.0 * k.ad * u * k.ih;
    }
    o.kg = sc3(o.e, o.kd);
    return true;
}

// Spiky derivative with argument warg along pd = p_j - p_i
// (pbf/kernel/cuda/detail/pbf_kernel.inl:50-63): f = K(r) e with
// K = K0 (1 - r warg)^2; df/dpd = Kp e e^T + (K/r)(I - e e^T).
struct SpikyAdj {
    double K, Kp, r;
    D3 e;
    __device__ __forceinline__ D3 f() const { return sc3(e, K); }
    __device__ __forceinline__ D3 H(D3 y) const {
        const double ey = dot3(e, y);
        return sc3(e, Kp * ey) + sc3(y - sc3(e, ey), K / r);
    }
};
__device__ __forceinline__ bool spiky_adj(D3 pd, const AdjConst& k, SpikyAdj& o) {
    const double r2 = dot3(pd, pd);
    if (!(r2 > 0.0)) return false;
    o.r = sqrt(r2);
    const double q = o.r * k.warg;
    if (q > 1.0) return false;
    o.e = sc3(pd, 1.0 / o.r);
    o.K = k.K0 * (1.0 - q) * (1.0 - q);
    o.Kp = -2.0 * k.K0 * k.warg * (1.0 - q);
    return true;
}

// poly6 with inv_h along pd (the XSPH weight, pbf_kernel.inl:34-48):
// W = alpha t^3, t = 1 - r^2/h^2; grad_pd W = gw pd.  False beyond the support.
__device__ __forceinline__ bool poly6_adj(D3 pd, const AdjConst& k, double& W, double& gw) {
    const double q2 = dot3(pd, pd) * k.ih * k.ih;
    if (q2 > 1.0) return false;
    const double t = 1.0 - q2;
    W = k.alpha_x * t * t * t;
    gw = -6.0 * k.alpha_x * k.ih * k.ih * t * t;
    return true;
}

// tensile correction of a pair and its derivatives (pbf_position_update.cu:131-188)
__device__ __forceinline__ double corr_of(const AdjConst& k, double W, double si, double sj) {
    if (k.mode == 0) {
        const double cl = (si < 0.0 ? -si : 0.0) + (sj < 0.0 ? -sj : 0.0);
        const double W2 = W * W;
        return cl * 10.0 * W2 * W2 * k.inv_k;
    }
    return k.sk * powid(W * k.inv_wdq, k.sn);
}
// dcorr/dW
__device__ __forceinline__ double corr_dW(const AdjConst& k, double W, double si, double sj) {
    if (k.mode == 0) {
        const double cl = (si < 0.0 ? -si : 0.0) + (sj < 0.0 ? -sj : 0.0);
        return cl * 40.0 * W * W * W * k.inv_k;
    }
    return k.sn > 0 ? k.sk * (double)k.sn * powid(W * k.inv_wdq, k.sn - 1) * k.inv_wdq : 0.0;
}
// dcorr/ds of the end whose s is `s` (REFERENCE form only)
__device__ __forceinline__ double corr_ds(const AdjConst& k, double W, double s) {
    if (k.mode != 0 || !(s < 0.0)) return 0.0;
    const double W2 = W * W;
    return -10.0 * W2 * W2 * k.inv_k;
}

__device__ __forceinline__ uint32_t gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }

// the 27 search cells around the cell of p, in z, y, x order
template <class F>
__device__ __forceinline__ void for_cells(const Grid& G, F3 p, F&& f) {
    int g[3];
    grid_pos(p, G.cw, g);
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) f(grid_hash(g[0] + dx, g[1] + dy, g[2] + dz, G), dx, dy, dz);
}

// How a thread takes part in a pass over a particle's candidate set: the
// light kernel runs one thread per particle not flagged heavy (bit), the
// heavy kernel one wave per listed particle, its lanes taking every 64th
// candidate and reducing their sums in a fixed order.
struct Role {
    uint32_t i, lane, stride;
    bool on;
};
template <bool kHeavy>
__device__ __forceinline__ Role role(const StepTape& t, const AdjBufs& a, uint32_t bit) {
    Role r;
    if (!kHeavy) {
        r.i = gtid();
        r.lane = 0;
        r.stride = 1;
        r.on = r.i < t.n_live && !(a.hflag[r.i] & bit);
    } else {
        const uint32_t w = gtid() >> 6;
        const uint32_t set = (uint32_t)__builtin_ctz(bit);
        r.lane = threadIdx.x & 63u;
        r.stride = 64;
        r.on = w < a.nh[set];
        r.i = r.on ? a.hlist[(size_t)set * a.cap + w] : 0u;
    }
    return r;
}
__device__ __forceinline__ void mark_heavy(const AdjBufs& a, uint32_t i, uint32_t set) {
    a.hlist[(size_t)set * a.cap + atomicAdd(&a.nh[set], 1u)] = i;
}
// sum over the wave (the heavy kernels; every lane holds the total)
__device__ __forceinline__ double wsum(double v) {
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}
__device__ __forceinline__ D3 wsum3(D3 v) { return md3(wsum(v.x), wsum(v.y), wsum(v.z)); }

// S(i): the slots of the 27 cells around the final cell of p (the forward
// sum's candidates, pbf_contribution.inl:7-57); this thread's share.
template <class F>
__device__ __forceinline__ void for_S(const StepTape& t, const Grid& G, F3 p, const Role& r, F&& f) {
    for_cells(G, p, [&](uint32_t c, int, int, int) {
        const uint32_t e = t.offset[c + 1];
        for (uint32_t j = t.offset[c] + r.lane; j < e; j += r.stride) f(j);
    });
}
// T(i) = {l : i in S(l)}: the particles whose final cell is one of the 27
// cells around the search cell of i (p0 = i's predicted position); this
// thread's share.  f gets the position q in final-cell order (slot
// l = cperm[q]); the pass's data of l sit at gP[q], gV[q], gD[q].
template <class F>
__device__ __forceinline__ void for_T(const AdjBufs& a, const Grid& G, F3 p0, const Role& r, F&& f) {
    for_cells(G, p0, [&](uint32_t c, int, int, int) {
        const uint32_t e = a.coffset[c + 1];
        for (uint32_t q = a.coffset[c] + r.lane; q < e; q += r.stride) f(q);
    });
}

// Candidate counts of the final passes' sets and of the search cells;
// particles above kAdjHeavy are flagged and listed for the heavy kernels
// (the reverse-list bit is set by k_adj_rev).  Also clears the per-particle
// shares of the parameter gradients.
__global__ __launch_bounds__(256) void k_adj_classify(StepTape t, AdjBufs a, Grid G) {
    const uint32_t i = gtid();
    if (i >= t.n_live) return;
    uint32_t cs = 0, ct = 0, cn = 0;
    for_cells(G, f3of(t.PI[i]), [&](uint32_t c, int, int, int) { cs += t.offset[c + 1] - t.offset[c]; });
    for_cells(G, f3of(t.Q[i]), [&](uint32_t c, int, int, int) {
        ct += a.coffset[c + 1] - a.coffset[c];
        cn += t.offset[c + 1] - t.offset[c];
    });
    uint32_t f = 0;
    const uint32_t cnt[3] = {cs, ct, cn};
    for (uint32_t set = 0; set < 3; ++set)
        if (cnt[set] > kAdjHeavy) {
            f |= 1u << set;
            mark_heavy(a, i, set);
        }
    a.hflag[i] = f;
    a.pc[i] = 0.0;
    a.pe[i] = 0.0;
}

// Gathers what a transposed pass reads of each l into final-cell order:
// gP = (final position, slot bits) and, by pass,
//   0 XSPH:        gV = (velocity after confinement, NaN guard), gD = dL/dv_out
//   1 confinement: gD = dL/deta
//   2 vorticity:   gV = velocity before confinement (as the forward forms it), gD = dL/domega
__global__ __launch_bounds__(256) void k_adj_gather(StepTape t, AdjBufs a, Phys P, uint32_t cap, int pass) {
    const uint32_t q = gtid();
    if (q >= t.n_live) return;
    const uint32_t l = a.cperm[q];
    const float4 p = t.PI[l];
    a.gP[q] = make_float4(p.x, p.y, p.z, __uint_as_float(l));
    if (pass == 0) {
        const float4 vc = t.R[2 * l + 1];
        a.gV[q] = make_float4(vc.x, vc.y, vc.z, a.xnan[l] ? 1.f : 0.f);
        a.gD[q] = make_double4(a.av[3 * l], a.av[3 * l + 1], a.av[3 * l + 2], 0.0);
    } else if (pass == 1) {
        a.gD[q] = a.etab[l];
    } else {
        const F3 v = mulsv(P.inv_dt, f3of(p) - ld3(t.X, cap, l));
        a.gV[q] = make_float4(v.x, v.y, v.z, 0.f);
        a.gD[q] = a.aOm[l];
    }
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
    const uint32_t r = wave_rank(a.ccount, k, active, k);
    if (active) {
        mark_tile(a.ctflag, k >> lgT);
        a.ckey[s] = k;
        a.crank[s] = r;
    }
}

__global__ __launch_bounds__(256) void k_adj_cscatter(StepTape t, AdjBufs a) {
    const uint32_t s = gtid();
    if (s >= t.n_live) return;
    const uint32_t k = a.ckey[s], r = a.crank[s], o0 = a.coffset[k];
    a.ctmp[o0 + r] = s;
    if (r == 0u) list_big_cell(a.coffset, k, o0, a.big, a.nbig, a.cap);
}

// slot order inside each cell: position = number of smaller slots in the
// cell -- in a cell sorted by k_sort_cells, found by binary search
__global__ __launch_bounds__(256) void k_adj_corder(StepTape t, AdjBufs a) {
    const uint32_t s = gtid();
    if (s >= t.n_live) return;
    const uint32_t k = a.ckey[s];
    const uint32_t o0 = a.coffset[k], o1 = a.coffset[k + 1];
    uint32_t r = 0;
    if (sorted_cell(o1 - o0)) {
        uint32_t lo = o0, hi = o1;  // the first entry >= s (it is s)
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (a.ctmp[mid] < s) lo = mid + 1u;
            else hi = mid;
        }
        r = lo - o0;
    } else {
        for (uint32_t q = o0; q < o1; ++q) r += a.ctmp[q] < s ? 1u : 0u;
    }
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
        const uint32_t jl = t.nlist[list_at(cap, l, n - 1u)];
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
// count into rcount (and flag reverse lists above kAdjHeavy); fill = true:
// write rlist.  A heavy wave takes the candidates 64 at a time and keeps the
// scan order with a ballot.
template <bool fill, bool kHeavy>
__global__ __launch_bounds__(256) void k_adj_rev(StepTape t, AdjBufs a, Grid G, Phys P) {
    const Role r = role<kHeavy>(t, a, 4u);
    if (!r.on) return;
    const uint32_t j = r.i;
    const F3 self = f3of(t.Q[j]);
    const uint32_t full = (uint32_t)P.cap;
    uint32_t cnt = 0;
    uint32_t* out = fill ? a.rlist + a.roff[j] : nullptr;
    for_cells(G, self, [&](uint32_t c, int dx, int dy, int dz) {
        const uint32_t b0 = t.offset[c], e = t.offset[c + 1];
        const uint32_t pos = (uint32_t)((1 - dz) * 9 + (1 - dy) * 3 + (1 - dx));  // j seen from l
        for (uint32_t base = b0; base < e; base += r.stride) {
            const uint32_t l = base + r.lane;
            bool listed = false;
            if (l < e && l != j) {
                const F3 d = f3of(t.Q[l]) - self;  // the search's d = p_l - p_j
                const float r2 = __builtin_fmaf(d.z, d.z, __builtin_fmaf(d.y, d.y, d.x * d.x));
                if (r2 < P.h2) {
                    const uint32_t n = t.ncnt[l];
                    listed = n > 0 && n < full;
                    if (n == full) {
                        const uint2 L = a.last[l];
                        listed = pos < L.x || (pos == L.x && j <= L.y);
                    }
                }
            }
            if (!kHeavy) {
                if (listed) {
                    if (fill) out[cnt] = l;
                    ++cnt;
                }
            } else {
                const uint64_t m = __ballot(listed);
                if (fill && listed) out[cnt + (uint32_t)__popcll(m & ((1ull << r.lane) - 1ull))] = l;
                cnt += (uint32_t)__popcll(m);
            }
        }
    });
    if (!fill && r.lane == 0) {
        a.rcount[j] = cnt;
        if (cnt) mark_tile(a.rtflag, j >> a.lgRT);
        if (cnt > kAdjHeavy) {
            atomicOr(&a.hflag[j], 8u);
            mark_heavy(a, j, 3);
        }
    }
}

// ------------------------------------------------------------ final stage
// Every pass below has a light (one thread per particle) and a heavy (one
// wave per particle) instantiation over the same body.

// Clears dL/dVc = dL/dv_out, dL/dP = dL/dx_out (the stage's starting point).
__global__ __launch_bounds__(256) void k_adj_final_init(StepTape t, AdjBufs a) {
    const uint32_t i = gtid();
    if (i >= t.n_live) return;
    st(a.aVc, i, ld3d(a.av, i));
    st(a.aP, i, ld3d(a.ax, i));
}

// XSPH, own pairs: dL/dVc_i, the position terms of i's own sum, dL/dc
// (v_out = Vc + gamma sum_j (Vc_j - Vc_i) W(|p_j - p_i|)).
template <bool kHeavy>
__global__ __launch_bounds__(256) void k_adj_xsph_own(StepTape t, AdjBufs a, Grid G, Phys P) {
    const Role r = role<kHeavy>(t, a, 1u);
    if (!r.on) return;
    const uint32_t i = r.i;
    const AdjConst k = adj_const(P);
    const F3 selff = f3of(t.PI[i]);
    const D3 self = wid3(selff), vi = wid(t.R[2 * i + 1]);
    const D3 gi = ld3d(a.av, i);
    double sumW = 0.0;
    D3 acc = dz3(), apos = dz3();
    for_S(t, G, selff, r, [&](uint32_t j) {
        if (j == i) return;
        const D3 pd = wid(t.PI[j]) - self;
        if (dot3(pd, pd) > (double)P.cut2_w) return;
        double W, gw;
        if (!poly6_adj(pd, k, W, gw)) return;
        const D3 e = wid(t.R[2 * j + 1]) - vi;
        acc = acc + sc3(e, W);
        sumW += W;
        apos = apos - sc3(pd, gw * k.gamma * dot3(gi, e));
    });
    if (kHeavy) {
        acc = wsum3(acc);
        sumW = wsum(sumW);
        apos = wsum3(apos);
        if (r.lane) return;
    }
    const bool nan = isnan(acc.x) || isnan(acc.y) || isnan(acc.z);
    a.xnan[i] = nan ? 1u : 0u;
    if (nan) return;  // velocity kept: v_out = Vc
    st(a.aVc, i, sc3(gi, 1.0 - k.gamma * sumW));
    st(a.aP, i, ld(a.aP, i) + apos);
    a.pc[i] = k.beta * dot3(gi, acc);
}

// XSPH, transposed pairs (l, i), i in S(l).
template <bool kHeavy>
__global__ __launch_bounds__(256) void k_adj_xsph_rev(StepTape t, AdjBufs a, Grid G, Phys P) {
    const Role r = role<kHeavy>(t, a, 2u);
    if (!r.on) return;
    const uint32_t i = r.i;
    const AdjConst k = adj_const(P);
    const D3 pi = wid(t.PI[i]), vi = wid(t.R[2 * i + 1]);
    D3 avc = dz3(), apos = dz3();
    for_T(a, G, f3of(t.Q[i]), r, [&](uint32_t q) {
        const float4 pl = a.gP[q];
        const float4 vl = a.gV[q];
        if (__float_as_uint(pl.w) == i || vl.w != 0.f) return;
        const D3 pd = pi - wid(pl);
        if (dot3(pd, pd) > (double)P.cut2_w) return;
        double W, gw;
        if (!poly6_adj(pd, k, W, gw)) return;
        const double4 g4 = a.gD[q];
        const D3 gl = md3(g4.x, g4.y, g4.z);
        avc = avc + sc3(gl, k.gamma * W);
        apos = apos + sc3(pd, gw * k.gamma * dot3(gl, vi - wid(vl)));
    });
    if (kHeavy) {
        avc = wsum3(avc);
        apos = wsum3(apos);
        if (r.lane) return;
    }
    st(a.aVc, i, ld(a.aVc, i) + avc);
    st(a.aP, i, ld(a.aP, i) + apos);
}

// Confinement, direct terms (Vc = V + kappa (N x omega_i), N = eta/|eta|,
// eta_i = sum_j |omega_j| K(r) e_ij): dL/domega_i, dL/deta_i, dL/d eps;
// dL/dV starts as dL/dVc.
template <bool kHeavy>
__global__ __launch_bounds__(256) void k_adj_conf_own(StepTape t, AdjBufs a, Grid G, Phys P) {
    const Role r = role<kHeavy>(t, a, 1u);
    if (!r.on) return;
    const uint32_t i = r.i;
    const AdjConst k = adj_const(P);
    const F3 selff = f3of(t.PI[i]);
    const D3 self = wid3(selff);
    D3 eta = dz3();
    for_S(t, G, selff, r, [&](uint32_t j) {
        SpikyAdj sd;
        if (j == i || !spiky_adj(wid(t.PI[j]) - self, k, sd)) return;
        const D3 om = wid(t.OM4[j]);
        eta = eta + sc3(sd.f(), sqrt(dot3(om, om)));
    });
    if (kHeavy) {
        eta = wsum3(eta);
        if (r.lane) return;
    }
    const D3 avc = ld(a.aVc, i);
    st(a.aV, i, avc);
    const double nr = sqrt(dot3(eta, eta));
    D3 aom = dz3(), eb = dz3();
    if (nr > 0.0) {
        const D3 Nh = sc3(eta, 1.0 / nr), om = wid(t.OM4[i]);
        a.pe[i] = dot3(avc, cross3(Nh, om)) * k.dpdvc;
        const D3 y = sc3(avc, k.kappa);
        aom = cross3(y, Nh);
        const D3 nb = cross3(om, y);
        eb = sc3(nb - sc3(Nh, dot3(Nh, nb)), 1.0 / nr);
    }
    st(a.aOm, i, aom);
    st(a.etab, i, eb);
}

// Confinement through eta, own pairs: the position terms of eta_i.
template <bool kHeavy>
__global__ __launch_bounds__(256) void k_adj_eta_own(StepTape t, AdjBufs a, Grid G, Phys P) {
    const Role r = role<kHeavy>(t, a, 1u);
    if (!r.on) return;
    const uint32_t i = r.i;
    const D3 eb = ld(a.etab, i);
    if (zero3(eb)) return;  // uniform over a heavy wave
    const AdjConst k = adj_const(P);
    const F3 selff = f3of(t.PI[i]);
    const D3 self = wid3(selff);
    D3 apos = dz3();
    for_S(t, G, selff, r, [&](uint32_t j) {
        SpikyAdj sd;
        if (j == i || !spiky_adj(wid(t.PI[j]) - self, k, sd)) return;
        const D3 om = wid(t.OM4[j]);
        apos = apos - sc3(sd.H(eb), sqrt(dot3(om, om)));
    });
    if (kHeavy) {
        apos = wsum3(apos);
        if (r.lane) return;
    }
    st(a.aP, i, ld(a.aP, i) + apos);
}

// Confinement through eta, transposed pairs (l, i): dL/d|omega_i| and the
// position terms.
template <bool kHeavy>
__global__ __launch_bounds__(256) void k_adj_eta_rev(StepTape t, AdjBufs a, Grid G, Phys P) {
    const Role r = role<kHeavy>(t, a, 2u);
    if (!r.on) return;
    const uint32_t i = r.i;
    const AdjConst k = adj_const(P);
    const D3 self = wid(t.PI[i]);
    const D3 omi = wid(t.OM4[i]);
    const double wi = sqrt(dot3(omi, omi));
    double mb = 0.0;
    D3 apos = dz3();
    for_T(a, G, f3of(t.Q[i]), r, [&](uint32_t q) {
        const float4 pl = a.gP[q];
        if (__float_as_uint(pl.w) == i) return;
        const D3 el = ld(a.gD, q);
        if (zero3(el)) return;
        SpikyAdj sd;
        if (!spiky_adj(self - wid(pl), k, sd)) return;
        mb += dot3(el, sd.f());
        apos = apos + sc3(sd.H(el), wi);
    });
    if (kHeavy) {
        mb = wsum(mb);
        apos = wsum3(apos);
        if (r.lane) return;
    }
    if (wi > 0.0) st(a.aOm, i, ld(a.aOm, i) + sc3(omi, mb / wi));
    st(a.aP, i, ld(a.aP, i) + apos);
}

// Vorticity (omega_i = beta sum_j (V_j - V_i) x (K(r) e_ij)), own pairs:
// dL/dV and the position terms.  V as the forward forms it (FP32).
template <bool kHeavy>
__global__ __launch_bounds__(256) void k_adj_vort_own(StepTape t, AdjBufs a, Grid G, Phys P, uint32_t cap) {
    const Role r = role<kHeavy>(t, a, 1u);
    if (!r.on) return;
    const uint32_t i = r.i;
    const D3 ab = ld(a.aOm, i);
    if (zero3(ab)) return;  // uniform over a heavy wave
    const AdjConst k = adj_const(P);
    const F3 selff = f3of(t.PI[i]);
    const D3 self = wid3(selff);
    const D3 vi = wid3(mulsv(P.inv_dt, selff - ld3(t.X, cap, i)));
    D3 av = dz3(), apos = dz3();
    for_S(t, G, selff, r, [&](uint32_t j) {
        SpikyAdj sd;
        const F3 pjf = f3of(t.PI[j]);
        if (j == i || !spiky_adj(wid3(pjf) - self, k, sd)) return;
        const D3 e = wid3(mulsv(P.inv_dt, pjf - ld3(t.X, cap, j))) - vi;
        av = av - sc3(cross3(sd.f(), ab), k.beta);
        apos = apos - sd.H(sc3(cross3(ab, e), k.beta));
    });
    if (kHeavy) {
        av = wsum3(av);
        apos = wsum3(apos);
        if (r.lane) return;
    }
    st(a.aV, i, ld(a.aV, i) + av);
    st(a.aP, i, ld(a.aP, i) + apos);
}

// Vorticity, transposed pairs (l, i).
template <bool kHeavy>
__global__ __launch_bounds__(256) void k_adj_vort_rev(StepTape t, AdjBufs a, Grid G, Phys P, uint32_t cap) {
    const Role r = role<kHeavy>(t, a, 2u);
    if (!r.on) return;
    const uint32_t i = r.i;
    const AdjConst k = adj_const(P);
    const F3 selff = f3of(t.PI[i]);
    const D3 self = wid3(selff);
    const D3 vi = wid3(mulsv(P.inv_dt, selff - ld3(t.X, cap, i)));
    D3 av = dz3(), apos = dz3();
    for_T(a, G, f3of(t.Q[i]), r, [&](uint32_t q) {
        const float4 pl = a.gP[q];
        if (__float_as_uint(pl.w) == i) return;
        const D3 al = ld(a.gD, q);
        if (zero3(al)) return;
        SpikyAdj sd;
        if (!spiky_adj(self - wid(pl), k, sd)) return;
        const D3 e = vi - wid(a.gV[q]);
        av = av + sc3(cross3(sd.f(), al), k.beta);
        apos = apos + sd.H(sc3(cross3(al, e), k.beta));
    });
    if (kHeavy) {
        av = wsum3(av);
        apos = wsum3(apos);
        if (r.lane) return;
    }
    st(a.aV, i, ld(a.aV, i) + av);
    st(a.aP, i, ld(a.aP, i) + apos);
}

// The update V = (P - X)/dt: dL/dP += dL/dV / dt, dL/dX = -dL/dV / dt.
__global__ __launch_bounds__(256) void k_adj_update(StepTape t, AdjBufs a, Phys P, int vort) {
    const uint32_t i = gtid();
    if (i >= t.n_live) return;
    const D3 av = ld(vort ? a.aV : a.aVc, i);
    const double inv_dt = P.inv_dt;
    st(a.aP, i, ld(a.aP, i) + sc3(av, inv_dt));
    st(a.aX, i, sc3(av, -inv_dt));
}

// ------------------------------------------------------------ Jacobi iteration
// Boundary responses of k_deltap's collide() with their Jacobian: every
// response that fires maps the point pp linearly around itself --
// planes: pp - ((pp - P).n) n, Jacobian I - n n^T (both cases of
// pbf_plane_boundary.cu:6-45 reduce to it); spheres (pbf_sphere_boundary.cu:8-64):
// pp + nudge cd - R cd/|cd| with cd = c - pp, Jacobian
// (1 - nudge) I + (R/|cd|)(I - u u^T), u = cd/|cd|.  The decisions and the
// points follow collide() exactly (FP32, same arithmetic); J (FP64)
// accumulates the product of the Jacobians of the responses that fired.
__device__ void collide_jac(const Phys& P, const Boundary& B, F3& dp, F3 pk, double J[9]) {
    auto apply = [&](const double M[9]) {
        double T[9];
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
        for (uint32_t q = 0; q < B.n_inner + B.n_outer; ++q) {
            const bool is_inner = q < B.n_inner;
            const float4 s = is_inner ? B.inner[q] : B.outer[q - B.n_inner];
            const bool inner = is_inner || P.outer_as_inner != 0;
            const F3 cdf = mk3(s.x, s.y, s.z) - pp;
            const float d2 = fdot(cdf, cdf);
            if (inner ? (d2 > s.w * s.w) : (d2 < s.w * s.w)) {
                const D3 cd = md3((double)s.x - pp.x, (double)s.y - pp.y, (double)s.z - pp.z);
                const double L = sqrt(dot3(cd, cd));
                const D3 u = sc3(cd, 1.0 / L);
                const double nudge = inner ? (double)1.0001f : (double)0.9999f;
                const double c0 = 1.0 - nudge, c1 = (double)s.w / L;
                const double M[9] = {c0 + c1 * (1.0 - u.x * u.x), -c1 * u.x * u.y, -c1 * u.x * u.z,
                                     -c1 * u.y * u.x, c0 + c1 * (1.0 - u.y * u.y), -c1 * u.y * u.z,
                                     -c1 * u.z * u.x, -c1 * u.z * u.y, c0 + c1 * (1.0 - u.z * u.z)};
                apply(M);
            }
            response_sphere(hit, dp, pp, mk3(s.x, s.y, s.z), s.w, inner);
            pp = pred_x + dp;
        }
        for (uint32_t q = 0; q < B.n_planes; ++q) {
            const F3 Pp = mk3(B.pts[3 * q], B.pts[3 * q + 1], B.pts[3 * q + 2]);
            const F3 nf = mk3(B.nrm[3 * q], B.nrm[3 * q + 1], B.nrm[3 * q + 2]);
            bool h = false;
            response_plane(h, dp, old_x, pp, Pp, nf);
            if (h) {
                hit = true;
                const D3 n = wid3(nf);
                const double M[9] = {1.0 - n.x * n.x, -n.x * n.y, -n.x * n.z, -n.y * n.x, 1.0 - n.y * n.y,
                                     -n.y * n.z, -n.z * n.x, -n.z * n.y, 1.0 - n.z * n.z};
                apply(M);
            }
            pp = pred_x + dp;
        }
        if (!hit) break;
    }
}

// dL/d(pre-collision position) = J^T dL/dP'.  dp is recomputed exactly as
// k_deltap computes it (FP32), so the responses replay the forward's
// decisions.
__global__ __launch_bounds__(256) void k_adj_coll(StepTape t, AdjBufs a, const float4* __restrict__ Q, Phys P,
                                                  Boundary B, uint32_t cap) {
    const uint32_t i = gtid();
    if (i >= t.n_live) return;
    D3 ag = ld(a.aP, i);
    if (B.n_planes + B.n_inner + B.n_outer > 0) {
        const float4 se = Q[i];
        const F3 self = f3of(se);
        const float si = se.w;
        const float self_cl = (si < 0.f) ? -si : 0.f;
        const uint32_t n = t.ncnt[i];
        F3 sum = mk3(0.f, 0.f, 0.f);
        for (uint32_t q = 0; q < n; ++q) {
            const float4 pe = Q[t.nlist[list_at(cap, i, q)]];
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
        double J[9] = {1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0};
        collide_jac(P, B, dp, self, J);
        ag = md3(J[0] * ag.x + J[3] * ag.y + J[6] * ag.z, J[1] * ag.x + J[4] * ag.y + J[7] * ag.z,
                 J[2] * ag.x + J[5] * ag.y + J[8] * ag.z);
    }
    st(a.abar, i, ag);
}

// delta p = beta sum_j A_ij kg_ij, A_ij = -(s_i + s_j + corr(W_ij)):
// dL/ds_i and the position terms through kg and corr, own pairs (i, j) and
// transposed pairs (l, i); then the coefficients of the lambda pass.
// Particles on long reverse lists run one wave each.
template <bool kHeavy>
__global__ __launch_bounds__(256) void k_adj_lam(StepTape t, AdjBufs a, const float4* __restrict__ Q, Phys P,
                                                 uint32_t cap) {
    const Role r = role<kHeavy>(t, a, 8u);
    if (!r.on) return;
    const uint32_t i = r.i;
    const AdjConst k = adj_const(P);
    const float4 se = Q[i];
    const D3 self = wid(se);
    const double si = se.w;
    const D3 ai = ld(a.abar, i);
    double sb = 0.0, kg2s = 0.0;
    D3 kgs = dz3(), apos = dz3();
    const uint32_t n = t.ncnt[i];
    for (uint32_t q = r.lane; q < n; q += r.stride) {
        const float4 pe = Q[t.nlist[list_at(cap, i, q)]];
        const D3 d = self - wid(pe);
        PairAdj pa;
        if (!pair_adj(d, k, pa)) continue;  // r = 0: W = a (constant), kg = 0
        kgs = kgs + pa.kg;
        kg2s += dot3(pa.kg, pa.kg);
        const double sj = pe.w;
        const double A = -(si + sj + corr_of(k, pa.W, si, sj));
        const double Ab = k.beta * dot3(pa.kg, ai);
        sb -= Ab;
        sb -= Ab * corr_ds(k, pa.W, si);
        const double Wb = -Ab * corr_dW(k, pa.W, si, sj);
        apos = apos + sc3(pa.H(ai), k.beta * A) + sc3(d, Wb * pa.gw);
    }
    const uint32_t r1 = a.roff[i + 1];
    for (uint32_t q = a.roff[i] + r.lane; q < r1; q += r.stride) {
        const uint32_t l = a.rlist[q];
        const float4 pe = Q[l];
        const D3 d = wid(pe) - self;  // the pair (l, i): d = p_l - p_i
        PairAdj pa;
        if (!pair_adj(d, k, pa)) continue;
        const double sl = pe.w;
        const D3 al = ld(a.abar, l);
        const double A = -(sl + si + corr_of(k, pa.W, sl, si));
        const double Ab = k.beta * dot3(pa.kg, al);
        sb -= Ab;
        sb -= Ab * corr_ds(k, pa.W, si);
        const double Wb = -Ab * corr_dW(k, pa.W, sl, si);
        apos = apos - (sc3(pa.H(al), k.beta * A) + sc3(d, Wb * pa.gw));
    }
    if (kHeavy) {
        sb = wsum(sb);
        kg2s = wsum(kg2s);
        kgs = wsum3(kgs);
        apos = wsum3(apos);
        if (r.lane) return;
    }
    // s = C / D, C = beta sum W - 1, D = m2r2 (|sum kg|^2 + sum |kg|^2) + eps
    const double D = k.m2r2 * (dot3(kgs, kgs) + kg2s) + k.relax;
    const double ksb = k.beta * sb / D;
    const double g2 = -2.0 * k.m2r2 * sb * si / D;
    a.coef[i] = make_double4(ksb, g2 * kgs.x, g2 * kgs.y, g2 * kgs.z);
    a.g2[i] = g2;
    st(a.aQ, i, ai + apos);
}

// lambda pass: the position terms through W and kg of s_i (own pairs) and
// s_l (transposed pairs); dL/dP of the iteration's input.
template <bool kHeavy>
__global__ __launch_bounds__(256) void k_adj_pos(StepTape t, AdjBufs a, const float4* __restrict__ Q, Phys P,
                                                 uint32_t cap) {
    const Role r = role<kHeavy>(t, a, 8u);
    if (!r.on) return;
    const uint32_t i = r.i;
    const AdjConst k = adj_const(P);
    const D3 self = wid(Q[i]);
    const double4 ci = a.coef[i];
    const double g2i = a.g2[i];
    D3 apos = dz3();
    if (ci.x != 0.0 || g2i != 0.0) {
        const uint32_t n = t.ncnt[i];
        for (uint32_t q = r.lane; q < n; q += r.stride) {
            const D3 d = self - wid(Q[t.nlist[list_at(cap, i, q)]]);
            PairAdj pa;
            if (!pair_adj(d, k, pa)) continue;
            apos = apos + sc3(d, ci.x * pa.gw) + pa.H(md3(ci.y, ci.z, ci.w) + sc3(pa.kg, g2i));
        }
    }
    const uint32_t r1 = a.roff[i + 1];
    for (uint32_t q = a.roff[i] + r.lane; q < r1; q += r.stride) {
        const uint32_t l = a.rlist[q];
        const double4 cl = a.coef[l];
        const double g2l = a.g2[l];
        if (cl.x == 0.0 && g2l == 0.0) continue;
        const D3 d = wid(Q[l]) - self;
        PairAdj pa;
        if (!pair_adj(d, k, pa)) continue;
        apos = apos - (sc3(d, cl.x * pa.gw) + pa.H(md3(cl.y, cl.z, cl.w) + sc3(pa.kg, g2l)));
    }
    if (kHeavy) {
        apos = wsum3(apos);
        if (r.lane) return;
    }
    st(a.aP, i, ld(a.aQ, i) + apos);
}

// ------------------------------------------------------------ predict + sort
// P0_s = x*_{sidx[s]}, x* = x + dt (v + g dt); X_s = x_{sidx[s]}.
__global__ __launch_bounds__(256) void k_adj_in(StepTape t, AdjBufs a, Phys P) {
    const uint32_t s = gtid();
    if (s >= t.n_in) return;
    const uint32_t i = t.sidx[s];
    D3 gx = dz3(), gv = dz3();
    if (s < t.n_live) {
        const D3 ap = ld(a.aP, s);
        gx = ap + ld(a.aX, s);
        gv = sc3(ap, (double)P.dt);
    }
    a.axin[3 * i] = gx.x;
    a.axin[3 * i + 1] = gx.y;
    a.axin[3 * i + 2] = gx.z;
    a.avin[3 * i] = gv.x;
    a.avin[3 * i + 1] = gv.y;
    a.avin[3 * i + 2] = gv.z;
}

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

// psum = (sum of pc, sum of pe) over the live slots, in a fixed order
__global__ __launch_bounds__(256) void k_adj_reduce(StepTape t, AdjBufs a) {
    __shared__ double sh[256];
    for (int w = 0; w < 2; ++w) {
        const double* src = w ? a.pe : a.pc;
        double v = 0.0;
        for (uint32_t q = threadIdx.x; q < t.n_live; q += blockDim.x) v += src[q];
        const double tot = block_sum(v, sh);
        if (threadIdx.x == 0) a.psum[w] = tot;
        __syncthreads();
    }
}

// largest |axin|, |avin| (NaN counts as infinity), as the bits of a double
// (non-negative doubles order like their bit patterns)
__global__ __launch_bounds__(256) void k_adj_absmax(AdjBufs a, uint32_t n3) {
    const uint32_t q = gtid();
    double m = 0.0;
    if (q < n3) {
        m = fmax(fabs(a.axin[q]), fabs(a.avin[q]));
        if (isnan(a.axin[q]) || isnan(a.avin[q])) m = __builtin_inf();
    }
    for (int d = 32; d > 0; d >>= 1) m = fmax(m, __shfl_xor(m, d, 64));
    if ((threadIdx.x & 63u) == 0 && m > 0.0)
        atomicMax(a.amax, (unsigned long long)__double_as_longlong(m));
}

__global__ __launch_bounds__(256) void k_adj_scale(double* x, double* y, uint32_t n, double f) {
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
                           hipStream_t s, uint32_t* h_nh) {
    const Grid& G = s0.G;
    const Phys& P = s0.P;
    const uint32_t nl = t.n_live;
    const dim3 blk(256);
    const dim3 gl(nblk(nl > 0 ? nl : 1));
    const uint32_t cap = b.cap;
    hipError_t e = hipSuccess;
    if (nl > 0) {
        // particles grouped by final cell
        hipLaunchKernelGGL(k_adj_ckey, gl, blk, 0, s, t, a, G, b.lgT);
        e = launch_count_scan(a.ccount, a.coffset, a.ctsum, a.ctflag, a.ctprefix, b.nT, b.lgT, b.CT, s);
        if (e != hipSuccess) return e;
        e = hipMemsetAsync(a.nbig, 0, 2 * sizeof(uint32_t), s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_adj_cscatter, gl, blk, 0, s, t, a);
        // the ranks are dead after the scatter: their array holds the huge cells' chunks
        e = launch_sort_cells(a.coffset, a.big, a.nbig, a.ctmp, a.crank, a.cap, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_adj_corder, gl, blk, 0, s, t, a);
        // heavy lists of the final passes and the search cells: their
        // lengths size the heavy launches
        e = hipMemsetAsync(a.nh, 0, 4 * sizeof(uint32_t), s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_adj_classify, gl, blk, 0, s, t, a, G);
        e = hipMemcpyAsync(h_nh, a.nh, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        auto ghv = [](uint32_t n) { return dim3(nblk(n * 64u > 0 ? n * 64u : 1u)); };
        const uint32_t hs = h_nh[0], ht = h_nh[1], hn = h_nh[2];
        // reverse lists
        hipLaunchKernelGGL(k_adj_last, gl, blk, 0, s, t, a, G, P, cap);
        hipLaunchKernelGGL((k_adj_rev<false, false>), gl, blk, 0, s, t, a, G, P);
        if (hn) hipLaunchKernelGGL((k_adj_rev<false, true>), ghv(hn), blk, 0, s, t, a, G, P);
        e = launch_count_scan(a.rcount, a.roff, a.rtsum, a.rtflag, a.rtprefix, a.nRT, a.lgRT, a.RT, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_adj_rev<true, false>), gl, blk, 0, s, t, a, G, P);
        if (hn) hipLaunchKernelGGL((k_adj_rev<true, true>), ghv(hn), blk, 0, s, t, a, G, P);
        e = hipMemcpyAsync(h_nh + 3, a.nh + 3, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        const uint32_t hr = h_nh[3];
        const dim3 ghS = ghv(hs), ghT = ghv(ht);
        // final stage
        hipLaunchKernelGGL(k_adj_final_init, gl, blk, 0, s, t, a);
        if (P.xsph_on) {
            hipLaunchKernelGGL(k_adj_xsph_own<false>, gl, blk, 0, s, t, a, G, P);
            if (hs) hipLaunchKernelGGL(k_adj_xsph_own<true>, ghS, blk, 0, s, t, a, G, P);
            hipLaunchKernelGGL(k_adj_gather, gl, blk, 0, s, t, a, P, cap, 0);
            hipLaunchKernelGGL(k_adj_xsph_rev<false>, gl, blk, 0, s, t, a, G, P);
            if (ht) hipLaunchKernelGGL(k_adj_xsph_rev<true>, ghT, blk, 0, s, t, a, G, P);
        }
        if (P.vort_on) {
            hipLaunchKernelGGL(k_adj_conf_own<false>, gl, blk, 0, s, t, a, G, P);
            if (hs) hipLaunchKernelGGL(k_adj_conf_own<true>, ghS, blk, 0, s, t, a, G, P);
            hipLaunchKernelGGL(k_adj_eta_own<false>, gl, blk, 0, s, t, a, G, P);
            if (hs) hipLaunchKernelGGL(k_adj_eta_own<true>, ghS, blk, 0, s, t, a, G, P);
            hipLaunchKernelGGL(k_adj_gather, gl, blk, 0, s, t, a, P, cap, 1);
            hipLaunchKernelGGL(k_adj_eta_rev<false>, gl, blk, 0, s, t, a, G, P);
            if (ht) hipLaunchKernelGGL(k_adj_eta_rev<true>, ghT, blk, 0, s, t, a, G, P);
            hipLaunchKernelGGL(k_adj_gather, gl, blk, 0, s, t, a, P, cap, 2);
            hipLaunchKernelGGL(k_adj_vort_own<false>, gl, blk, 0, s, t, a, G, P, cap);
            if (hs) hipLaunchKernelGGL(k_adj_vort_own<true>, ghS, blk, 0, s, t, a, G, P, cap);
            hipLaunchKernelGGL(k_adj_vort_rev<false>, gl, blk, 0, s, t, a, G, P, cap);
            if (ht) hipLaunchKernelGGL(k_adj_vort_rev<true>, ghT, blk, 0, s, t, a, G, P, cap);
        }
        hipLaunchKernelGGL(k_adj_update, gl, blk, 0, s, t, a, P, P.vort_on);
        // Jacobi iterations, last first
        for (int it = iters - 1; it >= 0; --it) {
            const float4* Q = t.Q + (size_t)it * cap;
            hipLaunchKernelGGL(k_adj_coll, gl, blk, 0, s, t, a, Q, P, s0.B, cap);
            hipLaunchKernelGGL(k_adj_lam<false>, gl, blk, 0, s, t, a, Q, P, cap);
            if (hr) hipLaunchKernelGGL(k_adj_lam<true>, ghv(hr), blk, 0, s, t, a, Q, P, cap);
            hipLaunchKernelGGL(k_adj_pos<false>, gl, blk, 0, s, t, a, Q, P, cap);
            if (hr) hipLaunchKernelGGL(k_adj_pos<true>, ghv(hr), blk, 0, s, t, a, Q, P, cap);
        }
    }
    if (t.n_in > 0) hipLaunchKernelGGL(k_adj_in, dim3(nblk(t.n_in)), blk, 0, s, t, a, P);
    hipLaunchKernelGGL(k_adj_reduce, dim3(1), blk, 0, s, t, a);
    return hipGetLastError();
}

hipError_t launch_adj_absmax(const AdjBufs& a, uint32_t n, hipStream_t s) {
    hipError_t e = hipMemsetAsync(a.amax, 0, sizeof(*a.amax), s);
    if (e != hipSuccess) return e;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adj_absmax, dim3(nblk(3 * n)), dim3(256), 0, s, a, 3 * n);
    return hipGetLastError();
}

hipError_t launch_adj_scale(double* x, double* y, uint32_t n, double f, hipStream_t s) {
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
