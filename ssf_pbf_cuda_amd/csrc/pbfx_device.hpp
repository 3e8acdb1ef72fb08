// pbfx_device.hpp -- device helpers shared by the solver's kernels
// (pbfx_kernels.hip) and the adjoint kernels (pbfx_adjoint.hip): vector math
// in GLM's operation order, the SPH kernels, the hashed grid, the scene
// accessors, the memo pair terms, the boundary responses and the wave/block
// primitives.  Citations are relative to cuda_pbf_with_ssf/src/ of
// yodaplusplus/ssf_pbf_cuda.
#pragma once

#include "pbfx_internal.hpp"

namespace pbfx {
namespace {

// ---------------------------------------------------------------- vector math
struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 mk3(float x, float y, float z) {
    F3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ void acc(F3& a, F3 b) {
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
}
__device__ __forceinline__ F3 mulvs(F3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ F3 mulsv(float s, F3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ F3 divvs(F3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
// GLM 0.9.7.1 dot/length/cross (external/include/glm/detail/func_geometric.inl:65-72,95-100,134-142)
__device__ __forceinline__ float glm_dot(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float glm_length(F3 a) { return __builtin_sqrtf(glm_dot(a, a)); }
__device__ __forceinline__ F3 glm_cross(F3 x, F3 y) {
    return mk3(x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y);
}
// the hand-fused dot of the solver kernels: fmaf(a.z,b.z, fmaf(a.y,b.y, a.x*b.x))
__device__ __forceinline__ float fdot(F3 a, F3 b) { return __builtin_fmaf(a.z, b.z, __builtin_fmaf(a.y, b.y, a.x * b.x)); }

__device__ __forceinline__ F3 ld3(const float* __restrict__ a, uint32_t cap, uint32_t i) {
    return mk3(a[i], a[cap + i], a[2 * cap + i]);
}
__device__ __forceinline__ void st3(float* __restrict__ a, uint32_t cap, uint32_t i, F3 v) {
    a[i] = v.x;
    a[cap + i] = v.y;
    a[2 * cap + i] = v.z;
}

// pbf/kernel/cuda/detail/pbf_kernel.inl:34-48
__device__ __forceinline__ float wpoly6(float r, float inv_h) {
    float q = r * inv_h;
    float alpha = 1.56668147065f * inv_h * inv_h * inv_h;
    if (q <= 1.f) {
        float q2 = q * q;
        float t = 1.f - q2;
        return alpha * t * t * t;
    }
    return 0.f;
}

// pbf/kernel/cuda/detail/pbf_kernel.inl:50-63
__device__ __forceinline__ float wspiky_deriv(float r, float inv_h) {
    float q = r * inv_h;
    float inv_h2 = inv_h * inv_h;
    if (q <= 1.f) {
        float t = 1.f - q;
        float inv_h4 = inv_h2 * inv_h2;
        return -14.3239448745f * inv_h4 * t * t;
    }
    return 0.f;
}

// float -> int with a defined result everywhere (NaN -> 0, saturation), the
// behaviour of the device conversion the reference relies on.
__device__ __forceinline__ int f2i(float f) {
    if (!(f == f)) return 0;
    if (f >= 2147483648.0f) return INT_MAX;
    if (f <= -2147483648.0f) return INT_MIN;
    return (int)f;
}

// pbf/interaction/cuda/pbf_grid.h:20-28 (IEEE division) and :54-61 (wrap)
__device__ __forceinline__ void grid_pos(F3 p, float cw, int g[3]) {
    g[0] = f2i(floorf(p.x / cw));
    g[1] = f2i(floorf(p.y / cw));
    g[2] = f2i(floorf(p.z / cw));
}
__device__ __forceinline__ uint32_t grid_hash(int gx, int gy, int gz, const Grid& G) {
    uint32_t x = (uint32_t)gx & (G.gx - 1u);
    uint32_t y = (uint32_t)gy & (G.gy - 1u);
    uint32_t z = (uint32_t)gz & (G.gz - 1u);
    return z * G.gx * G.gy + y * G.gx + x;
}

// cell key of a scene's grid position (the scene's range of the key space)
__device__ __forceinline__ uint32_t scene_key(int gx, int gy, int gz, const Grid& G) {
    return G.base + grid_hash(gx, gy, gz, G);
}

// Largest s < n with off[s] <= i (the scene owning index or key i).
__device__ __forceinline__ uint32_t find_scene(const uint32_t* __restrict__ off, uint32_t n, uint32_t i) {
    uint32_t lo = 0, hi = n - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1u;
    }
    return lo;
}

// The scenes of a launch: a single scene travels as a kernel argument;
// several are read from the device table.
template <bool kB>
struct Scenes {
    const Scene* table;
    Scene s0;
    __device__ __forceinline__ const Scene& operator[](uint32_t s) const { return kB ? table[s] : s0; }
};

template <bool kB>
__device__ __forceinline__ uint32_t scene_index(const uint32_t* __restrict__ off, uint32_t n, uint32_t i) {
    return kB ? find_scene(off, n, i) : 0u;
}

// Runs f(s) for the active lanes, one scene at a time: s is uniform over the
// lanes that call f (read from the first pending lane), so the scene's
// constants come through scalar loads.  f must not use wave-wide operations.
template <bool kB, class F>
__device__ __forceinline__ void by_scene(uint32_t sid, bool active, F&& f) {
    if (!kB) {
        if (active) f(0u);
        return;
    }
    bool todo = active;
    while (__any(todo)) {
        if (todo) {
            const uint32_t s = __builtin_amdgcn_readfirstlane(sid);
            if (sid == s) {
                f(s);
                todo = false;
            }
        }
    }
}

// t^n by binary exponentiation; n = 4 evaluates (t*t)*(t*t)
__device__ __forceinline__ float powi(float t, int n) {
    float result = 1.f, base = t;
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

// Element index of entry k of slot i's frozen list.  Lists are stored in
// groups of four entries: entries 4q..4q+3 of every slot form the 16-byte
// word q * cap + i, so that one wave reads the next four entries of its 64
// slots as one coalesced 1 KiB load.
__device__ __forceinline__ size_t list_at(uint32_t cap, uint32_t i, uint32_t k) {
    return (((size_t)(k >> 2) * cap + i) << 2) | (k & 3u);
}

// The word holding entries k..k+3 (k a multiple of 4) of slot i's list.
__device__ __forceinline__ uint4 list_word(const uint32_t* __restrict__ nl, uint32_t cap, uint32_t i, uint32_t k) {
    return reinterpret_cast<const uint4*>(nl)[(size_t)(k >> 2) * cap + i];
}

// The indices of word e = entries k..k+3 (k < n); entries at or past n are
// replaced by entry k, so every index returned is a written one.
__device__ __forceinline__ void list_group(uint4 e, uint32_t k, uint32_t n, uint32_t j[4]) {
    j[0] = e.x;
    j[1] = (k + 1u < n) ? e.y : e.x;
    j[2] = (k + 2u < n) ? e.z : e.x;
    j[3] = (k + 3u < n) ? e.w : e.x;
}

// Visits entries 0..n-1 of slot i's list in order, eight at a time: the
// two list words of a group and its eight gathers (load(j) -> float4) are in
// flight together, and the next group's words while this one is summed.
// Entries at or past n are replaced by the group's first one (a written
// index) and not passed to body(k, element).
template <class LoadF, class BodyF>
__device__ __forceinline__ void for_list8(const uint32_t* __restrict__ nl, uint32_t cap, uint32_t i, uint32_t n,
                                          LoadF&& load, BodyF&& body) {
    if (n == 0u) return;
    const uint32_t last = (n - 1u) & ~3u;  // the last word holding a written entry
    uint4 e0 = list_word(nl, cap, i, 0u), e1 = list_word(nl, cap, i, min(4u, last));
    for (uint32_t k = 0; k < n; k += 8u) {
        uint32_t j[8] = {e0.x, e0.y, e0.z, e0.w, e1.x, e1.y, e1.z, e1.w};
#pragma unroll
        for (uint32_t m = 1; m < 8u; ++m)
            if (k + m >= n) j[m] = j[0];
        float4 el[8];
#pragma unroll
        for (uint32_t m = 0; m < 8u; ++m) el[m] = load(j[m]);
        e0 = list_word(nl, cap, i, min(k + 8u, last));
        e1 = list_word(nl, cap, i, min(k + 12u, last));
#pragma unroll
        for (uint32_t m = 0; m < 8u; ++m)
            if (k + m < n) body(k + m, el[m]);
    }
}

// Gathers of 16-byte records through a buffer resource: one 32-bit byte
// offset per lane (no 64-bit address arithmetic), and an offset past the
// n records reads zeros instead of faulting.  Valid for n < kGatherMaxSlots.
constexpr uint32_t kGatherMaxSlots = 1u << 28;
struct Gather16 {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ Gather16(const float4* p, uint32_t n)
        : r(__builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)(n * 16u), 0x00020000)) {}
    __device__ __forceinline__ float4 operator()(uint32_t j) const {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)(j << 4), 0, 0);
        return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), __uint_as_float(v[3]));
    }
};

// for_list8 for a load that takes any index (Gather16): the entries past n
// in the last word are loaded as they are (whatever the word holds) and not
// passed to body, which saves the index selects.
template <class LoadF, class BodyF>
__device__ __forceinline__ void for_list8_any(const uint32_t* __restrict__ nl, uint32_t cap, uint32_t i, uint32_t n,
                                              LoadF&& load, BodyF&& body) {
    if (n == 0u) return;
    const uint32_t last = (n - 1u) & ~3u;  // the last word holding a written entry
    uint4 e0 = list_word(nl, cap, i, 0u), e1 = list_word(nl, cap, i, min(4u, last));
    for (uint32_t k = 0; k < n; k += 8u) {
        const uint32_t j[8] = {e0.x, e0.y, e0.z, e0.w, e1.x, e1.y, e1.z, e1.w};
        float4 el[8];
#pragma unroll
        for (uint32_t m = 0; m < 8u; ++m) el[m] = load(j[m]);
        e0 = list_word(nl, cap, i, min(k + 8u, last));
        e1 = list_word(nl, cap, i, min(k + 12u, last));
#pragma unroll
        for (uint32_t m = 0; m < 8u; ++m)
            if (k + m < n) body(k + m, el[m]);
    }
}

// The Jacobi passes' walk over slot i's list: buffer gathers when the
// records' byte offsets fit 32 bits, plain loads otherwise.
template <class BodyF>
__device__ __forceinline__ void walk_list(const uint32_t* __restrict__ nl, uint32_t cap, uint32_t i, uint32_t n,
                                          const float4* __restrict__ rec, BodyF&& body) {
    if (cap < kGatherMaxSlots)
        for_list8_any(nl, cap, i, n, Gather16(rec, cap), body);
    else
        for_list8(nl, cap, i, n, [&](uint32_t j) { return rec[j]; }, body);
}

// 1.f / sqrtf(r2) with both operations correctly rounded -- the IEEE
// result -- in about a third of the instructions of the IEEE sequences.  On
// [2^-96, FLT_MAX] (normal_r2): v_sqrt_f32, then the two-sided fma test of
// its neighbours (s - 1 ulp is taken when fma(-(s - ulp), s, r2) <= 0, s + 1
// ulp when fma(-(s + ulp), s, r2) > 0), then v_rcp_f32 and one Newton step
// y + y * fma(-s, y, 1).  Outside that range (0, below 2^-96, inf, NaN) the
// IEEE operations themselves.  pbfx_selftest_inv_sqrt compares this function
// with the IEEE form on every float of the range.
__device__ __forceinline__ bool normal_r2(float r2) { return r2 >= 0x1p-96f && r2 <= 0x1.fffffep+127f; }

__device__ __forceinline__ float inv_sqrt_normal(float r2) {
    const float s0 = __builtin_amdgcn_sqrtf(r2);
    const uint32_t si = __float_as_uint(s0);
    const float sd = __uint_as_float(si - 1u), su = __uint_as_float(si + 1u);
    const float vd = __builtin_fmaf(-sd, s0, r2), vu = __builtin_fmaf(-su, s0, r2);
    float s = (vd <= 0.f) ? sd : s0;
    s = (vu > 0.f) ? su : s;
    const float y = __builtin_amdgcn_rcpf(s);
    return __builtin_fmaf(__builtin_fmaf(-s, y, 1.f), y, y);
}

__device__ __forceinline__ float inv_sqrt(float r2) {
    if (__builtin_expect(normal_r2(r2), 1)) return inv_sqrt_normal(r2);
    return 1.f / __builtin_sqrtf(r2);
}

// pbf/solver/cuda/pbf_kernel_memorization.cu:15-56 (memorizeKernelPair):
// W and grad W of one listed pair from d = self - x_j and r2 = |d|^2
// (rsqrtf restated as 1/sqrtf), for any r2.
__device__ __forceinline__ void pair_core(F3 d, float r2, const Phys& P, float& W, F3& kg) {
    float inv_r = inv_sqrt(r2);
    float q2 = r2 * P.inv_h * P.inv_h;
    W = 0.f;
    if (q2 <= 1.f) {
        float t = 1.f - q2;
        W = P.a * t * t * t;
    }
    float r = r2 * inv_r;
    float u = __builtin_fmaf(-r, P.inv_h, 1.f);
    float kd = 0.f;
    if (P.inv_h <= inv_r) kd = P.ad * u * u;
    kg = mk3(0.f, 0.f, 0.f);
    if (r2 > 0.f) kg = mulsv(kd, mulvs(d, inv_r));
}

// pair_core for r2 in [2^-96, FLT_MAX], with the same results: the r2 > 0
// test always holds there, and t = max(1 - q2, 0) is 1 - q2 when q2 <= 1 and
// exactly 0 otherwise (q2 finite), so W = a t^3 needs no select.
__device__ __forceinline__ void pair_core_normal(F3 d, float r2, const Phys& P, float& W, F3& kg) {
    const float inv_r = inv_sqrt_normal(r2);
    const float q2 = r2 * P.inv_h * P.inv_h;
    const float t = __builtin_fmaxf(1.f - q2, 0.f);
    W = P.a * t * t * t;
    const float r = r2 * inv_r;
    const float u = __builtin_fmaf(-r, P.inv_h, 1.f);
    float kd = 0.f;
    if (P.inv_h <= inv_r) kd = P.ad * u * u;
    kg = mulsv(kd, mulvs(d, inv_r));
}

// W and grad W of the pair (self, pj).  A wave whose active lanes all have
// r2 in [2^-96, FLT_MAX] -- every pair but coincident particles and
// non-finite positions -- takes pair_core_normal.
__device__ __forceinline__ void pair_terms(F3 self, F3 pj, const Phys& P, float& W, F3& kg) {
    const F3 d = self - pj;
    const float r2 = __builtin_fmaf(d.z, d.z, __builtin_fmaf(d.y, d.y, d.x * d.x));
    if (__builtin_expect(__builtin_amdgcn_ballot_w64(!normal_r2(r2)) == 0ull, 1))
        pair_core_normal(d, r2, P, W, kg);
    else
        pair_core(d, r2, P, W, kg);
}

// pbf/solver/cuda/boundary/pbf_plane_boundary.cu:6-45
__device__ __forceinline__ void response_plane(bool& hit, F3& dp, F3 old_p, F3 pred, F3 Pp, F3 n) {
    F3 pa = old_p - Pp;
    F3 pb = pred - Pp;
    float pa_n = fdot(pa, n);
    float pb_n = fdot(pb, n);
    F3 I;
    if (pa_n == 0.f && pb_n == 0.f) {
        return;
    } else if (pa_n >= 0.f && pb_n < 0.f) {
        hit = true;
        float frac = fabsf(pa_n) / (fabsf(pa_n) + fabsf(pb_n));
        I = old_p + mulvs(pred - old_p, frac);
    } else if (pa_n < 0.f && pb_n < 0.f) {
        hit = true;
        float d = -fdot(Pp, n);
        float t = -(fdot(pred, n) + d) / fdot(n, n);
        I = pred + mulsv(t, n);
    } else {
        return;
    }
    F3 di = pred - I;
    float c = fdot(di, n);
    acc(dp, mulsv(-c, n));
}

// pbf/solver/cuda/boundary/pbf_sphere_boundary.cu:8-34 (inner), :36-64 (outer)
__device__ __forceinline__ void response_sphere(bool& hit, F3& dp, F3 pred, F3 center, float radius, bool inner) {
    F3 cd = center - pred;
    float d2 = fdot(cd, cd);
    float rinv = 1.f / __builtin_sqrtf(d2);
    bool coll = inner ? (d2 > radius * radius) : (d2 < radius * radius);
    if (!coll) return;
    hit = true;
    float nudge = inner ? 1.0001f : 0.9999f;
    dp.x = __builtin_fmaf(cd.x, nudge, dp.x);
    dp.y = __builtin_fmaf(cd.y, nudge, dp.y);
    dp.z = __builtin_fmaf(cd.z, nudge, dp.z);
    F3 tc = mulvs(cd, rinv);
    dp.x = __builtin_fmaf(-radius, tc.x, dp.x);
    dp.y = __builtin_fmaf(-radius, tc.y, dp.y);
    dp.z = __builtin_fmaf(-radius, tc.z, dp.z);
}

// pbf/solver/cuda/pbf_response_collision.cu:12-66: sweeps until none collides
// (capped).  Returns the sweeps used; `capped` when the cap stopped it.
__device__ __forceinline__ int collide(const Phys& P, const Boundary& B, F3& dp, F3 pk, bool& capped) {
    const F3 pred_x = pk, old_x = pk;
    int sweeps = 0;
    capped = false;
    while (true) {
        if (sweeps >= P.max_sweeps) {
            capped = true;
            break;
        }
        ++sweeps;
        F3 pp = pred_x + dp;
        bool hit = false;
        for (uint32_t i = 0; i < B.n_inner; ++i) {
            float4 s = B.inner[i];
            response_sphere(hit, dp, pp, mk3(s.x, s.y, s.z), s.w, true);
            pp = pred_x + dp;
        }
        for (uint32_t i = 0; i < B.n_outer; ++i) {
            float4 s = B.outer[i];
            response_sphere(hit, dp, pp, mk3(s.x, s.y, s.z), s.w, P.outer_as_inner != 0);
            pp = pred_x + dp;
        }
        for (uint32_t i = 0; i < B.n_planes; ++i) {
            F3 Pp = mk3(B.pts[3 * i], B.pts[3 * i + 1], B.pts[3 * i + 2]);
            F3 n = mk3(B.nrm[3 * i], B.nrm[3 * i + 1], B.nrm[3 * i + 2]);
            response_plane(hit, dp, old_x, pp, Pp, n);
            pp = pred_x + dp;
        }
        if (!hit) break;
    }
    return sweeps;
}

// The same for a boundary of exactly NP planes and no spheres (the box): the
// planes are loaded once, and the per-plane terms that do not change between
// sweeps -- (old_x - P).n, -(P.n) and n.n -- are computed once; each has the
// value the general form recomputes.
template <int NP>
__device__ __forceinline__ int collide_planes(const Phys& P, const Boundary& B, F3& dp, F3 pk, bool& capped) {
    F3 Pp[NP], n[NP];
    float pa_n[NP], dd[NP], nn[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        Pp[i] = mk3(B.pts[3 * i], B.pts[3 * i + 1], B.pts[3 * i + 2]);
        n[i] = mk3(B.nrm[3 * i], B.nrm[3 * i + 1], B.nrm[3 * i + 2]);
        pa_n[i] = fdot(pk - Pp[i], n[i]);
        dd[i] = -fdot(Pp[i], n[i]);
        nn[i] = fdot(n[i], n[i]);
    }
    int sweeps = 0;
    capped = false;
    while (true) {
        if (sweeps >= P.max_sweeps) {
            capped = true;
            break;
        }
        ++sweeps;
        bool hit = false;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            // response_plane with its constant terms hoisted
            const F3 pred = pk + dp;
            const float pb_n = fdot(pred - Pp[i], n[i]);
            F3 I;
            if (pa_n[i] == 0.f && pb_n == 0.f) {
                continue;
            } else if (pa_n[i] >= 0.f && pb_n < 0.f) {
                const float frac = fabsf(pa_n[i]) / (fabsf(pa_n[i]) + fabsf(pb_n));
                I = pk + mulvs(pred - pk, frac);
            } else if (pa_n[i] < 0.f && pb_n < 0.f) {
                const float t = -(fdot(pred, n[i]) + dd[i]) / nn[i];
                I = pred + mulsv(t, n[i]);
            } else {
                continue;
            }
            hit = true;
            const F3 di = pred - I;
            const float c = fdot(di, n[i]);
            acc(dp, mulsv(-c, n[i]));
        }
        if (!hit) break;
    }
    return sweeps;
}

// ------------------------------------------------------------ block helpers
__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Exclusive scan over the block (kBlock threads); `total` = block sum.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wsum, uint32_t& total) {
    const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    const uint32_t nw = blockDim.x >> 6;
    for (uint32_t w = 0; w < nw; ++w) {
        uint32_t s = wsum[w];
        if (w < wid) off += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return off + x - v;
}

// atomicAdd(&ctr[slot], 1) per active lane, where slot is where key k's
// counter is stored, returning a unique rank per lane (ranks among equal keys
// are arbitrary; the reorder makes them stable).  Runs of equal keys in
// consecutive lanes share one atomic, issued by the run's first lane; all
// runs' atomics are in flight together.  The active lanes must be a prefix of
// the wave.
__device__ __forceinline__ uint32_t wave_rank(uint32_t* ctr, uint32_t k, bool active, uint32_t slot) {
    const uint32_t lane = lane_id();
    const uint32_t nact = (uint32_t)__popcll(__ballot(active));
    const uint32_t kp = (uint32_t)__shfl_up((int)k, 1, 64);
    const bool head = active && (lane == 0 || kp != k);
    const uint64_t heads = __ballot(head);
    const uint64_t upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    uint32_t base = 0;
    if (head) {
        const uint64_t above = heads & ~upto;
        const uint32_t end = above ? (uint32_t)(__ffsll((unsigned long long)above) - 1) : nact;
        base = atomicAdd(&ctr[slot], end - lane);
    }
    const int hl = 63 - __clzll((unsigned long long)(heads & upto));
    base = (uint32_t)__shfl((int)base, hl < 0 ? 0 : hl, 64);
    return base + (lane - (uint32_t)(hl < 0 ? 0 : hl));
}

// wave_rank for keys in any lane order: all lanes holding one key share one
// atomic, issued by the key's lowest lane, which adds the key's lane count;
// a lane's rank is the number of lower lanes with its key.  The loop visits
// each distinct key of the wave once and touches no memory, so that the
// wave's atomics are all in flight together afterwards.
__device__ __forceinline__ uint32_t wave_rank_any(uint32_t* ctr, uint32_t k, bool active, uint32_t slot) {
    const uint32_t lane = lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    uint64_t todo = __ballot(active);
    uint32_t leader = lane, rank = 0, cnt = 0;
    while (todo) {
        const int l = __ffsll((unsigned long long)todo) - 1;
        const uint32_t kl = (uint32_t)__shfl((int)k, l, 64);
        const bool mine = active && k == kl;
        const uint64_t m = __ballot(mine);
        if (mine) {
            leader = (uint32_t)l;
            rank = (uint32_t)__popcll(m & below);
        }
        if (lane == (uint32_t)l) cnt = (uint32_t)__popcll(m);
        todo &= ~m;
    }
    uint32_t base = 0;
    if (active && leader == lane) base = atomicAdd(&ctr[slot], cnt);
    base = (uint32_t)__shfl((int)base, (int)leader, 64);
    return base + rank;
}

// Marks a scan tile as occupied: a read first, so that the common case (the
// tile is already marked) issues no store.
__device__ __forceinline__ void mark_tile(uint32_t* __restrict__ tflag, uint32_t t) {
    if (tflag[t] == 0u) tflag[t] = 1u;
}

__device__ __forceinline__ F3 f3of(float4 f) { return mk3(f.x, f.y, f.z); }
__device__ __forceinline__ F3 f3of(F3 f) { return f; }

// Whether the step sorts the segment of a cell of n particles.
__device__ __forceinline__ bool sorted_cell(uint32_t n) { return n > kSmallCell; }

// Lists cell k (its segment starting at o0) for the segment sort when
// sorted_cell: cells up to kSortMax at big[nbig[0]++], larger ones at
// big[huge_base(cap) + nbig[1]++]; bslot (when given) keeps that index at
// the cell's first slot.
__device__ __forceinline__ void list_big_cell(const uint32_t* __restrict__ offset, uint32_t k, uint32_t o0,
                                              uint32_t* __restrict__ big, uint32_t* __restrict__ nbig, uint32_t cap,
                                              uint32_t* __restrict__ bslot = nullptr) {
    const uint32_t n = offset[k + 1] - o0;
    uint32_t d = 0xFFFFFFFFu;
    if (n > kSortMax) d = huge_base(cap) + atomicAdd(&nbig[1], 1u);
    else if (n > kSmallCell) d = atomicAdd(&nbig[0], 1u);
    if (d != 0xFFFFFFFFu) {
        big[d] = k;
        if (bslot) bslot[o0] = d;
    }
}

// Block -> index range, contiguous per XCD (workgroup b runs on XCD b % 8),
// so that neighbouring slots share an L2.
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t nb = gridDim.x, b = blockIdx.x, nx = 8;
    if (nb < nx) return b;
    const uint32_t per = nb / nx, rem = nb % nx, x = b % nx;
    return x * per + min(x, rem) + b / nx;
}

}  // namespace
}  // namespace pbfx
