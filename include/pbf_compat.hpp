// pbf_compat.hpp -- the reference's PBF step interface, served by the pbfx C-ABI.
//
// For host code written against yodaplusplus/ssf_pbf_cuda (paths relative to
// cuda_pbf_with_ssf/src/): the types and entry points a simulator uses to run
// the step keep their names, fields and argument meaning --
//   pbf_phase_array, pbf_parameter, pbf_external, pbf_particle   (pbf/pbf_type.h:35-86)
//   pbf_buffer                                                    (pbf/pbf_type.h:88-100)
//   pbf_boundary                                                  (pbf/pbf_boundary.h:4-19)
//   pbf::one_step, with and without a boundary                    (pbf/solver/pbf_solve.h:57-70)
//   pbf::cuda::addParticle                                        (pbf/util/cuda/pbf_add.cu:12-28)
// so a caller such as pbf_sphere_sim::simulateOneStep
// (simulator/sphere/pbf_sphere_sim.cpp:384-389) recompiles against this header
// and links libpbfx.so instead of the CUDA objects.
//
// Differences a caller can see:
//  - dom_dim is a plain 12-byte {x, y, z} unless PBF_COMPAT_VEC3 names another
//    3-float type (glm::vec3 has the same layout); device arrays are HIP memory.
//  - pbf_particle::ns holds the pbfx context that replaces pbf::neighbor_search.
//  - pbf_buffer owns no arrays: the context holds all scratch; its allocate only
//    records max_pair_particle_num.
//  - Errors where the reference calls exit(-1) (parameters unset, capacity,
//    grid too large: pbf/pbf_type.cpp:42-44,155-158,167-170) print a message and
//    call exit(-1) as well; HIP errors, which the reference only checks in debug
//    builds, do the same here.
//  - Tensile correction, update and collision follow the reference as shipped:
//    REFERENCE s_corr, no XSPH or vorticity (pbf/solver/pbf_solve.cpp:125-145).
//    A particle's `options` (defaults from pbf_compat_options()) switch its
//    step to the full one (XSPH, vorticity, PAPER s_corr).
#ifndef PBF_COMPAT_HPP
#define PBF_COMPAT_HPP

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <utility>
#include <vector>

#include "pbfx.h"

#ifndef PBF_COMPAT_VEC3
struct pbf_compat_vec3 {
    float x, y, z;
};
struct pbf_compat_vec4 {
    float x, y, z, w;
};
#define PBF_COMPAT_VEC3 pbf_compat_vec3
#define PBF_COMPAT_VEC4 pbf_compat_vec4
#endif

typedef float scalar_t;
typedef PBF_COMPAT_VEC3 dom_dim;
typedef PBF_COMPAT_VEC4 dom_dim4;
static_assert(sizeof(dom_dim) == 3 * sizeof(float), "dom_dim must be three packed floats");
static_assert(sizeof(dom_dim4) == 4 * sizeof(float), "dom_dim4 must be four packed floats");

namespace pbf_compat_detail {
[[noreturn]] inline void fail(const char* what, pbfx_status st) {
    std::fprintf(stderr, "pbf: %s: %s\n", what, pbfx_status_string(st));
    std::exit(-1);
}
inline void check(const char* what, pbfx_status st) {
    if (st != PBFX_OK && st != PBFX_NEIGHBOR_TRUNCATED && st != PBFX_NOT_CONVERGED) fail(what, st);
}
inline void check_hip(const char* what, hipError_t e) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "pbf: %s: %s\n", what, hipGetErrorString(e));
        std::exit(-1);
    }
}
}  // namespace pbf_compat_detail

// Step options beyond the reference's fields; the defaults are the shipped step.
// Each pbf_particle carries its own copy (pbf_particle::options), taken from
// pbf_compat_options() when the particle is constructed, so simulators in one
// process can run different s_corr forms and passes (the reference keeps its
// s_corr constant in a function-static, pbf_solve.cpp:92).
struct pbf_compat_options_t {
    int32_t scorr_mode = PBFX_SCORR_REFERENCE;
    float scorr_k = 0.1f;
    int32_t scorr_n = 4;
    float scorr_dq_over_h = 0.2f;
    int32_t xsph_on = 0;
    int32_t vort_on = 0;
    int32_t max_sweeps = 32;
    hipStream_t stream = nullptr;
};
// The defaults new pbf_particle objects start from.
inline pbf_compat_options_t& pbf_compat_options() {
    static pbf_compat_options_t o;
    return o;
}

struct pbf_phase_array {
    uint32_t max_num = 0;
    uint32_t num = 0;
    uint32_t current_step = 0;
    dom_dim* x = nullptr;
    dom_dim* v = nullptr;
    void allocate(uint32_t n) {
        max_num = n;
        num = 0;
        current_step = 0;
        pbf_compat_detail::check_hip("allocate", hipMalloc((void**)&x, (n ? n : 1) * sizeof(dom_dim)));
        pbf_compat_detail::check_hip("allocate", hipMalloc((void**)&v, (n ? n : 1) * sizeof(dom_dim)));
        pbf_compat_detail::check_hip("allocate", hipMemset(x, 0, (n ? n : 1) * sizeof(dom_dim)));
        pbf_compat_detail::check_hip("allocate", hipMemset(v, 0, (n ? n : 1) * sizeof(dom_dim)));
    }
    void free() {
        max_num = 0;
        num = 0;
        if (x) (void)hipFree(x);
        if (v) (void)hipFree(v);
        x = v = nullptr;
    }
    void clear() {
        current_step = 0;
        num = 0;
    }
};

struct pbf_parameter {
    scalar_t stable_density = -1.f;    // kg/m^3
    scalar_t stable_distance = -1.f;   // m
    scalar_t particle_mass = -1.f;     // kg
    scalar_t smoothing_length = -1.f;  // m
    scalar_t time_step = -1.f;         // s
    scalar_t relaxation = 0.f;         // m^-2
    scalar_t xsph_parameter = 0.f;
    scalar_t vc_parameter = 0.f;
};

struct pbf_external {
    dom_dim body_force{0.f, 0.f, 0.f};  // m/s^2
};

struct pbf_particle {
    pbf_phase_array phase;
    std::shared_ptr<pbfx_ctx> ns;  // the pbfx context (replaces pbf::neighbor_search)
    pbf_parameter parameter;
    pbf_external external;
    pbf_compat_options_t options = pbf_compat_options();  // this simulatee's step options
    uint64_t boundary_seen = ~0ull;                        // version of the boundary its context holds
    scalar_t current_time() const { return static_cast<scalar_t>(phase.current_step) * parameter.time_step; }
    void allocate(uint32_t n, dom_dim domain_end);
    void free() {
        phase.free();
        ns.reset();
    }
};

// The reference's buffer holds the step's scratch arrays (pbf/pbf_type.cpp:80-101);
// here the context owns them and this records the neighbour slot count.
struct pbf_buffer {
    uint32_t max_pair_particle_num = 30;
    void allocate(uint32_t, uint32_t max_pair) { max_pair_particle_num = max_pair; }
    void free() {}
};

// Host copies of the boundary; the context receives them at the next step
// (the reference copies them to device arrays here, pbf/pbf_boundary.cpp:4-25).
struct pbf_boundary {
    std::vector<dom_dim4> inner_spheres;
    std::vector<dom_dim4> outer_spheres;
    std::vector<dom_dim> points_on_planes;
    std::vector<dom_dim> normals_on_planes;
    uint32_t inner_spheres_num = 0, outer_spheres_num = 0, planes_num = 0;
    uint64_t version = 0;
    void allocate(const std::vector<dom_dim4>& h_inner, const std::vector<dom_dim4>& h_outer,
                  const std::vector<dom_dim>& h_points, const std::vector<dom_dim>& h_normals) {
        inner_spheres = h_inner;
        outer_spheres = h_outer;
        points_on_planes = h_points;
        normals_on_planes = h_normals;
        inner_spheres_num = (uint32_t)h_inner.size();
        outer_spheres_num = (uint32_t)h_outer.size();
        planes_num = (uint32_t)(h_points.size() < h_normals.size() ? h_points.size() : h_normals.size());
        static std::atomic<uint64_t> counter{0};
        version = ++counter;
    }
    void free() {
        inner_spheres.clear();
        outer_spheres.clear();
        points_on_planes.clear();
        normals_on_planes.clear();
        inner_spheres_num = outer_spheres_num = planes_num = 0;
        version = 0;
    }
};

namespace pbf_compat_detail {
inline pbfx_params to_params(const pbf_particle& p, uint32_t max_pair) {
    pbfx_params q;
    pbfx_params_default(&q);
    q.stable_density = p.parameter.stable_density;
    q.stable_distance = p.parameter.stable_distance;
    q.particle_mass = p.parameter.particle_mass;
    q.smoothing_length = p.parameter.smoothing_length;
    q.time_step = p.parameter.time_step;
    q.relaxation = p.parameter.relaxation;
    q.xsph_parameter = p.parameter.xsph_parameter;
    q.vc_parameter = p.parameter.vc_parameter;
    q.body_force[0] = p.external.body_force.x;
    q.body_force[1] = p.external.body_force.y;
    q.body_force[2] = p.external.body_force.z;
    const pbf_compat_options_t& o = p.options;
    q.scorr_mode = o.scorr_mode;
    q.scorr_k = o.scorr_k;
    q.scorr_n = o.scorr_n;
    q.scorr_dq_over_h = o.scorr_dq_over_h;
    q.xsph_on = o.xsph_on;
    q.vort_on = o.vort_on;
    q.max_sweeps = o.max_sweeps;
    q.max_pair = (int32_t)max_pair;
    return q;
}
}  // namespace pbf_compat_detail

inline void pbf_particle::allocate(uint32_t n, dom_dim domain_end) {
    if (parameter.stable_density < 0.f || parameter.stable_distance < 0.f || parameter.particle_mass < 0.f ||
        parameter.smoothing_length < 0.f || parameter.time_step < 0.f) {
        std::fprintf(stderr, "pbf: parameters must be set before allocate\n");
        std::exit(-1);
    }
    phase.allocate(n);
    boundary_seen = ~0ull;
    const float end[3] = {domain_end.x, domain_end.y, domain_end.z};
    const pbfx_params q = pbf_compat_detail::to_params(*this, 30);
    pbfx_ctx* c = nullptr;
    pbf_compat_detail::check("allocate", pbfx_create(&q, n, end, &c));
    ns = std::shared_ptr<pbfx_ctx>(c, pbfx_destroy);
}

namespace pbf {

// pbf/solver/pbf_solve.cpp:182-216: one step with a boundary; x and v are
// left in the new cell-sorted order and phase.num shrinks by the particles
// whose predicted position left simulation_area.
inline void one_step(pbf_particle& simulatee, pbf_buffer& buffer, const pbf_boundary& boundary,
                     const std::pair<dom_dim, dom_dim> simulation_area, int num_solver_iteration) {
    pbfx_ctx* c = simulatee.ns.get();
    if (!c) pbf_compat_detail::fail("one_step", PBFX_INVALID_ARG);
    const pbfx_params q = pbf_compat_detail::to_params(simulatee, buffer.max_pair_particle_num);
    pbf_compat_detail::check("one_step", pbfx_set_params(c, &q));
    uint64_t& seen = simulatee.boundary_seen;
    if (seen != boundary.version) {
        pbf_compat_detail::check(
            "one_step",
            pbfx_set_boundary(c, reinterpret_cast<const float*>(boundary.inner_spheres.data()),
                              boundary.inner_spheres_num, reinterpret_cast<const float*>(boundary.outer_spheres.data()),
                              boundary.outer_spheres_num,
                              reinterpret_cast<const float*>(boundary.points_on_planes.data()),
                              reinterpret_cast<const float*>(boundary.normals_on_planes.data()), boundary.planes_num));
        seen = boundary.version;
    }
    const float lo[3] = {simulation_area.first.x, simulation_area.first.y, simulation_area.first.z};
    const float hi[3] = {simulation_area.second.x, simulation_area.second.y, simulation_area.second.z};
    uint32_t num = simulatee.phase.num;
    pbf_compat_detail::check("one_step",
                             pbfx_step(c, reinterpret_cast<float*>(simulatee.phase.x),
                                       reinterpret_cast<float*>(simulatee.phase.v), &num, lo, hi,
                                       num_solver_iteration, (void*)simulatee.options.stream));
    simulatee.phase.num = num;
}

// pbf/solver/pbf_solve.cpp:147-180: the overload without a boundary, whose
// collision is the fixed inner sphere centre (3,3,3), radius 3.
inline void one_step(pbf_particle& simulatee, pbf_buffer& buffer, const std::pair<dom_dim, dom_dim> simulation_area,
                     int num_solver_iteration) {
    pbfx_ctx* c = simulatee.ns.get();
    if (!c) pbf_compat_detail::fail("one_step", PBFX_INVALID_ARG);
    const pbfx_params q = pbf_compat_detail::to_params(simulatee, buffer.max_pair_particle_num);
    pbf_compat_detail::check("one_step", pbfx_set_params(c, &q));
    const float lo[3] = {simulation_area.first.x, simulation_area.first.y, simulation_area.first.z};
    const float hi[3] = {simulation_area.second.x, simulation_area.second.y, simulation_area.second.z};
    uint32_t num = simulatee.phase.num;
    pbf_compat_detail::check("one_step",
                             pbfx_step_no_boundary(c, reinterpret_cast<float*>(simulatee.phase.x),
                                                   reinterpret_cast<float*>(simulatee.phase.v), &num, lo, hi,
                                                   num_solver_iteration, (void*)simulatee.options.stream));
    simulatee.phase.num = num;
}

namespace cuda {

// pbf/util/cuda/pbf_add.cu:12-28: appends adding_num particles (device
// arrays) after the live ones and raises phase.num (no capacity check in the
// reference; here exceeding the capacity exits like its allocation errors).
inline void addParticle(pbf_particle& particle, const dom_dim* adding_position, const dom_dim* adding_velocity,
                        uint32_t adding_num) {
    pbfx_ctx* c = particle.ns.get();
    if (!c) pbf_compat_detail::fail("addParticle", PBFX_INVALID_ARG);
    uint32_t num = particle.phase.num;
    pbf_compat_detail::check("addParticle",
                             pbfx_add_particles(c, reinterpret_cast<float*>(particle.phase.x),
                                                reinterpret_cast<float*>(particle.phase.v), &num,
                                                reinterpret_cast<const float*>(adding_position),
                                                reinterpret_cast<const float*>(adding_velocity), adding_num,
                                                (void*)particle.options.stream));
    particle.phase.num = num;
}

}  // namespace cuda
}  // namespace pbf

#endif  // PBF_COMPAT_HPP
