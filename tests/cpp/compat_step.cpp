// compat_step -- runs pbf::one_step through include/pbf_compat.hpp, the
// reference-shaped C++ interface, for the GPU parity tests.
//
// usage: compat_step <input.bin> <output.bin>
// input:  u32 n, steps, iters, n_planes; f32 stable_density, stable_distance,
//         particle_mass, smoothing_length, time_step, relaxation,
//         xsph_parameter, vc_parameter, body_force[3], domain_end[3],
//         area_lo[3], area_hi[3], plane_points[3*n_planes],
//         plane_normals[3*n_planes], x[3n], v[3n]
// output: twice (u32 num, f32 x[3*num], v[3*num]) (sorted order, as the
//         reference leaves them): first the step as shipped, then the full
//         step (PAPER s_corr k = 0.1, n = 4, dq = 0.2h, XSPH, vorticity).
// The two simulatees live in one process and are stepped in turn, each with
// its own options.
#include <cstdio>
#include <vector>

#include "pbf_compat.hpp"

namespace {
template <class T>
bool rd(std::FILE* f, T* p, size_t n) {
    return std::fread(p, sizeof(T), n, f) == n;
}
}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <input.bin> <output.bin>\n", argv[0]);
        return 2;
    }
    std::FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    uint32_t hdr[4];
    float par[8], g[3], end[3], lo[3], hi[3];
    if (!rd(f, hdr, 4) || !rd(f, par, 8) || !rd(f, g, 3) || !rd(f, end, 3) || !rd(f, lo, 3) || !rd(f, hi, 3))
        return 2;
    const uint32_t n = hdr[0], steps = hdr[1], iters = hdr[2], np = hdr[3];
    std::vector<dom_dim> pts(np), nrm(np), x(n), v(n);
    if (!rd(f, (float*)pts.data(), 3 * np) || !rd(f, (float*)nrm.data(), 3 * np) || !rd(f, (float*)x.data(), 3 * n) ||
        !rd(f, (float*)v.data(), 3 * n))
        return 2;
    std::fclose(f);

    auto make = [&](pbf_particle& p, bool full) {
        p.parameter.stable_density = par[0];
        p.parameter.stable_distance = par[1];
        p.parameter.particle_mass = par[2];
        p.parameter.smoothing_length = par[3];
        p.parameter.time_step = par[4];
        p.parameter.relaxation = par[5];
        p.parameter.xsph_parameter = par[6];
        p.parameter.vc_parameter = par[7];
        p.external.body_force = dom_dim{g[0], g[1], g[2]};
        if (full) {
            p.options.scorr_mode = PBFX_SCORR_PAPER;
            p.options.scorr_k = 0.1f;
            p.options.scorr_n = 4;
            p.options.scorr_dq_over_h = 0.2f;
            p.options.xsph_on = 1;
            p.options.vort_on = 1;
        }
        p.allocate(n, dom_dim{end[0], end[1], end[2]});
        return hipMemcpy(p.phase.x, x.data(), n * sizeof(dom_dim), hipMemcpyHostToDevice) == hipSuccess &&
               hipMemcpy(p.phase.v, v.data(), n * sizeof(dom_dim), hipMemcpyHostToDevice) == hipSuccess;
    };
    pbf_particle shipped, full;
    if (!make(shipped, false) || !make(full, true)) return 3;
    shipped.phase.num = full.phase.num = n;
    pbf_buffer buf;
    buf.allocate(n, 30);
    pbf_boundary bnd;
    bnd.allocate({}, {}, pts, nrm);
    const std::pair<dom_dim, dom_dim> area{dom_dim{lo[0], lo[1], lo[2]}, dom_dim{hi[0], hi[1], hi[2]}};
    for (uint32_t s = 0; s < steps; ++s) {
        pbf::one_step(shipped, buf, bnd, area, (int)iters);
        pbf::one_step(full, buf, bnd, area, (int)iters);
    }

    std::FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    for (pbf_particle* p : {&shipped, &full}) {
        const uint32_t num = p->phase.num;
        std::vector<dom_dim> xo(num), vo(num);
        if (hipMemcpy(xo.data(), p->phase.x, num * sizeof(dom_dim), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(vo.data(), p->phase.v, num * sizeof(dom_dim), hipMemcpyDeviceToHost) != hipSuccess)
            return 3;
        std::fwrite(&num, sizeof(num), 1, o);
        std::fwrite(xo.data(), sizeof(dom_dim), num, o);
        std::fwrite(vo.data(), sizeof(dom_dim), num, o);
        p->free();
    }
    std::fclose(o);
    buf.free();
    bnd.free();
    return 0;
}
