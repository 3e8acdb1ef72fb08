// compat_step -- runs pbf::one_step through include/pbf_compat.hpp, the
// reference-shaped C++ interface, for the GPU parity tests.
//
// usage: compat_step <input.bin> <output.bin>
// input:  u32 n, steps, iters, n_planes; f32 stable_density, stable_distance,
//         particle_mass, smoothing_length, time_step, relaxation,
//         xsph_parameter, vc_parameter, body_force[3], domain_end[3],
//         area_lo[3], area_hi[3], plane_points[3*n_planes],
//         plane_normals[3*n_planes], x[3n], v[3n]
// output: u32 num, f32 x[3*num], v[3*num] (sorted order, as the reference leaves them)
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

    pbf_particle p;
    p.parameter.stable_density = par[0];
    p.parameter.stable_distance = par[1];
    p.parameter.particle_mass = par[2];
    p.parameter.smoothing_length = par[3];
    p.parameter.time_step = par[4];
    p.parameter.relaxation = par[5];
    p.parameter.xsph_parameter = par[6];
    p.parameter.vc_parameter = par[7];
    p.external.body_force = dom_dim{g[0], g[1], g[2]};
    p.allocate(n, dom_dim{end[0], end[1], end[2]});
    pbf_buffer buf;
    buf.allocate(n, 30);
    pbf_boundary bnd;
    bnd.allocate({}, {}, pts, nrm);
    if (hipMemcpy(p.phase.x, x.data(), n * sizeof(dom_dim), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p.phase.v, v.data(), n * sizeof(dom_dim), hipMemcpyHostToDevice) != hipSuccess)
        return 3;
    p.phase.num = n;
    const std::pair<dom_dim, dom_dim> area{dom_dim{lo[0], lo[1], lo[2]}, dom_dim{hi[0], hi[1], hi[2]}};
    for (uint32_t s = 0; s < steps; ++s) pbf::one_step(p, buf, bnd, area, (int)iters);

    const uint32_t num = p.phase.num;
    x.resize(num);
    v.resize(num);
    if (hipMemcpy(x.data(), p.phase.x, num * sizeof(dom_dim), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(v.data(), p.phase.v, num * sizeof(dom_dim), hipMemcpyDeviceToHost) != hipSuccess)
        return 3;
    std::FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    std::fwrite(&num, sizeof(num), 1, o);
    std::fwrite(x.data(), sizeof(dom_dim), num, o);
    std::fwrite(v.data(), sizeof(dom_dim), num, o);
    std::fclose(o);
    p.free();
    buf.free();
    bnd.free();
    return 0;
}
