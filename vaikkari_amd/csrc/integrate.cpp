// integrate.cpp -- gpe_integrate_in_time: the drop-in for the reference's
// single-device solver integrateInTime
// (koodit/GrossPitaevskii/GPUGrossPitaevskii_BCC.cu:218-499), with or
// without its SAVE_PICTURE / SAVE_VOLUME output (lines 424-468).
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "gpe_internal.hpp"
#include "../../include/vaikkari_gpe_output.h"

namespace {

// The picture and the volume map of the current even field, named by the
// iteration as the reference names them (GPUGrossPitaevskii_BCC.cu:440,466).
gpe_status write_output(gpe_solver *s, const gpe_output_options &o, const gpe_reference_run &r, double block_scale,
                        double fmax, uint32_t iter) {
    const std::string prefix = o.prefix ? o.prefix : "";
    const uint32_t wx = r.dims[0] + 2, wy = r.dims[1] + 2, wz = r.dims[2] + 2;
    if (o.save_picture) {
        std::vector<uint8_t> px((size_t)3 * wx * wy);
        gpe_status st = gpe_slice_picture(s, r.dims[2] / 2 + 1, 5.0, 0, px.data());
        if (st != GPE_OK) return st;
        st = gpe_write_bmp((prefix + "kuva" + std::to_string(iter) + ".bmp").c_str(), wx, wy, px.data());
        if (st != GPE_OK) return st;
    }
    if (o.save_volume) {
        std::vector<uint16_t> vol((size_t)wx * wy * wz);
        const double unit = 60000.0 / (GPE_CLASSES * fmax * fmax);
        gpe_status st = gpe_volume_map(s, unit, o.volume_flags, vol.data());
        if (st != GPE_OK) return st;
        const uint32_t dims[3] = {wx, wy, wz};
        const double h = block_scale * std::sqrt(8.0);
        const double spacing[3] = {h, h, h};
        st = gpe_write_volume_map((prefix + "volume" + std::to_string(iter) + ".mhd").c_str(), vol.data(), dims,
                                  spacing);
        if (st != GPE_OK) return st;
    }
    return GPE_OK;
}

gpe_status integrate(const gpe_vortex_state *vs, double block_scale, const double minp[3], const double maxp[3],
                     double iteration_period, uint32_t number_of_iterations, double trap_ratio, uint32_t precision,
                     const gpe_output_options *out, gpe_reference_run *run, double *even_out, double *odd_out) {
    gpe_reference_run r{};
    gpe_status st = gpe_reference_setup(vs, block_scale, minp, maxp, iteration_period, trap_ratio, 54363u, &r,
                                        nullptr, nullptr, nullptr);
    if (st != GPE_OK) return st;
    const size_t n = (size_t)r.dims[0] * r.dims[1] * r.dims[2] * GPE_CLASSES;
    std::vector<double> even(2 * n), odd(2 * n);
    // Same setup again, now with the fields (the reference seeds its noise
    // with 54363, GPUGrossPitaevskii_BCC.cu:326).
    st = gpe_reference_setup(vs, block_scale, minp, maxp, iteration_period, trap_ratio, 54363u, &r, even.data(),
                             odd.data(), nullptr);
    if (st != GPE_OK) return st;
    gpe_config cfg{};
    cfg.nx = r.dims[0];
    cfg.ny = r.dims[1];
    cfg.nz = r.dims[2];
    cfg.precision = precision;
    cfg.block_scale = block_scale;
    cfg.dt = r.dt;
    cfg.g = r.g;
    for (int a = 0; a < 3; a++) cfg.origin[a] = r.origin[a];
    // potentialV3 evaluated in-kernel: identical bits to the reference's
    // streamed pot array (GPUGrossPitaevskii_BCC.cu:260,307) without its 8 B
    // per node-update of HBM traffic.
    cfg.potential = GPE_POT_TRAP;
    cfg.trap_ratio = trap_ratio;
    cfg.mode = GPE_REAL_TIME;
    cfg.kernel = GPE_KERNEL_AUTO;
    gpe_solver *s = nullptr;
    st = gpe_create(&cfg, &s);
    if (st != GPE_OK) return st;
    st = gpe_set_state(s, even.data(), odd.data());
    const bool draw = out && (out->save_picture || out->save_volume);
    const double fmax = gpe_vortex_state_search_max(vs);
    // The reference's loop (GPUGrossPitaevskii_BCC.cu:421-482): output of the
    // current field, then one iteration of steps, N times, then the output
    // once more.
    for (uint32_t it = 0; st == GPE_OK; it++) {
        if (draw) st = write_output(s, *out, r, block_scale, fmax, it);
        if (st != GPE_OK || it == number_of_iterations) break;
        st = gpe_step(s, r.steps_per_iteration);
    }
    if (st == GPE_OK) st = gpe_synchronize(s);
    if (st == GPE_OK && (even_out || odd_out)) st = gpe_get_state(s, even_out, odd_out);
    gpe_destroy(s);
    if (run) *run = r;
    return st;
}

}  // namespace

extern "C" gpe_status gpe_integrate_in_time(const gpe_vortex_state *vs, double block_scale, const double minp[3],
                                            const double maxp[3], double iteration_period,
                                            uint32_t number_of_iterations, double trap_ratio, uint32_t precision,
                                            gpe_reference_run *run, double *even_out, double *odd_out) {
    return integrate(vs, block_scale, minp, maxp, iteration_period, number_of_iterations, trap_ratio, precision,
                     nullptr, run, even_out, odd_out);
}

extern "C" gpe_status gpe_integrate_in_time_output(const gpe_vortex_state *vs, double block_scale,
                                                   const double minp[3], const double maxp[3],
                                                   double iteration_period, uint32_t number_of_iterations,
                                                   double trap_ratio, uint32_t precision,
                                                   const gpe_output_options *opt, gpe_reference_run *run) {
    if (!opt) {
        gpe::set_error("gpe_integrate_in_time_output: null options");
        return GPE_EINVAL;
    }
    return integrate(vs, block_scale, minp, maxp, iteration_period, number_of_iterations, trap_ratio, precision,
                     opt, run, nullptr, nullptr);
}
