// integrate.cpp -- gpe_integrate_in_time: the drop-in for the reference's
// single-device solver integrateInTime
// (koodit/GrossPitaevskii/GPUGrossPitaevskii_BCC.cu:218-499).
#include <vector>

#include "gpe_internal.hpp"

extern "C" gpe_status gpe_integrate_in_time(const gpe_vortex_state *vs, double block_scale, const double minp[3],
                                            const double maxp[3], double iteration_period,
                                            uint32_t number_of_iterations, double trap_ratio, uint32_t precision,
                                            gpe_reference_run *run, double *even_out, double *odd_out) {
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
    for (uint32_t it = 0; it < number_of_iterations && st == GPE_OK; it++) st = gpe_step(s, r.steps_per_iteration);
    if (st == GPE_OK) st = gpe_synchronize(s);
    if (st == GPE_OK && (even_out || odd_out)) st = gpe_get_state(s, even_out, odd_out);
    gpe_destroy(s);
    if (run) *run = r;
    return st;
}
