// The drop-in from C++, the reference's own language, on the GPU: a caller that holds its
// state as the reference's VortexState members (koodit/GrossPitaevskii/VortexState.hpp:16-76)
// hands them over in memory (gpe_vortex_state_from_table, no file) and runs integrateInTime
// (GPUGrossPitaevskii_BCC.cu:218) through gpe_integrate_in_time.  Writes the final even and odd
// fields (raw float64, host order [k][j][i][l], re/im interleaved) and the state in the
// reference's file format, so that a checker can run the oracle on exactly the same inputs.
//
// usage: drop_in_gpu <output prefix>
// prints: dims X Y Z steps S
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "vaikkari_gpe.h"

namespace {

// The members a reference VortexState carries (VortexState.hpp:16-76), filled here with
// f(r, z) = r^kappa exp(-(r^2 + z^2)/2): this program's own data.
struct RefVortexState {
    unsigned kappa = 1;
    double g = 400.0;
    unsigned rsize = 60, zsize = 80;
    double rplug = 0.0, rstep = 0.1, zstep = 0.1, mu = 7.5;
    std::vector<double> f;
    RefVortexState() : f((size_t)rsize * zsize) {
        for (unsigned ir = 0; ir < rsize; ir++)
            for (unsigned iz = 0; iz < zsize; iz++) {
                const double r = ir * rstep, z = iz * zstep;
                f[(size_t)ir * zsize + iz] = std::pow(r, (double)kappa) * std::exp(-(r * r + z * z) / 2);
            }
    }
};

// What a maintainer of the reference would write in place of its integrateInTime: same
// arguments, the state passed in memory.
gpe_status integrateInTime(const RefVortexState &s, double block_scale, const double minp[3], const double maxp[3],
                           double iteration_period, unsigned number_of_iterations, gpe_reference_run *run,
                           std::vector<double> *even, std::vector<double> *odd, const char *save_state) {
    gpe_vortex_state *vs = nullptr;
    gpe_status st = gpe_vortex_state_from_table(s.kappa, s.g, s.rsize, s.zsize, s.rplug, s.rstep, s.zstep, s.mu,
                                                s.f.data(), &vs);
    if (st != GPE_OK) return st;
    uint32_t dims[3];
    double origin[3];
    gpe_reference_dims(block_scale, minp, maxp, dims, origin);
    const size_t n = (size_t)dims[0] * dims[1] * dims[2] * GPE_CLASSES;
    even->assign(2 * n, 0.0);
    odd->assign(2 * n, 0.0);
    st = gpe_integrate_in_time(vs, block_scale, minp, maxp, iteration_period, number_of_iterations, 0.7, GPE_FP64,
                               run, even->data(), odd->data());
    if (st == GPE_OK && save_state) st = gpe_vortex_state_save(vs, save_state);
    gpe_vortex_state_free(vs);
    return st;
}

bool write_raw(const std::string &path, const std::vector<double> &a) {
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) return false;
    const bool ok = std::fwrite(a.data(), sizeof(double), a.size(), fp) == a.size();
    return std::fclose(fp) == 0 && ok;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    const std::string prefix = argv[1];
    RefVortexState s;
    // extent as the reference's main derives it from the state (GPUGrossPitaevskii_BCC.cu:504-508)
    gpe_vortex_state *probe = nullptr;
    if (gpe_vortex_state_from_table(s.kappa, s.g, s.rsize, s.zsize, s.rplug, s.rstep, s.zstep, s.mu, s.f.data(),
                                    &probe) != GPE_OK)
        return 3;
    const double eps = 1e-5 * gpe_vortex_state_search_max(probe);
    const double mr = gpe_vortex_state_search_max_r(probe, eps), mz = gpe_vortex_state_search_max_z(probe, eps);
    gpe_vortex_state_free(probe);
    const double minp[3] = {-mr, -mr, -mz}, maxp[3] = {mr, mr, mz};
    gpe_reference_run run{};
    std::vector<double> even, odd;
    const gpe_status st =
        integrateInTime(s, 0.15, minp, maxp, 0.05, 2, &run, &even, &odd, (prefix + ".vs").c_str());
    if (st != GPE_OK) {
        std::fprintf(stderr, "integrateInTime: %s: %s\n", gpe_status_string(st), gpe_last_error());
        return 1;
    }
    if (!write_raw(prefix + ".even", even) || !write_raw(prefix + ".odd", odd)) return 4;
    std::printf("dims %u %u %u steps %u\n", run.dims[0], run.dims[1], run.dims[2], run.steps_per_iteration);
    return 0;
}
