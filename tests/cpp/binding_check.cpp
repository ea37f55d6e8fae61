// A C++ caller of the C ABI, as the reference's host code (C++) would bind it:
// every public header in one translation unit, linked against
// libvaikkari_gpe.so, calling host-side functions only (no GPU needed).
// Exit status 0 when every check holds; a failed check prints what failed.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "vaikkari_gpe.h"
#include "vaikkari_gpe_batch.h"
#include "vaikkari_gpe_diag.h"
#include "vaikkari_gpe_output.h"
#include "vaikkari_gpe_primal.h"
#include "vaikkari_gpe_recipes.h"
#include "vaikkari_gpe_serve.h"

static int failures = 0;
static void check(bool ok, const char *what) {
    if (!ok) {
        std::fprintf(stderr, "FAILED: %s\n", what);
        failures++;
    }
}

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    check(std::strlen(gpe_status_string(GPE_EINVAL)) > 0, "status string");

    // stencil: the kernels' table and the one derived from the positions alone hold the same
    // 48 (offset, class) entries per class
    int32_t table[192], derived[192];
    gpe_bcc_stencil(table);
    check(gpe_bcc_stencil_derived(derived) == 48, "48 derived neighbours");
    for (int l = 0; l < 12; l++)
        for (int f = 0; f < 4; f++) {
            const int32_t *e = table + 4 * (4 * l + f);
            bool found = false;
            for (int g = 0; g < 4 && !found; g++)
                found = std::memcmp(e, derived + 4 * (4 * l + g), 4 * sizeof(int32_t)) == 0;
            check(found, "kernel neighbour present in the derived set");
        }

    // integrateInTime's grid rule: ceil(extent / cube edge) cells, box centred on the range
    const double b = 0.05, a = std::sqrt(8.0) * b;
    const double lo[3] = {-1.0, -1.0, -2.0}, hi[3] = {1.0, 1.0, 2.0};
    uint32_t dims[3];
    double origin[3];
    gpe_reference_dims(b, lo, hi, dims, origin);
    for (int d = 0; d < 3; d++) {
        check(dims[d] == (uint32_t)std::ceil((hi[d] - lo[d]) / a), "cells per axis");
        check(std::fabs(origin[d] + 0.5 * dims[d] * a - 0.5 * (lo[d] + hi[d])) < 1e-12, "centred origin");
    }

    // VortexState: build a start shape, save it in the reference's format, load it back
    gpe_vortex_state *vs = nullptr, *back = nullptr;
    check(gpe_vortex_state_create(&vs) == GPE_OK, "create");
    gpe_vortex_state_set_kappa(vs, 1);
    gpe_vortex_state_set_g(vs, 250.0);
    check(gpe_vortex_state_set_dimensions(vs, 40, 0.1, 0.5, 30, 0.1) == GPE_OK, "set_dimensions");
    check(gpe_vortex_state_save(vs, argv[1]) == GPE_OK, "save");
    check(gpe_vortex_state_load(argv[1], &back) == GPE_OK, "load");
    double p0[8], p1[8], s0[2], s1[2];
    gpe_vortex_state_params(vs, p0);
    gpe_vortex_state_params(back, p1);
    check(std::memcmp(p0, p1, sizeof p0) == 0, "parameters survive save/load");
    gpe_vortex_state_psi(vs, 0.7, -0.3, 0.4, s0);
    gpe_vortex_state_psi(back, 0.7, -0.3, 0.4, s1);
    check(s0[0] == s1[0] && s0[1] == s1[1], "psi survives save/load");
    gpe_vortex_state_free(back);

    // the same state handed over in memory (no file): identical parameters and psi
    double ftab[40 * 30];
    check(gpe_vortex_state_table(vs, ftab) == GPE_OK, "table");
    gpe_vortex_state *mem = nullptr;
    check(gpe_vortex_state_from_table((uint32_t)p0[0], p0[1], (uint32_t)p0[2], (uint32_t)p0[3], p0[4], p0[5], p0[6],
                                      p0[7], ftab, &mem) == GPE_OK,
          "from_table");
    double p2[8], s2[2];
    gpe_vortex_state_params(mem, p2);
    gpe_vortex_state_psi(mem, 0.7, -0.3, 0.4, s2);
    check(std::memcmp(p0, p2, sizeof p0) == 0, "parameters survive the in-memory handoff");
    check(s0[0] == s2[0] && s0[1] == s2[1], "psi survives the in-memory handoff");
    gpe_vortex_state_free(mem);
    gpe_vortex_state_free(vs);

    std::printf("dims %u %u %u\n", dims[0], dims[1], dims[2]);
    return failures ? 1 : 0;
}
