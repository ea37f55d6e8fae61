// vortex_state.hpp -- the reference's VortexState: a stationary vortex
// f(r,z) e^{i kappa phi - i mu t} tabulated on a cylindrical (r, z) grid
// (koodit/GrossPitaevskii/VortexState.hpp:16-76).  Host-only, FP64.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

namespace gpe {

// ---------------------------------------------------------------------
// VortexState: f(r,z) on a cylindrical grid and its evaluation
// (koodit/GrossPitaevskii/VortexState.cpp).
// ---------------------------------------------------------------------
struct VortexState {
    uint32_t kappa = 0;
    double g = 0;
    uint32_t rsize = 0, zsize = 0;
    double rplug = 0, rstep = 0, zstep = 0, mu = 0;
    std::vector<double> f;

    // VortexState::getFunction, VortexState.cpp:220-238: bilinear in (r, |z|),
    // zero outside the table.
    double function(double r, double z) const {
        const double rr = (r - rplug) / rstep;
        if (rr < 0.0) return 0.0;
        const uint32_t ir = (uint32_t)rr;
        if (ir >= rsize) return 0.0;
        const double dr = rr - (double)ir;
        const double zz = std::fabs(z) / zstep;
        const uint32_t iz = (uint32_t)zz;
        if (iz >= zsize) return 0.0;
        const double dz = zz - (double)iz;
        const uint32_t ii = ir * zsize + iz;
        const bool rin = ir + 1 < rsize, zin = iz + 1 < zsize;
        const double f00 = f[ii];
        const double f10 = rin ? f[ii + zsize] : 0.0;
        const double f01 = zin ? f[ii + 1] : 0.0;
        const double f11 = (rin && zin) ? f[ii + zsize + 1] : 0.0;
        return (1.0 - dz) * ((1.0 - dr) * f00 + dr * f10) + dz * ((1.0 - dr) * f01 + dr * f11);
    }
    // VortexState::getPsi(Vector3), VortexState.cpp:204-212:
    // f(r,z) * ((x + i y)/r)^kappa, the power taken by repeated complex
    // multiplication in the order of Complex::operator*= (Complex.hpp:57-63).
    void psi(double x, double y, double z, double out[2]) const {
        const double r = std::sqrt(x * x + y * y);
        if (r < 1e-30) {
            out[0] = out[1] = 0.0;
            return;
        }
        double re = function(r, z), im = 0.0;
        const double ur = x / r, ui = y / r;
        for (uint32_t n = 0; n < kappa; n++) {
            const double t = re;
            re = t * ur - im * ui;
            im = im * ur + t * ui;
        }
        out[0] = re;
        out[1] = im;
    }
};

}  // namespace gpe

struct gpe_vortex_state {
    gpe::VortexState vs;
    uint32_t threads = 0;  // relaxation threads: 0 = automatic, 1 = serial
};
