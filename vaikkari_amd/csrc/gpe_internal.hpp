// gpe_internal.hpp -- shared host-side internals of libvaikkari_gpe.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/vaikkari_gpe.h"
#include "bcc_stencil.hpp"

namespace gpe {

// Thread-local last-error text behind gpe_last_error().
void set_error(const char *fmt, ...);

// Class positions q_l: node l sits at b*q_l/sqrt(8) inside its cell; every
// component is odd, so nodes never lie on cell faces
// (GPUGrossPitaevskii_BCC.cu:73-89).
extern const int kClassPos[GPE_CLASSES][3];

// Hodge weight of every dual edge: dual face area sqrt(8) b^2 over
// (primal edge b x node volume sqrt(32/9) b^3), times b^2
// (GPUGrossPitaevskii_BCC.cu:141).
constexpr double kHodge = 1.5;
// Node measure w = kVolume * b^3 = a^3/12 (GPUGrossPitaevskii_BCC.cu:71).
double node_measure(double block_scale);

// Position of node (i,j,k,l), written with the reference's operation order
// (GPUGrossPitaevskii_BCC.cu:258) so that device code evaluating the same
// expression produces identical bits.
void node_position(const double origin[3], double b, uint32_t i, uint32_t j, uint32_t k, int l, double p[3]);

// Largest trap potential over the nodes of a box (attained at a corner cell;
// node positions lie strictly inside the box).
double trap_vmax(const uint32_t dims[3], const double origin[3], double b, double ratio);

// Linear stability bound of both time integrators: the leapfrog
// (psi(t+dt/2) = psi(t-dt/2) - i dt H psi(t)) and the imaginary-time Euler
// step (psi <- (1 - dt H) psi) need dt * lambda_max(H) < 2, and
// lambda_max <= 6/b^2 + max V (Gershgorin on -L/2 = diagonal 3/b^2, four
// off-diagonals 0.75/b^2).  The nonlinear term adds g|psi|^2 on top.
bool linear_step_stable(double dt, double b, double vmax);

// potentialV3 (GPUGrossPitaevskii_BCC.cu:33-36) with RATIO -> ratio.
inline double trap_potential(const double p[3], double ratio) {
    return 0.5 * (p[0] * p[0] + p[1] * p[1] + ratio * ratio * p[2] * p[2]);
}

}  // namespace gpe
