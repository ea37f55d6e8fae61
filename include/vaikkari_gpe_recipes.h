/*
 * vaikkari_gpe_recipes.h -- host-side initial states and grid helpers of the
 * benchmark workloads.  The reference builds its only initial state from a
 * VortexState (GPUGrossPitaevskii_BCC.cu:245-348, exposed here as
 * gpe_reference_setup); the recipes below are this library's own
 * definitions for the Gaussian, Thomas-Fermi and vortex workloads and have
 * no counterpart in the reference.
 *
 * All arrays use the host order of vaikkari_gpe.h: [k][j][i][l], class
 * fastest, complex interleaved.  Node positions follow
 * GPUGrossPitaevskii_BCC.cu:258: p = origin + b*(i*sqrt(8) + q_l/sqrt(8)).
 */
#ifndef VAIKKARI_GPE_RECIPES_H
#define VAIKKARI_GPE_RECIPES_H

#include "vaikkari_gpe.h"

#ifdef __cplusplus
extern "C" {
#endif

/* origin of a cell box centred on 0: -dims*sqrt(8)*b/2 per axis. */
void gpe_centered_origin(const uint32_t dims[3], double block_scale, double origin[3]);
/* 3*nodes doubles (x, y, z per node). */
void gpe_node_positions(const uint32_t dims[3], const double origin[3], double block_scale, double *out);
/* Unscaled trap potential 0.5*(x^2 + y^2 + ratio^2 z^2) at every node. */
void gpe_trap_potential(const uint32_t dims[3], const double origin[3], double block_scale, double ratio,
                        double *V);
/* w * sum |psi|^2 with w = sqrt(32/9) b^3 (compensated). */
double gpe_field_norm(const uint32_t dims[3], double block_scale, const double *psi);

/* exp(-r^2/(2 sigma^2)) normalised to w*sum|psi|^2 = 1, then multiplied by
 * (1 + noise*(xr + i xi)) with xr, xi ~ N(0,1) drawn per node in host order
 * from mt19937_64(seed) by Box-Muller. */
gpe_status gpe_recipe_gaussian(const uint32_t dims[3], const double origin[3], double block_scale, double sigma,
                               double noise, uint64_t seed, double *psi);
/* sqrt(max(mu - V, 0)/g) for the trap with the given ratio, mu found by
 * bisection so that the discrete norm is 1.  mu_out may be NULL. */
gpe_status gpe_recipe_thomas_fermi(const uint32_t dims[3], const double origin[3], double block_scale, double g,
                                   double ratio, double *psi, double *mu_out);
/* psi <- |psi| ((x-x0) + i(y-y0))/rho: one straight vortex line along z. */
gpe_status gpe_recipe_vortex_line(const uint32_t dims[3], const double origin[3], double block_scale, double x0,
                                  double y0, double *psi);
/* psi <- |psi| ((rho'-R) + i z')/sqrt((rho'-R)^2 + z'^2) in ring-local
 * coordinates (z' along axis, rho' distance from the axis through center). */
gpe_status gpe_recipe_vortex_ring(const uint32_t dims[3], const double origin[3], double block_scale,
                                  const double center[3], const double axis[3], double radius, double *psi);
/* odd = even * exp(+i mu dt/2) (GPUGrossPitaevskii_BCC.cu:325,342-343). */
gpe_status gpe_recipe_phase_shift(const uint32_t dims[3], const double *even, double mu, double dt, double *odd);

#ifdef __cplusplus
}
#endif
#endif /* VAIKKARI_GPE_RECIPES_H */
