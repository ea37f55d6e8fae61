/*
 * vaikkari_gpe_primal.h -- Gross-Pitaevskii stepper with the 14-point
 * primal BCC operator.
 *
 * No reference stepper uses this discretisation (the reference's steppers
 * put psi on the 12 dual nodes of every cubic cell, vaikkari_gpe.h).  It is
 * the DEC Laplacian of 0-forms that the reference's mesh library builds on
 * a BCC grid (DelaunayMesh::createBccGrid, jukan_koodit/Mesh/
 * DelaunayMesh.cpp:178, Hodge weights Mesh::getEdgeHodge / getNodeHodge,
 * Mesh/Mesh.cpp:1479-1489): psi lives on the 2 N^3 BCC sites of N^3 cubic
 * cells of edge a -- the cube corners and the cube centres -- and each site
 * couples to its 8 diagonal neighbours (distance sqrt(3)a/2, weight 3a/8) and
 * its 6 axial neighbours (distance a, weight a/8) with dual volume a^3/2:
 *   (L psi)_i = (1/a^2) [ 3/4 sum_8 (psi_j - psi_i) + 1/4 sum_6 (psi_j - psi_i) ],
 *   H = -L/2 + V + g|psi|^2,  V = 0.5(x^2 + y^2 + ratio^2 z^2) (in-kernel).
 * Same integrators as the dual stepper: leapfrog (real time) and Euler with
 * FP64 renormalisation (imaginary time); psi = 0 outside the box.
 *
 * Host array order: [k][j][i][s], s = 0 corner site at origin + a(i, j, k),
 * s = 1 centre site at origin + a(i + 1/2, j + 1/2, k + 1/2); complex
 * interleaved (re, im).  2 nx ny nz complex values per field.
 */
#ifndef VAIKKARI_GPE_PRIMAL_H
#define VAIKKARI_GPE_PRIMAL_H

#include "vaikkari_gpe.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gpe_primal gpe_primal;

typedef struct gpe_primal_config {
    uint32_t nx, ny, nz;   /* cubic cells (2 nx ny nz sites) */
    uint32_t precision;    /* gpe_precision */
    double a;              /* cube edge */
    double dt;             /* time step (real) or dtau (imaginary), unscaled */
    double g;
    double origin[3];      /* corner site (0,0,0); non-finite: box centred on 0 */
    double trap_ratio;
    uint32_t mode;         /* gpe_mode */
    uint32_t tile;         /* 0: default for the precision; 1..gpe_primal_tiles() */
} gpe_primal_config;

/* GPE_EUNSTABLE when dt (7.5/a^2 + max V) >= 2, the leapfrog's linear
 * stability bound for this operator (Gershgorin: -L/2 has spectrum in
 * [0, 7.5/a^2]). */
gpe_status gpe_primal_create(const gpe_primal_config *cfg, gpe_primal **out);
void gpe_primal_destroy(gpe_primal *p);
/* even, odd: 2 * 2 nx ny nz doubles; odd may be NULL in imaginary time. */
gpe_status gpe_primal_set_state(gpe_primal *p, const double *even, const double *odd);
gpe_status gpe_primal_get_state(gpe_primal *p, double *even, double *odd);
gpe_status gpe_primal_set_mode(gpe_primal *p, uint32_t mode, double dt);
/* odd = even exp(i mu dt/2), the leapfrog start of the reference
 * (GPUGrossPitaevskii_BCC.cu:325,342-343). */
gpe_status gpe_primal_init_odd(gpe_primal *p, double mu);
gpe_status gpe_primal_step(gpe_primal *p, uint64_t nsteps);
/* As gpe_primal_step, timed with HIP events; launches = hot-kernel launches. */
gpe_status gpe_primal_time_steps(gpe_primal *p, uint64_t nsteps, double *elapsed_ms, uint64_t *launches);
/* {N, K, P, I} of field 0 (even) or 1 (odd), FP64, measure a^3/2 per site:
 * N = w sum|psi|^2, K = w sum 1/2 Re(conj(psi)(-L psi)), P = w sum V|psi|^2,
 * I = w sum g|psi|^4. */
gpe_status gpe_primal_observables(gpe_primal *p, int field, double out[4]);
gpe_status gpe_primal_synchronize(gpe_primal *p);
/* Number of tile shapes: one layer per block or a z-march of 4 to 16
 * layers, waves per block (2, 4, 8), block order over the XCDs, streaming
 * read-modify-write, and how a lane gets its x-neighbours (DPP lane shift,
 * cross-lane shuffle or a second load).
 *   1-20   one layer per block, or z-marches on 64-bit addresses;
 *   21-25  both half-updates of a real-time step in one out-of-place pass
 *          (temporal blocking, two more field buffers);
 *   26-65  z-marches that address the fields through buffer resources
 *          (32-bit offsets: fields below 2 GiB, else GPE_EINVAL at creation):
 *          26-36 one cell per lane with one (26-31) or two (32-36) layers of
 *          prefetch; 37-48 and 55-65 two cells per lane in FP32 (16-byte
 *          accesses; FP64 runs them one cell per lane), 43, 44, 48, 56 and 57
 *          with the registers held to 4 waves per SIMD, 55-57 loading the
 *          row-end values in the layer, 58-65 varying waves, z-chunk and
 *          block order; 49-54 z-chunks sized at launch so that one round of
 *          blocks fills the chip (54 one cell per lane, two layers of
 *          prefetch).
 * In FP64 every shape gives the same bits; in FP32 the buffer tiles add V*dt
 * as an x-y part and a z part, each rounded to FP32.
 * Tile 0 in the config picks by size: when both fields fit in 192 MB (the
 * Infinity Cache holds them) a cached read-modify-write in the buffer
 * z-march with a contiguous slab of layers per XCD, 16 layers per chunk
 * (tile 60; FP32 two cells per lane); below 2 GiB per
 * field the buffer z-march of 16 layers (tile 28 FP64, tile 37 FP32); larger
 * fields the z-march on 64-bit addresses (tile 13 FP64, tile 19 FP32). */
int gpe_primal_tiles(void);

/* Host helpers, the arithmetic the kernels use. */
/* {cD dt, cA dt, c0 dt, g dt}: cD = -3/(8a^2), cA = -1/(8a^2), c0 = 15/(4a^2). */
void gpe_primal_coefficients(double a, double dt, double g, double out[4]);
/* origin = -(a/2)(n - 1/2) per axis: the sites span a box centred on 0. */
void gpe_primal_centered_origin(const uint32_t dims[3], double a, double origin[3]);
/* 3 doubles per site, host order. */
void gpe_primal_positions(const uint32_t dims[3], const double origin[3], double a, double *xyz);
/* V per site, host order. */
void gpe_primal_trap_potential(const uint32_t dims[3], const double origin[3], double a, double ratio, double *V);

#ifdef __cplusplus
}
#endif
#endif /* VAIKKARI_GPE_PRIMAL_H */
