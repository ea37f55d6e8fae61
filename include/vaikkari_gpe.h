/*
 * vaikkari_gpe.h -- C ABI of the MI355X (gfx950) Gross-Pitaevskii stepper.
 *
 * The reference (markus-kivioja/vaikkari) has no plugin or FFI layer: its
 * single-device solver is the free function
 *   uint integrateInTime(const VortexState&, ddouble block_scale,
 *                        const Vector3& minp, const Vector3& maxp,
 *                        ddouble iteration_period, uint number_of_iterations)
 * (koodit/GrossPitaevskii/GPUGrossPitaevskii_BCC.cu:218) around one kernel
 *   update(PitchedPtr next, PitchedPtr prev, PitchedPtr pot, int2* lapind,
 *          double2 lapfacs, double g, uint3 dims)
 * (GPUGrossPitaevskii_BCC.cu:176).  This header replaces both: a solver object
 * owns the device state (the reference's cudaMalloc3D buffers,
 * GPUGrossPitaevskii_BCC.cu:269-287) and gpe_step() replaces the host loop of
 * update<<<>>> launches (GPUGrossPitaevskii_BCC.cu:475-481).
 * gpe_integrate_in_time() is the drop-in for integrateInTime itself.
 *
 * Discretisation (identical to the reference): body-centred-cubic Delaunay
 * mesh, psi on the 12 dual nodes (tetrahedron circumcentres) of every cubic
 * cell, 4-neighbour DEC Laplacian with Hodge weight 1.5/b^2, homogeneous
 * Dirichlet (zero) values outside the nx*ny*nz cell box.
 *
 * Host array order everywhere: [k][j][i][l] with the class l (0..11) fastest,
 * exactly the reference's Psi0/pot order without padding
 * (GPUGrossPitaevskii_BCC.cu:257).  Complex arrays interleave (re, im).
 *
 * Errors are returned as gpe_status; nothing exits the process.  The text of
 * the last error of the calling thread is gpe_last_error().  One solver must
 * not be used from two host threads at once.
 */
#ifndef VAIKKARI_GPE_H
#define VAIKKARI_GPE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPE_CLASSES 12
#define GPE_FACES 4

typedef enum gpe_status {
    GPE_OK = 0,
    GPE_EINVAL = 1,    /* bad argument or configuration */
    GPE_ENOMEM = 2,    /* device or host allocation failed */
    GPE_EDEVICE = 3,   /* HIP runtime error */
    GPE_EUNSTABLE = 4, /* non-finite values or a time step beyond the stability bound */
    GPE_EBUSY = 5,     /* resource budget exhausted (batch / serving admission) */
    GPE_ENOTSUP = 6,   /* combination not supported */
    GPE_EIO = 7,       /* file could not be read or written, or is not what it should be */
    GPE_ECANCELED = 8, /* serving: the request was cancelled (gpe_server_cancel) */
    GPE_ETIMEDOUT = 9  /* serving: the request outlived the server's timeout */
} gpe_status;

typedef enum gpe_precision {
    GPE_FP64 = 0,     /* complex FP64 storage and arithmetic (parity path) */
    GPE_FP32_BULK = 1 /* complex FP32 storage and stencil; FP64 reductions */
} gpe_precision;

typedef enum gpe_mode {
    GPE_REAL_TIME = 0, /* leapfrog, GPUGrossPitaevskii_BCC.cu:475-481 */
    GPE_IMAG_TIME = 1  /* Euler step + FP64 renormalisation every step */
} gpe_mode;

typedef enum gpe_potential_kind {
    GPE_POT_ARRAY = 0, /* V streamed from an array given by gpe_set_potential */
    GPE_POT_TRAP = 1   /* V = 0.5*(x^2 + y^2 + ratio^2 z^2) evaluated in-kernel
                          (potentialV3, GPUGrossPitaevskii_BCC.cu:33-36) */
} gpe_potential_kind;

/* Kernel family.  GPE_KERNEL_AUTO picks the fastest valid one. */
typedef enum gpe_kernel {
    GPE_KERNEL_AUTO = 0,
    GPE_KERNEL_GATHER = 1, /* one thread per cell, all neighbours via global loads */
    GPE_KERNEL_ZMARCH = 2, /* one thread per (i,j) column marching in z, register reuse */
    GPE_KERNEL_FUSED = 3   /* odd and even half-updates fused in one z-march pass */
} gpe_kernel;

typedef struct gpe_config {
    uint32_t nx, ny, nz;   /* cubic cells (reference xsize, ysize, zsize) */
    uint32_t precision;    /* gpe_precision */
    double block_scale;    /* b; cube edge a = sqrt(8)*b (GPUGrossPitaevskii_BCC.cu:70) */
    double dt;             /* time step (real time) or dtau (imaginary time), unscaled */
    double g;              /* interaction strength, unscaled */
    double origin[3];      /* p0: corner of cell (0,0,0) (GPUGrossPitaevskii_BCC.cu:227) */
    uint32_t potential;    /* gpe_potential_kind */
    uint32_t mode;         /* gpe_mode */
    double trap_ratio;     /* RATIO of the in-kernel trap (GPUGrossPitaevskii_BCC.cu:13) */
    uint32_t kernel;       /* gpe_kernel */
    uint32_t tile;         /* tile shape 1..gpe_variant_tiles(kernel), 0 = the family's defaults;
                              gather: by grid, precision and mode (real time with both fields in
                              256 MiB: tile 18 FP64, 19 FP32, or 23 on grids of at most 1024
                              one-wave blocks whose rows split into 8 y-bands; imaginary time with both fields in 256 MiB:
                              tile 19 FP64 from 256 four-wave blocks, 6 FP32 from 700;
                              FP64 real time beyond: tile 24 when the wave rows split into 8
                              y-bands of whole 6-wave blocks;
                              else the y-band tile with the most waves whose block rows divide
                              by 8: 13, 11, 12, 14 for 8 (FP64 imaginary time only), 4, 2, 1
                              waves; none: tile 7).  128^3 cells: FP64 tile 24 (imaginary
                              time 13), FP32 tile 11 */
} gpe_config;

typedef struct gpe_solver gpe_solver;

typedef struct gpe_solver_info {
    uint64_t nodes;          /* 12*nx*ny*nz */
    uint64_t device_bytes;   /* resident device memory of this solver */
    uint64_t steps_done;     /* steps taken since creation */
    uint32_t kernel;         /* gpe_kernel that gpe_step runs in the current mode */
    uint32_t tile;           /* its tile shape (1-based) */
    uint32_t launches_per_step;
    double bytes_per_node_step; /* algorithmic HBM bytes per node-step of the hot kernel */
} gpe_solver_info;

/* ---- solver lifecycle ------------------------------------------------ */
gpe_status gpe_create(const gpe_config *cfg, gpe_solver **out);
void gpe_destroy(gpe_solver *s);

/* Unscaled potential V (nodes doubles).  Switches the solver to
 * GPE_POT_ARRAY; the library multiplies by dt like
 * GPUGrossPitaevskii_BCC.cu:307. */
gpe_status gpe_set_potential(gpe_solver *s, const double *V);

/* even, odd: 2*nodes doubles.  odd may be NULL in imaginary-time mode. */
gpe_status gpe_set_state(gpe_solver *s, const double *even, const double *odd);
gpe_status gpe_get_state(gpe_solver *s, double *even, double *odd);

/* Switch mode and time step.  Entering real time requires an odd field:
 * set one with gpe_set_state or gpe_init_odd. */
gpe_status gpe_set_mode(gpe_solver *s, uint32_t mode, double dt);

/* odd = even * exp(+i*mu*dt/2): the reference's start of the leapfrog
 * (odd = even * getPhase(-dt/2), GPUGrossPitaevskii_BCC.cu:325,342-343). */
gpe_status gpe_init_odd(gpe_solver *s, double mu);

/* The even field set on the device, without a host array: the Gaussian
 * (pi sigma^2)^(-3/4) exp(-r^2/(2 sigma^2)) (1 + noise (xr + i xi)) centred on
 * 0, xr, xi ~ N(0,1) from a counter-based generator keyed by (seed, node),
 * scaled to unit discrete norm (FP64 reduction); the same recipe as a batch
 * instance's GPE_INIT_GAUSSIAN.  Real time: odd = even exp(i mu dt/2) as
 * gpe_init_odd, with mu = NaN meaning the state's own (K + P + I)/N. */
gpe_status gpe_init_gaussian(gpe_solver *s, double sigma, double noise, uint64_t seed, double mu);

/* nsteps time steps.  Real time: nsteps x (odd update, even update).
 * Imaginary time: nsteps x (Euler step, FP64 renormalisation). */
gpe_status gpe_step(gpe_solver *s, uint64_t nsteps);

/* As gpe_step, timed with HIP events on the solver's stream; elapsed_ms is
 * the device time of the whole enqueue, launches the number of hot-kernel
 * launches it contained. */
gpe_status gpe_time_steps(gpe_solver *s, uint64_t nsteps, double *elapsed_ms, uint64_t *launches);

/* FP64 observables of field 0 (even) or 1 (odd):
 * out = {N, K, P, I}: N = w*sum|psi|^2, K = w*sum 1/2 Re(conj(psi)(-L psi)),
 * P = w*sum V|psi|^2, I = w*sum g|psi|^4, w = sqrt(32/9) b^3.
 * Energy E = K + P + I/2; chemical potential mu = (K + P + I)/N. */
gpe_status gpe_observables(gpe_solver *s, int field, double out[4]);

/* Real time: nsteps (>= 1) leapfrog steps, with out = the observables of
 * gpe_observables of the even field as the last step reads it (the state
 * after nsteps - 1 steps), accumulated inside that step's odd half-update
 * from the values it already holds: no extra pass over the field.  The
 * steps are bit-identical to gpe_step.  GPE_ENOTSUP in imaginary time. */
gpe_status gpe_step_observe(gpe_solver *s, uint64_t nsteps, double out[4]);

/* Distance of field field_a (0 even, 1 odd) of solver a to field field_b of
 * solver b on the same cell grid, in any mix of precisions, reduced on the
 * device in FP64 without copying either field to the host:
 * out = {sum |a - b|^2, sum |b|^2, sum |a|^2} over the nodes, so the relative
 * L2 error of a against b is sqrt(out[0] / out[1]).  No counterpart in the
 * reference; it states the FP32-bulk error against an FP64 run at every
 * checkpoint of a precision study.  Waits for b's stream, runs on a's. */
gpe_status gpe_compare_fields(gpe_solver *a, int field_a, gpe_solver *b, int field_b, double out[3]);

/* Checkpoint and resume of a run.  No counterpart in the reference, which
 * saves only the VortexState that starts a run (VortexState.cpp:74-96) and
 * cannot resume one.  gpe_checkpoint_save writes the configuration in force
 * (mode and dt as last set, kernel family, and the tile chosen or that the
 * family's defaults are in use), the steps taken,
 * both time levels (the even one alone in imaginary time) and a streamed
 * potential to one file with a checksum; FP32 fields are stored exactly.
 * gpe_checkpoint_load creates a solver (on the current device) holding that
 * state: gpe_step on it continues the saved run bit for bit.  GPE_EIO for a
 * file that cannot be written or read, or whose magic, sizes or checksum are
 * wrong (checked before any device work). */
gpe_status gpe_checkpoint_save(gpe_solver *s, const char *path);
gpe_status gpe_checkpoint_load(const char *path, gpe_solver **out);
/* The configuration in force: as created, with the mode and dt of the last
 * gpe_set_mode, the potential kind of the last gpe_set_potential, and the
 * kernel family and tile (1-based) in use. */
gpe_status gpe_get_config(const gpe_solver *s, gpe_config *out);

gpe_status gpe_synchronize(gpe_solver *s);
gpe_status gpe_get_info(const gpe_solver *s, gpe_solver_info *info);

/* Device layout of one field ("class-row SoA"): node (i, j, k, class l) of an
 * nx x ny x nz cell grid is complex element origin + k*sz + j*sy + l*xp + i of
 * the buffer, each element (re, im) in 2 x 8 bytes (GPE_FP64) or 2 x 4 bytes
 * (GPE_FP32_BULK); the buffer holds elems elements, and every element outside
 * the box (the one-cell halo and the row padding) is zero.  No counterpart in
 * the reference, whose buffers are pitched arrays of 12 double2 per cell
 * (GPUGrossPitaevskii_BCC.cu:144-159,269-287). */
typedef struct gpe_field_layout {
    uint64_t elems;       /* complex elements of the whole buffer */
    uint64_t origin;      /* element of node (0, 0, 0, 0) */
    uint64_t xp, sy, sz;  /* class-row, cell-row and layer strides, in elements */
    uint32_t elem_bytes;  /* 16 or 8 */
    uint32_t reserved;
} gpe_field_layout;
gpe_status gpe_get_field_layout(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t precision, gpe_field_layout *out);
/* Device address of a solver's field 0 (even) or 1 (odd), in the layout
 * above.  It stays valid until the solver is destroyed; after gpe_step of the
 * fused family or gpe_set_variant ask again (those may move the time levels
 * between buffers). */
gpe_status gpe_get_field_device(gpe_solver *s, int field, void **dev);

/* gpe_observables for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The caller passes the field in device memory
 * in the layout of gpe_get_field_layout (its own buffer, or a solver's from
 * gpe_get_field_device) with cfg describing it (nx, ny, nz, precision,
 * block_scale, dt, g, origin, potential, trap_ratio; the other members are
 * ignored), pot_dt = V*dt in that layout with one real value of the field's
 * precision per element when cfg->potential is GPE_POT_ARRAY (else NULL), and
 * out = device memory for the 4 doubles {N, K, P, I} of gpe_observables, equal
 * to what gpe_observables gives for the same field.  The call only enqueues two
 * kernels on `stream` (a hipStream_t; NULL: the null stream) and returns: it
 * allocates nothing and does not synchronise.  The block partials live in an
 * 8 MiB device array of the library, one per GPU, held until the process ends:
 * grids of more than 2^18 blocks of 64 x 4 cells per layer (about 390^3
 * cells) are refused with GPE_EINVAL.  Calls that overlap on one GPU (other
 * streams, other threads) share that array and race. */
gpe_status gpe_observables_noalloc(const gpe_config *cfg, const void *field, const void *pot_dt, double *out,
                                   void *stream);

/* Kernel variants (the autotuner's search space).  gpe_set_variant switches
 * the solver's kernel family and tile shape between steps; results do not
 * depend on the variant (every variant performs the same arithmetic per
 * node).  gpe_variant_tiles gives the number of tile shapes of a family
 * (numbered 1..n; 0 selects the family's default), gpe_variant_shape the
 * shape {waves per block, z layers per block, flags} of tile 1..n with
 * flags bit 0 = XCD-contiguous block order, bit 1 = streaming
 * (non-temporal) access to the updated field, bit 2 = each XCD owns a
 * y-band of every z layer (grids whose block rows divide by 8; others run
 * in plain order), bit 3 = fused tile of one independent wavefront per cell
 * row (fields below 2 GiB only), bits 4-5 = gather tile of one thread per
 * node (1; waves per block = classes per block) or per cell and group of 4
 * classes (2).
 * GPE_KERNEL_FUSED needs two more field buffers (allocated on selection);
 * imaginary-time steps of a fused solver run the gather family's default
 * tile. */
gpe_status gpe_set_variant(gpe_solver *s, uint32_t kernel, uint32_t tile);
int gpe_variant_tiles(uint32_t kernel);
void gpe_variant_shape(uint32_t kernel, uint32_t tile, int32_t out[3]);

const char *gpe_status_string(gpe_status st);
const char *gpe_last_error(void);

/* ---- BCC dual-mesh geometry (host) ------------------------------------ */
/* Class positions q_l (node at b*q_l/sqrt(8) in the cell), 12x3 ints
 * (GPUGrossPitaevskii_BCC.cu:73-89). */
void gpe_bcc_positions(int32_t out[36]);
/* Neighbour table used by the kernels, 48 x (dx, dy, dz, class), in the
 * face order of GPUGrossPitaevskii_BCC.cu:93-140. */
void gpe_bcc_stencil(int32_t out[192]);
/* Neighbour sets derived from the positions alone (all nodes at distance b),
 * 48 x (dx, dy, dz, class), sorted per class.  Returns the number of
 * neighbours found (48 for a valid mesh). */
int gpe_bcc_stencil_derived(int32_t out[192]);

/* ---- reference initial state: VortexState (host) ---------------------- */
typedef struct gpe_vortex_state gpe_vortex_state;
/* VortexState::load (koodit/GrossPitaevskii/VortexState.cpp:49-72). */
gpe_status gpe_vortex_state_load(const char *path, gpe_vortex_state **out);
void gpe_vortex_state_free(gpe_vortex_state *vs);
/* A VortexState from memory, without a file: the fields of the reference's
 * object (koodit/GrossPitaevskii/VortexState.hpp:16-76, in the order load
 * reads them, VortexState.cpp:49-72) and its rsize x zsize table f(r, z),
 * r-major (f[ir * zsize + iz]), copied.  A caller holding a reference
 * VortexState passes its members here to call integrateInTime through
 * gpe_integrate_in_time (GPUGrossPitaevskii_BCC.cu:218). */
gpe_status gpe_vortex_state_from_table(uint32_t kappa, double g, uint32_t rsize, uint32_t zsize, double rplug,
                                       double rstep, double zstep, double mu, const double *f,
                                       gpe_vortex_state **out);
/* Copies the rsize x zsize table (r-major) into f. */
gpe_status gpe_vortex_state_table(const gpe_vortex_state *vs, double *f);
/* {kappa, g, rsize, zsize, rplug, rstep, zstep, mu} */
void gpe_vortex_state_params(const gpe_vortex_state *vs, double out[8]);
/* VortexState::getPsi(Vector3) (VortexState.cpp:204-212). */
void gpe_vortex_state_psi(const gpe_vortex_state *vs, double x, double y, double z, double out[2]);
/* VortexState::getPhase (VortexState.cpp:198-202). */
void gpe_vortex_state_phase(const gpe_vortex_state *vs, double time, double out[2]);
/* searchFunctionMax, searchMaxR, searchMaxZ (VortexState.cpp:240-280). */
double gpe_vortex_state_search_max(const gpe_vortex_state *vs);
double gpe_vortex_state_search_max_r(const gpe_vortex_state *vs, double eps);
double gpe_vortex_state_search_max_z(const gpe_vortex_state *vs, double eps);
/* integrateCurvature (VortexState.cpp:313-340). */
double gpe_vortex_state_curvature(const gpe_vortex_state *vs);

/* ---- generating a VortexState (host, FP64, serial) ---------------------
 * The reference's VortexState.cpp:5-47, 98-196, 250-257, 282-311, 342-361
 * and the state recipe of GPUGrossPitaevskii_BCC.cu:516-535.  Results are
 * bit-identical to the reference's host code built by gcc for x86-64 with
 * FMA contraction off. */
/* VortexState() (VortexState.cpp:5-10): kappa 2, g 0, 100 x 1 table, step 0.1. */
gpe_status gpe_vortex_state_create(gpe_vortex_state **out);
/* setKappa, setG (VortexState.hpp:23-27). */
void gpe_vortex_state_set_kappa(gpe_vortex_state *vs, uint32_t kappa);
void gpe_vortex_state_set_g(gpe_vortex_state *vs, double g);
/* setDimensions (VortexState.cpp:12-40): table and normalised start shape, mu = 1. */
gpe_status gpe_vortex_state_set_dimensions(gpe_vortex_state *vs, uint32_t rsize, double rstep, double rplug,
                                           uint32_t zsize, double zstep);
/* setRange (VortexState.cpp:42-47).  GPE_EINVAL for an empty range. */
gpe_status gpe_vortex_state_set_range(gpe_vortex_state *vs, double rmin, double rmax, double zmax, double rstep,
                                      double zstep);
/* initialize (VortexState.cpp:98-107): mu and f sampled from src. */
gpe_status gpe_vortex_state_initialize(gpe_vortex_state *vs, const gpe_vortex_state *src);

/* Potential V(r, z) of the relaxation, the reference's
 * ddouble (*potential)(ddouble r, ddouble z).  Where an entry point takes
 * one, NULL selects the trap 0.5(r^2 + trap_ratio^2 z^2) of potentialRZ
 * (GPUGrossPitaevskii_BCC.cu:28-31). */
typedef double (*gpe_rz_potential)(double r, double z, void *user);
/* Called where iterateSolution prints (VortexState.cpp:173). */
typedef void (*gpe_relax_progress)(uint32_t iteration, double diff, double prob, void *user);
typedef struct gpe_relax_report {
    uint32_t iterations;  /* sweeps run */
    uint32_t converged;   /* 1 when the last sweep's change fell below eps */
    double diff;          /* change of the last sweep */
    double prob;          /* <f,f> before the last normalisation */
    double mu;            /* final mu */
    double error;         /* integrateFunctionError of the result */
} gpe_relax_report;
/* Threads of gpe_vortex_state_iterate: 0 = automatic (one per hardware
 * thread, at most 16, for tables of 40000 entries or more, and serial when a
 * potential function is given), 1 = serial, n = n threads (the potential
 * function, if any, is then called from all of them).  The result does not
 * depend on it: rows run as a pipelined wavefront and every sum keeps the
 * serial order. */
void gpe_vortex_state_set_threads(gpe_vortex_state *vs, uint32_t threads);
/* iterateSolution (VortexState.cpp:109-196), without printing.  report may
 * be NULL.  GPE_EUNSTABLE when mu is not finite at the end. */
gpe_status gpe_vortex_state_iterate(gpe_vortex_state *vs, gpe_rz_potential V, double trap_ratio, void *user,
                                    uint32_t maxiters, double eps, gpe_relax_progress progress, void *progress_user,
                                    gpe_relax_report *report);
/* integrateFunctionError (VortexState.cpp:282-311). */
double gpe_vortex_state_function_error(const gpe_vortex_state *vs, gpe_rz_potential V, double trap_ratio,
                                       void *user);
/* getFunction (VortexState.cpp:220-238). */
double gpe_vortex_state_function(const gpe_vortex_state *vs, double r, double z);
/* searchMinR (VortexState.cpp:250-257). */
double gpe_vortex_state_search_min_r(const gpe_vortex_state *vs, double eps);
/* integrateZ (VortexState.cpp:342-361). */
double gpe_vortex_state_integrate_z(const gpe_vortex_state *vs, double vol);
/* save (VortexState.cpp:74-96), the format gpe_vortex_state_load reads. */
gpe_status gpe_vortex_state_save(const gpe_vortex_state *vs, const char *path);
/* main's state recipe with LOAD_STATE_FROM_DISK 0
 * (GPUGrossPitaevskii_BCC.cu:516-535): coarse state at step 0.2, support at
 * eps = 1e-5 max f, fine state at step 0.03 over that support, each relaxed
 * with `iters` iterations to eps_relax (the reference: 10000, 1e-29).
 * extent (may be NULL) = {minr, maxr, maxz} as main computes them. */
gpe_status gpe_vortex_state_reference_recipe(uint32_t kappa, double g, double trap_ratio, int32_t is_3d,
                                             uint32_t iters, double eps_relax, gpe_relax_progress progress,
                                             void *progress_user, gpe_vortex_state **out, double extent[3]);

/* Grid and origin of integrateInTime (GPUGrossPitaevskii_BCC.cu:223-227). */
void gpe_reference_dims(double block_scale, const double minp[3], const double maxp[3], uint32_t dims[3],
                        double origin[3]);

/* Scalars of the reference setup. */
typedef struct gpe_reference_run {
    uint32_t dims[3];
    uint32_t steps_per_iteration; /* GPUGrossPitaevskii_BCC.cu:298 */
    double origin[3];
    double dt;                    /* GPUGrossPitaevskii_BCC.cu:299 */
    double g;                     /* unscaled */
    double maxpot;                /* GPUGrossPitaevskii_BCC.cu:248-263 */
    double lapfac_dt, lapfac0_dt; /* GPUGrossPitaevskii_BCC.cu:291-306 */
} gpe_reference_run;

/* Setup of integrateInTime, GPUGrossPitaevskii_BCC.cu:245-348: psi0 from
 * the VortexState, V = potentialV3 with the given trap ratio, dt rule, the
 * seeded disk noise and odd = even*getPhase(-dt/2).  Any array may be NULL
 * (then only run is filled).  V is unscaled. */
gpe_status gpe_reference_setup(const gpe_vortex_state *vs, double block_scale, const double minp[3],
                               const double maxp[3], double iteration_period, double trap_ratio, uint32_t seed,
                               gpe_reference_run *run, double *even, double *odd, double *V);

/* Drop-in for integrateInTime (GPUGrossPitaevskii_BCC.cu:218-499) without the
 * picture/volume output: sets up as above, then runs number_of_iterations x
 * steps_per_iteration leapfrog steps on the GPU.  precision selects the
 * kernel family.  even_out/odd_out (may be NULL) receive the final fields. */
gpe_status gpe_integrate_in_time(const gpe_vortex_state *vs, double block_scale, const double minp[3],
                                 const double maxp[3], double iteration_period, uint32_t number_of_iterations,
                                 double trap_ratio, uint32_t precision, gpe_reference_run *run, double *even_out,
                                 double *odd_out);

#ifdef __cplusplus
}
#endif
#endif /* VAIKKARI_GPE_H */
