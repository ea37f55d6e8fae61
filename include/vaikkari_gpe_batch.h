/*
 * vaikkari_gpe_batch.h -- many independent condensates stepped together.
 *
 * No counterpart in the reference, which steps exactly one field per
 * process (GPUGrossPitaevskii_BCC.cu:475-481).  A batch keeps every
 * instance's state resident in HBM and advances all of them with one set of
 * persistent kernel launches per step: the two in-place leapfrog
 * half-updates for real-time instances, the Euler update plus a
 * per-instance FP64 renormalisation for imaginary-time instances.  A chunk of steps is
 * captured once into a hipGraph and replayed (or enqueued directly, see
 * gpe_batch_set_launch_mode).  Each instance's arithmetic
 * (tile decomposition, reduction order) depends only on that instance, so
 * its result is bit-identical whatever else shares the batch.
 *
 * Instances use the in-kernel harmonic trap V = 0.5(x^2 + y^2 + ratio^2 z^2).
 * Host array order as in vaikkari_gpe.h.
 */
#ifndef VAIKKARI_GPE_BATCH_H
#define VAIKKARI_GPE_BATCH_H

#include "vaikkari_gpe.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gpe_batch gpe_batch;

typedef enum gpe_init_kind {
    GPE_INIT_HOST = 0,    /* state supplied with gpe_batch_set_state */
    GPE_INIT_GAUSSIAN = 1 /* device recipe: (pi sigma^2)^(-3/4) exp(-r^2/(2 sigma^2))
                             * (1 + noise (xr + i xi)), xr, xi ~ N(0,1) from a
                             * counter-based generator keyed by (seed, node),
                             * then scaled to unit discrete norm w sum |psi|^2
                             * (FP64 reduction on the device: a box that clips
                             * the Gaussian would otherwise start below 1);
                             * odd = even exp(i mu0 dt/2); mu0 = NaN: the
                             * state's own mu = (K + P + I)/N (FP64, on the
                             * device), which starts the leapfrog without
                             * exciting its spurious mode */
} gpe_init_kind;

typedef struct gpe_instance_config {
    uint32_t nx, ny, nz;   /* cubic cells */
    uint32_t mode;         /* gpe_mode */
    double block_scale;    /* b = h/sqrt(8) */
    double dt;             /* time step (real) or dtau (imaginary) */
    double g;
    double trap_ratio;
    double origin[3];      /* corner of cell (0,0,0); non-finite -> box centred on 0 */
    uint64_t steps;        /* steps to run */
    uint32_t init;         /* gpe_init_kind */
    uint32_t zlayers;      /* imaginary time: layers per tile; 0 or 1 = one layer
                            * (the default, as the single solver's gather
                            * kernel), > 1 = a z-march of that many layers */
    double sigma, noise, mu0; /* GPE_INIT_GAUSSIAN parameters */
    uint64_t seed;
} gpe_instance_config;

typedef struct gpe_instance_info {
    uint64_t steps_done, steps_left;
    uint64_t nodes;
    uint64_t device_bytes;
    uint32_t mode;
    uint32_t active;       /* 1 while resident */
} gpe_instance_info;

/* precision: gpe_precision of every instance; hbm_budget_bytes: instances
 * are admitted while their resident bytes fit (0 = no limit). */
gpe_status gpe_batch_create(uint32_t precision, uint64_t hbm_budget_bytes, gpe_batch **out);
void gpe_batch_destroy(gpe_batch *b);

/* Adds an instance (allocating its fields); GPE_EBUSY when it would exceed
 * the budget.  *id >= 0 identifies it until gpe_batch_remove. */
gpe_status gpe_batch_add(gpe_batch *b, const gpe_instance_config *cfg, int32_t *id);
/* Bytes an instance of this configuration would hold resident. */
uint64_t gpe_batch_instance_bytes(const gpe_batch *b, const gpe_instance_config *cfg);
gpe_status gpe_batch_set_state(gpe_batch *b, int32_t id, const double *even, const double *odd);
gpe_status gpe_batch_get_state(gpe_batch *b, int32_t id, double *even, double *odd);
gpe_status gpe_batch_observables(gpe_batch *b, int32_t id, double out[4]);
gpe_status gpe_batch_instance(const gpe_batch *b, int32_t id, gpe_instance_info *info);
gpe_status gpe_batch_remove(gpe_batch *b, int32_t id);

/* Advances every active instance by min(chunk_steps, steps_left) steps,
 * `chunks` times in a row; the chunk is one captured hipGraph (re-captured
 * only when chunk_steps or the instance set changes).  A chunk that would
 * hold more than 16384 kernel launches (one per group of real-time instances
 * and half-step, two per imaginary-time step) runs as shorter chunks; where
 * chunks end changes no result.  With elapsed_ms non-NULL the device time of
 * the whole call is returned (HIP events). */
gpe_status gpe_batch_run(gpe_batch *b, uint32_t chunk_steps, uint32_t chunks, double *elapsed_ms);
/* Sum over active instances of nodes * min(chunk_steps, steps_left), i.e.
 * the node-steps the next chunk will perform. */
uint64_t gpe_batch_chunk_node_steps(const gpe_batch *b, uint32_t chunk_steps);
/* Tile of the real-time pass: waves per block (1, 2, 4 or 8), optionally
 * | GPE_BATCH_TILE_CACHED for a cached read-modify-write of the updated field
 * (without it the field streams past the caches, non-temporal) and
 * | GPE_BATCH_TILE_SLABS to give each XCD a contiguous eighth (a slab of
 * layers) of every instance's tiles, or | GPE_BATCH_TILE_GROUP_SLABS a
 * contiguous eighth of a whole group's tiles (whole instances, each in
 * layer order; small instances then keep their neighbour layers in one L2)
 * (without either y-bands of every layer when an instance's block rows
 * divide by 8, else plain order); only while the batch is empty.  A new
 * batch starts with 3 | GPE_BATCH_TILE_QUAD | GPE_BATCH_TILE_CACHED |
 * GPE_BATCH_TILE_GROUP_SLABS.  Results do not depend on it.  The autotuner
 * picks it. */
#define GPE_BATCH_TILE_CACHED 0x100u
#define GPE_BATCH_TILE_SLABS 0x200u
#define GPE_BATCH_TILE_GROUP_SLABS 0x400u
/* | GPE_BATCH_TILE_QUAD: one thread per cell and group of 4 classes (three
 * times the waves of one thread per cell, for about the same loads), waves
 * per block 3, 6 or 12; the single solver's gather tiles 19-24.  Same bits. */
#define GPE_BATCH_TILE_QUAD 0x800u
gpe_status gpe_batch_set_tile(gpe_batch *b, uint32_t tile);
/* How gpe_batch_run issues a chunk: GPE_BATCH_GRAPH replays a captured
 * hipGraph of the chunk (the default), GPE_BATCH_STREAM enqueues the chunk's
 * kernels onto the batch's stream directly.  Results are identical. */
typedef enum gpe_batch_launch_mode { GPE_BATCH_GRAPH = 0, GPE_BATCH_STREAM = 1 } gpe_batch_launch_mode;
gpe_status gpe_batch_set_launch_mode(gpe_batch *b, uint32_t mode);
/* Real-time instances advance group by group: contiguous runs of instances
 * (in id order; a removed instance's id is reused) whose fields total at
 * most `bytes` each run
 * the whole chunk before the next group starts, so that much of a group's
 * data is still in the GPU's last-level cache (256 MB Infinity Cache) when
 * the next half-update reads it.  0 puts every instance in one group (each
 * half-update sweeps the whole batch).  Default 256 MiB; may change at any
 * time (takes effect at the next gpe_batch_run).  Results do not depend on
 * it: instances are independent. */
gpe_status gpe_batch_set_group_bytes(gpe_batch *b, uint64_t bytes);
uint64_t gpe_batch_group_bytes(const gpe_batch *b);
/* Number of active instances; bytes resident. */
int32_t gpe_batch_active(const gpe_batch *b);
uint64_t gpe_batch_resident_bytes(const gpe_batch *b);

#ifdef __cplusplus
}
#endif
#endif /* VAIKKARI_GPE_BATCH_H */
