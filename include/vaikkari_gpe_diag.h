/*
 * vaikkari_gpe_diag.h -- measurement helpers.  No counterpart in the
 * reference (its only timer is clock() without a device synchronisation,
 * GPUGrossPitaevskii_BCC.cu:421,488); these exist to state the kernels'
 * bandwidth against the hardware.
 */
#ifndef VAIKKARI_GPE_DIAG_H
#define VAIKKARI_GPE_DIAG_H

#include "vaikkari_gpe.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Streams `bytes` from one device buffer to another `iters` times with a
 * 16-byte-per-lane copy kernel and returns the average device time of one
 * copy (bytes read + bytes written move per copy). */
gpe_status gpe_diag_copy_bandwidth(uint64_t bytes, uint32_t iters, double *ms_per_copy);

/* Bandwidth probes of other access patterns over buffers of `bytes` each,
 * average device time of one pass over `iters` passes:
 *   GPE_DIAG_COPY  as gpe_diag_copy_bandwidth (2 x bytes move per pass),
 *   GPE_DIAG_RMW   next += prev, next read and written non-temporally, one
 *                  16-byte element per lane (3 x bytes): the leapfrog
 *                  half-update's traffic without its stencil,
 *   GPE_DIAG_READ  one buffer read (1 x bytes). */
typedef enum gpe_diag_kind { GPE_DIAG_COPY = 0, GPE_DIAG_RMW = 1, GPE_DIAG_READ = 2 } gpe_diag_kind;
gpe_status gpe_diag_stream_bandwidth(uint32_t kind, uint64_t bytes, uint32_t iters, double *ms_per_pass);

#ifdef __cplusplus
}
#endif
#endif /* VAIKKARI_GPE_DIAG_H */
