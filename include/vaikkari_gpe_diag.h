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

#ifdef __cplusplus
}
#endif
#endif /* VAIKKARI_GPE_DIAG_H */
