/*
 * vaikkari_gpe_output.h -- pictures and volume maps of a solver's even
 * field, reduced on the GPU and written in the reference's file formats.
 *
 * The reference copies the whole padded even field to the host each
 * iteration (GPUGrossPitaevskii_BCC.cu:470-472, 192 bytes per cell in FP64)
 * and reduces it there (GPUGrossPitaevskii_BCC.cu:424-468).  Here the
 * reduction runs on the device and only its result crosses to the host:
 * 2 bytes per cell for a volume map, 3 bytes per pixel for a slice.
 *
 * Both cover the padded extent of the reference's host copy, (nx+2) x
 * (ny+2) x (nz+2) cells with x fastest, whose outer layer is the zero halo.
 */
#ifndef VAIKKARI_GPE_OUTPUT_H
#define VAIKKARI_GPE_OUTPUT_H

#include "vaikkari_gpe.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gpe_volume_flags {
    GPE_VOLUME_ALL_CLASSES = 0,     /* s = sum over the cell's 12 dual nodes of |psi_l|^2 */
    GPE_VOLUME_REFERENCE_CLASS0 = 1 /* the reference's arithmetic: |psi_0|^2 added 12 times
                                     * (GPUGrossPitaevskii_BCC.cu:456-459 reads values[0]
                                     * inside the loop over classes) */
} gpe_volume_flags;

/* Volume map of the even field (GPUGrossPitaevskii_BCC.cu:446-465): for each
 * padded cell, v = min(65535, trunc(unit * s)), s accumulated in FP64 in
 * class order; halo cells (and a NaN product) give 0.  The reference uses
 * unit = 60000 / (12 fmax^2), fmax = VortexState::searchFunctionMax.
 * out: (nx+2)(ny+2)(nz+2) values.  GPE_EINVAL for unit < 0 or not finite. */
gpe_status gpe_volume_map(gpe_solver *s, double unit, uint32_t flags, uint16_t *out);

/* Slice picture of the even field (GPUGrossPitaevskii_BCC.cu:426-442 drawn
 * with Picture::save, Output/Picture.cpp:81-137): pixel (i, j) of padded z
 * layer k (1..nz are the interior; the reference draws k = nz/2 + 1) has the
 * colour gain * (Re psi_0, |psi_0|, Im psi_0) of the cell's class-0 node,
 * mapped per channel to c = 256 colour (middle = 0, as the reference saves)
 * or c = 128 colour + 128 (middle = 1), byte = 0 if c < 1 or NaN, 255 if
 * c > 255, else trunc(c).  out: (nx+2)(ny+2) pixels, 3 bytes each in the
 * order blue, green, red, pixel (i, j) at 3 (j (nx+2) + i). */
gpe_status gpe_slice_picture(gpe_solver *s, uint32_t k, double gain, int32_t middle, uint8_t *bgr);

/* saveVolumeMap (GPUGrossPitaevskii_BCC.cu:38-67): writes path + ".raw"
 * (dims[0] dims[1] dims[2] little-endian uint16, x fastest) and the
 * MetaImage header at path, spacing = cell edge per axis, centred on 0. */
gpe_status gpe_write_volume_map(const char *path, const uint16_t *vol, const uint32_t dims[3],
                                const double spacing[3]);

/* Picture::save (Output/Picture.cpp:81-137): a 24-bit BMP of width x height
 * pixels given as bgr (3 bytes per pixel, row 0 first), rows padded to 4
 * bytes, header fields as the reference writes them (its size fields count
 * 4 bytes per pixel). */
gpe_status gpe_write_bmp(const char *path, uint32_t width, uint32_t height, const uint8_t *bgr);

/* Output during gpe_integrate_in_time_output: the reference's SAVE_PICTURE /
 * SAVE_VOLUME branches (GPUGrossPitaevskii_BCC.cu:17-18, 424-468), before
 * each iteration and after the last, as <prefix>kuva<iter>.bmp and
 * <prefix>volume<iter>.mhd (+ .raw). */
typedef struct gpe_output_options {
    uint32_t save_picture;   /* slice k = nz/2 + 1, gain 5, middle 0 */
    uint32_t save_volume;    /* unit = 60000/(12 fmax^2) */
    uint32_t volume_flags;   /* gpe_volume_flags */
    uint32_t reserved;
    const char *prefix;      /* directory and/or file-name prefix; NULL = "" */
} gpe_output_options;

/* gpe_integrate_in_time with the output above. */
gpe_status gpe_integrate_in_time_output(const gpe_vortex_state *vs, double block_scale, const double minp[3],
                                        const double maxp[3], double iteration_period,
                                        uint32_t number_of_iterations, double trap_ratio, uint32_t precision,
                                        const gpe_output_options *opt, gpe_reference_run *run);

#ifdef __cplusplus
}
#endif
#endif /* VAIKKARI_GPE_OUTPUT_H */
