/*
 * FFT volume operator of the dielectric VIE, C-ABI of libvie_hip.so.
 *
 * Replaces the reference's vie_fft.h:7-32 symbol for symbol.  The backend
 * this library implements is "hip": hand-written gfx950 pencil-FFT kernels
 * with the kernel spectrum built on the GPU in FP64.  "auto" selects it;
 * "builtin" and "fftw" name the reference's CPU backends, which this library
 * does not contain, so vie_fft_backend_set returns 0 for them (the reference
 * returns 0 for "fftw" when built without FFTW, vie_fft.c:73-76).
 *
 * Equal cubes on a common integer lattice; coordinates are nonnegative grid
 * indices, cell-major xyz.  Plans own device workspace: apply is not
 * reentrant on one plan.
 */
#ifndef VIE_FFT_H
#define VIE_FFT_H
#include "vie_iterative.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct vie_fft_plan vie_fft_plan;

int vie_fft_backend_set(const char *name);                 /* vie_fft.c:69-78 */
int vie_fft_fftw_available(void);                          /* vie_fft.c:79-86 */
const char *vie_fft_backend_requested_name(void);          /* vie_fft.c:87-88 */
const char *vie_fft_backend_name(const vie_fft_plan *plan);/* vie_fft.c:89-90 */
const char *vie_fft_backend_last_used_name(void);          /* vie_fft.c:91-92 */
void vie_fft_backend_clear_last_used(void);                /* vie_fft.c:93 */
/* vie_fft.c:378-441.  byte_limit covers the plan's host and device storage.
   k0 * pitch <= 1 is the certified kernel domain.  On failure *out is NULL. */
int vie_fft_create(vie_fft_plan **out, const size_t shape[3], size_t cells, const size_t *xyz,
                   double pitch, double k0, VIE_CDOUBLE epsilon_r, size_t byte_limit);
/* vie_fft.c:445-455: passive scalar permittivity per cell after creation. */
int vie_fft_set_permittivity(vie_fft_plan *plan, size_t cell, VIE_CDOUBLE epsilon_r);
void vie_fft_destroy(vie_fft_plan *plan);                  /* vie_fft.c:190-197 */
size_t vie_fft_bytes(const vie_fft_plan *plan);            /* vie_fft.c:198 */
size_t vie_fft_threads(const vie_fft_plan *plan);          /* vie_fft.c:199: host threads used (1) */
int vie_fft_parallel_available(void);                      /* vie_fft.c:200-215 */
void vie_fft_operator(vie_fft_plan *plan, vie_linear_operator *op); /* vie_fft.c:274-279 */

#ifdef __cplusplus
}
#endif
#endif
