/*
 * Restarted GMRES contract of the VIE solver, C-ABI of libvie_hip.so.
 *
 * Declares the reference's vie_iterative.h:11-62 types with its names, layouts
 * and enum values (the callback operator, options, report and status), so the
 * library's operators plug into nec2c's own vie_gmres, and this library's
 * vie_gmres: the device-resident solve of an operator produced by
 * vie_fft_operator or dielectric_fft_get_operator of this library, with the
 * reference's argument meaning, report and return convention.  Other callback
 * operators are refused (VIE_ITERATIVE_INVALID_INPUT): nec2c keeps its own
 * vie_gmres and vie_dense_gmres (vie_iterative.c) for them.
 */
#ifndef VIE_ITERATIVE_H
#define VIE_ITERATIVE_H

#include <stddef.h>
#ifdef __cplusplus
#define VIE_CDOUBLE double _Complex
extern "C" {
#else
#include <complex.h>
#define VIE_CDOUBLE complex double
#endif

/* Writes all n outputs, input != output, returns 0 on failure
   (vie_iterative.h:8-10 of the reference). */
typedef int (*vie_vector_map)(void *context, const VIE_CDOUBLE *input, VIE_CDOUBLE *output);

typedef struct {
  size_t order;
  vie_vector_map apply;
  void *context;
  vie_vector_map precondition; /* NULL: identity */
  void *precondition_context;
} vie_linear_operator;                       /* vie_iterative.h:11-18 */

typedef struct {
  size_t restart;
  size_t maximum_iterations;
  double relative_tolerance;
  double absolute_tolerance;
} vie_gmres_options;                         /* vie_iterative.h:20-26 */

typedef enum {
  VIE_ITERATIVE_CONVERGED,
  VIE_ITERATIVE_ITERATION_LIMIT,
  VIE_ITERATIVE_BREAKDOWN,
  VIE_ITERATIVE_CALLBACK_FAILED,
  VIE_ITERATIVE_INVALID_INPUT,
  VIE_ITERATIVE_ALLOCATION_FAILED,
  VIE_ITERATIVE_NONFINITE
} vie_iterative_status;                      /* vie_iterative.h:28-37 */

typedef struct {
  vie_iterative_status status;
  size_t iterations;
  size_t operator_applications;
  size_t restarts;
  double residual_norm;
  double relative_residual;
} vie_gmres_report;                          /* vie_iterative.h:39-47 */

/* restart 30, 300 iterations, rtol 1e-12, atol 0 (vie_iterative.c:27-33) */
void vie_gmres_default_options(vie_gmres_options *options);
/* vie_iterative.c:35-48 */
const char *vie_iterative_status_name(vie_iterative_status status);
/* Right-preconditioned restarted GMRES (vie_iterative.c:80-199) of this
   library's operators, on the device.  x is the initial guess and is updated
   even on failure; success only on a recomputed ||b - A x|| <=
   max(atol, rtol ||b||).  rhs and x must not alias.  Returns 0 with status
   VIE_ITERATIVE_INVALID_INPUT for an operator the library did not make. */
int vie_gmres(const vie_linear_operator *op, const VIE_CDOUBLE *rhs, VIE_CDOUBLE *x,
              const vie_gmres_options *options, vie_gmres_report *report);

#ifdef __cplusplus
}
#endif
#endif
