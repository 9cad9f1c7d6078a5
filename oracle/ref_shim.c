/*
 * Pointer-only C entry points over the reference's own hot-path sources
 * (vie_fft.c, vie_kernel.c, vie_iterative.c, dielectric_fft.c, vie_ground.c),
 * so Python can drive them through ctypes without passing complex values by
 * value.  Compiled together with those sources into oracle/_ref/libvieref.so
 * by oracle/Makefile.  TEST INFRASTRUCTURE ONLY: used by tests/ to pin the
 * oracle and by bench.py's CPU-baseline leg; never linked by the product.
 */
#include "dielectric_fft.h"
#include "vie_fft.h"
#include "vie_ground.h"
#include "vie_halfspace.h"
#include "vie_iterative.h"
#include "vie_kernel.h"
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int refshim_threads(int n)
{
#ifdef _OPENMP
  if(n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n; return 1;
#endif
}

/* vie_cuboid_n (vie_kernel.c:550) with explicit options; out = 18 doubles. */
int refshim_cuboid_n(const double tc[3], const double tsz[3], const double sc[3],
                     const double ssz[3], double k, double rtol, double atol,
                     int max_order, int far_order, double *out, int *order_used)
{
  vie_cuboid t, s; vie_kernel_options o; vie_kernel_report r; complex double m[3][3];
  int ok;
  memcpy(t.center, tc, 3*sizeof(double)); memcpy(t.size, tsz, 3*sizeof(double));
  memcpy(s.center, sc, 3*sizeof(double)); memcpy(s.size, ssz, 3*sizeof(double));
  vie_kernel_default_options(&o);
  o.relative_tolerance = rtol; o.absolute_tolerance = atol;
  o.maximum_order = max_order; o.far_order = far_order;
  ok = vie_cuboid_n(&t, &s, k, &o, m, &r);
  if(ok) memcpy(out, m, sizeof(m));
  if(order_used) *order_used = r.quadrature_order;
  return ok;
}

/* vie_cuboid_pec_n (vie_ground.c:20) with default options. */
int refshim_cuboid_pec_n(const double tc[3], const double tsz[3], const double sc[3],
                         const double ssz[3], double k, double *out)
{
  vie_cuboid t, s; complex double m[3][3]; int ok;
  memcpy(t.center, tc, 3*sizeof(double)); memcpy(t.size, tsz, 3*sizeof(double));
  memcpy(s.center, sc, 3*sizeof(double)); memcpy(s.size, ssz, 3*sizeof(double));
  ok = vie_cuboid_pec_n(&t, &s, k, 0, NULL, m, NULL);
  if(ok) memcpy(out, m, sizeof(m));
  return ok;
}

/* vie_halfspace_cuboids (vie_halfspace.c:230) with default options: the
   reflected tensor of a passive lower half-space eps = er + i ei (ei <= 0)
   between testing cuboid (tc, tsz) and source cuboid (sc, ssz); out = 18
   doubles (N), order_used the converged quadrature order. */
int refshim_halfspace_cuboids(const double tc[3], const double tsz[3], const double sc[3],
                              const double ssz[3], double k, double er, double ei, double *out,
                              int *order_used)
{
  vie_cuboid t, s; complex double n[3][3], h[3][3]; vie_halfspace_report r; int ok;
  memcpy(t.center, tc, 3*sizeof(double)); memcpy(t.size, tsz, 3*sizeof(double));
  memcpy(s.center, sc, 3*sizeof(double)); memcpy(s.size, ssz, 3*sizeof(double));
  ok = vie_halfspace_cuboids(&t, &s, k, er + I*ei, NULL, n, h, &r);
  if(ok) memcpy(out, n, sizeof(n));
  if(order_used) *order_used = r.order;
  return ok;
}

/* vie_halfspace_fresnel (vie_halfspace.c:14): out = rs (2 doubles), rp (2). */
int refshim_halfspace_fresnel(double er, double ei, double u, double *out)
{
  complex double rs, rp; int ok = vie_halfspace_fresnel(er + I*ei, u, &rs, &rp);
  if(ok) { out[0] = creal(rs); out[1] = cimag(rs); out[2] = creal(rp); out[3] = cimag(rp); }
  return ok;
}

int refshim_backend_set(const char *name) { return vie_fft_backend_set(name); }

vie_fft_plan *refshim_create(const size_t shape[3], size_t cells, const size_t *xyz,
                             double h, double k, double er_re, double er_im, size_t limit)
{
  vie_fft_plan *p = NULL;
  if(!vie_fft_create(&p, shape, cells, xyz, h, k, er_re+I*er_im, limit)) return NULL;
  return p;
}

int refshim_set_eps(vie_fft_plan *p, size_t cell, double re, double im)
{ return vie_fft_set_permittivity(p, cell, re+I*im); }

void refshim_destroy(vie_fft_plan *p) { vie_fft_destroy(p); }
size_t refshim_bytes(const vie_fft_plan *p) { return vie_fft_bytes(p); }

int refshim_apply(vie_fft_plan *p, const double *x, double *y)
{
  vie_linear_operator op; vie_fft_operator(p, &op);
  return op.apply(op.context, (const complex double *)x, (complex double *)y);
}

int refshim_precondition(vie_fft_plan *p, const double *x, double *y)
{
  vie_linear_operator op; vie_fft_operator(p, &op);
  return op.precondition(op.precondition_context, (const complex double *)x, (complex double *)y);
}

/* report: status, iterations, applications, restarts, residual, relative. */
static void pack_report(const vie_gmres_report *r, double *out)
{
  out[0] = r->status; out[1] = (double)r->iterations; out[2] = (double)r->operator_applications;
  out[3] = (double)r->restarts; out[4] = r->residual_norm; out[5] = r->relative_residual;
}

int refshim_gmres(vie_fft_plan *p, const double *b, double *x, size_t restart,
                  size_t maxit, double rtol, double atol, double *report)
{
  vie_linear_operator op; vie_gmres_options o; vie_gmres_report r; int ok;
  vie_fft_operator(p, &op);
  o.restart = restart; o.maximum_iterations = maxit;
  o.relative_tolerance = rtol; o.absolute_tolerance = atol;
  ok = vie_gmres(&op, (const complex double *)b, (complex double *)x, &o, &r);
  pack_report(&r, report);
  return ok;
}

/* Hybrid conductor/volume system (dielectric_fft.c:89-185). Cells are given
   as centers/sizes/material ids; materials as (id, eps_r, eps_loss,
   conductivity) rows. */
typedef struct { dielectric_fft_system *s; dielectric_model m; dielectric_mesh_result mesh;
                 dielectric_material *mat; dielectric_cell *cells; dielectric_sample *samples; } refshim_system;

refshim_system *refshim_system_prepare(size_t ncells, const double *centers, const double *sizes,
                                       const int *material_ids, size_t nmat, const int *mat_ids,
                                       const double *mat_props, size_t conductors, double k0,
                                       size_t extra, int pec, char *message, size_t message_len)
{
  refshim_system *r = calloc(1, sizeof(*r)); dielectric_diagnostic d; size_t i; int ok;
  if(!r) return NULL;
  r->mat = calloc(nmat ? nmat : 1, sizeof(*r->mat));
  r->cells = calloc(ncells ? ncells : 1, sizeof(*r->cells));
  for(i = 0; i < nmat; i++) {
    r->mat[i].id = mat_ids[i]; r->mat[i].epsilon_r = mat_props[3*i];
    r->mat[i].epsilon_loss = mat_props[3*i+1]; r->mat[i].conductivity = mat_props[3*i+2];
  }
  for(i = 0; i < ncells; i++) {
    memcpy(r->cells[i].center, centers+3*i, 3*sizeof(double));
    memcpy(r->cells[i].size, sizes+3*i, 3*sizeof(double));
    r->cells[i].material_id = material_ids[i];
  }
  r->m.materials = r->mat; r->m.material_count = nmat;
  r->mesh.cells = r->cells; r->mesh.cell_count = ncells;
  memset(&d, 0, sizeof(d));
  ok = (pec ? dielectric_fft_prepare_pec : dielectric_fft_prepare)(&r->s, &r->m, &r->mesh,
        conductors, k0, extra, &d);
  if(message && message_len) { strncpy(message, d.message, message_len-1); message[message_len-1] = 0; }
  if(!ok) { free(r->mat); free(r->cells); free(r); return NULL; }
  return r;
}

/* refshim_system_prepare with tabulated materials: material i is tabulated
   when sample_counts[i] > 0, its samples 4 doubles each (frequency MHz,
   eps_r, loss, conductivity_kind) at samples + 4 * (sum of earlier counts). */
refshim_system *refshim_system_prepare_tab(size_t ncells, const double *centers, const double *sizes,
                                           const int *material_ids, size_t nmat, const int *mat_ids,
                                           const double *mat_props, const size_t *sample_counts,
                                           const double *samples, size_t conductors, double k0, size_t extra,
                                           int pec, char *message, size_t message_len)
{
  refshim_system *r = calloc(1, sizeof(*r)); dielectric_diagnostic d; size_t i, j, total = 0, at = 0; int ok;
  if(!r) return NULL;
  for(i = 0; i < nmat; i++) total += sample_counts[i];
  r->mat = calloc(nmat ? nmat : 1, sizeof(*r->mat));
  r->cells = calloc(ncells ? ncells : 1, sizeof(*r->cells));
  r->samples = calloc(total ? total : 1, sizeof(*r->samples));
  for(i = 0; i < nmat; i++) {
    r->mat[i].id = mat_ids[i]; r->mat[i].epsilon_r = mat_props[3*i];
    r->mat[i].epsilon_loss = mat_props[3*i+1]; r->mat[i].conductivity = mat_props[3*i+2];
    if(sample_counts[i]) {
      r->mat[i].tabulated = 1; r->mat[i].samples = r->samples + at; r->mat[i].sample_count = sample_counts[i];
      for(j = 0; j < sample_counts[i]; j++, at++) {
        r->samples[at].material_id = mat_ids[i];
        r->samples[at].frequency_mhz = samples[4*at]; r->samples[at].epsilon_r = samples[4*at+1];
        r->samples[at].loss = samples[4*at+2]; r->samples[at].conductivity_kind = (int)samples[4*at+3];
      }
    }
  }
  for(i = 0; i < ncells; i++) {
    memcpy(r->cells[i].center, centers+3*i, 3*sizeof(double));
    memcpy(r->cells[i].size, sizes+3*i, 3*sizeof(double));
    r->cells[i].material_id = material_ids[i];
  }
  r->m.materials = r->mat; r->m.material_count = nmat;
  r->mesh.cells = r->cells; r->mesh.cell_count = ncells;
  memset(&d, 0, sizeof(d));
  ok = (pec ? dielectric_fft_prepare_pec : dielectric_fft_prepare)(&r->s, &r->m, &r->mesh,
        conductors, k0, extra, &d);
  if(message && message_len) { strncpy(message, d.message, message_len-1); message[message_len-1] = 0; }
  if(!ok) { free(r->mat); free(r->cells); free(r->samples); free(r); return NULL; }
  return r;
}

void refshim_system_free(refshim_system *r)
{ if(r) { dielectric_fft_free(r->s); free(r->mat); free(r->cells); free(r->samples); free(r); } }

int refshim_system_set(refshim_system *r, size_t row, size_t col, double re, double im)
{
  complex double *e = dielectric_fft_entry(r->s, row, col);
  if(!e) return 0;
  *e = re+I*im; return 1;
}

/* All conductor blocks at once through dielectric_fft_entry: C c x c, B c x v,
   D v x c, row-major (interleaved re, im). */
int refshim_system_set_blocks(refshim_system *r, size_t c, size_t v, const double *C, const double *B,
                              const double *D)
{
  size_t i, j;
  const complex double *cc = (const complex double *)C, *bb = (const complex double *)B,
                       *dd = (const complex double *)D;
  for(i = 0; i < c; i++) {
    for(j = 0; j < c; j++) *dielectric_fft_entry(r->s, i, j) = cc[i*c+j];
    for(j = 0; j < v; j++) *dielectric_fft_entry(r->s, i, c+j) = bb[i*v+j];
  }
  for(i = 0; i < v; i++) for(j = 0; j < c; j++) *dielectric_fft_entry(r->s, c+i, j) = dd[i*c+j];
  return 1;
}

size_t refshim_system_storage(refshim_system *r) { return dielectric_fft_storage(r->s); }

int refshim_system_apply(refshim_system *r, const double *x, double *y)
{
  vie_linear_operator op; dielectric_fft_get_operator(r->s, &op);
  return op.apply(op.context, (const complex double *)x, (complex double *)y);
}

int refshim_system_precondition(refshim_system *r, const double *x, double *y)
{
  vie_linear_operator op; dielectric_fft_get_operator(r->s, &op);
  return op.precondition(op.precondition_context, (const complex double *)x, (complex double *)y);
}

int refshim_system_gmres(refshim_system *r, const double *b, double *x, size_t restart,
                         size_t maxit, double rtol, double atol, double *report)
{
  vie_linear_operator op; vie_gmres_options o; vie_gmres_report rep; int ok;
  dielectric_fft_get_operator(r->s, &op);
  o.restart = restart; o.maximum_iterations = maxit;
  o.relative_tolerance = rtol; o.absolute_tolerance = atol;
  ok = vie_gmres(&op, (const complex double *)b, (complex double *)x, &o, &rep);
  pack_report(&rep, report);
  return ok;
}

/* Dense-matrix GMRES adapter (vie_iterative.c:228-244). */
int refshim_dense_gmres(const double *a, size_t n, const double *b, double *x,
                        size_t restart, size_t maxit, double rtol, double atol, double *report)
{
  vie_gmres_options o; vie_gmres_report rep; int ok;
  o.restart = restart; o.maximum_iterations = maxit;
  o.relative_tolerance = rtol; o.absolute_tolerance = atol;
  ok = vie_dense_gmres((const complex double *)a, n, (const complex double *)b,
                       (complex double *)x, &o, &rep);
  pack_report(&rep, report);
  return ok;
}

/* Coupled [C B; D V] operator over a reference volume plan built with an
   unlimited budget (vie_fft_create).  The reference's own dielectric_fft
   prepare caps storage at 512 MiB (dielectric_fft.h:5), which the 64^3
   coupled workload exceeds; the block products below restate
   dielectric_fft.c:53-64 (same loops, same OpenMP split) around the
   reference's volume apply / precondition.  C: c x c, B: c x v, D: v x c,
   all row-major, volume index 3*cell+component. */
typedef struct { vie_fft_plan *p; size_t c, v; const complex double *C, *B, *D; } refshim_coupled;

refshim_coupled *refshim_coupled_create(vie_fft_plan *p, size_t c, const double *C, const double *B, const double *D)
{
  refshim_coupled *r = calloc(1, sizeof(*r));
  vie_linear_operator op;
  if(!r) return NULL;
  vie_fft_operator(p, &op);
  r->p = p; r->c = c; r->v = op.order;
  r->C = (const complex double *)C; r->B = (const complex double *)B; r->D = (const complex double *)D;
  return r;
}

void refshim_coupled_free(refshim_coupled *r) { free(r); }

static int coupled_apply(void *ctx, const complex double *x, complex double *y)
{
  refshim_coupled *r = ctx; vie_linear_operator op; size_t i, j, c = r->c, v = r->v;
  vie_fft_operator(r->p, &op);
  if(!op.apply(op.context, x+c, y+c)) return 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if(c >= 2)
#endif
  for(i = 0; i < c; i++) {
    complex double s = 0;
    for(j = 0; j < c; j++) s += r->C[i*c+j]*x[j];
    for(j = 0; j < v; j++) s += r->B[i*v+j]*x[c+j];
    y[i] = s;
  }
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if(v >= 1024)
#endif
  for(i = 0; i < v; i++) for(j = 0; j < c; j++) y[c+i] += r->D[i*c+j]*x[j];
  return 1;
}

static int coupled_precondition(void *ctx, const complex double *x, complex double *y)
{
  refshim_coupled *r = ctx; vie_linear_operator op; size_t i;
  vie_fft_operator(r->p, &op);
  if(!op.precondition(op.precondition_context, x+r->c, y+r->c)) return 0;
  for(i = 0; i < r->c; i++) y[i] = x[i]/r->C[i*r->c+i];
  return 1;
}

int refshim_coupled_apply(refshim_coupled *r, const double *x, double *y)
{ return coupled_apply(r, (const complex double *)x, (complex double *)y); }

int refshim_coupled_gmres(refshim_coupled *r, const double *b, double *x, size_t restart, size_t maxit,
                          double rtol, double atol, double *report)
{
  vie_linear_operator op = {r->c+r->v, coupled_apply, r, coupled_precondition, r};
  vie_gmres_options o; vie_gmres_report rep; int ok;
  o.restart = restart; o.maximum_iterations = maxit;
  o.relative_tolerance = rtol; o.absolute_tolerance = atol;
  ok = vie_gmres(&op, (const complex double *)b, (complex double *)x, &o, &rep);
  pack_report(&rep, report);
  return ok;
}
