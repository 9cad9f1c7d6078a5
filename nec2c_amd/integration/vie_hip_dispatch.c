/*
 * "--fft-backend hip" for nec2c as a backend added beside builtin and fftw.
 *
 * This file is compiled into nec2c unchanged next to the reference's own
 * vie_fft.c, dielectric_fft.c and vie_iterative.c, which stay in the build
 * and keep serving "auto", "builtin" and "fftw".  It needs no edit of those
 * files: nec2c's link line routes the hot-path entry points through it with
 * GNU ld's --wrap (every reference from the other objects to SYMBOL goes to
 * __wrap_SYMBOL here, and __real_SYMBOL names nec2c's own definition):
 *
 *   -Wl,--wrap=vie_fft_backend_set -Wl,--wrap=vie_fft_backend_requested_name
 *   -Wl,--wrap=vie_fft_backend_last_used_name -Wl,--wrap=vie_fft_backend_clear_last_used
 *   -Wl,--wrap=dielectric_fft_prepare -Wl,--wrap=dielectric_fft_prepare_pec
 *   -Wl,--wrap=dielectric_fft_free -Wl,--wrap=dielectric_fft_entry
 *   -Wl,--wrap=dielectric_fft_get_operator -Wl,--wrap=dielectric_fft_storage
 *   -Wl,--wrap=dielectric_fft_threads -Wl,--wrap=dielectric_fft_backend
 *   -Wl,--wrap=vie_gmres -ldl
 *
 * With "hip" selected (vie_fft_backend_set, vie_fft.c:69-78 of the
 * reference), libvie_hip.so is opened with dlopen(RTLD_LOCAL) and its entry
 * points are taken with dlsym on that handle, so its definitions of the same
 * names never mix with nec2c's (the library binds its own references
 * internally, -Bsymbolic).  Systems prepared while "hip" is selected are the
 * library's GPU systems; the dispatcher hands nec2c a handle of its own for
 * them and forwards every call on that handle to the library.  vie_gmres on
 * the operator of such a system runs the library's device-resident solve
 * (dielectric_fft_solve); on any other operator it is nec2c's vie_gmres.
 * Every other backend value reaches nec2c's code unchanged.
 *
 * Loading the library does not start the GPU runtime: the device is first
 * used by a prepare.  The library is looked up as $VIE_HIP_LIBRARY, else as
 * "libvie_hip.so" on the loader's path.
 *
 * Timing: after every call it forwards, the dispatcher passes the library's
 * own timing phases that advanced ("fft_build", "fft_upload", "fft_gmres",
 * "fft_matvec", ...; vie_hip_phases) to nec2c's performance_add
 * (performance.h:8-17), so --timing-json lists them beside nec2c's phases.
 */
#include "dielectric_fft.h"
#include "performance.h"
#include "vie_fft.h"
#include "vie_iterative.h"

#include <dlfcn.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* nec2c's own definitions (the reference's vie_fft.c, dielectric_fft.c,
   vie_iterative.c) under --wrap. */
int __real_vie_fft_backend_set(const char *name);
const char *__real_vie_fft_backend_requested_name(void);
const char *__real_vie_fft_backend_last_used_name(void);
void __real_vie_fft_backend_clear_last_used(void);
int __real_dielectric_fft_prepare(dielectric_fft_system **, const dielectric_model *,
                                  const dielectric_mesh_result *, size_t, double, size_t,
                                  dielectric_diagnostic *);
int __real_dielectric_fft_prepare_pec(dielectric_fft_system **, const dielectric_model *,
                                      const dielectric_mesh_result *, size_t, double, size_t,
                                      dielectric_diagnostic *);
void __real_dielectric_fft_free(dielectric_fft_system *);
complex double *__real_dielectric_fft_entry(dielectric_fft_system *, size_t, size_t);
void __real_dielectric_fft_get_operator(dielectric_fft_system *, vie_linear_operator *);
size_t __real_dielectric_fft_storage(const dielectric_fft_system *);
size_t __real_dielectric_fft_threads(const dielectric_fft_system *);
const char *__real_dielectric_fft_backend(const dielectric_fft_system *);
int __real_vie_gmres(const vie_linear_operator *, const complex double *, complex double *,
                     const vie_gmres_options *, vie_gmres_report *);

/* The library's entry points used here (include/dielectric_fft.h and
   include/vie_hip.h of libvie_hip.so declare them; the model, mesh and
   diagnostic structs are layout-identical to nec2c's dielectric.h). */
typedef int (*prepare_fn)(void **, const dielectric_model *, const dielectric_mesh_result *, size_t,
                          double, size_t, dielectric_diagnostic *);
typedef struct {
  const char *name;
  double seconds;
  unsigned long long calls;
} hip_phase;                   /* vie_hip_phase */
#define HIP_PHASES 16
static struct {
  void *handle;
  prepare_fn prepare, prepare_pec;
  int (*prepare_ground)(void **, const dielectric_model *, const dielectric_mesh_result *, size_t, double, size_t,
                        double, double, dielectric_diagnostic *);
  void (*release)(void *);
  complex double *(*entry)(void *, size_t, size_t);
  void (*get_operator)(void *, vie_linear_operator *);
  size_t (*storage)(const void *);
  int (*solve)(void *, const complex double *, complex double *, const vie_gmres_options *,
               vie_gmres_report *);
  const char *(*last_error)(void);
  size_t (*phases)(hip_phase *, size_t);
} lib;

typedef struct hip_system {
  void *system;                /* the library's dielectric_fft_system */
  vie_linear_operator op;      /* the library's operator of it */
  struct hip_system *next;
} hip_system;

static pthread_mutex_t lock = PTHREAD_MUTEX_INITIALIZER;
static hip_system *systems;    /* every live GPU system handed to nec2c */
static int hip_requested, hip_last_used;
static hip_phase reported[HIP_PHASES];
static size_t reported_count;

/* Hands the library's phases that advanced since the last report to
   nec2c's timing document. */
static void report_phases(void)
{
  hip_phase now[HIP_PHASES];
  size_t n = lib.phases(now, HIP_PHASES), i, j;
  if(n > HIP_PHASES) n = HIP_PHASES;
  for(i = 0; i < n; i++) {
    double before = 0;
    unsigned long long calls = 0;
    for(j = 0; j < reported_count; j++)
      if(!strcmp(reported[j].name, now[i].name)) { before = reported[j].seconds; calls = reported[j].calls; }
    if(now[i].calls > calls) performance_add(now[i].name, now[i].seconds - before);
  }
  memcpy(reported, now, n * sizeof(*now));
  reported_count = n;
}

static int load_library(void)
{
  const char *path = getenv("VIE_HIP_LIBRARY");
  void *h;
  if(lib.handle) return 1;
  h = dlopen(path && *path ? path : "libvie_hip.so", RTLD_NOW | RTLD_LOCAL);
  if(!h) return 0;
#define TAKE(field, name) (*(void **)&lib.field = dlsym(h, name))
  if(!TAKE(prepare, "dielectric_fft_prepare") || !TAKE(prepare_pec, "dielectric_fft_prepare_pec") ||
     !TAKE(prepare_ground, "dielectric_fft_prepare_ground") ||
     !TAKE(release, "dielectric_fft_free") || !TAKE(entry, "dielectric_fft_entry") ||
     !TAKE(get_operator, "dielectric_fft_get_operator") || !TAKE(storage, "dielectric_fft_storage") ||
     !TAKE(solve, "dielectric_fft_solve") || !TAKE(last_error, "vie_hip_last_error") ||
     !TAKE(phases, "vie_hip_phases")) {
    dlclose(h);
    memset(&lib, 0, sizeof(lib));
    return 0;
  }
#undef TAKE
  lib.handle = h;
  return 1;
}

static hip_system *find(const void *s)
{
  hip_system *p;
  pthread_mutex_lock(&lock);
  for(p = systems; p && (const void *)p != s; p = p->next) ;
  pthread_mutex_unlock(&lock);
  return p;
}

/* ---- backend selection (vie_fft.c:69-93) ---- */
int __wrap_vie_fft_backend_set(const char *name)
{
  if(name && !strcmp(name, "hip")) {
    if(!load_library()) return 0;
    hip_requested = 1;
    return 1;
  }
  if(!__real_vie_fft_backend_set(name)) return 0;
  hip_requested = 0;
  return 1;
}
const char *__wrap_vie_fft_backend_requested_name(void)
{ return hip_requested ? "hip" : __real_vie_fft_backend_requested_name(); }
const char *__wrap_vie_fft_backend_last_used_name(void)
{ return hip_last_used ? "hip" : __real_vie_fft_backend_last_used_name(); }
void __wrap_vie_fft_backend_clear_last_used(void)
{ hip_last_used = 0; __real_vie_fft_backend_clear_last_used(); }

/* ---- systems (dielectric_fft.c:27-40, 84-88, 177-185) ---- */
static int prepare_hip(prepare_fn fn, dielectric_fft_system **out, const dielectric_model *model,
                       const dielectric_mesh_result *mesh, size_t conductors, double k0, size_t extra,
                       dielectric_diagnostic *diagnostic)
{
  hip_system *h = calloc(1, sizeof(*h));
  *out = NULL;
  if(!h) {
    memset(diagnostic, 0, sizeof(*diagnostic));
    snprintf(diagnostic->message, sizeof(diagnostic->message), "FFT: dispatch allocation failed");
    return 0;
  }
  if(!fn(&h->system, model, mesh, conductors, k0, extra, diagnostic)) {
    free(h);
    report_phases();
    return 0;
  }
  report_phases();
  lib.get_operator(h->system, &h->op);
  pthread_mutex_lock(&lock);
  h->next = systems;
  systems = h;
  hip_last_used = 1;
  pthread_mutex_unlock(&lock);
  *out = (dielectric_fft_system *)h;
  return 1;
}

int __wrap_dielectric_fft_prepare(dielectric_fft_system **out, const dielectric_model *model,
                                  const dielectric_mesh_result *mesh, size_t conductors, double k0,
                                  size_t extra, dielectric_diagnostic *diagnostic)
{
  if(hip_requested)
    return prepare_hip(lib.prepare, out, model, mesh, conductors, k0, extra, diagnostic);
  return __real_dielectric_fft_prepare(out, model, mesh, conductors, k0, extra, diagnostic);
}
int __wrap_dielectric_fft_prepare_pec(dielectric_fft_system **out, const dielectric_model *model,
                                      const dielectric_mesh_result *mesh, size_t conductors, double k0,
                                      size_t extra, dielectric_diagnostic *diagnostic)
{
  if(hip_requested)
    return prepare_hip(lib.prepare_pec, out, model, mesh, conductors, k0, extra, diagnostic);
  return __real_dielectric_fft_prepare_pec(out, model, mesh, conductors, k0, extra, diagnostic);
}

/* The FFT system over a real ground (GN 2), which nec2c's own FFT path
   rejects (dielectric_ground.c:28-29): with "hip" selected, the library's
   system with the reflected half-space tensor; ground permittivity as
   dielectric_ground_epsilon(frequency) gives it.  Not wrapped: nec2c calls it
   only where the patch at dielectric_wire.c:1121-1123 does. */
int dielectric_fft_prepare_ground(dielectric_fft_system **out, const dielectric_model *model,
                                  const dielectric_mesh_result *mesh, size_t conductors, double k0, size_t extra,
                                  complex double ground, dielectric_diagnostic *diagnostic)
{
  hip_system *h;
  *out = NULL;
  if(!hip_requested) {
    memset(diagnostic, 0, sizeof(*diagnostic));
    snprintf(diagnostic->message, sizeof(diagnostic->message),
             "FFT: real-ground FFT requires --fft-backend hip");
    return 0;
  }
  h = calloc(1, sizeof(*h));
  if(!h) {
    memset(diagnostic, 0, sizeof(*diagnostic));
    snprintf(diagnostic->message, sizeof(diagnostic->message), "FFT: dispatch allocation failed");
    return 0;
  }
  if(!lib.prepare_ground(&h->system, model, mesh, conductors, k0, extra, creal(ground), cimag(ground), diagnostic)) {
    free(h);
    report_phases();
    return 0;
  }
  report_phases();
  lib.get_operator(h->system, &h->op);
  pthread_mutex_lock(&lock);
  h->next = systems;
  systems = h;
  hip_last_used = 1;
  pthread_mutex_unlock(&lock);
  *out = (dielectric_fft_system *)h;
  return 1;
}

void __wrap_dielectric_fft_free(dielectric_fft_system *s)
{
  hip_system *h = find(s), **p;
  if(!h) { __real_dielectric_fft_free(s); return; }
  pthread_mutex_lock(&lock);
  for(p = &systems; *p != h; p = &(*p)->next) ;
  *p = h->next;
  pthread_mutex_unlock(&lock);
  lib.release(h->system);
  free(h);
}

complex double *__wrap_dielectric_fft_entry(dielectric_fft_system *s, size_t row, size_t column)
{
  hip_system *h = find(s);
  return h ? lib.entry(h->system, row, column) : __real_dielectric_fft_entry(s, row, column);
}

/* The operator of a GPU system: host-pointer products by the library (block
   writes made through dielectric_fft_entry before this call reach the device
   with the next product or solve). */
static int hip_apply(void *context, const complex double *x, complex double *y)
{
  hip_system *h = context;
  int ok = h->op.apply(h->op.context, x, y);
  report_phases();
  return ok;
}
static int hip_precondition(void *context, const complex double *x, complex double *y)
{
  hip_system *h = context;
  int ok = h->op.precondition(h->op.precondition_context, x, y);
  report_phases();
  return ok;
}

void __wrap_dielectric_fft_get_operator(dielectric_fft_system *s, vie_linear_operator *op)
{
  hip_system *h = find(s);
  if(!h) { __real_dielectric_fft_get_operator(s, op); return; }
  lib.get_operator(h->system, &h->op);
  op->order = h->op.order;
  op->apply = hip_apply;
  op->context = h;
  op->precondition = h->op.precondition ? hip_precondition : NULL;
  op->precondition_context = h;
}

size_t __wrap_dielectric_fft_storage(const dielectric_fft_system *s)
{
  hip_system *h = find(s);
  return h ? lib.storage(h->system) : __real_dielectric_fft_storage(s);
}
/* Host threads a GPU system uses inside a call: one. */
size_t __wrap_dielectric_fft_threads(const dielectric_fft_system *s)
{ return find(s) ? 1 : __real_dielectric_fft_threads(s); }
const char *__wrap_dielectric_fft_backend(const dielectric_fft_system *s)
{ return find(s) ? "hip" : __real_dielectric_fft_backend(s); }

/* ---- the solve (vie_iterative.c:80-199) ---- */
int __wrap_vie_gmres(const vie_linear_operator *op, const complex double *rhs, complex double *x,
                     const vie_gmres_options *options, vie_gmres_report *report)
{
  if(op && op->apply == hip_apply && find(op->context)) {
    hip_system *h = op->context;
    int ok = lib.solve(h->system, rhs, x, options, report);
    report_phases();
    return ok;
  }
  return __real_vie_gmres(op, rhs, x, options, report);
}

/* Last error text of the library ("" when none or not loaded). */
const char *vie_hip_dispatch_last_error(void)
{ return lib.handle ? lib.last_error() : ""; }
