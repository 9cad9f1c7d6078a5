/* CPU oracle of the mesher's relationship checks -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/ and bench.py's CPU-baseline leg load this code (through
 * liboracle.so), and only as the checker; the product library never uses it.
 *
 * Plain-C restatement, loop for loop, of
 *   validate_box_relationships  dielectric_mesh.c:75-123 (interior overlap of
 *                               every box pair, first pair in loop order;
 *                               the region lookups are left out)
 *   evaluate_clearance          dielectric_mesh.c:125-316 (segment-box and
 *                               point-box distances, worst margin and
 *                               smallest gap over boxes x wires x patches)
 * Parity status: pinned in tests/test_mesh_check_oracle.py against the
 * reference's own dielectric_mesh_build (oracle/_ref/libmeshref.so), the
 * reference tests' overlap / touching / clearance cases, and brute-force
 * distances by dense sampling.
 */
#include <float.h>
#include <math.h>
#include <stddef.h>

#define ORACLE_MESH_TOLERANCE 1.0e-12 /* dielectric_mesh.c:8 */
#define ORACLE_NEC_PI 3.141592654     /* nec2c.h */

static int meet(double amin, double amax, double bmin, double bmax) { return fmax(amin, bmin) < fmin(amax, bmax); }

/* first pair (left, right) with overlapping interiors, or -1, -1 */
void oracle_box_overlap(const double *mn, const double *mx, size_t count, size_t first_left, size_t last_left,
                        long long *left, long long *right) {
  *left = *right = -1;
  for (size_t i = first_left; i < count && i < last_left; i++)
    for (size_t j = i + 1; j < count; j++) {
      int in = 1;
      for (int a = 0; a < 3; a++) in &= meet(mn[3 * i + a], mx[3 * i + a], mn[3 * j + a], mx[3 * j + a]);
      if (in) {
        *left = (long long)i;
        *right = (long long)j;
        return;
      }
    }
}

/* squared distance from segment s..e to the box of centre c, size h
   (dielectric_mesh.c:141-213) */
double oracle_segment_box_d2(const double s[3], const double e[3], const double c[3], const double h[3]) {
  double br[8], best = DBL_MAX;
  int nb = 2;
  br[0] = 0.0;
  br[1] = 1.0;
  for (int a = 0; a < 3; a++) {
    double d = e[a] - s[a], lo = c[a] - 0.5 * h[a], hi = c[a] + 0.5 * h[a];
    if (d != 0.0) {
      double p = (lo - s[a]) / d, q = (hi - s[a]) / d;
      if (p > 0.0 && p < 1.0) br[nb++] = p;
      if (q > 0.0 && q < 1.0) br[nb++] = q;
    }
  }
  for (int k = 1; k < nb; k++) {
    double v = br[k];
    int m = k;
    while (m > 0 && br[m - 1] > v) {
      br[m] = br[m - 1];
      m--;
    }
    br[m] = v;
  }
  for (int k = 0; k + 1 < nb; k++) {
    double lower = br[k], upper = br[k + 1], mid = 0.5 * (lower + upper), qa = 0.0, qb = 0.0, qc = 0.0, u, v;
    for (int a = 0; a < 3; a++) {
      double d = e[a] - s[a], at = s[a] + d * mid, lo = c[a] - 0.5 * h[a], hi = c[a] + 0.5 * h[a], bound;
      if (at >= lo && at <= hi) continue;
      bound = at < lo ? lo : hi;
      qa += d * d;
      qb += 2.0 * d * (s[a] - bound);
      qc += (s[a] - bound) * (s[a] - bound);
    }
    u = qa > 0.0 ? -qb / (2.0 * qa) : lower;
    if (u < lower) u = lower;
    if (u > upper) u = upper;
    v = qa * u * u + qb * u + qc;
    if (v < best) best = v;
  }
  return best < 0.0 && best > -ORACLE_MESH_TOLERANCE ? 0.0 : best;
}

static double point_box_d2(const double p[3], const double c[3], const double h[3]) {
  double d2 = 0.0;
  for (int a = 0; a < 3; a++) {
    double lo = c[a] - 0.5 * h[a], hi = c[a] + 0.5 * h[a];
    double d = p[a] < lo ? lo - p[a] : p[a] > hi ? p[a] - hi : 0.0;
    d2 += d * d;
  }
  return d2;
}

/* out[0..4] = worst margin, worst gap, worst required, minimum gap, its
   required; idx[0..2] = worst box (1-based), worst conductor (1-based), is
   patch.  Boxes [first_box, last_box) only (a bounded sample). */
void oracle_clearance(const double *mn, const double *mx, const double *cell_longest, size_t boxes, size_t first_box,
                      size_t last_box, int n, const double *x, const double *y, const double *z, const double *si,
                      const double *cab, const double *sab, const double *salp, const double *bi, int m,
                      const double *pts, const double *area, double *out, long long *idx) {
  double worst = DBL_MAX, wgap = 0.0, wreq = 0.0, mgap = DBL_MAX, mreq = 0.0;
  long long wbox = 0, wcond = 0, wpatch = 0;
  for (size_t b = first_box; b < boxes && b < last_box; b++) {
    double c[3], h[3];
    if (!(cell_longest[b] > 0.0)) continue;
    for (int a = 0; a < 3; a++) {
      c[a] = 0.5 * (mn[3 * b + a] + mx[3 * b + a]);
      h[a] = mx[3 * b + a] - mn[3 * b + a];
    }
    for (int k = 0; k < n; k++) {
      double s[3] = {x[k] - 0.5 * si[k] * cab[k], y[k] - 0.5 * si[k] * sab[k], z[k] - 0.5 * si[k] * salp[k]};
      double e[3] = {x[k] + 0.5 * si[k] * cab[k], y[k] + 0.5 * si[k] * sab[k], z[k] + 0.5 * si[k] * salp[k]};
      double gap = sqrt(oracle_segment_box_d2(s, e, c, h)) - bi[k], req;
      if (gap < 0.0) gap = 0.0;
      req = fmax(cell_longest[b], si[k]);
      if (gap - req < worst) {
        worst = gap - req;
        wgap = gap;
        wreq = req;
        wbox = (long long)b + 1;
        wcond = k + 1;
        wpatch = 0;
      }
      if (gap < mgap) {
        mgap = gap;
        mreq = req;
      }
    }
    for (int k = 0; k < m; k++) {
      double p[3] = {pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]};
      double radius = sqrt(area[k] / ORACLE_NEC_PI);
      double gap = sqrt(point_box_d2(p, c, h)) - radius, req = fmax(cell_longest[b], sqrt(area[k]));
      if (gap < 0.0) gap = 0.0;
      if (gap - req < worst) {
        worst = gap - req;
        wgap = gap;
        wreq = req;
        wbox = (long long)b + 1;
        wcond = k + 1;
        wpatch = 1;
      }
      if (gap < mgap) {
        mgap = gap;
        mreq = req;
      }
    }
  }
  out[0] = worst;
  out[1] = wgap;
  out[2] = wreq;
  out[3] = mgap;
  out[4] = mreq;
  idx[0] = wbox;
  idx[1] = wcond;
  idx[2] = wpatch;
}
