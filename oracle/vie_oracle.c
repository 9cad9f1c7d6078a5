/*
 * CPU oracle for the dielectric VIE hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is a plain-C restatement of the reference algorithms that build
 * the Green-tensor kernel of the FFT volume operator.  It is compiled into
 * oracle/_build/liboracle.so and is only ever loaded by tests/, by the smoke
 * check and by bench.py's CPU-baseline leg, as the checker.  The product
 * library (nec2c_amd/lib/libvie_hip.so) never links or calls it.
 *
 * What is restated (reference file:line for each):
 *   - Gauss-Legendre rules                     vie_kernel.c:144-181, vie_fft.c:284-298
 *   - cuboid relation classifier / overlap     vie_kernel.c:85-142
 *   - difference-coordinate segment splitting  vie_kernel.c:318-407
 *   - tensor and Duffy sector quadrature       vie_kernel.c:183-316
 *   - segment products, radial-power grading   vie_kernel.c:409-458
 *   - face-pair reduction of curl curl         vie_kernel.c:460-523
 *   - order ladder and acceptance test         vie_kernel.c:525-641
 *   - far-pair triangular-weight Gauss rule    vie_fft.c:299-335
 *   - per-offset table used by the spectrum    vie_fft.c:338-374
 *
 * Pinning: tests/test_oracle_kernel.py checks this file against the six
 * independent NumPy tensors of the reference's tests/vie-reference.csv
 * (mirrored as tests/golden/vie_reference_tensors.json), against the static
 * cube self term diag(2/3), against reciprocity / translation / permutation
 * invariances, and entry-for-entry against oracle/_ref (the reference's own
 * vie_kernel.c and vie_fft.c compiled from source) where that build exists.
 */
#include <complex.h>
#include <float.h>
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846264338327950288
#define ORC_MAX_ORDER 64
#define ORC_MAX_PIECES 5
#define ORC_GEOM_TOL 64.0

typedef struct { double center[3], size[3]; } orc_box;

/* One 1-D piece of a difference coordinate u = a - b (vie_kernel.c:17-26). */
typedef struct {
  int overlap;          /* both a and b vary: weight is the overlap length */
  double lo, hi;        /* integration range of u */
  double amin, amax, bmin, bmax;
} orc_piece;

typedef struct {
  int dims;             /* number of integrated (varying) axes */
  orc_piece piece[3];
  double fixed2;        /* squared contribution of non-varying axes */
  double k;
  const double *node, *wt;
  int order, power;     /* radial grading power for the Duffy map */
} orc_reduced;

/* ---- Gauss-Legendre on [-1,1], Newton from the Tricomi start
        (vie_kernel.c:144-181). ---- */
static int orc_gauss_pm1(int n, double *x, double *w)
{
  int r;
  if(n < 2 || n > ORC_MAX_ORDER) return 0;
  for(r = 0; r < (n+1)/2; r++) {
    double z = cos(ORC_PI*((double)r+0.75)/((double)n+0.5)), z0, dp = 0.0;
    int it = 0;
    do {
      double p0 = 1.0, p1 = z; int d;
      z0 = z;
      for(d = 2; d <= n; d++) {
        double p2 = ((2.0*d-1.0)*z*p1-((double)d-1.0)*p0)/(double)d;
        p0 = p1; p1 = p2;
      }
      dp = (double)n*(z*p1-p0)/(z*z-1.0);
      z = z0-p1/dp;
      it++;
    } while(fabs(z-z0) > 4.0*DBL_EPSILON && it < 32);
    if(it == 32) return 0;
    x[r] = -z; x[n-1-r] = z;
    w[r] = w[n-1-r] = 2.0/((1.0-z*z)*dp*dp);
  }
  return 1;
}

/* ---- Gauss-Legendre mapped to [0,1] as used by the far-pair rule
        (vie_fft.c:284-298): nodes (1-z)/2 then (1+z)/2, unnormalised
        weights 1/((1-z^2) P'^2) which sum to 1. ---- */
static void orc_gauss_01(int n, double *x, double *w)
{
  int i, j;
  for(i = 0; i < (n+1)/2; i++) {
    double z = cos(ORC_PI*(i+.75)/(n+.5)), z0, dp = 0;
    do {
      double p = 1, q = 0;
      z0 = z;
      for(j = 1; j <= n; j++) { double pr = q; q = p; p = ((2*j-1)*z*q-(j-1)*pr)/j; }
      dp = n*(z*p-q)/(z*z-1); z = z0-p/dp;
    } while(fabs(z-z0) > 2e-15);
    x[i] = (1-z)/2; x[n-1-i] = (1+z)/2;
    w[i] = w[n-1-i] = 1/((1-z*z)*dp*dp);
  }
}

static double orc_max3(double a, double b, double c) { return fmax(a, fmax(b, c)); }

static complex double orc_green(double r, double k)
{ return cexp(-I*k*r)/(4.0*ORC_PI*r); }

static double orc_piece_weight(const orc_piece *p, double u)
{
  double lo, hi;
  if(!p->overlap) return 1.0;
  lo = fmax(p->amin, p->bmin+u); hi = fmin(p->amax, p->bmax+u);
  return hi > lo ? hi-lo : 0.0;
}

/* Plain tensor-product Gauss rule over the reduced box (vie_kernel.c:200-233). */
static complex double orc_tensor(const orc_reduced *R, int d, double *u, double acc)
{
  const orc_piece *p; double mid, half; int i; complex double s = 0.0;
  if(d < 0 || d >= R->dims || d >= 3) return 0.0;
  p = &R->piece[d]; mid = 0.5*(p->lo+p->hi); half = 0.5*(p->hi-p->lo);
  for(i = 0; i < R->order; i++) {
    double v = mid+half*R->node[i];
    double w = acc*half*R->wt[i]*orc_piece_weight(p, v);
    u[d] = v;
    if(d+1 == R->dims) {
      double r2 = R->fixed2; int a;
      for(a = 0; a < R->dims; a++) r2 += u[a]*u[a];
      s += w*orc_green(sqrt(r2), R->k);
    } else s += orc_tensor(R, d+1, u, w);
  }
  return s;
}

/* Duffy sector map about the corner nearest the origin, radial grading
   u_s -> u_s^power (vie_kernel.c:235-279). */
static complex double orc_duffy(const orc_reduced *R, int sector, int d, double *q, double acc)
{
  int i; complex double s = 0.0;
  for(i = 0; i < R->order; i++) {
    double uq = 0.5*(1.0+R->node[i]);
    double w = acc*0.5*R->wt[i];
    q[d] = uq;
    if(d+1 == R->dims) {
      double c[3] = {0, 0, 0}, rp = q[sector];
      double t = pow(rp, (double)R->power), r2 = R->fixed2, jac = 1.0, pw = 1.0;
      int a;
      for(a = 0; a < R->dims; a++) {
        const orc_piece *p = &R->piece[a];
        double ext = fabs(p->hi-p->lo);
        double org = fabs(p->lo) <= fabs(p->hi) ? p->lo : p->hi;
        double sg = org == p->lo ? 1.0 : -1.0;
        double m = a == sector ? t : t*q[a];
        c[a] = org+sg*ext*m;
        r2 += c[a]*c[a];
        jac *= ext;
        pw *= orc_piece_weight(p, c[a]);
      }
      jac *= pow(t, (double)(R->dims-1))*(double)R->power*pow(rp, (double)(R->power-1));
      s += w*jac*pw*orc_green(sqrt(r2), R->k);
    } else s += orc_duffy(R, sector, d+1, q, w);
  }
  return s;
}

/* Choose tensor or Duffy quadrature for one reduced box (vie_kernel.c:286-316). */
static complex double orc_reduced_box(const orc_reduced *R)
{
  double u[3] = {0, 0, 0}, near2 = R->fixed2, big = 0.0;
  int sing = 1, a;
  for(a = 0; a < R->dims; a++) {
    const orc_piece *p = &R->piece[a];
    double nr = fmin(fabs(p->lo), fabs(p->hi));
    near2 += nr*nr;
    big = fmax(big, p->hi-p->lo);
    sing &= (p->lo == 0.0 || p->hi == 0.0);
  }
  if(!sing && sqrt(near2) >= 0.1*big) return orc_tensor(R, 0, u, 1.0);
  {
    complex double s = 0.0;
    for(a = 0; a < R->dims; a++) { double q[3] = {0, 0, 0}; s += orc_duffy(R, a, 0, q, 1.0); }
    return s;
  }
}

/* Break list with coincidence merging (vie_kernel.c:318-345). */
static int orc_add_break(double *b, int *n, double v, double tol)
{
  int i;
  if(fabs(v) <= tol) v = 0.0;
  for(i = 0; i < *n; i++) if(fabs(v-b[i]) <= tol) return 1;
  if(*n >= 6) return 0;
  b[(*n)++] = v;
  return 1;
}

static void orc_sort(double *b, int n)
{
  int i;
  for(i = 1; i < n; i++) {
    double v = b[i]; int c = i;
    while(c > 0 && b[c-1] > v) { b[c] = b[c-1]; c--; }
    b[c] = v;
  }
}

/* Split one axis' difference coordinate into pieces (vie_kernel.c:347-407). */
static int orc_pieces(double amin, double amax, int av, double bmin, double bmax, int bv,
                      orc_piece *out, int *count, double *fixed)
{
  double br[6];
  double tol = ORC_GEOM_TOL*DBL_EPSILON*fmax(fmax(fabs(amin), fabs(amax)), fmax(fabs(bmin), fabs(bmax)));
  int nb = 0, i;
  *count = 0;
  memset(out, 0, ORC_MAX_PIECES*sizeof(*out));
  if(!av && !bv) {
    *fixed = amin-bmin;
    if(fabs(*fixed) <= tol) *fixed = 0.0;
    return 1;
  }
  if(av && bv) {
    if(!orc_add_break(br, &nb, amin-bmax, tol) || !orc_add_break(br, &nb, amin-bmin, tol) ||
       !orc_add_break(br, &nb, amax-bmax, tol) || !orc_add_break(br, &nb, amax-bmin, tol))
      return 0;
  } else if(av) {
    orc_add_break(br, &nb, amin-bmin, tol);
    orc_add_break(br, &nb, amax-bmin, tol);
  } else {
    orc_add_break(br, &nb, amin-bmax, tol);
    orc_add_break(br, &nb, amin-bmin, tol);
  }
  orc_sort(br, nb);
  if(br[0] < 0.0 && br[nb-1] > 0.0) { orc_add_break(br, &nb, 0.0, tol); orc_sort(br, nb); }
  for(i = 0; i+1 < nb; i++) {
    orc_piece *p;
    if(br[i+1] == br[i]) continue;
    if(*count >= ORC_MAX_PIECES) return 0;
    p = &out[(*count)++];
    p->overlap = av && bv;
    p->lo = br[i]; p->hi = br[i+1];
    p->amin = amin; p->amax = amax; p->bmin = bmin; p->bmax = bmax;
  }
  return 1;
}

/* Sum over the product of per-axis pieces with radial grading
   (vie_kernel.c:409-458). */
static complex double orc_piece_product(orc_piece P[3][ORC_MAX_PIECES], const int cnt[3],
                                        const double fixed[3], double k, int order,
                                        const double *node, const double *wt)
{
  complex double s = 0.0; int i, j, l;
  for(i = 0; i < (cnt[0] ? cnt[0] : 1); i++)
    for(j = 0; j < (cnt[1] ? cnt[1] : 1); j++)
      for(l = 0; l < (cnt[2] ? cnt[2] : 1); l++) {
        int sel[3] = {i, j, l}, a;
        orc_reduced R;
        double big = 0.0, near2, ratio;
        memset(&R, 0, sizeof(R));
        R.k = k; R.order = order; R.node = node; R.wt = wt; R.power = 1;
        for(a = 0; a < 3; a++) {
          if(cnt[a]) R.piece[R.dims++] = P[a][sel[a]];
          else R.fixed2 += fixed[a]*fixed[a];
        }
        near2 = R.fixed2;
        for(a = 0; a < R.dims; a++) {
          double nr = fmin(fabs(R.piece[a].lo), fabs(R.piece[a].hi));
          near2 += nr*nr;
          big = fmax(big, R.piece[a].hi-R.piece[a].lo);
        }
        ratio = sqrt(near2)/big;
        while(ratio > 0.0 && ratio < 0.1 && R.power < 8) { R.power++; ratio *= 10.0; }
        s += orc_reduced_box(&R);
      }
  return s;
}

/* Integral of g over one test face x one source face (vie_kernel.c:460-489). */
static complex double orc_face_pair(const orc_box *t, const orc_box *s, int ta, int ts,
                                    int sa, int ss, double k, int order,
                                    const double *node, const double *wt)
{
  orc_piece P[3][ORC_MAX_PIECES]; int cnt[3]; double fixed[3] = {0, 0, 0}; int a;
  for(a = 0; a < 3; a++) {
    int tv = a != ta, sv = a != sa;
    double tmin = tv ? -0.5*t->size[a] : 0.5*ts*t->size[a];
    double tmax = tv ? 0.5*t->size[a] : tmin;
    double sc = s->center[a]-t->center[a];
    double smin = sv ? sc-0.5*s->size[a] : sc+0.5*ss*s->size[a];
    double smax = sv ? sc+0.5*s->size[a] : smin;
    if(!orc_pieces(tmin, tmax, tv, smin, smax, sv, P[a], &cnt[a], &fixed[a]))
      return NAN+I*NAN;
  }
  return orc_piece_product(P, cnt, fixed, k, order, node, wt);
}

/* Galerkin N at one quadrature order by the face-pair reduction
   (vie_kernel.c:491-523). */
static int orc_matrix_order(const orc_box *t, const orc_box *s, double k, int order,
                            complex double out[3][3])
{
  double node[ORC_MAX_ORDER], wt[ORC_MAX_ORDER];
  double vol = t->size[0]*t->size[1]*t->size[2];
  int ta, sa, ts, ss;
  memset(out, 0, 9*sizeof(complex double));
  if(!orc_gauss_pm1(order, node, wt)) return 0;
  for(ta = 0; ta < 3; ta++)
    for(sa = 0; sa < 3; sa++)
      for(ts = -1; ts <= 1; ts += 2)
        for(ss = -1; ss <= 1; ss += 2) {
          complex double v = orc_face_pair(t, s, ta, ts, sa, ss, k, order, node, wt)/vol;
          if(!isfinite(creal(v)) || !isfinite(cimag(v))) return 0;
          if(ta == sa) {
            int c;
            for(c = 0; c < 3; c++) if(c != ta) out[c][c] += (double)(ts*ss)*v;
          } else out[sa][ta] -= (double)(ts*ss)*v;
        }
  return 1;
}

static int orc_valid(const orc_box *b)
{
  int a;
  if(!b) return 0;
  for(a = 0; a < 3; a++)
    if(!isfinite(b->center[a]) || !isfinite(b->size[a]) || b->size[a] <= 0.0) return 0;
  return 1;
}

/* 0 self, 1 face, 2 edge, 3 corner, 4 near, 5 far (vie_kernel.c:85-123). */
int oracle_relation(const double tc[3], const double tsz[3], const double sc[3], const double ssz[3])
{
  int a, same = 1, sep = 0, touch = 0; double d2 = 0.0;
  double scale = fmax(orc_max3(tsz[0], tsz[1], tsz[2]), orc_max3(ssz[0], ssz[1], ssz[2]));
  for(a = 0; a < 3; a++) {
    double cd = fabs(tc[a]-sc[a]), hs = 0.5*(tsz[a]+ssz[a]);
    double tol = ORC_GEOM_TOL*DBL_EPSILON*fmax(1.0, fmax(cd, hs)), gap = cd-hs;
    if(cd > tol || fabs(tsz[a]-ssz[a]) > tol) same = 0;
    if(gap > tol) { sep++; d2 += gap*gap; }
    else if(fabs(gap) <= tol) touch++;
  }
  if(same) return 0;
  if(sep == 0) return touch == 1 ? 1 : touch == 2 ? 2 : touch == 3 ? 3 : 4;
  return sqrt(d2) > 2.0*scale ? 5 : 4;
}

static int orc_overlap(const orc_box *t, const orc_box *s)
{
  int a, same = 1, ov = 1;
  for(a = 0; a < 3; a++) {
    double cd = fabs(t->center[a]-s->center[a]), hs = 0.5*(t->size[a]+s->size[a]);
    double tol = ORC_GEOM_TOL*DBL_EPSILON*fmax(1.0, fmax(cd, hs));
    same &= cd <= tol && fabs(t->size[a]-s->size[a]) <= tol;
    ov &= cd < hs-tol;
  }
  return ov && !same;
}

/*
 * Volume-averaged Galerkin tensor of curl curl (integral of g) between two
 * axis-aligned cuboids, with the reference's order ladder {4,...,64}
 * (vie_kernel.c:550-641).  out is row-major 3x3 complex (18 doubles).
 * Returns 1 on success; *order_used receives the accepted order.
 */
int oracle_cuboid_n(const double tc[3], const double tsz[3], const double sc[3], const double ssz[3],
                    double k, double rtol, double atol, int max_order, int far_order,
                    double *out, int *order_used)
{
  static const int ladder[] = {4, 6, 8, 12, 16, 24, 32, 40, 48, 56, 64};
  orc_box t, s; complex double prev[3][3], cur[3][3]; int have = 0, i, rel;
  memcpy(t.center, tc, sizeof(t.center)); memcpy(t.size, tsz, sizeof(t.size));
  memcpy(s.center, sc, sizeof(s.center)); memcpy(s.size, ssz, sizeof(s.size));
  if(order_used) *order_used = 0;
  if(!orc_valid(&t) || !orc_valid(&s) || !isfinite(k) || k < 0.0 || !out) return 0;
  if(k*orc_max3(t.size[0], t.size[1], t.size[2]) > 1.0+ORC_GEOM_TOL*DBL_EPSILON ||
     k*orc_max3(s.size[0], s.size[1], s.size[2]) > 1.0+ORC_GEOM_TOL*DBL_EPSILON) return 0;
  if(orc_overlap(&t, &s)) return 0;
  if(!isfinite(rtol) || !isfinite(atol) || rtol <= 0 || atol <= 0 || max_order < 4 ||
     max_order > ORC_MAX_ORDER || far_order < 2 || far_order > max_order) return 0;
  rel = oracle_relation(tc, tsz, sc, ssz);
  if(rel == 5) {
    if(!orc_matrix_order(&t, &s, k, far_order, cur)) return 0;
    memcpy(out, cur, sizeof(cur));
    if(order_used) *order_used = far_order;
    return 1;
  }
  for(i = 0; i < (int)(sizeof(ladder)/sizeof(ladder[0])); i++) {
    int order = ladder[i];
    if(order > max_order) break;
    if(!orc_matrix_order(&t, &s, k, order, cur)) return 0;
    if(have) {
      double worst = 0.0; int e;
      for(e = 0; e < 9; e++) {
        complex double l = (&prev[0][0])[e], r = (&cur[0][0])[e];
        double sc2 = atol+rtol*fmax(cabs(l), cabs(r)), q = cabs(l-r)/sc2;
        if(q > worst) worst = q;
      }
      if(worst <= 1.0) {
        memcpy(out, cur, sizeof(cur));
        if(order_used) *order_used = order;
        return 1;
      }
    }
    memcpy(prev, cur, sizeof(prev));
    have = 1;
  }
  return 0;
}

/* Far-pair difference-coordinate rule at one order (vie_fft.c:299-318). */
static void orc_far_order(const long v[3], double h, double k, int order, complex double out[3][3])
{
  double x[12], w[12]; int a, b, c, m, i, j;
  orc_gauss_01(order, x, w);
  memset(out, 0, 9*sizeof(complex double));
  for(a = 0; a < order; a++) for(b = 0; b < order; b++) for(c = 0; c < order; c++) {
    double t[3] = {x[a], x[b], x[c]};
    double wt = h*h*h*w[a]*w[b]*w[c]*(1-t[0])*(1-t[1])*(1-t[2]);
    for(m = 0; m < 8; m++) {
      double r[3], r2 = 0, rr; complex double g, dg, rad;
      for(i = 0; i < 3; i++) { r[i] = h*((double)v[i]+((m & (1<<i)) ? -t[i] : t[i])); r2 += r[i]*r[i]; }
      rr = sqrt(r2); g = wt*cexp(-I*k*rr)/(4*ORC_PI*rr);
      dg = g*(k*k-I*k/rr-1/r2);
      rad = g*(-k*k+3*I*k/rr+3/r2)/r2;
      for(i = 0; i < 3; i++) for(j = i; j < 3; j++) out[i][j] += (i == j ? dg : 0)+rad*r[i]*r[j];
    }
  }
  for(i = 0; i < 3; i++) for(j = i+1; j < 3; j++) out[j][i] = out[i][j];
}

/* Far-pair ladder {4,6,8,12}; 0 when the pair is not far or does not settle
   (vie_fft.c:319-335).  out: 18 doubles. */
int oracle_far_kernel(const long v[3], double h, double k, double *out)
{
  static const int orders[] = {4, 6, 8, 12};
  complex double prev[3][3], cur[3][3]; double gap2 = 0; int st, i, j;
  for(i = 0; i < 3; i++) { double g = v[i] > 1 ? (double)v[i]-1 : 0; gap2 += g*g; }
  if(gap2 <= 4) return 0;
  orc_far_order(v, h, k, orders[0], prev);
  for(st = 1; st < 4; st++) {
    int ok = 1;
    orc_far_order(v, h, k, orders[st], cur);
    for(i = 0; i < 3; i++) for(j = 0; j < 3; j++)
      if(!isfinite(creal(cur[i][j])) || !isfinite(cimag(cur[i][j])) ||
         cabs(cur[i][j]-prev[i][j]) > 2e-12+2e-9*cabs(cur[i][j])) ok = 0;
    if(ok) { memcpy(out, cur, sizeof(cur)); return 1; }
    memcpy(prev, cur, sizeof(prev));
  }
  return 0;
}

/*
 * Per sorted offset v = (a >= b >= c) the symmetrised tensor that the
 * spectrum build spreads over sign/permutation images (vie_fft.c:338-362):
 * far rule first, else the cuboid integral at rtol 5e-7, atol 5e-10,
 * far order 16; off-diagonals averaged and zeroed on a zero offset axis.
 * offsets: count x 3 longs; out: count x 18 doubles.  Returns the number of
 * offsets done (== count on success).
 */
long oracle_offset_table(const long *offsets, long count, double h, double k, double *out)
{
  long n, failed = count;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4)
#endif
  for(n = 0; n < count; n++) {
    const long *v = offsets+3*n;
    complex double T[3][3]; int i, j, ok = 1;
    if(!oracle_far_kernel(v, h, k, (double *)T)) {
      double tc[3], tsz[3] = {h, h, h}, sc[3] = {0, 0, 0};
      for(i = 0; i < 3; i++) tc[i] = h*(double)v[i];
      ok = oracle_cuboid_n(tc, tsz, sc, tsz, k, 5e-7, 5e-10, 64, 16, (double *)T, NULL);
    }
    if(!ok) {
#ifdef _OPENMP
#pragma omp critical
#endif
      if(n < failed) failed = n;
      continue;
    }
    for(i = 0; i < 3; i++) for(j = i+1; j < 3; j++) {
      complex double avg = (T[i][j]+T[j][i])*0.5;
      T[i][j] = T[j][i] = (v[i] && v[j]) ? avg : 0;
    }
    memcpy(out+18*n, T, sizeof(T));
  }
  return failed;
}
