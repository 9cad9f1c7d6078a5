/*
 * CPU oracle for the conductor-to-volume coupling -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the field that NEC's classic (NEC-2) wire engine
 * radiates from frozen basis coefficients, averaged over each volume cell,
 * which fills the volume rows of the conductor columns of the coupled
 * system.  It is compiled into oracle/_build/liboracle.so and is only loaded
 * by tests/ (as the checker) and by bench.py's CPU-baseline leg; the product
 * library never links or calls it.
 *
 * What is restated (reference file:line for each):
 *   - gf, test                    calculations.c:497-520, :913-935
 *   - intx (Romberg, ij != 0)     calculations.c:734-896
 *   - gx                          fields.c:581-595
 *   - eksc (thin-wire kernel)     fields.c:349-394
 *   - efld, classic engine, free
 *     space and PEC image         fields.c:38-235 (ksymp 1, or 2 with iperf 1)
 *   - unere (patch field)         fields.c:1713-1760 (same grounds)
 *   - nefld (ax search, segment
 *     and patch sums, iexk 0)     fields.c:1123-1370
 *   - 6^3 cell average            dw_average_wire_field, dielectric_wire.c:316-357
 *   - column x cell tasks         dw_wire_to_volume_range, dielectric_wire.c:399-420
 * NEC's own rounded constants (nec2c.h:36-59: PI, TP, ETA, CONST1, CONST2)
 * are kept as NEC defines them, since the fields depend on them.
 *
 * Pinning: tests/test_wire_volume_oracle.py checks the thin-wire kernel
 * against brute-force quadrature of the reduced-kernel field of sine, cosine
 * and constant currents, the lumped element against a short dipole, the
 * patch field against an elementary current, image theory over a PEC plane,
 * the Romberg integral against a fixed high-order rule, and the cell average
 * against numpy's Gauss-Legendre rule.
 */
#include <complex.h>
#include <float.h>
#include <math.h>
#include <stddef.h>
#include <string.h>

/* NEC's constants as nec2c.h defines them */
#define NEC_PI 3.141592654
#define NEC_TP 6.283185308
#define NEC_ETA 376.73
#define NEC_CONST1 (4.771341189 * I)
#define NEC_CONST2 4.771341188
#define ORC_EPS0 8.8541878128e-12
#define ORC_C0 299792458.0
#define ORC_PIF 3.141592653589793238462643383279502884

typedef struct {
  int n, m;
  const double *x, *y, *z, *si, *bi, *cab, *sab, *salp;
  const double *pp, *pbi, *t1, *t2; /* patch centres, areas, tangents [3 m] */
  double rkh;
  int ground; /* 0 free space, 1 PEC plane */
} wv_geom;

typedef struct {
  double zpk, rkb2;
} wv_kernel;

/* exp(-i rk) / rk at the Romberg abscissa, field-point form (ij != 0) */
static void wv_gf(const wv_kernel *k, double zk, double *co, double *si)
{
  const double d = zk - k->zpk, rk = sqrt(k->rkb2 + d * d);
  *si = sin(rk) / rk;
  *co = cos(rk) / rk;
}

static void wv_test(double f1r, double f2r, double *tr, double f1i, double f2i, double *ti)
{
  double den = fabs(f2r);
  if (den < fabs(f2i)) den = fabs(f2i);
  if (den < 1.0e-37) {
    *tr = *ti = 0.0;
    return;
  }
  *tr = fabs((f1r - f2r) / den);
  *ti = fabs((f1i - f2i) / den);
}

/* Romberg integral of gf over [el1, el2] with NEC's step control: 3- then
   5-point estimates, step halving down to s / 65536, doubling after four
   accepted steps; cint + i sint. */
static void wv_intx(const wv_kernel *k, double el1, double el2, double *cint, double *sint)
{
  const int nma = 65536;
  const double rx = 1.0e-4, s = el2 - el1, ep = s / (10.0 * nma), zend = el2 - ep;
  double z = el1, sr = 0.0, si = 0.0, dz = 0.0, dzot = 0.0;
  double g1r, g1i, g3r = 0, g3i = 0, g5r = 0, g5i = 0;
  int ns = 1, nt = 0, fresh = 1;
  wv_gf(k, z, &g1r, &g1i);
  for (;;) {
    if (fresh) {
      dz = s / ns;
      if (z + dz > el2) {
        dz = el2 - z;
        if (fabs(dz) <= ep) break;
      }
      dzot = 0.5 * dz;
      wv_gf(k, z + dzot, &g3r, &g3i);
      wv_gf(k, z + dz, &g5r, &g5i);
    }
    const double t00r = (g1r + g5r) * dzot, t00i = (g1i + g5i) * dzot;
    const double t01r = (t00r + dz * g3r) * 0.5, t01i = (t00i + dz * g3i) * 0.5;
    const double t10r = (4.0 * t01r - t00r) / 3.0, t10i = (4.0 * t01i - t00i) / 3.0;
    double er, ei;
    wv_test(t01r, t10r, &er, t01i, t10i, &ei);
    double addr, addi;
    int steps;
    if (ei <= rx && er <= rx) {
      addr = t10r;
      addi = t10i;
      steps = 2;
    } else {
      double g2r, g2i, g4r, g4i;
      wv_gf(k, z + 0.25 * dz, &g2r, &g2i);
      wv_gf(k, z + 0.75 * dz, &g4r, &g4i);
      const double t02r = (t01r + dzot * (g2r + g4r)) * 0.5, t02i = (t01i + dzot * (g2i + g4i)) * 0.5;
      const double t11r = (4.0 * t02r - t01r) / 3.0, t11i = (4.0 * t02i - t01i) / 3.0;
      const double t20r = (16.0 * t11r - t10r) / 15.0, t20i = (16.0 * t11i - t10i) / 15.0;
      wv_test(t11r, t20r, &er, t11i, t20i, &ei);
      if (ei > rx || er > rx) {
        nt = 0;
        if (ns < nma) {
          /* halve the step; the old midpoint and quarter point carry over */
          ns *= 2;
          dz = s / ns;
          dzot = 0.5 * dz;
          g5r = g3r;
          g5i = g3i;
          g3r = g2r;
          g3i = g2i;
          fresh = 0;
          continue;
        }
        /* step size limit reached: NEC accepts the 5-point value */
      }
      addr = t20r;
      addi = t20i;
      steps = 1;
    }
    sr += addr;
    si += addi;
    nt += steps;
    z += dz;
    if (z >= zend) break;
    g1r = g5r;
    g1i = g5i;
    if (nt >= 4 && ns > 1) {
      ns /= 2;
      nt = 1;
    }
    fresh = 1;
  }
  *cint = sr;
  *sint = si;
}

static void wv_gx(double zz, double rh, double xk, complex double *gz, complex double *gzp)
{
  const double r2 = zz * zz + rh * rh, r = sqrt(r2), rkz = xk * r;
  *gz = (cos(rkz) - I * sin(rkz)) / r;
  *gzp = -(1.0 + I * rkz) * *gz / r2;
}

/* axial (ez*) and radial (er*) fields of sine, cosine and constant current
   filaments on a segment of length s, observed at axial offset z and reduced
   radius rh (all in wavelengths, xk = 2 pi) */
static void wv_eksc(double s, double z, double rh, double xk, complex double e[6])
{
  wv_kernel k = {xk * z, (xk * rh) * (xk * rh)};
  const double sh = 0.5 * s, shk = xk * sh, ss = sin(shk), cs = cos(shk);
  const double z2a = sh - z, z1a = -(sh + z);
  complex double gz1, gz2, gp1, gp2;
  wv_gx(z1a, rh, xk, &gz1, &gp1);
  wv_gx(z2a, rh, xk, &gz2, &gp2);
  complex double gzp1 = gp1 * z1a, gzp2 = gp2 * z2a;
  double cint, sint;
  /* e = {ezs, ers, ezc, erc, ezk, erk} */
  e[0] = NEC_CONST1 * ((gz2 - gz1) * cs * xk - (gzp2 + gzp1) * ss);
  e[2] = -NEC_CONST1 * ((gz2 + gz1) * ss * xk + (gzp2 - gzp1) * cs);
  e[5] = NEC_CONST1 * (gp2 - gp1) * rh;
  wv_intx(&k, -shk, shk, &cint, &sint);
  e[4] = -NEC_CONST1 * (gzp2 - gzp1 + xk * xk * (cint - I * sint));
  gzp1 *= z1a;
  gzp2 *= z2a;
  if (rh >= 1.0e-10) {
    e[1] = -NEC_CONST1 * ((gzp2 + gzp1 + gz2 + gz1) * ss - (z2a * gz2 - z1a * gz1) * cs * xk) / rh;
    e[3] = -NEC_CONST1 * ((gzp2 - gzp1 + gz2 - gz1) * cs + (z2a * gz2 + z1a * gz1) * ss * xk) / rh;
  } else {
    e[1] = e[3] = 0.0;
  }
}

/* field at (xi, yi, zi) of segment j's constant (k), sine (s) and cosine (c)
   currents: f[0..2] k, f[3..5] s, f[6..8] c (x, y, z each); ai the kernel
   radius (fields.c:38-235, classic engine) */
static void wv_efld(const wv_geom *g, int j, double xi, double yi, double zi, double ai, complex double f[9])
{
  const double xij = xi - g->x[j], yij = yi - g->y[j];
  const double cabj = g->cab[j], sabj = g->sab[j], s = g->si[j];
  for (int ip = 0; ip < (g->ground ? 2 : 1); ip++) {
    const double rfl = ip ? -1.0 : 1.0;
    const double salpr = g->salp[j] * rfl, zij = zi - rfl * g->z[j];
    const double zp = xij * cabj + yij * sabj + zij * salpr;
    double rhox = xij - cabj * zp, rhoy = yij - sabj * zp, rhoz = zij - salpr * zp;
    const double rh = sqrt(rhox * rhox + rhoy * rhoy + rhoz * rhoz + ai * ai);
    if (rh <= 1.0e-10) {
      rhox = rhoy = rhoz = 0.0;
    } else {
      rhox /= rh;
      rhoy /= rh;
      rhoz /= rh;
    }
    const double r = sqrt(zp * zp + rh * rh);
    complex double tezk, terk, tezc, terc, t[9];
    if (r >= g->rkh) {
      /* lumped current element beyond rkh wavelengths */
      double rmag = NEC_TP * r;
      const double cth = zp / r, px = rh / r, py = NEC_TP * r * r;
      const complex double e0 = cos(rmag) - I * sin(rmag);
      const complex double ty = NEC_ETA * cth * e0 * (1.0 - I / rmag) / py;
      const complex double tz = NEC_ETA * px * e0 * (1.0 + I * (rmag - 1.0 / rmag)) / (2.0 * py);
      tezk = ty * cth - tz * px;
      terk = ty * px + tz * cth;
      rmag = sin(NEC_PI * s) / NEC_PI;
      tezc = tezk * rmag;
      terc = terk * rmag;
      tezk *= s;
      terk *= s;
      t[3] = t[4] = t[5] = 0.0;
    } else {
      complex double e[6];
      wv_eksc(s, zp, rh, NEC_TP, e);
      tezk = e[4];
      terk = e[5];
      tezc = e[2];
      terc = e[3];
      t[3] = e[0] * cabj + e[1] * rhox;
      t[4] = e[0] * sabj + e[1] * rhoy;
      t[5] = e[0] * salpr + e[1] * rhoz;
    }
    t[0] = tezk * cabj + terk * rhox;
    t[1] = tezk * sabj + terk * rhoy;
    t[2] = tezk * salpr + terk * rhoz;
    t[6] = tezc * cabj + terc * rhox;
    t[7] = tezc * sabj + terc * rhoy;
    t[8] = tezc * salpr + terc * rhoz;
    for (int q = 0; q < 9; q++) f[q] = ip ? f[q] - t[q] : t[q]; /* PEC: frati = 1 */
  }
}

/* field of patch i's two surface-current components (t1: f[0..2], t2:
   f[3..5]) at the point; image = 1 for the mirrored patch over PEC */
static void wv_unere(const wv_geom *g, int i, int image, double xob, double yob, double zob, complex double f[6])
{
  const double *c = g->pp + 3 * i, *t1 = g->t1 + 3 * i, *t2 = g->t2 + 3 * i;
  const double sgn = image ? -1.0 : 1.0;
  const double zr = sgn * c[2], t1zr = sgn * t1[2], t2zr = sgn * t2[2];
  const double rx = xob - c[0], ry = yob - c[1], rz = zob - zr;
  const double r2 = rx * rx + ry * ry + rz * rz;
  if (r2 <= 1.0e-20) {
    for (int q = 0; q < 6; q++) f[q] = 0.0;
    return;
  }
  const double r = sqrt(r2), tt1 = -NEC_TP * r, tt2 = tt1 * tt1, rt = r2 * r;
  const complex double er = (sin(tt1) - I * cos(tt1)) * (NEC_CONST2 * g->pbi[i]);
  const complex double q1 = ((tt2 - 1.0) + I * tt1) * er / rt;
  const complex double q2 = ((3.0 - tt2) - 3.0 * I * tt1) * er / (rt * r2);
  complex double e1 = q2 * (t1[0] * rx + t1[1] * ry + t1zr * rz);
  f[0] = q1 * t1[0] + e1 * rx;
  f[1] = q1 * t1[1] + e1 * ry;
  f[2] = q1 * t1zr + e1 * rz;
  e1 = q2 * (t2[0] * rx + t2[1] * ry + t2zr * rz);
  f[3] = q1 * t2[0] + e1 * rx;
  f[4] = q1 * t2[1] + e1 * ry;
  f[5] = q1 * t2zr + e1 * rz;
  if (image)
    for (int q = 0; q < 6; q++) f[q] = -f[q]; /* PEC */
}

/* NEC's near field (wavelength units) of the frozen source coefficients:
   src[3 j .. 3 j + 2] = (a, b, c) of segment j, src[3 n + 3 i ..] the
   patch current of patch i */
static void wv_nefld(const wv_geom *g, const complex double *src, double xob, double yob, double zob,
                     complex double e[3])
{
  e[0] = e[1] = e[2] = 0.0;
  double ax = 0.0;
  for (int i = 0; i < g->n; i++) {
    const double xj = xob - g->x[i], yj = yob - g->y[i], zj = zob - g->z[i];
    const double zp = g->cab[i] * xj + g->sab[i] * yj + g->salp[i] * zj;
    if (fabs(zp) > 0.5001 * g->si[i]) continue;
    if (xj * xj + yj * yj + zj * zj - zp * zp > 0.9 * g->bi[i] * g->bi[i]) continue;
    ax = g->bi[i];
    break;
  }
  for (int i = 0; i < g->n; i++) {
    const complex double a = src[3 * i], b = src[3 * i + 1], c = src[3 * i + 2];
    if (a == 0.0 && b == 0.0 && c == 0.0) continue;
    complex double f[9];
    wv_efld(g, i, xob, yob, zob, ax, f);
    for (int p = 0; p < 3; p++) e[p] += f[p] * a + f[3 + p] * b + f[6 + p] * c;
  }
  for (int i = 0; i < g->m; i++) {
    const complex double *cur = src + 3 * g->n + 3 * i;
    const double *t1 = g->t1 + 3 * i, *t2 = g->t2 + 3 * i;
    const complex double acx = t1[0] * cur[0] + t1[1] * cur[1] + t1[2] * cur[2];
    const complex double bcx = t2[0] * cur[0] + t2[1] * cur[1] + t2[2] * cur[2];
    for (int image = 0; image < (g->ground ? 2 : 1); image++) {
      complex double f[6];
      wv_unere(g, i, image, xob, yob, zob, f);
      for (int p = 0; p < 3; p++) e[p] += acx * f[p] + bcx * f[3 + p];
    }
  }
}

static int wv_gauss(int order, double *x, double *w)
{
  for (int root = 0; root < (order + 1) / 2; ++root) {
    double z = cos(ORC_PIF * (root + 0.75) / (order + 0.5)), previous, derivative = 0.0;
    int it = 0;
    do {
      double p0 = 1.0, p1 = z;
      for (int deg = 2; deg <= order; ++deg) {
        double p2 = ((2.0 * deg - 1.0) * z * p1 - (deg - 1.0) * p0) / deg;
        p0 = p1;
        p1 = p2;
      }
      derivative = order * (z * p1 - p0) / (z * z - 1.0);
      previous = z;
      z = previous - p1 / derivative;
      ++it;
    } while (fabs(z - previous) > 4.0 * DBL_EPSILON && it < 32);
    if (it == 32) return 0;
    x[root] = -z;
    x[order - 1 - root] = z;
    w[root] = 2.0 / ((1.0 - z * z) * derivative * derivative);
    w[order - 1 - root] = w[root];
  }
  return 1;
}

/* Near field of one source state at a point given in metres (the public
   check of the restated NEC routines). */
int oracle_nec_near_field(int n, const double *x, const double *y, const double *z, const double *si,
                          const double *bi, const double *cab, const double *sab, const double *salp, int m,
                          const double *pp, const double *pbi, const double *t1, const double *t2, double rkh,
                          int ground, double wlam, const double complex *src, const double point[3],
                          double complex e[3])
{
  const wv_geom g = {n, m, x, y, z, si, bi, cab, sab, salp, pp, pbi, t1, t2, rkh, ground};
  wv_nefld(&g, src, point[0] / wlam, point[1] / wlam, point[2] / wlam, e);
  return 1;
}

/* One thin-wire kernel evaluation (the public check of eksc): e = {ezs,
   ers, ezc, erc, ezk, erk} */
void oracle_nec_eksc(double s, double z, double rh, double complex e[6]) { wv_eksc(s, z, rh, NEC_TP, e); }

/* Romberg integral of exp(-i rk)/rk, rk = sqrt(rkb2 + (t - zpk)^2), over
   [el1, el2] (the public check of intx). */
void oracle_nec_intx(double zpk, double rkb2, double el1, double el2, double *cint, double *sint)
{
  const wv_kernel k = {zpk, rkb2};
  wv_intx(&k, el1, el2, cint, sint);
}

/*
 * coupling[column * 3 cells + 3 cell + q] = -i omega eps0 chi[cell] <E_q>,
 * <E> the 6^3-point average over the cell of the field of column `column`'s
 * sources (sources + column * (3 n + 3 m)).  Segment and patch data in
 * wavelengths (NEC's), cells in metres, wlam the wavelength in metres,
 * omega = k0 c0.  OpenMP over the (column, cell) tasks with `threads`.
 */
int oracle_conductor_volume(double k0, double wlam, size_t cells, const double *centers, const double *sizes,
                            const double complex *chi, int n, const double *x, const double *y, const double *z,
                            const double *si, const double *bi, const double *cab, const double *sab,
                            const double *salp, int m, const double *pp, const double *pbi, const double *t1,
                            const double *t2, size_t columns, const double complex *sources, double rkh,
                            int ground, double complex *coupling, int threads)
{
  double node[6], weight[6];
  if (!wv_gauss(6, node, weight)) return 0;
  const wv_geom g = {n, m, x, y, z, si, bi, cab, sab, salp, pp, pbi, t1, t2, rkh, ground};
  const double omega = k0 * ORC_C0;
  const size_t stride = 3 * (size_t)n + 3 * (size_t)m;
  const long long tasks = (long long)(columns * cells);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1) if (tasks >= 8)
  for (long long task = 0; task < tasks; ++task) {
    const size_t column = (size_t)task / cells, cell = (size_t)task % cells;
    const double *c = centers + 3 * cell, *s = sizes + 3 * cell;
    complex double field[3] = {0, 0, 0};
    for (int ix = 0; ix < 6; ix++)
      for (int iy = 0; iy < 6; iy++)
        for (int iz = 0; iz < 6; iz++) {
          const double w = weight[ix] * weight[iy] * weight[iz] / 8.0;
          const double p[3] = {c[0] + 0.5 * s[0] * node[ix], c[1] + 0.5 * s[1] * node[iy],
                               c[2] + 0.5 * s[2] * node[iz]};
          complex double e[3];
          wv_nefld(&g, sources + column * stride, p[0] / wlam, p[1] / wlam, p[2] / wlam, e);
          for (int q = 0; q < 3; q++) field[q] += w * e[q];
        }
    for (int q = 0; q < 3; q++)
      coupling[column * 3 * cells + 3 * cell + q] = -I * omega * ORC_EPS0 * chi[cell] * field[q];
  }
  return 1;
}
