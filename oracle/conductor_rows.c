/*
 * CPU oracle for the volume-to-conductor rows -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the reference's point-tested fields of a constant
 * cuboid current and the wire/patch rows they fill.  It is compiled into
 * oracle/_build/liboracle.so and is only loaded by tests/ (as a second
 * restatement beside oracle/conductor_rows.py) and by bench.py's CPU-baseline
 * leg; the product library never links or calls it.
 *
 * What is restated (reference file:line for each):
 *   - Gauss-Legendre rule by Newton on P_n          dielectric_wire.c:70-98
 *   - E dyadic of a cuboid at a point (N operator)  dielectric_wire.c:101-142
 *   - H dyadic (grad g x e_q)                       dielectric_wire.c:144-181
 *   - free space / PEC image                        dielectric_wire.c:183-214
 *   - wire rows and patch rows, OpenMP over the
 *     (conductor, cell) tasks as the reference runs dielectric_wire.c:1164-1262
 *
 * Pinning: tests/test_conductor_rows_oracle.py checks oracle/conductor_rows.py
 * against numpy's Gauss-Legendre rule, finite differences of the scalar Green
 * function, the point-source limit and image theory, and this file entry for
 * entry against that module.
 */
#include <complex.h>
#include <float.h>
#include <math.h>
#include <stddef.h>
#include <string.h>

#define ORC_PI 3.141592653589793238462643383279502884
#define ORC_C0 299792458.0
#define ORC_EPS0 8.8541878128e-12
enum { ORC_ORDER = 10 };

static int orc_rule(double *x, double *w)
{
  for (int root = 0; root < (ORC_ORDER + 1) / 2; ++root) {
    double z = cos(ORC_PI * (root + 0.75) / (ORC_ORDER + 0.5)), previous, derivative = 0.0;
    int it = 0;
    do {
      double p0 = 1.0, p1 = z;
      for (int deg = 2; deg <= ORC_ORDER; ++deg) {
        double p2 = ((2.0 * deg - 1.0) * z * p1 - (deg - 1.0) * p0) / deg;
        p0 = p1;
        p1 = p2;
      }
      derivative = ORC_ORDER * (z * p1 - p0) / (z * z - 1.0);
      previous = z;
      z = previous - p1 / derivative;
      ++it;
    } while (fabs(z - previous) > 4.0 * DBL_EPSILON && it < 32);
    if (it == 32) return 0;
    x[root] = -z;
    x[ORC_ORDER - 1 - root] = z;
    w[root] = 2.0 / ((1.0 - z * z) * derivative * derivative);
    w[ORC_ORDER - 1 - root] = w[root];
  }
  return 1;
}

/* out[p][q]: E dyadic (magnetic == 0) or H dyadic of the cuboid at obs. */
static int orc_cuboid(const double obs[3], const double c[3], const double s[3], double k0, int magnetic,
                      const double *x, const double *w, complex double out[3][3])
{
  memset(out, 0, 9 * sizeof(complex double));
  for (int i = 0; i < ORC_ORDER; ++i)
    for (int j = 0; j < ORC_ORDER; ++j)
      for (int l = 0; l < ORC_ORDER; ++l) {
        const int idx[3] = {i, j, l};
        double d[3], r2 = 0.0, wt = 1.0;
        for (int a = 0; a < 3; ++a) {
          d[a] = obs[a] - (c[a] + 0.5 * s[a] * x[idx[a]]);
          r2 += d[a] * d[a];
          wt *= 0.5 * s[a] * w[idx[a]];
        }
        if (!(r2 > 0.0)) return 0;
        const double r = sqrt(r2);
        const complex double g = cexp(-I * k0 * r) / (4.0 * ORC_PI * r);
        if (!magnetic) {
          const complex double dg = g * (k0 * k0 - I * k0 / r - 1.0 / r2);
          const complex double rg = g * (-k0 * k0 + 3.0 * I * k0 / r + 3.0 / r2);
          for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) out[p][q] += wt * (p == q ? dg : 0.0) + wt * rg * d[p] * d[q] / r2;
        } else {
          const complex double f = wt * g * (-I * k0 - 1.0 / r) / r;
          const complex double gr[3] = {f * d[0], f * d[1], f * d[2]};
          out[1][0] += gr[2];
          out[2][0] -= gr[1];
          out[0][1] -= gr[2];
          out[2][1] += gr[0];
          out[0][2] += gr[1];
          out[1][2] -= gr[0];
        }
      }
  return 1;
}

static int orc_background(const double obs[3], const double c[3], const double s[3], double k0, int magnetic,
                          int ground, const double *x, const double *w, complex double out[3][3])
{
  if (!orc_cuboid(obs, c, s, k0, magnetic, x, w, out)) return 0;
  if (ground == 1) {
    const double ci[3] = {c[0], c[1], -c[2]};
    complex double im[3][3];
    if (!orc_cuboid(obs, ci, s, k0, magnetic, x, w, im)) return 0;
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q) out[p][q] += (q == 2 ? 1.0 : -1.0) * im[p][q];
  }
  return 1;
}

/*
 * Same arguments and row layout as the library's vie_hip_volume_conductor_rows
 * (row w for wire w, rows wires + 2p and wires + 2p + 1 for patch p, entry
 * (row, 3 cell + q) at rows[row ld + 3 cell + q]).  threads > 0 sets the
 * OpenMP team.  Returns 1, or 0 with *failed = the first failing task * cells
 * + cell in the reference's order.
 */
int oracle_conductor_rows(double k0, size_t cells, const double *centers, const double *sizes, int wires,
                          const double *wp, const double *wt, int patches, const double *pp, const double *t1,
                          const double *t2, const double *ps, int ground, double complex *rows, size_t ld,
                          int threads, long long *failed)
{
  double x[ORC_ORDER], w[ORC_ORDER];
  if (!orc_rule(x, w)) return 0;
  const double omega = k0 * ORC_C0;
  const long long tasks = (long long)(wires + patches) * (long long)cells;
  long long first = -1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1) if (tasks >= 8)
  for (long long task = 0; task < tasks; ++task) {
    const int k = (int)(task / (long long)cells);
    const size_t cell = (size_t)(task % (long long)cells);
    complex double b[3][3];
    const double *c = centers + 3 * cell, *s = sizes + 3 * cell;
    if (k < wires) {
      if (!orc_background(wp + 3 * k, c, s, k0, 0, ground, x, w, b)) {
#pragma omp critical
        if (first < 0 || task < first) first = task;
        continue;
      }
      for (int q = 0; q < 3; ++q) {
        complex double t = 0.0;
        for (int p = 0; p < 3; ++p) t += wt[3 * k + p] * b[p][q];
        rows[(size_t)k * ld + 3 * cell + q] = t / (I * omega * ORC_EPS0);
      }
    } else {
      const int p = k - wires;
      if (!orc_background(pp + 3 * p, c, s, k0, 1, ground, x, w, b)) {
#pragma omp critical
        if (first < 0 || task < first) first = task;
        continue;
      }
      for (int q = 0; q < 3; ++q) {
        complex double a1 = 0.0, a2 = 0.0;
        for (int e = 0; e < 3; ++e) {
          a1 += t1[3 * p + e] * b[e][q];
          a2 += t2[3 * p + e] * b[e][q];
        }
        rows[(size_t)(wires + 2 * p) * ld + 3 * cell + q] = -ps[p] * a2;
        rows[(size_t)(wires + 2 * p + 1) * ld + 3 * cell + q] = -ps[p] * a1;
      }
    }
  }
  if (failed) *failed = first;
  return first < 0;
}
