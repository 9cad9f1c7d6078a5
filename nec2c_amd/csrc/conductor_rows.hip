// Volume-to-conductor rows of the coupled wire/patch + volume system
// (dielectric_wire.c:1164-1262): the point-tested field of each cell's
// constant current, integrated over the cell with the reference's 10-point
// Gauss-Legendre rule per axis (dw_point_cuboid_n / _h, :101-181), free space
// or over a PEC plane (the mirrored cell, dw_point_background :183-214).
//
// One thread per (conductor task, cell): wire rows need the E dyadic tested
// by the wire tangent, t.N[:, q] = sum w [g_d t_q + g_r (t.r)(r_q)/r^2], so the
// thread keeps three complex sums; a patch's two rows need the H dyadic,
// t.H[:, q] = (t x G)_q with G = sum w g' r, so it keeps G.  Each thread runs
// the 1000 quadrature nodes with a reciprocal square root and a phase per
// node: FP64-ALU-bound.  The phase exp(-i k r) splits into exp(-i k r0) at
// the cell centre, once per cell, and exp(-i k (r - r0)), a short Taylor
// series, since |k (r - r0)| stays below k |size| / 2 (sincos at every node
// for cells too large for the series).
#include "vie_kernels.h"

namespace vie {

namespace {

constexpr int RWG = 256;
#ifndef VIE_ROWS_UNROLL
#define VIE_ROWS_UNROLL 2  // nodes of the z loop in flight per thread
#endif
constexpr int kOrder = 10;
constexpr double kPi = 3.141592653589793238462643383279502884;

struct GaussRule {
  double x[kOrder], w[kOrder];
};

struct Conductors {
  int nw, np;
  const double* wp;  // wire observation points [3 nw]
  const double* wt;  // wire tangents [3 nw]
  const double* pp;  // patch points [3 np]
  const double* t1;  // patch tangents [3 np]
  const double* t2;
  const double* ps;  // patch scale (the reference's psalp) [np]
};

// r = sqrt(a) and 1/r from the hardware reciprocal square root and two
// coupled Newton (Goldschmidt) steps: g -> sqrt(a), h -> 1/(2 sqrt(a)).
__device__ inline void root_and_inverse(double a, double& r, double& ir) {
  const double y = __builtin_amdgcn_rsq(a);
  double g = a * y, h = 0.5 * y;
  double e = fma(-h, g, 0.5);
  g = fma(g, e, g);
  h = fma(h, e, h);
  e = fma(-h, g, 0.5);
  g = fma(g, e, g);
  h = fma(h, e, h);
  r = g;
  ir = h + h;
}

// (-1)^m / n! with n = 2m (cosine) or 2m + 1 (sine), folded at compile time.
constexpr double series_coefficient(int m, bool sine) {
  double f = 1.0;
  for (int n = 2; n <= 2 * m + (sine ? 1 : 0); n++) f *= (double)n;
  return (m % 2 ? -1.0 : 1.0) / f;
}

// cos x and sin x by their Taylor series (Horner in x^2).  DEG 8 (terms to
// x^16 and x^17) serves |x| <= 1, DEG 5 (to x^10 and x^11) |x| <= 1/4: the
// first dropped term is below 2e-16 in both cases.
template <int DEG> __device__ inline void cos_sin_series(double x, double& c, double& s) {
  const double x2 = x * x;
  double cc = series_coefficient(DEG, false), ss = series_coefficient(DEG, true);
#pragma unroll
  for (int m = DEG - 1; m >= 0; m--) {
    cc = fma(cc, x2, series_coefficient(m, false));
    ss = fma(ss, x2, series_coefficient(m, true));
  }
  c = cc;
  s = ss * x;
}

// Phase evaluation of exp(-i k r) inside one cell.
//   kSeries1 / kSeries4: exp(-i k r0) exp(-i k (r - r0)) with r0 the distance
//     to the cell centre; the second factor by the series above, valid since
//     |k (r - r0)| <= k |node - centre| < k |size| / 2 <= 1 (or 1/4); the
//     constant first factor leaves the sums and multiplies them once.
//   kDirect: sincos(k r) at every node (cells large against the wavelength).
enum Phase { kDirect = 0, kSeries1 = 1, kSeries4 = 2 };

// Sums over the cuboid's 10^3 quadrature nodes for one observation point,
// scaled by the cell volume / (4 pi): the three tested E components
// t.N[:, q] (MAGNETIC false) or the vector G = sum w grad g (MAGNETIC true,
// H[:, q] = G x e_q).  The nodes of one (x, y) column share dx, dy, so the
// z loop accumulates F = sum f and Fz = sum f dz of each node's vector
// coefficient f and the column adds (dx F, dy F, Fz).  Returns false when a
// node coincides with the observation point (or the geometry is not
// finite), where the reference's evaluation fails.
template <bool MAGNETIC, int PHASE>
__device__ inline bool cell_sums(const double obs[3], const double t[3], const double c[3], const double s[3],
                                 double k0, const GaussRule& q, c64 out[3]) {
  const double hx = 0.5 * s[0], hy = 0.5 * s[1], hz = 0.5 * s[2];
  const double ox = obs[0] - c[0], oy = obs[1] - c[1], oz = obs[2] - c[2];
  double kr0 = 0.0;
  if (PHASE != kDirect) kr0 = k0 * sqrt(ox * ox + oy * oy + oz * oz);
  const double k2 = k0 * k0;
  c64 S = {0, 0}, V[3] = {{0, 0}, {0, 0}, {0, 0}};
  bool ok = true;
  for (int ix = 0; ix < kOrder; ix++) {
    const double dx = fma(-hx, q.x[ix], ox);
    for (int iy = 0; iy < kOrder; iy++) {
      const double dy = fma(-hy, q.x[iy], oy);
      const double wxy = q.w[ix] * q.w[iy];
      const double rho2 = dx * dx + dy * dy;
      const double txy = MAGNETIC ? 0.0 : t[0] * dx + t[1] * dy;
      c64 F = {0, 0}, Fz = {0, 0};
#pragma unroll VIE_ROWS_UNROLL
      for (int iz = 0; iz < kOrder; iz++) {
        const double dz = fma(-hz, q.x[iz], oz);
        const double w = wxy * q.w[iz];
        const double r2 = fma(dz, dz, rho2);
        ok = ok && r2 > 0.0;
        double r, ir;
        root_and_inverse(r2, r, ir);
        const double ir2 = ir * ir;
        double cs, sn;
        if (PHASE == kDirect) sincos(k0 * r, &sn, &cs);
        else cos_sin_series<PHASE == kSeries1 ? 8 : 5>(fma(k0, r, -kr0), cs, sn);
        const c64 e = {cs, -sn};  // exp(-i k r) [/ exp(-i k r0)]
        c64 f;
        if (!MAGNETIC) {
          // g = e / r; diag = g (k^2 - 1/r^2 - i k/r), radial = g (3/r^2 - k^2 + 3 i k/r)
          const c64 g = {e.re * ir, e.im * ir};
          const c64 dg = g * c64{k2 - ir2, -k0 * ir};
          S = S + w * dg;
          const c64 rg = g * c64{fma(3.0, ir2, -k2), 3.0 * k0 * ir};
          f = (w * fma(t[2], dz, txy) * ir2) * rg;
        } else {
          // w g (-i k - 1/r) / r = (w / r^2) e (-1/r - i k)
          f = (w * ir2) * (e * c64{-ir, -k0});
        }
        F = F + f;
        Fz = Fz + dz * f;
      }
      V[0] = V[0] + dx * F;
      V[1] = V[1] + dy * F;
      V[2] = V[2] + Fz;
    }
  }
  // volume factor of the nodes' weights and 1/(4 pi), and exp(-i k r0)
  const double vol = hx * hy * hz / (4.0 * kPi);
  c64 scale = {vol, 0.0};
  if (PHASE != kDirect) {
    double sn, cs;
    sincos(kr0, &sn, &cs);
    scale = c64{vol * cs, -vol * sn};
  }
  for (int k = 0; k < 3; k++) out[k] = scale * (MAGNETIC ? V[k] : t[k] * S + V[k]);
  return ok;
}

// Phase variant for a cell: the series when k |size| / 2 bounds |k (r - r0)|
// within its range.
__device__ inline int phase_of(const double s[3], double k0) {
  const double b = 0.5 * k0 * sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  return b <= 0.25 ? kSeries4 : b <= 1.0 ? kSeries1 : kDirect;
}

// FIXED >= 0: every cell of the launch takes that variant (the host found
// them all within its range); -1: each cell picks its own.
template <bool MAGNETIC, int FIXED>
__device__ inline bool cell_sums_any(int phase, const double obs[3], const double t[3], const double c[3],
                                     const double s[3], double k0, const GaussRule& q, c64 out[3]) {
  if (FIXED >= 0) return cell_sums<MAGNETIC, (FIXED < 0 ? 0 : FIXED)>(obs, t, c, s, k0, q, out);
  if (phase == kSeries4) return cell_sums<MAGNETIC, kSeries4>(obs, t, c, s, k0, q, out);
  if (phase == kSeries1) return cell_sums<MAGNETIC, kSeries1>(obs, t, c, s, k0, q, out);
  return cell_sums<MAGNETIC, kDirect>(obs, t, c, s, k0, q, out);
}

__device__ inline void cross(const double a[3], const c64 b[3], c64 o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// grid: x over cells, y over tasks task0 + blockIdx.y (wires, then patches).
// failed: the smallest task * cells + cell whose evaluation failed (the first
// failure in the reference's order: wire tasks precede patch tasks).
template <int FIXED>
__global__ __launch_bounds__(RWG) void k_conductor_rows(Conductors cd, const double* __restrict__ centers,
                                                        const double* __restrict__ sizes, size_t cells, double k0,
                                                        double omega_eps0, int ground, GaussRule q, int task0,
                                                        c64* __restrict__ rows, size_t ld,
                                                        unsigned long long* failed) {
  const size_t cell = (size_t)blockIdx.x * RWG + threadIdx.x;
  const int task = task0 + (int)blockIdx.y;
  if (cell >= cells) return;
  double c[3] = {centers[3 * cell], centers[3 * cell + 1], centers[3 * cell + 2]};
  const double s[3] = {sizes[3 * cell], sizes[3 * cell + 1], sizes[3 * cell + 2]};
  const double csign[3] = {-1.0, -1.0, 1.0};  // image columns (q == 2 ? +1 : -1)
  const int phase = FIXED >= 0 ? FIXED : phase_of(s, k0);
  bool ok;
  if (task < cd.nw) {
    const double obs[3] = {cd.wp[3 * task], cd.wp[3 * task + 1], cd.wp[3 * task + 2]};
    const double t[3] = {cd.wt[3 * task], cd.wt[3 * task + 1], cd.wt[3 * task + 2]};
    c64 e[3], ei[3];
    ok = cell_sums_any<false, FIXED>(phase, obs, t, c, s, k0, q, e);
    if (ground == 1) {
      c[2] = -c[2];
      ok = cell_sums_any<false, FIXED>(phase, obs, t, c, s, k0, q, ei) && ok;
      for (int k = 0; k < 3; k++) e[k] = e[k] + csign[k] * ei[k];
    }
    // tested / (i omega eps0)
    c64* row = rows + (size_t)task * ld + 3 * cell;
    for (int k = 0; k < 3; k++) row[k] = c64{e[k].im / omega_eps0, -e[k].re / omega_eps0};
  } else {
    const int p = task - cd.nw;
    const double obs[3] = {cd.pp[3 * p], cd.pp[3 * p + 1], cd.pp[3 * p + 2]};
    const double t1[3] = {cd.t1[3 * p], cd.t1[3 * p + 1], cd.t1[3 * p + 2]};
    const double t2[3] = {cd.t2[3 * p], cd.t2[3 * p + 1], cd.t2[3 * p + 2]};
    const double unused[3] = {0, 0, 0};
    c64 G[3], Gi[3], a[3], b[3];
    ok = cell_sums_any<true, FIXED>(phase, obs, unused, c, s, k0, q, G);
    cross(t2, G, a);
    cross(t1, G, b);
    if (ground == 1) {
      c[2] = -c[2];
      ok = cell_sums_any<true, FIXED>(phase, obs, unused, c, s, k0, q, Gi) && ok;
      c64 ai[3], bi[3];
      cross(t2, Gi, ai);
      cross(t1, Gi, bi);
      for (int k = 0; k < 3; k++) {
        a[k] = a[k] + csign[k] * ai[k];
        b[k] = b[k] + csign[k] * bi[k];
      }
    }
    const double scale = -cd.ps[p];
    c64* r0 = rows + (size_t)(cd.nw + 2 * p) * ld + 3 * cell;
    c64* r1 = r0 + ld;
    for (int k = 0; k < 3; k++) {
      r0[k] = scale * a[k];
      r1[k] = scale * b[k];
    }
  }
  if (!ok) atomicMin(failed, (unsigned long long)task * cells + cell);
}

// Gauss-Legendre nodes and weights by Newton's method on P_n from the
// Chebyshev-like start cos(pi (i + 3/4) / (n + 1/2)) (the rule of
// dielectric_wire.c:70-98, which the row entries must share).
GaussRule gauss_rule() {
  GaussRule g{};
  const int n = kOrder;
  for (int root = 0; root < (n + 1) / 2; root++) {
    double z = cos(kPi * (root + 0.75) / (n + 0.5)), previous, derivative = 0;
    int it = 0;
    do {
      double p0 = 1.0, p1 = z;
      for (int deg = 2; deg <= n; deg++) {
        const double p2 = ((2.0 * deg - 1.0) * z * p1 - (deg - 1.0) * p0) / deg;
        p0 = p1;
        p1 = p2;
      }
      derivative = n * (z * p1 - p0) / (z * z - 1.0);
      previous = z;
      z = previous - p1 / derivative;
      ++it;
    } while (fabs(z - previous) > 4.0 * 2.220446049250313e-16 && it < 32);
    g.x[root] = -z;
    g.x[n - 1 - root] = z;
    g.w[root] = 2.0 / ((1.0 - z * z) * derivative * derivative);
    g.w[n - 1 - root] = g.w[root];
  }
  return g;
}

}  // namespace

int conductor_rows(double k0, const double* centers, const double* sizes, size_t cells, double max_size, int nw,
                   const double* wp, const double* wt, int np, const double* pp, const double* t1, const double* t2,
                   const double* ps, int ground, c64* rows, size_t ld, unsigned long long* failed, hipStream_t s) {
  const int tasks = nw + np;
  if (tasks == 0 || cells == 0) return 1;
  Conductors cd{nw, np, wp, wt, pp, t1, t2, ps};
  // omega eps0 with omega = k0 c0 (dielectric_wire.c:33-35, :869)
  const double omega_eps0 = k0 * 299792458.0 * 8.8541878128e-12;
  const GaussRule rule = gauss_rule();
  const unsigned blocks = (unsigned)((cells + RWG - 1) / RWG);
  // one phase variant for the launch when every cell is within its range
  // (a slightly larger bound than the cells' own test keeps the choice safe)
  const double b = 0.5 * k0 * max_size * (1.0 + 1e-12);
  auto kernel = b <= 0.25 ? k_conductor_rows<kSeries4> : b <= 1.0 ? k_conductor_rows<kSeries1> : k_conductor_rows<-1>;
  for (int t0 = 0; t0 < tasks; t0 += 65535) {
    const unsigned chunk = (unsigned)(tasks - t0 < 65535 ? tasks - t0 : 65535);
    hipLaunchKernelGGL(kernel, dim3(blocks, chunk), dim3(RWG), 0, s, cd, centers, sizes, cells, k0, omega_eps0,
                       ground, rule, t0, rows, ld, failed);
  }
  return hipGetLastError() == hipSuccess;
}

}  // namespace vie
