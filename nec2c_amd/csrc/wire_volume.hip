// Conductor-to-volume coupling: the volume rows of every conductor column of
// the coupled system (dw_assemble_wire_to_volume, dielectric_wire.c:316-508).
// For each conductor unknown the reference freezes NEC's basis coefficients
// (constant, sine and cosine current per segment; a surface current per
// patch), evaluates NEC's near field (nefld, fields.c:1123) at the 6^3 Gauss
// points of every cell, averages, and scales by -i omega eps0 chi.
//
// The field of one segment at one point does not depend on the column, and
// the average is linear, so the work splits in two kernels:
//   k_source_fields: one thread per (cell, radiating segment or patch)
//     averages that source's 9 (segment: constant, sine and cosine current,
//     x y z) or 6 (patch: t1 and t2 current) field components over the
//     cell's 216 points -> F[j][3 cell + q], j the basis term;
//   k_columns: coupling[col][3 cell + q] = -i omega eps0 chi_cell
//     sum_j coef[col][j] F[j][3 cell + q], the terms of a segment skipped in
//     a column where all three of its coefficients are zero (as nefld skips
//     it).
// NEC's classic engine only (thin-wire kernel eksc with its Romberg integral,
// the lumped element beyond rkh wavelengths), free space or a PEC plane.
// Every (point, source) evaluation runs NEC's adaptive integration: FP64
// ALU-bound, with per-thread step counts.
#include "vie_kernels.h"

namespace vie {

namespace {

constexpr int SWG = 128;
// NEC's constants as nec2c.h defines them (nec2c.h:36-59): the fields depend
// on these roundings.
constexpr double kNecPi = 3.141592654;
constexpr double kNecTp = 6.283185308;
constexpr double kNecEta = 376.73;
constexpr double kNecConst1 = 4.771341189;  // CONST1 = i 4.771341189
constexpr double kNecConst2 = 4.771341188;
constexpr double kPiExact = 3.141592653589793238462643383279502884;

struct Gauss6 {
  double x[6], w[6];
};

struct Sources {
  int n, m;  // segments, patches
  const double *x, *y, *z, *si, *bi, *cab, *sab, *salp;  // [n], wavelengths
  const double *pp, *pbi, *t1, *t2;                      // patches: [3 m], [m], [3 m], [3 m]
  double rkh;
};

// sin and cos of x for |x| < 8192 (arguments here are 2 pi times distances in
// wavelengths): x - k pi/2 with pi/2 in two parts (an FMA each), then the
// Taylor series of sin and cos on [-pi/4, pi/4] to x^17 and x^16 (first
// dropped term below 1e-17); larger arguments take the library routine.
constexpr double taylor(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; i++) f *= (double)i;
  return 1.0 / f;
}

__device__ inline void sin_cos(double x, double* s, double* c) {
  if (!(fabs(x) < 8192.0)) {
    sincos(x, s, c);
    return;
  }
  constexpr double kHalfPi1 = 1.5707963267948966;     // pi/2 rounded to double
  constexpr double kHalfPi2 = 6.123233995736766e-17;  // pi/2 - kHalfPi1
  const double k = rint(x * 0.6366197723675814);      // 2/pi
  double r = fma(-k, kHalfPi1, x);
  r = fma(-k, kHalfPi2, r);
  const double r2 = r * r;
  double cc = taylor(16), ss = -taylor(17);
#pragma unroll
  for (int m = 7; m >= 0; m--) {
    cc = fma(cc, r2, (m % 2 ? -1.0 : 1.0) * taylor(2 * m));
    ss = fma(ss, r2, (m % 2 ? -1.0 : 1.0) * taylor(2 * m + 1));
  }
  ss *= r;
  const int q = (int)((long long)k & 3);
  *s = q == 0 ? ss : q == 1 ? cc : q == 2 ? -ss : -cc;
  *c = q == 0 ? cc : q == 1 ? -ss : q == 2 ? -cc : ss;
}

// i * CONST1 as a multiplier: CONST1 * v = (-4.77 v.im, 4.77 v.re)
__device__ inline c64 mul_const1(c64 v) { return c64{-kNecConst1 * v.im, kNecConst1 * v.re}; }

// Romberg integral of exp(-i rk)/rk, rk = sqrt(rkb2 + (t - zpk)^2), over
// [el1, el2] with NEC's step control (intx, calculations.c:734-896, field
// point form): 3- then 5-point estimates accepted below 1e-4 relative change,
// step halving down to s / 65536 (where NEC accepts the 5-point value),
// doubling after four accepted steps.
__device__ inline void gf(double zpk, double rkb2, double t, double& co, double& si) {
  const double d = t - zpk, rk = sqrt(rkb2 + d * d), inv = 1.0 / rk;
  double sn, cs;
  sin_cos(rk, &sn, &cs);
  si = sn * inv;
  co = cs * inv;
}

__device__ inline bool converged(double f1r, double f2r, double f1i, double f2i) {
  const double den = fmax(fabs(f2r), fabs(f2i));
  if (den < 1.0e-37) return true;
  return fabs((f1r - f2r) / den) <= 1.0e-4 && fabs((f1i - f2i) / den) <= 1.0e-4;
}

// g_start, g_end: the integrand at el1 and el2 when the caller has them
// (they are the phases of the segment ends that gx evaluates as well).
__device__ c64 romberg(double zpk, double rkb2, double el1, double el2, c64 g_start, c64 g_end) {
  const int nma = 65536;
  const double s = el2 - el1, ep = s / (10.0 * nma), zend = el2 - ep;
  double z = el1, sr = 0.0, si = 0.0, dz = 0.0, dzot = 0.0;
  double g1r, g1i, g3r = 0, g3i = 0, g5r = 0, g5i = 0;
  int ns = 1, nt = 0;
  bool fresh = true, first = true;
  g1r = g_start.re;
  g1i = g_start.im;
  for (;;) {
    if (fresh) {
      dz = s / ns;
      if (z + dz > el2) {
        dz = el2 - z;
        if (fabs(dz) <= ep) break;
      }
      dzot = 0.5 * dz;
      gf(zpk, rkb2, z + dzot, g3r, g3i);
      if (first) {
        // the first step spans the whole interval
        g5r = g_end.re;
        g5i = g_end.im;
      } else {
        gf(zpk, rkb2, z + dz, g5r, g5i);
      }
      first = false;
    }
    const double t00r = (g1r + g5r) * dzot, t00i = (g1i + g5i) * dzot;
    const double t01r = (t00r + dz * g3r) * 0.5, t01i = (t00i + dz * g3i) * 0.5;
    const double t10r = (4.0 * t01r - t00r) / 3.0, t10i = (4.0 * t01i - t00i) / 3.0;
    double addr, addi;
    int steps;
    if (converged(t01r, t10r, t01i, t10i)) {
      addr = t10r;
      addi = t10i;
      steps = 2;
    } else {
      double g2r, g2i, g4r, g4i;
      gf(zpk, rkb2, z + 0.25 * dz, g2r, g2i);
      gf(zpk, rkb2, z + 0.75 * dz, g4r, g4i);
      const double t02r = (t01r + dzot * (g2r + g4r)) * 0.5, t02i = (t01i + dzot * (g2i + g4i)) * 0.5;
      const double t11r = (4.0 * t02r - t01r) / 3.0, t11i = (4.0 * t02i - t01i) / 3.0;
      const double t20r = (16.0 * t11r - t10r) / 15.0, t20i = (16.0 * t11i - t10i) / 15.0;
      if (!converged(t11r, t20r, t11i, t20i)) {
        nt = 0;
        if (ns < nma) {
          ns *= 2;
          dz = s / ns;
          dzot = 0.5 * dz;
          g5r = g3r;
          g5i = g3i;
          g3r = g2r;
          g3i = g2i;
          fresh = false;
          continue;
        }
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
    fresh = true;
  }
  return c64{sr, si};
}

__device__ inline void gx(double zz, double rh, c64& gz, c64& gzp, double& r) {
  const double r2 = zz * zz + rh * rh, inv = 1.0 / (r = sqrt(r2));
  double sn, cs;
  sin_cos(kNecTp * r, &sn, &cs);
  gz = c64{cs * inv, -sn * inv};
  // -(1 + i k r) gz / r^2
  const c64 a = c64{1.0, kNecTp * r} * gz;
  const double inv2 = inv * inv;
  gzp = c64{-a.re * inv2, -a.im * inv2};
}

// Thin-wire kernel of one segment (eksc, fields.c:349-394): axial and radial
// fields of its sine, cosine and constant currents at axial offset z and
// reduced radius rh; ss, cs = sin, cos of k s / 2.
struct Kernel {
  c64 ezs, ers, ezc, erc, ezk, erk;
};

__device__ inline Kernel thin_wire(double s, double ss, double cs, double z, double rh) {
  const double xk = kNecTp, sh = 0.5 * s, shk = xk * sh;
  const double z2a = sh - z, z1a = -(sh + z);
  c64 gz1, gz2, gp1, gp2;
  double r1, r2;
  gx(z1a, rh, gz1, gp1, r1);
  gx(z2a, rh, gz2, gp2, r2);
  // exp(-i k r)/r at the ends is the Romberg integrand exp(-i rk)/rk there,
  // conjugated and divided by k
  const c64 it = romberg(xk * z, (xk * rh) * (xk * rh), -shk, shk, c64{gz1.re / xk, -gz1.im / xk},
                         c64{gz2.re / xk, -gz2.im / xk});
  c64 gzp1 = z1a * gp1, gzp2 = z2a * gp2;
  Kernel k;
  k.ezs = mul_const1((cs * xk) * (gz2 - gz1) - ss * (gzp2 + gzp1));
  const c64 c1 = (ss * xk) * (gz2 + gz1) + cs * (gzp2 - gzp1);
  k.ezc = c64{-c1.re, -c1.im};
  k.ezc = mul_const1(k.ezc);
  k.erk = mul_const1(rh * (gp2 - gp1));
  // -(gzp2 - gzp1 + k^2 (cint - i sint))
  const c64 ek = (gzp2 - gzp1) + (xk * xk) * c64{it.re, -it.im};
  k.ezk = mul_const1(c64{-ek.re, -ek.im});
  gzp1 = z1a * gzp1;
  gzp2 = z2a * gzp2;
  if (rh >= 1.0e-10) {
    const c64 a = ss * (gzp2 + gzp1 + gz2 + gz1) - (cs * xk) * (z2a * gz2 - z1a * gz1);
    const c64 b = cs * (gzp2 - gzp1 + gz2 - gz1) + (ss * xk) * (z2a * gz2 + z1a * gz1);
    k.ers = mul_const1(c64{-a.re / rh, -a.im / rh});
    k.erc = mul_const1(c64{-b.re / rh, -b.im / rh});
  } else {
    k.ers = k.erc = c64{0, 0};
  }
  return k;
}

// w times the field of segment j's constant (acc[0..2]), sine (acc[3..5])
// and cosine (acc[6..8]) currents at the point, added to acc; classic engine
// (efld, fields.c:38-235): free space, or minus the image segment's field
// over a PEC plane.  ss, cs: sin, cos of k s / 2; mc = sin(pi s) / pi (NEC's pi).
template <bool PEC>
__device__ inline void segment_field(const Sources& S, int j, double ss, double cs, double mc, double xi, double yi,
                                     double zi, double ai, double w, c64 acc[9]) {
  const double xij = xi - S.x[j], yij = yi - S.y[j];
  const double cabj = S.cab[j], sabj = S.sab[j], s = S.si[j];
#pragma unroll 1
  for (int ip = 0; ip < (PEC ? 2 : 1); ip++) {
    const double rfl = ip ? -1.0 : 1.0;
    const double salpr = S.salp[j] * rfl, zij = zi - rfl * S.z[j];
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
    const double wi = ip ? -w : w;
    c64 tezk, terk, tezc, terc;
    if (r >= S.rkh) {
      // lumped current element
      const double rmag = kNecTp * r, cth = zp / r, px = rh / r, py = kNecTp * r * r;
      double sn, cn;
      sin_cos(rmag, &sn, &cn);
      const c64 e0 = {cn, -sn};
      const c64 ty = ((kNecEta * cth) / py) * (e0 * c64{1.0, -1.0 / rmag});
      const c64 tz = ((kNecEta * px) / (2.0 * py)) * (e0 * c64{1.0, rmag - 1.0 / rmag});
      tezk = cth * ty - px * tz;
      terk = px * ty + cth * tz;
      tezc = mc * tezk;
      terc = mc * terk;
      tezk = s * tezk;
      terk = s * terk;
    } else {
      const Kernel k = thin_wire(s, ss, cs, zp, rh);
      tezk = k.ezk;
      terk = k.erk;
      tezc = k.ezc;
      terc = k.erc;
      acc[3] = acc[3] + wi * (cabj * k.ezs + rhox * k.ers);
      acc[4] = acc[4] + wi * (sabj * k.ezs + rhoy * k.ers);
      acc[5] = acc[5] + wi * (salpr * k.ezs + rhoz * k.ers);
    }
    acc[0] = acc[0] + wi * (cabj * tezk + rhox * terk);
    acc[1] = acc[1] + wi * (sabj * tezk + rhoy * terk);
    acc[2] = acc[2] + wi * (salpr * tezk + rhoz * terk);
    acc[6] = acc[6] + wi * (cabj * tezc + rhox * terc);
    acc[7] = acc[7] + wi * (sabj * tezc + rhoy * terc);
    acc[8] = acc[8] + wi * (salpr * tezc + rhoz * terc);
  }
}

// w times the field of patch i's t1 (acc[0..2]) and t2 (acc[3..5]) surface
// currents (unere, fields.c:1713-1760), added to acc; over a PEC plane plus
// the negated field of the mirrored patch.
template <bool PEC>
__device__ inline void patch_field(const Sources& S, int i, double xob, double yob, double zob, double w,
                                   c64 acc[9]) {
  const double* c = S.pp + 3 * i;
  const double* t1 = S.t1 + 3 * i;
  const double* t2 = S.t2 + 3 * i;
#pragma unroll
  for (int image = 0; image < (PEC ? 2 : 1); image++) {
    const double sgn = image ? -1.0 : 1.0;
    const double zr = sgn * c[2], t1zr = sgn * t1[2], t2zr = sgn * t2[2];
    const double rx = xob - c[0], ry = yob - c[1], rz = zob - zr;
    const double r2 = rx * rx + ry * ry + rz * rz;
    if (r2 <= 1.0e-20) continue;
    const double r = sqrt(r2), tt1 = -kNecTp * r, tt2 = tt1 * tt1, rt = r2 * r;
    double sn, cn;
    sin_cos(tt1, &sn, &cn);
    const double amp = kNecConst2 * S.pbi[i];
    const c64 er = {sn * amp, -cn * amp};
    const c64 q1 = (1.0 / rt) * (c64{tt2 - 1.0, tt1} * er);
    const c64 q2 = (1.0 / (rt * r2)) * (c64{3.0 - tt2, -3.0 * tt1} * er);
    const c64 e1 = (t1[0] * rx + t1[1] * ry + t1zr * rz) * q2;
    const c64 e2 = (t2[0] * rx + t2[1] * ry + t2zr * rz) * q2;
    const c64 g[6] = {t1[0] * q1 + rx * e1, t1[1] * q1 + ry * e1, t1zr * q1 + rz * e1,
                      t2[0] * q1 + rx * e2, t2[1] * q1 + ry * e2, t2zr * q1 + rz * e2};
    const double wi = image ? -w : w;
#pragma unroll
    for (int q = 0; q < 6; q++) acc[q] = acc[q] + wi * g[q];
  }
}

// grid: x over cells, y over the radiating sources (active[y]: a segment
// index, or n + patch index).  F[(jbase + k) * ldf + 3 cell + q].
// Two waves per SIMD: the PEC variant would otherwise take every register
// and run one (measured 1.4x faster at two).
template <bool PEC>
__global__ __launch_bounds__(SWG) __attribute__((amdgpu_waves_per_eu(2))) void k_source_fields(Sources S, const int* __restrict__ active,
                                                       const int* __restrict__ jbase,
                                                       const double* __restrict__ centers,
                                                       const double* __restrict__ sizes, size_t cell0,
                                                       size_t cells, double wlam, Gauss6 g, c64* __restrict__ F,
                                                       size_t ldf) {
  const size_t cl = (size_t)blockIdx.x * SWG + threadIdx.x;
  if (cl >= cells) return;
  const size_t cell = cell0 + cl;
  const int src = active[blockIdx.y];
  const double c[3] = {centers[3 * cell], centers[3 * cell + 1], centers[3 * cell + 2]};
  const double h[3] = {0.5 * sizes[3 * cell], 0.5 * sizes[3 * cell + 1], 0.5 * sizes[3 * cell + 2]};
  // Can a point of this cell fall inside a wire (within 0.9 of its radius
  // of the axis, along its length)?  Only then does the kernel radius
  // search of nefld run per point.
  bool near = false;
  {
    const double hd = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]) / wlam;
    const double cx = c[0] / wlam, cy = c[1] / wlam, cz = c[2] / wlam;
    for (int i = 0; i < S.n && !near; i++) {
      const double dx = cx - S.x[i], dy = cy - S.y[i], dz = cz - S.z[i];
      const double lim = (hd + sqrt(0.5001 * 0.5001 * S.si[i] * S.si[i] + 0.9 * S.bi[i] * S.bi[i])) * (1.0 + 1e-9);
      near = dx * dx + dy * dy + dz * dz <= lim * lim;
    }
  }
  const bool seg = src < S.n;
  double ss = 0.0, cs = 0.0, mc = 0.0;
  if (seg) {
    sin_cos(kNecTp * (0.5 * S.si[src]), &ss, &cs);
    mc = sin(kNecPi * S.si[src]) / kNecPi;
  }
  c64 acc[9];
#pragma unroll
  for (int q = 0; q < 9; q++) acc[q] = c64{0, 0};
#pragma unroll 1
  for (int ix = 0; ix < 6; ix++)
#pragma unroll 1
    for (int iy = 0; iy < 6; iy++)
#pragma unroll 1
      for (int iz = 0; iz < 6; iz++) {
        const double w = g.w[ix] * g.w[iy] * g.w[iz] / 8.0;
        const double xo = (c[0] + h[0] * g.x[ix]) / wlam, yo = (c[1] + h[1] * g.x[iy]) / wlam,
                     zo = (c[2] + h[2] * g.x[iz]) / wlam;
        if (seg) {
          double ax = 0.0;
          if (near)
            for (int i = 0; i < S.n; i++) {
              const double xj = xo - S.x[i], yj = yo - S.y[i], zj = zo - S.z[i];
              const double zp = S.cab[i] * xj + S.sab[i] * yj + S.salp[i] * zj;
              if (fabs(zp) > 0.5001 * S.si[i]) continue;
              if (xj * xj + yj * yj + zj * zj - zp * zp > 0.9 * S.bi[i] * S.bi[i]) continue;
              ax = S.bi[i];
              break;
            }
          segment_field<PEC>(S, src, ss, cs, mc, xo, yo, zo, ax, w, acc);
        } else {
          patch_field<PEC>(S, src - S.n, xo, yo, zo, w, acc);
        }
      }
  const size_t j = (size_t)jbase[blockIdx.y];
#pragma unroll
  for (int k = 0; k < 3; k++)
    if (k < 2 || seg)
#pragma unroll
      for (int q = 0; q < 3; q++) F[(j + k) * ldf + 3 * cl + q] = acc[3 * k + q];
}

// coupling[col * ldc + 3 cell + q] for up to CB columns per thread:
// -i omega eps0 chi sum_j coef[col][j] F[j][3 cell + q].  skip[col * nseg + s]
// marks a segment whose three coefficients are all zero in that column.
constexpr int CB = 8;
__global__ __launch_bounds__(256) void k_columns(const c64* __restrict__ F, size_t ldf, int terms,
                                                 const int* __restrict__ term_segment,
                                                 const c64* __restrict__ coef, const unsigned char* __restrict__ skip,
                                                 int nseg, int col0, int ncols, const c64* __restrict__ chi,
                                                 size_t cell0, size_t cells, double omega_eps0,
                                                 c64* __restrict__ coupling, size_t ldc) {
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;  // 3 cell + q within the chunk
  if (idx >= 3 * cells) return;
  c64 acc[CB];
#pragma unroll
  for (int b = 0; b < CB; b++) acc[b] = c64{0, 0};
  for (int j = 0; j < terms; j++) {
    const c64 f = F[(size_t)j * ldf + idx];
    const int sg = term_segment[j];
#pragma unroll
    for (int b = 0; b < CB; b++) {
      const int col = col0 + (b < ncols ? b : 0);
      const bool use = b < ncols && !(sg >= 0 && skip[(size_t)col * nseg + sg]);
      if (use) acc[b] = acc[b] + coef[(size_t)col * terms + j] * f;
    }
  }
  const size_t cell = cell0 + idx / 3;
  const c64 x = chi[cell];
  // -i omega eps0 chi
  const c64 scale = {omega_eps0 * x.im, -omega_eps0 * x.re};
#pragma unroll
  for (int b = 0; b < CB; b++)
    if (b < ncols) coupling[(size_t)(col0 + b) * ldc + 3 * cell0 + idx] = scale * acc[b];
}

Gauss6 gauss6() {
  // dw_gauss (dielectric_wire.c:70-98) at order 6
  Gauss6 g{};
  const int n = 6;
  for (int root = 0; root < (n + 1) / 2; root++) {
    double z = cos(kPiExact * (root + 0.75) / (n + 0.5)), previous, derivative = 0;
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

int conductor_volume_fields(const ConductorSet& cs, const int* active, const int* jbase, int nactive,
                            const double* centers, const double* sizes, size_t cell0, size_t cells, double wlam,
                            int ground, c64* F, size_t ldf, hipStream_t s) {
  if (nactive == 0 || cells == 0) return 1;
  Sources S{cs.n, cs.m, cs.x, cs.y, cs.z, cs.si, cs.bi, cs.cab, cs.sab, cs.salp, cs.pp, cs.pbi, cs.t1, cs.t2, cs.rkh};
  const dim3 grid((unsigned)((cells + SWG - 1) / SWG), (unsigned)nactive);
  if (ground == 1)
    hipLaunchKernelGGL(k_source_fields<true>, grid, dim3(SWG), 0, s, S, active, jbase, centers, sizes, cell0, cells,
                       wlam, gauss6(), F, ldf);
  else
    hipLaunchKernelGGL(k_source_fields<false>, grid, dim3(SWG), 0, s, S, active, jbase, centers, sizes, cell0,
                       cells, wlam, gauss6(), F, ldf);
  return hipGetLastError() == hipSuccess;
}

int conductor_volume_columns(const c64* F, size_t ldf, int terms, const int* term_segment, const c64* coef,
                             const unsigned char* skip, int nseg, int columns, const c64* chi, size_t cell0,
                             size_t cells, double k0, c64* coupling, size_t ldc, hipStream_t s) {
  if (columns == 0 || cells == 0) return 1;
  const double omega_eps0 = k0 * 299792458.0 * 8.8541878128e-12;  // omega = k0 c0 (dielectric_wire.c:869)
  const unsigned blocks = (unsigned)((3 * cells + 255) / 256);
  for (int c0 = 0; c0 < columns; c0 += CB)
    hipLaunchKernelGGL(k_columns, dim3(blocks), dim3(256), 0, s, F, ldf, terms, term_segment, coef, skip, nseg, c0,
                       columns - c0 < CB ? columns - c0 : CB, chi, cell0, cells, omega_eps0, coupling, ldc);
  return hipGetLastError() == hipSuccess;
}

}  // namespace vie
