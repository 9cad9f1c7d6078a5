// Scattered far field of a volume polarisation current for many directions
// at once (dielectric_solver.c:124-158, dielectric_ground.c:133-144).
//
// For every direction d the reference forms
//   I(d) = sum_cells V_c sinc3(d, size_c) exp(i k0 d.r_c) J_c
//          [+ R(d) sum_cells V_c sinc3(d, size_c) exp(i k0 d.r'_c) J_c   over a ground,
//             r'_c the cell mirrored in z = 0, R(d) the reflection of the current]
//   a(d) = i k0 c0 mu0 / (4 pi) ((I x d) x d).
// Work is directions x cells with a sine/cosine per pair: compute-bound FP64.
// k_farfield_sum: workgroups own a tile of DT directions and a slice of the
// cells; every thread accumulates its cells' terms for the DT directions in
// registers, the workgroup reduces them (DPP wave sums + LDS) to one partial
// per direction, component and slice.  k_farfield_finish adds the slices in
// a fixed order, applies R(d) and the transverse projection.
#include "vie_kernels.h"

namespace vie {

namespace {

constexpr int FWG = 256;
constexpr double kMu0 = 1.25663706212e-6;  // dielectric_solver.c:20-23
constexpr double kC0 = 299792458.0;
constexpr double kPi = 3.141592653589793238462643383279502884;

__device__ inline double sinc_real(double x) {
  // dielectric_solver.c:47-55
  if (fabs(x) < 1.0e-5) {
    const double x2 = x * x;
    return 1.0 - x2 / 6.0 + x2 * x2 / 120.0;
  }
  return sin(x) / x;
}

template <int DT, bool IMAGE, bool UNIFORM>
__global__ __launch_bounds__(FWG) void k_farfield_sum(const double* __restrict__ centers,
                                                      const double* __restrict__ sizes, const c64* __restrict__ current,
                                                      size_t n, const double* __restrict__ dirs, size_t ndir,
                                                      double k0, c64* __restrict__ partial) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const size_t d0 = (size_t)blockIdx.y * DT;
  const unsigned slices = gridDim.x, slice = blockIdx.x;
  double dx[DT], dy[DT], dz[DT], avg_u[DT];
#pragma unroll
  for (int q = 0; q < DT; q++) {
    const size_t d = d0 + q < ndir ? d0 + q : ndir - 1;  // a short last tile repeats its last direction
    dx[q] = dirs[3 * d];
    dy[q] = dirs[3 * d + 1];
    dz[q] = dirs[3 * d + 2];
    if (UNIFORM)
      avg_u[q] = sinc_real(0.5 * k0 * dx[q] * sizes[0]) * sinc_real(0.5 * k0 * dy[q] * sizes[1]) *
                 sinc_real(0.5 * k0 * dz[q] * sizes[2]);
  }
  c64 acc[DT][3], img[IMAGE ? DT : 1][3];
#pragma unroll
  for (int q = 0; q < DT; q++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      acc[q][c] = c64{0, 0};
      if (IMAGE) img[q][c] = c64{0, 0};
    }
  for (size_t cell = (size_t)slice * FWG + tid; cell < n; cell += (size_t)slices * FWG) {
    const double cx = centers[3 * cell], cy = centers[3 * cell + 1], cz = centers[3 * cell + 2];
    const double sx = UNIFORM ? sizes[0] : sizes[3 * cell], sy = UNIFORM ? sizes[1] : sizes[3 * cell + 1],
                 sz = UNIFORM ? sizes[2] : sizes[3 * cell + 2];
    const double vol = sx * sy * sz;
    const c64 J0 = current[3 * cell], J1 = current[3 * cell + 1], J2 = current[3 * cell + 2];
#pragma unroll
    for (int q = 0; q < DT; q++) {
      const double avg =
          UNIFORM ? avg_u[q]
                  : sinc_real(0.5 * k0 * dx[q] * sx) * sinc_real(0.5 * k0 * dy[q] * sy) * sinc_real(0.5 * k0 * dz[q] * sz);
      double sn, cs;
      sincos(k0 * (dx[q] * cx + dy[q] * cy + dz[q] * cz), &sn, &cs);
      const c64 w = {vol * avg * cs, vol * avg * sn};
      acc[q][0] = acc[q][0] + J0 * w;
      acc[q][1] = acc[q][1] + J1 * w;
      acc[q][2] = acc[q][2] + J2 * w;
      if (IMAGE) {
        sincos(k0 * (dx[q] * cx + dy[q] * cy - dz[q] * cz), &sn, &cs);
        const c64 wi = {vol * avg * cs, vol * avg * sn};
        img[q][0] = img[q][0] + J0 * wi;
        img[q][1] = img[q][1] + J1 * wi;
        img[q][2] = img[q][2] + J2 * wi;
      }
    }
  }
  // workgroup reduction: one partial per (direction, image, component, slice)
  constexpr int NV = DT * 3 * (IMAGE ? 2 : 1);
  __shared__ c64 red[NV][FWG / 64];
#pragma unroll
  for (int q = 0; q < DT; q++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const c64 a = wave_sum_last(acc[q][c]);
      if (lane == 63) red[q * 3 + c][wave] = a;
      if (IMAGE) {
        const c64 b = wave_sum_last(img[q][c]);
        if (lane == 63) red[DT * 3 + q * 3 + c][wave] = b;
      }
    }
  __syncthreads();
  for (int v = tid; v < NV; v += FWG) {
    const int im = v / (DT * 3), q = (v % (DT * 3)) / 3, c = v % 3;
    const size_t d = d0 + q;
    if (d >= ndir) continue;
    c64 t = red[v][0];
    for (int w = 1; w < FWG / 64; w++) t = t + red[v][w];
    partial[(((size_t)d * 2 + im) * 3 + c) * slices + slice] = t;
  }
}

// a(d) from the slice partials: direct sum, plus R(d) times the image sum
// (reflection of dielectric_ground.c:133-144 with rs, rp; PEC: -1, +1).
__global__ void k_farfield_finish(const c64* __restrict__ partial, unsigned slices, size_t ndir,
                                  const double* __restrict__ dirs, double k0, int ground, const c64* rs,
                                  const c64* rp, c64* __restrict__ amplitude) {
  const size_t d = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndir) return;
  c64 I[3], M[3];
  for (int c = 0; c < 3; c++) {
    c64 s = {0, 0}, t = {0, 0};
    for (unsigned k = 0; k < slices; k++) {
      s = s + partial[(((size_t)d * 2 + 0) * 3 + c) * slices + k];
      if (ground) t = t + partial[(((size_t)d * 2 + 1) * 3 + c) * slices + k];
    }
    I[c] = s;
    M[c] = t;
  }
  const double ux = dirs[3 * d], uy = dirs[3 * d + 1], uz = dirs[3 * d + 2];
  if (ground) {
    // down = (ux, uy, -uz); reflected = rs te (te.M) + rp pu (pd.M)
    const double u = hypot(ux, uy);
    const double c = u > 1e-14 ? ux / u : 1.0, s = u > 1e-14 ? uy / u : 0.0, w = fabs(uz);
    const double te[3] = {-s, c, 0.0}, pd[3] = {-w * c, -w * s, -u}, pu[3] = {w * c, w * s, -u};
    const c64 r_s = ground == 1 ? c64{-1, 0} : rs[d], r_p = ground == 1 ? c64{1, 0} : rp[d];
    c64 es = {0, 0}, ep = {0, 0};
    for (int p = 0; p < 3; p++) {
      es = es + te[p] * M[p];
      ep = ep + pd[p] * M[p];
    }
    for (int p = 0; p < 3; p++) I[p] = I[p] + te[p] * (r_s * es) + pu[p] * (r_p * ep);
  }
  // (I x d) x d
  const c64 x0 = uz * I[1] - uy * I[2], x1 = ux * I[2] - uz * I[0], x2 = uy * I[0] - ux * I[1];
  const c64 t0 = uz * x1 - uy * x2, t1 = ux * x2 - uz * x0, t2 = uy * x0 - ux * x1;
  const double f = k0 * kC0 * kMu0 / (4.0 * kPi);  // i f t
  amplitude[3 * d] = c64{-f * t0.im, f * t0.re};
  amplitude[3 * d + 1] = c64{-f * t1.im, f * t1.re};
  amplitude[3 * d + 2] = c64{-f * t2.im, f * t2.re};
}

template <int DT, bool IMAGE, bool UNIFORM>
void launch_sum(const double* centers, const double* sizes, const c64* current, size_t n, const double* dirs,
                size_t ndir, double k0, c64* partial, unsigned slices, hipStream_t s) {
  const unsigned tiles = (unsigned)((ndir + DT - 1) / DT);
  hipLaunchKernelGGL((k_farfield_sum<DT, IMAGE, UNIFORM>), dim3(slices, tiles), dim3(FWG), 0, s, centers, sizes,
                     current, n, dirs, ndir, k0, partial);
}

}  // namespace

// ---- far field of the wires and patches (ffld, radiation.c:36-338; fflds
// :340-373), classic engine, free space or the PEC image (rrv = rrh = -1) ----
namespace {
constexpr double kNecPi = 3.141592654;  // NEC's rounded constants (nec2c.h)
constexpr double kNecTp = 6.283185308;
constexpr double kNecConst3 = -29.97922085;  // CONST3 = i (-29.979...)
constexpr double kNecConst4 = 188.365;       // CONST4 = i 188.365

// sum over segments of exp(i 2 pi r.d) (rr + i ri) t, the segment's current
// a + b sin + c cos integrated in closed form (radiation.c:102-140)
__device__ inline void segment_sum(const ConductorFar& cf, double rox, double roy, double roz, c64 ci[3]) {
  ci[0] = ci[1] = ci[2] = c64{0, 0};
  for (int i = 0; i < cf.n; i++) {
    const double omega = -(rox * cf.cab[i] + roy * cf.sab[i] + roz * cf.salp[i]);
    const double el = kNecPi * cf.si[i], sill = omega * el, top = el + sill, bot = el - sill;
    const double a = fabs(omega) >= 1.0e-7 ? 2.0 * sin(sill) / omega : (2.0 - omega * omega * el * el / 3.0) * el;
    const double too = fabs(top) >= 1.0e-7 ? sin(top) / top : 1.0 - top * top / 6.0;
    const double boo = fabs(bot) >= 1.0e-7 ? sin(bot) / bot : 1.0 - bot * bot / 6.0;
    const double b = el * (boo - too), c = el * (boo + too);
    const c64 A = cf.cur[3 * i], B = cf.cur[3 * i + 1], C = cf.cur[3 * i + 2];
    const double rr = a * A.re + b * B.im + c * C.re, ri = a * A.im - b * B.re + c * C.im;
    double sn, cs;
    sincos(kNecTp * (cf.x[i] * rox + cf.y[i] * roy + cf.z[i] * roz), &sn, &cs);
    const c64 exa = c64{cs, sn} * c64{rr, ri};
    ci[0] = ci[0] + cf.cab[i] * exa;
    ci[1] = ci[1] + cf.sab[i] * exa;
    ci[2] = ci[2] + cf.salp[i] * exa;
  }
}

// fflds: CONST4 ((e.r) r - e), e = sum J area exp(i 2 pi r.p)
__device__ inline void patch_sum(const ConductorFar& cf, double rox, double roy, double roz, c64 e[3]) {
  e[0] = e[1] = e[2] = c64{0, 0};
  for (int j = 0; j < cf.m; j++) {
    double sn, cs;
    sincos(kNecTp * (rox * cf.pp[3 * j] + roy * cf.pp[3 * j + 1] + roz * cf.pp[3 * j + 2]), &sn, &cs);
    const c64 ct = {cs * cf.pbi[j], sn * cf.pbi[j]};
    for (int q = 0; q < 3; q++) e[q] = e[q] + cf.pcur[3 * j + q] * ct;
  }
  const c64 dot = rox * e[0] + roy * e[1] + roz * e[2];
  const double r[3] = {rox, roy, roz};
  for (int q = 0; q < 3; q++) {
    const c64 v = r[q] * dot - e[q];
    e[q] = c64{-kNecConst4 * v.im, kNecConst4 * v.re};
  }
}

__global__ void k_conductor_far(ConductorFar cf, int pec, size_t count, const double* __restrict__ theta,
                                const double* __restrict__ phi, c64* __restrict__ et, c64* __restrict__ ep) {
  const size_t d = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= count) return;
  double sth, cth, sph, cph;
  sincos(theta[d], &sth, &cth);
  sincos(phi[d], &sph, &cph);
  const double phx = -sph, phy = cph, roz = cth, thx = roz * phy, thy = -roz * phx, thz = -sth;
  const double rox = -thz * phy, roy = thz * phx;
  c64 ci[3] = {{0, 0}, {0, 0}, {0, 0}};
  if (cf.n) {
    segment_sum(cf, rox, roy, roz, ci);
    if (pec) {
      c64 im[3];
      segment_sum(cf, rox, roy, -roz, im);
      ci[0] = ci[0] - im[0];
      ci[1] = ci[1] - im[1];
      ci[2] = ci[2] + im[2];
    }
  }
  c64 e[3] = {{0, 0}, {0, 0}, {0, 0}};
  if (cf.m) {
    patch_sum(cf, rox, roy, roz, e);
    if (pec) {
      c64 g[3];
      patch_sum(cf, rox, roy, -roz, g);  // the image patch, negated over PEC
      e[0] = e[0] - g[0];
      e[1] = e[1] - g[1];
      e[2] = e[2] + g[2];
    }
  }
  // + CONST3 ci
  for (int q = 0; q < 3; q++) e[q] = e[q] + c64{-kNecConst3 * ci[q].im, kNecConst3 * ci[q].re};
  et[d] = thx * e[0] + thy * e[1] + thz * e[2];
  ep[d] = phx * e[0] + phy * e[1];
}
}  // namespace

int conductor_far_field(const ConductorFar& cf, int pec, size_t count, const double* theta, const double* phi,
                        c64* et, c64* ep, hipStream_t s) {
  if (count == 0) return 1;
  hipLaunchKernelGGL(k_conductor_far, dim3((unsigned)((count + 127) / 128)), dim3(128), 0, s, cf, pec, count, theta,
                     phi, et, ep);
  return hipGetLastError() == hipSuccess;
}

unsigned farfield_slices(size_t n, size_t ndir, int ground) {
  const size_t dt = ground ? 2 : 4;  // directions per workgroup (launch_sum below)
  const size_t tiles = (ndir + dt - 1) / dt;
  size_t per = (n + FWG - 1) / FWG;  // workgroups that give each thread one cell
  size_t want = tiles >= 2048 ? 1 : (2048 + tiles - 1) / tiles;  // at least ~2048 workgroups in all
  size_t sl = per < want ? per : want;
  if (sl > 256) sl = 256;
  return (unsigned)(sl < 1 ? 1 : sl);
}

int farfield_amplitudes(const double* centers, const double* sizes, int uniform, const c64* current, size_t n,
                        const double* dirs, size_t ndir, double k0, int ground, const c64* rs, const c64* rp,
                        c64* partial, unsigned slices, c64* amplitude, hipStream_t s) {
  if (ndir == 0) return 1;
  if (ground) {
    if (uniform) launch_sum<2, true, true>(centers, sizes, current, n, dirs, ndir, k0, partial, slices, s);
    else launch_sum<2, true, false>(centers, sizes, current, n, dirs, ndir, k0, partial, slices, s);
  } else {
    if (uniform) launch_sum<4, false, true>(centers, sizes, current, n, dirs, ndir, k0, partial, slices, s);
    else launch_sum<4, false, false>(centers, sizes, current, n, dirs, ndir, k0, partial, slices, s);
  }
  hipLaunchKernelGGL(k_farfield_finish, dim3((unsigned)((ndir + 255) / 256)), dim3(256), 0, s, partial, slices,
                     ndir, dirs, k0, ground, rs, rp, amplitude);
  return hipGetLastError() == hipSuccess;
}

}  // namespace vie
