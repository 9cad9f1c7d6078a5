// FP64 Green-tensor kernel spectrum on the GPU.
//
// For every sorted lattice offset v = (a >= b >= c) that some axis
// permutation places inside the kernel shape (vie_fft.c:347-353):
//   * far pairs (gap^2 > 4) use the triangular-weight difference-coordinate
//     Gauss rule at orders 4, 6, 8, 12 until successive orders agree to
//     2e-12 + 2e-9 |N| (vie_fft.c:299-335) -- one wavefront per offset;
//   * the rest (self, touching and near pairs, and any far pair that does not
//     settle) use the face-pair-reduced, Duffy-regularised cuboid integral of
//     vie_kernel.c:550-641 with the order ladder {4,...,64}, rtol 5e-7,
//     atol 5e-10, far order 16 (vie_fft.c:344).  The host splits each face
//     pair into reduced boxes (geometry only); every quadrature sum runs on
//     the GPU in FP64, one workgroup per reduced box.
// Off-diagonals are averaged and zeroed on a zero offset axis (vie_fft.c:356-359).
//
// The padded tensor is even (diagonal components) or odd (off-diagonal, along
// each of its two axes) under the sign images the reference spreads
// (vie_fft.c:363-373), so its DFT is the separable cosine / (-2i) sine sum
// over the positive octant.  Only the (Px/2+1)(Py/2+1)(Pz/2+1) octant of the
// spectrum is stored; the matvec folds the signs back.
#include <math.h>
#include <string.h>
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <vector>
#include "vie_kernels.h"

namespace vie {

static const double kPi = 3.141592653589793238462643383279502884;

// ------------------------------------------------------------------ Gauss
// Legendre nodes on [-1,1] by Newton from the Tricomi start, the recipe of
// vie_kernel.c:144-181 (4 eps stop, 32 iterations).
static bool gauss_pm1(int n, double* x, double* w) {
  if (n < 2 || n > 64) return false;
  for (int r = 0; r < (n + 1) / 2; r++) {
    double z = cos(kPi * ((double)r + 0.75) / ((double)n + 0.5)), z0, dp = 0;
    int it = 0;
    do {
      double p0 = 1.0, p1 = z;
      for (int d = 2; d <= n; d++) {
        double p2 = ((2.0 * d - 1.0) * z * p1 - ((double)d - 1.0) * p0) / (double)d;
        p0 = p1;
        p1 = p2;
      }
      z0 = z;
      dp = (double)n * (z * p1 - p0) / (z * z - 1.0);
      z = z0 - p1 / dp;
      it++;
    } while (fabs(z - z0) > 4.0 * 2.220446049250313e-16 && it < 32);
    if (it == 32) return false;
    x[r] = -z;
    x[n - 1 - r] = z;
    w[r] = w[n - 1 - r] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
  return true;
}

// Legendre nodes mapped to [0,1] with weights summing to one, as the far
// rule uses them (vie_fft.c:284-298, 2e-15 stop).
static void gauss_01(int n, double* x, double* w) {
  for (int i = 0; i < (n + 1) / 2; i++) {
    double z = cos(kPi * (i + .75) / (n + .5)), z0, dp = 0;
    do {
      double p = 1, q = 0;
      z0 = z;
      for (int j = 1; j <= n; j++) {
        double pr = q;
        q = p;
        p = ((2 * j - 1) * z * q - (j - 1) * pr) / j;
      }
      dp = n * (z * p - q) / (z * z - 1);
      z = z0 - p / dp;
    } while (fabs(z - z0) > 2e-15);
    x[i] = (1 - z) / 2;
    x[n - 1 - i] = (1 + z) / 2;
    w[i] = w[n - 1 - i] = 1 / ((1 - z * z) * dp * dp);
  }
}

// ------------------------------------------------------------------ far rule
struct FarRule {
  double x[4][12], w[4][12];  // orders 4, 6, 8, 12 on [0,1]
};
__constant__ FarRule c_far;
static const int kFarOrders[4] = {4, 6, 8, 12};

__device__ inline double wave_sum(double v) {
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// One 64-lane wavefront per offset.  out: 6 complex (xx yy zz xy xz yz).
// status: accepted order or 0 when the ladder does not settle.
__global__ __launch_bounds__(256) void k_far(const int* offs, int count, double h, double k,
                                             c64* out, int* status) {
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (wave >= count) return;
  const int v[3] = {offs[3 * wave], offs[3 * wave + 1], offs[3 * wave + 2]};
  double prev[12], cur[12];
  int accepted = 0;
  for (int st = 0; st < 4; st++) {
    const int n = kFarOrders[st];
    const int pts = n * n * n * 8;
    double acc[12];
#pragma unroll
    for (int q = 0; q < 12; q++) acc[q] = 0;
    for (int p = lane; p < pts; p += 64) {
      int m = p & 7, t3 = p >> 3;
      int c = t3 % n, b = (t3 / n) % n, a = t3 / (n * n);
      double t[3] = {c_far.x[st][a], c_far.x[st][b], c_far.x[st][c]};
      double wt = h * h * h * c_far.w[st][a] * c_far.w[st][b] * c_far.w[st][c] * (1 - t[0]) * (1 - t[1]) *
                  (1 - t[2]);
      double r[3], r2 = 0;
#pragma unroll
      for (int i = 0; i < 3; i++) {
        r[i] = h * ((double)v[i] + ((m & (1 << i)) ? -t[i] : t[i]));
        r2 += r[i] * r[i];
      }
      double rr = sqrt(r2), sn, cs;
      sincos(k * rr, &sn, &cs);
      double gs = wt / (4 * kPi * rr);
      c64 g = {gs * cs, -gs * sn};
      c64 dg = g * c64{k * k - 1 / r2, -k / rr};
      c64 rad = g * c64{(-k * k + 3 / r2) / r2, (3 * k / rr) / r2};
      // xx yy zz xy xz yz
      acc[0] += dg.re + rad.re * r[0] * r[0];
      acc[1] += dg.im + rad.im * r[0] * r[0];
      acc[2] += dg.re + rad.re * r[1] * r[1];
      acc[3] += dg.im + rad.im * r[1] * r[1];
      acc[4] += dg.re + rad.re * r[2] * r[2];
      acc[5] += dg.im + rad.im * r[2] * r[2];
      acc[6] += rad.re * r[0] * r[1];
      acc[7] += rad.im * r[0] * r[1];
      acc[8] += rad.re * r[0] * r[2];
      acc[9] += rad.im * r[0] * r[2];
      acc[10] += rad.re * r[1] * r[2];
      acc[11] += rad.im * r[1] * r[2];
    }
#pragma unroll
    for (int q = 0; q < 12; q++) cur[q] = wave_sum(acc[q]);
    if (st > 0) {
      bool ok = true;
      for (int q = 0; q < 6; q++) {
        double re = cur[2 * q], im = cur[2 * q + 1];
        double dr = re - prev[2 * q], di = im - prev[2 * q + 1];
        if (!isfinite(re) || !isfinite(im) || hypot(dr, di) > 2e-12 + 2e-9 * hypot(re, im)) ok = false;
      }
      if (ok) {
        accepted = n;
        break;
      }
    }
#pragma unroll
    for (int q = 0; q < 12; q++) prev[q] = cur[q];
  }
  if (lane == 0) {
    status[wave] = accepted;
    if (accepted)
      for (int q = 0; q < 6; q++) out[6 * wave + q] = c64{cur[2 * q], cur[2 * q + 1]};
  }
}

// ------------------------------------------------------------------ near jobs
// One reduced box of the face-pair integral (vie_kernel.c:200-316).
struct Piece {
  int overlap;
  double lo, hi, amin, amax, bmin, bmax;
};
struct QuadJob {
  int dims, order, duffy, sector, power, node0;
  double fixed2;
  Piece p[3];
};

__device__ inline double piece_weight(const Piece& p, double u) {
  if (!p.overlap) return 1.0;
  double lo = fmax(p.amin, p.bmin + u), hi = fmin(p.amax, p.bmax + u);
  return hi > lo ? hi - lo : 0.0;
}

__global__ __launch_bounds__(256) void k_quad(const QuadJob* jobs, const double* nodes, const double* weights,
                                              double k, c64* out) {
  __shared__ double red[2][256];
  const QuadJob J = jobs[blockIdx.x];
  const double* xn = nodes + J.node0;
  const double* wn = weights + J.node0;
  int n = J.order, pts = n;
  for (int d = 1; d < J.dims; d++) pts *= n;
  double sr = 0, si = 0;
  for (int p = threadIdx.x; p < pts; p += blockDim.x) {
    int idx[3] = {0, 0, 0}, rem = p;
    for (int d = J.dims - 1; d >= 0; d--) {
      idx[d] = rem % n;
      rem /= n;
    }
    double r2 = J.fixed2, wgt = 1.0;
    if (!J.duffy) {
      for (int d = 0; d < J.dims; d++) {
        const Piece& pc = J.p[d];
        double mid = 0.5 * (pc.lo + pc.hi), half = 0.5 * (pc.hi - pc.lo);
        double u = mid + half * xn[idx[d]];
        wgt *= half * wn[idx[d]] * piece_weight(pc, u);
        r2 += u * u;
      }
    } else {
      double q[3];
      for (int d = 0; d < J.dims; d++) {
        q[d] = 0.5 * (1.0 + xn[idx[d]]);
        wgt *= 0.5 * wn[idx[d]];
      }
      double rp = q[J.sector];
      double t = pow(rp, (double)J.power), jac = 1.0, pw = 1.0;
      for (int d = 0; d < J.dims; d++) {
        const Piece& pc = J.p[d];
        double ext = fabs(pc.hi - pc.lo);
        double org = fabs(pc.lo) <= fabs(pc.hi) ? pc.lo : pc.hi;
        double sg = org == pc.lo ? 1.0 : -1.0;
        double m = d == J.sector ? t : t * q[d];
        double c = org + sg * ext * m;
        r2 += c * c;
        jac *= ext;
        pw *= piece_weight(pc, c);
      }
      jac *= pow(t, (double)(J.dims - 1)) * (double)J.power * pow(rp, (double)(J.power - 1));
      wgt *= jac * pw;
    }
    double r = sqrt(r2), sn, cs;
    sincos(k * r, &sn, &cs);
    double gs = wgt / (4.0 * kPi * r);
    sr += gs * cs;
    si -= gs * sn;
  }
  red[0][threadIdx.x] = sr;
  red[1][threadIdx.x] = si;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = c64{red[0][0], red[1][0]};
}

// Host geometry: split one axis of a face pair into pieces (vie_kernel.c:318-407).
static const double kGeomTol = 64.0 * 2.220446049250313e-16;

static bool add_break(double* b, int& n, double v, double tol) {
  if (fabs(v) <= tol) v = 0.0;
  for (int i = 0; i < n; i++)
    if (fabs(v - b[i]) <= tol) return true;
  if (n >= 6) return false;
  b[n++] = v;
  return true;
}

static void sort_small(double* b, int n) {
  for (int i = 1; i < n; i++) {
    double v = b[i];
    int c = i;
    while (c > 0 && b[c - 1] > v) {
      b[c] = b[c - 1];
      c--;
    }
    b[c] = v;
  }
}

static bool axis_pieces(double amin, double amax, bool av, double bmin, double bmax, bool bv,
                        std::vector<Piece>& out, double& fixed) {
  double tol = kGeomTol * std::max(std::max(fabs(amin), fabs(amax)), std::max(fabs(bmin), fabs(bmax)));
  double br[6];
  int nb = 0;
  out.clear();
  if (!av && !bv) {
    fixed = amin - bmin;
    if (fabs(fixed) <= tol) fixed = 0.0;
    return true;
  }
  if (av && bv) {
    if (!add_break(br, nb, amin - bmax, tol) || !add_break(br, nb, amin - bmin, tol) ||
        !add_break(br, nb, amax - bmax, tol) || !add_break(br, nb, amax - bmin, tol))
      return false;
  } else if (av) {
    add_break(br, nb, amin - bmin, tol);
    add_break(br, nb, amax - bmin, tol);
  } else {
    add_break(br, nb, amin - bmax, tol);
    add_break(br, nb, amin - bmin, tol);
  }
  sort_small(br, nb);
  if (br[0] < 0.0 && br[nb - 1] > 0.0) {
    add_break(br, nb, 0.0, tol);
    sort_small(br, nb);
  }
  for (int i = 0; i + 1 < nb; i++) {
    if (br[i + 1] == br[i]) continue;
    if (out.size() >= 5) return false;
    out.push_back(Piece{av && bv, br[i], br[i + 1], amin, amax, bmin, bmax});
  }
  return true;
}

struct Cube {
  double c[3], s[3];
};

// Appends the reduced boxes of one face pair; returns false on a geometry
// failure (reported by the reference as a NaN integral).
static bool face_pair_jobs(const Cube& t, const Cube& s, int ta, int ts, int sa, int ss, int order, int node0,
                           std::vector<QuadJob>& jobs) {
  std::vector<Piece> P[3];
  double fixed[3] = {0, 0, 0};
  for (int a = 0; a < 3; a++) {
    bool tv = a != ta, sv = a != sa;
    double tmin = tv ? -0.5 * t.s[a] : 0.5 * ts * t.s[a];
    double tmax = tv ? 0.5 * t.s[a] : tmin;
    double sc = s.c[a] - t.c[a];
    double smin = sv ? sc - 0.5 * s.s[a] : sc + 0.5 * ss * s.s[a];
    double smax = sv ? sc + 0.5 * s.s[a] : smin;
    if (!axis_pieces(tmin, tmax, tv, smin, smax, sv, P[a], fixed[a])) return false;
  }
  int cnt[3];
  for (int a = 0; a < 3; a++) cnt[a] = (int)P[a].size();
  for (int i = 0; i < (cnt[0] ? cnt[0] : 1); i++)
    for (int j = 0; j < (cnt[1] ? cnt[1] : 1); j++)
      for (int l = 0; l < (cnt[2] ? cnt[2] : 1); l++) {
        int sel[3] = {i, j, l};
        QuadJob J;
        memset(&J, 0, sizeof(J));
        J.order = order;
        J.node0 = node0;
        J.power = 1;
        for (int a = 0; a < 3; a++) {
          if (cnt[a]) J.p[J.dims++] = P[a][sel[a]];
          else J.fixed2 += fixed[a] * fixed[a];
        }
        double near2 = J.fixed2, big = 0.0;
        int sing = 1;
        for (int a = 0; a < J.dims; a++) {
          double nr = std::min(fabs(J.p[a].lo), fabs(J.p[a].hi));
          near2 += nr * nr;
          big = std::max(big, J.p[a].hi - J.p[a].lo);
          sing &= (J.p[a].lo == 0.0 || J.p[a].hi == 0.0);
        }
        // radial grading (vie_kernel.c:448-453)
        double ratio = sqrt(near2) / big;
        while (ratio > 0.0 && ratio < 0.1 && J.power < 8) {
          J.power++;
          ratio *= 10.0;
        }
        // tensor rule unless singular or close (vie_kernel.c:303-304)
        if (!sing && sqrt(near2) >= 0.1 * big) {
          J.duffy = 0;
          jobs.push_back(J);
        } else {
          J.duffy = 1;
          for (int sct = 0; sct < J.dims; sct++) {
            J.sector = sct;
            jobs.push_back(J);
          }
        }
      }
  return true;
}

// Per face pair bookkeeping: which jobs belong to it and how it feeds N.
struct PairTag {
  int first, count, ta, sa, sign;
};

// Evaluates N at one quadrature order for several offsets in one launch.
static bool near_matrices(const std::vector<Cube>& tests, const Cube& src, int order, double k,
                          std::vector<std::array<c64, 9>>& result, hipStream_t s) {
  double xn[64], wn[64];
  if (!gauss_pm1(order, xn, wn)) return false;
  std::vector<QuadJob> jobs;
  std::vector<std::vector<PairTag>> tags(tests.size());
  for (size_t o = 0; o < tests.size(); o++) {
    for (int ta = 0; ta < 3; ta++)
      for (int sa = 0; sa < 3; sa++)
        for (int ts = -1; ts <= 1; ts += 2)
          for (int ss = -1; ss <= 1; ss += 2) {
            PairTag tg{(int)jobs.size(), 0, ta, sa, ts * ss};
            if (!face_pair_jobs(tests[o], src, ta, ts, sa, ss, order, 0, jobs)) return false;
            tg.count = (int)jobs.size() - tg.first;
            tags[o].push_back(tg);
          }
  }
  result.assign(tests.size(), std::array<c64, 9>{});
  if (jobs.empty()) return true;
  QuadJob* dj = nullptr;
  double *dx = nullptr, *dw = nullptr;
  c64* dout = nullptr;
  bool ok = hipMalloc(&dj, jobs.size() * sizeof(QuadJob)) == hipSuccess &&
            hipMalloc(&dx, 64 * sizeof(double)) == hipSuccess && hipMalloc(&dw, 64 * sizeof(double)) == hipSuccess &&
            hipMalloc(&dout, jobs.size() * sizeof(c64)) == hipSuccess;
  std::vector<c64> hout(jobs.size());
  if (ok) {
    ok = hipMemcpyAsync(dj, jobs.data(), jobs.size() * sizeof(QuadJob), hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(dx, xn, order * sizeof(double), hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(dw, wn, order * sizeof(double), hipMemcpyHostToDevice, s) == hipSuccess;
    if (ok) {
      hipLaunchKernelGGL(k_quad, dim3((unsigned)jobs.size()), dim3(256), 0, s, dj, dx, dw, k, dout);
      ok = hipGetLastError() == hipSuccess &&
           hipMemcpyAsync(hout.data(), dout, jobs.size() * sizeof(c64), hipMemcpyDeviceToHost, s) == hipSuccess &&
           hipStreamSynchronize(s) == hipSuccess;
    }
  }
  (void)hipFree(dj);
  (void)hipFree(dx);
  (void)hipFree(dw);
  (void)hipFree(dout);
  if (!ok) return false;
  // face-pair reduction into N (vie_kernel.c:500-521)
  for (size_t o = 0; o < tests.size(); o++) {
    const Cube& t = tests[o];
    double vol = t.s[0] * t.s[1] * t.s[2];
    auto& N = result[o];
    for (const PairTag& tg : tags[o]) {
      c64 sum = {0, 0};
      for (int q = 0; q < tg.count; q++) sum = sum + hout[tg.first + q];
      c64 v = {sum.re / vol, sum.im / vol};
      if (!std::isfinite(v.re) || !std::isfinite(v.im)) return false;
      if (tg.ta == tg.sa) {
        for (int c = 0; c < 3; c++)
          if (c != tg.ta) N[c * 3 + c] = N[c * 3 + c] + (double)tg.sign * v;
      } else {
        N[tg.sa * 3 + tg.ta] = N[tg.sa * 3 + tg.ta] - (double)tg.sign * v;
      }
    }
  }
  return true;
}

// Order ladder with the reference's acceptance test (vie_kernel.c:607-640);
// offsets that settle drop out of later launches.
static bool near_ladder(const std::vector<Cube>& tests, const Cube& src, double k, double rtol, double atol,
                        std::vector<std::array<c64, 9>>& out, int* max_order, hipStream_t s) {
  static const int ladder[] = {4, 6, 8, 12, 16, 24, 32, 40, 48, 56, 64};
  size_t n = tests.size();
  out.assign(n, std::array<c64, 9>{});
  std::vector<std::array<c64, 9>> prev(n);
  std::vector<int> pending(n);
  for (size_t i = 0; i < n; i++) pending[i] = (int)i;
  bool have = false;
  for (int order : ladder) {
    if (pending.empty()) break;
    std::vector<Cube> sub;
    for (int i : pending) sub.push_back(tests[i]);
    std::vector<std::array<c64, 9>> cur;
    if (!near_matrices(sub, src, order, k, cur, s)) return false;
    std::vector<int> still;
    for (size_t q = 0; q < pending.size(); q++) {
      int i = pending[q];
      bool accept = false;
      if (have) {
        double worst = 0;
        for (int e = 0; e < 9; e++) {
          c64 l = prev[i][e], r = cur[q][e];
          double scale = atol + rtol * std::max(hypot(l.re, l.im), hypot(r.re, r.im));
          double ratio = hypot(l.re - r.re, l.im - r.im) / scale;
          if (ratio > worst) worst = ratio;
        }
        accept = worst <= 1.0;
      }
      if (accept) {
        out[i] = cur[q];
        if (max_order && order > *max_order) *max_order = order;
      } else {
        prev[i] = cur[q];
        still.push_back(i);
      }
    }
    pending.swap(still);
    have = true;
  }
  return pending.empty();
}

// ------------------------------------------------------------------ spread + transforms
// Writes the positive-octant tensor Ts[6][kx][ky][kz] for each offset in the
// reference's permutation order (later permutations win, vie_fft.c:363-373).
__global__ void k_spread(const int* offs, const c64* tens, int count, int sx, int sy, int sz, c64* Ts) {
  int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= count) return;
  const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  const int v[3] = {offs[3 * o], offs[3 * o + 1], offs[3 * o + 2]};
  // full symmetric 3x3 from packed xx yy zz xy xz yz
  c64 n[3][3];
  const c64* t = tens + 6 * o;
  n[0][0] = t[0]; n[1][1] = t[1]; n[2][2] = t[2];
  n[0][1] = n[1][0] = t[3];
  n[0][2] = n[2][0] = t[4];
  n[1][2] = n[2][1] = t[5];
  const size_t vol = (size_t)sx * sy * sz;
  const int pairs[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
  for (int p = 0; p < 6; p++) {
    const int* q = perm[p];
    int d0 = v[q[0]], d1 = v[q[1]], d2 = v[q[2]];
    if (d0 >= sx || d1 >= sy || d2 >= sz) continue;
    size_t g = ((size_t)d0 * sy + d1) * sz + d2;
    for (int c = 0; c < 6; c++) Ts[c * vol + g] = n[q[pairs[c][0]]][q[pairs[c][1]]];
  }
}

// Even/odd DFT along one axis of a [6][n0][n1][n2] array (axis ax) into
// [6][..H..] where H = P/2+1 along ax.  odd_mask bit c: component c odd on ax.
__global__ void k_axis_transform(const c64* in, c64* out, int n0, int n1, int n2, int ax, int P, int odd_mask,
                                 const double* cosT, const double* sinT) {
  int nin[3] = {n0, n1, n2};
  int nout[3] = {n0, n1, n2};
  nout[ax] = P / 2 + 1;
  size_t total = (size_t)6 * nout[0] * nout[1] * nout[2];
  size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= total) return;
  size_t r = id;
  int o2 = r % nout[2]; r /= nout[2];
  int o1 = r % nout[1]; r /= nout[1];
  int o0 = r % nout[0]; r /= nout[0];
  int c = (int)r;
  int oi[3] = {o0, o1, o2};
  int kk = oi[ax];
  size_t stride = ax == 2 ? 1 : ax == 1 ? (size_t)nin[2] : (size_t)nin[1] * nin[2];
  oi[ax] = 0;
  size_t base = (size_t)c * nin[0] * nin[1] * nin[2] + ((size_t)oi[0] * nin[1] + oi[1]) * nin[2] + oi[2];
  bool odd = (odd_mask >> c) & 1;
  double sr = 0, si = 0;
  int n = nin[ax];
  if (!odd) {
    c64 v0 = in[base];
    sr = v0.re;
    si = v0.im;
    for (int d = 1; d < n; d++) {
      c64 v = in[base + d * stride];
      double w = 2.0 * cosT[((long)kk * d) % P];
      sr += w * v.re;
      si += w * v.im;
    }
  } else {
    double ar = 0, ai = 0;
    for (int d = 1; d < n; d++) {
      c64 v = in[base + d * stride];
      double w = 2.0 * sinT[((long)kk * d) % P];
      ar += w * v.re;
      ai += w * v.im;
    }
    // -i (ar + i ai) = ai - i ar
    sr = ai;
    si = -ar;
  }
  out[id] = c64{sr, si};
}

__global__ void k_finish_spectrum(const c64* in, size_t n, double scale, c64* s64, c32* s32) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  c64 v = {in[i].re * scale, in[i].im * scale};
  if (s64) s64[i] = v;
  if (s32) s32[i] = c32{(float)v.re, (float)v.im};
}

static void trig_table(int P, std::vector<double>& c, std::vector<double>& s) {
  c.resize(P);
  s.resize(P);
  for (int m = 0; m < P; m++) {
    // exact values on the quadrant points so odd components vanish exactly
    if (4 * m % P == 0) {
      int q = (4 * m) / P;
      static const double C[4] = {1, 0, -1, 0}, S[4] = {0, 1, 0, -1};
      c[m] = C[q & 3];
      s[m] = S[q & 3];
    } else {
      c[m] = cos(2 * kPi * m / P);
      s[m] = sin(2 * kPi * m / P);
    }
  }
}

int build_spectrum(const int ks[3], const int P[3], double h, double k, c64* spec64, c32* spec32,
                   SpectrumResult* res, hipStream_t s) {
  SpectrumResult local;
  if (!res) res = &local;
  *res = SpectrumResult();
  // unique sorted offsets that fit some permutation (vie_fft.c:347-353)
  static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  int mx = std::max(ks[0], std::max(ks[1], ks[2]));
  std::vector<int> offs;
  std::vector<int> far_idx, near_idx;
  for (int a = 0; a < mx; a++)
    for (int b = 0; b <= a; b++)
      for (int c = 0; c <= b; c++) {
        int v[3] = {a, b, c};
        bool useful = false;
        for (int p = 0; p < 6; p++)
          if (v[perm[p][0]] < ks[0] && v[perm[p][1]] < ks[1] && v[perm[p][2]] < ks[2]) useful = true;
        if (!useful) continue;
        double gap2 = 0;
        for (int i = 0; i < 3; i++) {
          double gp = v[i] > 1 ? (double)v[i] - 1 : 0;
          gap2 += gp * gp;
        }
        int idx = (int)(offs.size() / 3);
        offs.insert(offs.end(), v, v + 3);
        (gap2 > 4 ? far_idx : near_idx).push_back(idx);
      }
  const int count = (int)(offs.size() / 3);
  std::vector<c64> tens((size_t)count * 6);
  // far offsets on the GPU
  FarRule fr;
  for (int st = 0; st < 4; st++) gauss_01(kFarOrders[st], fr.x[st], fr.w[st]);
  int* d_offs = nullptr;
  c64* d_tens = nullptr;
  int* d_status = nullptr;
  bool ok = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_far), &fr, sizeof(fr), 0, hipMemcpyHostToDevice, s) == hipSuccess;
  std::vector<int> far_offs;
  for (int i : far_idx) far_offs.insert(far_offs.end(), &offs[3 * i], &offs[3 * i] + 3);
  std::vector<int> status(far_idx.size(), 0);
  std::vector<c64> far_out(far_idx.size() * 6);
  if (ok && !far_idx.empty()) {
    size_t nf = far_idx.size();
    ok = hipMalloc(&d_offs, nf * 3 * sizeof(int)) == hipSuccess &&
         hipMalloc(&d_tens, nf * 6 * sizeof(c64)) == hipSuccess &&
         hipMalloc(&d_status, nf * sizeof(int)) == hipSuccess &&
         hipMemcpyAsync(d_offs, far_offs.data(), nf * 3 * sizeof(int), hipMemcpyHostToDevice, s) == hipSuccess;
    if (ok) {
      unsigned blocks = (unsigned)((nf * 64 + 255) / 256);
      hipLaunchKernelGGL(k_far, dim3(blocks), dim3(256), 0, s, d_offs, (int)nf, h, k, d_tens, d_status);
      ok = hipGetLastError() == hipSuccess &&
           hipMemcpyAsync(status.data(), d_status, nf * sizeof(int), hipMemcpyDeviceToHost, s) == hipSuccess &&
           hipMemcpyAsync(far_out.data(), d_tens, nf * 6 * sizeof(c64), hipMemcpyDeviceToHost, s) == hipSuccess &&
           hipStreamSynchronize(s) == hipSuccess;
    }
  }
  (void)hipFree(d_offs);
  (void)hipFree(d_tens);
  (void)hipFree(d_status);
  if (!ok) return 0;
  std::vector<int> cuboid_idx = near_idx;
  for (size_t q = 0; q < far_idx.size(); q++) {
    if (status[q]) {
      for (int c = 0; c < 6; c++) tens[6 * (size_t)far_idx[q] + c] = far_out[6 * q + c];
    } else {
      cuboid_idx.push_back(far_idx[q]);
      res->fallback_offsets++;
    }
  }
  res->far_offsets = (int)far_idx.size() - res->fallback_offsets;
  res->near_offsets = (int)near_idx.size();
  // near / fallback offsets: certified cuboid integral (vie_fft.c:344, 355)
  if (!cuboid_idx.empty()) {
    std::vector<Cube> tests, far_tests;
    std::vector<int> near_list, farclass_list;
    Cube src = {{0, 0, 0}, {h, h, h}};
    for (int i : cuboid_idx) {
      Cube t = {{h * offs[3 * i], h * offs[3 * i + 1], h * offs[3 * i + 2]}, {h, h, h}};
      // relation FAR when the separation exceeds twice the cell (vie_kernel.c:122)
      double d2 = 0;
      int sep = 0;
      for (int a = 0; a < 3; a++) {
        double cd = fabs(t.c[a]), hs = h, tol = kGeomTol * std::max(1.0, std::max(cd, hs)), gap = cd - hs;
        if (gap > tol) {
          sep++;
          d2 += gap * gap;
        }
      }
      if (sep && sqrt(d2) > 2.0 * h) {
        far_tests.push_back(t);
        farclass_list.push_back(i);
      } else {
        tests.push_back(t);
        near_list.push_back(i);
      }
    }
    std::vector<std::array<c64, 9>> N;
    if (!tests.empty()) {
      if (!near_ladder(tests, src, k, 5e-7, 5e-10, N, &res->max_near_order, s)) return 0;
      for (size_t q = 0; q < tests.size(); q++) {
        c64* t = &tens[6 * (size_t)near_list[q]];
        const auto& M = N[q];
        t[0] = M[0]; t[1] = M[4]; t[2] = M[8];
        t[3] = 0.5 * (M[1] + M[3]);
        t[4] = 0.5 * (M[2] + M[6]);
        t[5] = 0.5 * (M[5] + M[7]);
      }
    }
    if (!far_tests.empty()) {
      // classified far inside vie_cuboid_n: single evaluation at far order 16
      if (!near_matrices(far_tests, src, 16, k, N, s)) return 0;
      for (size_t q = 0; q < far_tests.size(); q++) {
        c64* t = &tens[6 * (size_t)farclass_list[q]];
        const auto& M = N[q];
        t[0] = M[0]; t[1] = M[4]; t[2] = M[8];
        t[3] = 0.5 * (M[1] + M[3]);
        t[4] = 0.5 * (M[2] + M[6]);
        t[5] = 0.5 * (M[5] + M[7]);
      }
    }
  }
  // symmetrise and zero on zero axes (vie_fft.c:356-359); far values are
  // symmetric by construction, their average is themselves.
  for (int o = 0; o < count; o++) {
    const int* v = &offs[3 * o];
    c64* t = &tens[6 * (size_t)o];
    if (!(v[0] && v[1])) t[3] = c64{0, 0};
    if (!(v[0] && v[2])) t[4] = c64{0, 0};
    if (!(v[1] && v[2])) t[5] = c64{0, 0};
    if (v[0] == 0) {
      res->self[0] = t[0];
      res->self[1] = t[1];
      res->self[2] = t[2];
    }
  }
  // spread, then separable even/odd transforms z, y, x
  const size_t kvol = (size_t)ks[0] * ks[1] * ks[2];
  const int H[3] = {P[0] / 2 + 1, P[1] / 2 + 1, P[2] / 2 + 1};
  size_t s1 = 6 * (size_t)ks[0] * ks[1] * H[2];
  size_t s2 = 6 * (size_t)ks[0] * H[1] * H[2];
  size_t s3 = 6 * (size_t)H[0] * H[1] * H[2];
  size_t big = std::max(std::max(6 * kvol, s1), std::max(s2, s3));
  c64 *w0 = nullptr, *w1 = nullptr, *d_t = nullptr;
  int* d_o = nullptr;
  double* d_trig = nullptr;
  std::vector<double> trig;
  size_t trig_off[3];
  for (int a = 0; a < 3; a++) {
    std::vector<double> c, sn;
    trig_table(P[a], c, sn);
    trig_off[a] = trig.size();
    trig.insert(trig.end(), c.begin(), c.end());
    trig.insert(trig.end(), sn.begin(), sn.end());
  }
  ok = hipMalloc(&w0, big * sizeof(c64)) == hipSuccess && hipMalloc(&w1, big * sizeof(c64)) == hipSuccess &&
       hipMalloc(&d_t, (size_t)count * 6 * sizeof(c64)) == hipSuccess &&
       hipMalloc(&d_o, (size_t)count * 3 * sizeof(int)) == hipSuccess &&
       hipMalloc(&d_trig, trig.size() * sizeof(double)) == hipSuccess;
  if (ok) {
    ok = hipMemsetAsync(w0, 0, 6 * kvol * sizeof(c64), s) == hipSuccess &&
         hipMemcpyAsync(d_t, tens.data(), (size_t)count * 6 * sizeof(c64), hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(d_o, offs.data(), (size_t)count * 3 * sizeof(int), hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(d_trig, trig.data(), trig.size() * sizeof(double), hipMemcpyHostToDevice, s) == hipSuccess;
  }
  if (ok) {
    hipLaunchKernelGGL(k_spread, dim3((count + 127) / 128), dim3(128), 0, s, d_o, d_t, count, ks[0], ks[1], ks[2], w0);
    // odd masks per axis: xy(3) odd in x,y; xz(4) odd in x,z; yz(5) odd in y,z
    const int odd_x = (1 << 3) | (1 << 4), odd_y = (1 << 3) | (1 << 5), odd_z = (1 << 4) | (1 << 5);
    hipLaunchKernelGGL(k_axis_transform, dim3((unsigned)((s1 + 255) / 256)), dim3(256), 0, s, w0, w1, ks[0], ks[1],
                       ks[2], 2, P[2], odd_z, d_trig + trig_off[2], d_trig + trig_off[2] + P[2]);
    hipLaunchKernelGGL(k_axis_transform, dim3((unsigned)((s2 + 255) / 256)), dim3(256), 0, s, w1, w0, ks[0], ks[1],
                       H[2], 1, P[1], odd_y, d_trig + trig_off[1], d_trig + trig_off[1] + P[1]);
    hipLaunchKernelGGL(k_axis_transform, dim3((unsigned)((s3 + 255) / 256)), dim3(256), 0, s, w0, w1, ks[0], H[1],
                       H[2], 0, P[0], odd_x, d_trig + trig_off[0], d_trig + trig_off[0] + P[0]);
    double scale = 1.0 / ((double)P[0] * P[1] * P[2]);
    hipLaunchKernelGGL(k_finish_spectrum, dim3((unsigned)((s3 + 255) / 256)), dim3(256), 0, s, w1, s3, scale, spec64,
                       spec32);
    ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(s) == hipSuccess;
  }
  (void)hipFree(w0);
  (void)hipFree(w1);
  (void)hipFree(d_t);
  (void)hipFree(d_o);
  (void)hipFree(d_trig);
  return ok ? 1 : 0;
}

// ------------------------------------------------------------------ real ground
// Reflected tensor of a passive lower half-space z < 0 (the reference's
// vie_halfspace_cuboids, vie_halfspace.c:73-203, which its dense and GMRES
// paths add to the free-space tensor for GN 2, dielectric_ground.c:116-131),
// for every pair of cells of a lattice of cubes of edge h above the ground.
// The spectral integral over the transverse wavenumber u (propagating: u =
// sin t, w = cos t, t in [0, pi/2]; evanescent: u = hypot(1, ev), w = -i ev)
// and the azimuth phi has an integrand that factors into
//   (terms in u, phi) x exp(-i k u (cos phi dx + sin phi dy) h) x exp(-i k w S) sinc^2(k w h / 2),
// dx, dy the lateral offset in cells and S = z_test + z_source.  So one set
// of nodes serves every pair: per node u the five azimuthal integrals
// J(dx, dy) of {sin^2, cos^2, sin cos, cos, sin} (trapezoidal in phi, enough
// points for the lattice's widest lateral phase), then the tensor for every
// S as a sum over u.  The u nodes are Gauss-Legendre on the reference's
// segments (t in [0, pi/2], then ev in [0, 1, 4, 16, 64] / scale and the
// substrate branch point sqrt(Re eps - 1)), with scale = k x (the smallest
// gap S - h of the lattice): the integrand of every larger S decays faster
// on the same nodes.  Components xx, yy, zz, xy, xz, yz; zx = -xz and zy = -yz
// (the reflected tensor is not symmetric).

struct GroundNode {
  double u;
  c64 w, rs, rp;
  c64 weight;   // -i k^3 / (8 pi^2) h^3 x (radial Jacobian x Gauss weight)
  int nphi, phi0;  // azimuth points and their offset in the per-node tables
};

static std::complex<double> to_std(c64 a) { return {a.re, a.im}; }
static c64 from_std(std::complex<double> a) { return c64{a.real(), a.imag()}; }

// Fresnel coefficients of the lower half-space for transverse wavenumber u/k
// (vie_halfspace.c:14-28 of the reference: outgoing/decaying branch of the
// substrate's normal wavenumber).
static bool ground_fresnel(std::complex<double> e, double u, std::complex<double>* rs, std::complex<double>* rp) {
  using C = std::complex<double>;
  if (e == C(1.0, 0.0)) {
    *rs = *rp = 0.0;
    return true;
  }
  C w = u <= 1 ? C(sqrt((1 - u) * (1 + u)), 0) : C(0, -sqrt((u - 1) * (u + 1)));
  C v = std::sqrt(e - u * u);
  if (v.imag() > 0) v = -v;
  C a = w + v, b = e * w + v;
  if (std::abs(a) == 0 || std::abs(b) == 0) return false;
  *rs = (w - v) / a;
  *rp = (e * w - v) / b;
  return std::isfinite(rs->real()) && std::isfinite(rs->imag()) && std::isfinite(rp->real()) &&
         std::isfinite(rp->imag());
}

// Gauss-Legendre nodes and weights on [-1, 1] (Newton on the Legendre recurrence).
static void gauss_nodes(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0);
  w.assign(n, 0);
  for (int j = 0; j < (n + 1) / 2; j++) {
    double z = cos(kPi * (j + 0.75) / (n + 0.5)), d = 0;
    for (int it = 0; it < 100; it++) {
      double p0 = 1, p1 = z;
      for (int m = 2; m <= n; m++) {
        double p2 = ((2 * m - 1) * z * p1 - (m - 1) * p0) / m;
        p0 = p1;
        p1 = p2;
      }
      d = n * (z * p1 - p0) / (z * z - 1);
      double dz = p1 / d;
      z -= dz;
      if (fabs(dz) < 1e-16) break;
    }
    x[j] = -z;
    x[n - 1 - j] = z;
    w[j] = w[n - 1 - j] = 2 / ((1 - z * z) * d * d);
  }
}

// J_f(node, dx, dy) for the five azimuthal weights, combined into the six
// tensor components (times the node's weight):
//   xx = rs J_ss - rp w^2 J_cc, yy = rs J_cc - rp w^2 J_ss, zz = rp u^2 (J_ss + J_cc),
//   xy = -(rs + rp w^2) J_sc, xz = -rp u w J_c, yz = -rp u w J_s.
// One thread per (node, dx, dy); phi tables hold (k u h cos, k u h sin) and
// dphi shape(u, phi) x {s^2, c^2, s c, c, s} per azimuth point.
__global__ __launch_bounds__(256) void k_ground_angular(const GroundNode* nodes, int U, const double* phase_tab,
                                                        const double* wt_tab, int nx, int ny, c64* G) {
  const size_t lat = (size_t)nx * ny;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)U * lat) return;
  const int q = (int)(id / lat);
  const int l = (int)(id % lat);
  const int dx = l / ny, dy = l % ny;
  const GroundNode nd = nodes[q];
  double jr[5] = {0, 0, 0, 0, 0}, ji[5] = {0, 0, 0, 0, 0};
  for (int j = 0; j < nd.nphi; j++) {
    const double* ph = phase_tab + 2 * (size_t)(nd.phi0 + j);
    const double* wt = wt_tab + 5 * (size_t)(nd.phi0 + j);
    double sn, cs;
    sincos(ph[0] * dx + ph[1] * dy, &sn, &cs);
#pragma unroll
    for (int f = 0; f < 5; f++) {
      jr[f] += wt[f] * cs;
      ji[f] -= wt[f] * sn;
    }
  }
  c64 J[5];
  for (int f = 0; f < 5; f++) J[f] = c64{jr[f], ji[f]};
  const c64 w2 = nd.w * nd.w, uw = nd.u * nd.w, rpw2 = nd.rp * w2, rpuw = nd.rp * uw;
  c64 out[6];
  out[0] = nd.rs * J[0] - rpw2 * J[1];
  out[1] = nd.rs * J[1] - rpw2 * J[0];
  out[2] = (nd.u * nd.u) * (nd.rp * (J[0] + J[1]));
  out[3] = c64{0, 0} - (nd.rs + rpw2) * J[2];
  out[4] = c64{0, 0} - rpuw * J[3];
  out[5] = c64{0, 0} - rpuw * J[4];
  for (int c = 0; c < 6; c++) G[((size_t)q * 6 + c) * lat + l] = nd.weight * out[c];
}

// T[c][dx][dy][s] = sum_q G[q][c][dx][dy] Z[q][s]; one thread per output, s fastest.
__global__ __launch_bounds__(256) void k_ground_sum(const c64* G, const c64* Z, int U, int lat, int L, c64* T) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)6 * lat * L) return;
  const int sidx = (int)(id % L);
  const size_t cl = id / L;  // c * lat + l
  const int c = (int)(cl / lat), l = (int)(cl % lat);
  double ar = 0, ai = 0;
  for (int q = 0; q < U; q++) {
    const c64 g = G[((size_t)q * 6 + c) * lat + l], z = Z[(size_t)q * L + sidx];
    ar += g.re * z.re - g.im * z.im;
    ai += g.re * z.im + g.im * z.re;
  }
  T[id] = c64{ar, ai};
}

// Full DFT along the last axis: out[r][k] = sum_{s < L} in[r][s] exp(-2 pi i k s / P), k < P.
__global__ __launch_bounds__(256) void k_dft_last(const c64* in, size_t rows, int L, int P, const double* cosT,
                                                  const double* sinT, c64* out) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= rows * P) return;
  const int k = (int)(id % P);
  const size_t r = id / P;
  double ar = 0, ai = 0;
  for (int s = 0; s < L; s++) {
    const c64 v = in[r * L + s];
    const int m = (int)(((long)k * s) % P);
    const double cw = cosT[m], sw = sinT[m];  // exp(-i theta) = cw - i sw
    ar += v.re * cw + v.im * sw;
    ai += v.im * cw - v.re * sw;
  }
  out[id] = c64{ar, ai};
}

// Host: the quadrature nodes of the reflected tensor for a lattice of nx x ny
// lateral cells of edge h whose lowest cell centre is at height z0.
static bool ground_nodes(int nx, int ny, double h, double z0, double k, c64 eps, int order,
                         std::vector<GroundNode>& nodes, std::vector<double>& phase_tab, std::vector<double>& wt_tab) {
  using C = std::complex<double>;
  const C e = to_std(eps);
  const double gap = 2 * z0 - h, scale = k * gap;
  if (!(gap > 0) || !std::isfinite(scale) || !(scale > 0) || order < 8) return false;
  std::vector<double> cut = {0, 1, 4, 16, 64};
  const double branch = scale * sqrt(std::max(e.real() - 1, 0.0));
  if (branch > 0 && branch < 64) cut.push_back(branch);
  std::sort(cut.begin(), cut.end());
  std::vector<double> gx, gw;
  gauss_nodes(order, gx, gw);
  const double rho = h * hypot((double)std::max(nx - 1, 0), (double)std::max(ny - 1, 0));
  const C pref = C(0, -1) * (k * k * k / (8 * kPi * kPi)) * (h * h * h);
  nodes.clear();
  phase_tab.clear();
  wt_tab.clear();
  for (int seg = -1; seg < (int)cut.size() - 1; seg++) {
    const double lo = seg < 0 ? 0 : cut[seg], hi = seg < 0 ? kPi / 2 : cut[seg + 1];
    if (hi == lo) continue;
    for (int j = 0; j < order; j++) {
      const double t = 0.5 * ((hi - lo) * gx[j] + hi + lo), rw = 0.5 * (hi - lo) * gw[j];
      double u;
      C w, radial;
      if (seg < 0) {
        u = sin(t);
        w = cos(t);
        radial = sin(t) * rw;
      } else {
        const double ev = t / scale;
        u = hypot(1.0, ev);
        w = C(0, -ev);
        radial = C(0, 1) * rw / scale;
      }
      C rs, rp;
      if (!ground_fresnel(e, u, &rs, &rp)) return false;
      GroundNode nd;
      nd.u = u;
      nd.w = from_std(w);
      nd.rs = from_std(rs);
      nd.rp = from_std(rp);
      nd.weight = from_std(pref * radial);
      // azimuth: periodic trapezoid, enough points for the widest lateral phase k u rho
      int np = (int)ceil(1.25 * k * u * rho + 48.0);
      np = std::max(np, 2 * order);
      np = (np + 3) / 4 * 4;
      nd.nphi = np;
      nd.phi0 = (int)(phase_tab.size() / 2);
      const double dphi = 2 * kPi / np;
      for (int a = 0; a < np; a++) {
        // quadrant points exact so the parity of the lattice sums is exact
        double c, sn;
        if ((4 * a) % np == 0) {
          int qd = 4 * a / np;
          c = qd == 0 ? 1 : qd == 2 ? -1 : 0;
          sn = qd == 1 ? 1 : qd == 3 ? -1 : 0;
        } else {
          c = cos(dphi * a);
          sn = sin(dphi * a);
        }
        auto sinc = [](double v) { return fabs(v) < 1e-5 ? 1 - v * v / 6 + v * v * v * v / 120 : sin(v) / v; };
        const double sx = sinc(0.5 * k * u * c * h), sy = sinc(0.5 * k * u * sn * h);
        const double shape = sx * sx * sy * sy * dphi;
        phase_tab.push_back(k * u * h * c);
        phase_tab.push_back(k * u * h * sn);
        wt_tab.push_back(shape * sn * sn);
        wt_tab.push_back(shape * c * c);
        wt_tab.push_back(shape * sn * c);
        wt_tab.push_back(shape * c);
        wt_tab.push_back(shape * sn);
      }
      nodes.push_back(nd);
    }
  }
  return true;
}

// exp(-i k w S) sinc^2(k w h / 2) for the S of every layer sum s (the product
// of the two cells' z averages, vie_halfspace.c:61-69).
static void ground_z(const std::vector<GroundNode>& nodes, double k, double h, double z0, int L, std::vector<c64>& Z) {
  using C = std::complex<double>;
  Z.resize(nodes.size() * (size_t)L);
  for (size_t q = 0; q < nodes.size(); q++) {
    const C kw = k * to_std(nodes[q].w);
    const C a = 0.5 * kw * h;
    const C sc = std::abs(a) < 1e-4 ? 1.0 - a * a / 6.0 + a * a * a * a / 120.0 : std::sin(a) / a;
    for (int s = 0; s < L; s++) {
      const double S = 2 * z0 + s * h;
      Z[q * L + s] = from_std(std::exp(C(0, -1) * kw * S) * sc * sc);
    }
  }
}

int ground_table(int nx, int ny, int nz, double h, double z0, double k, c64 eps, int order, c64* T_dev,
                 hipStream_t s) {
  std::vector<GroundNode> nodes;
  std::vector<double> ph, wt;
  if (!ground_nodes(nx, ny, h, z0, k, eps, order, nodes, ph, wt)) return 0;
  const int U = (int)nodes.size(), L = 2 * nz - 1;
  const size_t lat = (size_t)nx * ny;
  std::vector<c64> Z;
  ground_z(nodes, k, h, z0, L, Z);
  GroundNode* d_nodes = nullptr;
  double *d_ph = nullptr, *d_wt = nullptr;
  c64 *d_G = nullptr, *d_Z = nullptr;
  bool ok = hipMalloc(&d_nodes, nodes.size() * sizeof(GroundNode)) == hipSuccess &&
            hipMalloc(&d_ph, ph.size() * sizeof(double)) == hipSuccess &&
            hipMalloc(&d_wt, wt.size() * sizeof(double)) == hipSuccess &&
            hipMalloc(&d_G, (size_t)U * 6 * lat * sizeof(c64)) == hipSuccess &&
            hipMalloc(&d_Z, Z.size() * sizeof(c64)) == hipSuccess;
  ok = ok && hipMemcpyAsync(d_nodes, nodes.data(), nodes.size() * sizeof(GroundNode), hipMemcpyHostToDevice, s) ==
                 hipSuccess &&
       hipMemcpyAsync(d_ph, ph.data(), ph.size() * sizeof(double), hipMemcpyHostToDevice, s) == hipSuccess &&
       hipMemcpyAsync(d_wt, wt.data(), wt.size() * sizeof(double), hipMemcpyHostToDevice, s) == hipSuccess &&
       hipMemcpyAsync(d_Z, Z.data(), Z.size() * sizeof(c64), hipMemcpyHostToDevice, s) == hipSuccess;
  if (ok) {
    const size_t n1 = (size_t)U * lat, n2 = 6 * lat * (size_t)L;
    hipLaunchKernelGGL(k_ground_angular, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, s, d_nodes, U, d_ph, d_wt,
                       nx, ny, d_G);
    hipLaunchKernelGGL(k_ground_sum, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, s, d_G, d_Z, U, (int)lat, L,
                       T_dev);
    ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(s) == hipSuccess;
  }
  (void)hipFree(d_nodes);
  (void)hipFree(d_ph);
  (void)hipFree(d_wt);
  (void)hipFree(d_G);
  (void)hipFree(d_Z);
  return ok ? 1 : 0;
}

int build_ground_spectrum(int nx, int ny, int nz, const int padded[3], double h, double z0, double k, c64 eps,
                          int order, c64* rspec64, c32* rspec32, hipStream_t s) {
  const int L = 2 * nz - 1, Px = padded[0], Py = padded[1], Pz = padded[2];
  if (L > Pz) return 0;
  const int Hx = Px / 2 + 1, Hy = Py / 2 + 1;
  const size_t s0 = 6 * (size_t)nx * ny * L, s1 = 6 * (size_t)nx * ny * Pz, s2 = 6 * (size_t)Hx * ny * Pz,
               s3 = 6 * (size_t)Hx * Hy * Pz;
  const size_t big = std::max(std::max(s0, s1), std::max(s2, s3));
  std::vector<double> trig;
  size_t off[3];
  const int P[3] = {Px, Py, Pz};
  for (int a = 0; a < 3; a++) {
    std::vector<double> c, sn;
    trig_table(P[a], c, sn);
    off[a] = trig.size();
    trig.insert(trig.end(), c.begin(), c.end());
    trig.insert(trig.end(), sn.begin(), sn.end());
  }
  c64 *w0 = nullptr, *w1 = nullptr;
  double* d_trig = nullptr;
  bool ok = hipMalloc(&w0, big * sizeof(c64)) == hipSuccess && hipMalloc(&w1, big * sizeof(c64)) == hipSuccess &&
            hipMalloc(&d_trig, trig.size() * sizeof(double)) == hipSuccess &&
            hipMemcpyAsync(d_trig, trig.data(), trig.size() * sizeof(double), hipMemcpyHostToDevice, s) == hipSuccess;
  ok = ok && ground_table(nx, ny, nz, h, z0, k, eps, order, w0, s);
  if (ok) {
    // z: full DFT of the layer sums s = 0 .. 2 nz - 2 (no parity)
    hipLaunchKernelGGL(k_dft_last, dim3((unsigned)((s1 + 255) / 256)), dim3(256), 0, s, w0, 6 * (size_t)nx * ny, L, Pz,
                       d_trig + off[2], d_trig + off[2] + Pz, w1);
    // x, y: even / odd lateral transforms (xy odd in x and y, xz odd in x, yz odd in y)
    const int odd_x = (1 << 3) | (1 << 4), odd_y = (1 << 3) | (1 << 5);
    hipLaunchKernelGGL(k_axis_transform, dim3((unsigned)((s2 + 255) / 256)), dim3(256), 0, s, w1, w0, nx, ny, Pz, 0, Px,
                       odd_x, d_trig + off[0], d_trig + off[0] + Px);
    hipLaunchKernelGGL(k_axis_transform, dim3((unsigned)((s3 + 255) / 256)), dim3(256), 0, s, w0, w1, Hx, ny, Pz, 1, Py,
                       odd_y, d_trig + off[1], d_trig + off[1] + Py);
    const double scale = 1.0 / ((double)Px * Py * Pz);
    hipLaunchKernelGGL(k_finish_spectrum, dim3((unsigned)((s3 + 255) / 256)), dim3(256), 0, s, w1, s3, scale, rspec64,
                       rspec32);
    ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(s) == hipSuccess;
  }
  (void)hipFree(w0);
  (void)hipFree(w1);
  (void)hipFree(d_trig);
  return ok ? 1 : 0;
}

}  // namespace vie
