// Mesh relationship checks of a dielectric model on the GPU: the pairwise
// box-interior overlap test and the dielectric-to-conductor clearance of the
// reference's mesher (dielectric_mesh.c:75-123 validate_box_relationships,
// dielectric_mesh.c:125-316 evaluate_clearance).  Both are all-pairs scans
// whose answer is the first pair in the reference's loop order, so each
// kernel reduces an ordering key with the value it is chosen by and no
// result depends on the launch shape.
//
// Arithmetic is FP64 with contraction off, the same operations in the same
// order as the reference, so distances, gaps and margins are bit-identical.
#include "vie_kernels.h"

#include <cfloat>

#pragma clang fp contract(off)

namespace vie {

namespace {

constexpr int kOverlapTile = 256;  // boxes per side of a pair tile (threads per block)
constexpr int kClearWG = 256;
constexpr double kMeshTolerance = 1.0e-12;  // dielectric_mesh.c:8
constexpr double kNecPi = 3.141592654;      // NEC's PI (nec2c.h)

// Interiors meet on one axis (dielectric_mesh.c:75-79).
__device__ __forceinline__ bool interiors_meet(double amin, double amax, double bmin, double bmax) {
  return fmax(amin, bmin) < fmin(amax, bmax);
}

// Upper-triangle tile (ti <= tj) of linear block index b, rows of T - ti tiles.
__device__ __forceinline__ void triangle_tile(unsigned long long b, unsigned long long T, unsigned long long& ti,
                                              unsigned long long& tj) {
  const double t2 = 2.0 * (double)T + 1.0;
  long long r = (long long)floor((t2 - sqrt(t2 * t2 - 8.0 * (double)b)) * 0.5);
  if (r < 0) r = 0;
  auto start = [&](unsigned long long row) { return row * T - row * (row - 1) / 2; };
  while (r > 0 && start((unsigned long long)r) > b) r--;
  while ((unsigned long long)r + 1 < T && start((unsigned long long)r + 1) <= b) r++;
  ti = (unsigned long long)r;
  tj = ti + (b - start(ti));
}

// first[0] <- min over overlapping pairs (i < j) of i * count + j.  Block b
// covers the pair tile (ti, tj); thread t owns box i = ti * TILE + t and scans
// the tile's boxes j in order, so its first hit is its smallest j.
__global__ __launch_bounds__(kOverlapTile) void k_box_overlap(const double* __restrict__ mn,
                                                              const double* __restrict__ mx, unsigned long long count,
                                                              unsigned long long tiles, unsigned long long block0,
                                                              unsigned long long* first) {
  __shared__ double bj[6][kOverlapTile];
  unsigned long long ti, tj;
  triangle_tile(block0 + blockIdx.x, tiles, ti, tj);
  const int t = threadIdx.x;
  const unsigned long long j0 = tj * kOverlapTile, i = ti * kOverlapTile + t;
  if (j0 + t < count) {
    const unsigned long long j = j0 + t;
    for (int a = 0; a < 3; a++) {
      bj[a][t] = mn[3 * j + a];
      bj[3 + a][t] = mx[3 * j + a];
    }
  }
  __syncthreads();
  if (i >= count) return;
  const double a0 = mn[3 * i], a1 = mn[3 * i + 1], a2 = mn[3 * i + 2];
  const double b0 = mx[3 * i], b1 = mx[3 * i + 1], b2 = mx[3 * i + 2];
  const unsigned long long jend = j0 + kOverlapTile < count ? j0 + kOverlapTile : count;
  unsigned long long q0 = i + 1 > j0 ? i + 1 - j0 : 0;
  for (unsigned long long q = q0; j0 + q < jend; q++) {
    const bool hit = interiors_meet(a0, b0, bj[0][q], bj[3][q]) & interiors_meet(a1, b1, bj[1][q], bj[4][q]) &
                     interiors_meet(a2, b2, bj[2][q], bj[5][q]);
    if (hit) {
      atomicMin(first, i * count + j0 + q);
      break;
    }
  }
}

// ---------------------------------------------------------------- clearance
// Squared distance from the segment start..end to the box (dielectric_mesh.c:
// 141-213): the segment is cut where it crosses the box's slab planes; on each
// piece every axis is inside its slab or on one fixed side, so the squared
// distance is a quadratic in the parameter, minimised on the piece.
__device__ double segment_box_d2(const double s[3], const double e[3], const double c[3], const double h[3]) {
  double br[8];
  int nb = 2;
  br[0] = 0.0;
  br[1] = 1.0;
  for (int a = 0; a < 3; a++) {
    const double d = e[a] - s[a];
    const double lo = c[a] - 0.5 * h[a], hi = c[a] + 0.5 * h[a];
    if (d != 0.0) {
      const double p = (lo - s[a]) / d, q = (hi - s[a]) / d;
      if (p > 0.0 && p < 1.0) br[nb++] = p;
      if (q > 0.0 && q < 1.0) br[nb++] = q;
    }
  }
  for (int k = 1; k < nb; k++) {  // insertion sort of the cut parameters
    const double v = br[k];
    int m = k;
    while (m > 0 && br[m - 1] > v) {
      br[m] = br[m - 1];
      m--;
    }
    br[m] = v;
  }
  double best = DBL_MAX;
  for (int k = 0; k + 1 < nb; k++) {
    const double lower = br[k], upper = br[k + 1], mid = 0.5 * (lower + upper);
    double qa = 0.0, qb = 0.0, qc = 0.0;
    for (int a = 0; a < 3; a++) {
      const double d = e[a] - s[a];
      const double at = s[a] + d * mid;
      const double lo = c[a] - 0.5 * h[a], hi = c[a] + 0.5 * h[a];
      if (at >= lo && at <= hi) continue;
      const double bound = at < lo ? lo : hi;
      qa += d * d;
      qb += 2.0 * d * (s[a] - bound);
      qc += (s[a] - bound) * (s[a] - bound);
    }
    double u = qa > 0.0 ? -qb / (2.0 * qa) : lower;
    if (u < lower) u = lower;
    if (u > upper) u = upper;
    const double v = qa * u * u + qb * u + qc;
    if (v < best) best = v;
  }
  return best < 0.0 && best > -kMeshTolerance ? 0.0 : best;
}

// Squared distance from a point to the box (dielectric_mesh.c:125-139).
__device__ double point_box_d2(const double p[3], const double c[3], const double h[3]) {
  double d2 = 0.0;
  for (int a = 0; a < 3; a++) {
    const double lo = c[a] - 0.5 * h[a], hi = c[a] + 0.5 * h[a];
    const double d = p[a] < lo ? lo - p[a] : p[a] > hi ? p[a] - hi : 0.0;
    d2 += d * d;
  }
  return d2;
}

// gap and required clearance of pair L = box * (n + m) + k, k < n a wire
// segment, else patch k - n (dielectric_mesh.c:250-303).
__device__ void clearance_pair(const ClearanceArgs& A, unsigned long long L, double* gap_out, double* req_out) {
  const unsigned long long per = (unsigned long long)(A.n + A.m);
  const size_t b = (size_t)(L / per);
  const long long k = (long long)(L - (unsigned long long)b * per);
  double c[3], h[3];
  for (int a = 0; a < 3; a++) {
    c[a] = 0.5 * (A.minimum[3 * b + a] + A.maximum[3 * b + a]);
    h[a] = A.maximum[3 * b + a] - A.minimum[3 * b + a];
  }
  double gap, req;
  if (k < A.n) {
    const double si = A.si[k];
    const double s[3] = {A.x[k] - 0.5 * si * A.cab[k], A.y[k] - 0.5 * si * A.sab[k], A.z[k] - 0.5 * si * A.salp[k]};
    const double e[3] = {A.x[k] + 0.5 * si * A.cab[k], A.y[k] + 0.5 * si * A.sab[k], A.z[k] + 0.5 * si * A.salp[k]};
    gap = sqrt(segment_box_d2(s, e, c, h)) - A.bi[k];
    if (gap < 0.0) gap = 0.0;
    req = fmax(A.cell_longest[b], si);
  } else {
    const long long p = k - A.n;
    const double pt[3] = {A.px[p], A.py[p], A.pz[p]};
    const double radius = sqrt(A.pbi[p] / kNecPi);
    gap = sqrt(point_box_d2(pt, c, h)) - radius;
    req = fmax(A.cell_longest[b], sqrt(A.pbi[p]));
    if (gap < 0.0) gap = 0.0;
  }
  *gap_out = gap;
  *req_out = req;
}

struct Best {
  double v;
  unsigned long long key;
};
// (v, key) ordering of the reference's strict-less scans: a smaller value
// wins, an equal one keeps the earlier pair; values that are not below the
// starting DBL_MAX (and NaN) never win.
__device__ __forceinline__ Best better(Best a, Best b) {
  const bool bv = b.v < DBL_MAX, av = a.v < DBL_MAX;
  if (!bv) return a;
  if (!av) return b;
  if (b.v < a.v || (b.v == a.v && b.key < a.key)) return b;
  return a;
}
__device__ __forceinline__ Best shfl_best(Best x, int m) {
  Best y;
  y.v = __shfl_xor(x.v, m, 64);
  y.key = (unsigned long long)__shfl_xor((long long)x.key, m, 64);
  return y;
}
__device__ Best block_best(Best x, Best* lds) {
  for (int m = 32; m >= 1; m >>= 1) x = better(x, shfl_best(x, m));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) lds[wave] = x;
  __syncthreads();
  Best r = lds[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = better(r, lds[w]);
  __syncthreads();
  return r;
}

// Per block: the worst margin (gap - required) and the smallest gap over its
// pairs, with their pair keys.
__global__ __launch_bounds__(kClearWG) void k_clearance(ClearanceArgs A, unsigned long long pairs, Best* part) {
  __shared__ Best lds[kClearWG / 64];
  Best worst = {DBL_MAX, ~0ull}, near = {DBL_MAX, ~0ull};
  const unsigned long long per = (unsigned long long)(A.n + A.m);
  for (unsigned long long L = (unsigned long long)blockIdx.x * kClearWG + threadIdx.x; L < pairs;
       L += (unsigned long long)gridDim.x * kClearWG) {
    const size_t b = (size_t)(L / per);
    if (!(A.cell_longest[b] > 0.0)) continue;  // boxes without cells are skipped
    double gap, req;
    clearance_pair(A, L, &gap, &req);
    worst = better(worst, Best{gap - req, L});
    near = better(near, Best{gap, L});
  }
  worst = block_best(worst, lds);
  near = block_best(near, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = worst;
    part[2 * blockIdx.x + 1] = near;
  }
}

// One block: reduce the partials, then evaluate the two chosen pairs again
// for the values the reference reports with them.
__global__ __launch_bounds__(kClearWG) void k_clearance_finish(ClearanceArgs A, const Best* part, unsigned nblk,
                                                               ClearanceResult* out) {
  __shared__ Best lds[kClearWG / 64];
  Best worst = {DBL_MAX, ~0ull}, near = {DBL_MAX, ~0ull};
  for (unsigned k = threadIdx.x; k < nblk; k += kClearWG) {
    worst = better(worst, part[2 * k]);
    near = better(near, part[2 * k + 1]);
  }
  worst = block_best(worst, lds);
  near = block_best(near, lds);
  if (threadIdx.x != 0) return;
  ClearanceResult r;
  r.worst_margin = DBL_MAX;
  r.worst_gap = 0.0;
  r.worst_required = 0.0;
  r.worst_box = 0;
  r.worst_conductor = 0;
  r.worst_is_patch = 0;
  r.minimum_gap = DBL_MAX;
  r.required_gap = 0.0;
  const unsigned long long per = (unsigned long long)(A.n + A.m);
  if (worst.v < DBL_MAX) {
    double gap, req;
    clearance_pair(A, worst.key, &gap, &req);
    const long long k = (long long)(worst.key % per);
    r.worst_margin = gap - req;
    r.worst_gap = gap;
    r.worst_required = req;
    r.worst_box = (long long)(worst.key / per) + 1;
    r.worst_is_patch = k >= A.n;
    r.worst_conductor = (k >= A.n ? k - A.n : k) + 1;
  }
  if (near.v < DBL_MAX) {
    double gap, req;
    clearance_pair(A, near.key, &gap, &req);
    r.minimum_gap = gap;
    r.required_gap = req;
  }
  *out = r;
}

}  // namespace

int box_overlap(const double* mn, const double* mx, size_t count, unsigned long long* first, hipStream_t s) {
  unsigned long long init = ~0ull;
  if (hipMemcpyAsync(first, &init, sizeof(init), hipMemcpyHostToDevice, s) != hipSuccess) return 0;
  if (count < 2) return 1;
  const unsigned long long T = (count + kOverlapTile - 1) / kOverlapTile;
  const unsigned long long blocks = T * (T + 1) / 2;
  const unsigned long long chunk = 1ull << 30;  // blocks per launch
  for (unsigned long long b0 = 0; b0 < blocks; b0 += chunk) {
    const unsigned long long nb = blocks - b0 < chunk ? blocks - b0 : chunk;
    hipLaunchKernelGGL(k_box_overlap, dim3((unsigned)nb), dim3(kOverlapTile), 0, s, mn, mx,
                       (unsigned long long)count, T, b0, first);
  }
  return hipGetLastError() == hipSuccess;
}

size_t clearance_partials(size_t pairs) {
  size_t b = (pairs + kClearWG - 1) / kClearWG;
  return b < 1 ? 1 : (b > 4096 ? 4096 : b);
}

int conductor_clearance(const ClearanceArgs& A, void* partials, ClearanceResult* out, hipStream_t s) {
  const unsigned long long pairs = (unsigned long long)A.boxes * (unsigned long long)(A.n + A.m);
  const unsigned nb = (unsigned)clearance_partials((size_t)pairs);
  Best* part = static_cast<Best*>(partials);
  hipLaunchKernelGGL(k_clearance, dim3(nb), dim3(kClearWG), 0, s, A, pairs, part);
  hipLaunchKernelGGL(k_clearance_finish, dim3(1), dim3(kClearWG), 0, s, A, (const Best*)part, nb, out);
  return hipGetLastError() == hipSuccess;
}

}  // namespace vie
