// Device-resident vector kernels for the restarted GMRES and the conductor
// coupling blocks.
//
// Reductions are deterministic and atomic-free: every block writes its
// partial sums (key-major, partial[key * kMaxDotBlocks + block]) and a
// separate finishing launch adds them in block order, so a result never
// depends on scheduling or on what else runs on the device.  Dot products
// and norms accumulate in FP64 whatever the storage precision of the Krylov
// basis.
#include "krylov.h"
#include "vie_kernels.h"

namespace vie {

constexpr int KWG = 256;  // threads per block
constexpr int KEPT = 4;   // elements per thread (elementwise reductions)
constexpr int KCH = KWG * KEPT;
constexpr int kMaxDotBlocks = 1024;  // grid-stride cap: partials per key

size_t red_blocks(size_t n) { return (n + KCH - 1) / KCH; }

template <typename T> struct DotShape;
template <typename T> static unsigned dot_blocks(size_t n) {
  size_t ch = (size_t)64 * DotShape<T>::CPL;
  size_t b = (n + ch - 1) / ch;
  return (unsigned)(b < 1 ? 1 : (b > kMaxDotBlocks ? kMaxDotBlocks : b));
}
// Elementwise grid-stride kernels (no partials) take a larger grid: at 4096
// workgroups the delayed-CGS2 update of a 6.3M-element basis runs 2% faster
// than at 1024 (0.609 against 0.621 ms at column 66 in FP32).
constexpr unsigned kMaxUpdateBlocks = 4096;
static unsigned update_blocks(size_t n, int ept = 2) {
  size_t ch = (size_t)KWG * ept;
  size_t b = (n + ch - 1) / ch;
  return (unsigned)(b < 1 ? 1 : (b > kMaxUpdateBlocks ? kMaxUpdateBlocks : b));
}

// Delayed-CGS2 streaming shapes: basis elements per lane per key in the
// reduction (64 bytes per lane either way), elements per thread in the fused
// update.
template <typename T> struct DcShape {
  static constexpr int CPL = sizeof(T) == 4 ? 8 : 4;  // 64 bytes per lane per key either way
  static constexpr int UE = 2;
};
// the padded basis covers whole reduction chunks and update groups
static_assert(kBasisAlign % (64 * DcShape<float>::CPL) == 0 && kBasisAlign % (64 * DcShape<double>::CPL) == 0,
              "basis padding must cover whole delayed-CGS2 reduction chunks");
static_assert(kBasisAlign % (KWG * DcShape<float>::UE) == 0 && kBasisAlign % (KWG * DcShape<double>::UE) == 0,
              "basis padding must cover whole delayed-CGS2 update groups");
template <typename T> static unsigned dc_blocks(size_t n) {
  size_t ch = (size_t)64 * DcShape<T>::CPL;
  size_t b = (n + ch - 1) / ch;
  return (unsigned)(b < 1 ? 1 : (b > kMaxDotBlocks ? kMaxDotBlocks : b));
}

__device__ inline c64 wave_reduce(c64 v) {
  for (int m = 32; m >= 1; m >>= 1) {
    v.re += __shfl_xor(v.re, m, 64);
    v.im += __shfl_xor(v.im, m, 64);
  }
  return v;
}


// Reduces per-wave values through LDS; returns the block total on thread 0.
// Every thread of the block calls it.
__device__ inline c64 block_reduce(c64 v, c64* lds) {
  v = wave_sum_last(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 63) lds[wave] = v;
  __syncthreads();
  c64 t = {0, 0};
  if (threadIdx.x == 0)
    for (int q = 0; q < KWG / 64; q++) t = t + lds[q];
  return t;
}

// ---------------------------------------------------------------- dots
// partial[blk * ldp + k] = block's share of sum_i op(V[k][i]) z[i], where
// op is conj (Krylov projections) or identity (the conductor rows B z), and
// z = w, or D^{-1} w (Jacobi on the fly) with JAC.  With self_norm, key nk
// receives ||w||^2.  The vector is cut into chunks of 64 * CPL elements;
// every wave of a block holds the chunk of z in registers and takes whole
// keys (k = wave, wave + 4, ...), so one shuffle reduction covers CPL loads
// per lane.  Grid-stride over chunks; LDS holds the block's per-key sums.
template <typename T> struct DotShape { static constexpr int CPL = 128 / (int)sizeof(cx<T>); };

template <typename T, bool CONJ, bool JAC>
__device__ __forceinline__ void k_dots_body(const cx<T>* V, size_t ldv, int nk, const cx<T>* w, size_t n,
                                              int self_norm, c64* partial, int ldp, JacobiLite<T> J,
                                              DevIndex di, unsigned bx, unsigned nbx) {
  constexpr int CPL = DotShape<T>::CPL;
  constexpr int WAVES = KWG / 64;
  constexpr int CH = 64 * CPL;
  __shared__ c64 lds[kMaxKeys];
  __shared__ c64 zs[CH];
  if (di.skipped()) return;
  w += di.off();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tot = nk + (self_norm ? 1 : 0);
  for (int k = threadIdx.x; k < tot; k += KWG) lds[k] = c64{0, 0};
  const size_t nchunks = (n + CH - 1) / CH;
  for (size_t ch = bx; ch < nchunks; ch += nbx) {
    const size_t base = ch * CH;
    cx<T> a[CPL], an[CPL];
    auto load_key = [&](int k, cx<T>* dst) {
      const cx<T>* vk = V + (size_t)k * ldv + base + lane;
#pragma unroll
      for (int e = 0; e < CPL; e++) dst[e] = base + lane + (size_t)e * 64 < n ? ld_stream(vk + e * 64) : cx<T>{0, 0};
    };
    // the wave's first key is in flight while the chunk of z is staged
    if (wave < nk) load_key(wave, a);
    __syncthreads();
    for (int e = threadIdx.x; e < CH; e += KWG) {
      size_t i = base + e;
      cx<T> v = i < n ? w[i] : cx<T>{0, 0};
      if (JAC && i < n) v = J.apply(i, v);
      zs[e] = c64{(double)v.re, (double)v.im};
    }
    __syncthreads();
    for (int k = wave; k < tot; k += WAVES) {
      c64 s = {0, 0};
      if (k < nk) {
        if (k + WAVES < nk) load_key(k + WAVES, an);
#pragma unroll
        for (int e = 0; e < CPL; e++) {
          const c64 z = zs[lane + e * 64];
          double ar = a[e].re, ai = CONJ ? -(double)a[e].im : (double)a[e].im;
          s.re += ar * z.re - ai * z.im;
          s.im += ar * z.im + ai * z.re;
        }
#pragma unroll
        for (int e = 0; e < CPL; e++) a[e] = an[e];
      } else {
#pragma unroll
        for (int e = 0; e < CPL; e++) {
          const c64 z = zs[lane + e * 64];
          s.re += z.re * z.re + z.im * z.im;
        }
      }
      s = wave_sum_last(s);
      if (lane == 63) lds[k] = lds[k] + s;  // keys are owned by one wave
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < tot; k += KWG) partial[(size_t)k * ldp + bx] = lds[k];
}
template <typename T, bool CONJ, bool JAC>
__global__ __launch_bounds__(KWG, 4) void k_dots(const cx<T>* V, size_t ldv, int nk, const cx<T>* w, size_t n,
                                              int self_norm, c64* partial, int ldp, JacobiLite<T> J,
                                              DevIndex di) {
  k_dots_body<T, CONJ, JAC>(V, ldv, nk, w, n, self_norm, partial, ldp, J, di, blockIdx.x, gridDim.x);
}

// out[k] = sum_b partial[k * ldp + b] (key-major, contiguous per key): one
// wave per key, four loads in flight per lane, fixed summation order.
__device__ __forceinline__ c64 sum_partials(const c64* p, unsigned nblk, int lane) {
  c64 s0 = {0, 0}, s1 = {0, 0}, s2 = {0, 0}, s3 = {0, 0};
  unsigned b = lane;
  for (; b + 192 < nblk; b += 256) {
    s0 = s0 + p[b];
    s1 = s1 + p[b + 64];
    s2 = s2 + p[b + 128];
    s3 = s3 + p[b + 192];
  }
  for (; b < nblk; b += 64) s0 = s0 + p[b];
  return wave_reduce((s0 + s1) + (s2 + s3));
}

__global__ __launch_bounds__(KWG) void k_finish(const c64* partial, int ldp, unsigned nblk, int nk, c64* out) {
  const int key = blockIdx.x * (KWG / 64) + (threadIdx.x >> 6);
  if (key >= nk) return;  // whole waves leave together
  const int lane = threadIdx.x & 63;
  const c64 s = sum_partials(partial + (size_t)key * ldp, nblk, lane);
  if (lane == 0) out[key] = s;
}

// k_finish for the delayed-CGS2 reduction, key count read from the state.
__device__ __forceinline__ void k_finish_dc_body(const c64* partial, int ldp, unsigned nblk,
                                                   const ArnoldiState* st, int has_w, c64* out, unsigned bx, unsigned nbx) {
  if (st->done) return;
  const int j = st->j, nk = has_w ? 2 * j + 3 : j + 1;
  const int key = bx * (KWG / 64) + (threadIdx.x >> 6);
  if (key >= nk) return;
  const int lane = threadIdx.x & 63;
  const c64 s = sum_partials(partial + (size_t)key * ldp, nblk, lane);
  if (lane == 0) out[key] = s;
}
__global__ __launch_bounds__(KWG) void k_finish_dc(const c64* partial, int ldp, unsigned nblk,
                                                   const ArnoldiState* st, int has_w, c64* out) {
  k_finish_dc_body(partial, ldp, nblk, st, has_w, out, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------- update
// w -= sum_k h[k] V[k] (h from device memory, FP64 accumulation), with the
// block's share of ||w||^2 in partial[blk * ldp] when want_norm.  Two
// elements per thread and four basis vectors per loop step keep eight loads
// in flight per lane.
constexpr int UEPT = 2;

template <typename T>
__global__ __launch_bounds__(KWG) void k_update(const cx<T>* V, size_t ldv, int nk, const c64* h, cx<T>* w,
                                                size_t n, int want_norm, c64* partial, int ldp) {
  __shared__ c64 hs[kMaxKeys];
  __shared__ c64 lds[KWG / 64];
  for (int k = threadIdx.x; k < nk; k += KWG) hs[k] = h[k];
  __syncthreads();
  c64 s = {0, 0};
  const size_t step = (size_t)gridDim.x * KWG * UEPT;
  for (size_t base = (size_t)blockIdx.x * KWG * UEPT + threadIdx.x; base < n; base += step) {
    double ar[UEPT], ai[UEPT];
    bool in[UEPT];
#pragma unroll
    for (int e = 0; e < UEPT; e++) {
      size_t i = base + (size_t)e * KWG;
      in[e] = i < n;
      cx<T> v = in[e] ? w[i] : cx<T>{0, 0};
      ar[e] = v.re;
      ai[e] = v.im;
    }
    int k = 0;
    for (; k + 4 <= nk; k += 4) {
      cx<T> a[4][UEPT];
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int e = 0; e < UEPT; e++) {
          size_t i = base + (size_t)e * KWG;
          a[q][e] = in[e] ? ld_stream(V + (size_t)(k + q) * ldv + i) : cx<T>{0, 0};
        }
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const double hr = hs[k + q].re, hi = hs[k + q].im;
#pragma unroll
        for (int e = 0; e < UEPT; e++) {
          ar[e] -= hr * a[q][e].re - hi * a[q][e].im;
          ai[e] -= hr * a[q][e].im + hi * a[q][e].re;
        }
      }
    }
    for (; k < nk; k++) {
      const double hr = hs[k].re, hi = hs[k].im;
#pragma unroll
      for (int e = 0; e < UEPT; e++) {
        size_t i = base + (size_t)e * KWG;
        if (in[e]) {
          cx<T> a = ld_stream(V + (size_t)k * ldv + i);
          ar[e] -= hr * a.re - hi * a.im;
          ai[e] -= hr * a.im + hi * a.re;
        }
      }
    }
#pragma unroll
    for (int e = 0; e < UEPT; e++) {
      size_t i = base + (size_t)e * KWG;
      if (in[e]) {
        cx<T> r = {(T)ar[e], (T)ai[e]};
        w[i] = r;
        double rr = r.re, ri = r.im;
        s.re += rr * rr + ri * ri;
      }
    }
  }
  if (!want_norm) return;
  s = block_reduce(s, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// ---------------------------------------------------------------- Arnoldi tail
// Finishes column j: H(:,j) = h1 + h2, H(j+1,j) = ||w||, applies the previous
// Givens rotations, forms rotation j, updates g (vie_iterative.c:160-177).
__device__ inline void givens(c64 a, c64 b, double* c, c64* s) {
  double aa = hypot(a.re, a.im), bb = hypot(b.re, b.im);
  if (bb == 0.0) {
    *c = 1.0;
    *s = c64{0, 0};
  } else if (aa == 0.0) {
    *c = 0.0;
    *s = c64{b.re / bb, -b.im / bb};
  } else {
    double len = hypot(aa, bb);
    *c = aa / len;
    c64 u = {a.re / aa, a.im / aa}, v = {b.re / len, -b.im / len};
    *s = u * v;
  }
}

__global__ __launch_bounds__(KWG) void k_arnoldi_tail(ArnoldiState* st, const c64* h1, const c64* h2,
                                                      const c64* before2, const c64* norm_partial, int ldp,
                                                      unsigned nblk, double breakdown_eps) {
  __shared__ c64 col[kMaxRestart + 2];
  __shared__ c64 sn[kMaxRestart];
  __shared__ double cs[kMaxRestart];
  __shared__ c64 lds[KWG / 64];
  if (st->done) return;  // a step issued after the cycle ended
  const int j = st->j, m = st->m;
  for (int k = threadIdx.x; k <= j; k += KWG) col[k] = h1[k] + (h2 ? h2[k] : c64{0, 0});
  for (int k = threadIdx.x; k < j; k += KWG) {
    sn[k] = st->sn[k];
    cs[k] = st->cs[k];
  }
  c64 s = {0, 0};
  for (unsigned b = threadIdx.x; b < nblk; b += KWG) s = s + norm_partial[b];
  s = block_reduce(s, lds);  // contains the barrier that publishes col / sn / cs
  if (threadIdx.x != 0) return;
  c64* H = st->H;  // row-major (m+1) x m
  double before = sqrt(before2->re), next = sqrt(s.re);
  st->nonfinite = !(isfinite(before) && isfinite(next));
  st->breakdown = next <= breakdown_eps * before;
  st->inv_next = st->breakdown || next == 0 ? 0.0 : 1.0 / next;
  col[j + 1] = c64{next, 0};
  for (int k = 0; k < j; k++) {
    c64 f = col[k], g = col[k + 1];
    col[k] = cs[k] * f + sn[k] * g;
    col[k + 1] = cs[k] * g - conj(sn[k]) * f;
  }
  double c;
  c64 z;
  givens(col[j], col[j + 1], &c, &z);
  st->cs[j] = c;
  st->sn[j] = z;
  col[j] = c * col[j] + z * col[j + 1];
  col[j + 1] = c64{0, 0};
  for (int k = 0; k <= j + 1; k++) H[k * m + j] = col[k];
  c64 gj = st->g[j];
  st->g[j + 1] = c64{0, 0} - conj(z) * gj;
  st->g[j] = c * gj;
  st->gabs = hypot(st->g[j + 1].re, st->g[j + 1].im);
  st->used = j + 1;
  st->j = j + 1;
  st->done = st->breakdown || st->nonfinite || st->gabs <= st->inner_target || j + 1 >= m;
}

// y = H(base:used,base:used)^{-1} g by back substitution
// (vie_iterative.c:179-186): one row dot product per step, spread over the
// block.  A recycled cycle also forms the U coefficients of its update,
// uc = beta zc - B y, with B = H(0:base, base:used) (unrotated rows).
__global__ __launch_bounds__(KWG) void k_backsolve(ArnoldiState* st) {
  __shared__ c64 y[kMaxRestart];
  __shared__ c64 lds[KWG / 64];
  __shared__ int bad;
  const int used = st->used, m = st->m, base = st->base;
  const c64* H = st->H;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  for (int i = used - 1; i >= base; i--) {
    c64 part = {0, 0};
    for (int k = i + 1 + threadIdx.x; k < used; k += KWG) part = part + H[i * m + k] * y[k];
    part = block_reduce(part, lds);
    if (threadIdx.x == 0) {
      c64 d = H[i * m + i];
      if (d.re == 0.0 && d.im == 0.0) bad = 1;
      else y[i] = cdiv(st->g[i] - part, d);
    }
    __syncthreads();
    if (bad) break;
  }
  __syncthreads();
  if (threadIdx.x == 0) st->singular = bad;
  if (bad) return;
  for (int k = base + threadIdx.x; k < used; k += KWG) st->y[k] = y[k];
  const double beta = st->beta;
  for (int l = threadIdx.x; l < base; l += KWG) {
    c64 s = beta * st->zc[l];
    for (int c = base; c < used; c++) s = s - H[l * m + c] * y[c];
    st->uc[l] = s;
  }
}

// Starts a cycle at column base: g = beta e_base, H = 0, j = base, and the
// columns of Hraw that stand for the recycled images zeroed (they have no
// Arnoldi relation of their own).
__global__ void k_cycle_init(ArnoldiState* st, const c64* beta2, double inner_target, int base) {
  const int m = st->m;
  for (int i = threadIdx.x; i < (m + 1) * m; i += blockDim.x) st->H[i] = c64{0, 0};
  for (int i = threadIdx.x; i <= m; i += blockDim.x) st->g[i] = c64{0, 0};
  for (int i = threadIdx.x; i < (m + 1) * base; i += blockDim.x) st->Hraw[i] = c64{0, 0};
  __syncthreads();
  if (threadIdx.x == 0) {
    double beta = sqrt(beta2->re);
    st->g[base] = c64{beta, 0};
    st->beta = beta;
    st->inv_beta = beta > 0 ? 1.0 / beta : 0.0;
    st->j = base;
    st->base = base;
    st->step = 0;
    st->used = base;
    st->breakdown = 0;
    st->nonfinite = !isfinite(beta);
    st->done = 0;
    st->inner_target = inner_target;
  }
}

// ---------------------------------------------------------------- delayed CGS2
// Keys of the single reduction of step j:
//   a_k = q_k^H u, b_k = q_k^H w (k < j), nu = u^H u, c = u^H w, d = w^H w.
// Layout: [a_0..a_{j-1}, b_0..b_{j-1}, nu, c, d]; without w only [a, nu].
template <typename T>
__device__ __forceinline__ void k_dots_dc_body(const cx<T>* Q, size_t ldv, const ArnoldiState* st, int has_w,
                                                 size_t n, c64* partial, int ldp, unsigned bx, unsigned nbx) {
  constexpr int CPL = DcShape<T>::CPL;  // u, w and q in flight per lane
  constexpr int WAVES = KWG / 64;
  constexpr int CH = 64 * CPL;
  __shared__ c64 lds[kMaxKeys];
  __shared__ c64 us[CH], ws[CH];
  if (st->done) return;
  const int j = st->j;
  const cx<T>* u = Q + (size_t)j * ldv;
  const cx<T>* w = has_w ? Q + (size_t)(j + 1) * ldv : nullptr;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool hw = has_w != 0;
  const int tot = hw ? 2 * j + 3 : j + 1;
  for (int k = threadIdx.x; k < tot; k += KWG) lds[k] = c64{0, 0};
  // whole chunks: the basis padding past n is zero (krylov.h, kBasisAlign)
  const size_t nchunks = (n + CH - 1) / CH;
  for (size_t ch = bx; ch < nchunks; ch += nbx) {
    const size_t base = ch * CH;
    __syncthreads();
    for (int e = threadIdx.x; e < CH; e += KWG) {
      size_t i = base + e;
      cx<T> a = u[i], b = hw ? w[i] : cx<T>{0, 0};
      us[e] = c64{(double)a.re, (double)a.im};
      ws[e] = c64{(double)b.re, (double)b.im};
    }
    __syncthreads();
    // basis vectors: one load feeds both projections; the next key's load is
    // issued before this key's reduction (software pipelining)
    cx<T> q[CPL], qn[CPL];
#define VIE_DC_LOAD(k, dst)                                                                                \
  do {                                                                                                     \
    const cx<T>* qk_ = Q + (size_t)(k) * ldv + base + lane;                                                \
    _Pragma("unroll") for (int e_ = 0; e_ < CPL; e_++) dst[e_] = ld_stream(qk_ + e_ * 64);                 \
  } while (0)
    // one key's two projections, reduced over the wave and added to the block's sums
#define VIE_DC_KEY(qb, k)                                                                                  \
  do {                                                                                                     \
    c64 sa = {0, 0}, sb = {0, 0};                                                                          \
    _Pragma("unroll") for (int e = 0; e < CPL; e++) {                                                      \
      const c64 uu = us[lane + e * 64];                                                                    \
      const double qr = qb[e].re, qi = qb[e].im;                                                           \
      sa.re += qr * uu.re + qi * uu.im;                                                                    \
      sa.im += qr * uu.im - qi * uu.re;                                                                    \
      if (hw) {                                                                                            \
        const c64 ww = ws[lane + e * 64];                                                                  \
        sb.re += qr * ww.re + qi * ww.im;                                                                  \
        sb.im += qr * ww.im - qi * ww.re;                                                                  \
      }                                                                                                    \
    }                                                                                                      \
    sa = wave_sum_last(sa);                                                                                \
    if (hw) sb = wave_sum_last(sb);                                                                        \
    if (lane == 63) {                                                                                      \
      lds[k] = lds[k] + sa;                                                                                \
      if (hw) lds[j + k] = lds[j + k] + sb;                                                                \
    }                                                                                                      \
  } while (0)
    // Two register sets in turn: the next key's loads are issued before this
    // key's arithmetic and stay in flight through it (the loop body holds no
    // conditional load and no register copy, so the wait before the
    // arithmetic covers only the older set).
    int k = wave;
    if (k < j) VIE_DC_LOAD(k, q);
    for (; k + 2 * WAVES < j; k += 2 * WAVES) {
      VIE_DC_LOAD(k + WAVES, qn);
      VIE_DC_KEY(q, k);
      VIE_DC_LOAD(k + 2 * WAVES, q);
      VIE_DC_KEY(qn, k + WAVES);
    }
    if (k + WAVES < j) {
      VIE_DC_LOAD(k + WAVES, qn);
      VIE_DC_KEY(q, k);
      VIE_DC_KEY(qn, k + WAVES);
    } else if (k < j) {
      VIE_DC_KEY(q, k);
    }
#undef VIE_DC_KEY
#undef VIE_DC_LOAD
    // the scalars nu, c, d: one wave each
    const int nscal = hw ? 3 : 1;
    for (int q = wave; q < nscal; q += WAVES) {
      c64 s = {0, 0};
#pragma unroll
      for (int e = 0; e < CPL; e++) {
        const c64 uu = us[lane + e * 64], ww = ws[lane + e * 64];
        if (q == 0) s.re += uu.re * uu.re + uu.im * uu.im;
        else if (q == 1) {
          s.re += uu.re * ww.re + uu.im * ww.im;
          s.im += uu.re * ww.im - uu.im * ww.re;
        } else s.re += ww.re * ww.re + ww.im * ww.im;
      }
      s = wave_sum_last(s);
      int key = (hw ? 2 * j : j) + q;
      if (lane == 63) lds[key] = lds[key] + s;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < tot; k += KWG) partial[(size_t)k * ldp + bx] = lds[k];
}
template <typename T>
__global__ __launch_bounds__(KWG, 4) void k_dots_dc(const cx<T>* Q, size_t ldv, const ArnoldiState* st, int has_w,
                                                 size_t n, c64* partial, int ldp) {
  k_dots_dc_body<T>(Q, ldv, st, has_w, n, partial, ldp, blockIdx.x, gridDim.x);
}

// End of an Arnoldi step (thread 0): status block to the host ring slot of
// this step, then j advances unless the cycle is done (later steps are then
// no-ops); ju keeps the column the step's update finishes.
__device__ __forceinline__ void end_step(ArnoldiState* st, StepStatus* ring, int j) {
  StepStatus* r = ring + (st->step & 1);
  r->gabs = st->gabs;
  r->beta = st->beta;
  r->breakdown = st->breakdown;
  r->nonfinite = st->nonfinite;
  r->singular = st->singular;
  r->used = st->used;
  r->done = st->done;
  r->pad_ = st->step;
  __threadfence_system();
  st->step = st->step + 1;
  st->ju = j;
  if (!st->done) st->j = j + 1;
}

// Small solve of the delayed step (one block).  With u = alpha q_j + Q a:
//  1. alpha^2 = nu - |a|^2 (Q orthonormal);
//  2. column j-1 becomes final: Hraw(0:j, j-1) += a, Hraw(j, j-1) = alpha;
//     it is rotated into H, the residual estimate |g_j| formed and the cycle
//     marked done on convergence or breakdown (vie_iterative.c:160-177);
//  3. A M^{-1} q_j = (w - Q_{j+1} t)/alpha with t = Hraw(:, 0:j-1) a, so its
//     projections are h = [(b - t)/alpha ; ((c - a^H b)/alpha - t_j)/alpha];
//  4. next candidate u' = (w - Q_j e - f u)/alpha with s = t + alpha h,
//     e = s(0:j-1) - (s_j/alpha) a, f = s_j/alpha; q_j = (u - Q_j a)/alpha.
__device__ __forceinline__ void k_dc_mid_body(ArnoldiState* st, int has_w, const c64* dots,
                                                double breakdown_eps, StepStatus* ring, unsigned bx, unsigned nbx) {
  __shared__ c64 a[kMaxRestart], b[kMaxRestart], t[kMaxRestart + 1], col[kMaxRestart + 1];
  __shared__ c64 tpart[2][kMaxRestart + 1];
  __shared__ c64 sn[kMaxRestart];
  __shared__ double cs[kMaxRestart];
  __shared__ c64 red[KWG / 64];
  __shared__ c64 sc[4];        // nu, c, d, g_{j-1}
  __shared__ double sd[2];     // dc_before, inner_target
  __shared__ double alpha_s;
  __shared__ int stop;
  if (st->done) {  // a step queued after the cycle ended: status only
    if (threadIdx.x == 0) end_step(st, ring, st->j);
    return;
  }
  const int j = st->j, base = st->base;
  const int m = st->m, ld = m + 1;  // Hraw is column-major: Hraw[l * (m + 1) + k] = row k, column l
  const int tid = threadIdx.x;
  c64* Hraw = st->Hraw;
  // every operand in one round of independent loads: j <= kMaxRestart <= KWG,
  // so each thread takes one entry (threads past j read entry 0 and keep
  // nothing); no load sits under a lane-divergent branch
  {
    static_assert(kMaxRestart <= KWG, "one Arnoldi column entry per thread");
    const int k = tid < j ? tid : 0;
    const c64 va = dots[k];
    const c64 vb = has_w ? dots[j + k] : c64{0, 0};
    const c64 vsn = st->sn[k];
    const double vcs = st->cs[k];
    const c64 vcol = Hraw[(size_t)(j > base ? j - 1 : 0) * ld + k];  // old column j-1, rows 0..j-1
    const c64 s0 = dots[has_w ? 2 * j : j];
    const c64 s1 = has_w ? dots[2 * j + 1] : c64{0, 0};
    const c64 s2 = has_w ? dots[2 * j + 2] : c64{0, 0};
    const c64 s3 = st->g[j > base ? j - 1 : base];
    const double d0 = st->dc_before, d1 = st->inner_target;
    if (tid < j) {
      a[tid] = va;
      b[tid] = vb;
      sn[tid] = vsn;
      cs[tid] = vcs;
      col[tid] = vcol;
    }
    if (tid == 0) {
      sc[0] = s0;
      sc[1] = s1;
      sc[2] = s2;
      sc[3] = s3;
      sd[0] = d0;
      sd[1] = d1;
    }
  }
  __syncthreads();
  // |a|^2 and a^H b in parallel
  c64 p2 = {0, 0}, pab = {0, 0};
  for (int k = tid; k < j; k += KWG) {
    p2.re += a[k].re * a[k].re + a[k].im * a[k].im;
    if (has_w) pab = pab + conj(a[k]) * b[k];
  }
  p2 = block_reduce(p2, red);
  if (tid == 0) {
    double al2 = sc[0].re - p2.re;
    alpha_s = al2 > 0 ? sqrt(al2) : 0.0;
  }
  __syncthreads();
  pab = block_reduce(pab, red);  // thread 0 holds a^H b
  const double alpha = alpha_s;
  // column j-1 becomes final: Hraw(0:j, j-1) += a, Hraw(j, j-1) = alpha
  if (j > base) {
    for (int k = tid; k <= j; k += KWG) {
      c64 v = k < j ? col[k] + a[k] : c64{alpha, 0};
      Hraw[(size_t)(j - 1) * ld + k] = v;
      col[k] = v;
    }
  }
  __syncthreads();
  if (tid == 0) {
    stop = 0;
    if (!isfinite(alpha)) {
      st->nonfinite = 1;
      stop = 1;
    } else if (j == base) {
      // q_base = (u - Q a) / alpha: the least-squares right-hand side is
      // beta * alpha.  In a recycled cycle u was projected on C once and a
      // holds the second projection; an alpha at rounding level means the
      // residual lies in span(C), which the U update alone then removes.
      st->g[base] = alpha * sc[3];
      if (base > 0 && alpha <= breakdown_eps) {
        st->breakdown = 1;
        st->used = base;
        stop = 1;
      }
    } else {
      const int c = j - 1;
      const int brk = alpha <= breakdown_eps * sd[0];
      st->breakdown = brk;
      // earlier rotations in turn; the rotated entry k + 1 is carried in a
      // register, so the chain waits on no LDS round trip
      c64 f = col[base];
      for (int k = base; k < c; k++) {
        const c64 gg = col[k + 1];
        col[k] = cs[k] * f + sn[k] * gg;
        f = cs[k] * gg - conj(sn[k]) * f;
      }
      col[c] = f;
      double cr;
      c64 z;
      givens(col[c], col[c + 1], &cr, &z);
      st->cs[c] = cr;
      st->sn[c] = z;
      col[c] = cr * col[c] + z * col[c + 1];
      col[c + 1] = c64{0, 0};
      const c64 gc = sc[3];
      const c64 gn = c64{0, 0} - conj(z) * gc;
      st->g[c + 1] = gn;
      st->g[c] = cr * gc;
      const double gabs = hypot(gn.re, gn.im);
      st->gabs = gabs;
      st->used = j;
      st->j = j;
      if (brk || gabs <= sd[1] || j >= m) stop = 1;
    }
    if (!has_w) stop = 1;
    if (stop) st->done = 1;
  }
  __syncthreads();
  if (j > base)
    for (int k = tid; k <= j; k += KWG) st->H[k * m + j - 1] = col[k];
  if (j == base)
    for (int k = tid; k < base; k += KWG) st->zc[k] = st->zc[k] + a[k];
  if (stop) {
    // the last basis vector is left as u = alpha q_j + Q_j a: keep a and alpha
    // so the host can still form q_j (recycled-subspace construction)
    for (int k = tid; k < j; k += KWG) st->dc_a[k] = a[k];
    if (tid == 0) {
      st->dc_alpha = alpha;
      end_step(st, ring, j);
    }
    return;
  }
  // t = Hraw(0:j, 0:j-1) a (Hessenberg: row k only has columns >= k-1).
  // Threads own rows (consecutive rows are contiguous in a column, so each
  // column read is coalesced); the two halves of the block split the columns.
  {
    const int half = tid >> 7, r0 = tid & 127;
    const int mid = j / 2;
    for (int k = r0; k <= j; k += 128) {
      const int lo = half == 0 ? (k > 0 ? k - 1 : 0) : (mid > k - 1 ? mid : (k > 0 ? k - 1 : 0));
      const int hi = half == 0 ? (mid < j ? mid : j) : j;
      // eight columns per iteration: eight independent loads in flight
      c64 sp[8];
#pragma unroll
      for (int q = 0; q < 8; q++) sp[q] = c64{0, 0};
      int l = lo;
      for (; l + 7 < hi; l += 8) {
        c64 h[8];
#pragma unroll
        for (int q = 0; q < 8; q++) h[q] = Hraw[(size_t)(l + q) * ld + k];
#pragma unroll
        for (int q = 0; q < 8; q++) sp[q] = sp[q] + h[q] * a[l + q];
      }
      for (; l < hi; l++) sp[0] = sp[0] + Hraw[(size_t)l * ld + k] * a[l];
      tpart[half][k] = ((sp[0] + sp[1]) + (sp[2] + sp[3])) + ((sp[4] + sp[5]) + (sp[6] + sp[7]));
    }
  }
  __syncthreads();
  for (int k = tid; k <= j; k += KWG) t[k] = tpart[0][k] + tpart[1][k];
  __syncthreads();
  const double ia = 1.0 / alpha;
  __shared__ c64 f_s;
  if (tid == 0) {
    const c64 cc = sc[1], dd = sc[2];
    c64 hj = ia * (ia * (cc - pab) - t[j]);
    c64 sj = t[j] + alpha * hj;  // f = s_j / alpha
    f_s = ia * sj;
    st->dc_alpha = alpha;
    st->dc_f_re = f_s.re;
    st->dc_f_im = f_s.im;
    const double before = sqrt(dd.re) * ia;
    st->dc_before = before;
    Hraw[(size_t)j * ld + j] = hj;
    const int bad = !(isfinite(hj.re) && isfinite(hj.im) && isfinite(before));
    st->nonfinite = bad;
    if (bad) st->done = 1;
  }
  __syncthreads();
  const c64 f = f_s;
  for (int k = tid; k < j; k += KWG) {
    c64 hk = ia * (b[k] - t[k]);
    Hraw[(size_t)j * ld + k] = hk;
    c64 sk = t[k] + alpha * hk;
    st->dc_a[k] = a[k];
    st->dc_e[k] = sk - f * a[k];
  }
  // rows below the new column stay zero
  for (int k = j + 1 + tid; k <= m; k += KWG) Hraw[(size_t)j * ld + k] = c64{0, 0};
  if (tid == 0) end_step(st, ring, j);
}
__global__ __launch_bounds__(KWG) void k_dc_mid(ArnoldiState* st, int has_w, const c64* dots,
                                                double breakdown_eps, StepStatus* ring) {
  k_dc_mid_body(st, has_w, dots, breakdown_eps, ring, blockIdx.x, gridDim.x);
}

// Fused update of step j: slot j <- (u - Q_j a)/alpha, slot j+1 <- (w - Q_j e - f u)/alpha.
template <typename T>
__device__ __forceinline__ void k_update_dc_body(cx<T>* Q, size_t ldv, size_t n, const ArnoldiState* st, unsigned bx, unsigned nbx) {
  __shared__ c64 ca[kMaxRestart], ce[kMaxRestart];
  if (st->done) return;
  const int j = st->ju;
  constexpr int UE = DcShape<T>::UE;
  for (int k = threadIdx.x; k < j; k += KWG) {
    ca[k] = st->dc_a[k];
    ce[k] = st->dc_e[k];
  }
  __syncthreads();
  const double ia = 1.0 / st->dc_alpha;
  const double fr = st->dc_f_re, fi = st->dc_f_im;
  cx<T>* u = Q + (size_t)j * ldv;
  cx<T>* w = Q + (size_t)(j + 1) * ldv;
  // whole KWG * UE groups: the basis padding past n is zero and stays zero
  const size_t npad = (n + KWG * UE - 1) / (KWG * UE) * (KWG * UE);
  const size_t step = (size_t)nbx * KWG * UE;
  for (size_t base = (size_t)bx * KWG * UE + threadIdx.x; base < npad; base += step) {
    double ur[UE], ui[UE], wr[UE], wi[UE];
#pragma unroll
    for (int e = 0; e < UE; e++) {
      size_t i = base + (size_t)e * KWG;
      cx<T> uu = u[i], ww = w[i];
      ur[e] = uu.re;
      ui[e] = uu.im;
      // w - f u uses the original u
      wr[e] = ww.re - (fr * uu.re - fi * uu.im);
      wi[e] = ww.im - (fr * uu.im + fi * uu.re);
    }
    // groups of KG basis vectors, two register sets in turn: the next group's
    // loads are in flight while this group's are applied
    constexpr int KG = 4;
#define VIE_UP_LOAD(k0, qb)                                                                                \
  do {                                                                                                     \
    _Pragma("unroll") for (int p = 0; p < KG; p++) _Pragma("unroll") for (int e = 0; e < UE; e++)          \
        qb[p][e] = ld_stream(Q + (size_t)((k0) + p) * ldv + base + (size_t)e * KWG);                       \
  } while (0)
#define VIE_UP_USE(k0, qb)                                                                                 \
  do {                                                                                                     \
    _Pragma("unroll") for (int p = 0; p < KG; p++) {                                                       \
      const double ar = ca[(k0) + p].re, ai = ca[(k0) + p].im, er = ce[(k0) + p].re, ei = ce[(k0) + p].im; \
      _Pragma("unroll") for (int e = 0; e < UE; e++) {                                                     \
        const double qr = qb[p][e].re, qi = qb[p][e].im;                                                   \
        ur[e] -= ar * qr - ai * qi;                                                                        \
        ui[e] -= ar * qi + ai * qr;                                                                        \
        wr[e] -= er * qr - ei * qi;                                                                        \
        wi[e] -= er * qi + ei * qr;                                                                        \
      }                                                                                                    \
    }                                                                                                      \
  } while (0)
    cx<T> qa[KG][UE], qb[KG][UE];
    const int jg = j / KG * KG;
    int k = 0;
    if (k < jg) VIE_UP_LOAD(k, qa);
    for (; k + 2 * KG < jg; k += 2 * KG) {
      VIE_UP_LOAD(k + KG, qb);
      VIE_UP_USE(k, qa);
      VIE_UP_LOAD(k + 2 * KG, qa);
      VIE_UP_USE(k + KG, qb);
    }
    if (k + KG < jg) {
      VIE_UP_LOAD(k + KG, qb);
      VIE_UP_USE(k, qa);
      VIE_UP_USE(k + KG, qb);
      k += 2 * KG;
    } else if (k < jg) {
      VIE_UP_USE(k, qa);
      k += KG;
    }
#undef VIE_UP_LOAD
#undef VIE_UP_USE
    for (; k < j; k++) {
      const double ar = ca[k].re, ai = ca[k].im, er = ce[k].re, ei = ce[k].im;
#pragma unroll
      for (int e = 0; e < UE; e++) {
        size_t i = base + (size_t)e * KWG;
        cx<T> qq = ld_stream(Q + (size_t)k * ldv + i);
        double qr = qq.re, qi = qq.im;
        ur[e] -= ar * qr - ai * qi;
        ui[e] -= ar * qi + ai * qr;
        wr[e] -= er * qr - ei * qi;
        wi[e] -= er * qi + ei * qr;
      }
    }
#pragma unroll
    for (int e = 0; e < UE; e++) {
      size_t i = base + (size_t)e * KWG;
      u[i] = cx<T>{(T)(ur[e] * ia), (T)(ui[e] * ia)};
      w[i] = cx<T>{(T)(wr[e] * ia), (T)(wi[e] * ia)};
    }
  }
}
template <typename T>
__global__ __launch_bounds__(KWG) void k_update_dc(cx<T>* Q, size_t ldv, size_t n, const ArnoldiState* st) {
  k_update_dc_body<T>(Q, ldv, n, st, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------- elementwise
// Entries [n, ldv) of column blockIdx.x.
template <typename T> __global__ void k_zero_pad(cx<T>* V, size_t ldv, size_t n) {
  cx<T>* col = V + (size_t)blockIdx.x * ldv;
  for (size_t i = n + threadIdx.x; i < ldv; i += blockDim.x) col[i] = cx<T>{0, 0};
}

template <typename T> __global__ void k_zero(cx<T>* x, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = cx<T>{0, 0};
}

template <typename To, typename From> __global__ void k_cast(cx<To>* y, const cx<From>* x, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = cast<To>(x[i]);
}

// V0 = r * (1/beta) in the basis precision.
template <typename T> __global__ void k_scale_from(cx<T>* v, const c64* r, size_t n, const ArnoldiState* st) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double s = st->inv_beta;
    v[i] = cx<T>{(T)(r[i].re * s), (T)(r[i].im * s)};
  }
}

// v *= 1/||w|| in place (skipped on breakdown, where inv_next == 0 leaves the
// slot unused by later columns).
template <typename T> __global__ void k_scale_next(cx<T>* v, size_t n, const ArnoldiState* st) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double s = st->inv_next;
  if (i < n && s != 0.0) v[i] = cx<T>{(T)(v[i].re * s), (T)(v[i].im * s)};
}

// Jacobi inverse of element i of a system vector [c | 3 nbox].
__device__ inline c64 jacobi_inv(const JacobiView& J, size_t i, c64 v) {
  if (i < J.c) return cdiv(v, J.cdiag[i]);
  size_t q = i - J.c;
  size_t comp = q / J.nbox, cell = q - comp * J.nbox;
  c64 e = J.eps[cell], one = {1, 0};
  return cdiv(v, e - (e - one) * J.self[comp]);
}

// x += M^{-1} (sum_k y_k V[k] + sum_l uc_l U[l])   (vie_iterative.c:187-188
// with Z = M^{-1} V; the U term only in a recycled cycle)
template <typename T>
__global__ void k_combine(c64* x, const cx<T>* V, size_t ldv, size_t n, const ArnoldiState* st, JacobiView J,
                          int precond, const cx<T>* U) {
  __shared__ c64 ys[kMaxRestart], us[kMaxRecycle];
  if (st->singular) return;  // no least-squares solution: x stays as it is
  const int used = st->used, base = U ? st->base : 0;
  for (int k = base + threadIdx.x; k < used; k += blockDim.x) ys[k] = st->y[k];
  for (int k = threadIdx.x; k < base; k += blockDim.x) us[k] = st->uc[k];
  __syncthreads();
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  c64 s = {0, 0};
  for (int k = base; k < used; k++) {
    cx<T> a = ld_stream(V + (size_t)k * ldv + i);
    s = s + ys[k] * c64{(double)a.re, (double)a.im};
  }
  for (int k = 0; k < base; k++) {
    cx<T> a = ld_stream(U + (size_t)k * ldv + i);
    s = s + us[k] * c64{(double)a.re, (double)a.im};
  }
  if (precond) s = jacobi_inv(J, i, s);
  x[i] = x[i] + s;
}

// r = b - Ax and the block's share of ||r||^2 (partial[blk * ldp]).
__global__ __launch_bounds__(KWG) void k_residual(const c64* b, const c64* ax, c64* r, size_t n, c64* partial,
                                                  int ldp) {
  __shared__ c64 lds[KWG / 64];
  c64 s = {0, 0};
  const size_t step = (size_t)gridDim.x * KWG;
  for (size_t i = (size_t)blockIdx.x * KWG + threadIdx.x; i < n; i += step) {
    c64 v = b[i] - ax[i];
    r[i] = v;
    s.re += v.re * v.re + v.im * v.im;
  }
  s = block_reduce(s, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

template <typename T>
__global__ __launch_bounds__(KWG) void k_norm2(const cx<T>* x, size_t n, c64* partial, int ldp) {
  __shared__ c64 lds[KWG / 64];
  c64 s = {0, 0};
  const size_t step = (size_t)gridDim.x * KWG;
  for (size_t i = (size_t)blockIdx.x * KWG + threadIdx.x; i < n; i += step) {
    double a = x[i].re, b = x[i].im;
    s.re += a * a + b * b;
  }
  s = block_reduce(s, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// ---------------------------------------------------------------- conductor coupling
// z_c = v_c / C_ii (Jacobi) or v_c   (dielectric_fft.c:77-81), for conductors
// j0 .. j0 + cnt - 1, into LDS.
template <typename T>
__device__ __forceinline__ void stage_zc(cx<T>* zs, const cx<T>* vc, const cx<T>* Cm, int c, int j0, int cnt,
                                         int precond) {
  for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
    const int jj = j0 + q;
    zs[q] = precond ? cdiv(vc[jj], Cm[(size_t)jj * c + jj]) : vc[jj];
  }
}

// y_v += sum_j D[j] z_c[j]   (dielectric_fft.c:63-64), D stored column-wise,
// and, in the first ceil(c / 4) blocks (one wave per conductor row),
// y_c = C z_c + B z_v   (dielectric_fft.c:56-60) with B z_v summed from the
// partials of the preceding B reduction (or taken from bz when given).
template <typename T>
__device__ __forceinline__ void k_couple_out_body(const cx<T>* vc, const cx<T>* Cm, int c, int precond,
                                                    const cx<T>* D, cx<T>* yv, size_t nv, const c64* bpart, int ldp,
                                                    unsigned nblk, const c64* bz, cx<T>* yc, DevIndex di, unsigned bx, unsigned nbx) {
  __shared__ cx<T> zs[kMaxConductorsShared];
  if (di.skipped()) return;
  vc += di.off();
  yv += di.off();
  yc += di.off();
  const size_t i = (size_t)bx * 256 + threadIdx.x;
  if ((size_t)bx * 256 < nv) {
    c64 s = {0, 0};
    if (i < nv) s = c64{(double)yv[i].re, (double)yv[i].im};
    for (int j0 = 0; j0 < c; j0 += kMaxConductorsShared) {
      const int cnt = c - j0 < kMaxConductorsShared ? c - j0 : kMaxConductorsShared;
      __syncthreads();
      stage_zc(zs, vc, Cm, c, j0, cnt, precond);
      __syncthreads();
      if (i < nv)
        for (int j = 0; j < cnt; j++) {
          const cx<T> a = ld_stream(D + (size_t)(j0 + j) * nv + i), z = zs[j];
          s = s + c64{a.re, a.im} * c64{z.re, z.im};
        }
    }
    if (i < nv) yv[i] = cx<T>{(T)s.re, (T)s.im};
  }
  if ((int)bx * 4 < c) {
    const int lane = threadIdx.x & 63, row = bx * 4 + (threadIdx.x >> 6);
    c64 acc = {0, 0};
    if (row < c) acc = bz ? bz[row] : sum_partials(bpart + (size_t)row * ldp, nblk, lane);
    c64 part = {0, 0};
    for (int j0 = 0; j0 < c; j0 += kMaxConductorsShared) {
      const int cnt = c - j0 < kMaxConductorsShared ? c - j0 : kMaxConductorsShared;
      __syncthreads();
      stage_zc(zs, vc, Cm, c, j0, cnt, precond);
      __syncthreads();
      if (row < c)
        for (int j = lane; j < cnt; j += 64) {
          const cx<T> a = Cm[(size_t)row * c + j0 + j], z = zs[j];
          part = part + c64{a.re, a.im} * c64{z.re, z.im};
        }
    }
    part = wave_reduce(part);
    if (row < c && lane == 0) yc[row] = cx<T>{(T)(acc.re + part.re), (T)(acc.im + part.im)};
  }
}
template <typename T>
__global__ __launch_bounds__(256) void k_couple_out(const cx<T>* vc, const cx<T>* Cm, int c, int precond,
                                                    const cx<T>* D, cx<T>* yv, size_t nv, const c64* bpart, int ldp,
                                                    unsigned nblk, const c64* bz, cx<T>* yc, DevIndex di) {
  k_couple_out_body<T>(vc, Cm, c, precond, D, yv, nv, bpart, ldp, nblk, bz, yc, di, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------- layout
// Scatter a host-order vector [c | 3 cells interleaved] into [c | 3][nbox].
template <typename T>
__global__ void k_to_box(const c64* user, size_t c, size_t cells, const size_t* map, size_t nbox, cx<T>* sys) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < c) sys[i] = cast<T>(user[i]);
  else if (i < c + 3 * cells) {
    size_t q = i - c, cell = q / 3, comp = q - 3 * cell;
    sys[c + comp * nbox + map[cell]] = cast<T>(user[i]);
  }
}

template <typename T>
__global__ void k_from_box(const cx<T>* sys, size_t c, size_t cells, const size_t* map, size_t nbox, c64* user) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < c) user[i] = cast<double>(sys[i]);
  else if (i < c + 3 * cells) {
    size_t q = i - c, cell = q / 3, comp = q - 3 * cell;
    user[i] = cast<double>(sys[c + comp * nbox + map[cell]]);
  }
}

// Conductor columns from the reference's row-major v x c block (row 3 cell +
// comp, column r) into the device layout [r][comp * nbox + map[cell]].
__global__ void k_columns_to_box(const c64* __restrict__ host_order, size_t c, size_t cells,
                                 const size_t* __restrict__ map, size_t nbox, c64* __restrict__ D) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * cells * c) return;
  const size_t ui = i / c, r = i - ui * c, cell = ui / 3, comp = ui - 3 * cell;
  D[r * 3 * nbox + comp * nbox + map[cell]] = host_order[i];
}

// Jacobi preconditioner application on a system vector (compatibility path).
__global__ void k_precondition(const c64* x, c64* y, size_t n, JacobiView J) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = jacobi_inv(J, i, x[i]);
}

// out[:, i] = sum_c in[:, c] coef[c + i * nin]: the recycled subspace U, C as
// combinations of a finished cycle's basis.  Each thread owns one element and
// keeps BC_OUT outputs in registers; the coefficients are uniform (scalar
// loads).  out and in do not overlap.
constexpr int BC_OUT = 16;
template <typename T>
__global__ __launch_bounds__(256) void k_block_combine(cx<T>* out, int nout, const cx<T>* in, int nin, size_t ldv,
                                                       const c64* coef, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  for (int i0 = 0; i0 < nout; i0 += BC_OUT) {
    const int cnt = nout - i0 < BC_OUT ? nout - i0 : BC_OUT;
    c64 acc[BC_OUT];
#pragma unroll
    for (int q = 0; q < BC_OUT; q++) acc[q] = c64{0, 0};
    for (int c = 0; c < nin; c++) {
      const cx<T> a = in[(size_t)c * ldv + i];
      const c64 v = {(double)a.re, (double)a.im};
#pragma unroll
      for (int q = 0; q < BC_OUT; q++)
        if (q < cnt) acc[q] = acc[q] + coef[(size_t)(i0 + q) * nin + c] * v;
    }
#pragma unroll
    for (int q = 0; q < BC_OUT; q++)
      if (q < cnt) out[(size_t)(i0 + q) * ldv + i] = cx<T>{(T)acc[q].re, (T)acc[q].im};
  }
}

// Streaming 16-byte copy (HBM traffic calibration probe).
__global__ void k_copy16(float4* __restrict__ dst, const float4* __restrict__ src, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

// ================================================================ host side
void copy16(void* dst, const void* src, size_t n16, hipStream_t s) {
  if (n16) hipLaunchKernelGGL(k_copy16, dim3(4096), dim3(256), 0, s, (float4*)dst, (const float4*)src, n16);
}

static inline unsigned gblocks(size_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

template <typename T> void vec_zero(cx<T>* x, size_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL((k_zero<T>), dim3(gblocks(n)), dim3(256), 0, s, x, n);
}
template <typename To, typename From> void vec_cast(cx<To>* y, const cx<From>* x, size_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL((k_cast<To, From>), dim3(gblocks(n)), dim3(256), 0, s, y, x, n);
}
template void vec_zero<float>(c32*, size_t, hipStream_t);
template void vec_zero<double>(c64*, size_t, hipStream_t);
template void vec_cast<float, double>(c32*, const c64*, size_t, hipStream_t);
template void vec_cast<double, float>(c64*, const c32*, size_t, hipStream_t);
template void vec_cast<double, double>(c64*, const c64*, size_t, hipStream_t);

template <typename T>
void multidot(const cx<T>* V, size_t ldv, int nk, const cx<T>* w, size_t n, int self_norm, ReduceScratch& rs,
              c64* dots, hipStream_t s) {
  unsigned nb = dot_blocks<T>(n);
  JacobiLite<T> none{};
  hipLaunchKernelGGL((k_dots<T, true, false>), dim3(nb), dim3(KWG), 0, s, V, ldv, nk, w, n, self_norm, rs.partial,
                     kMaxDotBlocks, none, DevIndex{});
  int tot = nk + (self_norm ? 1 : 0);
  hipLaunchKernelGGL(k_finish, dim3((tot + 3) / 4), dim3(KWG), 0, s, rs.partial, kMaxDotBlocks, nb, tot, dots);
}
template <typename T>
unsigned update(const cx<T>* V, size_t ldv, int nk, const c64* h, cx<T>* w, size_t n, int want_norm,
                ReduceScratch& rs, hipStream_t s) {
  // its norm partials are laid out kMaxDotBlocks per key
  unsigned nb = std::min(update_blocks(n), (unsigned)kMaxDotBlocks);
  hipLaunchKernelGGL((k_update<T>), dim3(nb), dim3(KWG), 0, s, V, ldv, nk, h, w, n, want_norm, rs.partial, kMaxDotBlocks);
  return nb;
}
static unsigned elem_blocks(size_t n) {
  size_t b = (n + KWG - 1) / KWG;
  return (unsigned)(b < 1 ? 1 : (b > kMaxDotBlocks ? kMaxDotBlocks : b));
}
template <typename T> void norm2(const cx<T>* x, size_t n, ReduceScratch& rs, c64* out, hipStream_t s) {
  unsigned nb = elem_blocks(n);
  hipLaunchKernelGGL((k_norm2<T>), dim3(nb), dim3(KWG), 0, s, x, n, rs.partial, kMaxDotBlocks);
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(KWG), 0, s, rs.partial, kMaxDotBlocks, nb, 1, out);
}
void residual(const c64* b, const c64* ax, c64* r, size_t n, ReduceScratch& rs, c64* out, hipStream_t s) {
  unsigned nb = elem_blocks(n);
  hipLaunchKernelGGL(k_residual, dim3(nb), dim3(KWG), 0, s, b, ax, r, n, rs.partial, kMaxDotBlocks);
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(KWG), 0, s, rs.partial, kMaxDotBlocks, nb, 1, out);
}
void arnoldi_tail(ArnoldiState* st, const c64* h1, const c64* h2, const c64* before2, ReduceScratch& rs,
                  unsigned nblk, double breakdown_eps, hipStream_t s) {
  hipLaunchKernelGGL(k_arnoldi_tail, dim3(1), dim3(KWG), 0, s, st, h1, h2, before2, rs.partial, kMaxDotBlocks, nblk,
                     breakdown_eps);
}
void backsolve(ArnoldiState* st, hipStream_t s) { hipLaunchKernelGGL(k_backsolve, dim3(1), dim3(KWG), 0, s, st); }

template <typename T>
void dcgs2_step(cx<T>* Q, size_t ldv, bool has_w, size_t n, ArnoldiState* st, ReduceScratch& rs, c64* dots,
                double breakdown_eps, int max_j, StepStatus* ring, hipStream_t s) {
  const unsigned nb = dc_blocks<T>(n);
  const int hw = has_w ? 1 : 0;
  hipLaunchKernelGGL((k_dots_dc<T>), dim3(nb), dim3(KWG), 0, s, Q, ldv, (const ArnoldiState*)st, hw, n, rs.partial,
                     kMaxDotBlocks);
  const int max_keys = 2 * max_j + 3;
  hipLaunchKernelGGL(k_finish_dc, dim3((max_keys + 3) / 4), dim3(KWG), 0, s, rs.partial, kMaxDotBlocks, nb,
                     (const ArnoldiState*)st, hw, dots);
  hipLaunchKernelGGL(k_dc_mid, dim3(1), dim3(KWG), 0, s, st, hw, dots, breakdown_eps, ring);
  if (has_w) hipLaunchKernelGGL((k_update_dc<T>), dim3(update_blocks(n, DcShape<T>::UE)), dim3(KWG), 0, s, Q, ldv,
                                n, (const ArnoldiState*)st);
}
// Timing hook: a state at column j with finite update coefficients, and a
// basis whose first j + 2 columns hold a constant, so the two streaming
// kernels of a delayed-CGS2 step can be timed on their own.
__global__ void k_time_state(ArnoldiState* st, int j) {
  for (int k = threadIdx.x; k < j; k += blockDim.x) {
    st->dc_a[k] = c64{1e-3, 0};
    st->dc_e[k] = c64{1e-3, 0};
  }
  if (threadIdx.x == 0) {
    st->j = j;
    st->ju = j;
    st->done = 0;
    st->dc_alpha = 1.0;
    st->dc_f_re = 0.0;
    st->dc_f_im = 0.0;
  }
}

template <typename T> __global__ void k_fill(cx<T>* x, size_t n, T v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = cx<T>{v, (T)0};
}

template <typename T>
int dcgs2_time(cx<T>* Q, size_t ldv, size_t n, ArnoldiState* st, ReduceScratch& rs, int j, int reps, hipStream_t s,
               float* ms_reduce, float* ms_update) {
  const unsigned nb = dc_blocks<T>(n);
  const size_t cells = (size_t)(j + 2) * ldv;
  hipLaunchKernelGGL((k_fill<T>), dim3(gblocks(cells)), dim3(256), 0, s, Q, cells, (T)1e-3);
  hipLaunchKernelGGL(k_time_state, dim3(1), dim3(256), 0, s, st, j);
  hipEvent_t e[3];
  for (auto& x : e)
    if (hipEventCreate(&x) != hipSuccess) return 0;
  for (int r = 0; r < 2; r++) {
    hipLaunchKernelGGL((k_dots_dc<T>), dim3(nb), dim3(KWG), 0, s, Q, ldv, (const ArnoldiState*)st, 1, n, rs.partial,
                       kMaxDotBlocks);
    hipLaunchKernelGGL((k_update_dc<T>), dim3(update_blocks(n, DcShape<T>::UE)), dim3(KWG), 0, s, Q, ldv, n,
                       (const ArnoldiState*)st);
  }
  (void)hipEventRecord(e[0], s);
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL((k_dots_dc<T>), dim3(nb), dim3(KWG), 0, s, Q, ldv, (const ArnoldiState*)st, 1, n, rs.partial,
                       kMaxDotBlocks);
  (void)hipEventRecord(e[1], s);
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL((k_update_dc<T>), dim3(update_blocks(n, DcShape<T>::UE)), dim3(KWG), 0, s, Q, ldv, n,
                       (const ArnoldiState*)st);
  (void)hipEventRecord(e[2], s);
  int ok = hipEventSynchronize(e[2]) == hipSuccess;
  ok = ok && hipEventElapsedTime(ms_reduce, e[0], e[1]) == hipSuccess &&
       hipEventElapsedTime(ms_update, e[1], e[2]) == hipSuccess;
  *ms_reduce /= reps;
  *ms_update /= reps;
  for (auto& x : e) (void)hipEventDestroy(x);
  return ok && hipGetLastError() == hipSuccess;
}

void cycle_init(ArnoldiState* st, const c64* beta2, double inner_target, int base, hipStream_t s) {
  hipLaunchKernelGGL(k_cycle_init, dim3(1), dim3(256), 0, s, st, beta2, inner_target, base);
}
template <typename T>
void recycle_project(cx<T>* Q, size_t ldv, int k, size_t n, ArnoldiState* st, ReduceScratch& rs, hipStream_t s) {
  c64* zc = st->zc;  // device address of the state's projection coefficients
  multidot<T>(Q, ldv, k, Q + (size_t)k * ldv, n, 0, rs, zc, s);
  update<T>(Q, ldv, k, zc, Q + (size_t)k * ldv, n, 0, rs, s);
}
template <typename T>
void block_combine(cx<T>* out, int nout, const cx<T>* in, int nin, size_t ldv, const c64* coef, size_t n,
                   hipStream_t s) {
  if (n && nout > 0 && nin > 0)
    hipLaunchKernelGGL((k_block_combine<T>), dim3(gblocks(n)), dim3(256), 0, s, out, nout, in, nin, ldv, coef, n);
}
template <typename T> void zero_padding(cx<T>* V, size_t ldv, size_t n, size_t cols, hipStream_t s) {
  if (cols && ldv > n) hipLaunchKernelGGL((k_zero_pad<T>), dim3((unsigned)cols), dim3(256), 0, s, V, ldv, n);
}
template <typename T> void scale_from(cx<T>* v, const c64* r, size_t n, const ArnoldiState* st, hipStream_t s) {
  hipLaunchKernelGGL((k_scale_from<T>), dim3(gblocks(n)), dim3(256), 0, s, v, r, n, st);
}
template <typename T> void scale_next(cx<T>* v, size_t n, const ArnoldiState* st, hipStream_t s) {
  hipLaunchKernelGGL((k_scale_next<T>), dim3(gblocks(n)), dim3(256), 0, s, v, n, st);
}
template <typename T>
void combine(c64* x, const cx<T>* V, size_t ldv, size_t n, const ArnoldiState* st, const JacobiView& J, int precond,
             hipStream_t s, const cx<T>* U) {
  hipLaunchKernelGGL((k_combine<T>), dim3(gblocks(n)), dim3(256), 0, s, x, V, ldv, n, st, J, precond, U);
}
template <typename T>
void to_box(const c64* user, size_t c, size_t cells, const size_t* map, size_t nbox, cx<T>* sys, hipStream_t s) {
  size_t n = c + 3 * cells;
  if (n) hipLaunchKernelGGL((k_to_box<T>), dim3(gblocks(n)), dim3(256), 0, s, user, c, cells, map, nbox, sys);
}
template <typename T>
void from_box(const cx<T>* sys, size_t c, size_t cells, const size_t* map, size_t nbox, c64* user, hipStream_t s) {
  size_t n = c + 3 * cells;
  if (n) hipLaunchKernelGGL((k_from_box<T>), dim3(gblocks(n)), dim3(256), 0, s, sys, c, cells, map, nbox, user);
}
void columns_to_box(const c64* host_order, size_t c, size_t cells, const size_t* map, size_t nbox, c64* D,
                    hipStream_t s) {
  const size_t n = 3 * cells * c;
  if (n) hipLaunchKernelGGL(k_columns_to_box, dim3(gblocks(n)), dim3(256), 0, s, host_order, c, cells, map, nbox, D);
}
void precondition_apply(const c64* x, c64* y, size_t n, const JacobiView& J, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_precondition, dim3(gblocks(n)), dim3(256), 0, s, x, y, n, J);
}

template <typename T>
void coupling_apply(const CouplingView<T>& cv, const cx<T>* vc, const cx<T>* vv, cx<T>* yc, cx<T>* yv, int precond,
                    ReduceScratch& rs, hipStream_t s, DevIndex di) {
  const int c = cv.c;
  if (c == 0) return;
  // B z_v: one reduction launch for up to kMaxKeys conductor rows, whose
  // partials the output kernel sums; more rows go through bz chunk by chunk
  const c64* bpart = rs.partial;
  const c64* bz = nullptr;
  unsigned nb = 0;
  if (cv.nv) {
    JacobiLite<T> J{cv.nbox, cv.eps, {cv.self[0], cv.self[1], cv.self[2]}};
    nb = dot_blocks<T>(cv.nv);
    for (int r0 = 0; r0 < c; r0 += kMaxKeys) {
      int rows = c - r0 < kMaxKeys ? c - r0 : kMaxKeys;
      const cx<T>* Br = cv.B + (size_t)r0 * cv.nv;
      if (precond)
        hipLaunchKernelGGL((k_dots<T, false, true>), dim3(nb), dim3(KWG), 0, s, Br, cv.nv, rows, vv, cv.nv, 0,
                           rs.partial, kMaxDotBlocks, J, di);
      else
        hipLaunchKernelGGL((k_dots<T, false, false>), dim3(nb), dim3(KWG), 0, s, Br, cv.nv, rows, vv, cv.nv, 0,
                           rs.partial, kMaxDotBlocks, J, di);
      if (c > kMaxKeys) {
        hipLaunchKernelGGL(k_finish, dim3((rows + 3) / 4), dim3(KWG), 0, s, rs.partial, kMaxDotBlocks, nb, rows,
                           cv.bz + r0);
        bz = cv.bz;
      }
    }
  } else {
    hipLaunchKernelGGL((k_zero<double>), dim3(gblocks(c)), dim3(256), 0, s, cv.bz, (size_t)c);
    bz = cv.bz;
  }
  const size_t vblocks = (cv.nv + 255) / 256, cblocks = ((size_t)c + 3) / 4;
  hipLaunchKernelGGL((k_couple_out<T>), dim3((unsigned)(vblocks > cblocks ? vblocks : cblocks)), dim3(256), 0, s, vc,
                     cv.C, c, precond, cv.D, yv, cv.nv, bpart, kMaxDotBlocks, nb, bz, yc, di);
}

// ---------------------------------------------------------------- multi-job step
template <typename T> void step_blocks_mj(MjStep<T>& m) {
  m.nb_dots = m.c > 0 && m.nv > 0 ? dot_blocks<T>(m.nv) : 0;
  m.nb_out = m.c > 0 ? (unsigned)std::max<size_t>((m.nv + 255) / 256, ((size_t)m.c + 3) / 4) : 0;
  m.nb_dc = dc_blocks<T>(m.n);
  m.nb_finish = (unsigned)((2 * m.max_j + 3 + 3) / 4);
  m.nb_update = update_blocks(m.n, DcShape<T>::UE);
}

// One job's step arguments from the table, pointers marked global.
template <typename T> __device__ __forceinline__ MjStep<T> step_global(const MjStep<T>& t) {
  MjStep<T> m = t;
  if (m.c > 0) {
    m.C = to_global(m.C);
    m.B = to_global(m.B);
    m.D = to_global(m.D);
  }
  m.x = to_global(m.x);
  m.y = to_global(m.y);
  m.J.eps = to_global(m.J.eps);
  m.di = to_global(m.di);
  m.Q = to_global(m.Q);
  m.st = to_global(m.st);
  m.partial = to_global(m.partial);
  m.dots = to_global(m.dots);
  m.ring = to_global(m.ring);
  return m;
}

// (the multi-job reductions carry no minimum occupancy: at four waves per
// SIMD their chunk registers would spill to scratch)
template <typename T>
__global__ __launch_bounds__(KWG) void k_dots_mj(const MjStep<T>* __restrict__ tab, const unsigned* __restrict__ pre,
                                                 int count) {
  const int q = mj_job(pre, count, blockIdx.x);
  const MjStep<T> m = step_global(tab[q]);
  k_dots_body<T, false, true>(m.B, m.nv, m.c, m.x + m.c, m.nv, 0, m.partial, kMaxDotBlocks, m.J, m.di,
                              blockIdx.x - pre[q], m.nb_dots);
}
template <typename T>
__global__ __launch_bounds__(256) void k_couple_out_mj(const MjStep<T>* __restrict__ tab,
                                                       const unsigned* __restrict__ pre, int count) {
  const int q = mj_job(pre, count, blockIdx.x);
  const MjStep<T> m = step_global(tab[q]);
  k_couple_out_body<T>(m.x, m.C, m.c, 1, m.D, m.y + m.c, m.nv, m.partial, kMaxDotBlocks, m.nb_dots, nullptr, m.y,
                       m.di, blockIdx.x - pre[q], m.nb_out);
}
template <typename T>
__global__ __launch_bounds__(KWG) void k_dots_dc_mj(const MjStep<T>* __restrict__ tab,
                                                    const unsigned* __restrict__ pre, int count) {
  const int q = mj_job(pre, count, blockIdx.x);
  const MjStep<T> m = step_global(tab[q]);
  k_dots_dc_body<T>(m.Q, m.ldv, m.st, 1, m.n, m.partial, kMaxDotBlocks, blockIdx.x - pre[q], m.nb_dc);
}
template <typename T>
__global__ __launch_bounds__(KWG) void k_finish_dc_mj(const MjStep<T>* __restrict__ tab,
                                                      const unsigned* __restrict__ pre, int count) {
  const int q = mj_job(pre, count, blockIdx.x);
  const MjStep<T> m = step_global(tab[q]);
  k_finish_dc_body(m.partial, kMaxDotBlocks, m.nb_dc, m.st, 1, m.dots, blockIdx.x - pre[q], m.nb_finish);
}
template <typename T> __global__ __launch_bounds__(KWG) void k_dc_mid_mj(const MjStep<T>* __restrict__ tab) {
  const MjStep<T> m = step_global(tab[blockIdx.x]);
  k_dc_mid_body(m.st, 1, m.dots, m.eps_b, m.ring, 0, 1);
}
template <typename T>
__global__ __launch_bounds__(KWG) void k_update_dc_mj(const MjStep<T>* __restrict__ tab,
                                                      const unsigned* __restrict__ pre, int count) {
  const int q = mj_job(pre, count, blockIdx.x);
  const MjStep<T> m = step_global(tab[q]);
  k_update_dc_body<T>(m.Q, m.ldv, m.n, m.st, blockIdx.x - pre[q], m.nb_update);
}

template <typename T>
void step_mj(const MjStep<T>* tab, int count, const unsigned* pre, const unsigned total[5], hipStream_t s) {
  if (count <= 0) return;
  const size_t P = (size_t)count + 1;
  if (total[0]) hipLaunchKernelGGL((k_dots_mj<T>), dim3(total[0]), dim3(KWG), 0, s, tab, pre, count);
  if (total[1]) hipLaunchKernelGGL((k_couple_out_mj<T>), dim3(total[1]), dim3(256), 0, s, tab, pre + P, count);
  hipLaunchKernelGGL((k_dots_dc_mj<T>), dim3(total[2]), dim3(KWG), 0, s, tab, pre + 2 * P, count);
  hipLaunchKernelGGL((k_finish_dc_mj<T>), dim3(total[3]), dim3(KWG), 0, s, tab, pre + 3 * P, count);
  hipLaunchKernelGGL((k_dc_mid_mj<T>), dim3((unsigned)count), dim3(KWG), 0, s, tab);
  hipLaunchKernelGGL((k_update_dc_mj<T>), dim3(total[4]), dim3(KWG), 0, s, tab, pre + 4 * P, count);
}

#define VIE_INST(T)                                                                                              \
  template void multidot<T>(const cx<T>*, size_t, int, const cx<T>*, size_t, int, ReduceScratch&, c64*,            \
                            hipStream_t);                                                                        \
  template unsigned update<T>(const cx<T>*, size_t, int, const c64*, cx<T>*, size_t, int, ReduceScratch&,          \
                              hipStream_t);                                                                      \
  template void norm2<T>(const cx<T>*, size_t, ReduceScratch&, c64*, hipStream_t);                               \
  template void scale_from<T>(cx<T>*, const c64*, size_t, const ArnoldiState*, hipStream_t);                     \
  template void zero_padding<T>(cx<T>*, size_t, size_t, size_t, hipStream_t);                                    \
  template void scale_next<T>(cx<T>*, size_t, const ArnoldiState*, hipStream_t);                                 \
  template void combine<T>(c64*, const cx<T>*, size_t, size_t, const ArnoldiState*, const JacobiView&, int,       \
                           hipStream_t, const cx<T>*);                                                           \
  template void recycle_project<T>(cx<T>*, size_t, int, size_t, ArnoldiState*, ReduceScratch&, hipStream_t);      \
  template void block_combine<T>(cx<T>*, int, const cx<T>*, int, size_t, const c64*, size_t, hipStream_t);        \
  template void to_box<T>(const c64*, size_t, size_t, const size_t*, size_t, cx<T>*, hipStream_t);                \
  template void from_box<T>(const cx<T>*, size_t, size_t, const size_t*, size_t, c64*, hipStream_t);              \
  template void coupling_apply<T>(const CouplingView<T>&, const cx<T>*, const cx<T>*, cx<T>*, cx<T>*, int,         \
                                  ReduceScratch&, hipStream_t, DevIndex);                                        \
  template void dcgs2_step<T>(cx<T>*, size_t, bool, size_t, ArnoldiState*, ReduceScratch&, c64*, double, int,    \
                              StepStatus*, hipStream_t);                                                         \
  template int dcgs2_time<T>(cx<T>*, size_t, size_t, ArnoldiState*, ReduceScratch&, int, int, hipStream_t, float*, \
                             float*);                                                                            \
  template void step_blocks_mj<T>(MjStep<T>&);                                                                   \
  template void step_mj<T>(const MjStep<T>*, int, const unsigned*, const unsigned*, hipStream_t);
VIE_INST(float)
VIE_INST(double)

}  // namespace vie
