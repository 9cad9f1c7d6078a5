// Fused five-pass FFT volume operator for gfx950 (kernels and launch
// dispatch; instantiated per precision by vie_passes_f32.hip / _f64.hip).
//
// Computes, for the unknown box [3][nx][ny][nz] (component-major, z fastest),
//   y = eps z - (eps - 1) IFFT( T^ . FFT( pad(z) ) ),   z = D^{-1} v or v,
// which is the reference's apply (vie_fft.c:239-264) with its Jacobi
// preconditioner (vie_fft.c:265-273) optionally fused on the input.  Passes:
//   1. Z forward : scatter + zero-pad (+ PEC image) + Jacobi, FFT along z     -> A[3][nx][ny][Pz]
//   2. Y forward : FFT along y of the ny nonzero rows                         -> B[3][nx][Py][Pz]
//   3. X         : FFT along x, 3x3 symmetric spectral multiply, inverse FFT,
//                  keep the nx physical rows                                  -> B (in place)
//   4. Y inverse : inverse FFT along y, keep ny rows                          -> A
//   5. Z inverse : inverse FFT along z, gather physical z, contrast           -> y
// Only the region that can hold nonzeros is ever read or written, so each
// pass touches the minimum the pruned pencil pipeline allows.  The 1/(PxPyPz)
// inverse scaling lives in the spectrum.
#pragma once
#include <algorithm>
#include <type_traits>

#include "fft_core.h"
#include "vie_kernels.h"

namespace vie {


template <typename T> __device__ __forceinline__ cx<T> jacobi_scale(cx<T> v, cx<T> e, cx<T> s) {
  // v / (eps - (eps - 1) self)   (vie_fft.c:271)
  cx<T> one = {(T)1, (T)0};
  return cdiv(v, e - (e - one) * s);
}

template <typename T> __device__ __forceinline__ cx<T> ld(const cx<T>* p) { return *p; }

// ---------------------------------------------------------------- pass 1
// All three components' inputs are loaded into registers before the first
// transform so their HBM latency overlaps; each transformed component is
// stored as soon as it is done.
template <typename T, int PZ, int LW, int WG>
__device__ __forceinline__ void zfwd_body(const OpArgs<T>& a, unsigned bx, unsigned nbx) {
  constexpr int E = LW * PZ, EPT = (E + WG - 1) / WG;
  __shared__ cx<T> lds[E];
  __shared__ cx<T> tw[PZ];
  if (a.di.skipped()) return;
  const int tid = threadIdx.x;
  const Geom& g = a.g;
  const long lines = (long)g.nx * g.ny;  // (x, y) pencils
  const long line0 = (long)bx * LW;
  const cx<T>* twz = a.tw + g.Px + g.Py;
  for (int i = tid; i < PZ; i += WG) tw[i] = twz[i];
  const size_t nbox = g.nbox();
  const cx<T>* xin = a.xin + a.di.off();
  cx<T> in[3][EPT];
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int f = tid + i * WG;
    const int l = f / PZ, e = f % PZ;
    const long line = line0 + l;
    int z = -1;
    bool image = false;
    if ((E % WG == 0 || f < E) && line < lines) {
      if (e >= g.zoff && e - g.zoff < g.nz) z = e - g.zoff;
      else if (g.pec && e < g.nz) {
        z = g.nz - 1 - e;
        image = true;
      }
    }
    cx<T> ev = {(T)1, (T)0};
    size_t cell = z >= 0 ? (size_t)line * g.nz + z : 0;
    if (z >= 0 && a.precond) ev = a.eps[cell];
#pragma unroll
    for (int comp = 0; comp < 3; comp++) {
      cx<T> v = {(T)0, (T)0};
      if (z >= 0) {
        v = xin[comp * nbox + cell];
        if (a.precond) v = jacobi_scale(v, ev, a.self[comp]);
        if (image && comp != 2) v = cx<T>{-v.re, -v.im};
      }
      in[comp][i] = v;
    }
  }
#pragma unroll
  for (int comp = 0; comp < 3; comp++) {
#pragma unroll
    for (int i = 0; i < EPT; i++) {
      const int f = tid + i * WG;
      if (E % WG == 0 || f < E) lds[f] = in[comp][i];  // f = l * PZ + e
    }
    __syncthreads();
    fft_lines<T, PZ, LW, WG, 1, PZ, false>(lds, tw, tid);
    for (int f = tid; f < E; f += WG) {
      int l = f / PZ, e = f % PZ;
      long line = line0 + l;
      if (line < lines) a.A[((size_t)comp * lines + line) * PZ + e] = lds[l * PZ + e];
    }
    __syncthreads();
  }
}

template <typename T, int PZ, int LW, int WG>
__global__ __launch_bounds__(WG) void k_zfwd(OpArgs<T> a) {
  zfwd_body<T, PZ, LW, WG>(a, blockIdx.x, gridDim.x);
}
template <typename T, int PZ, int LW, int WG>
__global__ __launch_bounds__(WG) void k_zfwd_mj(const OpArgs<T>* __restrict__ ops, const unsigned* __restrict__ blocks,
                                                const int* __restrict__ list,
                                                const unsigned* __restrict__ pre, int count) {
  const int i = mj_job(pre, count, blockIdx.x), q = list[i];
  zfwd_body<T, PZ, LW, WG>(op_args_global(ops[q]), blockIdx.x - pre[i], blocks[5 * q + 0]);
}


// ---------------------------------------------------------------- pass 2 / 4
// Lines are (comp, x, kz) with kz fastest; a workgroup takes LW consecutive
// lines so each y row it loads is LW contiguous values.
template <typename T, int PY, int LW, int WG, bool INV>
__device__ __forceinline__ void ypass_body(const OpArgs<T>& a, unsigned bx, unsigned nbx) {
  constexpr int E = LW * PY;
  __shared__ cx<T> lds[E];
  __shared__ cx<T> tw[PY];
  if (a.di.skipped()) return;
  const int tid = threadIdx.x;
  const Geom& g = a.g;
  const long lines = 3L * g.nx * g.Pz;
  const long line0 = (long)bx * LW;
  const cx<T>* twy = a.tw + g.Px;
  for (int i = tid; i < PY; i += WG) tw[i] = twy[i];
  const int nin = INV ? PY : g.ny;      // rows read
  const cx<T>* src = INV ? a.B : a.A;
  const int srows = INV ? g.Py : g.ny;  // rows per (comp, x) in the source
  for (int f = tid; f < E; f += WG) {
    int e = f / LW, l = f % LW;
    long line = line0 + l;
    cx<T> v = {(T)0, (T)0};
    if (line < lines && e < nin) {
      long cxi = line / g.Pz, kz = line % g.Pz;
      v = src[((size_t)cxi * srows + e) * g.Pz + kz];
      if (INV) v = conj(v);
    }
    lds[e * LW + l] = v;
  }
  __syncthreads();
  fft_lines<T, PY, LW, WG, LW, 1, true>(lds, tw, tid);
  cx<T>* dst = INV ? a.A : a.B;
  const int nout = INV ? g.ny : PY;
  const int drows = INV ? g.ny : g.Py;
  for (int f = tid; f < E; f += WG) {
    int e = f / LW, l = f % LW;
    long line = line0 + l;
    if (line < lines && e < nout) {
      long cxi = line / g.Pz, kz = line % g.Pz;
      cx<T> v = lds[e * LW + l];
      dst[((size_t)cxi * drows + e) * g.Pz + kz] = INV ? conj(v) : v;
    }
  }
}

template <typename T, int PY, int LW, int WG, bool INV>
__global__ __launch_bounds__(WG) void k_ypass(OpArgs<T> a) {
  ypass_body<T, PY, LW, WG, INV>(a, blockIdx.x, gridDim.x);
}
template <typename T, int PY, int LW, int WG, bool INV>
__global__ __launch_bounds__(WG) void k_ypass_mj(const OpArgs<T>* __restrict__ ops, const unsigned* __restrict__ blocks,
                                                const int* __restrict__ list,
                                                 const unsigned* __restrict__ pre, int count) {
  const int i = mj_job(pre, count, blockIdx.x), q = list[i];
  ypass_body<T, PY, LW, WG, INV>(op_args_global(ops[q]), blockIdx.x - pre[i], blocks[5 * q + (INV ? 3 : 1)]);
}


// ---------------------------------------------------------------- 16-point DFT in registers
template <typename T> __device__ __forceinline__ cx<T> w16(cx<T> x, int e) {
  // x * exp(-2 pi i e / 16) for the e of a 4 x 4 split (e in 0..9)
  const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173, r2 = (T)0.70710678118654752440;
  switch (e) {
    case 0: return x;
    case 1: return cx<T>{c1 * x.re + s1 * x.im, c1 * x.im - s1 * x.re};
    case 2: return cx<T>{r2 * (x.re + x.im), r2 * (x.im - x.re)};
    case 3: return cx<T>{s1 * x.re + c1 * x.im, s1 * x.im - c1 * x.re};
    case 4: return mul_mi(x);
    case 6: return cx<T>{r2 * (x.im - x.re), -r2 * (x.re + x.im)};
    default: return cx<T>{-c1 * x.re - s1 * x.im, s1 * x.re - c1 * x.im};  // 9
  }
}

// 16-point DFT of v in place, natural order, split n = 4a + b, k = c + 4d.
// IN8: v[8..15] are zero on entry; OUT8: only X[0..7] are formed.
template <typename T, bool IN8, bool OUT8> __device__ __forceinline__ void dft16(cx<T>* v) {
  cx<T> u[4][4];
#pragma unroll
  for (int b = 0; b < 4; b++) {
    if constexpr (IN8) {
      const cx<T> x0 = v[b], x1 = v[4 + b];
      u[b][0] = x0 + x1;
      u[b][1] = x0 + mul_mi(x1);
      u[b][2] = x0 - x1;
      u[b][3] = x0 - mul_mi(x1);
    } else {
      cx<T> w[4] = {v[b], v[4 + b], v[8 + b], v[12 + b]};
      dft4(w);
#pragma unroll
      for (int c = 0; c < 4; c++) u[b][c] = w[c];
    }
  }
#pragma unroll
  for (int b = 1; b < 4; b++)
#pragma unroll
    for (int c = 1; c < 4; c++) u[b][c] = w16(u[b][c], b * c);
#pragma unroll
  for (int c = 0; c < 4; c++) {
    if constexpr (OUT8) {
      const cx<T> s02 = u[0][c] + u[2][c], d02 = u[0][c] - u[2][c];
      const cx<T> s13 = u[1][c] + u[3][c], d13 = u[1][c] - u[3][c];
      v[c] = s02 + s13;
      v[c + 4] = d02 + mul_mi(d13);
    } else {
      cx<T> w[4] = {u[0][c], u[1][c], u[2][c], u[3][c]};
      dft4(w);
#pragma unroll
      for (int d = 0; d < 4; d++) v[c + 4 * d] = w[d];
    }
  }
}

// ---------------------------------------------------------------- pass 2 / 4, registers
// The y passes for P = 256 as the two-level 16 x 16 transform of the
// register-resident x pass (one component per line, so one 16-entry register
// set per thread): thread (l, c) serves line l of LW consecutive kz lines.
//   forward: inputs y = c + 16 n1 with n1 < 8 (rows below ny <= P/2), DFT over
//     n1, twiddle, LDS transpose, DFT over n2; stores X[c + 16 k2];
//   inverse: loads X[c + 16 k2], DFT over k2, twiddle, LDS transpose, DFT
//     over k1 forming only the rows m1 < 8; stores the ny physical rows.
template <typename T, int LW, bool INV>
__global__ __launch_bounds__(16 * LW) void k_ypass_r16(OpArgs<T> a) {
  constexpr int PY = 256;
  __shared__ cx<T> lds[LW * PY];
  __shared__ cx<T> tw[16][16];
  if (a.di.skipped()) return;
  const int tid = threadIdx.x, l = tid % LW, c = tid / LW;
  const Geom& g = a.g;
  const long lines = 3L * g.nx * g.Pz;
  const long line = (long)blockIdx.x * LW + l;
  const bool live = line < lines;
  const size_t cxi = live ? (size_t)(line / g.Pz) : 0, kz = live ? (size_t)(line % g.Pz) : 0;
  for (int i = tid; i < 256; i += 16 * LW) tw[i / 16][i % 16] = a.tw[g.Px + (i / 16) * (i % 16)];
  const cx<T>* src = INV ? a.B : a.A;
  cx<T>* dst = INV ? a.A : a.B;
  const int srows = INV ? g.Py : g.ny, drows = INV ? g.ny : g.Py;
  cx<T> v[16];
  if constexpr (!INV) {
#pragma unroll
    for (int n1 = 0; n1 < 8; n1++) {
      const int e = c + 16 * n1;
      const bool ok = live && e < g.ny;
      const cx<T> val = src[ok ? ((cxi * srows + e) * g.Pz + kz) : 0];
      v[n1] = ok ? val : cx<T>{(T)0, (T)0};
    }
#pragma unroll
    for (int n1 = 8; n1 < 16; n1++) v[n1] = cx<T>{(T)0, (T)0};
    __syncthreads();  // twiddle table
    dft16<T, true, false>(v);
  } else {
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) v[k2] = conj(src[(cxi * srows + c + 16 * k2) * g.Pz + kz]);
    __syncthreads();  // twiddle table
    dft16<T, false, false>(v);
  }
#pragma unroll
  for (int k = 1; k < 16; k++) v[k] = v[k] * tw[c][k];
#pragma unroll
  for (int k = 0; k < 16; k++) lds[(k * 16 + c) * LW + l] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) v[k] = lds[(c * 16 + k) * LW + l];
  if constexpr (!INV) {
    dft16<T, false, false>(v);
    if (live)
#pragma unroll
      for (int k2 = 0; k2 < 16; k2++) dst[(cxi * drows + c + 16 * k2) * g.Pz + kz] = v[k2];
  } else {
    dft16<T, false, true>(v);
#pragma unroll
    for (int m1 = 0; m1 < 8; m1++) {
      const int e = c + 16 * m1;
      if (live && e < g.ny) dst[(cxi * drows + e) * g.Pz + kz] = conj(v[m1]);
    }
  }
}

// ---------------------------------------------------------------- passes 1+2 and 4+5 fused
// For lattices whose padded y and z extents are at most 128, one workgroup
// owns a whole x plane of one component, so the z and y transforms of that
// plane meet in LDS and the [3][nx][ny][Pz] workspace between them is never
// written or read (a round trip of 3 nx ny Pz values per direction, and one
// launch fewer).  Forward: the ny rows of the plane (scatter, Jacobi scaling,
// PEC image as pass 1) are loaded and z-transformed in LDS; then C kz columns
// at a time are copied into a padded scratch (line stride PY + 1 against bank
// conflicts), zero-extended to PY and y-transformed, and stored as kz-
// contiguous rows of B.  Inverse: C kz columns of B at a time (the next
// chunk's loads in flight) are y-transformed and their ny physical rows kept
// in LDS; then the rows are z-transformed and gathered with the contrast as
// pass 5.  The butterflies are those of the separate passes (same radix plan
// and twiddles), so the products agree with them to rounding.
// One workgroup holds a whole plane, so a plane of 128 x 128 takes the CU
// alone: it gets 1024 threads (FP32, 32 columns per y chunk) or 512 (FP64,
// whose plane leaves LDS room for 8 columns) to keep the CU's SIMDs busy;
// smaller planes run 256 threads and several workgroups per CU.
template <typename T, int PZ, int PY> struct FusedShape {
  static constexpr bool big = PZ * PY >= 128 * 128;
  static constexpr int C = std::is_same<T, float>::value ? (big ? 32 : 16) : 8;
  static constexpr int WG = big ? (std::is_same<T, float>::value ? 1024 : 512) : 256;
};

template <typename T, int PZ, int PY>
__device__ __forceinline__ void zy_fwd_body(const OpArgs<T>& a, unsigned bx) {
  constexpr int NY = PY / 2, C = FusedShape<T, PZ, PY>::C, WG = FusedShape<T, PZ, PY>::WG, LS = PY + 1;
  constexpr int E1 = NY * PZ, EPT1 = (E1 + WG - 1) / WG, E2 = C * PY, EPT2 = (E2 + WG - 1) / WG;
  __shared__ cx<T> zs[E1];
  __shared__ cx<T> ys[C * LS];
  __shared__ cx<T> twz[PZ], twy[PY];
  if (a.di.skipped()) return;
  const int tid = threadIdx.x;
  const Geom& g = a.g;
  const int comp = (int)(bx / g.nx), x = (int)(bx % g.nx);
  for (int i = tid; i < PZ; i += WG) twz[i] = a.tw[g.Px + g.Py + i];
  for (int i = tid; i < PY; i += WG) twy[i] = a.tw[g.Px + i];
  const size_t nbox = g.nbox();
  const cx<T>* xin = a.xin + a.di.off();
  const cx<T> sf = a.self[comp];
  cx<T> in[EPT1];
#pragma unroll
  for (int i = 0; i < EPT1; i++) {
    const int f = tid + i * WG;
    const int y = f / PZ, e = f % PZ;
    int z = -1;
    bool image = false;
    if ((E1 % WG == 0 || f < E1) && y < g.ny) {
      if (e >= g.zoff && e - g.zoff < g.nz) z = e - g.zoff;
      else if (g.pec && e < g.nz) {
        z = g.nz - 1 - e;
        image = true;
      }
    }
    cx<T> v = {(T)0, (T)0};
    if (z >= 0) {
      const size_t cell = ((size_t)x * g.ny + y) * g.nz + z;
      v = xin[comp * nbox + cell];
      if (a.precond) v = jacobi_scale(v, a.eps[cell], sf);
      if (image && comp != 2) v = cx<T>{-v.re, -v.im};
    }
    in[i] = v;
  }
#pragma unroll
  for (int i = 0; i < EPT1; i++) {
    const int f = tid + i * WG;
    if (E1 % WG == 0 || f < E1) zs[f] = in[i];  // row y at zs[y PZ]
  }
  __syncthreads();
  fft_lines<T, PZ, NY, WG, 1, PZ, false>(zs, twz, tid);
  cx<T>* out = a.B + ((size_t)comp * g.nx + x) * PY * PZ;
  for (int k0 = 0; k0 < PZ; k0 += C) {
#pragma unroll
    for (int i = 0; i < EPT2; i++) {
      const int f = tid + i * WG;
      if (E2 % WG == 0 || f < E2) {
        const int c = f % C, y = f / C;
        ys[c * LS + y] = y < g.ny ? zs[y * PZ + k0 + c] : cx<T>{(T)0, (T)0};
      }
    }
    __syncthreads();
    fft_lines<T, PY, C, WG, 1, LS, true>(ys, twy, tid);
#pragma unroll
    for (int i = 0; i < EPT2; i++) {
      const int f = tid + i * WG;
      if (E2 % WG == 0 || f < E2) {
        const int c = f % C, ky = f / C;
        out[(size_t)ky * PZ + k0 + c] = ys[c * LS + ky];
      }
    }
    __syncthreads();
  }
}

template <typename T, int PY, int PZ>
__device__ __forceinline__ void yz_inv_body(const OpArgs<T>& a, unsigned bx) {
  constexpr int NY = PY / 2, C = FusedShape<T, PZ, PY>::C, WG = FusedShape<T, PZ, PY>::WG, LS = PY + 1;
  constexpr int E2 = C * PY, EPT2 = (E2 + WG - 1) / WG;
  __shared__ cx<T> zs[NY * PZ];
  __shared__ cx<T> ys[C * LS];
  __shared__ cx<T> twz[PZ], twy[PY];
  if (a.di.skipped()) return;
  const int tid = threadIdx.x;
  const Geom& g = a.g;
  const int comp = (int)(bx / g.nx), x = (int)(bx % g.nx);
  for (int i = tid; i < PZ; i += WG) twz[i] = a.tw[g.Px + g.Py + i];
  for (int i = tid; i < PY; i += WG) twy[i] = a.tw[g.Px + i];
  const cx<T>* src = a.B + ((size_t)comp * g.nx + x) * PY * PZ;
  // y inverse of C kz columns at a time: conj(FFT(conj B)); its conjugate is
  // what the z inverse takes in, so the rows are kept unconjugated
  cx<T> nxt[EPT2];
  auto load = [&](int k0) {
#pragma unroll
    for (int i = 0; i < EPT2; i++) {
      const int f = tid + i * WG;
      const int c = f % C, ky = f / C;
      nxt[i] = (E2 % WG == 0 || f < E2) ? conj(src[(size_t)ky * PZ + k0 + c]) : cx<T>{(T)0, (T)0};
    }
  };
  load(0);
  for (int k0 = 0; k0 < PZ; k0 += C) {
#pragma unroll
    for (int i = 0; i < EPT2; i++) {
      const int f = tid + i * WG;
      if (E2 % WG == 0 || f < E2) ys[(f % C) * LS + f / C] = nxt[i];
    }
    if (k0 + C < PZ) load(k0 + C);
    __syncthreads();
    fft_lines<T, PY, C, WG, 1, LS, true>(ys, twy, tid);
#pragma unroll
    for (int i = 0; i < EPT2; i++) {
      const int f = tid + i * WG;
      if (E2 % WG == 0 || f < E2) {
        const int c = f % C, y = f / C;
        if (y < g.ny) zs[y * PZ + k0 + c] = ys[c * LS + y];
      }
    }
    __syncthreads();
  }
  // rows y >= ny are not written and not read back: they only feed lines
  // whose results are discarded
  fft_lines<T, PZ, NY, WG, 1, PZ, false>(zs, twz, tid);
  // gather + contrast: y = eps z - (eps - 1) dst   (vie_fft.c:261-262)
  const size_t nbox = g.nbox();
  const cx<T> one = {(T)1, (T)0}, sf = a.self[comp];
  const cx<T>* xin = a.xin + a.di.off();
  cx<T>* yout = a.yout + a.di.off();
  for (int f = tid; f < g.ny * g.nz; f += WG) {
    const int y = f / g.nz, z = f % g.nz;
    const size_t cell = ((size_t)x * g.ny + y) * g.nz + z;
    const cx<T> e = a.eps[cell];
    cx<T> v = xin[comp * nbox + cell];
    if (a.precond) v = jacobi_scale(v, e, sf);
    const cx<T> d = conj(zs[y * PZ + g.zoff + z]);
    yout[comp * nbox + cell] = e * v - (e - one) * d;
  }
}

template <typename T, int PZ, int PY>
__global__ __launch_bounds__((FusedShape<T, PZ, PY>::WG)) void k_zyfwd(OpArgs<T> a) {
  zy_fwd_body<T, PZ, PY>(a, blockIdx.x);
}
template <typename T, int PY, int PZ>
__global__ __launch_bounds__((FusedShape<T, PZ, PY>::WG)) void k_yzinv(OpArgs<T> a) {
  yz_inv_body<T, PY, PZ>(a, blockIdx.x);
}
template <typename T, int PZ, int PY>
__global__ __launch_bounds__((FusedShape<T, PZ, PY>::WG)) void k_zyfwd_mj(const OpArgs<T>* __restrict__ ops,
                                                                const int* __restrict__ list,
                                                                const unsigned* __restrict__ pre, int count) {
  const int i = mj_job(pre, count, blockIdx.x), q = list[i];
  zy_fwd_body<T, PZ, PY>(op_args_global(ops[q]), blockIdx.x - pre[i]);
}
template <typename T, int PY, int PZ>
__global__ __launch_bounds__((FusedShape<T, PZ, PY>::WG)) void k_yzinv_mj(const OpArgs<T>* __restrict__ ops,
                                                                const int* __restrict__ list,
                                                                const unsigned* __restrict__ pre, int count) {
  const int i = mj_job(pre, count, blockIdx.x), q = list[i];
  yz_inv_body<T, PY, PZ>(op_args_global(ops[q]), blockIdx.x - pre[i]);
}

// ---------------------------------------------------------------- pass 3
// Lines are (ky, kz) with kz fastest.  Forward x FFT of each component is
// staged through LDS and parked in registers; the 3x3 multiply runs in
// registers; each product component goes back through LDS for the inverse.
// A workgroup reads LW consecutive kz of each x row (one 128-byte segment in
// FP32 at LW = 16).  The octant spectrum (13 MB at 128^3 in FP32) is read by
// the eight mirror points of each entry from different workgroups; it stays
// resident in the last-level cache, so those re-reads cost no HBM traffic.
// Line (ky Pz + kz) of local line l of workgroup blk.  Unpaired: LW
// consecutive lines.  Paired (kz mirrors): lines l < LW/2 are kz = gz LW/2 + l
// of one ky and lines LW/2 + l their mirrors Pz - kz (kz = 0 pairs with Pz/2),
// so both halves multiply by the same octant spectrum entries, which a wave
// then loads once for both (needs Pz % LW == 0; the same workgroup count).
template <bool PAIR, int LW>
__device__ __forceinline__ long xline(const Geom& g, long blk, int l) {
  if constexpr (!PAIR) {
    return blk * LW + l;
  } else {
    constexpr int H = LW / 2;
    const long groups = g.Pz / LW;
    const long ky = blk / groups;
    int kz = (int)(blk % groups) * H + (l < H ? l : l - H);
    if (l >= H) kz = kz ? g.Pz - kz : g.Pz / 2;
    return ky * g.Pz + kz;
  }
}

// Tile order of the x pass.  Workgroups are dispatched round-robin over the
// eight XCDs, each with its own L2: workgroup b runs on XCD b % 8 in that
// XCD's (b / 8)-th slot.  Giving consecutive slots of one XCD the tiles that
// read the same octant spectrum entries (rows ky and Py - ky; unpaired kz
// tiles t and T - 1 - t, which hold each other's kz mirrors) lets every
// reader after the first hit that XCD's L2 instead of the last-level cache or
// HBM.  Returns the natural tile index (row-major over (ky, tile)); grids the
// order does not cover keep the natural order.
template <bool PAIR, int LW>
__device__ __forceinline__ long xcd_tile(const Geom& g, long b, long nb) {
  const int T = g.Pz / LW;  // tiles (paired: kz groups) per ky row
  if (T < 1 || g.Pz % LW || g.Py < 2 || nb % 8 || nb != (long)g.Py * T) return b;
  const long v = (b % 8) * (nb / 8) + b / 8;
  const int TT = (!PAIR && T >= 2) ? 2 : 1;  // tiles of one row in a group
  const int TP = T / TT;
  const long rest = v / (2 * TT);
  const int inner = (int)(v % (2 * TT));
  const int tp = (int)(rest % TP), p = (int)(rest / TP);
  const int ri = inner / TT, ti = inner % TT;
  // row pairs (0, Py/2) and (p, Py - p): the same |ky| fold, bar the first
  const int row = p == 0 ? (ri ? g.Py / 2 : 0) : (ri ? g.Py - p : p);
  return (long)row * T + (ti ? T - 1 - tp : tp);
}

template <typename T, int PX, int LW, int WG, bool PAIR, bool XCD>
__device__ __forceinline__ void xpass_body(const OpArgs<T>& a, unsigned bx, unsigned nbx) {
  constexpr int E = LW * PX, EPT = (E + WG - 1) / WG;
  __shared__ cx<T> lds[E];
  __shared__ cx<T> tw[PX];
  if (a.di.skipped()) return;
  const int tid = threadIdx.x;
  const Geom& g = a.g;
  const long lines = (long)g.Py * g.Pz;
  const long blk = XCD ? xcd_tile<PAIR, LW>(g, bx, nbx) : (long)bx;
  const size_t plane = (size_t)g.Py * g.Pz;
  for (int i = tid; i < PX; i += WG) tw[i] = a.tw[i];
  cx<T> hold[3][EPT], next[EPT];
  // element f = e * LW + l of component comp (zero beyond the nx physical rows)
#define VIE_XLOAD(comp, i)                                                                   \
  do {                                                                                       \
    const int f_ = tid + (i) * WG, e_ = f_ / LW, l_ = f_ % LW;                               \
    const long line_ = xline<PAIR, LW>(g, blk, l_);                                          \
    next[i] = (E % WG == 0 || f_ < E) && line_ < lines && e_ < g.nx                          \
                  ? (XCD ? ld_stream(&a.B[((size_t)(comp) * g.nx + e_) * plane + line_])      \
                         : a.B[((size_t)(comp) * g.nx + e_) * plane + line_])                 \
                  : cx<T>{(T)0, (T)0};                                                       \
  } while (0)
#pragma unroll
  for (int i = 0; i < EPT; i++) VIE_XLOAD(0, i);
#pragma unroll
  for (int comp = 0; comp < 3; comp++) {
#pragma unroll
    for (int i = 0; i < EPT; i++) {
      const int f = tid + i * WG;
      if (E % WG == 0 || f < E) lds[f] = next[i];  // f = e * LW + l
    }
    __syncthreads();
    if (comp < 2) {  // next component in flight during this transform
#pragma unroll
      for (int i = 0; i < EPT; i++) VIE_XLOAD(comp + 1, i);
    }
    fft_lines<T, PX, LW, WG, LW, 1, true>(lds, tw, tid);
#pragma unroll
    for (int i = 0; i < EPT; i++) {
      int f = tid + i * WG;
      if (E % WG == 0 || f < E) hold[comp][i] = lds[f];
    }
    __syncthreads();
  }
#undef VIE_XLOAD
  // spectral multiply: dst_i = sum_j T^_ij src_j   (vie_fft.c:253-256)
  const int Hx = g.Px / 2 + 1, Hy = g.Py / 2 + 1, Hz = g.Pz / 2 + 1;
  const size_t oct = (size_t)Hx * Hy * Hz;
  // real ground: + R^(k) X(kx, ky, -kz), the reflected tensor being a
  // correlation in z (the z-flipped source's transform is X at -kz).  The
  // line of -kz is in this group (paired halves, or whole ky rows); its
  // element of the same kx sits in the lane l' - l away.
  int mlane = 0;
  if (g.ground) {
    const int l = tid % LW;
    int lm;
    if constexpr (PAIR) {
      constexpr int H = LW / 2;
      lm = ((blk % (g.Pz / LW)) == 0 && (l % H) == 0) ? l : (l ^ H);
    } else {
      const int kz = l % g.Pz;
      lm = l - kz + ((g.Pz - kz) & (g.Pz - 1));
    }
    mlane = (tid - l + lm) & 63;
  }
  const size_t rstride = (size_t)Hx * Hy * g.Pz;
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    int f = tid + i * WG;
    cx<T> m0 = {(T)0, (T)0}, m1 = m0, m2 = m0;
    if (g.ground) {  // uniform branch: every lane takes part in the shuffles
      m0 = cx<T>{__shfl(hold[0][i].re, mlane), __shfl(hold[0][i].im, mlane)};
      m1 = cx<T>{__shfl(hold[1][i].re, mlane), __shfl(hold[1][i].im, mlane)};
      m2 = cx<T>{__shfl(hold[2][i].re, mlane), __shfl(hold[2][i].im, mlane)};
    }
    if (E % WG == 0 || f < E) {
      int kx = f / LW, l = f % LW;
      long line = xline<PAIR, LW>(g, blk, l);
      if (line < lines) {
        int ky = (int)(line / g.Pz), kz = (int)(line % g.Pz);
        int ox = kx <= g.Px / 2 ? kx : g.Px - kx;
        int oy = ky <= g.Py / 2 ? ky : g.Py - ky;
        int oz = kz <= g.Pz / 2 ? kz : g.Pz - kz;
        T sx = kx <= g.Px / 2 ? (T)1 : (T)-1;
        T sy = ky <= g.Py / 2 ? (T)1 : (T)-1;
        T sz = kz <= g.Pz / 2 ? (T)1 : (T)-1;
        size_t o = ((size_t)ox * Hy + oy) * Hz + oz;
        cx<T> txx = a.spec[o], tyy = a.spec[oct + o], tzz = a.spec[2 * oct + o];
        cx<T> txy = (sx * sy) * a.spec[3 * oct + o];
        cx<T> txz = (sx * sz) * a.spec[4 * oct + o];
        cx<T> tyz = (sy * sz) * a.spec[5 * oct + o];
        cx<T> s0 = hold[0][i], s1 = hold[1][i], s2 = hold[2][i];
        cx<T> d0 = txx * s0 + txy * s1 + txz * s2;
        cx<T> d1 = txy * s0 + tyy * s1 + tyz * s2;
        cx<T> d2 = txz * s0 + tyz * s1 + tzz * s2;
        if (g.ground) {
          // reflected tensor: xy odd in x and y, xz odd in x, yz odd in y;
          // zx = -xz, zy = -yz
          const size_t r = ((size_t)ox * Hy + oy) * g.Pz + kz;
          cx<T> rxx = a.rspec[r], ryy = a.rspec[rstride + r], rzz = a.rspec[2 * rstride + r];
          cx<T> rxy = (sx * sy) * a.rspec[3 * rstride + r];
          cx<T> rxz = sx * a.rspec[4 * rstride + r];
          cx<T> ryz = sy * a.rspec[5 * rstride + r];
          d0 = d0 + rxx * m0 + rxy * m1 + rxz * m2;
          d1 = d1 + rxy * m0 + ryy * m1 + ryz * m2;
          d2 = d2 - rxz * m0 - ryz * m1 + rzz * m2;
        }
        hold[0][i] = d0;
        hold[1][i] = d1;
        hold[2][i] = d2;
      }
    }
  }
#pragma unroll
  for (int comp = 0; comp < 3; comp++) {
#pragma unroll
    for (int i = 0; i < EPT; i++) {
      int f = tid + i * WG;
      if (E % WG == 0 || f < E) lds[f] = conj(hold[comp][i]);
    }
    __syncthreads();
    fft_lines<T, PX, LW, WG, LW, 1, true>(lds, tw, tid);
    for (int f = tid; f < E; f += WG) {
      int e = f / LW, l = f % LW;
      long line = xline<PAIR, LW>(g, blk, l);
      if (line < lines && e < g.nx) a.B[((size_t)comp * g.nx + e) * plane + line] = conj(lds[e * LW + l]);
    }
    __syncthreads();
  }
}

template <typename T, int PX, int LW, int WG, bool PAIR = false, bool XCD = false>
__global__ __launch_bounds__(WG) void k_xpass(OpArgs<T> a) {
  xpass_body<T, PX, LW, WG, PAIR, XCD>(a, blockIdx.x, gridDim.x);
}
template <typename T, int PX, int LW, int WG, bool PAIR, bool XCD>
__global__ __launch_bounds__(WG) void k_xpass_mj(const OpArgs<T>* __restrict__ ops, const unsigned* __restrict__ blocks,
                                                const int* __restrict__ list,
                                                 const unsigned* __restrict__ pre, int count) {
  const int i = mj_job(pre, count, blockIdx.x), q = list[i];
  xpass_body<T, PX, LW, WG, PAIR, XCD>(op_args_global(ops[q]), blockIdx.x - pre[i], blocks[5 * q + 2]);
}


// ---------------------------------------------------------------- pass 3, registers
// The x pass for P = 256 in FP32 as a two-level 16 x 16 transform held in
// registers.  Thread (l, c) of a workgroup of 16 LW threads serves line l:
//   forward, x index n = c + 16 n1, frequency k = k1 + 16 k2:
//     a 16-point DFT over n1 (inputs n1 >= 8 are zero: n < P/2 >= nx),
//     twiddle W256^(c k1), one LDS transpose, a 16-point DFT over n2 that
//     leaves the thread holding X[c + 16 k2] for every k2;
//   the 3x3 spectral multiply on those registers (the same entries as k_xpass);
//   inverse, n = m2 + 16 m1: a 16-point DFT over k2 in the same registers,
//     twiddle, one LDS transpose, a 16-point DFT over k1 forming only the
//     outputs m1 < 8 (the nx physical rows live below P/2).
// So each component crosses LDS once per direction (one transpose, two
// barriers) instead of once per radix stage.  Inverse transforms use
// conj(DFT(conj x)); the 1/P scaling is in the spectrum.
// WPE waves per SIMD: the register budget (168 at three, 256 at two) for the
// three held components (96 registers) and one transform's temporaries.
template <int LW, bool PAIR, bool XCD, int WPE>
__global__ __launch_bounds__(16 * LW) __attribute__((amdgpu_waves_per_eu(WPE)))
void k_xpass_r16(OpArgs<float> a) {
  using T = float;
  constexpr int PX = 256, WG = 16 * LW;
  __shared__ cx<T> lds[LW * PX];
  __shared__ cx<T> tw[16][16];  // tw[c][k] = W256^(c k): a thread's twiddles are one row
  if (a.di.skipped()) return;
  const int tid = threadIdx.x, l = tid % LW, c = tid / LW;
  const Geom& g = a.g;
  const long lines = (long)g.Py * g.Pz;
  const long blk = XCD ? xcd_tile<PAIR, LW>(g, blockIdx.x, gridDim.x) : (long)blockIdx.x;
  // element offsets fit 32 bits (checked on the host), so loads take a
  // scalar base and one offset register
  const unsigned plane = (unsigned)g.Py * g.Pz;
  if (tid < 256) tw[tid / 16][tid % 16] = a.tw[(tid / 16) * (tid % 16)];
  const long line = xline<PAIR, LW>(g, blk, l);
  const bool live = line < lines;
  const unsigned uline = live ? (unsigned)line : 0u;
  cx<T> hold[3][16], nxt[8];
  // loads are unconditional (from element 0 where there is no input) and
  // selected afterwards: a branch per load would force spills around it
  auto load = [&](int comp) {
#pragma unroll
    for (int n1 = 0; n1 < 8; n1++) {
      const int n = c + 16 * n1;
      const bool ok = live && n < g.nx;
      const cx<T>* src = a.B + (ok ? ((unsigned)comp * g.nx + n) * plane + uline : 0u);
      const cx<T> val = XCD ? ld_stream(src) : *src;
      nxt[n1] = ok ? val : cx<T>{(T)0, (T)0};
    }
  };
  load(0);
#pragma unroll
  for (int comp = 0; comp < 3; comp++) {
    cx<T> v[16];
#pragma unroll
    for (int n1 = 0; n1 < 8; n1++) v[n1] = nxt[n1];
#pragma unroll
    for (int n1 = 8; n1 < 16; n1++) v[n1] = cx<T>{(T)0, (T)0};
    if (comp < 2) load(comp + 1);  // next component in flight during this one
    if (comp == 0) __syncthreads();  // twiddle table
    dft16<T, true, false>(v);
#pragma unroll
    for (int k1 = 1; k1 < 16; k1++) v[k1] = v[k1] * tw[c][k1];
#pragma unroll
    for (int k1 = 0; k1 < 16; k1++) lds[(k1 * 16 + c) * LW + l] = v[k1];
    __syncthreads();
#pragma unroll
    for (int n2 = 0; n2 < 16; n2++) hold[comp][n2] = lds[(c * 16 + n2) * LW + l];
    __syncthreads();
    dft16<T, false, false>(hold[comp]);  // hold[comp][k2] = X[c + 16 k2]
  }
  // spectral multiply: dst_i = sum_j T^_ij src_j   (vie_fft.c:253-256)
  const int Hx = g.Px / 2 + 1, Hy = g.Py / 2 + 1, Hz = g.Pz / 2 + 1;
  const unsigned oct = (unsigned)Hx * Hy * Hz;
  int mlane = 0;
  if (g.ground) {  // the -kz line of this lane's line (see k_xpass)
    int lm;
    if constexpr (PAIR) {
      constexpr int H = LW / 2;
      lm = ((blk % (g.Pz / LW)) == 0 && (l % H) == 0) ? l : (l ^ H);
    } else {
      const int kz = l % g.Pz;
      lm = l - kz + ((g.Pz - kz) & (g.Pz - 1));
    }
    mlane = (tid - l + lm) & 63;
  }
  const unsigned rstride = (unsigned)Hx * Hy * g.Pz;
  // lanes past the last line compute on line 0 and store nothing
  const int ky = live ? (int)(line / g.Pz) : 0, kz = live ? (int)(line % g.Pz) : 0;
  const int oy = ky <= g.Py / 2 ? ky : g.Py - ky, oz = kz <= g.Pz / 2 ? kz : g.Pz - kz;
  const T sy = ky <= g.Py / 2 ? (T)1 : (T)-1, sz = kz <= g.Pz / 2 ? (T)1 : (T)-1;
#pragma unroll
  for (int k2 = 0; k2 < 16; k2++) {
    const int kx = c + 16 * k2;
    const int ox = kx <= PX / 2 ? kx : PX - kx;
    const T sx = kx <= PX / 2 ? (T)1 : (T)-1;
    const unsigned o = ((unsigned)ox * Hy + oy) * Hz + oz;
    // every load of the step is issued before any of its arithmetic, so the
    // step waits for memory once (the reflected loads of a ground are not
    // held back behind the octant arithmetic)
    cx<T> t[6], rr[6] = {};
#pragma unroll
    for (int q = 0; q < 6; q++) t[q] = a.spec[q * oct + o];
    cx<T> m0 = {(T)0, (T)0}, m1 = m0, m2 = m0;
    if (g.ground) {  // uniform branch: every lane takes part in the shuffles
      const unsigned r = ((unsigned)ox * Hy + oy) * g.Pz + kz;
#pragma unroll
      for (int q = 0; q < 6; q++) rr[q] = a.rspec[q * rstride + r];
      m0 = cx<T>{__shfl(hold[0][k2].re, mlane), __shfl(hold[0][k2].im, mlane)};
      m1 = cx<T>{__shfl(hold[1][k2].re, mlane), __shfl(hold[1][k2].im, mlane)};
      m2 = cx<T>{__shfl(hold[2][k2].re, mlane), __shfl(hold[2][k2].im, mlane)};
    }
    {
      cx<T> txy = (sx * sy) * t[3], txz = (sx * sz) * t[4], tyz = (sy * sz) * t[5];
      cx<T> s0 = hold[0][k2], s1 = hold[1][k2], s2 = hold[2][k2];
      cx<T> d0 = t[0] * s0 + txy * s1 + txz * s2;
      cx<T> d1 = txy * s0 + t[1] * s1 + tyz * s2;
      cx<T> d2 = txz * s0 + tyz * s1 + t[2] * s2;
      if (g.ground) {
        cx<T> rxy = (sx * sy) * rr[3], rxz = sx * rr[4], ryz = sy * rr[5];
        d0 = d0 + rr[0] * m0 + rxy * m1 + rxz * m2;
        d1 = d1 + rxy * m0 + rr[1] * m1 + ryz * m2;
        d2 = d2 - rxz * m0 - ryz * m1 + rr[2] * m2;
      }
      hold[0][k2] = d0;
      hold[1][k2] = d1;
      hold[2][k2] = d2;
    }
    // keep each k2's spectrum loads within its own step: hoisting them all
    // would need far more registers than the held components leave
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int comp = 0; comp < 3; comp++) {
    cx<T>* v = hold[comp];
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) v[k2] = conj(v[k2]);
    dft16<T, false, false>(v);  // conj Z[m2] for this k1 = c
#pragma unroll
    for (int m2 = 1; m2 < 16; m2++) v[m2] = v[m2] * tw[c][m2];
#pragma unroll
    for (int m2 = 0; m2 < 16; m2++) lds[(m2 * 16 + c) * LW + l] = v[m2];
    __syncthreads();
    cx<T> u[16];
#pragma unroll
    for (int k1 = 0; k1 < 16; k1++) u[k1] = lds[(c * 16 + k1) * LW + l];
    __syncthreads();
    dft16<T, false, true>(u);  // conj x[c + 16 m1], m1 < 8
#pragma unroll
    for (int m1 = 0; m1 < 8; m1++) {
      const int n = c + 16 * m1;
      if (live && n < g.nx) a.B[((unsigned)comp * g.nx + n) * plane + uline] = conj(u[m1]);
    }
  }
}

// ---------------------------------------------------------------- pass 5
template <typename T, int PZ, int LW, int WG>
__device__ __forceinline__ void zinv_body(const OpArgs<T>& a, unsigned bx, unsigned nbx) {
  constexpr int E = LW * PZ, EPT = (E + WG - 1) / WG;
  __shared__ cx<T> lds[E];
  __shared__ cx<T> tw[PZ];
  if (a.di.skipped()) return;
  const int tid = threadIdx.x;
  const Geom& g = a.g;
  const long lines = (long)g.nx * g.ny;
  const long line0 = (long)bx * LW;
  const cx<T>* twz = a.tw + g.Px + g.Py;
  for (int i = tid; i < PZ; i += WG) tw[i] = twz[i];
  const size_t nbox = g.nbox();
  const cx<T> one = {(T)1, (T)0};
  const cx<T>* xin = a.xin + a.di.off();
  cx<T>* yout = a.yout + a.di.off();
  cx<T> in[3][EPT];
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int f = tid + i * WG;
    const long line = line0 + f / PZ;
    const bool ok = (E % WG == 0 || f < E) && line < lines;
#pragma unroll
    for (int comp = 0; comp < 3; comp++)
      in[comp][i] = ok ? conj(a.A[((size_t)comp * lines + line) * PZ + f % PZ]) : cx<T>{(T)0, (T)0};
  }
#pragma unroll
  for (int comp = 0; comp < 3; comp++) {
#pragma unroll
    for (int i = 0; i < EPT; i++) {
      const int f = tid + i * WG;
      if (E % WG == 0 || f < E) lds[f] = in[comp][i];
    }
    __syncthreads();
    fft_lines<T, PZ, LW, WG, 1, PZ, false>(lds, tw, tid);
    // gather + contrast: y = eps z - (eps - 1) dst   (vie_fft.c:261-262)
    for (int f = tid; f < LW * g.nz; f += WG) {
      int l = f / g.nz, z = f % g.nz;
      long line = line0 + l;
      if (line < lines) {
        size_t cell = (size_t)line * g.nz + z;
        cx<T> e = a.eps[cell];
        cx<T> v = xin[comp * nbox + cell];
        if (a.precond) v = jacobi_scale(v, e, a.self[comp]);
        cx<T> d = conj(lds[l * PZ + g.zoff + z]);
        yout[comp * nbox + cell] = e * v - (e - one) * d;
      }
    }
    __syncthreads();
  }
}

template <typename T, int PZ, int LW, int WG>
__global__ __launch_bounds__(WG) void k_zinv(OpArgs<T> a) {
  zinv_body<T, PZ, LW, WG>(a, blockIdx.x, gridDim.x);
}
template <typename T, int PZ, int LW, int WG>
__global__ __launch_bounds__(WG) void k_zinv_mj(const OpArgs<T>* __restrict__ ops, const unsigned* __restrict__ blocks,
                                                const int* __restrict__ list,
                                                const unsigned* __restrict__ pre, int count) {
  const int i = mj_job(pre, count, blockIdx.x), q = list[i];
  zinv_body<T, PZ, LW, WG>(op_args_global(ops[q]), blockIdx.x - pre[i], blocks[5 * q + 4]);
}


// ---------------------------------------------------------------- launch
// Launch shapes of a pass over lines of length P.  Variant 0 stages the
// widest tile the LDS budget allows (L0 lines, 256 threads); variants 1 and 2
// keep the elements per thread and shrink the workgroup to 128 and 64
// threads (more, smaller workgroups: better fill for small grids); variant 3
// halves the elements per thread at 256 threads (fewer live registers).
// A per-geometry timing tuner picks one variant per pass.
__host__ __device__ constexpr int tile_l0(int E, int P, int cap) { return E / P < cap ? E / P : cap; }
__host__ __device__ constexpr int tile_lw(int l0, int v) {
  return v == 0 ? l0 : v == 2 ? (l0 / 4 > 0 ? l0 / 4 : 1) : (l0 / 2 > 0 ? l0 / 2 : 1);
}
__host__ __device__ constexpr int tile_wg(int v) { return v == 1 ? 128 : v == 2 ? 64 : 256; }
constexpr int kZCap = 8, kYCap = 16, kXCap = 32;

template <typename T, int P> struct PassLauncher {
  static constexpr int E = PassShape<T>::E;
  template <int V, bool PAIR = false, bool XCD = false>
  static void run(int pass, const OpArgs<T>& a, hipStream_t s) {
    constexpr int WG = tile_wg(V);
    constexpr int LZ = tile_lw(tile_l0(E, P, kZCap), V);
    constexpr int LY = tile_lw(tile_l0(E, P, kYCap), V);
    // the x pass holds three components per line: twice the elements and
    // threads of the other passes, for wider HBM segments per line group
    constexpr int LX = tile_lw(tile_l0(2 * E, P, kXCap), V);
    const long zl = (long)a.g.nx * a.g.ny, yl = 3L * a.g.nx * a.g.Pz, xl = (long)a.g.Py * a.g.Pz;
    switch (pass) {
      case 0: hipLaunchKernelGGL((k_zfwd<T, P, LZ, WG>), dim3((zl + LZ - 1) / LZ), dim3(WG), 0, s, a); break;
      case 1: hipLaunchKernelGGL((k_ypass<T, P, LY, WG, false>), dim3((yl + LY - 1) / LY), dim3(WG), 0, s, a); break;
      case 2: hipLaunchKernelGGL((k_xpass<T, P, LX, 2 * WG, PAIR, XCD>), dim3((xl + LX - 1) / LX), dim3(2 * WG), 0, s, a); break;
      case 3: hipLaunchKernelGGL((k_ypass<T, P, LY, WG, true>), dim3((yl + LY - 1) / LY), dim3(WG), 0, s, a); break;
      default: hipLaunchKernelGGL((k_zinv<T, P, LZ, WG>), dim3((zl + LZ - 1) / LZ), dim3(WG), 0, s, a); break;
    }
  }
  // x-pass line group of variant v (for the paired mapping's Pz % LX test)
  static constexpr int x_lines(int v) { return tile_lw(tile_l0(2 * E, P, kXCap), v & 3); }
  template <int V> static void run_x(int pass, int v, const OpArgs<T>& a, hipStream_t s) {
    switch (pass == 2 ? (v >> 2) & 3 : 0) {
      case 1: run<V, true>(pass, a, s); return;
      case 2: run<V, false, true>(pass, a, s); return;
      case 3: run<V, true, true>(pass, a, s); return;
      default: run<V>(pass, a, s); return;
    }
  }
  template <int LW, bool PAIR, bool XCD> static void run_r16(int v, const OpArgs<T>& a, hipStream_t s) {
    if constexpr (std::is_same<T, float>::value && P == 256) {
      const long xl = (long)a.g.Py * a.g.Pz;
      if (LW == 16 && !(v & 2))
        hipLaunchKernelGGL((k_xpass_r16<16, PAIR, XCD, 3>), dim3((xl + 15) / 16), dim3(256), 0, s, a);
      else
        hipLaunchKernelGGL((k_xpass_r16<LW, PAIR, XCD, 2>), dim3((xl + LW - 1) / LW), dim3(16 * LW), 0, s, a);
    }
  }
  // Multi-job launches at the job's own tuned variant (the LDS kernels:
  // shapes 0-3, and for the x pass pairing and the XCD order), for lines of
  // at most kMjMaxP points; the register-resident kernels (P = 256) are not
  // offered.  Same workgroup count per job as its single-job launch.
  template <int V, bool PAIR = false, bool XCD = false> static unsigned blocks_v(int pass, const Geom& g) {
    constexpr int LZ = tile_lw(tile_l0(E, P, kZCap), V), LY = tile_lw(tile_l0(E, P, kYCap), V);
    constexpr int LX = tile_lw(tile_l0(2 * E, P, kXCap), V);
    const long zl = (long)g.nx * g.ny, yl = 3L * g.nx * g.Pz, xl = (long)g.Py * g.Pz;
    switch (pass) {
      case 0: case 4: return (unsigned)((zl + LZ - 1) / LZ);
      case 1: case 3: return (unsigned)((yl + LY - 1) / LY);
      default: return (unsigned)((xl + LX - 1) / LX);
    }
  }
  static unsigned blocks_mj(int pass, int v, const Geom& g) {
    if (v == kFuseSkip) return 0;
    if (v >= kFuseBase) return 3u * (unsigned)g.nx;
    switch (v & 3) {
      case 1: return blocks_v<1>(pass, g);
      case 2: return blocks_v<2>(pass, g);
      case 3: return blocks_v<3>(pass, g);
      default: return blocks_v<0>(pass, g);
    }
  }
  template <int V, bool PAIR, bool XCD>
  static void run_mj(const MjClass& c, const OpArgs<T>* ops, const unsigned* blocks, const int* lists,
                     const unsigned* pres, hipStream_t s) {
    if constexpr (P <= kMjMaxP) {
      constexpr int WG = tile_wg(V);
      constexpr int LZ = tile_lw(tile_l0(E, P, kZCap), V), LY = tile_lw(tile_l0(E, P, kYCap), V);
      constexpr int LX = tile_lw(tile_l0(2 * E, P, kXCap), V);
      const dim3 grid(c.total_blocks);
      const int* list = lists + c.offset;
      const unsigned* pre = pres + c.pre;
      const int n = c.count;
      switch (c.pass) {
        case 0: hipLaunchKernelGGL((k_zfwd_mj<T, P, LZ, WG>), grid, dim3(WG), 0, s, ops, blocks, list, pre, n); break;
        case 1:
          hipLaunchKernelGGL((k_ypass_mj<T, P, LY, WG, false>), grid, dim3(WG), 0, s, ops, blocks, list, pre, n);
          break;
        case 2:
          hipLaunchKernelGGL((k_xpass_mj<T, P, LX, 2 * WG, PAIR, XCD>), grid, dim3(2 * WG), 0, s, ops, blocks, list,
                             pre, n);
          break;
        case 3:
          hipLaunchKernelGGL((k_ypass_mj<T, P, LY, WG, true>), grid, dim3(WG), 0, s, ops, blocks, list, pre, n);
          break;
        default: hipLaunchKernelGGL((k_zinv_mj<T, P, LZ, WG>), grid, dim3(WG), 0, s, ops, blocks, list, pre, n); break;
      }
    }
  }
  template <int V>
  static void run_mj_x(const MjClass& c, const OpArgs<T>* ops, const unsigned* blocks, const int* lists,
                       const unsigned* pres, hipStream_t s) {
    switch (c.pass == 2 ? (c.variant >> 2) & 3 : 0) {
      case 1: run_mj<V, true, false>(c, ops, blocks, lists, pres, s); return;
      case 2: run_mj<V, false, true>(c, ops, blocks, lists, pres, s); return;
      case 3: run_mj<V, true, true>(c, ops, blocks, lists, pres, s); return;
      default: run_mj<V, false, false>(c, ops, blocks, lists, pres, s); return;
    }
  }
  static void launch_mj(const MjClass& c, const OpArgs<T>* ops, const unsigned* blocks, const int* lists,
                        const unsigned* pres, hipStream_t s) {
    if (c.variant >= kFuseBase) {
      (void)fused_launch(c.pass, c.variant, nullptr, &c, ops, lists, pres, s);
      return;
    }
    switch (c.variant & 3) {
      case 1: run_mj_x<1>(c, ops, blocks, lists, pres, s); return;
      case 2: run_mj_x<2>(c, ops, blocks, lists, pres, s); return;
      case 3: run_mj_x<3>(c, ops, blocks, lists, pres, s); return;
      default: run_mj_x<0>(c, ops, blocks, lists, pres, s); return;
    }
  }
  // x-pass variant bits: 0-1 the shape, 2 kz-mirror pairing, 3 the
  // XCD-grouped tile order (with streaming loads of the pencils), 4 the
  // register-resident 16 x 16 kernel (FP32, P = 256; bit 0 then picks 32
  // lines per workgroup instead of 16, bit 1 two waves per SIMD instead of
  // three at 16 lines)
  template <int LW> static void run_y16(int pass, const OpArgs<T>& a, hipStream_t s) {
    if constexpr (P == 256) {
      const long yl = 3L * a.g.nx * a.g.Pz;
      if (pass == 1)
        hipLaunchKernelGGL((k_ypass_r16<T, LW, false>), dim3((yl + LW - 1) / LW), dim3(16 * LW), 0, s, a);
      else
        hipLaunchKernelGGL((k_ypass_r16<T, LW, true>), dim3((yl + LW - 1) / LW), dim3(16 * LW), 0, s, a);
    }
  }
  // y-pass variants 4 and 5: the register-resident 16 x 16 kernel (P = 256)
  // at 16 or 32 lines per workgroup in FP32, 8 or 16 in FP64
  static constexpr int kY16 = std::is_same<T, float>::value ? 16 : 8;
  // Fused z+y passes (pass 0 forward, pass 4 inverse; P is Pz), one
  // workgroup per (component, x plane), the y extent from the variant code.
  template <int PY> static void run_fused(int pass, const OpArgs<T>& a, hipStream_t s) {
    if constexpr (P == 32 || P == 64 || P == 128) {
      const dim3 grid(3u * (unsigned)a.g.nx), blk(FusedShape<T, P, PY>::WG);
      if (pass == 0) hipLaunchKernelGGL((k_zyfwd<T, P, PY>), grid, blk, 0, s, a);
      else hipLaunchKernelGGL((k_yzinv<T, PY, P>), grid, blk, 0, s, a);
    }
  }
  template <int PY>
  static void run_fused_mj(const MjClass& c, const OpArgs<T>* ops, const int* lists, const unsigned* pres,
                           hipStream_t s) {
    if constexpr (P == 32 || P == 64 || P == 128) {
      const dim3 grid(c.total_blocks), blk(FusedShape<T, P, PY>::WG);
      const int* list = lists + c.offset;
      const unsigned* pre = pres + c.pre;
      if (c.pass == 0) hipLaunchKernelGGL((k_zyfwd_mj<T, P, PY>), grid, blk, 0, s, ops, list, pre, c.count);
      else hipLaunchKernelGGL((k_yzinv_mj<T, PY, P>), grid, blk, 0, s, ops, list, pre, c.count);
    }
  }
  static bool fused_launch(int pass, int v, const OpArgs<T>* a, const MjClass* c, const OpArgs<T>* ops,
                           const int* lists, const unsigned* pres, hipStream_t s) {
    if (v == kFuseSkip) return true;  // done by the fused pass 0 or 4
    if (v < kFuseBase + 5 || v > kFuseBase + 7) return false;
    switch (v - kFuseBase) {
      case 5: a ? run_fused<32>(pass, *a, s) : run_fused_mj<32>(*c, ops, lists, pres, s); break;
      case 6: a ? run_fused<64>(pass, *a, s) : run_fused_mj<64>(*c, ops, lists, pres, s); break;
      default: a ? run_fused<128>(pass, *a, s) : run_fused_mj<128>(*c, ops, lists, pres, s); break;
    }
    return true;
  }
  static void launch(int pass, int v, const OpArgs<T>& a, hipStream_t s) {
    if (v >= kFuseBase && fused_launch(pass, v, &a, nullptr, nullptr, nullptr, nullptr, s)) return;
    if ((pass == 1 || pass == 3) && v >= 4) {
      v == 4 ? run_y16<kY16>(pass, a, s) : run_y16<2 * kY16>(pass, a, s);
      return;
    }
    if (pass == 2 && v >= 16) {
      switch (v & 13) {
        case 0: run_r16<16, false, false>(v, a, s); return;
        case 1: run_r16<32, false, false>(v, a, s); return;
        case 4: run_r16<16, true, false>(v, a, s); return;
        case 5: run_r16<32, true, false>(v, a, s); return;
        case 8: run_r16<16, false, true>(v, a, s); return;
        case 9: run_r16<32, false, true>(v, a, s); return;
        case 12: run_r16<16, true, true>(v, a, s); return;
        default: run_r16<32, true, true>(v, a, s); return;
      }
    }
    if constexpr (P >= 16) {
      switch (v & 3) {
        case 1: run_x<1>(pass, v, a, s); return;
        case 2: run_x<2>(pass, v, a, s); return;
        case 3: run_x<3>(pass, v, a, s); return;
        default: run_x<0>(pass, v, a, s); return;
      }
    } else {
      run_x<0>(pass, v, a, s);
    }
  }
};

template <typename T, typename F> static bool dispatch_p(int P, F&& f) {
  switch (P) {
    case 1: f(PassLauncher<T, 1>()); return true;
    case 2: f(PassLauncher<T, 2>()); return true;
    case 4: f(PassLauncher<T, 4>()); return true;
    case 8: f(PassLauncher<T, 8>()); return true;
    case 16: f(PassLauncher<T, 16>()); return true;
    case 32: f(PassLauncher<T, 32>()); return true;
    case 64: f(PassLauncher<T, 64>()); return true;
    case 128: f(PassLauncher<T, 128>()); return true;
    case 256: f(PassLauncher<T, 256>()); return true;
    case 512: f(PassLauncher<T, 512>()); return true;
    case 1024: f(PassLauncher<T, 1024>()); return true;
    default: return false;
  }
}


static int pass_extent(const Geom& g, int pass) { return pass == 2 ? g.Px : (pass == 1 || pass == 3) ? g.Py : g.Pz; }

template <typename T>
int pass_launch(const VolumeView<T>& v, int pass, int variant, const cx<T>* xin, cx<T>* yout, int precond,
                hipStream_t s) {
  OpArgs<T> a = op_args(v, xin, yout, precond);
  bool ok = dispatch_p<T>(pass_extent(v.g, pass), [&](auto L) { L.launch(pass, variant, a, s); });
  return ok && hipGetLastError() == hipSuccess;
}

template <typename T> int pass_variant_count(const Geom& g, int pass) {
  const int P = pass_extent(g, pass);
  if ((pass == 1 || pass == 3) && P == 256) return kPassVariants + 2;
  if (pass == 2) return std::is_same<T, float>::value && P == 256 ? 8 * kPassVariants : 4 * kPassVariants;
  return P >= 16 ? kPassVariants : 1;
}

template <typename T> bool pass_variant_ok(const Geom& g, int pass, int v) {
  const int P = pass_extent(g, pass);
  // fused z+y passes: pass 0 / 4 carry both transforms, 1 / 3 launch nothing
  // (whether the two agree is variants_consistent's to check)
  if (v == kFuseSkip) return (pass == 1 || pass == 3) && fused_ok(g);
  if (v >= kFuseBase) return (pass == 0 || pass == 4) && fused_ok(g) && v == fused_code(g);
  if (v < 0 || v >= pass_variant_count<T>(g, pass)) return false;
  if (pass != 2) return true;
  if (P < 16 && (v & 3) != 0) return false;
  int lx = 0;
  if (v >= 16) {
    if ((v & 3) == 3) return false;  // 32 lines run at two waves only
    // element offsets of the pencils, spectrum and reflected spectrum in 32 bits
    const double big = 3.0 * g.nx * g.Py * g.Pz + 6.0 * (g.Px / 2 + 1) * (g.Py / 2 + 1) * g.Pz;
    if (big * 8 >= 4294967296.0) return false;
    lx = (v & 1) ? 32 : 16;
  } else {
    dispatch_p<T>(P, [&](auto L) { lx = L.x_lines(v & 3); });
  }
  const bool paired = (v & 4) != 0;
  if (paired && !(lx >= 2 && g.Pz >= lx && g.Pz % lx == 0)) return false;
  if (g.ground && !paired && lx % g.Pz != 0) return false;
  return true;
}

template <typename T> unsigned pass_blocks_mj(const Geom& g, int pass, int variant) {
  unsigned nb = 0;
  dispatch_p<T>(pass_extent(g, pass), [&](auto L) { nb = L.blocks_mj(pass, variant, g); });
  return nb;
}

template <typename T> bool pass_mj_ok(const Geom& g, int pass, int variant) {
  if (variant >= kFuseBase) return pass_variant_ok<T>(g, pass, variant);
  return pass_extent(g, pass) <= kMjMaxP && pass_variant_ok<T>(g, pass, variant) &&
         variant < (pass == 2 ? 4 * kPassVariants : kPassVariants);
}



template <typename T>
int volume_apply_mj(const OpArgs<T>* ops, const unsigned* blocks, const int* lists, const unsigned* pres,
                    const MjClass* classes, int nclasses, hipStream_t s) {
  bool ok = true;
  // classes are in pass order, so every job's passes run in sequence
  for (int i = 0; i < nclasses && ok; i++)
    if (classes[i].count > 0 && classes[i].total_blocks > 0)
      ok = dispatch_p<T>(classes[i].P, [&](auto L) { L.launch_mj(classes[i], ops, blocks, lists, pres, s); });
  return ok && hipGetLastError() == hipSuccess;
}

template <typename T>
int volume_apply(const VolumeView<T>& v, const cx<T>* xin, cx<T>* yout, int precond, hipStream_t s, hipEvent_t* ev,
                 DevIndex di) {
  OpArgs<T> a = op_args(v, xin, yout, precond, di);
  // A unit axis pads to P = 1, whose DFT is the identity (zero stages).
  bool ok = true;
  for (int pass = 0; pass < 5 && ok; pass++) {
    if (ev) (void)hipEventRecord(ev[pass], s);
    ok = dispatch_p<T>(pass_extent(v.g, pass), [&](auto L) { L.launch(pass, v.var[pass], a, s); });
  }
  if (ev) (void)hipEventRecord(ev[5], s);
  return ok && hipGetLastError() == hipSuccess;
}

}  // namespace vie
