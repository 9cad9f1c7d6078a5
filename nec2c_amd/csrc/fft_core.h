// LDS-staged Stockham pencil FFT for gfx950 (64-wide wavefronts).
//
// A workgroup of WG threads transforms LW lines of length N held in LDS
// (LW * N == E elements, E/WG elements per thread).  Each stage is a
// radix-8/4/2 Stockham step: a thread gathers the R inputs of its butterflies
// into registers, applies twiddles exp(-2 pi i k r / (Ns R)) from an LDS table
// and the radix-R DFT, barrier, then scatters the R outputs back in place.
// The result is in natural order, so no bit reversal pass is needed.
//
// Forward transform only: X(k) = sum_n x(n) exp(-2 pi i k n / N), the sign
// convention of the reference's roots (vie_fft.c:433-436).  Inverse
// transforms use conj(FFT(conj(x))), i.e. the caller conjugates on load and
// store (no 1/N: the operator folds 1/(Px Py Pz) into the kernel spectrum).
#pragma once
#include "vie_common.h"

namespace vie {

template <typename T> __device__ __forceinline__ void dft2(cx<T>* v) {
  cx<T> a = v[0], b = v[1];
  v[0] = a + b;
  v[1] = a - b;
}

template <typename T> __device__ __forceinline__ void dft4(cx<T>* v) {
  cx<T> b0 = v[0] + v[2], b2 = v[0] - v[2];
  cx<T> b1 = v[1] + v[3], b3 = mul_mi(v[1] - v[3]);
  v[0] = b0 + b1;
  v[2] = b0 - b1;
  v[1] = b2 + b3;
  v[3] = b2 - b3;
}

// Radix-8 as one radix-2 split (with w8 twiddles) into two radix-4 DFTs.
template <typename T> __device__ __forceinline__ void dft8(cx<T>* v) {
  const T s = (T)0.70710678118654752440084436210484904;
  cx<T> e[4], o[4];
#pragma unroll
  for (int n = 0; n < 4; n++) {
    e[n] = v[n] + v[n + 4];
    o[n] = v[n] - v[n + 4];
  }
  // o[n] *= exp(-2 pi i n / 8)
  o[1] = cx<T>{s * (o[1].re + o[1].im), s * (o[1].im - o[1].re)};
  o[2] = mul_mi(o[2]);
  o[3] = cx<T>{s * (o[3].im - o[3].re), -s * (o[3].re + o[3].im)};
  dft4(e);
  dft4(o);
#pragma unroll
  for (int m = 0; m < 4; m++) {
    v[2 * m] = e[m];
    v[2 * m + 1] = o[m];
  }
}

template <int R, typename T> __device__ __forceinline__ void dftR(cx<T>* v) {
  if constexpr (R == 2) dft2(v);
  else if constexpr (R == 4) dft4(v);
  else dft8(v);
}

// Radix of stage s for a length-N transform: 8s first, then a 4 or 2 tail
// (256 = 8*8*4, 128 = 8*4*4, 64 = 8*8, 32 = 8*4, 16 = 4*4, 8, 4, 2).
__host__ __device__ constexpr int log2i(int n) { return n <= 1 ? 0 : 1 + log2i(n / 2); }
__host__ __device__ constexpr int stage_count(int N) {
  return N <= 1 ? 0 : log2i(N) <= 3 ? 1 : (log2i(N) % 3 == 0 ? log2i(N) / 3 : log2i(N) / 3 + 1);
}
__host__ __device__ constexpr int stage_radix(int N, int s) {
  // total bits L = log2 N; q = L/3 radix-8 stages, remainder r = L%3:
  // r == 0: all 8; r == 1: (q-1) x 8 then 4, 4; r == 2: q x 8 then 4.
  return log2i(N) <= 3 ? N
       : (log2i(N) % 3 == 0) ? 8
       : (log2i(N) % 3 == 2) ? (s < log2i(N) / 3 ? 8 : 4)
       : (s < log2i(N) / 3 - 1 ? 8 : 4);
}

__host__ __device__ constexpr int stage_span(int N, int S) {
  return S == 0 ? 1 : stage_span(N, S - 1) * stage_radix(N, S - 1);
}

// Element (line l, index e) lives at lds[e * ES + l * LS].
// LINE_FAST selects the butterfly->thread mapping: lines vary fastest across
// lanes (for pencils whose lines are contiguous in HBM), or butterflies do.
template <typename T, int N, int LW, int WG, int ES, int LS, bool LINE_FAST, int S>
__device__ __forceinline__ void fft_stage(cx<T>* lds, const cx<T>* tw, int tid) {
  constexpr int R = stage_radix(N, S);
  constexpr int NR = N / R;
  constexpr int BF = LW * NR;               // butterflies in the workgroup
  constexpr int PER = (BF + WG - 1) / WG;   // butterflies per thread
  // span of already-combined sub-transforms before this stage
  constexpr int Ns = stage_span(N, S);
  cx<T> v[PER][R];
#pragma unroll
  for (int p = 0; p < PER; p++) {
    int b = tid + p * WG;
    if (BF % WG == 0 || b < BF) {
      int l = LINE_FAST ? (b % LW) : (b / NR);
      int j = LINE_FAST ? (b / LW) : (b % NR);
      int k = j & (Ns - 1);
#pragma unroll
      for (int r = 0; r < R; r++) v[p][r] = lds[(j + r * NR) * ES + l * LS];
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; r++) v[p][r] = v[p][r] * tw[(k * r) * (N / (Ns * R))];
      }
      dftR<R>(v[p]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; p++) {
    int b = tid + p * WG;
    if (BF % WG == 0 || b < BF) {
      int l = LINE_FAST ? (b % LW) : (b / NR);
      int j = LINE_FAST ? (b / LW) : (b % NR);
      int k = j & (Ns - 1);
      int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) lds[(base + r * Ns) * ES + l * LS] = v[p][r];
    }
  }
  __syncthreads();
}

template <typename T, int N, int LW, int WG, int ES, int LS, bool LINE_FAST, int S = 0>
__device__ __forceinline__ void fft_lines(cx<T>* lds, const cx<T>* tw, int tid) {
  if constexpr (S < stage_count(N)) {
    fft_stage<T, N, LW, WG, ES, LS, LINE_FAST, S>(lds, tw, tid);
    fft_lines<T, N, LW, WG, ES, LS, LINE_FAST, S + 1>(lds, tw, tid);
  }
}

}  // namespace vie
