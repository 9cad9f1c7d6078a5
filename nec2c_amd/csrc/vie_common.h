// Shared device/host definitions for the MI355X (gfx950) VIE FFT-GMRES library.
//
// Complex values are interleaved (re, im) pairs of float or double, the same
// memory layout as C99 `complex float` / `complex double`, so the C-ABI can
// hand host `complex double` arrays straight to hipMemcpy.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace vie {

template <typename T> struct alignas(2 * sizeof(T)) cx { T re, im; };

using c32 = cx<float>;
using c64 = cx<double>;

template <typename T> __host__ __device__ inline cx<T> mk(T r, T i) { return cx<T>{r, i}; }
template <typename T> __host__ __device__ inline cx<T> operator+(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T> __host__ __device__ inline cx<T> operator-(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T> __host__ __device__ inline cx<T> operator*(cx<T> a, cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T> __host__ __device__ inline cx<T> operator*(T s, cx<T> a) { return {s * a.re, s * a.im}; }
template <typename T> __host__ __device__ inline cx<T> conj(cx<T> a) { return {a.re, -a.im}; }
// multiply by -i
template <typename T> __host__ __device__ inline cx<T> mul_mi(cx<T> a) { return {a.im, -a.re}; }
template <typename T> __host__ __device__ inline cx<T> cdiv(cx<T> a, cx<T> b) {
  // Smith's algorithm keeps the quotient finite for widely scaled operands.
  if ((b.re < 0 ? -b.re : b.re) >= (b.im < 0 ? -b.im : b.im)) {
    T r = b.im / b.re, d = b.re + b.im * r;
    return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  } else {
    T r = b.re / b.im, d = b.im + b.re * r;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
  }
}
template <typename To, typename From> __host__ __device__ inline cx<To> cast(cx<From> a) {
  return {(To)a.re, (To)a.im};
}

// Streaming loads for data read once per pass (Krylov basis, coupling
// blocks): the nontemporal hint keeps them from displacing reused data in
// the caches and lets the memory pipeline stream at a higher rate.
// (device memory only: the load is a global one)
template <typename T> __device__ __forceinline__ cx<T> ld_stream(const cx<T>* p) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  typedef __attribute__((address_space(1))) const v2 gv2;
  v2 v = __builtin_nontemporal_load((gv2*)(uintptr_t)p);
  return cx<T>{v.x, v.y};
}

// Elements staged in LDS per workgroup by the pencil-FFT passes, and the
// workgroup size.  FP32: 4096 x 8 B = 32 KiB; FP64: 2048 x 16 B = 32 KiB.
template <typename T> struct PassShape;
template <> struct PassShape<float> { static constexpr int E = 4096; static constexpr int WG = 256; };
template <> struct PassShape<double> { static constexpr int E = 2048; static constexpr int WG = 256; };

// Device-side operand indexing, so one launch sequence can be recorded once
// (hipGraph) and replayed for every Arnoldi step: a kernel given a DevIndex
// offsets its vector operands by (*j) * stride elements, and returns at once
// when *skip is nonzero.  A default DevIndex means plain pointers.
struct DevIndex {
  const int* j = nullptr;
  size_t stride = 0;
  const int* skip = nullptr;
  __device__ inline size_t off() const { return j ? (size_t)(*j) * stride : 0; }
  __device__ inline bool skipped() const { return skip && *skip; }
};

// A pointer read from device memory is generic (flat) to the compiler; one
// that comes from an integer through the global address space is known to
// be global, so accesses through it are global instructions with a scalar
// base (fewer registers than flat ones).  For device pointers only.
template <typename P> __device__ __forceinline__ P* to_global(P* p) {
  typedef __attribute__((address_space(1))) P gP;
  return (P*)(gP*)(uintptr_t)p;
}
__device__ __forceinline__ DevIndex to_global(DevIndex d) {
  d.j = d.j ? to_global(d.j) : nullptr;
  d.skip = d.skip ? to_global(d.skip) : nullptr;
  return d;
}

// Multi-job launches run a compact 1-D grid: pre[i] is the first workgroup
// of job i and pre[count] the total, so workgroup b serves job
// mj_job(pre, count, b) (the last i with pre[i] <= b; jobs without work have
// pre[i] == pre[i + 1] and are never chosen) as its block b - pre[i].  b is
// uniform across the workgroup, so the search runs on scalar loads.
__device__ __forceinline__ int mj_job(const unsigned* pre, int count, unsigned b) {
  int lo = 0, hi = count;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pre[mid] <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Wave sum by DPP row operations (no LDS crossbar), total in lane 63 only:
// quad swaps (xor 1, xor 2), half-row and row mirrors give every row its sum,
// then row_bcast15 / row_bcast31 carry row sums upward (GCN DPP controls
// 0xB1, 0x4E, 0x141, 0x140, 0x142 with rows 1 and 3, 0x143 with rows 2 and 3).
// Every lane of the wave must be active.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_add_src(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWS, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double wave_sum_last(double v) {
  v += dpp_add_src<0xB1, 0xf>(v);
  v += dpp_add_src<0x4E, 0xf>(v);
  v += dpp_add_src<0x141, 0xf>(v);
  v += dpp_add_src<0x140, 0xf>(v);
  v += dpp_add_src<0x142, 0xa>(v);
  v += dpp_add_src<0x143, 0xc>(v);
  return v;
}
__device__ __forceinline__ c64 wave_sum_last(c64 v) { return c64{wave_sum_last(v.re), wave_sum_last(v.im)}; }

// Geometry of one FFT operator instance (one lattice).
struct Geom {
  int nx, ny, nz;        // unknown box (physical cells; holes carry eps = 1, x = 0)
  int Px, Py, Pz;        // padded FFT extents, powers of two >= 2 n_kernel - 1
  int kz;                // kernel z extent (nz, or 2 nz + gap for the PEC image)
  int pec;               // 1: image cells mirrored below the physical box
  int zoff;              // z index of the first physical plane inside the FFT grid
  int ground = 0;        // 1: a reflected (real-ground) spectrum is added in the x pass
  __host__ __device__ size_t nbox() const { return (size_t)nx * ny * nz; }
};

}  // namespace vie
