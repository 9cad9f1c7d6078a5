// Internal interfaces between the C-ABI layer and the HIP kernels.
#pragma once
#include <vector>
#include "vie_common.h"

namespace vie {

// Device views of one FFT volume operator at precision T.
template <typename T> struct VolumeView {
  Geom g;
  cx<T>* A;           // [3][nx][ny][Pz] workspace
  cx<T>* B;           // [3][nx][Py][Pz] workspace
  const cx<T>* spec;  // [6][Px/2+1][Py/2+1][Pz/2+1], scaled by 1/(Px Py Pz)
  const cx<T>* rspec = nullptr;  // real ground: [6][Px/2+1][Py/2+1][Pz] reflected spectrum
  const cx<T>* eps;   // [nbox]
  const cx<T>* tw;    // twiddle tables, x then y then z
  cx<T> self[3];
  unsigned char var[5] = {0, 0, 0, 0, 0};  // launch-shape variant of each pass
};

inline bool max_fft_extent_ok(int P) { return P >= 1 && P <= 1024 && (P & (P - 1)) == 0; }
// Launch-shape variants available to a pass over lines of length P.
constexpr int kPassVariants = 4;
inline int pass_variants(int P) { return P >= 16 ? kPassVariants : 1; }
// One pass (0 Z forward, 1 Y forward, 2 X, 3 Y inverse, 4 Z inverse) at a
// given variant, for the tuner.
template <typename T>
int pass_launch(const VolumeView<T>& v, int pass, int variant, const cx<T>* xin, cx<T>* yout, int precond,
                hipStream_t s);
// Launch variants of a pass: the four shapes (one below 16 points), and for
// the x pass also those shapes with kz-mirror pairing (variants 4-7).  A
// variant applies to a geometry when pass_variant_ok: paired groups must
// divide Pz, and over a real ground every x-pass group must hold the kz
// mirror of each of its lines (paired, or whole ky rows).
template <typename T> int pass_variant_count(const Geom& g, int pass);
template <typename T> bool pass_variant_ok(const Geom& g, int pass, int variant);

// Fused-pass variant codes: pass 0 (forward) or 4 (inverse) at kFuseBase +
// log2(Py) does both of its transforms, and passes 1 and 3 at kFuseSkip then
// launch nothing.  The y extent rides in the code so that multi-job classes
// (keyed by pass extent and variant) hold one (Pz, Py) pair each.
constexpr int kFuseBase = 96, kFuseSkip = 120;
__host__ __device__ inline bool fused_extent_ok(int P) { return P == 32 || P == 64 || P == 128; }
inline bool fused_ok(const Geom& g) {
  return fused_extent_ok(g.Pz) && fused_extent_ok(g.Py) && 2 * g.ny <= g.Py && g.nx >= 1;
}
inline int fused_code(const Geom& g) { return kFuseBase + (g.Py == 32 ? 5 : g.Py == 64 ? 6 : 7); }

// A fused forward (pass 0) goes with a skipped pass 1 and a fused inverse
// (pass 4) with a skipped pass 3, and the other way round.
inline bool variants_consistent(const unsigned char* var) {
  const bool f0 = var[0] >= kFuseBase && var[0] != kFuseSkip, s1 = var[1] == kFuseSkip;
  const bool f4 = var[4] >= kFuseBase && var[4] != kFuseSkip, s3 = var[3] == kFuseSkip;
  return f0 == s1 && f4 == s3;
}

// Launch arguments of the five volume passes (one job).
template <typename T> struct OpArgs {
  Geom g;
  const cx<T>* xin;   // [3][nbox] input (v when precond, else z)
  cx<T>* yout;        // [3][nbox] output
  cx<T>* A;           // [3][nx][ny][Pz]
  cx<T>* B;           // [3][nx][Py][Pz]
  const cx<T>* spec;  // [6][Px/2+1][Py/2+1][Pz/2+1] octant spectrum / G
  const cx<T>* rspec; // real ground: [6][Px/2+1][Py/2+1][Pz] reflected spectrum / G
  const cx<T>* eps;   // [nbox] relative permittivity (1 in holes)
  const cx<T>* tw;    // twiddles: x table (Px), then y (Py), then z (Pz)
  cx<T> self[3];      // self-term diagonal N_ii(0)
  int precond;
  DevIndex di;        // xin / yout offset and skip flag (Arnoldi step replay)
};

// A job's pass arguments as read from a multi-job table, its pointers marked global.
template <typename T> __device__ __forceinline__ OpArgs<T> op_args_global(const OpArgs<T>& t) {
  OpArgs<T> a = t;
  a.xin = to_global(a.xin);
  a.yout = to_global(a.yout);
  a.A = to_global(a.A);
  a.B = to_global(a.B);
  a.spec = to_global(a.spec);
  a.rspec = a.rspec ? to_global(a.rspec) : nullptr;
  a.eps = to_global(a.eps);
  a.tw = to_global(a.tw);
  a.di = to_global(a.di);
  return a;
}

template <typename T>
inline OpArgs<T> op_args(const VolumeView<T>& v, const cx<T>* xin, cx<T>* yout, int precond, DevIndex di = {}) {
  OpArgs<T> a;
  a.di = di;
  a.g = v.g;
  a.xin = xin;
  a.yout = yout;
  a.A = v.A;
  a.B = v.B;
  a.spec = v.spec;
  a.rspec = v.rspec;
  a.eps = v.eps;
  a.tw = v.tw;
  for (int i = 0; i < 3; i++) a.self[i] = v.self[i];
  a.precond = precond;
  return a;
}

// Multi-job launches: one launch per pass covers every job whose lines of
// that pass have the same length P and the same launch variant (a class),
// for P up to kMjMaxP and the variants of the LDS kernels (pass_mj_ok).
// ops[q] and blocks[5 q + pass] describe job q; lists + offset holds the
// job indices of a class, blockIdx.y walking them.
// pres + pre holds the class's workgroup prefix (count + 1 entries; see mj_job).
constexpr int kMjMaxP = 128;
struct MjClass {
  int pass, P, variant, count, offset, pre;
  unsigned total_blocks;
};
// Workgroups of pass `pass` at `variant` for geometry g (as its single-job launch).
template <typename T> unsigned pass_blocks_mj(const Geom& g, int pass, int variant);
template <typename T> bool pass_mj_ok(const Geom& g, int pass, int variant);
template <typename T>
int volume_apply_mj(const OpArgs<T>* ops, const unsigned* blocks, const int* lists, const unsigned* pres,
                    const MjClass* classes, int nclasses, hipStream_t s);

// y = eps z - (eps-1) conv(z), z = precond ? D^{-1} xin : xin.  Volume part only.
// ev (optional, 6 events): recorded before each of the five passes and after the last.
// di (optional): xin / yout offset by (*di.j) * di.stride on the device, whole
// apply skipped while *di.skip is set.
template <typename T>
int volume_apply(const VolumeView<T>& v, const cx<T>* xin, cx<T>* yout, int precond, hipStream_t s,
                 hipEvent_t* ev = nullptr, DevIndex di = {});

// ---- kernel spectrum (FP64) ----
struct SpectrumResult {
  c64 self[3];
  int near_offsets = 0, far_offsets = 0, fallback_offsets = 0, max_near_order = 0;
};
// Builds the octant spectrum [6][Px/2+1][Py/2+1][Pz/2+1] (FP64, scaled by
// 1/G) into spec64 and its FP32 rounding into spec32 (either may be null).
// kshape: lattice extent of the kernel (z includes the PEC image).
int build_spectrum(const int kshape[3], const int padded[3], double h, double k, c64* spec64,
                   c32* spec32, SpectrumResult* out, hipStream_t s);

// ---- reflected (real-ground) tensor (FP64) ----
// T[6][nx][ny][2 nz - 1] on the device: the reflected tensor of a lower
// half-space of permittivity eps between cells (dx, dy) apart laterally whose
// heights add to 2 z0 + s h (z0: centre of the lowest cell layer), at
// Gauss-Legendre order `order` per u segment.
int ground_table(int nx, int ny, int nz, double h, double z0, double k, c64 eps, int order, c64* T_dev,
                 hipStream_t s);
// Its spectrum [6][Px/2+1][Py/2+1][Pz] (lateral octant, full kz), scaled by
// 1/(Px Py Pz), into rspec64 and its FP32 rounding into rspec32.
int build_ground_spectrum(int nx, int ny, int nz, const int padded[3], double h, double z0, double k, c64 eps,
                          int order, c64* rspec64, c32* rspec32, hipStream_t s);

// ---- scattered far field (farfield.hip) ----
// Slices of the cells per direction tile (the partial sums per direction are
// 2 x 3 x slices values).
unsigned farfield_slices(size_t n, size_t ndir, int ground);
// amplitude[3 d] for the unit directions dirs[3 d] (device arrays; centers and
// sizes [3 n], sizes of one cell only when uniform, current [3 n] cell-major);
// ground 0 free space, 1 PEC plane, 2 per-direction rs / rp.  partial holds
// 6 ndir slices values.
int farfield_amplitudes(const double* centers, const double* sizes, int uniform, const c64* current, size_t n,
                        const double* dirs, size_t ndir, double k0, int ground, const c64* rs, const c64* rp,
                        c64* partial, unsigned slices, c64* amplitude, hipStream_t s);

// Far field of the solved wires and patches (classic engine): segments
// x y z si cab sab salp [n] in wavelengths with currents cur[3 n] (a, b, c
// per segment, NEC's amperes per wavelength); patches pp[3 m], pbi[m] with
// currents pcur[3 m].  et/ep: NEC's E_theta, E_phi (r in wavelengths) for
// each (theta, phi); pec: the image over a PEC plane.
struct ConductorFar {
  int n, m;
  const double *x, *y, *z, *si, *cab, *sab, *salp;
  const c64* cur;
  const double *pp, *pbi;
  const c64* pcur;
};
int conductor_far_field(const ConductorFar& cf, int pec, size_t count, const double* theta, const double* phi,
                        c64* et, c64* ep, hipStream_t s);

// ---- volume-to-conductor rows (conductor_rows.hip) ----
// Rows of the conductor equations over the volume unknowns (device arrays):
// wire task w writes row w, patch task p rows nw + 2p and nw + 2p + 1; entry
// (row, 3 cell + q) at rows[row ld + 3 cell + q].  ground 0 free space, 1 PEC
// plane.  max_size: the largest |size| of a cell (the launch picks the phase
// evaluation from it; NaN or inf selects the per-cell choice).  failed
// (initialised to ~0ull by the caller) receives the smallest task * cells +
// cell whose quadrature met the observation point.
int conductor_rows(double k0, const double* centers, const double* sizes, size_t cells, double max_size, int nw,
                   const double* wp, const double* wt, int np, const double* pp, const double* t1, const double* t2,
                   const double* ps, int ground, c64* rows, size_t ld, unsigned long long* failed, hipStream_t s);

// ---- conductor-to-volume coupling (wire_volume.hip) ----
// NEC's segment and patch arrays on the device, in wavelengths: segments
// x y z si bi cab sab salp [n]; patches pp [3 m] (centres), pbi [m] (areas),
// t1 t2 [3 m]; rkh the lumped-element range.
struct ConductorSet {
  int n, m;
  const double *x, *y, *z, *si, *bi, *cab, *sab, *salp;
  const double *pp, *pbi, *t1, *t2;
  double rkh;
};
// F[(jbase[a] + k) * ldf + 3 (cell - cell0) + q]: the cell-averaged field of
// basis term k of source active[a] (a segment index, or n + patch index:
// segments have 3 terms, constant / sine / cosine current, patches 2, t1 /
// t2 current) over cells [cell0, cell0 + cells); ground 0 free, 1 PEC.
int conductor_volume_fields(const ConductorSet& cs, const int* active, const int* jbase, int nactive,
                            const double* centers, const double* sizes, size_t cell0, size_t cells, double wlam,
                            int ground, c64* F, size_t ldf, hipStream_t s);
// coupling[col * ldc + 3 cell + q] = -i omega eps0 chi[cell] sum_j
// coef[col * terms + j] F[j][3 (cell - cell0) + q] for the chunk's cells; a
// term of segment term_segment[j] >= 0 is skipped where skip[col * nseg +
// segment] is set.
int conductor_volume_columns(const c64* F, size_t ldf, int terms, const int* term_segment, const c64* coef,
                             const unsigned char* skip, int nseg, int columns, const c64* chi, size_t cell0,
                             size_t cells, double k0, c64* coupling, size_t ldc, hipStream_t s);

// ---- vector kernels (krylov.hip) ----
template <typename T> void vec_zero(cx<T>* x, size_t n, hipStream_t s);
template <typename To, typename From> void vec_cast(cx<To>* y, const cx<From>* x, size_t n, hipStream_t s);

// ---- mesh relationship checks (mesh_check.hip) ----
// first <- i * count + j of the first pair (i < j, i outer) of boxes
// [mn, mx] (count x 3 each) whose interiors overlap, or ~0 when none does.
int box_overlap(const double* mn, const double* mx, size_t count, unsigned long long* first, hipStream_t s);
// Boxes (count x 3 corners, cell_longest <= 0 for a box without cells) against
// n wire segments (NEC's centre, length, direction cosines, radius) and m
// patches (centre, area), all device pointers.
struct ClearanceArgs {
  const double *minimum, *maximum, *cell_longest;
  size_t boxes;
  long long n, m;
  const double *x, *y, *z, *si, *cab, *sab, *salp, *bi;
  const double *px, *py, *pz, *pbi;
};
struct ClearanceResult {
  double worst_margin, worst_gap, worst_required, minimum_gap, required_gap;
  long long worst_box, worst_conductor;  // 1-based, 0 when nothing was evaluated
  int worst_is_patch;
};
size_t clearance_partials(size_t pairs);  // partial records (16 bytes, two per block) to provide
int conductor_clearance(const ClearanceArgs& a, void* partials, ClearanceResult* out, hipStream_t s);

}  // namespace vie
