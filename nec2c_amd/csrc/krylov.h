// Device-resident GMRES building blocks (krylov.hip).
#pragma once
#include "vie_common.h"

namespace vie {

constexpr int kMaxRestart = 128;           // largest Krylov cycle the device solver keeps
constexpr int kMaxRecycle = 32;            // largest recycled subspace (GCRO cycles)
constexpr int kMaxConductorsShared = 256;  // conductor coefficients staged in LDS

// Small per-solve state kept in device memory.  The host reads the leading
// status words after each Arnoldi step.
struct ArnoldiState {
  // --- status block read by the host
  double gabs;      // |g_{j+1}|, the residual estimate after column j
  double beta;      // ||r|| at the start of the cycle
  int breakdown;    // ||w|| <= eps_b ||A z||
  int nonfinite;
  int singular;     // zero pivot in back substitution
  int used;         // columns in the current cycle
  int done;         // cycle finished: later Arnoldi tails leave the state alone
  int pad_;
  // --- working data
  int j, m;
  int step;  // Arnoldi steps issued in this cycle (status ring slot)
  // first Krylov column of the cycle: columns 0..base-1 of the basis hold the
  // recycled images C (A M^{-1} U = C, orthonormal) and are never expanded;
  // Hessenberg rows 0..base-1 are the projections C^H A M^{-1} q_j (unrotated)
  int base;
  int ju;  // column that the step's fused update finishes (its reduction's j)
  double inv_beta, inv_next;
  double inner_target;  // cycle stops once |g_{j+1}| <= inner_target
  // delayed (DCGS2) orthogonalisation: coefficients of the fused update,
  // the unrotated Hessenberg and the norm estimate of the last column
  double dc_alpha, dc_f_re, dc_f_im, dc_before;
  c64 dc_a[kMaxRestart];
  c64 dc_e[kMaxRestart];
  c64 Hraw[(kMaxRestart + 1) * kMaxRestart];
  c64 g[kMaxRestart + 1];
  double cs[kMaxRestart];
  c64 sn[kMaxRestart];
  c64 y[kMaxRestart];
  c64 H[(kMaxRestart + 1) * kMaxRestart];
  // recycled cycles: projection of the cycle's start residual on C (in units
  // of beta), and the coefficients of U in the cycle's solution update
  c64 zc[kMaxRecycle];
  c64 uc[kMaxRecycle];
};

// The leading status words of ArnoldiState, as the host reads them.
struct StepStatus {
  double gabs, beta;
  int breakdown, nonfinite, singular, used, done, pad_;
};

constexpr int kMaxKeys = 320;  // sums per reduction launch (>= 2 kMaxRestart + 3)
// Krylov basis columns are kBasisAlign-element aligned and zero past the
// vector length n: the delayed-CGS2 kernels stream whole 512-element chunks
// with no per-element bounds test (the padding adds zeros to every sum and is
// rewritten as zeros by the update).
constexpr size_t kBasisAlign = 512;
inline size_t basis_ld(size_t n) { return (n + kBasisAlign - 1) / kBasisAlign * kBasisAlign; }

struct ReduceScratch {
  c64* partial = nullptr;  // kMaxDotBlocksHost x kMaxKeys partial sums
};
constexpr size_t kReduceScratchBytes = (size_t)1024 * kMaxKeys * sizeof(c64);

// Jacobi scaling of the volume part of a system vector, for kernels that
// read v and need z = D^{-1} v (index i counts from the first volume entry).
template <typename T> struct JacobiLite {
  size_t nbox;
  const cx<T>* eps;
  cx<T> self[3];
  __device__ inline cx<T> apply(size_t i, cx<T> v) const {
    size_t comp = i / nbox, cell = i - comp * nbox;
    cx<T> e = eps[cell], one = {(T)1, (T)0};
    // selects, not an indexed load: a copy of the struct stays in registers
    const cx<T> sf = comp == 0 ? self[0] : comp == 1 ? self[1] : self[2];
    return cdiv(v, e - (e - one) * sf);
  }
};

struct JacobiView {
  size_t c, nbox;
  const c64* cdiag;  // conductor diagonal C_ii
  const c64* eps;    // [nbox]
  c64 self[3];
};

template <typename T> struct CouplingView {
  int c;
  size_t nv, nbox;
  const cx<T>* C;  // c x c row-major
  const cx<T>* B;  // c rows of nv (volume -> conductor)
  const cx<T>* D;  // c columns of nv (conductor -> volume)
  c64* bz;         // c scratch (B z_v when c exceeds one reduction launch)
  const cx<T>* eps;
  cx<T> self[3];
};

size_t red_blocks(size_t n);
void copy16(void* dst, const void* src, size_t n16, hipStream_t s);

// dots[k] = V[k]^H w (k < nk), dots[nk] = ||w||^2 when self_norm.
template <typename T>
void multidot(const cx<T>* V, size_t ldv, int nk, const cx<T>* w, size_t n, int self_norm, ReduceScratch& rs,
              c64* dots, hipStream_t s);
// w -= sum_k h[k] V[k]; with want_norm the ||w||^2 partials stay in rs for
// arnoldi_tail.  Returns the number of partials.
template <typename T>
unsigned update(const cx<T>* V, size_t ldv, int nk, const c64* h, cx<T>* w, size_t n, int want_norm,
                ReduceScratch& rs, hipStream_t s);
template <typename T> void norm2(const cx<T>* x, size_t n, ReduceScratch& rs, c64* out, hipStream_t s);
void residual(const c64* b, const c64* ax, c64* r, size_t n, ReduceScratch& rs, c64* out, hipStream_t s);
void arnoldi_tail(ArnoldiState* st, const c64* h1, const c64* h2, const c64* before2, ReduceScratch& rs,
                  unsigned nblk, double breakdown_eps, hipStream_t s);
void backsolve(ArnoldiState* st, hipStream_t s);
// Zeroes entries [n, ldv) of the first cols columns of a basis.
template <typename T> void zero_padding(cx<T>* V, size_t ldv, size_t n, size_t cols, hipStream_t s);
// Delayed CGS2 Arnoldi step j = st->j, read on the device (slot j holds the
// once-projected candidate u, slot j+1 holds w = A M^{-1} u, or nothing
// without has_w: the closing step of a cycle).  One reduction pass, one small
// solve, one fused update pass; the small solve also writes the status block
// to ring[st->step & 1] (device-visible host memory) and advances j unless the
// cycle is done.  The launch sequence does not depend on j, so it can be
// recorded once and replayed for every step.  max_j bounds j (the restart).
// ldv = basis_ld(n) and the padding of every column read is zero.
template <typename T>
void dcgs2_step(cx<T>* Q, size_t ldv, bool has_w, size_t n, ArnoldiState* st, ReduceScratch& rs, c64* dots,
                double breakdown_eps, int max_j, StepStatus* ring, hipStream_t s);
// Times the reduction and the fused update of a delayed-CGS2 step at column
// j (averages over reps launches); overwrites the state and columns 0..j+1.
template <typename T>
int dcgs2_time(cx<T>* Q, size_t ldv, size_t n, ArnoldiState* st, ReduceScratch& rs, int j, int reps, hipStream_t s,
               float* ms_reduce, float* ms_update);
// Starts a cycle at basis column `base` (0 for plain GMRES cycles).
void cycle_init(ArnoldiState* st, const c64* beta2, double inner_target, int base, hipStream_t s);
// Recycled cycle start: with the scaled residual u in column k of Q and the
// recycled images C in columns 0..k-1, st->zc = C^H u and u -= C zc.
template <typename T>
void recycle_project(cx<T>* Q, size_t ldv, int k, size_t n, ArnoldiState* st, ReduceScratch& rs, hipStream_t s);
// out[:, i] = sum_c in[:, c] coef[c + i * nin] for i < nout (column stride ldv).
template <typename T>
void block_combine(cx<T>* out, int nout, const cx<T>* in, int nin, size_t ldv, const c64* coef, size_t n,
                   hipStream_t s);
template <typename T> void scale_from(cx<T>* v, const c64* r, size_t n, const ArnoldiState* st, hipStream_t s);
template <typename T> void scale_next(cx<T>* v, size_t n, const ArnoldiState* st, hipStream_t s);
// x += M^{-1} (sum_k y_k V[k] + sum_l uc_l U[l]); U (columns of stride ldv)
// is read only in a recycled cycle (st->base > 0).
template <typename T>
void combine(c64* x, const cx<T>* V, size_t ldv, size_t n, const ArnoldiState* st, const JacobiView& J, int precond,
             hipStream_t s, const cx<T>* U = nullptr);
template <typename T>
void to_box(const c64* user, size_t c, size_t cells, const size_t* map, size_t nbox, cx<T>* sys, hipStream_t s);
template <typename T>
void from_box(const cx<T>* sys, size_t c, size_t cells, const size_t* map, size_t nbox, c64* user, hipStream_t s);
// D[r][comp * nbox + map[cell]] = host_order[(3 cell + comp) * c + r] (the
// reference's v x c conductor-column block into the device layout).
void columns_to_box(const c64* host_order, size_t c, size_t cells, const size_t* map, size_t nbox, c64* D,
                    hipStream_t s);
void precondition_apply(const c64* x, c64* y, size_t n, const JacobiView& J, hipStream_t s);
// di offsets vc, vv, yc, yv on the device and skips the whole product.
template <typename T>
void coupling_apply(const CouplingView<T>& cv, const cx<T>* vc, const cx<T>* vv, cx<T>* yc, cx<T>* yv, int precond,
                    ReduceScratch& rs, hipStream_t s, DevIndex di = {});

// ---- multi-job Arnoldi steps ----
// One job's share of a multi-job delayed-CGS2 matvec step: the arguments its
// single-job coupling (Jacobi-scaled input) and delayed-CGS2 launches take,
// and its own workgroup counts.  Every kernel of the step takes the job
// table and walks the jobs with blockIdx.y; workgroups past a job's own count
// return at once.
template <typename T> struct MjStep {
  // conductor coupling: c conductors (0: none, at most kMaxKeys), nv volume
  // unknowns; x / y the step's input and output columns (offset by di)
  int c;
  size_t nv;
  const cx<T>* C;
  const cx<T>* B;
  const cx<T>* D;
  const cx<T>* x;
  cx<T>* y;
  JacobiLite<T> J;
  DevIndex di;
  // delayed CGS2 over the cycle basis Q (n unknowns, leading dimension ldv)
  cx<T>* Q;
  size_t ldv, n;
  ArnoldiState* st;
  c64* partial;  // reduction partials, kMaxDotBlocks per key
  c64* dots;
  StepStatus* ring;
  double eps_b;
  int max_j;
  // workgroups: coupling reduction, coupling output, delayed-CGS2 reduction,
  // its finishing sums, the fused update
  unsigned nb_dots, nb_out, nb_dc, nb_finish, nb_update;
};
// Workgroup counts of one job (fills the nb_ fields from c, nv, n, max_j).
template <typename T> void step_blocks_mj(MjStep<T>& m);
// The coupling and delayed-CGS2 kernels of one step for count jobs (tab a
// device array); pre: five workgroup prefixes of count + 1 entries (the nb_
// fields in order, see mj_job), total the five totals.
template <typename T>
void step_mj(const MjStep<T>* tab, int count, const unsigned* pre, const unsigned total[5], hipStream_t s);

}  // namespace vie
