// C-ABI of libvie_hip.so: the reference's vie_fft.h / vie_iterative.h /
// dielectric_fft.h entry points plus the device-resident solve (vie_hip.h).
//
// Host code here owns validation, budgets, layout maps and control flow; all
// arithmetic on the operator and the Krylov vectors runs in the HIP kernels of
// vie_passes.hip, spectrum.hip and krylov.hip.  There is no CPU compute path:
// without a usable GPU every create/prepare fails with a diagnostic.
#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <functional>
#include <array>
#include <chrono>
#include <map>
#include <thread>
#include <memory>
#include <cmath>
#include <complex>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/vie_hip.h"
#include "krylov.h"
#include "small_dense.h"
#include "vie_kernels.h"

using vie::c32;
using vie::c64;

// ------------------------------------------------------------------ errors
static thread_local std::string g_last_error;
static void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}
extern "C" const char* vie_hip_last_error(void) { return g_last_error.c_str(); }

#define HIP_OK(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
      return 0;                                                                   \
    }                                                                             \
  } while (0)

// ------------------------------------------------------------------ globals
enum { BACKEND_AUTO, BACKEND_HIP };
static int g_requested = BACKEND_AUTO;
static int g_last_used = -1;
static vie_hip_solve_options g_solve = {VIE_HIP_FP64, 0.0, 0};
static bool g_solve_init = false;
static size_t g_budget = 0;
static std::mutex g_mutex;

static const double kPi = 3.141592653589793238462643383279502884;

static void init_solve_defaults() {
  if (g_solve_init) return;
  g_solve_init = true;
  const char* p = getenv("VIE_HIP_PRECISION");
  if (p && (!strcmp(p, "mixed") || !strcmp(p, "MIXED"))) g_solve.precision = VIE_HIP_MIXED;
}

// ------------------------------------------------------------------ timing phases
// Wall seconds and calls of the library's hot-path work, cumulative over the
// process (or since vie_hip_phases_reset), for timing reports such as nec2c's
// performance_add phases.
enum {
  PHASE_BUILD, PHASE_UPLOAD, PHASE_GMRES, PHASE_MATVEC, PHASE_ROWS, PHASE_COLUMNS, PHASE_FAR_FIELD, PHASE_MESH,
  PHASE_COUNT
};
static const char* const kPhaseNames[PHASE_COUNT] = {
    "fft_build", "fft_upload", "fft_gmres", "fft_matvec", "fft_conductor_rows", "fft_conductor_columns",
    "fft_far_field", "fft_mesh_checks"};
static double g_phase_s[PHASE_COUNT];
static unsigned long long g_phase_n[PHASE_COUNT];
static std::mutex g_phase_mu;
static thread_local int g_phase_depth = 0;  // an outer timed call owns nested work

struct PhaseTimer {
  int id;
  bool own;
  std::chrono::steady_clock::time_point t0;
  explicit PhaseTimer(int i) : id(i), own(g_phase_depth == 0), t0(std::chrono::steady_clock::now()) {
    g_phase_depth++;
  }
  ~PhaseTimer() {
    g_phase_depth--;
    if (!own) return;
    double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::lock_guard<std::mutex> g(g_phase_mu);
    g_phase_s[id] += dt;
    g_phase_n[id]++;
  }
};

extern "C" size_t vie_hip_phases(vie_hip_phase* out, size_t max) {
  std::lock_guard<std::mutex> g(g_phase_mu);
  size_t k = 0;
  for (int i = 0; i < PHASE_COUNT; i++) {
    if (!g_phase_n[i]) continue;
    if (out && k < max) out[k] = vie_hip_phase{kPhaseNames[i], g_phase_s[i], g_phase_n[i]};
    k++;
  }
  return k;
}
extern "C" void vie_hip_phases_reset(void) {
  std::lock_guard<std::mutex> g(g_phase_mu);
  for (int i = 0; i < PHASE_COUNT; i++) g_phase_s[i] = 0, g_phase_n[i] = 0;
}

extern "C" int vie_hip_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

extern "C" int vie_hip_memory_info(size_t* free_bytes, size_t* total_bytes) {
  if (vie_hip_device_count() <= 0) return 0;
  size_t f = 0, t = 0;
  if (hipMemGetInfo(&f, &t) != hipSuccess) return 0;
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return 1;
}

static int ensure_device() {
  int n = vie_hip_device_count();
  if (n <= 0) {
    set_error("no HIP device available");
    return 0;
  }
  return 1;
}

// ------------------------------------------------------------------ device buffers
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  int alloc(size_t b) {
    if (bytes >= b && p) return 1;
    release();
    if (b == 0) b = 16;
    if (hipMalloc(&p, b) != hipSuccess) {
      p = nullptr;
      set_error("hipMalloc of %zu bytes failed", b);
      return 0;
    }
    bytes = b;
    return 1;
  }
  template <typename T> T* as() const { return (T*)p; }
};

static bool mul_ok(size_t a, size_t b, size_t* out) {
  if (b && a > SIZE_MAX / b) return false;
  *out = a * b;
  return true;
}
static bool add_ok(size_t a, size_t b, size_t* out) {
  if (a > SIZE_MAX - b) return false;
  *out = a + b;
  return true;
}

// ------------------------------------------------------------------ plan
struct vie_fft_plan {
  vie::Geom g{};
  unsigned long long gen = 1;  // permittivity generation: bumped by each upload
  size_t cells = 0;
  size_t bytes = 0;
  double h = 0, k = 0;
  std::vector<size_t> map;       // cell -> box index
  std::vector<c64> eps;          // per cell
  c64 self[3] = {};
  vie::SpectrumResult spec_info;
  bool eps_dirty = true;
  hipStream_t stream = nullptr;
  // device
  DevBuf spec64, spec32, work, tw64, tw32, eps64, eps32, map_dev;
  DevBuf rspec64, rspec32;                // real ground: reflected spectrum [6][Px/2+1][Py/2+1][Pz]
  c64 ground_eps = {1, 0};
  double ground_z0 = 0;                   // centre height of the lowest cell layer
  DevBuf io_user, io_sys_in, io_sys_out;  // host-pointer apply staging
  dielectric_fft_system* wrap = nullptr;  // c = 0 system around this plan, for solves
  unsigned char var[2][5] = {};          // pass launch variants: [0] FP64, [1] FP32
  // the variants of multi-job launches: the separate passes (many jobs per
  // launch hide the small passes' latency, so the fused ones, tuned on one
  // job, measured slower there)
  unsigned char var_mj[2][5] = {};
  bool tuned[2] = {false, false};
  ~vie_fft_plan();
  size_t nbox() const { return g.nbox(); }
  size_t oct() const { return (size_t)(g.Px / 2 + 1) * (g.Py / 2 + 1) * (g.Pz / 2 + 1); }
  template <typename T> vie::VolumeView<T> view();
};

template <> vie::VolumeView<double> vie_fft_plan::view<double>() {
  vie::VolumeView<double> v;
  v.g = g;
  size_t a_elems = 3 * (size_t)g.nx * g.ny * g.Pz;
  v.A = work.as<c64>();
  v.B = work.as<c64>() + a_elems;
  v.spec = spec64.as<c64>();
  v.rspec = g.ground ? rspec64.as<c64>() : nullptr;
  v.eps = eps64.as<c64>();
  v.tw = tw64.as<c64>();
  for (int i = 0; i < 3; i++) v.self[i] = self[i];
  for (int i = 0; i < 5; i++) v.var[i] = var[0][i];
  return v;
}
template <> vie::VolumeView<float> vie_fft_plan::view<float>() {
  vie::VolumeView<float> v;
  v.g = g;
  size_t a_elems = 3 * (size_t)g.nx * g.ny * g.Pz;
  v.A = work.as<c32>();
  v.B = work.as<c32>() + a_elems;
  v.spec = spec32.as<c32>();
  v.rspec = g.ground ? rspec32.as<c32>() : nullptr;
  v.eps = eps32.as<c32>();
  v.tw = tw32.as<c32>();
  for (int i = 0; i < 3; i++) v.self[i] = c32{(float)self[i].re, (float)self[i].im};
  for (int i = 0; i < 5; i++) v.var[i] = var[1][i];
  return v;
}

static size_t pad_extent(size_t n) {
  size_t p = 1;
  while (p < 2 * n - 1) p *= 2;
  return p;
}

static bool passive(c64 e) { return std::isfinite(e.re) && std::isfinite(e.im) && e.re >= 1 && e.im <= 0; }

// Device bytes of a plan, counted before anything is allocated.
static bool plan_bytes(const vie::Geom& g, size_t cells, size_t* out) {
  size_t nbox = g.nbox(), oct = (size_t)(g.Px / 2 + 1) * (g.Py / 2 + 1) * (g.Pz / 2 + 1), t, b = sizeof(vie_fft_plan);
  size_t work = 3 * (size_t)g.nx * g.ny * g.Pz + 3 * (size_t)g.nx * g.Py * g.Pz;
  if (!mul_ok(oct, 6 * (16 + 8), &t) || !add_ok(b, t, &b)) return false;
  if (!mul_ok(work, 16, &t) || !add_ok(b, t, &b)) return false;
  if (!mul_ok(nbox, 16 + 8, &t) || !add_ok(b, t, &b)) return false;
  if (!mul_ok(cells, 2 * sizeof(size_t) + sizeof(c64), &t) || !add_ok(b, t, &b)) return false;
  if (!add_ok(b, ((size_t)g.Px + g.Py + g.Pz) * 24, &b)) return false;
  // host-pointer apply staging: user vector + two system vectors (FP64)
  if (!mul_ok(3 * cells + 6 * nbox, 16, &t) || !add_ok(b, t, &b)) return false;
  *out = b;
  return true;
}

static void twiddles(int P, std::vector<c64>& out) {
  for (int m = 0; m < P; m++) {
    double s, c;
    if (4 * m % P == 0) {
      static const double C[4] = {1, 0, -1, 0}, S[4] = {0, -1, 0, 1};
      int q = (4 * m / P) & 3;
      c = C[q];
      s = S[q];
    } else {
      c = cos(2 * kPi * m / P);
      s = -sin(2 * kPi * m / P);
    }
    out.push_back(c64{c, s});
  }
}

static int plan_upload_eps(vie_fft_plan* p) {
  if (!p->eps_dirty) return 1;
  size_t nbox = p->nbox();
  std::vector<c64> box(nbox, c64{1, 0});
  for (size_t c = 0; c < p->cells; c++) box[p->map[c]] = p->eps[c];
  std::vector<c32> box32(nbox);
  for (size_t i = 0; i < nbox; i++) box32[i] = c32{(float)box[i].re, (float)box[i].im};
  HIP_OK(hipMemcpyAsync(p->eps64.p, box.data(), nbox * sizeof(c64), hipMemcpyHostToDevice, p->stream));
  HIP_OK(hipMemcpyAsync(p->eps32.p, box32.data(), nbox * sizeof(c32), hipMemcpyHostToDevice, p->stream));
  HIP_OK(hipStreamSynchronize(p->stream));
  p->eps_dirty = false;
  p->gen++;
  return 1;
}

// ------------------------------------------------------------------ pass-shape tuner
// Picks, per pass and precision, the launch variant (lines per workgroup,
// workgroup size) with the shortest measured time on this geometry.  Results
// are cached per (precision, box, padded extents, PEC) for the life of the
// process, so every plan of one size class is tuned once.  VIE_HIP_TUNE=0
// keeps variant 0 everywhere.
namespace {
struct TuneKey {
  int prec, nx, ny, nz, Px, Py, Pz, pec, fuse;
  bool operator<(const TuneKey& o) const {
    return memcmp(this, &o, sizeof(TuneKey)) < 0;
  }
};
std::mutex g_tune_mu;
std::map<TuneKey, std::array<unsigned char, 10>> g_tune_cache;  // var, then var_mj
}  // namespace

static bool tuning_enabled() {
  const char* e = getenv("VIE_HIP_TUNE");
  return !(e && !strcmp(e, "0"));
}

template <typename T> static int plan_tune(vie_fft_plan* p) {
  const int slot = sizeof(T) == sizeof(double) ? 0 : 1;
  if (p->tuned[slot]) return 1;
  if (!tuning_enabled()) {
    p->tuned[slot] = true;
    return 1;
  }
  TuneKey key;
  memset(&key, 0, sizeof(key));
  const char* fz = getenv("VIE_HIP_FUSE");
  const bool fuse = !(fz && !strcmp(fz, "0"));
  const char* mt = getenv("VIE_HIP_MJ_TUNED");
  const bool mj_tuned = mt && !strcmp(mt, "1");
  key = TuneKey{slot, p->g.nx, p->g.ny, p->g.nz, p->g.Px, p->g.Py, p->g.Pz, p->g.pec + 2 * p->g.ground,
                fuse + 2 * mj_tuned};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      memcpy(p->var[slot], it->second.data(), 5);
      memcpy(p->var_mj[slot], it->second.data() + 5, 5);
      p->tuned[slot] = true;
      return 1;
    }
  }
  const size_t n = 3 * p->nbox();
  DevBuf xin, yout;
  if (!xin.alloc(n * sizeof(vie::cx<T>)) || !yout.alloc(n * sizeof(vie::cx<T>))) return 0;
  hipStream_t st = p->stream;
  HIP_OK(hipMemsetAsync(xin.p, 0, n * sizeof(vie::cx<T>), st));
  for (int pass = 0; pass < 5; pass++) {  // a valid start (the x pass over a ground needs paired groups)
    int v0 = 0;
    while (!vie::pass_variant_ok<T>(p->g, pass, v0)) v0++;
    p->var[slot][pass] = (unsigned char)v0;
  }
  vie::VolumeView<T> v = p->view<T>();
  // a zero input leaves zeros in the workspace, so every pass maps zeros to zeros
  if (!vie::volume_apply<T>(v, xin.as<vie::cx<T>>(), yout.as<vie::cx<T>>(), 1, st)) return 0;
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  const int reps = 10;
  bool ok = true;
  for (int pass = 0; pass < 5 && ok; pass++) {
    int nv = vie::pass_variant_count<T>(p->g, pass);
    float best = 0;
    int best_v = -1;
    for (int var = 0; var < nv && ok; var++) {
      if (!vie::pass_variant_ok<T>(p->g, pass, var)) continue;
      for (int r = 0; r < 2; r++)
        ok = ok && vie::pass_launch<T>(v, pass, var, xin.as<vie::cx<T>>(), yout.as<vie::cx<T>>(), 1, st);
      ok = ok && hipEventRecord(e0, st) == hipSuccess;
      for (int r = 0; r < reps; r++)
        ok = ok && vie::pass_launch<T>(v, pass, var, xin.as<vie::cx<T>>(), yout.as<vie::cx<T>>(), 1, st);
      ok = ok && hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
      float ms = 0;
      ok = ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
      // a later variant must win by 2% to displace an earlier one
      if (ok && (best_v < 0 || ms < 0.98f * best)) {
        best = ms;
        best_v = var;
      }
    }
    p->var[slot][pass] = (unsigned char)best_v;
  }
  // Multi-job launches take shape 1 (the widest tile's elements per thread
  // at 128 threads) where it applies: their grids are the sum of many jobs',
  // so smaller workgroups pack the CUs better than the shapes tuned on a
  // lone job (C4: 2.37 against 2.45 s mixed and 3.80 against 3.92 s FP64;
  // shape 0 measured 2.35 / 3.98, shape 2 2.61 / 4.02, shape 3 2.44 / 4.31).
  // VIE_HIP_MJ_TUNED=1 keeps the solo-tuned shapes (so that batched and solo
  // solves run the same arithmetic, as tests compare them bit for bit).
  memcpy(p->var_mj[slot], p->var[slot], 5);
  if (!mj_tuned)
    for (int pass = 0; pass < 5; pass++)
      if (vie::pass_mj_ok<T>(p->g, pass, 1)) p->var_mj[slot][pass] = 1;
  // the fused z+y passes against the chosen pass pairs (0 + 1 forward,
  // 3 + 4 inverse): the fused one must win by 2% as well (VIE_HIP_FUSE=0
  // keeps the separate passes)
  if (ok && vie::fused_ok(p->g) && fuse) {
    const int code = vie::fused_code(p->g);
    for (int dir = 0; dir < 2 && ok; dir++) {
      const int pa = dir ? 3 : 0, pb = dir ? 4 : 1, fused = dir ? 4 : 0;
      float ms[2] = {0, 0};
      for (int alt = 0; alt < 2 && ok; alt++) {
        auto run = [&]() {
          if (alt == 0) {
            ok = ok && vie::pass_launch<T>(v, pa, p->var[slot][pa], xin.as<vie::cx<T>>(), yout.as<vie::cx<T>>(), 1, st);
            ok = ok && vie::pass_launch<T>(v, pb, p->var[slot][pb], xin.as<vie::cx<T>>(), yout.as<vie::cx<T>>(), 1, st);
          } else {
            ok = ok && vie::pass_launch<T>(v, fused, code, xin.as<vie::cx<T>>(), yout.as<vie::cx<T>>(), 1, st);
          }
        };
        for (int r = 0; r < 2; r++) run();
        ok = ok && hipEventRecord(e0, st) == hipSuccess;
        for (int r = 0; r < reps; r++) run();
        ok = ok && hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
        ok = ok && hipEventElapsedTime(&ms[alt], e0, e1) == hipSuccess;
      }
      if (ok && ms[1] < 0.98f * ms[0]) {
        p->var[slot][fused] = (unsigned char)code;
        p->var[slot][dir ? 3 : 1] = (unsigned char)vie::kFuseSkip;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!ok) {
    set_error("pass tuning failed");
    return 0;
  }
  p->tuned[slot] = true;
  std::array<unsigned char, 10> a;
  memcpy(a.data(), p->var[slot], 5);
  memcpy(a.data() + 5, p->var_mj[slot], 5);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_cache[key] = a;
  return 1;
}

static int plan_tune_prec(vie_fft_plan* p, int prec) {
  if (!p) return 1;
  return prec == VIE_HIP_MIXED ? plan_tune<float>(p) : plan_tune<double>(p);
}

extern "C" int vie_fft_tune(vie_fft_plan* p, int* variants) {
  if (!p || !plan_upload_eps(p) || !plan_tune<double>(p) || !plan_tune<float>(p)) return 0;
  if (variants)
    for (int i = 0; i < 5; i++) {
      variants[i] = p->var[0][i];
      variants[5 + i] = p->var[1][i];
    }
  return 1;
}

extern "C" int vie_hip_ground_table(int nx, int ny, int nz, double h, double z0, double k0, double eps_re,
                                    double eps_im, int order, double _Complex* out) {
  g_last_error.clear();
  if (nx < 1 || ny < 1 || nz < 1 || !(h > 0) || !(z0 - 0.5 * h > 0) || !(k0 > 0) || !std::isfinite(eps_re) ||
      !std::isfinite(eps_im) || eps_re < 1 || eps_im > 0 || order < 8 || order > 256 || !out ||
      (size_t)nx * ny * nz > ((size_t)1 << 26)) {
    set_error("vie_hip_ground_table: invalid arguments");
    return 0;
  }
  if (!ensure_device()) return 0;
  const size_t n = 6 * (size_t)nx * ny * (2 * nz - 1);
  DevBuf t;
  hipStream_t st;
  HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  int ok = t.alloc(n * sizeof(c64)) &&
           vie::ground_table(nx, ny, nz, h, z0, k0, c64{eps_re, eps_im}, order, t.as<c64>(), st) &&
           hipMemcpyAsync(out, t.p, n * sizeof(c64), hipMemcpyDeviceToHost, st) == hipSuccess &&
           hipStreamSynchronize(st) == hipSuccess;
  (void)hipStreamDestroy(st);
  if (!ok && g_last_error.empty()) set_error("vie_hip_ground_table: quadrature or device failure");
  return ok;
}

extern "C" int vie_fft_set_variants(vie_fft_plan* p, const int* variants) {
  if (!p || !variants) return 0;
  for (int i = 0; i < 10; i++) {
    const int pass = i % 5;
    const bool valid = i < 5 ? vie::pass_variant_ok<double>(p->g, pass, variants[i])
                             : vie::pass_variant_ok<float>(p->g, pass, variants[i]);
    if (!valid) {
      set_error("vie_fft_set_variants: variant %d of pass %d does not apply to this %s plan", variants[i], pass,
                i < 5 ? "FP64" : "FP32");
      return 0;
    }
  }
  for (int k = 0; k < 2; k++) {
    unsigned char v[5];
    for (int i = 0; i < 5; i++) v[i] = (unsigned char)variants[5 * k + i];
    if (!vie::variants_consistent(v)) {
      set_error("vie_fft_set_variants: a fused z+y pass needs its y pass skipped, and only then");
      return 0;
    }
  }
  for (int i = 0; i < 5; i++) {
    p->var[0][i] = (unsigned char)variants[i];
    p->var[1][i] = (unsigned char)variants[5 + i];
    for (int k = 0; k < 2; k++)  // multi-job launches run forced fused passes as the separate variant 0
      p->var_mj[k][i] = p->var[k][i] >= vie::kFuseBase ? 0 : p->var[k][i];
  }
  p->tuned[0] = p->tuned[1] = true;
  return 1;
}

// Internal creation shared by vie_fft_create (pec = 0) and the PEC system.
// xyz: box coordinates of the unknown cells; kshape_z: kernel z extent.
// Real ground of a plan: the lower half-space's permittivity and the centre
// height of the lattice's lowest cell layer above it.
struct GroundSpec {
  c64 eps;
  double z0;
};
static const int kGroundOrder = 48;  // Gauss-Legendre points per transverse-wavenumber segment

static vie_fft_plan* plan_create(const size_t box[3], size_t kshape_z, int pec, size_t zoff, size_t cells,
                                 const size_t* xyz, double h, double k, c64 er, size_t limit,
                                 const GroundSpec* ground = nullptr) {
  if (!cells || cells > SIZE_MAX / 3 || !std::isfinite(h) || h <= 0 || !std::isfinite(k) || k < 0 || !passive(er) ||
      !std::isfinite(k * h) || k * h > 1) {
    set_error("invalid plan arguments");
    return nullptr;
  }
  size_t kshape[3] = {box[0], box[1], kshape_z};
  size_t padded[3];
  for (int d = 0; d < 3; d++) {
    if (!kshape[d] || kshape[d] > (1u << 20) || !box[d]) {
      set_error("lattice extent out of range");
      return nullptr;
    }
    padded[d] = pad_extent(kshape[d]);
    if (!vie::max_fft_extent_ok((int)padded[d])) {
      set_error("padded FFT extent %zu exceeds the 1024-point pencil kernels", padded[d]);
      return nullptr;
    }
  }
  vie::Geom g;
  g.nx = (int)box[0];
  g.ny = (int)box[1];
  g.nz = (int)box[2];
  g.Px = (int)padded[0];
  g.Py = (int)padded[1];
  g.Pz = (int)padded[2];
  g.kz = (int)kshape_z;
  g.pec = pec;
  g.zoff = (int)zoff;
  g.ground = ground ? 1 : 0;
  if (ground && (pec || g.Pz < 2 * g.nz - 1)) {
    set_error("a real ground needs a free-space plan");
    return nullptr;
  }
  size_t bytes;
  const size_t rspec_n = ground ? 6 * (size_t)(g.Px / 2 + 1) * (g.Py / 2 + 1) * g.Pz : 0;
  if (!plan_bytes(g, cells, &bytes) || bytes + rspec_n * 24 > limit) {
    set_error("plan storage exceeds the byte limit");
    return nullptr;
  }
  // lattice map with duplicate detection (vie_fft.c:418-425)
  std::vector<size_t> map(cells);
  {
    std::vector<unsigned char> seen(g.nbox(), 0);
    for (size_t c = 0; c < cells; c++) {
      for (int d = 0; d < 3; d++)
        if (xyz[3 * c + d] >= box[d]) {
          set_error("cell outside the lattice");
          return nullptr;
        }
      size_t idx = (xyz[3 * c] * box[1] + xyz[3 * c + 1]) * box[2] + xyz[3 * c + 2];
      if (seen[idx]) {
        set_error("duplicate lattice cell");
        return nullptr;
      }
      seen[idx] = 1;
      map[c] = idx;
    }
  }
  if (!ensure_device()) return nullptr;
  vie_fft_plan* p = new (std::nothrow) vie_fft_plan();
  if (!p) return nullptr;
  p->g = g;
  p->cells = cells;
  p->bytes = bytes;
  p->h = h;
  p->k = k;
  p->map.swap(map);
  p->eps.assign(cells, er);
  bool ok = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) == hipSuccess;
  size_t nbox = g.nbox(), oct = p->oct();
  size_t work = 3 * (size_t)g.nx * g.ny * g.Pz + 3 * (size_t)g.nx * g.Py * g.Pz;
  size_t ntw = (size_t)g.Px + g.Py + g.Pz;
  ok = ok && p->spec64.alloc(6 * oct * sizeof(c64)) && p->spec32.alloc(6 * oct * sizeof(c32)) &&
       p->work.alloc(work * sizeof(c64)) && p->tw64.alloc(ntw * sizeof(c64)) && p->tw32.alloc(ntw * sizeof(c32)) &&
       p->eps64.alloc(nbox * sizeof(c64)) && p->eps32.alloc(nbox * sizeof(c32)) &&
       p->map_dev.alloc(cells * sizeof(size_t));
  if (ok) {
    std::vector<c64> tw;
    twiddles(g.Px, tw);
    twiddles(g.Py, tw);
    twiddles(g.Pz, tw);
    std::vector<c32> tw32(tw.size());
    for (size_t i = 0; i < tw.size(); i++) tw32[i] = c32{(float)tw[i].re, (float)tw[i].im};
    ok = hipMemcpyAsync(p->tw64.p, tw.data(), tw.size() * sizeof(c64), hipMemcpyHostToDevice, p->stream) ==
             hipSuccess &&
         hipMemcpyAsync(p->tw32.p, tw32.data(), tw32.size() * sizeof(c32), hipMemcpyHostToDevice, p->stream) ==
             hipSuccess &&
         hipMemcpyAsync(p->map_dev.p, p->map.data(), cells * sizeof(size_t), hipMemcpyHostToDevice, p->stream) ==
             hipSuccess &&
         hipStreamSynchronize(p->stream) == hipSuccess;
  }
  if (ok) {
    int ks[3] = {g.nx, g.ny, g.kz}, P[3] = {g.Px, g.Py, g.Pz};
    ok = vie::build_spectrum(ks, P, h, k, p->spec64.as<c64>(), p->spec32.as<c32>(), &p->spec_info, p->stream) != 0;
    if (!ok && g_last_error.empty()) set_error("kernel spectrum quadrature failed");
  }
  if (ok && ground) {
    p->ground_eps = ground->eps;
    p->ground_z0 = ground->z0;
    p->bytes += rspec_n * 24;
    int P[3] = {g.Px, g.Py, g.Pz};
    ok = p->rspec64.alloc(rspec_n * sizeof(c64)) && p->rspec32.alloc(rspec_n * sizeof(c32)) &&
         vie::build_ground_spectrum(g.nx, g.ny, g.nz, P, h, ground->z0, k, ground->eps, kGroundOrder,
                                    p->rspec64.as<c64>(), p->rspec32.as<c32>(), p->stream) != 0;
    if (!ok && g_last_error.empty()) set_error("reflected (real-ground) spectrum failed");
    // the x pass must keep each line's kz mirror in its workgroup
    for (int slot = 0; slot < 2 && ok; slot++) {
      int v = 0;
      const int nv = slot ? vie::pass_variant_count<float>(g, 2) : vie::pass_variant_count<double>(g, 2);
      while (v < nv && !(slot ? vie::pass_variant_ok<float>(g, 2, v) : vie::pass_variant_ok<double>(g, 2, v))) v++;
      if (v == nv) {
        set_error("no x-pass launch shape keeps the kz mirrors of this grid together (Pz = %d)", g.Pz);
        ok = false;
      }
      p->var[slot][2] = (unsigned char)v;
    }
  }
  if (ok) {
    for (int i = 0; i < 3; i++) p->self[i] = p->spec_info.self[i];
    // every cell's Jacobi diagonal must be usable (vie_fft.c:438, 450-453)
    for (int i = 0; i < 3 && ok; i++) {
      std::complex<double> e(er.re, er.im), s(p->self[i].re, p->self[i].im), d = e - (e - 1.0) * s;
      if (!std::isfinite(std::abs(d)) || std::abs(d) < 1e-30) ok = false;
    }
    ok = ok && plan_upload_eps(p);
  }
  if (!ok) {
    delete p;
    return nullptr;
  }
  return p;
}

extern "C" int vie_fft_backend_set(const char* name) {
  if (!name || !strcmp(name, "auto")) g_requested = BACKEND_AUTO;
  else if (!strcmp(name, "hip")) g_requested = BACKEND_HIP;
  else return 0;  // "builtin" / "fftw": the reference's CPU backends, not in this library
  return 1;
}
extern "C" int vie_fft_fftw_available(void) { return 0; }
extern "C" const char* vie_fft_backend_requested_name(void) { return g_requested == BACKEND_HIP ? "hip" : "auto"; }
extern "C" const char* vie_fft_backend_name(const vie_fft_plan* p) { return p ? "hip" : "hip"; }
extern "C" const char* vie_fft_backend_last_used_name(void) { return g_last_used < 0 ? nullptr : "hip"; }
extern "C" void vie_fft_backend_clear_last_used(void) { g_last_used = -1; }
extern "C" int vie_fft_parallel_available(void) { return 1; }
extern "C" size_t vie_fft_bytes(const vie_fft_plan* p) { return p ? p->bytes : 0; }
extern "C" size_t vie_fft_threads(const vie_fft_plan* p) { return p ? 1 : 0; }
extern "C" void vie_fft_destroy(vie_fft_plan* p) { delete p; }

extern "C" int vie_fft_create(vie_fft_plan** out, const size_t shape[3], size_t cells, const size_t* xyz, double h,
                              double k, double _Complex er, size_t limit) {
  PhaseTimer phase_timer_(PHASE_BUILD);
  if (!out) return 0;
  *out = nullptr;
  g_last_error.clear();
  if (!shape || !xyz) {
    set_error("invalid plan arguments");
    return 0;
  }
  c64 e = {__real__ er, __imag__ er};
  vie_fft_plan* p = plan_create(shape, shape[2], 0, 0, cells, xyz, h, k, e, limit);
  if (!p) return 0;
  g_last_used = BACKEND_HIP;
  *out = p;
  return 1;
}

extern "C" int vie_fft_set_permittivity(vie_fft_plan* p, size_t cell, double _Complex er) {
  c64 e = {__real__ er, __imag__ er};
  if (!p || cell >= p->cells || !passive(e)) return 0;
  std::complex<double> ec(e.re, e.im);
  for (int i = 0; i < 3; i++) {
    std::complex<double> d = ec - (ec - 1.0) * std::complex<double>(p->self[i].re, p->self[i].im);
    if (!std::isfinite(std::abs(d)) || std::abs(d) < 1e-30) return 0;
  }
  p->eps[cell] = e;
  p->eps_dirty = true;
  return 1;
}

extern "C" int vie_fft_set_permittivities(vie_fft_plan* p, const double _Complex* er, size_t count) {
  if (!p || !er || count != p->cells) return 0;
  const c64* e = (const c64*)er;
  for (size_t i = 0; i < count; i++) {
    if (!passive(e[i])) return 0;
    std::complex<double> ec(e[i].re, e[i].im);
    for (int q = 0; q < 3; q++) {
      std::complex<double> d = ec - (ec - 1.0) * std::complex<double>(p->self[q].re, p->self[q].im);
      if (!std::isfinite(std::abs(d)) || std::abs(d) < 1e-30) return 0;
    }
  }
  std::copy(e, e + count, p->eps.begin());
  p->eps_dirty = true;
  return 1;
}

extern "C" int vie_hip_synchronize(void) { return hipDeviceSynchronize() == hipSuccess; }

extern "C" double vie_hip_copy_probe(size_t bytes, int repeats) {
  if (!ensure_device() || repeats < 1) return -1;
  size_t n16 = bytes / 16;
  DevBuf a, b;
  if (!a.alloc(n16 * 16) || !b.alloc(n16 * 16)) return -1;
  if (hipMemset(a.p, 0, n16 * 16) != hipSuccess) return -1;
  hipEvent_t t0, t1;
  if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess) return -1;
  vie::copy16(b.p, a.p, n16, nullptr);
  (void)hipEventRecord(t0, nullptr);
  for (int r = 0; r < repeats; r++) vie::copy16(b.p, a.p, n16, nullptr);
  (void)hipEventRecord(t1, nullptr);
  (void)hipEventSynchronize(t1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, t0, t1);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  return ms / repeats;
}

// ------------------------------------------------------------------ system
using StatusBlock = vie::StepStatus;

// Host-visible landing area of one solve (pinned, written by async copies).
struct PinnedStatus {
  StatusBlock ring[2];  // status after Arnoldi steps j (slot j & 1)
  StatusBlock fin;      // status after back substitution
  c64 scal[4];          // ||b||^2, ||x||^2, ||r||^2
};

struct dielectric_fft_system;

// One device-resident GMRES solve: its buffers and a resumable state machine
// (every host wait is an event query, so many solves can be interleaved).
struct SolverWork {
  size_t n = 0, ldv = 0;
  int m = 0, prec = -1;
  DevBuf V, x, b, r, ax, state, dots, partial, scal;
  PinnedStatus* pin = nullptr;
  PinnedStatus* pin_dev = nullptr;  // the same memory as the device addresses it
  hipEvent_t ev_ring[2] = {nullptr, nullptr}, ev_misc = nullptr;
  // recorded Arnoldi step (matvec + delayed CGS2 step) of this workspace
  hipGraphExec_t step_exec = nullptr;
  std::vector<const void*> step_key;
  bool graph_off = false;
  c64* host_io = nullptr;  // pinned user-layout staging of a batched solve
  size_t host_io_cap = 0;
  // job state
  dielectric_fft_system* sys = nullptr;
  vie_gmres_options o{};
  vie_hip_solve_options so{};
  vie_gmres_report rep{};
  int phase = 0, m_eff = 0, steps = 0, issued = 0, checked = 0, result = 0;
  bool after_cycle = false, breakdown = false, cycle_done = false, dcgs2 = true;
  double cycle_beta = 0;    // true residual at the start of the current cycle
  bool escalated = false;   // a MIXED solve that switched to FP64 cycles
  bool resid_known = false; // the residual of PH_RESID is already in pin->scal[2]
  bool check_singular = false;  // PH_RESID follows a cycle end: test its back substitution first
  double col_sum = 0;       // sum of basis columns j over the matvec Arnoldi steps issued
  unsigned long long col_steps = 0;
  double bnorm = 0, target = 0, beta = 0, reduction = 1e-5, t_submit = 0, t_end = 0;
  size_t it0 = 0;
  // Recycled subspace.  Basis columns (stride ldv) are laid out as
  // [U (rk) | C (rk) | Krylov]: the first cycle builds its basis from column
  // 2 rk on, U and C are then formed from it in place, and every later cycle
  // sees C as its columns 0..rk-1 (basis pointer at column rk).
  int rk = 0;              // recycled vectors of this solve (0: plain restarted GMRES)
  int rk_req = 0;          // the recycled vectors the solve was set up with
  int cyc_base = 0;        // first Krylov column of the current cycle
  int cycles = 0;          // cycles started in this solve
  int rec_used = 0;        // columns of the first cycle
  bool rec_breakdown = false, rec_ready = false;
  DevBuf rcoef;            // combination coefficients of U and C
  c64* rec_host = nullptr; // pinned: first cycle's Hraw, dc_a, dc_alpha, then the coefficients
  // Multi-job steps (a batch's step groups, by cycle precision: 0 FP64, 1 FP32)
  struct StepGroup* groups[2] = {nullptr, nullptr};
  // Sharing of the recycled subspace among a batch's jobs of one system: the
  // first job whose first cycle ends publishes its U and C; a job of the same
  // system, operator generation, precision and k that starts later adopts
  // them, so its first cycle is already a recycled one.
  bool share = false, adopted = false, published = false;
  unsigned long long op_gen = 0, op_eps_gen = 0;  // the operator this job solves
  bool grouped = false;         // the group issues and checks this cycle's matvec steps
  bool group_nonfinite = false; // a grouped step reported non-finite values
  DevBuf fftws;                 // own FFT workspace (group members run concurrently)
  hipEvent_t ev_join = nullptr; // end of the cycle start, for the group stream
  ~SolverWork() {
    if (step_exec) (void)hipGraphExecDestroy(step_exec);
    if (pin) (void)hipHostFree(pin);
    if (rec_host) (void)hipHostFree(rec_host);
    if (host_io) (void)hipHostFree(host_io);
    for (auto& e : ev_ring)
      if (e) (void)hipEventDestroy(e);
    if (ev_misc) (void)hipEventDestroy(ev_misc);
    if (ev_join) (void)hipEventDestroy(ev_join);
  }
};

struct dielectric_fft_system {
  size_t c = 0, v = 0, cells = 0, bytes = 0;
  bool pec = false;
  bool owns_plan = true;
  // reference layout (dielectric_fft.c:34-40), zero-initialised by calloc so
  // that the pages of a block nobody writes from the host are never touched
  c64* blocks = nullptr;
  vie_fft_plan* plan = nullptr;
  bool blocks_dirty = true;
  // B (rows) / D (columns) assembled on the device: the device copy is the
  // newer one and the host block is stale until a dielectric_fft_entry call
  // brings it back
  bool rows_on_device = false, cols_on_device = false;
  hipStream_t stream = nullptr;  // the plan's stream, or an own one without cells
  bool own_stream = false;
  DevBuf C64, B64, D64, C32, B32, D32, bz, cdiag, stage;
  SolverWork sw;
  bool rhs_loaded = false;
  // Operator generation: bumped whenever the device blocks change (uploads,
  // rows or columns assembled on the device).
  unsigned long long gen = 1;
  // Recycled subspace published by a batch job for the other right-hand
  // sides of this system, per cycle precision (0 FP64, 1 FP32): [U | C],
  // 2 k columns of stride ldv, for the operator generations it was built on.
  struct SharedSubspace {
    DevBuf uc;
    int k = 0;
    size_t ldv = 0;
    unsigned long long gen = 0, eps_gen = 0;
    bool ok = false;
  } shared[2];
  ~dielectric_fft_system() {
    free(blocks);
    if (owns_plan) delete plan;
    if (own_stream && stream) (void)hipStreamDestroy(stream);
  }
  size_t nbox() const { return plan ? plan->nbox() : 0; }
  size_t nsys() const { return c + 3 * nbox(); }  // device system vector length
};

vie_fft_plan::~vie_fft_plan() {
  delete wrap;
  if (stream) (void)hipStreamDestroy(stream);
}

vie_gmres_options dielectric_gmres_settings = {30, 300, 1e-12, 0};

extern "C" size_t dielectric_fft_budget(void) {
  if (!g_budget) {
    const char* e = getenv("VIE_HIP_BUDGET_BYTES");
    g_budget = e ? strtoull(e, nullptr, 10) : ((size_t)64 << 30);
    if (!g_budget) g_budget = (size_t)64 << 30;
  }
  return g_budget;
}
extern "C" void dielectric_fft_set_budget(size_t bytes) { g_budget = bytes; }

static int diag_error(dielectric_diagnostic* d, const char* msg) {
  if (d) {
    memset(d, 0, sizeof(*d));
    snprintf(d->message, sizeof(d->message), "FFT: %s", msg);
  }
  set_error("FFT: %s", msg);
  return 0;
}

// Untabulated and tabulated material permittivity, exp(+i omega t)
// (dielectric.h:36-56).
static c64 material_epsilon(const dielectric_material* m, double k0) {
  double er = m->epsilon_r, ei = m->epsilon_loss, sigma = m->conductivity;
  if (m->tabulated) {
    double f = k0 * 299792458.0 / (2 * 3.14159265358979323846 * 1e6), t;
    size_t j = 1;
    const double eps16 = 16 * 2.220446049250313e-16 * std::max(1.0, fabs(f));
    if (m->sample_count < 2 || !std::isfinite(f)) return c64{NAN, NAN};
    if (f < m->samples[0].frequency_mhz - eps16 || f > m->samples[m->sample_count - 1].frequency_mhz + eps16)
      return c64{NAN, NAN};
    f = std::max(m->samples[0].frequency_mhz, std::min(f, m->samples[m->sample_count - 1].frequency_mhz));
    while (j + 1 < m->sample_count && f > m->samples[j].frequency_mhz) ++j;
    const dielectric_sample *a = &m->samples[j - 1], *b = &m->samples[j];
    t = (f - a->frequency_mhz) / (b->frequency_mhz - a->frequency_mhz);
    er = (1 - t) * a->epsilon_r + t * b->epsilon_r;
    ei = (1 - t) * a->loss + t * b->loss;
    sigma = 0;
    if (a->conductivity_kind) {
      sigma = ei;
      ei = 0;
    }
  }
  return c64{er, -(ei + (sigma == 0 ? 0 : sigma / (k0 * 299792458.0 * 8.8541878128e-12)))};
}

static int prepare(dielectric_fft_system** out, const dielectric_model* model, const dielectric_mesh_result* mesh,
                   size_t conductors, double k0, size_t extra, dielectric_diagnostic* diag, int pec,
                   const c64* ground = nullptr) {
  PhaseTimer phase_timer_(PHASE_BUILD);
  g_last_error.clear();
  *out = nullptr;
  size_t cells = mesh->cell_count, v, n, bytes = sizeof(dielectric_fft_system), blocks = 0, t;
  const size_t m = dielectric_gmres_settings.restart;
  if (cells > SIZE_MAX / 3 || conductors > SIZE_MAX - 3 * cells) return diag_error(diag, "unknown count overflow");
  v = 3 * cells;
  n = v + conductors;
  // storage inventory (dielectric_fft.c:98-108) for this backend: host blocks,
  // FP64 + FP32 device copies, FP64 + FP32 Krylov bases and solver vectors.
  bool ok = mul_ok(conductors, conductors, &t) && add_ok(blocks, t, &blocks) && mul_ok(conductors, v, &t) &&
            add_ok(blocks, 2 * t, &blocks) && mul_ok(blocks ? blocks : 1, 16 + 24, &t) && add_ok(bytes, t, &bytes) &&
            mul_ok(n, (m + 1) * 24 + 5 * 16, &t) && add_ok(bytes, t, &bytes) &&
            mul_ok(cells, sizeof(dielectric_cell), &t) && add_ok(bytes, t, &bytes) && add_ok(bytes, extra, &bytes);
  if (!ok || bytes >= dielectric_fft_budget()) return diag_error(diag, "solver storage exceeds the HIP memory budget");
  std::vector<size_t> xyz;
  size_t shape[3] = {1, 1, 1};
  double origin[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, h = 0;
  c64 er = {0, 0};
  std::vector<c64> cell_eps(cells);
  std::unordered_map<int, size_t> by_id;
  for (size_t q = model->material_count; q-- > 0;) by_id[model->materials[q].id] = q;
  if (cells) {
    h = mesh->cells[0].size[0];
    for (size_t i = 0; i < cells; i++) {
      c64 value = {0, 0};
      bool found = false;
      if (pec && mesh->cells[i].center[2] - .5 * mesh->cells[i].size[2] < 0)
        return diag_error(diag, "PEC FFT cells must lie in z >= 0");
      if (ground && !(mesh->cells[i].center[2] - .5 * mesh->cells[i].size[2] > 0))
        return diag_error(diag, "real ground requires strictly positive clearance for every volume cell");
      auto it = by_id.find(mesh->cells[i].material_id);
      if (it != by_id.end()) {
        value = material_epsilon(&model->materials[it->second], k0);
        found = true;
      }
      if (!found || (value.re == 0 && value.im == 0) || !passive(value))
        return diag_error(diag, "unknown or invalid passive material");
      cell_eps[i] = value;
      if (i == 0) er = value;
      for (int d = 0; d < 3; d++) {
        if (!std::isfinite(h) || h <= 0 || !std::isfinite(mesh->cells[i].size[d]) ||
            !std::isfinite(mesh->cells[i].center[d]) || fabs(mesh->cells[i].size[d] - h) > 1e-10 * h)
          return diag_error(diag, "requires equal cubic cells");
        origin[d] = std::min(origin[d], mesh->cells[i].center[d]);
      }
    }
    xyz.resize(3 * cells);
    for (size_t i = 0; i < cells; i++)
      for (int d = 0; d < 3; d++) {
        double index = (mesh->cells[i].center[d] - origin[d]) / h, nearest = round(index);
        if (!std::isfinite(index) || nearest < 0 || nearest >= (double)(SIZE_MAX / 2) || fabs(index - nearest) > 1e-8)
          return diag_error(diag, "cell centers do not share a common lattice");
        xyz[3 * i + d] = (size_t)nearest;
        if (xyz[3 * i + d] >= shape[d]) shape[d] = xyz[3 * i + d] + 1;
      }
  }
  size_t kz = shape[2], zoff = 0;
  if (pec && cells) {
    // image embedding (dielectric_fft.c:137-154)
    double gap_index = 2 * (origin[2] - .5 * h) / h, nearest = round(gap_index);
    if (!std::isfinite(gap_index) || nearest < 0 || nearest >= (double)(SIZE_MAX / 4) ||
        fabs(gap_index - nearest) > 1e-8)
      return diag_error(diag, "PEC FFT requires nonnegative bottom gap on the half-cell lattice");
    size_t gap = (size_t)nearest;
    if (shape[2] > SIZE_MAX / 2 || gap > SIZE_MAX - 2 * shape[2]) return diag_error(diag, "PEC image embedding overflow");
    kz = 2 * shape[2] + gap;
    zoff = shape[2] + gap;
  }
  dielectric_fft_system* s = new (std::nothrow) dielectric_fft_system();
  if (!s) return diag_error(diag, "system allocation failed");
  s->c = conductors;
  s->v = v;
  s->cells = cells;
  s->pec = pec != 0;
  s->blocks = (c64*)calloc(blocks ? blocks : 1, sizeof(c64));
  if (!s->blocks) {
    delete s;
    return diag_error(diag, "conductor block allocation failed");
  }
  if (cells) {
    size_t limit = dielectric_fft_budget() - bytes;
    GroundSpec gs{ground ? *ground : c64{1, 0}, origin[2]};
    s->plan = plan_create(shape, kz, pec, zoff, cells, xyz.data(), h, k0, er, limit, ground ? &gs : nullptr);
    if (!s->plan) {
      std::string why = g_last_error;
      delete s;
      std::string msg = "kernel/lattice build failed or padded FFT exceeds memory budget";
      if (!why.empty()) msg += " (" + why + ")";
      return diag_error(diag, msg.c_str());
    }
    if (!vie_fft_set_permittivities(s->plan, (const double _Complex*)cell_eps.data(), cells)) {
      delete s;
      return diag_error(diag, "kernel/lattice build failed or padded FFT exceeds memory budget");
    }
    s->stream = s->plan->stream;
  } else {
    if (!ensure_device()) {
      delete s;
      return diag_error(diag, "no HIP device available");
    }
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
      delete s;
      return diag_error(diag, "stream creation failed");
    }
    s->own_stream = true;
  }
  s->bytes = bytes + (s->plan ? s->plan->bytes : 0);
  g_last_used = BACKEND_HIP;
  *out = s;
  return 1;
}

extern "C" int dielectric_fft_prepare(dielectric_fft_system** out, const dielectric_model* model,
                                      const dielectric_mesh_result* mesh, size_t conductors, double k0, size_t extra,
                                      dielectric_diagnostic* diag) {
  return prepare(out, model, mesh, conductors, k0, extra, diag, 0);
}
extern "C" int dielectric_fft_prepare_pec(dielectric_fft_system** out, const dielectric_model* model,
                                          const dielectric_mesh_result* mesh, size_t conductors, double k0,
                                          size_t extra, dielectric_diagnostic* diag) {
  return prepare(out, model, mesh, conductors, k0, extra, diag, 1);
}
extern "C" int dielectric_fft_prepare_ground(dielectric_fft_system** out, const dielectric_model* model,
                                             const dielectric_mesh_result* mesh, size_t conductors, double k0,
                                             size_t extra, double ground_re, double ground_im,
                                             dielectric_diagnostic* diag) {
  PhaseTimer phase_timer_(PHASE_BUILD);
  if (out) *out = nullptr;
  if (!std::isfinite(ground_re) || !std::isfinite(ground_im) || ground_re < 1 || ground_im > 0)
    return diag_error(diag, "real ground requires a passive finite permittivity (Re >= 1, Im <= 0)");
  const c64 e{ground_re, ground_im};
  return prepare(out, model, mesh, conductors, k0, extra, diag, 0, &e);
}
extern "C" void dielectric_fft_free(dielectric_fft_system* s) { delete s; }
extern "C" size_t dielectric_fft_storage(const dielectric_fft_system* s) { return s ? s->bytes : 0; }
extern "C" size_t dielectric_fft_threads(const dielectric_fft_system* s) { return s ? 1 : 0; }
extern "C" const char* dielectric_fft_backend(const dielectric_fft_system* s) { return s ? "hip" : "none"; }
extern "C" vie_fft_plan* dielectric_fft_plan(dielectric_fft_system* s) { return s ? s->plan : nullptr; }
extern "C" void dielectric_fft_blocks_changed(dielectric_fft_system* s) {
  if (s) s->blocks_dirty = true;
}

// Brings device-assembled B / D blocks back into the host storage (reference
// layout) and makes the host copy the one that uploads again.
static int sync_cross_to_host(dielectric_fft_system* s) {
  if (!s->rows_on_device && !s->cols_on_device) return 1;
  const size_t c = s->c, v = s->v, nbox = s->nbox(), nv = 3 * nbox;
  HIP_OK(hipStreamSynchronize(s->stream));
  for (int which = 0; which < 2; which++) {
    if (which == 0 ? !s->rows_on_device : !s->cols_on_device) continue;
    std::vector<c64> dev(c * nv);
    HIP_OK(hipMemcpy(dev.data(), which == 0 ? s->B64.p : s->D64.p, c * nv * 16, hipMemcpyDeviceToHost));
    for (size_t cell = 0; cell < s->cells; cell++)
      for (int comp = 0; comp < 3; comp++) {
        const size_t bi = comp * nbox + s->plan->map[cell], ui = 3 * cell + comp;
        for (size_t r = 0; r < c; r++) {
          if (which == 0) s->blocks[r * (c + v) + c + ui] = dev[r * nv + bi];
          else s->blocks[c * (c + v) + ui * c + r] = dev[r * nv + bi];
        }
      }
  }
  s->rows_on_device = s->cols_on_device = false;
  return 1;
}

extern "C" double _Complex* dielectric_fft_entry(dielectric_fft_system* s, size_t row, size_t col) {
  size_t c = s->c, v = s->v;
  // any host access may write through the returned pointer into B or D
  if (!sync_cross_to_host(s)) return nullptr;
  s->blocks_dirty = true;
  if (row < c) return (double _Complex*)&s->blocks[row * (c + v) + col];
  if (col < c) return (double _Complex*)&s->blocks[c * (c + v) + (row - c) * c + col];
  return nullptr;
}

// Reorders C, B, D into device layouts (B rows and D columns in box order)
// and keeps FP64 and FP32 copies.
static int upload_blocks(dielectric_fft_system* s) {
  if (!s->blocks_dirty) return 1;
  PhaseTimer phase_timer_(PHASE_UPLOAD);
  const size_t c = s->c, v = s->v, nbox = s->nbox(), nv = 3 * nbox;
  hipStream_t st = s->stream;
  if (c) {
    std::vector<c64> C(c * c);
    for (size_t r = 0; r < c; r++)
      for (size_t q = 0; q < c; q++) C[r * c + q] = s->blocks[r * (c + v) + q];
    if (!(s->C64.alloc(c * c * 16) && s->C32.alloc(c * c * 8) &&
          s->bz.alloc(c * 16) && s->cdiag.alloc(c * 16)))
      return 0;
    std::vector<c64> cd(c);
    for (size_t r = 0; r < c; r++) {
      cd[r] = C[r * c + r];
      if (!std::isfinite(hypot(cd[r].re, cd[r].im)) || hypot(cd[r].re, cd[r].im) < DBL_MIN) {
        set_error("conductor diagonal unusable for the preconditioner");
        return 0;
      }
    }
    HIP_OK(hipMemcpyAsync(s->C64.p, C.data(), c * c * 16, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(s->cdiag.p, cd.data(), c * 16, hipMemcpyHostToDevice, st));
    vie::vec_cast<float, double>(s->C32.as<c32>(), s->C64.as<c64>(), c * c, st);
    if (nv) {
      if (!(s->B64.alloc(c * nv * 16) && s->D64.alloc(c * nv * 16) && s->B32.alloc(c * nv * 8) &&
            s->D32.alloc(c * nv * 8)))
        return 0;
      // a block assembled on the device stays as it is there; a host block
      // goes up in the reference layout and is placed into the box order on
      // the device
      const bool up_b = !s->rows_on_device, up_d = !s->cols_on_device;
      if ((up_b || up_d) && !s->stage.alloc(c * v * 16)) return 0;
      const size_t* map = s->plan->map_dev.as<size_t>();
      if (up_b) {
        HIP_OK(hipMemcpy2DAsync(s->stage.p, v * 16, s->blocks + c, (c + v) * 16, v * 16, c, hipMemcpyHostToDevice,
                                st));
        HIP_OK(hipMemsetAsync(s->B64.p, 0, c * nv * 16, st));
        for (size_t r = 0; r < c; r++)
          vie::to_box<double>(s->stage.as<c64>() + r * v, 0, s->cells, map, nbox, s->B64.as<c64>() + r * nv, st);
        vie::vec_cast<float, double>(s->B32.as<c32>(), s->B64.as<c64>(), c * nv, st);
        HIP_OK(hipStreamSynchronize(st));  // the staging buffer is read before it is refilled
      }
      if (up_d) {
        HIP_OK(hipMemcpyAsync(s->stage.p, s->blocks + c * (c + v), c * v * 16, hipMemcpyHostToDevice, st));
        HIP_OK(hipMemsetAsync(s->D64.p, 0, c * nv * 16, st));
        vie::columns_to_box(s->stage.as<c64>(), c, s->cells, map, nbox, s->D64.as<c64>(), st);
        vie::vec_cast<float, double>(s->D32.as<c32>(), s->D64.as<c64>(), c * nv, st);
      }
    }
    HIP_OK(hipStreamSynchronize(st));
  }
  s->blocks_dirty = false;
  s->gen++;
  return 1;
}

// ------------------------------------------------------------------ operators on device vectors
template <typename T> struct Prec;
template <> struct Prec<double> {
  static DevBuf& C(dielectric_fft_system* s) { return s->C64; }
  static DevBuf& B(dielectric_fft_system* s) { return s->B64; }
  static DevBuf& D(dielectric_fft_system* s) { return s->D64; }
  static DevBuf& eps(vie_fft_plan* p) { return p->eps64; }
  static constexpr double eps_b = 32 * 2.220446049250313e-16;
};
template <> struct Prec<float> {
  static DevBuf& C(dielectric_fft_system* s) { return s->C32; }
  static DevBuf& B(dielectric_fft_system* s) { return s->B32; }
  static DevBuf& D(dielectric_fft_system* s) { return s->D32; }
  static DevBuf& eps(vie_fft_plan* p) { return p->eps32; }
  static constexpr double eps_b = 32 * 1.1920928955078125e-07;
};

// y = A z, z = precond ? M^{-1} x : x, on system vectors [c | 3][nbox].
template <typename T>
static int sys_apply(dielectric_fft_system* s, const vie::cx<T>* x, vie::cx<T>* y, int precond,
                     vie::ReduceScratch& rs, hipEvent_t* ev = nullptr, vie::DevIndex di = {}) {
  const size_t c = s->c;
  if (s->plan) {
    vie::VolumeView<T> vv = s->plan->view<T>();
    if (!vie::volume_apply<T>(vv, x + c, y + c, precond, s->stream, ev, di)) {
      set_error("volume kernels failed to launch");
      return 0;
    }
  }
  if (c) {
    vie::CouplingView<T> cv;
    cv.c = (int)c;
    cv.nbox = s->nbox();
    cv.nv = 3 * cv.nbox;
    cv.C = Prec<T>::C(s).template as<vie::cx<T>>();
    cv.B = Prec<T>::B(s).template as<vie::cx<T>>();
    cv.D = Prec<T>::D(s).template as<vie::cx<T>>();
    cv.bz = s->bz.as<c64>();
    cv.eps = s->plan ? Prec<T>::eps(s->plan).template as<vie::cx<T>>() : nullptr;
    for (int i = 0; i < 3; i++)
      cv.self[i] = s->plan ? vie::cx<T>{(T)s->plan->self[i].re, (T)s->plan->self[i].im} : vie::cx<T>{0, 0};
    vie::coupling_apply<T>(cv, x, x + c, y, y + c, precond, rs, s->stream, di);
    if (ev) HIP_OK(hipEventRecord(ev[6], s->stream));
  }
  return hipGetLastError() == hipSuccess;
}

// sys_apply with the volume passes on a caller-given FFT workspace
// ([3][nx][ny][Pz] then [3][nx][Py][Pz]) instead of the plan's own.
template <typename T>
static int sys_apply_ws(dielectric_fft_system* s, const vie::cx<T>* x, vie::cx<T>* y, int precond,
                        vie::ReduceScratch& rs, vie::cx<T>* work) {
  const size_t c = s->c;
  if (s->plan) {
    vie::VolumeView<T> vv = s->plan->view<T>();
    vv.A = work;
    vv.B = work + 3 * (size_t)vv.g.nx * vv.g.ny * vv.g.Pz;
    if (!vie::volume_apply<T>(vv, x + c, y + c, precond, s->stream)) {
      set_error("volume kernels failed to launch");
      return 0;
    }
  }
  if (c) {
    vie::CouplingView<T> cv;
    cv.c = (int)c;
    cv.nbox = s->nbox();
    cv.nv = 3 * cv.nbox;
    cv.C = Prec<T>::C(s).template as<vie::cx<T>>();
    cv.B = Prec<T>::B(s).template as<vie::cx<T>>();
    cv.D = Prec<T>::D(s).template as<vie::cx<T>>();
    cv.bz = s->bz.as<c64>();
    cv.eps = s->plan ? Prec<T>::eps(s->plan).template as<vie::cx<T>>() : nullptr;
    for (int i = 0; i < 3; i++)
      cv.self[i] = s->plan ? vie::cx<T>{(T)s->plan->self[i].re, (T)s->plan->self[i].im} : vie::cx<T>{0, 0};
    vie::coupling_apply<T>(cv, x, x + c, y, y + c, precond, rs, s->stream, {});
  }
  return hipGetLastError() == hipSuccess;
}

static vie::JacobiView jacobi_view(dielectric_fft_system* s) {
  vie::JacobiView J;
  J.c = s->c;
  J.nbox = s->nbox();
  J.cdiag = s->cdiag.as<c64>();
  J.eps = s->plan ? s->plan->eps64.as<c64>() : nullptr;
  for (int i = 0; i < 3; i++) J.self[i] = s->plan ? s->plan->self[i] : c64{0, 0};
  return J;
}

static int ensure_scratch(SolverWork& w) {
  if (!w.partial.alloc(vie::kReduceScratchBytes)) return 0;
  if (!w.scal.alloc(16 * sizeof(c64))) return 0;
  if (!w.pin) {
    // coherent: the step kernels write the status ring straight into it
    HIP_OK(hipHostMalloc((void**)&w.pin, sizeof(PinnedStatus), hipHostMallocMapped | hipHostMallocCoherent));
    HIP_OK(hipHostGetDevicePointer((void**)&w.pin_dev, w.pin, 0));
    for (auto& e : w.ev_ring) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&w.ev_misc, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&w.ev_join, hipEventDisableTiming));
  }
  return 1;
}

// Pinned host area of the recycled-subspace construction (c64 entries).
static size_t rec_host_entries() {
  const size_t m = vie::kMaxRestart;
  return (m + 1) * m + m + 1 + 2 * (m + 1) * vie::kMaxRecycle;
}

// Host and device areas of the recycled-subspace construction.
static int ensure_recycle(SolverWork& w) {
  if (!w.rec_host)
    HIP_OK(hipHostMalloc((void**)&w.rec_host, rec_host_entries() * sizeof(c64), hipHostMallocDefault));
  return w.rcoef.alloc(2 * (vie::kMaxRestart + 1) * vie::kMaxRecycle * sizeof(c64));
}

static int ensure_solver(dielectric_fft_system* s, SolverWork& w, int m, int prec, int rk = 0) {
  size_t n = s->nsys();
  size_t ldv = vie::basis_ld(n);
  size_t es = prec == VIE_HIP_MIXED ? sizeof(c32) : sizeof(c64);
  if (!ensure_scratch(w)) return 0;
  if (rk > 0 && !ensure_recycle(w)) return 0;
  // a larger basis from an earlier reserve or FP64 solve is reused as is
  if (!(w.V.alloc((size_t)(m + 1 + 2 * rk) * ldv * es) && w.x.alloc(n * 16) && w.b.alloc(n * 16) &&
        w.r.alloc(n * 16) &&
        w.ax.alloc(n * 16) && w.state.alloc(sizeof(vie::ArnoldiState)) &&
        w.dots.alloc(2 * (vie::kMaxRestart + 2) * sizeof(c64))))
    return 0;
  w.sys = s;
  w.n = n;
  w.ldv = ldv;
  w.m = m;
  w.prec = prec;
  w.rk = rk;
  return 1;
}

static size_t solver_bytes(size_t n, int m, int prec, int rk = 0) {
  size_t ldv = vie::basis_ld(n), es = prec == VIE_HIP_MIXED ? sizeof(c32) : sizeof(c64);
  return (size_t)(m + 1 + 2 * rk) * ldv * es + 4 * n * 16 + vie::kReduceScratchBytes + sizeof(vie::ArnoldiState);
}

// Recycled vectors a solve keeps: the option's k (0: default per precision,
// -1: none), capped at kMaxRecycle and at half the restart so every later
// cycle still has a Krylov space of its own.
constexpr int kDefaultRecycleMixed = 16;
static int recycle_count(const vie_hip_solve_options* so, int prec, int m) {
  int k = so ? so->recycle : 0;
  if (k < 0) return 0;
  if (k == 0) k = prec == VIE_HIP_MIXED ? kDefaultRecycleMixed : 0;
  if (k > vie::kMaxRecycle) k = vie::kMaxRecycle;
  if (k > m / 2) k = m / 2;
  return k;
}

// ------------------------------------------------------------------ device-resident GMRES
// The restarted GMRES of vie_iterative.c:80-199 with every vector operation on
// the GPU.  T is the storage / matvec precision of the Krylov cycle; the
// solution, right-hand side, true residual and all reductions are FP64.
// FP64: the reference's algorithm (reorthogonalised Gram-Schmidt, Givens,
// true residual acceptance).  MIXED: the same outer loop with FP32 cycles that
// stop once |g| <= max(target, inner_reduction * beta); each restart's FP64
// true residual refines the solution.
enum { PH_IDLE, PH_NORMS, PH_RESID, PH_CYCLE, PH_FINISHED };

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void job_finish(SolverWork& w, vie_iterative_status st, int ok) {
  w.rep.status = st;
  w.result = ok;
  w.phase = PH_FINISHED;
  w.t_end = now_s();
}

// 1: event complete; 0: still running (non-blocking); -1: runtime error.
static int job_ready(hipEvent_t e, bool blocking) {
  hipError_t r = blocking ? hipEventSynchronize(e) : hipEventQuery(e);
  if (r == hipSuccess) return 1;
  if (r == hipErrorNotReady) return 0;
  set_error("event wait failed: %s", hipGetErrorString(r));
  return -1;
}

static int job_issue_residual(SolverWork& w) {
  dielectric_fft_system* s = w.sys;
  vie::ReduceScratch rs;
  rs.partial = w.partial.as<c64>();
  w.rep.operator_applications++;
  if (!sys_apply<double>(s, w.x.as<c64>(), w.ax.as<c64>(), 0, rs)) return 0;
  vie::residual(w.b.as<c64>(), w.ax.as<c64>(), w.r.as<c64>(), w.n, rs, w.scal.as<c64>() + 2, s->stream);
  HIP_OK(hipMemcpyAsync(&w.pin->scal[2], w.scal.as<c64>() + 2, sizeof(c64), hipMemcpyDeviceToHost, s->stream));
  HIP_OK(hipEventRecord(w.ev_misc, s->stream));
  return 1;
}

// Queues the first work of a solve whose b and x are staged in w (x_zero:
// the caller has just zeroed x, so its norm is not computed).
static int job_start(SolverWork& w, const vie_gmres_options* o, const vie_hip_solve_options* so,
                     bool x_zero = false) {
  dielectric_fft_system* s = w.sys;
  memset(&w.rep, 0, sizeof(w.rep));
  w.rep.status = VIE_ITERATIVE_INVALID_INPUT;
  w.rep.residual_norm = w.rep.relative_residual = INFINITY;
  w.o = *o;
  w.so = *so;
  const char* ortho = getenv("VIE_HIP_ORTHO");
  w.dcgs2 = !(ortho && !strcmp(ortho, "cgs2"));
  if (!w.dcgs2) w.rk = 0;  // the recycled cycles are built on the delayed-CGS2 step
  w.rk_req = w.rk;
  // with a recycled subspace a restart loses little, so FP32 cycles may end earlier
  w.reduction = so->inner_reduction > 0 ? so->inner_reduction : (w.rk ? 1e-4 : 1e-5);
  w.cyc_base = w.cycles = w.rec_used = 0;
  w.rec_breakdown = w.rec_ready = false;
  w.after_cycle = w.breakdown = w.cycle_done = w.escalated = w.resid_known = w.check_singular = false;
  w.cycle_beta = INFINITY;
  w.result = 0;
  w.t_submit = now_s();
  vie::ReduceScratch rs;
  rs.partial = w.partial.as<c64>();
  int mm = w.m;
  HIP_OK(hipMemcpyAsync((char*)w.state.p + offsetof(vie::ArnoldiState, m), &mm, sizeof(int), hipMemcpyHostToDevice,
                        s->stream));
  vie::norm2<double>(w.b.as<c64>(), w.n, rs, w.scal.as<c64>() + 0, s->stream);
  if (x_zero) {
    w.pin->scal[1] = c64{0, 0};
    HIP_OK(hipMemcpyAsync(&w.pin->scal[0], w.scal.p, sizeof(c64), hipMemcpyDeviceToHost, s->stream));
  } else {
    vie::norm2<double>(w.x.as<c64>(), w.n, rs, w.scal.as<c64>() + 1, s->stream);
    HIP_OK(hipMemcpyAsync(&w.pin->scal[0], w.scal.p, 2 * sizeof(c64), hipMemcpyDeviceToHost, s->stream));
  }
  HIP_OK(hipEventRecord(w.ev_misc, s->stream));
  w.phase = PH_NORMS;
  return 1;
}

// VIE_HIP_TRACE=1: one line per GMRES cycle on stderr (diagnostics).
static bool trace_enabled() {
  static int on = -1;
  if (on < 0) {
    const char* e = getenv("VIE_HIP_TRACE");
    on = e && !strcmp(e, "1");
  }
  return on == 1;
}

// Replaying a recorded step (VIE_HIP_GRAPH=1) is opt-in: on the workloads
// measured here the device, not launch issue, bounds the step rate, and the
// direct launches run the same kernels.
static bool graphs_enabled() {
  const char* e = getenv("VIE_HIP_GRAPH");
  return e && !strcmp(e, "1");
}

// Column 0 of the current cycle's basis (the layout of SolverWork::rk).
template <typename T> static vie::cx<T>* cycle_basis(SolverWork& w) {
  vie::cx<T>* V = w.V.as<vie::cx<T>>();
  if (!w.rk) return V;
  return V + (size_t)(w.rec_ready ? w.rk : 2 * w.rk) * w.ldv;
}

// Everything a recorded step depends on: if any of it changes, record again.
template <typename T> static std::vector<const void*> step_graph_key(SolverWork& w) {
  dielectric_fft_system* s = w.sys;
  std::vector<const void*> k = {s, (const void*)(uintptr_t)sizeof(T), cycle_basis<T>(w), w.partial.p, w.dots.p,
                                w.state.p,
                                w.pin_dev, (const void*)(uintptr_t)w.n, (const void*)(uintptr_t)w.ldv,
                                (const void*)(uintptr_t)w.m, s->stream, s->C64.p, s->B64.p, s->D64.p, s->C32.p,
                                s->B32.p, s->D32.p, s->bz.p};
  if (s->plan) {
    vie_fft_plan* p = s->plan;
    const void* more[] = {p->work.p, p->spec64.p, p->spec32.p, p->eps64.p, p->eps32.p, p->tw64.p, p->tw32.p};
    k.insert(k.end(), more, more + 7);
    const int slot = sizeof(T) == sizeof(double) ? 0 : 1;
    for (int i = 0; i < 5; i++) k.push_back((const void*)(uintptr_t)p->var[slot][i]);
  }
  return k;
}

// The launch sequence of one matvec Arnoldi step; j is read on the device.
template <typename T> static int issue_matvec_step(SolverWork& w) {
  dielectric_fft_system* s = w.sys;
  vie::ReduceScratch rs;
  rs.partial = w.partial.as<c64>();
  vie::cx<T>* V = cycle_basis<T>(w);
  vie::ArnoldiState* sd = w.state.as<vie::ArnoldiState>();
  vie::DevIndex di;
  di.j = &sd->j;
  di.stride = w.ldv;
  di.skip = &sd->done;
  if (!sys_apply<T>(s, V, V + w.ldv, 1, rs, nullptr, di)) return 0;
  vie::dcgs2_step<T>(V, w.ldv, true, w.n, sd, rs, w.dots.as<c64>(), Prec<T>::eps_b, w.m, w.pin_dev->ring,
                     s->stream);
  return hipGetLastError() == hipSuccess;
}

// Records the matvec step once per workspace / system / precision.
template <typename T> static int ensure_step_graph(SolverWork& w) {
  std::vector<const void*> key = step_graph_key<T>(w);
  if (w.step_exec && key == w.step_key) return 1;
  if (w.step_exec) (void)hipGraphExecDestroy(w.step_exec);
  w.step_exec = nullptr;
  hipStream_t st = w.sys->stream;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return 0;
  int ok = issue_matvec_step<T>(w);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(st, &g);
  if (!ok || e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return 0;
  }
  e = hipGraphInstantiate(&w.step_exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    w.step_exec = nullptr;
    (void)hipGetLastError();
    return 0;
  }
  w.step_key.swap(key);
  if (trace_enabled()) fprintf(stderr, "[vie_hip] step graph recorded\n");
  return 1;
}

// One Arnoldi step j of the current cycle, followed by its status.
template <typename T> static int job_issue_step(SolverWork& w, int j) {
  dielectric_fft_system* s = w.sys;
  hipStream_t st = s->stream;
  vie::ReduceScratch rs;
  rs.partial = w.partial.as<c64>();
  vie::cx<T>* V = cycle_basis<T>(w);
  vie::ArnoldiState* sd = w.state.as<vie::ArnoldiState>();
  c64* dots1 = w.dots.as<c64>();
  c64* dots2 = dots1 + vie::kMaxRestart + 2;
  const size_t ldv = w.ldv, n = w.n;
  vie::cx<T>* vj = V + (size_t)j * ldv;
  vie::cx<T>* wv = V + (size_t)(j + 1) * ldv;
  if (w.dcgs2) {
    // delayed CGS2: m_eff matvec steps (recorded once, replayed), then one
    // closing step that only finalises the last column
    if (j < w.m_eff) {
      w.rep.operator_applications++;
      w.col_sum += j + w.cyc_base;
      w.col_steps++;
      if (!w.graph_off && graphs_enabled()) {
        if (!ensure_step_graph<T>(w)) {
          w.graph_off = true;  // record failed: launch directly from now on
          if (trace_enabled()) fprintf(stderr, "[vie_hip] step graph could not be recorded: direct launches\n");
        }
      }
      if (w.step_exec && !w.graph_off && graphs_enabled()) {
        HIP_OK(hipGraphLaunch(w.step_exec, st));
      } else if (!issue_matvec_step<T>(w)) {
        return 0;
      }
    } else {
      vie::dcgs2_step<T>(V, ldv, false, n, sd, rs, dots1, Prec<T>::eps_b, w.m, w.pin_dev->ring, st);
    }
  } else {
    if (!sys_apply<T>(s, vj, wv, 1, rs)) return 0;
    w.rep.operator_applications++;
    // two classical Gram-Schmidt passes (same role as the reference's two
    // MGS passes, vie_iterative.c:151-159); ||w|| before and after
    vie::multidot<T>(V, ldv, j + 1, wv, n, 1, rs, dots1, st);
    vie::update<T>(V, ldv, j + 1, dots1, wv, n, 0, rs, st);
    vie::multidot<T>(V, ldv, j + 1, wv, n, 0, rs, dots2, st);
    unsigned nb = vie::update<T>(V, ldv, j + 1, dots2, wv, n, 1, rs, st);
    vie::arnoldi_tail(sd, dots1, dots2, dots1 + j + 1, rs, nb, Prec<T>::eps_b, st);
    vie::scale_next<T>(wv, n, sd, st);
    HIP_OK(hipMemcpyAsync(&w.pin->ring[j & 1], sd, sizeof(StatusBlock), hipMemcpyDeviceToHost, st));
  }
  HIP_OK(hipEventRecord(w.ev_ring[j & 1], st));
  return 1;
}

// ---------------------------------------------------------------- multi-job steps
// A batch advances the Arnoldi steps of its jobs in groups, one per cycle
// precision: a job whose cycle starts joins its group, and every group step
// is one launch per kernel over all members -- per pass and line length for
// the FFT passes (vie::volume_apply_mj), then the coupling and delayed-CGS2
// kernels (vie::step_mj) -- instead of eleven launches per job.  Each job
// keeps its own state machine: the group only issues the matvec steps of
// running cycles and reads their statuses.  A job leaves when its cycle is
// marked done, when it needs the closing step of a full cycle (its own
// launch), or on non-finite values; it rejoins at its next cycle.  Members
// run concurrently, so each uses its own FFT workspace.  Real-ground
// operators, more than kMaxKeys conductors, the plain-CGS2 step and recorded
// steps keep the single-job path.  VIE_HIP_MULTIJOB=0 turns grouping off.
static bool multijob_enabled() {
  static int on = -1;
  if (on < 0) {
    const char* e = getenv("VIE_HIP_MULTIJOB");
    on = !(e && !strcmp(e, "0"));
  }
  return on == 1;
}

struct StepGroup {
  bool fp32 = false;
  hipStream_t stream = nullptr;
  std::vector<SolverWork*> members;    // joined jobs
  std::vector<SolverWork*> table;      // jobs of the device tables, in table order
  std::vector<SolverWork*> stepped[2]; // jobs of the last two issued steps
  std::vector<hipEvent_t> waits;       // cycle starts the next step waits for
  hipEvent_t ev[2] = {nullptr, nullptr};
  int issued = 0, checked = 0;
  DevBuf dev;  // ops | blocks | lists | steps
  struct Stage {
    char* host = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
    bool used = false;
  };
  Stage stage[3];
  int stage_next = 0;
  std::vector<vie::MjClass> classes;
  unsigned total_nb[5] = {0, 0, 0, 0, 0};  // workgroups of the five step kernels
  size_t off_blocks = 0, off_lists = 0, off_steps = 0, off_pre = 0, step_pre = 0;
  unsigned long long launches = 0, steps = 0;  // group steps issued, jobs x steps covered
  bool stale = true;  // a member joined: its cycle basis may have moved (recycled columns)
  int init() {
    HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    for (auto& e : ev) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto& st : stage) HIP_OK(hipEventCreateWithFlags(&st.done, hipEventDisableTiming));
    return 1;
  }
  ~StepGroup() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& st : stage) {
      if (st.host) (void)hipHostFree(st.host);
      if (st.done) (void)hipEventDestroy(st.done);
    }
    if (stream) (void)hipStreamDestroy(stream);
  }
};

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

template <typename T> static bool mj_eligible(const SolverWork& w) {
  const dielectric_fft_system* s = w.sys;
  if (!(w.groups[sizeof(T) == sizeof(float)] && multijob_enabled() && w.dcgs2 && !graphs_enabled() && s->plan &&
        !s->plan->g.ground && s->c <= (size_t)vie::kMaxKeys))
    return false;
  const int slot = sizeof(T) == sizeof(double) ? 0 : 1;
  for (int p = 0; p < 5; p++)
    if (!vie::pass_mj_ok<T>(s->plan->g, p, s->plan->var_mj[slot][p])) return false;
  return true;
}

// Called when a cycle of w has been set up on its system stream.
template <typename T> static int group_join(SolverWork& w) {
  StepGroup& g = *w.groups[sizeof(T) == sizeof(float)];
  const vie::Geom& ge = w.sys->plan->g;
  const size_t ws = (3 * (size_t)ge.nx * ge.ny * ge.Pz + 3 * (size_t)ge.nx * ge.Py * ge.Pz) * sizeof(vie::cx<T>);
  if (!w.fftws.alloc(ws)) return 0;
  HIP_OK(hipEventRecord(w.ev_join, w.sys->stream));
  g.waits.push_back(w.ev_join);
  w.grouped = true;
  w.group_nonfinite = false;
  g.members.push_back(&w);
  g.stale = true;
  return 1;
}

// Device tables of the jobs in g.table (host staging, then one copy on the
// group stream, ordered after the steps that read the previous tables).
template <typename T> static int group_upload(StepGroup& g) {
  const size_t n = g.table.size();
  const size_t b_ops = align256(n * sizeof(vie::OpArgs<T>)), b_blocks = align256(5 * n * sizeof(unsigned));
  const size_t b_lists = align256(5 * n * sizeof(int)), b_steps = align256(n * sizeof(vie::MjStep<T>));
  // workgroup prefixes: at most 5 n classes of count + 1 entries, then the step's five
  const size_t b_pre = align256((5 * n + 5 * n + 5 * (n + 1)) * sizeof(unsigned));
  const size_t total = b_ops + b_blocks + b_lists + b_steps + b_pre;
  g.off_blocks = b_ops;
  g.off_lists = b_ops + b_blocks;
  g.off_steps = b_ops + b_blocks + b_lists;
  g.off_pre = g.off_steps + b_steps;
  StepGroup::Stage& st = g.stage[g.stage_next];
  g.stage_next = (g.stage_next + 1) % 3;
  if (st.used) HIP_OK(hipEventSynchronize(st.done));  // its last copy has read the staging area
  if (st.cap < total) {
    if (st.host) HIP_OK(hipHostFree(st.host));
    st.host = nullptr;
    st.cap = 0;
    HIP_OK(hipHostMalloc((void**)&st.host, total, hipHostMallocDefault));
    st.cap = total;
  }
  if (g.dev.bytes < total) {
    // steps in flight still read the old tables
    HIP_OK(hipStreamSynchronize(g.stream));
    if (!g.dev.alloc(2 * total)) return 0;
  }
  auto* ops = (vie::OpArgs<T>*)st.host;
  auto* blocks = (unsigned*)(st.host + g.off_blocks);
  auto* lists = (int*)(st.host + g.off_lists);
  auto* steps = (vie::MjStep<T>*)(st.host + g.off_steps);
  auto* pre = (unsigned*)(st.host + g.off_pre);
  for (size_t q = 0; q < n; q++) {
    SolverWork& w = *g.table[q];
    dielectric_fft_system* s = w.sys;
    vie::VolumeView<T> vv = s->plan->view<T>();
    const vie::Geom& ge = vv.g;
    vv.A = w.fftws.as<vie::cx<T>>();
    vv.B = vv.A + 3 * (size_t)ge.nx * ge.ny * ge.Pz;
    vie::ArnoldiState* sd = w.state.as<vie::ArnoldiState>();
    vie::DevIndex di;
    di.j = &sd->j;
    di.stride = w.ldv;
    di.skip = &sd->done;
    vie::cx<T>* V = cycle_basis<T>(w);
    const size_t c = s->c;
    ops[q] = vie::op_args<T>(vv, V + c, V + w.ldv + c, 1, di);
    for (int p = 0; p < 5; p++)
      blocks[5 * q + p] = vie::pass_blocks_mj<T>(ge, p, s->plan->var_mj[sizeof(T) == sizeof(float)][p]);
    vie::MjStep<T> m{};
    m.c = (int)c;
    m.nv = 3 * s->nbox();
    m.C = Prec<T>::C(s).template as<vie::cx<T>>();
    m.B = Prec<T>::B(s).template as<vie::cx<T>>();
    m.D = Prec<T>::D(s).template as<vie::cx<T>>();
    m.x = V;
    m.y = V + w.ldv;
    m.J.nbox = s->nbox();
    m.J.eps = Prec<T>::eps(s->plan).template as<vie::cx<T>>();
    for (int i = 0; i < 3; i++) m.J.self[i] = vie::cx<T>{(T)s->plan->self[i].re, (T)s->plan->self[i].im};
    m.di = di;
    m.Q = V;
    m.ldv = w.ldv;
    m.n = w.n;
    m.st = sd;
    m.partial = w.partial.as<c64>();
    m.dots = w.dots.as<c64>();
    m.ring = w.pin_dev->ring;
    m.eps_b = Prec<T>::eps_b;
    m.max_j = w.m;
    vie::step_blocks_mj<T>(m);
    steps[q] = m;
  }
  // classes: per pass, the jobs with the same line length and variant, in pass order
  g.classes.clear();
  int off = 0, npre = 0;
  const int slot = sizeof(T) == sizeof(double) ? 0 : 1;
  for (int p = 0; p < 5; p++) {
    std::map<std::pair<int, int>, std::vector<int>> byP;
    for (size_t q = 0; q < n; q++) {
      const vie_fft_plan* pl = g.table[q]->sys->plan;
      const vie::Geom& ge = pl->g;
      byP[{p == 2 ? ge.Px : (p == 1 || p == 3) ? ge.Py : ge.Pz, (int)pl->var_mj[slot][p]}].push_back((int)q);
    }
    for (auto& kv : byP) {
      vie::MjClass c{p, kv.first.first, kv.first.second, (int)kv.second.size(), off, npre, 0};
      for (int q : kv.second) {
        lists[off++] = q;
        pre[npre++] = c.total_blocks;
        c.total_blocks += blocks[5 * q + p];
      }
      pre[npre++] = c.total_blocks;
      g.classes.push_back(c);
    }
  }
  // the step kernels' prefixes over all jobs, in table order
  g.step_pre = (size_t)npre;
  for (int k = 0; k < 5; k++) {
    unsigned acc = 0;
    for (size_t q = 0; q < n; q++) {
      pre[npre++] = acc;
      const vie::MjStep<T>& m = steps[q];
      acc += k == 0 ? m.nb_dots : k == 1 ? m.nb_out : k == 2 ? m.nb_dc : k == 3 ? m.nb_finish : m.nb_update;
    }
    pre[npre++] = acc;
    g.total_nb[k] = acc;
  }
  HIP_OK(hipMemcpyAsync(g.dev.p, st.host, total, hipMemcpyHostToDevice, g.stream));
  HIP_OK(hipEventRecord(st.done, g.stream));
  st.used = true;
  return 1;
}

// Reads the statuses of completed group steps, lets finished members go and
// issues the next step.  Returns the number of state changes.
template <typename T> static int group_progress(StepGroup& g) {
  int changed = 0;
  while (g.checked < g.issued) {
    hipError_t r = hipEventQuery(g.ev[g.checked & 1]);
    if (r == hipErrorNotReady) break;
    if (r != hipSuccess) {
      set_error("group step wait failed: %s", hipGetErrorString(r));
      for (SolverWork* w : g.stepped[g.checked & 1]) w->group_nonfinite = true;  // ends the member's solve
    }
    for (SolverWork* w : g.stepped[g.checked & 1]) {
      const StatusBlock& sb = w->pin->ring[w->checked & 1];
      if (sb.nonfinite) w->group_nonfinite = true;
      if (sb.done) w->cycle_done = true;
      w->checked++;
    }
    g.checked++;
    changed++;
  }
  // members whose issued steps are all checked leave when the group has no
  // more matvec steps for them
  size_t keep = 0;
  for (size_t i = 0; i < g.members.size(); i++) {
    SolverWork* w = g.members[i];
    const bool more = !w->cycle_done && !w->group_nonfinite && w->issued < w->m_eff;
    if (!more && w->checked == w->issued) {
      w->grouped = false;
      if (w->group_nonfinite) job_finish(*w, VIE_ITERATIVE_NONFINITE, 0);
      changed++;
      continue;
    }
    g.members[keep++] = w;
  }
  g.members.resize(keep);
  if (g.issued - g.checked >= 2) return changed;
  std::vector<SolverWork*> next;
  for (SolverWork* w : g.members)
    if (!w->cycle_done && !w->group_nonfinite && w->issued < w->m_eff) next.push_back(w);
  if (next.empty()) return changed;
  if (g.stale || next != g.table) {
    g.stale = false;
    g.table.swap(next);
    if (!group_upload<T>(g)) {
      for (SolverWork* w : g.table) w->group_nonfinite = true;
      g.table.clear();
      return changed + 1;
    }
  }
  for (hipEvent_t e : g.waits) HIP_OK(hipStreamWaitEvent(g.stream, e, 0));
  g.waits.clear();
  const char* base = g.dev.as<char>();
  const unsigned* pres = (const unsigned*)(base + g.off_pre);
  const bool ok = vie::volume_apply_mj<T>((const vie::OpArgs<T>*)base, (const unsigned*)(base + g.off_blocks),
                                          (const int*)(base + g.off_lists), pres, g.classes.data(),
                                          (int)g.classes.size(), g.stream) != 0;
  vie::step_mj<T>((const vie::MjStep<T>*)(base + g.off_steps), (int)g.table.size(), pres + g.step_pre, g.total_nb,
                  g.stream);
  if (!ok || hipGetLastError() != hipSuccess || hipEventRecord(g.ev[g.issued & 1], g.stream) != hipSuccess) {
    set_error("group step failed to launch");
    for (SolverWork* w : g.table) w->group_nonfinite = true;
    g.table.clear();
    return changed + 1;
  }
  g.stepped[g.issued & 1] = g.table;
  for (SolverWork* w : g.table) {
    w->rep.operator_applications++;
    w->col_sum += w->issued + w->cyc_base;
    w->col_steps++;
    w->issued++;
  }
  g.issued++;
  g.launches++;
  g.steps += g.table.size();
  return changed + 1;
}

// The shared subspace of w's system and precision, when it was built for the
// operator w solves with w's k and basis stride.
template <typename T> static dielectric_fft_system::SharedSubspace* shared_for(SolverWork& w) {
  dielectric_fft_system::SharedSubspace& sh = w.sys->shared[sizeof(T) == sizeof(float)];
  const bool match = sh.ok && sh.k <= w.rk && sh.ldv == w.ldv && sh.gen == w.op_gen && sh.eps_gen == w.op_eps_gen;
  return match ? &sh : nullptr;
}

// Publishes w's freshly built U and C (basis columns 0 .. 2 rk - 1) for the
// other jobs of its system, unless an equal subspace is already there.  The
// copy is queued on the system stream, which every job of the system uses
// for its cycle starts, so an adopting job reads it complete.
template <typename T> static void shared_publish(SolverWork& w) {
  if (shared_for<T>(w)) return;
  dielectric_fft_system::SharedSubspace& sh = w.sys->shared[sizeof(T) == sizeof(float)];
  sh.ok = false;
  const size_t bytes = (size_t)2 * w.rk * w.ldv * sizeof(vie::cx<T>);
  if (!sh.uc.alloc(bytes) ||
      hipMemcpyAsync(sh.uc.p, w.V.p, bytes, hipMemcpyDeviceToDevice, w.sys->stream) != hipSuccess)
    return;
  sh.k = w.rk;
  sh.ldv = w.ldv;
  sh.gen = w.op_gen;
  sh.eps_gen = w.op_eps_gen;
  sh.ok = true;
  w.published = true;
}

// At the start of w's first cycle: take the system's shared U and C, if any
// fit, so that this cycle minimises over span(U) + its Krylov space already.
// A published subspace may hold fewer vectors than w reserved (a short seed
// cycle): w then recycles that many.
template <typename T> static void shared_adopt(SolverWork& w) {
  dielectric_fft_system::SharedSubspace* sh = shared_for<T>(w);
  if (!sh) return;
  const size_t bytes = (size_t)2 * sh->k * w.ldv * sizeof(vie::cx<T>);
  if (hipMemcpyAsync(w.V.p, sh->uc.p, bytes, hipMemcpyDeviceToDevice, w.sys->stream) != hipSuccess) return;
  w.rk = sh->k;
  w.rec_ready = true;
  w.adopted = true;
}

// Builds the recycled subspace from the finished first cycle (its basis from
// column 2 rk, its unrotated Hessenberg, and the last step's a and alpha in
// w.rec_host): P = the rk harmonic Ritz vectors of smallest modulus,
// Hbar P = Q R (thin QR), C = V_{J+1} Q, U = V_J P R^{-1}, so that
// A M^{-1} U = C with C orthonormal (Parks, de Sturler, Mackey, Johnson, Maiti,
// "Recycling Krylov subspaces for sequences of linear systems", SIAM J. Sci.
// Comput. 28 (2006), the GCRO-DR start; here the subspace then stays fixed).
// The first cycle leaves its last basis vector as u = alpha q_J + Q_J a, so
// the coefficients of column J are folded onto u and the earlier columns.
// False (the solve continues with plain cycles) when the cycle is too short
// or the small problems fail.
template <typename T> static bool recycle_build(SolverWork& w) {
  namespace D = vie::dense;
  const int J = w.rec_used, m = w.m, k0 = w.rk;
  // a job that shares its subspace takes what a short first cycle offers
  const int k = (w.share && J < k0 + 1) ? J - 1 : k0;
  if (k < 1 || J < k + 1 || w.rec_breakdown) return false;
  const double t0 = now_s();
  const c64* hr = w.rec_host;  // column-major, ld m + 1
  const c64* a = hr + (size_t)(m + 1) * m;
  const double alpha = hr[(size_t)(m + 1) * m + m].re;
  if (!(alpha > 0) || !std::isfinite(alpha)) return false;
  D::Mat Hb(J + 1, J);
  for (int c = 0; c < J; c++)
    for (int r = 0; r <= c + 1 && r <= J; r++) {
      const c64 v = hr[(size_t)c * (m + 1) + r];
      Hb(r, c) = D::zd(v.re, v.im);
    }
  D::Mat P, G(J + 1, k), Qg, R, Ri;
  if (!D::harmonic_ritz(Hb, k, P)) return false;
  for (int i = 0; i < k; i++)
    for (int r = 0; r <= J; r++) {
      D::zd acc(0, 0);
      for (int c = (r > 0 ? r - 1 : 0); c < J; c++) acc = acc + Hb(r, c) * P(c, i);
      G(r, i) = acc;
    }
  if (!D::thin_qr(G, Qg, R) || !D::upper_inverse(R, Ri)) return false;
  c64* cc = w.rec_host + (size_t)(m + 1) * m + m + 1;  // C coefficients, (J+1) x k
  c64* uc = cc + (size_t)(J + 1) * k;                  // U coefficients, J x k
  bool finite = true;
  for (int i = 0; i < k; i++) {
    const D::zd last = (1.0 / alpha) * Qg(J, i);
    for (int c = 0; c < J; c++) {
      const D::zd v = Qg(c, i) - D::zd(a[c].re, a[c].im) * last;
      cc[(size_t)i * (J + 1) + c] = c64{v.re, v.im};
    }
    cc[(size_t)i * (J + 1) + J] = c64{last.re, last.im};
    for (int c = 0; c < J; c++) {
      D::zd acc(0, 0);
      for (int l = 0; l <= i; l++) acc = acc + P(c, l) * Ri(l, i);
      uc[(size_t)i * J + c] = c64{acc.re, acc.im};
      finite = finite && std::isfinite(acc.re) && std::isfinite(acc.im);
    }
  }
  if (!finite) return false;
  hipStream_t st = w.sys->stream;
  const size_t nc = (size_t)(J + 1) * k + (size_t)J * k;
  if (hipMemcpyAsync(w.rcoef.p, cc, nc * sizeof(c64), hipMemcpyHostToDevice, st) != hipSuccess) return false;
  vie::cx<T>* V = w.V.as<vie::cx<T>>();
  const size_t ldv = w.ldv;
  const vie::cx<T>* first = V + (size_t)2 * k0 * ldv;  // the first cycle's basis
  vie::block_combine<T>(V + (size_t)k * ldv, k, first, J + 1, ldv, w.rcoef.as<c64>(), w.n, st);
  vie::block_combine<T>(V, k, first, J, ldv, w.rcoef.as<c64>() + (size_t)(J + 1) * k, w.n, st);
  if (hipGetLastError() != hipSuccess) return false;
  w.rk = k;
  w.rec_ready = true;
  if (trace_enabled())
    fprintf(stderr, "[vie_hip] recycled subspace of %d vectors from a %d-column cycle (host %.2f ms)\n", k, J,
            (now_s() - t0) * 1e3);
  if (w.share) shared_publish<T>(w);
  return true;
}

// Advances the solve as far as it can without waiting (or, when blocking, to
// the end).  Returns 1 once the solve has finished.
template <typename T> static int job_advance(SolverWork& w, bool blocking) {
  dielectric_fft_system* s = w.sys;
  hipStream_t st = s->stream;
  vie_gmres_report& rep = w.rep;
  vie::ArnoldiState* sd = w.state.as<vie::ArnoldiState>();
  const bool mixed = sizeof(T) == sizeof(float);
  for (;;) {
    switch (w.phase) {
      case PH_FINISHED:
        return 1;
      case PH_NORMS: {
        int r = job_ready(w.ev_misc, blocking);
        if (r <= 0) {
          if (r < 0) job_finish(w, VIE_ITERATIVE_CALLBACK_FAILED, 0);
          return r < 0;
        }
        w.bnorm = sqrt(w.pin->scal[0].re);
        double xnorm = sqrt(w.pin->scal[1].re);
        w.target = std::max(w.o.absolute_tolerance, w.o.relative_tolerance * w.bnorm);
        if (!std::isfinite(w.bnorm) || !std::isfinite(xnorm) || !std::isfinite(w.target)) {
          job_finish(w, VIE_ITERATIVE_NONFINITE, 0);
          return 1;
        }
        if (xnorm == 0) {
          // x = 0 exactly: r = b and ||r|| = ||b|| (the norm kernels sum in
          // the same order), so the start needs no operator application and
          // no second round trip
          // (the cycle start reads ||r||^2 from the device copy, scal[2])
          if (hipMemcpyAsync(w.r.p, w.b.p, w.n * sizeof(c64), hipMemcpyDeviceToDevice, st) != hipSuccess ||
              hipMemcpyAsync(w.scal.as<c64>() + 2, w.scal.as<c64>(), sizeof(c64), hipMemcpyDeviceToDevice, st) !=
                  hipSuccess) {
            job_finish(w, VIE_ITERATIVE_CALLBACK_FAILED, 0);
            return 1;
          }
          w.pin->scal[2] = w.pin->scal[0];
          w.phase = PH_RESID;
          w.resid_known = true;
          break;
        }
        if (!job_issue_residual(w)) {
          job_finish(w, VIE_ITERATIVE_CALLBACK_FAILED, 0);
          return 1;
        }
        w.phase = PH_RESID;
        break;
      }
      case PH_RESID: {
        if (!w.resid_known) {
          int r = job_ready(w.ev_misc, blocking);
          if (r <= 0) {
            if (r < 0) job_finish(w, VIE_ITERATIVE_CALLBACK_FAILED, 0);
            return r < 0;
          }
        }
        w.resid_known = false;
        if (w.check_singular) {
          w.check_singular = false;
          if (w.pin->fin.singular) {
            rep.operator_applications--;  // the residual of the unchanged x was not needed
            job_finish(w, VIE_ITERATIVE_BREAKDOWN, 0);
            return 1;
          }
        }
        w.beta = sqrt(w.pin->scal[2].re);
        rep.residual_norm = w.beta;
        if (!std::isfinite(w.beta)) {
          job_finish(w, VIE_ITERATIVE_NONFINITE, 0);
          return 1;
        }
        rep.relative_residual = w.bnorm > 0 ? w.beta / w.bnorm : w.beta;
        if (trace_enabled() && w.after_cycle)
          fprintf(stderr, "[vie_hip] %s cycle end: iterations %zu (cycle %zu) true relative residual %.3e "
                          "estimate %.3e\n", mixed ? "fp32" : "fp64", (size_t)rep.iterations,
                  (size_t)(rep.iterations - w.it0), rep.relative_residual,
                  w.bnorm > 0 ? w.pin->fin.gabs / w.bnorm : 0.0);
        if (w.after_cycle) {
          // an FP32 breakdown is a precision artefact: FP64 cycles take over below
          if (w.breakdown && w.beta > w.target && !mixed) {
            job_finish(w, VIE_ITERATIVE_BREAKDOWN, 0);
            return 1;
          }
          if (w.beta > w.target && rep.iterations < w.o.maximum_iterations) rep.restarts++;
        }
        if (w.beta <= w.target) {
          // a first cycle that converged still gives the other jobs of its
          // system their recycled subspace
          if (w.share && w.rk && !w.rec_ready && w.cycles == 1 && !shared_for<T>(w)) (void)recycle_build<T>(w);
          job_finish(w, VIE_ITERATIVE_CONVERGED, 1);
          return 1;
        }
        if (rep.iterations >= w.o.maximum_iterations) {
          job_finish(w, VIE_ITERATIVE_ITERATION_LIMIT, 0);
          return 1;
        }
        // An FP32 cycle that gained less than 2x on the FP64 true residual has
        // hit the FP32 floor of this system (its solution is too large against
        // the tolerance for FP32 products): the remaining cycles run in FP64.
        if (mixed && w.after_cycle && (w.breakdown || w.beta > 0.5 * w.cycle_beta)) {
          // The FP32 recycled subspace does not carry over, but the FP64
          // cycles recycle their own: they are what the hard systems need
          // (restarted GMRES alone can stagnate on them).  The first FP64
          // cycle is a first cycle again: it builds its subspace from
          // column 2 k, and a shared FP64 subspace of the system is adopted.
          // a system FP32 could not advance is a hard one: the FP64 cycles
          // keep the largest subspace the basis allows
          if (w.rk_req > 0) w.rk_req = std::max(w.rk_req, std::min<int>(vie::kMaxRecycle, w.m / 2));
          if (!w.V.alloc((size_t)(w.m + 1 + 2 * w.rk_req) * w.ldv * sizeof(c64))) {
            job_finish(w, VIE_ITERATIVE_ALLOCATION_FAILED, 0);
            return 1;
          }
          w.prec = VIE_HIP_FP64;
          w.escalated = true;
          w.rk = w.rk_req;
          w.rec_ready = false;
          w.cycles = 0;
          w.rec_used = 0;
          w.rec_breakdown = false;
          if (trace_enabled())
            fprintf(stderr, "[vie_hip] fp32 cycle stalled at %.3e: continuing in fp64\n", rep.relative_residual);
          w.cycle_beta = w.beta;
          w.breakdown = false;
          w.phase = PH_RESID;  // re-entered in FP64: the event has completed
          w.after_cycle = false;
          return job_advance<double>(w, blocking);
        }
        w.cycle_beta = w.beta;
        if (w.share && w.rk && !w.rec_ready && w.cycles == 0) shared_adopt<T>(w);
        // new cycle; the second one first forms the recycled subspace
        if (w.rk && !w.rec_ready && w.cycles == 1 && !recycle_build<T>(w)) {
          if (trace_enabled()) fprintf(stderr, "[vie_hip] no recycled subspace: plain cycles\n");
          w.rec_used = 0;
        }
        const int base = w.rec_ready ? w.rk : 0;
        const double inner_target = mixed ? std::max(w.target, w.reduction * w.beta) : w.target;
        vie::cycle_init(sd, w.scal.as<c64>() + 2, inner_target, base, st);
        // the streaming Arnoldi kernels need zero column padding; the buffer
        // may last have held the other precision's columns
        vie::zero_padding<T>(w.V.as<vie::cx<T>>(), w.ldv, w.n, w.V.bytes / (w.ldv * sizeof(vie::cx<T>)), st);
        {
          vie::cx<T>* Q = cycle_basis<T>(w);
          vie::scale_from<T>(Q + (size_t)base * w.ldv, w.r.as<c64>(), w.n, sd, st);
          if (base) {
            vie::ReduceScratch rs;
            rs.partial = w.partial.as<c64>();
            vie::recycle_project<T>(Q, w.ldv, base, w.n, sd, rs, st);
          }
        }
        w.cyc_base = base;
        w.cycles++;
        w.it0 = rep.iterations;
        // columns this cycle may add (the reference counts Arnoldi steps, vie_iterative.c:134)
        w.m_eff = (int)std::min<size_t>((size_t)(w.m - base), w.o.maximum_iterations - w.it0);
        w.steps = w.dcgs2 ? w.m_eff + 1 : w.m_eff;
        w.issued = w.checked = 0;
        w.cycle_done = false;
        w.phase = PH_CYCLE;
        if (mj_eligible<T>(w) && w.m_eff > 0 && !group_join<T>(w)) {
          job_finish(w, VIE_ITERATIVE_ALLOCATION_FAILED, 0);
          return 1;
        }
        break;
      }
      case PH_CYCLE: {
        if (w.grouped) return 0;  // the batch's step group issues and checks the matvec steps
        // one step of lookahead: step j is queued before the status of step
        // j-1 is read, so the GPU never idles on the host; the device marks
        // the cycle done and later queued steps leave the state untouched
        bool progressed = true;
        while (progressed) {
          progressed = false;
          if (!w.cycle_done && w.issued < w.steps && w.issued - w.checked < 2) {
            if (!job_issue_step<T>(w, w.issued)) {
              job_finish(w, VIE_ITERATIVE_CALLBACK_FAILED, 0);
              return 1;
            }
            w.issued++;
            progressed = true;
            continue;
          }
          if (w.checked < w.issued) {
            int r = job_ready(w.ev_ring[w.checked & 1], blocking);
            if (r < 0) {
              job_finish(w, VIE_ITERATIVE_CALLBACK_FAILED, 0);
              return 1;
            }
            if (r == 0) return 0;
            const StatusBlock& sb = w.pin->ring[w.checked & 1];
            if (sb.nonfinite) {
              job_finish(w, VIE_ITERATIVE_NONFINITE, 0);
              return 1;
            }
            if (sb.done) w.cycle_done = true;
            w.checked++;
            progressed = true;
          }
        }
        // all queued steps are complete: the last status holds the cycle's end
        const StatusBlock& last = w.pin->ring[(w.issued - 1) & 1];
        rep.iterations = w.it0 + (size_t)(last.used - w.cyc_base);
        w.breakdown = last.breakdown != 0;
        vie::backsolve(sd, st);
        HIP_OK(hipMemcpyAsync(&w.pin->fin, sd, sizeof(StatusBlock), hipMemcpyDeviceToHost, st));
        if (w.rk && !w.rec_ready && w.cycles == 1) {
          // the first cycle's Hessenberg and last-step coefficients, for the
          // recycled subspace should the solve need another cycle
          const int J = last.used, m = w.m;
          w.rec_used = J;
          w.rec_breakdown = last.breakdown != 0;
          const char* sp = (const char*)sd;
          if (J > 0) {
            HIP_OK(hipMemcpyAsync(w.rec_host, sp + offsetof(vie::ArnoldiState, Hraw),
                                  (size_t)J * (m + 1) * sizeof(c64), hipMemcpyDeviceToHost, st));
            HIP_OK(hipMemcpyAsync(w.rec_host + (size_t)(m + 1) * m, sp + offsetof(vie::ArnoldiState, dc_a),
                                  (size_t)J * sizeof(c64), hipMemcpyDeviceToHost, st));
            HIP_OK(hipMemcpyAsync(w.rec_host + (size_t)(m + 1) * m + m, sp + offsetof(vie::ArnoldiState, dc_alpha),
                                  sizeof(double), hipMemcpyDeviceToHost, st));
          }
        }
        // The solution update and the true residual are queued behind the
        // back substitution at once (one host round trip per cycle end):
        // k_combine leaves x as it is when the back substitution found no
        // least-squares solution, and PH_RESID reports that case first.
        vie::combine<T>(w.x.as<c64>(), cycle_basis<T>(w), w.ldv, w.n, sd, jacobi_view(s), 1, st,
                        w.cyc_base ? w.V.as<vie::cx<T>>() : nullptr);
        if (!job_issue_residual(w)) {
          job_finish(w, VIE_ITERATIVE_CALLBACK_FAILED, 0);
          return 1;
        }
        w.after_cycle = true;
        w.check_singular = true;
        w.phase = PH_RESID;
        break;
      }
      default:
        job_finish(w, VIE_ITERATIVE_INVALID_INPUT, 0);
        return 1;
    }
  }
}

static int job_advance_any(SolverWork& w, bool blocking) {
  return w.prec == VIE_HIP_MIXED ? job_advance<float>(w, blocking) : job_advance<double>(w, blocking);
}

static int validate_options(const vie_gmres_options* o) {
  return o && o->restart && o->maximum_iterations && std::isfinite(o->relative_tolerance) &&
         o->relative_tolerance >= 0 && std::isfinite(o->absolute_tolerance) && o->absolute_tolerance >= 0 &&
         !(o->relative_tolerance == 0 && o->absolute_tolerance == 0);
}

static void report_init(vie_gmres_report* r) {
  memset(r, 0, sizeof(*r));
  r->status = VIE_ITERATIVE_INVALID_INPUT;
  r->residual_norm = r->relative_residual = INFINITY;
}

// Restart length the device solver uses for these options (0 when too long).
static size_t device_restart(size_t n, const vie_gmres_options* o) {
  size_t m = o->restart;
  if (m > n) m = n;
  if (m > o->maximum_iterations) m = o->maximum_iterations;
  if (m > (size_t)vie::kMaxRestart) {
    set_error("restart %zu exceeds the device solver's %d", m, vie::kMaxRestart);
    return 0;
  }
  return m;
}

// Readies w for a solve of s (basis sized for the options, blocks and
// permittivities on the device).  0 with rep->status set on failure.
static int job_prepare(dielectric_fft_system* s, SolverWork& w, const vie_gmres_options* o,
                       const vie_hip_solve_options* so, vie_gmres_report* rep) {
  size_t m = device_restart(s->c + s->v, o);
  int prec = so->precision == VIE_HIP_MIXED ? VIE_HIP_MIXED : VIE_HIP_FP64;
  if (!m || !ensure_solver(s, w, (int)m, prec, recycle_count(so, prec, (int)m)) || !upload_blocks(s) ||
      (s->plan && !plan_upload_eps(s->plan))) {
    rep->status = VIE_ITERATIVE_ALLOCATION_FAILED;
    return 0;
  }
  if (!plan_tune_prec(s->plan, prec) || !plan_tune_prec(s->plan, VIE_HIP_FP64)) {
    rep->status = VIE_ITERATIVE_CALLBACK_FAILED;
    return 0;
  }
  w.op_gen = s->gen;
  w.op_eps_gen = s->plan ? s->plan->gen : 0;
  w.share = w.adopted = w.published = false;
  return 1;
}

// Solve with b already in the solver's device buffer (system layout) and x
// holding the initial guess.
static int solve_resident(dielectric_fft_system* s, const vie_gmres_options* options,
                          const vie_hip_solve_options* so, vie_gmres_report* rep) {
  PhaseTimer phase_timer_(PHASE_GMRES);
  vie_gmres_options defaults = {30, 300, 1e-12, 0};
  if (!options) options = &defaults;
  if (!validate_options(options)) return 0;
  if (!so) {
    init_solve_defaults();
    so = &g_solve;
  }
  SolverWork& w = s->sw;
  if (!job_prepare(s, w, options, so, rep)) return 0;
  if (!job_start(w, options, so)) {
    rep->status = VIE_ITERATIVE_CALLBACK_FAILED;
    return 0;
  }
  job_advance_any(w, true);
  *rep = w.rep;
  return w.result;
}

static int stage_vector(dielectric_fft_system* s, const double _Complex* user, c64* dst) {
  // user layout [c | 3 cells interleaved] -> system layout (holes zero)
  size_t nu = s->c + s->v, n = s->nsys();
  DevBuf& tmp = s->plan ? s->plan->io_user : s->sw.ax;
  if (!tmp.alloc(std::max<size_t>(nu, 1) * 16)) return 0;
  HIP_OK(hipMemcpyAsync(tmp.p, user, nu * 16, hipMemcpyHostToDevice, s->stream));
  vie::vec_zero<double>(dst, n, s->stream);
  vie::to_box<double>(tmp.as<c64>(), s->c, s->cells, s->plan ? s->plan->map_dev.as<size_t>() : nullptr, s->nbox(), dst,
                      s->stream);
  HIP_OK(hipStreamSynchronize(s->stream));
  return 1;
}

static int unstage_vector(dielectric_fft_system* s, const c64* src, double _Complex* user) {
  size_t nu = s->c + s->v;
  DevBuf& tmp = s->plan ? s->plan->io_user : s->sw.r;
  if (!tmp.alloc(std::max<size_t>(nu, 1) * 16)) return 0;
  vie::from_box<double>(src, s->c, s->cells, s->plan ? s->plan->map_dev.as<size_t>() : nullptr, s->nbox(),
                        tmp.as<c64>(), s->stream);
  HIP_OK(hipMemcpyAsync(user, tmp.p, nu * 16, hipMemcpyDeviceToHost, s->stream));
  HIP_OK(hipStreamSynchronize(s->stream));
  return 1;
}

static int ensure_vectors(dielectric_fft_system* s) {
  size_t n = s->nsys();
  return s->sw.x.alloc(n * 16) && s->sw.b.alloc(n * 16) && s->sw.r.alloc(n * 16) && s->sw.ax.alloc(n * 16);
}

static int solve_host(dielectric_fft_system* s, const double _Complex* b, double _Complex* x,
                      const vie_gmres_options* options, const vie_hip_solve_options* so, vie_gmres_report* report) {
  vie_gmres_report local;
  if (!report) report = &local;
  report_init(report);
  g_last_error.clear();
  if (!s || !b || !x || (const void*)b == (const void*)x) return 0;
  if (!validate_options(options ? options : &dielectric_gmres_settings)) return 0;
  if (!ensure_vectors(s) || !stage_vector(s, b, s->sw.b.as<c64>()) || !stage_vector(s, x, s->sw.x.as<c64>())) {
    report->status = VIE_ITERATIVE_ALLOCATION_FAILED;
    return 0;
  }
  int ok = solve_resident(s, options, so, report);
  // x is updated even on failure (vie_iterative.h:48-50)
  if (report->status != VIE_ITERATIVE_INVALID_INPUT && !unstage_vector(s, s->sw.x.as<c64>(), x)) return 0;
  return ok;
}

static dielectric_fft_system* plan_system(vie_fft_plan* p) {
  if (!p->wrap) {
    dielectric_fft_system* s = new (std::nothrow) dielectric_fft_system();
    if (!s) return nullptr;
    s->c = 0;
    s->v = 3 * p->cells;
    s->cells = p->cells;
    s->plan = p;
    s->owns_plan = false;
    s->stream = p->stream;
    s->blocks = (c64*)calloc(1, sizeof(c64));
    if (!s->blocks) {
      delete s;
      return nullptr;
    }
    s->blocks_dirty = false;
    p->wrap = s;
  }
  return p->wrap;
}

extern "C" int dielectric_fft_solve(dielectric_fft_system* s, const double _Complex* b, double _Complex* x,
                                    const vie_gmres_options* options, vie_gmres_report* report) {
  return solve_host(s, b, x, options, nullptr, report);
}

extern "C" int vie_fft_solve(vie_fft_plan* p, const double _Complex* b, double _Complex* x,
                             const vie_gmres_options* options, vie_gmres_report* report) {
  if (!p) {
    if (report) report_init(report);
    return 0;
  }
  dielectric_fft_system* s = plan_system(p);
  if (!s) return 0;
  return solve_host(s, b, x, options, nullptr, report);
}

extern "C" void vie_hip_set_solve_options(const vie_hip_solve_options* o) {
  init_solve_defaults();
  if (o) g_solve = *o;
}
extern "C" void vie_hip_get_solve_options(vie_hip_solve_options* o) {
  init_solve_defaults();
  if (o) *o = g_solve;
}

extern "C" int dielectric_fft_upload_rhs(dielectric_fft_system* s, const double _Complex* b) {
  if (!s || !b || !ensure_vectors(s) || !stage_vector(s, b, s->sw.b.as<c64>())) return 0;
  s->rhs_loaded = true;
  return 1;
}

extern "C" int dielectric_fft_reserve(dielectric_fft_system* s, const vie_gmres_options* options) {
  vie_gmres_options defaults = {30, 300, 1e-12, 0};
  if (!s) return 0;
  if (!options) options = &defaults;
  if (!validate_options(options)) return 0;
  size_t n = s->c + s->v, m = options->restart;
  if (m > n) m = n;
  if (m > options->maximum_iterations) m = options->maximum_iterations;
  if (m > (size_t)vie::kMaxRestart) return 0;
  // FP64-sized basis serves both precisions (and a MIXED one with its recycled subspace)
  if (!ensure_solver(s, s->sw, (int)m, VIE_HIP_FP64) || !ensure_recycle(s->sw) || !upload_blocks(s) ||
      (s->plan && !plan_upload_eps(s->plan)) ||
      !plan_tune_prec(s->plan, VIE_HIP_FP64) || !plan_tune_prec(s->plan, VIE_HIP_MIXED))
    return 0;
  HIP_OK(hipStreamSynchronize(s->stream));
  return 1;
}

extern "C" int dielectric_fft_solve_resident(dielectric_fft_system* s, const vie_gmres_options* options,
                                             const vie_hip_solve_options* so, vie_gmres_report* report) {
  vie_gmres_report local;
  if (!report) report = &local;
  report_init(report);
  if (!s || !s->rhs_loaded) return 0;
  vie::vec_zero<double>(s->sw.x.as<c64>(), s->nsys(), s->stream);  // x0 = 0 as the callers do
  return solve_resident(s, options, so, report);
}

extern "C" int dielectric_fft_download_solution(dielectric_fft_system* s, double _Complex* x) {
  if (!s || !x || !s->sw.x.p) return 0;
  return unstage_vector(s, s->sw.x.as<c64>(), x);
}

// ------------------------------------------------------------------ host-pointer operator callbacks
// Compatibility path of the reference's operator contract: each call stages
// through device memory.  vie_gmres recognises these callbacks and runs the
// whole solve on the device instead.
static int apply_host(dielectric_fft_system* s, const double _Complex* in, double _Complex* out, bool precond) {
  PhaseTimer phase_timer_(PHASE_MATVEC);
  if (!s || !in || !out) return 0;
  size_t n = s->nsys();
  if (!s->sw.x.alloc(n * 16) || !s->sw.ax.alloc(n * 16) || !upload_blocks(s) ||
      (s->plan && !plan_upload_eps(s->plan)) || !ensure_scratch(s->sw) || !plan_tune_prec(s->plan, VIE_HIP_FP64))
    return 0;
  if (!stage_vector(s, in, s->sw.x.as<c64>())) return 0;
  if (precond) {
    vie::precondition_apply(s->sw.x.as<c64>(), s->sw.ax.as<c64>(), n, jacobi_view(s), s->stream);
  } else {
    vie::ReduceScratch rs;
    rs.partial = s->sw.partial.as<c64>();
    if (!sys_apply<double>(s, s->sw.x.as<c64>(), s->sw.ax.as<c64>(), 0, rs)) return 0;
  }
  return unstage_vector(s, s->sw.ax.as<c64>(), out);
}

extern "C" int dielectric_fft_apply_ex(dielectric_fft_system* s, vie_hip_precision prec, int jacobi,
                                       const double _Complex* in, double _Complex* out) {
  PhaseTimer phase_timer_(PHASE_MATVEC);
  if (!s || !in || !out) return 0;
  size_t n = s->nsys();
  if (!s->sw.x.alloc(n * 16) || !s->sw.ax.alloc(n * 16) || !upload_blocks(s) ||
      (s->plan && !plan_upload_eps(s->plan)) || !ensure_scratch(s->sw) || !plan_tune_prec(s->plan, prec))
    return 0;
  if (!stage_vector(s, in, s->sw.x.as<c64>())) return 0;
  vie::ReduceScratch rs;
  rs.partial = s->sw.partial.as<c64>();
  if (prec == VIE_HIP_MIXED) {
    DevBuf a, b;
    if (!a.alloc(n * 8) || !b.alloc(n * 8)) return 0;
    vie::vec_cast<float, double>(a.as<c32>(), s->sw.x.as<c64>(), n, s->stream);
    if (!sys_apply<float>(s, a.as<c32>(), b.as<c32>(), jacobi, rs)) return 0;
    vie::vec_cast<double, float>(s->sw.ax.as<c64>(), b.as<c32>(), n, s->stream);
    HIP_OK(hipStreamSynchronize(s->stream));
  } else if (!sys_apply<double>(s, s->sw.x.as<c64>(), s->sw.ax.as<c64>(), jacobi, rs)) {
    return 0;
  }
  return unstage_vector(s, s->sw.ax.as<c64>(), out);
}

// ------------------------------------------------------------------ allocation-free apply
// Device scratch of dielectric_fft_apply_noalloc: one array per GPU for the
// life of the process, so that entry point never allocates.  Host code reaches
// it only through the per-device address hipGetSymbolAddress returns.
constexpr size_t kNoAllocScratchBytes = (size_t)96 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevicesNoAlloc = 64;
static std::mutex noalloc_mu;
static void* noalloc_ptr[kMaxDevicesNoAlloc];

// Address of the current device's scratch array (looked up once per device).
static unsigned char* noalloc_scratch() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevicesNoAlloc) {
    set_error("dielectric_fft_apply_noalloc: no usable current device");
    return nullptr;
  }
  std::lock_guard<std::mutex> g(noalloc_mu);
  if (!noalloc_ptr[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error("dielectric_fft_apply_noalloc: hipGetSymbolAddress failed: %s", hipGetErrorString(e));
      return nullptr;
    }
    noalloc_ptr[dev] = ptr;
  }
  return (unsigned char*)noalloc_ptr[dev];
}

// Bytes of scratch one allocation-free apply of this system needs, and the
// offsets of its parts (user staging, box-order input and output, FFT
// workspace, reduction partials), each rounded up to 256 bytes.
struct NoAllocLayout {
  size_t user, xin, yout, work, partial, total;
};
static NoAllocLayout noalloc_layout(const dielectric_fft_system* s, bool mixed) {
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t es = mixed ? 8 : 16;
  const size_t nu = std::max<size_t>(s->c + s->v, 1), n = s->nsys();
  size_t work = 0;
  if (s->plan) {
    const vie::Geom& g = s->plan->g;
    work = (3 * (size_t)g.nx * g.ny * g.Pz + 3 * (size_t)g.nx * g.Py * g.Pz) * es;
  }
  NoAllocLayout L;
  L.user = 0;
  L.xin = L.user + up(nu * 16);
  L.yout = L.xin + up(n * es);
  L.work = L.yout + up(n * es);
  L.partial = L.work + up(work);
  L.total = L.partial + (s->c ? vie::kReduceScratchBytes : 0);
  return L;
}

extern "C" size_t dielectric_fft_apply_noalloc_bytes(const dielectric_fft_system* s, vie_hip_precision prec) {
  return s ? noalloc_layout(s, prec == VIE_HIP_MIXED).total : 0;
}

// dielectric_fft_apply_ex with every temporary in the per-device scratch
// array: no device memory is allocated or freed.  The system's resident data
// (spectrum, permittivity, C/B/D) must already be on the device.
extern "C" int dielectric_fft_apply_noalloc(dielectric_fft_system* s, vie_hip_precision prec, int jacobi,
                                            const double _Complex* in, double _Complex* out) {
  PhaseTimer phase_timer_(PHASE_MATVEC);
  if (!s || !in || !out || (prec != VIE_HIP_FP64 && prec != VIE_HIP_MIXED)) {
    set_error("dielectric_fft_apply_noalloc: invalid arguments");
    return 0;
  }
  const bool mixed = prec == VIE_HIP_MIXED;
  const NoAllocLayout L = noalloc_layout(s, mixed);
  if (L.total > kNoAllocScratchBytes) {
    set_error("dielectric_fft_apply_noalloc: the system needs %zu bytes of scratch, more than the %zu-byte "
              "maximum; use dielectric_fft_apply_ex",
              L.total, kNoAllocScratchBytes);
    return 0;
  }
  if (s->blocks_dirty || (s->plan && (s->plan->eps_dirty || !s->plan->map_dev.p))) {
    set_error("dielectric_fft_apply_noalloc: the system's device copies are not current; call "
              "dielectric_fft_reserve (or any apply or solve) after the last entry or permittivity change");
    return 0;
  }
  unsigned char* base = noalloc_scratch();
  if (!base) return 0;
  const size_t nu = s->c + s->v, n = s->nsys();
  const size_t* map = s->plan ? s->plan->map_dev.as<size_t>() : nullptr;
  c64* user = (c64*)(base + L.user);
  vie::ReduceScratch rs;
  rs.partial = (c64*)(base + L.partial);
  HIP_OK(hipMemcpyAsync(user, in, nu * 16, hipMemcpyHostToDevice, s->stream));
  int ok;
  if (mixed) {
    c32* x = (c32*)(base + L.xin);
    c32* y = (c32*)(base + L.yout);
    vie::vec_zero<float>(x, n, s->stream);
    vie::to_box<float>(user, s->c, s->cells, map, s->nbox(), x, s->stream);
    ok = sys_apply_ws<float>(s, x, y, jacobi, rs, (c32*)(base + L.work));
    if (ok) vie::from_box<float>(y, s->c, s->cells, map, s->nbox(), user, s->stream);
  } else {
    c64* x = (c64*)(base + L.xin);
    c64* y = (c64*)(base + L.yout);
    vie::vec_zero<double>(x, n, s->stream);
    vie::to_box<double>(user, s->c, s->cells, map, s->nbox(), x, s->stream);
    ok = sys_apply_ws<double>(s, x, y, jacobi, rs, (c64*)(base + L.work));
    if (ok) vie::from_box<double>(y, s->c, s->cells, map, s->nbox(), user, s->stream);
  }
  if (!ok) {
    (void)hipStreamSynchronize(s->stream);
    return 0;
  }
  HIP_OK(hipMemcpyAsync(out, user, nu * 16, hipMemcpyDeviceToHost, s->stream));
  HIP_OK(hipStreamSynchronize(s->stream));
  return 1;
}

static int cb_plan_apply(void* ctx, const double _Complex* in, double _Complex* out) {
  vie_fft_plan* p = (vie_fft_plan*)ctx;
  return p ? apply_host(plan_system(p), in, out, false) : 0;
}
static int cb_plan_precondition(void* ctx, const double _Complex* in, double _Complex* out) {
  vie_fft_plan* p = (vie_fft_plan*)ctx;
  return p ? apply_host(plan_system(p), in, out, true) : 0;
}
static int cb_system_apply(void* ctx, const double _Complex* in, double _Complex* out) {
  return apply_host((dielectric_fft_system*)ctx, in, out, false);
}
static int cb_system_precondition(void* ctx, const double _Complex* in, double _Complex* out) {
  return apply_host((dielectric_fft_system*)ctx, in, out, true);
}

extern "C" void vie_fft_operator(vie_fft_plan* p, vie_linear_operator* op) {
  memset(op, 0, sizeof(*op));
  if (!p) return;
  op->order = 3 * p->cells;
  op->apply = cb_plan_apply;
  op->context = p;
  op->precondition = cb_plan_precondition;
  op->precondition_context = p;
}

extern "C" void dielectric_fft_get_operator(dielectric_fft_system* s, vie_linear_operator* op) {
  op->order = s->c + s->v;
  op->apply = cb_system_apply;
  op->context = s;
  op->precondition = cb_system_precondition;
  op->precondition_context = s;
  s->blocks_dirty = true;  // callers write entries before asking for the operator
}

// ------------------------------------------------------------------ vie_gmres
extern "C" void vie_gmres_default_options(vie_gmres_options* o) {
  o->restart = 30;
  o->maximum_iterations = 300;
  o->relative_tolerance = 1e-12;
  o->absolute_tolerance = 0.0;
}

extern "C" const char* vie_iterative_status_name(vie_iterative_status s) {
  switch (s) {
    case VIE_ITERATIVE_CONVERGED: return "converged";
    case VIE_ITERATIVE_ITERATION_LIMIT: return "iteration limit";
    case VIE_ITERATIVE_BREAKDOWN: return "Krylov breakdown";
    case VIE_ITERATIVE_CALLBACK_FAILED: return "operator/preconditioner failure";
    case VIE_ITERATIVE_INVALID_INPUT: return "invalid input";
    case VIE_ITERATIVE_ALLOCATION_FAILED: return "workspace allocation/size failure";
    case VIE_ITERATIVE_NONFINITE: return "nonfinite arithmetic";
  }
  return "unknown status";
}

extern "C" int vie_gmres(const vie_linear_operator* op, const double _Complex* b, double _Complex* x,
                         const vie_gmres_options* options, vie_gmres_report* report) {
  vie_gmres_report local;
  vie_gmres_options defaults;
  if (!report) report = &local;
  report_init(report);
  vie_gmres_default_options(&defaults);
  if (!options) options = &defaults;
  if (!op || !op->apply || op->order == 0 || !b || !x || (const void*)b == (const void*)x ||
      !validate_options(options))
    return 0;
  // this library's operators: the device-resident solve
  if (op->apply == cb_system_apply && op->precondition == cb_system_precondition && op->context)
    return solve_host((dielectric_fft_system*)op->context, b, x, options, nullptr, report);
  if (op->apply == cb_plan_apply && op->precondition == cb_plan_precondition && op->context) {
    dielectric_fft_system* s = plan_system((vie_fft_plan*)op->context);
    if (!s) return 0;
    return solve_host(s, b, x, options, nullptr, report);
  }
  // Other callback operators are not this library's to iterate: nec2c keeps
  // its own vie_gmres (vie_iterative.c:80-199) for them, and the dispatcher
  // that adds the "hip" backend sends only this library's operators here.
  report->status = VIE_ITERATIVE_INVALID_INPUT;
  set_error("vie_gmres: the operator was not made by vie_fft_operator or dielectric_fft_get_operator "
            "of this library; solve it with the host program's own vie_gmres");
  return 0;
}

// ------------------------------------------------------------------ measurement hooks
extern "C" int dielectric_fft_time_apply(dielectric_fft_system* s, vie_hip_precision prec, int repeats,
                                         double* ms_per_apply, double* pass_ms) {
  if (!s || repeats < 1) return 0;
  size_t n = s->nsys();
  if (!ensure_vectors(s) || !upload_blocks(s) || (s->plan && !plan_upload_eps(s->plan)) ||
      !ensure_scratch(s->sw) || !plan_tune_prec(s->plan, prec))
    return 0;
  vie::ReduceScratch rs;
  rs.partial = s->sw.partial.as<c64>();
  const bool mixed = prec == VIE_HIP_MIXED;
  // resident input: a fixed nonzero vector in the system layout
  DevBuf in, out;
  size_t es = mixed ? 8 : 16;
  if (!in.alloc(n * es) || !out.alloc(n * es)) return 0;
  {
    std::vector<c64> h(n);
    for (size_t i = 0; i < n; i++) h[i] = c64{sin(0.7 * (double)i), cos(0.3 * (double)i)};
    HIP_OK(hipMemcpyAsync(s->sw.r.p, h.data(), n * 16, hipMemcpyHostToDevice, s->stream));
    if (mixed) vie::vec_cast<float, double>(in.as<c32>(), s->sw.r.as<c64>(), n, s->stream);
    else vie::vec_cast<double, double>(in.as<c64>(), s->sw.r.as<c64>(), n, s->stream);
    HIP_OK(hipStreamSynchronize(s->stream));
  }
  auto run = [&](hipEvent_t* ev) -> int {
    return mixed ? sys_apply<float>(s, in.as<c32>(), out.as<c32>(), 1, rs, ev)
                 : sys_apply<double>(s, in.as<c64>(), out.as<c64>(), 1, rs, ev);
  };
  hipEvent_t ev[7], t0, t1;
  for (int i = 0; i < 7; i++) HIP_OK(hipEventCreate(&ev[i]));
  HIP_OK(hipEventCreate(&t0));
  HIP_OK(hipEventCreate(&t1));
  int ok = run(nullptr) && run(nullptr);  // warm
  HIP_OK(hipStreamSynchronize(s->stream));
  HIP_OK(hipEventRecord(t0, s->stream));
  for (int r = 0; r < repeats && ok; r++) ok = run(nullptr);
  HIP_OK(hipEventRecord(t1, s->stream));
  HIP_OK(hipEventSynchronize(t1));
  float total = 0;
  HIP_OK(hipEventElapsedTime(&total, t0, t1));
  if (ms_per_apply) *ms_per_apply = total / repeats;
  if (pass_ms) {
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int r = 0; r < repeats && ok; r++) {
      if (!s->plan) break;
      ok = run(ev);
      HIP_OK(hipStreamSynchronize(s->stream));
      for (int q = 0; q < 5; q++) {
        float t;
        HIP_OK(hipEventElapsedTime(&t, ev[q], ev[q + 1]));
        acc[q] += t;
      }
      if (s->c) {
        float t;
        HIP_OK(hipEventElapsedTime(&t, ev[5], ev[6]));
        acc[5] += t;
      }
    }
    for (int q = 0; q < 6; q++) pass_ms[q] = acc[q] / repeats;
  }
  for (int i = 0; i < 7; i++) (void)hipEventDestroy(ev[i]);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  return ok;
}

// 1 when the system's last resident MIXED solve switched to FP64 cycles.
extern "C" int dielectric_fft_last_solve_escalated(const dielectric_fft_system* s) {
  return s && s->sw.escalated ? 1 : 0;
}

// ------------------------------------------------------------------ scattered far field
// The reference's post-processing of a solved volume current
// (dielectric_solver.c:124-217) for many directions at once on the device.
namespace {
struct FarfieldWork {
  std::mutex mu;
  DevBuf centers, sizes, current, dirs, rs, rp, partial, amp;
  hipStream_t stream = nullptr;
};
FarfieldWork& farfield_work() {
  static FarfieldWork* w = new FarfieldWork();  // lives for the process (device buffers)
  return *w;
}
}  // namespace

static int farfield_run(size_t cells, const double* centers, const double* sizes, const double _Complex* current,
                        double k0, int ground, const double _Complex* rs, const double _Complex* rp, size_t count,
                        const double* directions, double _Complex* amplitude) {
  g_last_error.clear();
  if ((cells && (!centers || !sizes || !current)) || (count && (!directions || !amplitude)) || ground < 0 ||
      ground > 2 || (ground == 2 && count && (!rs || !rp)) || !std::isfinite(k0) || !(k0 > 0)) {
    set_error("radiation amplitudes: invalid argument");
    return 0;
  }
  size_t b3c, b3d, bj, bpart;
  if (!mul_ok(cells, 3 * sizeof(double), &b3c) || !mul_ok(count, 3 * sizeof(double), &b3d) ||
      !mul_ok(cells, 3 * sizeof(c64), &bj)) {
    set_error("radiation amplitudes: size overflow");
    return 0;
  }
  if (count == 0) return 1;
  if (!ensure_device()) return 0;
  // one cell size for all (a lattice) lets the kernel keep the cell factor per direction
  int uniform = 1;
  for (size_t c = 1; c < cells && uniform; c++)
    for (int a = 0; a < 3; a++)
      if (sizes[3 * c + a] != sizes[a]) uniform = 0;
  const unsigned slices = vie::farfield_slices(cells, count, ground);
  if (!mul_ok(count, 6 * (size_t)slices * sizeof(c64), &bpart)) {
    set_error("radiation amplitudes: size overflow");
    return 0;
  }
  FarfieldWork& w = farfield_work();
  std::lock_guard<std::mutex> lock(w.mu);
  if (!w.stream) HIP_OK(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking));
  const size_t bsz = uniform ? 3 * sizeof(double) : b3c;
  if (!(w.centers.alloc(b3c ? b3c : 16) && w.sizes.alloc(bsz ? bsz : 16) && w.current.alloc(bj ? bj : 16) &&
        w.dirs.alloc(b3d) && w.partial.alloc(bpart) && w.amp.alloc(count * 3 * sizeof(c64)) &&
        (ground != 2 || (w.rs.alloc(count * sizeof(c64)) && w.rp.alloc(count * sizeof(c64))))))
    return 0;
  hipStream_t st = w.stream;
  if (cells) {
    HIP_OK(hipMemcpyAsync(w.centers.p, centers, b3c, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(w.sizes.p, sizes, bsz, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(w.current.p, current, bj, hipMemcpyHostToDevice, st));
  }
  HIP_OK(hipMemcpyAsync(w.dirs.p, directions, b3d, hipMemcpyHostToDevice, st));
  if (ground == 2) {
    HIP_OK(hipMemcpyAsync(w.rs.p, rs, count * sizeof(c64), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(w.rp.p, rp, count * sizeof(c64), hipMemcpyHostToDevice, st));
  }
  if (!vie::farfield_amplitudes(w.centers.as<double>(), w.sizes.as<double>(), uniform, w.current.as<c64>(), cells,
                                w.dirs.as<double>(), count, k0, ground, w.rs.as<c64>(), w.rp.as<c64>(),
                                w.partial.as<c64>(), slices, w.amp.as<c64>(), st)) {
    set_error("far-field kernels failed to launch");
    return 0;
  }
  HIP_OK(hipMemcpyAsync(amplitude, w.amp.p, count * 3 * sizeof(c64), hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  return 1;
}

extern "C" int vie_hip_radiation_amplitudes(size_t cells, const double* centers, const double* sizes,
                                            const double _Complex* current, double k0, int ground,
                                            const double _Complex* rs, const double _Complex* rp, size_t count,
                                            const double* directions, double _Complex* amplitude) {
  PhaseTimer phase_timer_(PHASE_FAR_FIELD);
  return farfield_run(cells, centers, sizes, current, k0, ground, rs, rp, count, directions, amplitude);
}

extern "C" int vie_hip_scattered_power(size_t cells, const double* centers, const double* sizes,
                                       const double _Complex* current, double k0, int ground, double* power) {
  PhaseTimer phase_timer_(PHASE_FAR_FIELD);
  // the golden-spiral rule of scattered_power (dielectric_solver.c:160-217):
  // the sphere in free space, the upper hemisphere over a PEC ground
  if (!power || (ground != 0 && ground != 1)) {
    set_error("scattered power: invalid argument");
    return 0;
  }
  const size_t count = 974;
  const double golden = kPi * (3.0 - sqrt(5.0));
  std::vector<double> dirs(3 * count);
  for (size_t i = 0; i < count; i++) {
    const double z = 1.0 - (ground ? 1.0 : 2.0) * ((double)i + 0.5) / (double)count;
    const double radial = sqrt(std::max(0.0, 1.0 - z * z)), az = golden * (double)i;
    dirs[3 * i] = radial * cos(az);
    dirs[3 * i + 1] = radial * sin(az);
    dirs[3 * i + 2] = z;
  }
  std::vector<c64> amp(3 * count);
  if (!farfield_run(cells, centers, sizes, current, k0, ground, nullptr, nullptr, count, dirs.data(),
                    (double _Complex*)amp.data()))
    return 0;
  double sum = 0;
  for (size_t i = 0; i < 3 * count; i++) sum += amp[i].re * amp[i].re + amp[i].im * amp[i].im;
  const double eta0 = sqrt(1.25663706212e-6 / 8.8541878128e-12);  // dielectric_solver.c:20-23, 166
  *power = ((ground ? 2.0 : 4.0) * kPi / (double)count) * sum / (2.0 * eta0);
  return 1;
}

// ------------------------------------------------------------------ volume-to-conductor rows
// The wire and patch rows of the coupled system over the volume unknowns
// (dielectric_wire.c:1164-1262), every (conductor, cell) pair at once.
namespace {
struct RowsWork {
  std::mutex mu;
  DevBuf centers, sizes, cond, rows, failed;
  hipStream_t stream = nullptr;
};
RowsWork& rows_work() {
  static RowsWork* w = new RowsWork();  // lives for the process (device buffers)
  return *w;
}
struct RowsArgs {
  double k0;
  size_t cells;
  const double *centers, *sizes;
  int wires;
  const double *wire_points, *wire_tangents;
  int patches;
  const double *patch_points, *patch_t1, *patch_t2, *patch_scale;
  int ground;
};
}  // namespace

// Uploads the geometry and fills the device rows; with repeats > 0 the kernel
// runs that many more times between HIP events and *ms gets the mean.  With
// host_rows it downloads the rows into host_rows (leading dimension ld).
// A consumer of the device result (rows [nrows][3 cells], or columns
// [columns][3 cells]) on the computing stream, before the call returns.
using DeviceSink = std::function<bool(const c64*, hipStream_t)>;

static int rows_run(const RowsArgs& a, double _Complex* host_rows, size_t ld, int repeats, float* ms,
                    const DeviceSink& sink = nullptr) {
  g_last_error.clear();
  const size_t nrows = (size_t)(a.wires > 0 ? a.wires : 0) + 2 * (size_t)(a.patches > 0 ? a.patches : 0);
  if (a.wires < 0 || a.patches < 0 || a.ground < 0 || a.ground > 2 || !std::isfinite(a.k0) || !(a.k0 > 0) ||
      (a.cells && (!a.centers || !a.sizes)) || (a.wires && (!a.wire_points || !a.wire_tangents)) ||
      (a.patches && (!a.patch_points || !a.patch_t1 || !a.patch_t2 || !a.patch_scale)) ||
      (nrows && a.cells && repeats <= 0 && !sink && (!host_rows || ld < 3 * a.cells))) {
    set_error("volume-to-conductor rows: invalid argument");
    return 0;
  }
  if (a.ground == 2) {
    set_error("volume-to-conductor rows: a finite ground (Sommerfeld half space) is not supported on the device");
    return 0;
  }
  size_t b3c, row_bytes, rows_bytes, pitch;
  if (!mul_ok(a.cells, 3 * sizeof(double), &b3c) || !mul_ok(a.cells, 3 * sizeof(c64), &row_bytes) ||
      !mul_ok(nrows, row_bytes, &rows_bytes) || !mul_ok(ld, sizeof(c64), &pitch) ||
      (size_t)a.wires + 2 * (size_t)a.patches > (size_t)INT_MAX) {
    set_error("volume-to-conductor rows: size overflow");
    return 0;
  }
  if (nrows == 0 || a.cells == 0) {
    if (ms) *ms = 0.0f;
    return 1;
  }
  if (!ensure_device()) return 0;
  const size_t nw = (size_t)a.wires, np = (size_t)a.patches;
  const size_t ncond = 6 * nw + 10 * np;  // points, tangents; points, t1, t2, scale
  RowsWork& w = rows_work();
  std::lock_guard<std::mutex> lock(w.mu);
  if (!w.stream) HIP_OK(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking));
  if (!(w.centers.alloc(b3c) && w.sizes.alloc(b3c) && w.cond.alloc(ncond * sizeof(double)) &&
        w.rows.alloc(rows_bytes) && w.failed.alloc(sizeof(unsigned long long))))
    return 0;
  hipStream_t st = w.stream;
  double* cond = w.cond.as<double>();
  double *wp = cond, *wt = wp + 3 * nw, *pp = wt + 3 * nw, *t1 = pp + 3 * np, *t2 = t1 + 3 * np, *ps = t2 + 3 * np;
  HIP_OK(hipMemcpyAsync(w.centers.p, a.centers, b3c, hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(w.sizes.p, a.sizes, b3c, hipMemcpyHostToDevice, st));
  if (nw) {
    HIP_OK(hipMemcpyAsync(wp, a.wire_points, 3 * nw * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(wt, a.wire_tangents, 3 * nw * sizeof(double), hipMemcpyHostToDevice, st));
  }
  if (np) {
    HIP_OK(hipMemcpyAsync(pp, a.patch_points, 3 * np * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(t1, a.patch_t1, 3 * np * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(t2, a.patch_t2, 3 * np * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(ps, a.patch_scale, np * sizeof(double), hipMemcpyHostToDevice, st));
  }
  const unsigned long long none = ~0ull;
  HIP_OK(hipMemcpyAsync(w.failed.p, &none, sizeof(none), hipMemcpyHostToDevice, st));
  // largest cell diagonal (NaN when a size is NaN: the kernel then picks per cell)
  double max_size = 0.0;
  for (size_t c = 0; c < a.cells; c++) {
    const double* z = a.sizes + 3 * c;
    const double d = sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
    max_size = std::isnan(d) || std::isnan(max_size) ? NAN : std::max(max_size, d);
  }
  auto launch = [&]() {
    return vie::conductor_rows(a.k0, w.centers.as<double>(), w.sizes.as<double>(), a.cells, max_size, a.wires, wp, wt,
                               a.patches, pp, t1, t2, ps, a.ground, w.rows.as<c64>(), 3 * a.cells,
                               w.failed.as<unsigned long long>(), st);
  };
  if (!launch()) {
    set_error("volume-to-conductor row kernel failed to launch");
    return 0;
  }
  if (repeats > 0) {
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    int ok = hipEventRecord(e0, st) == hipSuccess;
    for (int r = 0; r < repeats && ok; r++) ok = launch();
    ok = ok && hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
    float t = 0;
    if (ok) ok = hipEventElapsedTime(&t, e0, e1) == hipSuccess;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!ok) {
      set_error("volume-to-conductor row timing failed");
      return 0;
    }
    if (ms) *ms = t / (float)repeats;
  }
  unsigned long long failed = none;
  HIP_OK(hipMemcpyAsync(&failed, w.failed.p, sizeof(failed), hipMemcpyDeviceToHost, st));
  if (sink && !sink(w.rows.as<c64>(), st)) {
    set_error("volume-to-conductor rows: placing the rows failed");
    return 0;
  }
  if (host_rows)
    HIP_OK(hipMemcpy2DAsync(host_rows, pitch, w.rows.p, row_bytes, row_bytes, nrows,
                            hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  if (failed != none) {
    // the reference's message for its first failed task (dielectric_wire.c:1200-1207, :1249-1256)
    const size_t task = (size_t)(failed / a.cells), cell = (size_t)(failed % a.cells);
    if (task < nw)
      set_error("wire %zu, volume cell %zu: volume-to-wire interaction failed", task + 1, cell + 1);
    else
      set_error("patch %zu, volume cell %zu: volume-to-patch interaction failed", task - nw + 1, cell + 1);
    return 0;
  }
  return 1;
}

extern "C" int vie_hip_volume_conductor_rows(double k0, size_t cells, const double* centers, const double* sizes,
                                             int wires, const double* wire_points, const double* wire_tangents,
                                             int patches, const double* patch_points, const double* patch_t1,
                                             const double* patch_t2, const double* patch_scale, int ground,
                                             double _Complex* rows, size_t ld) {
  PhaseTimer phase_timer_(PHASE_ROWS);
  const RowsArgs a{k0, cells, centers, sizes, wires, wire_points, wire_tangents,
                   patches, patch_points, patch_t1, patch_t2, patch_scale, ground};
  return rows_run(a, rows, ld, 0, nullptr);
}

extern "C" int vie_hip_time_volume_conductor_rows(double k0, size_t cells, const double* centers,
                                                  const double* sizes, int wires, const double* wire_points,
                                                  const double* wire_tangents, int patches,
                                                  const double* patch_points, const double* patch_t1,
                                                  const double* patch_t2, const double* patch_scale, int ground,
                                                  int repeats, float* ms) {
  const RowsArgs a{k0, cells, centers, sizes, wires, wire_points, wire_tangents,
                   patches, patch_points, patch_t1, patch_t2, patch_scale, ground};
  if (repeats < 1 || !ms) {
    set_error("volume-to-conductor row timing: invalid argument");
    return 0;
  }
  return rows_run(a, nullptr, 0, repeats, ms);
}

// ------------------------------------------------------------------ conductor-to-volume coupling
// The volume rows of the conductor columns (dw_assemble_wire_to_volume,
// dielectric_wire.c:316-508): NEC's near field of each column's frozen basis
// coefficients averaged over every cell.
namespace {
struct ColumnsWork {
  std::mutex mu;
  DevBuf geom, cells, chi, active, jbase, term_seg, coef, skip, F, out;
  hipStream_t stream = nullptr;
};
ColumnsWork& columns_work() {
  static ColumnsWork* w = new ColumnsWork();  // lives for the process (device buffers)
  return *w;
}
struct ColumnsArgs {
  double k0, wlam;
  size_t cells;
  const double *centers, *sizes;
  const double _Complex* chi;
  int n;
  const double *x, *y, *z, *si, *bi, *cab, *sab, *salp;
  int m;
  const double *pp, *pbi, *t1, *t2;
  size_t columns;
  const double _Complex* sources;
  double rkh;
  int ground;
};
constexpr size_t kColumnsFieldBudget = (size_t)1 << 30;  // bytes of averaged fields per cell chunk
}  // namespace

static int columns_run(const ColumnsArgs& a, double _Complex* coupling, int repeats, float* ms,
                       const DeviceSink& sink = nullptr) {
  g_last_error.clear();
  const bool any = a.cells && a.columns && (a.n > 0 || a.m > 0);
  if (a.n < 0 || a.m < 0 || !std::isfinite(a.k0) || !(a.k0 > 0) || !std::isfinite(a.wlam) || !(a.wlam > 0) ||
      std::isnan(a.rkh) || a.ground < 0 || a.ground > 2 || (a.cells && (!a.centers || !a.sizes || !a.chi)) ||
      (a.n && (!a.x || !a.y || !a.z || !a.si || !a.bi || !a.cab || !a.sab || !a.salp)) ||
      (a.m && (!a.pp || !a.pbi || !a.t1 || !a.t2)) || (a.columns && (a.n || a.m) && !a.sources) ||
      (any && repeats <= 0 && !coupling && !sink)) {
    set_error("conductor-to-volume coupling: invalid argument");
    return 0;
  }
  if (a.ground == 2) {
    set_error("conductor-to-volume coupling: a finite ground (Sommerfeld) is not supported on the device");
    return 0;
  }
  // finite geometry and positive segment lengths: NEC's Romberg integration
  // advances by a fraction of the segment length, which must be finite
  {
    auto finite = [](const double* v, size_t count) {
      for (size_t i = 0; i < count; i++)
        if (!std::isfinite(v[i])) return false;
      return true;
    };
    bool ok = finite(a.centers, 3 * a.cells) && finite(a.sizes, 3 * a.cells);
    const double* seg[8] = {a.x, a.y, a.z, a.si, a.bi, a.cab, a.sab, a.salp};
    for (int k = 0; k < 8 && a.n; k++) ok = ok && finite(seg[k], (size_t)a.n);
    for (int i = 0; i < a.n && ok; i++) ok = a.si[i] > 0 && a.bi[i] >= 0;
    if (a.m) ok = ok && finite(a.pp, 3 * (size_t)a.m) && finite(a.pbi, (size_t)a.m) &&
                  finite(a.t1, 3 * (size_t)a.m) && finite(a.t2, 3 * (size_t)a.m);
    if (!ok) {
      set_error("conductor-to-volume coupling: non-finite geometry or a segment length that is not positive");
      return 0;
    }
  }
  const size_t n = (size_t)a.n, m = (size_t)a.m, stride = 3 * n + 3 * m;
  size_t bcells, bout;
  if (!mul_ok(a.cells, 3 * sizeof(double), &bcells) || !mul_ok(a.columns, 3 * a.cells, &bout) ||
      !mul_ok(bout, sizeof(c64), &bout) || n + m > (size_t)INT_MAX / 4 || a.columns > (size_t)INT_MAX) {
    set_error("conductor-to-volume coupling: size overflow");
    return 0;
  }
  if (!any) {
    if (coupling && a.columns && a.cells) memset(coupling, 0, bout);
    if (ms) *ms = 0.0f;
    return 1;
  }
  // radiating sources: segments with a nonzero coefficient, patches with a
  // nonzero current, in some column; their basis terms and coefficients
  const c64* src = (const c64*)a.sources;
  const size_t cols = a.columns;
  std::vector<int> active, jbase, term_seg;
  std::vector<unsigned char> skip(cols * (n ? n : 1), 0);
  for (size_t s = 0; s < n; s++) {
    bool used = false;
    for (size_t c = 0; c < cols; c++) {
      const c64* v = src + c * stride + 3 * s;
      const bool zero = v[0].re == 0 && v[0].im == 0 && v[1].re == 0 && v[1].im == 0 && v[2].re == 0 && v[2].im == 0;
      skip[c * n + s] = zero;
      used = used || !zero;
    }
    if (used) {
      active.push_back((int)s);
      jbase.push_back((int)term_seg.size());
      for (int k = 0; k < 3; k++) term_seg.push_back((int)s);
    }
  }
  for (size_t p = 0; p < m; p++) {
    bool used = false;
    for (size_t c = 0; c < cols && !used; c++) {
      const c64* v = src + c * stride + 3 * n + 3 * p;
      for (int k = 0; k < 3; k++) used = used || v[k].re != 0 || v[k].im != 0;
    }
    if (used) {
      active.push_back((int)(n + p));
      jbase.push_back((int)term_seg.size());
      term_seg.push_back(-1);
      term_seg.push_back(-1);
    }
  }
  const size_t terms = term_seg.size();
  if (terms == 0) {
    if (coupling) memset(coupling, 0, bout);
    if (ms) *ms = 0.0f;
    return 1;
  }
  // coefficient of each term per column: (a, b, c) of a segment; t1.J and t2.J of a patch
  std::vector<c64> coef(cols * terms);
  for (size_t c = 0; c < cols; c++)
    for (size_t q = 0; q < active.size(); q++) {
      const size_t j = (size_t)jbase[q];
      if ((size_t)active[q] < n) {
        const c64* v = src + c * stride + 3 * (size_t)active[q];
        for (int k = 0; k < 3; k++) coef[c * terms + j + k] = v[k];
      } else {
        const size_t p = (size_t)active[q] - n;
        const c64* v = src + c * stride + 3 * n + 3 * p;
        const double* u1 = a.t1 + 3 * p;
        const double* u2 = a.t2 + 3 * p;
        c64 s1 = {0, 0}, s2 = {0, 0};
        for (int k = 0; k < 3; k++) {
          s1 = s1 + u1[k] * v[k];
          s2 = s2 + u2[k] * v[k];
        }
        coef[c * terms + j] = s1;
        coef[c * terms + j + 1] = s2;
      }
    }
  size_t chunk = kColumnsFieldBudget / (terms * 3 * sizeof(c64));
  if (chunk < 1) chunk = 1;
  if (chunk > a.cells) chunk = a.cells;
  if (!ensure_device()) return 0;
  ColumnsWork& w = columns_work();
  std::lock_guard<std::mutex> lock(w.mu);
  if (!w.stream) HIP_OK(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking));
  const size_t ngeom = 8 * n + 10 * m;
  if (!(w.geom.alloc((ngeom ? ngeom : 1) * sizeof(double)) && w.cells.alloc(2 * bcells) &&
        w.chi.alloc(a.cells * sizeof(c64)) && w.active.alloc(active.size() * sizeof(int)) &&
        w.jbase.alloc(jbase.size() * sizeof(int)) && w.term_seg.alloc(terms * sizeof(int)) &&
        w.coef.alloc(coef.size() * sizeof(c64)) && w.skip.alloc(skip.size()) &&
        w.F.alloc(terms * 3 * chunk * sizeof(c64)) && w.out.alloc(bout)))
    return 0;
  hipStream_t st = w.stream;
  double* g = w.geom.as<double>();
  const double* seg_src[8] = {a.x, a.y, a.z, a.si, a.bi, a.cab, a.sab, a.salp};
  for (int k = 0; k < 8 && n; k++)
    HIP_OK(hipMemcpyAsync(g + k * n, seg_src[k], n * sizeof(double), hipMemcpyHostToDevice, st));
  double* pg = g + 8 * n;  // pp [3m], pbi [m], t1 [3m], t2 [3m]
  if (m) {
    HIP_OK(hipMemcpyAsync(pg, a.pp, 3 * m * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(pg + 3 * m, a.pbi, m * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(pg + 4 * m, a.t1, 3 * m * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(pg + 7 * m, a.t2, 3 * m * sizeof(double), hipMemcpyHostToDevice, st));
  }
  double* dcent = w.cells.as<double>();
  double* dsize = dcent + 3 * a.cells;
  HIP_OK(hipMemcpyAsync(dcent, a.centers, bcells, hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(dsize, a.sizes, bcells, hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(w.chi.p, a.chi, a.cells * sizeof(c64), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(w.active.p, active.data(), active.size() * sizeof(int), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(w.jbase.p, jbase.data(), jbase.size() * sizeof(int), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(w.term_seg.p, term_seg.data(), terms * sizeof(int), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(w.coef.p, coef.data(), coef.size() * sizeof(c64), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(w.skip.p, skip.data(), skip.size(), hipMemcpyHostToDevice, st));
  vie::ConductorSet cs{a.n, a.m, g, g + n, g + 2 * n, g + 3 * n, g + 4 * n, g + 5 * n, g + 6 * n, g + 7 * n,
                       pg, pg + 3 * m, pg + 4 * m, pg + 7 * m, a.rkh};
  auto pass = [&]() {
    for (size_t c0 = 0; c0 < a.cells; c0 += chunk) {
      const size_t nc = a.cells - c0 < chunk ? a.cells - c0 : chunk;
      if (!vie::conductor_volume_fields(cs, w.active.as<int>(), w.jbase.as<int>(), (int)active.size(), dcent, dsize,
                                        c0, nc, a.wlam, a.ground, w.F.as<c64>(), 3 * nc, st) ||
          !vie::conductor_volume_columns(w.F.as<c64>(), 3 * nc, (int)terms, w.term_seg.as<int>(), w.coef.as<c64>(),
                                         w.skip.as<unsigned char>(), a.n > 0 ? a.n : 1, (int)cols, w.chi.as<c64>(),
                                         c0, nc, a.k0, w.out.as<c64>(), 3 * a.cells, st))
        return false;
    }
    return true;
  };
  if (!pass()) {
    set_error("conductor-to-volume kernels failed to launch");
    return 0;
  }
  if (repeats > 0) {
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    bool ok = hipEventRecord(e0, st) == hipSuccess;
    for (int r = 0; r < repeats && ok; r++) ok = pass();
    ok = ok && hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
    float t = 0;
    if (ok) ok = hipEventElapsedTime(&t, e0, e1) == hipSuccess;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!ok) {
      set_error("conductor-to-volume timing failed");
      return 0;
    }
    if (ms) *ms = t / (float)repeats;
  }
  if (sink && !sink(w.out.as<c64>(), st)) {
    set_error("conductor-to-volume coupling: placing the columns failed");
    return 0;
  }
  if (coupling) HIP_OK(hipMemcpyAsync(coupling, w.out.p, bout, hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  return 1;
}

extern "C" int vie_hip_conductor_volume_columns(double k0, double wlam, size_t cells, const double* centers,
                                                const double* sizes, const double _Complex* chi, int segments,
                                                const double* x, const double* y, const double* z, const double* si,
                                                const double* bi, const double* cab, const double* sab,
                                                const double* salp, int patches, const double* patch_points,
                                                const double* patch_area, const double* patch_t1,
                                                const double* patch_t2, size_t columns,
                                                const double _Complex* sources, double rkh, int ground,
                                                double _Complex* coupling) {
  PhaseTimer phase_timer_(PHASE_COLUMNS);
  const ColumnsArgs a{k0,  wlam, cells, centers, sizes,    chi,          segments,   x,        y,
                      z,   si,   bi,    cab,     sab,      salp,         patches,    patch_points, patch_area,
                      patch_t1, patch_t2, columns, sources, rkh, ground};
  return columns_run(a, coupling, 0, nullptr);
}

extern "C" int vie_hip_time_conductor_volume_columns(double k0, double wlam, size_t cells, const double* centers,
                                                     const double* sizes, const double _Complex* chi, int segments,
                                                     const double* x, const double* y, const double* z,
                                                     const double* si, const double* bi, const double* cab,
                                                     const double* sab, const double* salp, int patches,
                                                     const double* patch_points, const double* patch_area,
                                                     const double* patch_t1, const double* patch_t2, size_t columns,
                                                     const double _Complex* sources, double rkh, int ground,
                                                     int repeats, float* ms) {
  const ColumnsArgs a{k0,  wlam, cells, centers, sizes,    chi,          segments,   x,        y,
                      z,   si,   bi,    cab,     sab,      salp,         patches,    patch_points, patch_area,
                      patch_t1, patch_t2, columns, sources, rkh, ground};
  if (repeats < 1 || !ms) {
    set_error("conductor-to-volume timing: invalid argument");
    return 0;
  }
  return columns_run(a, nullptr, repeats, ms);
}

// ------------------------------------------------------------------ cross blocks straight into a system
// Places [c][3 cells] results (reference cell order) into the system's device
// block (box order) and its FP32 copy, the block zeroed first.
static int system_block_target(dielectric_fft_system* s, size_t cells, const char* what, DevBuf*& b64,
                               DevBuf*& b32, bool rows) {
  if (!s || !s->plan || s->c == 0 || cells != s->cells) {
    set_error("%s: the system has no conductors or volume cells, or the cell count differs", what);
    return 0;
  }
  const size_t nv = 3 * s->nbox(), bytes = s->c * nv;
  b64 = rows ? &s->B64 : &s->D64;
  b32 = rows ? &s->B32 : &s->D32;
  if (!(b64->alloc(bytes * 16) && b32->alloc(bytes * 8))) return 0;
  HIP_OK(hipStreamSynchronize(s->stream));  // nothing of the system's still reads the block
  HIP_OK(hipMemset(b64->p, 0, bytes * 16));
  HIP_OK(hipMemset(b32->p, 0, bytes * 8));
  return 1;
}

static DeviceSink system_block_sink(dielectric_fft_system* s, DevBuf* b64, DevBuf* b32) {
  return [s, b64, b32](const c64* res, hipStream_t st) {
    const size_t c = s->c, nbox = s->nbox(), nv = 3 * nbox, cells = s->cells;
    for (size_t r = 0; r < c; r++)
      vie::to_box<double>(res + r * 3 * cells, 0, cells, s->plan->map_dev.as<size_t>(), nbox,
                          b64->as<c64>() + r * nv, st);
    vie::vec_cast<float, double>(b32->as<c32>(), b64->as<c64>(), c * nv, st);
    return hipGetLastError() == hipSuccess;
  };
}

extern "C" int dielectric_fft_assemble_conductor_rows(dielectric_fft_system* s, double k0, const double* centers,
                                                      const double* sizes, int wires, const double* wire_points,
                                                      const double* wire_tangents, int patches,
                                                      const double* patch_points, const double* patch_t1,
                                                      const double* patch_t2, const double* patch_scale,
                                                      int ground) {
  PhaseTimer phase_timer_(PHASE_ROWS);
  g_last_error.clear();
  DevBuf *b64 = nullptr, *b32 = nullptr;
  if (!s || wires < 0 || patches < 0 || (size_t)wires + 2 * (size_t)patches != (s ? s->c : 0)) {
    set_error("conductor rows into a system: wires + 2 patches must equal the system's conductor count");
    return 0;
  }
  if (!system_block_target(s, s->cells, "conductor rows into a system", b64, b32, true)) return 0;
  const RowsArgs a{k0, s->cells, centers, sizes, wires, wire_points, wire_tangents,
                   patches, patch_points, patch_t1, patch_t2, patch_scale, ground};
  // until the rows are in place the block is neither the host's nor the device's
  s->rows_on_device = true;
  s->gen++;  // the device operator changes
  if (!rows_run(a, nullptr, 0, 0, nullptr, system_block_sink(s, b64, b32))) {
    s->rows_on_device = false;
    s->blocks_dirty = true;  // the next upload restores the host block
    return 0;
  }
  return 1;
}

extern "C" int dielectric_fft_assemble_conductor_columns(
    dielectric_fft_system* s, double k0, double wlam, const double* centers, const double* sizes,
    const double _Complex* chi, int segments, const double* x, const double* y, const double* z, const double* si,
    const double* bi, const double* cab, const double* sab, const double* salp, int patches,
    const double* patch_points, const double* patch_area, const double* patch_t1, const double* patch_t2,
    const double _Complex* sources, double rkh, int ground) {
  PhaseTimer phase_timer_(PHASE_COLUMNS);
  g_last_error.clear();
  DevBuf *b64 = nullptr, *b32 = nullptr;
  if (!s) {
    set_error("conductor columns into a system: no system");
    return 0;
  }
  if (!system_block_target(s, s->cells, "conductor columns into a system", b64, b32, false)) return 0;
  const ColumnsArgs a{k0,  wlam, s->cells, centers, sizes,    chi,          segments,   x,        y,
                      z,   si,   bi,       cab,     sab,      salp,         patches,    patch_points, patch_area,
                      patch_t1, patch_t2, s->c, sources, rkh, ground};
  s->cols_on_device = true;
  s->gen++;  // the device operator changes
  if (!columns_run(a, nullptr, 0, nullptr, system_block_sink(s, b64, b32))) {
    s->cols_on_device = false;
    s->blocks_dirty = true;
    return 0;
  }
  return 1;
}

// ------------------------------------------------------------------ far field of the coupled solve
// dw_far_field (dielectric_wire.c:768-791): the wires' and patches' field
// (ffld) times NEC's wavelength plus the volume amplitude projected on theta
// and phi; dw_radiated_power (:793-840) over the golden spiral.
namespace {
struct TotalWork {
  std::mutex mu;
  DevBuf geom, cur, pcur, ang, e;
  hipStream_t stream = nullptr;
};
TotalWork& total_work() {
  static TotalWork* w = new TotalWork();  // lives for the process (device buffers)
  return *w;
}
}  // namespace

static int total_run(double k0, double wlam, int n, const double* x, const double* y, const double* z,
                     const double* si, const double* cab, const double* sab, const double* salp,
                     const double _Complex* seg_cur, int m, const double* pp, const double* pbi,
                     const double _Complex* pat_cur, size_t cells, const double* centers, const double* sizes,
                     const double _Complex* vol_cur, int ground, size_t count, const double* theta,
                     const double* phi, double _Complex* et, double _Complex* ep) {
  g_last_error.clear();
  if (n < 0 || m < 0 || !std::isfinite(wlam) || !(wlam > 0) || (ground != 0 && ground != 1) ||
      (n && (!x || !y || !z || !si || !cab || !sab || !salp || !seg_cur)) || (m && (!pp || !pbi || !pat_cur)) ||
      (count && (!theta || !phi || !et || !ep))) {
    set_error("total far field: invalid argument");
    return 0;
  }
  if (count == 0) return 1;
  const size_t N = (size_t)n, M = (size_t)m;
  std::vector<c64> cet(count, c64{0, 0}), cep(count, c64{0, 0});
  if (n || m) {
    if (!ensure_device()) return 0;
    TotalWork& w = total_work();
    std::lock_guard<std::mutex> lock(w.mu);
    if (!w.stream) HIP_OK(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking));
    if (!(w.geom.alloc((7 * N + 4 * M + 1) * sizeof(double)) && w.cur.alloc((3 * N + 1) * sizeof(c64)) &&
          w.pcur.alloc((3 * M + 1) * sizeof(c64)) && w.ang.alloc(2 * count * sizeof(double)) &&
          w.e.alloc(2 * count * sizeof(c64))))
      return 0;
    hipStream_t st = w.stream;
    double* g = w.geom.as<double>();
    const double* segs[7] = {x, y, z, si, cab, sab, salp};
    for (int k = 0; k < 7 && n; k++) HIP_OK(hipMemcpyAsync(g + k * N, segs[k], N * 8, hipMemcpyHostToDevice, st));
    if (m) {
      HIP_OK(hipMemcpyAsync(g + 7 * N, pp, 3 * M * 8, hipMemcpyHostToDevice, st));
      HIP_OK(hipMemcpyAsync(g + 7 * N + 3 * M, pbi, M * 8, hipMemcpyHostToDevice, st));
      HIP_OK(hipMemcpyAsync(w.pcur.p, pat_cur, 3 * M * sizeof(c64), hipMemcpyHostToDevice, st));
    }
    if (n) HIP_OK(hipMemcpyAsync(w.cur.p, seg_cur, 3 * N * sizeof(c64), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(w.ang.p, theta, count * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(w.ang.as<double>() + count, phi, count * 8, hipMemcpyHostToDevice, st));
    vie::ConductorFar cf{n,           m,          g,          g + N,   g + 2 * N,          g + 3 * N, g + 4 * N,
                         g + 5 * N,   g + 6 * N,  w.cur.as<c64>(), g + 7 * N, g + 7 * N + 3 * M, w.pcur.as<c64>()};
    if (!vie::conductor_far_field(cf, ground, count, w.ang.as<double>(), w.ang.as<double>() + count, w.e.as<c64>(),
                                  w.e.as<c64>() + count, st)) {
      set_error("conductor far-field kernel failed to launch");
      return 0;
    }
    HIP_OK(hipMemcpyAsync(cet.data(), w.e.p, count * sizeof(c64), hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(cep.data(), w.e.as<c64>() + count, count * sizeof(c64), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
  }
  std::vector<c64> amp;
  if (cells) {
    std::vector<double> dirs(3 * count);
    for (size_t d = 0; d < count; d++) {
      dirs[3 * d] = sin(theta[d]) * cos(phi[d]);
      dirs[3 * d + 1] = sin(theta[d]) * sin(phi[d]);
      dirs[3 * d + 2] = cos(theta[d]);
    }
    amp.resize(3 * count);
    if (!farfield_run(cells, centers, sizes, vol_cur, k0, ground, nullptr, nullptr, count, dirs.data(),
                      (double _Complex*)amp.data()))
      return 0;
  }
  c64* ot = (c64*)et;
  c64* op = (c64*)ep;
  for (size_t d = 0; d < count; d++) {
    c64 t_ = wlam * cet[d], p_ = wlam * cep[d];
    if (cells) {
      const double th = theta[d], ph = phi[d];
      const double tv[3] = {cos(th) * cos(ph), cos(th) * sin(ph), -sin(th)}, pv[3] = {-sin(ph), cos(ph), 0.0};
      for (int a = 0; a < 3; a++) {
        t_ = t_ + tv[a] * amp[3 * d + a];
        p_ = p_ + pv[a] * amp[3 * d + a];
      }
    }
    ot[d] = t_;
    op[d] = p_;
  }
  return 1;
}

extern "C" int vie_hip_total_far_field(double k0, double wlam, int segments, const double* x, const double* y,
                                       const double* z, const double* si, const double* cab, const double* sab,
                                       const double* salp, const double _Complex* segment_current, int patches,
                                       const double* patch_points, const double* patch_area,
                                       const double _Complex* patch_current, size_t cells, const double* centers,
                                       const double* sizes, const double _Complex* volume_current, int ground,
                                       size_t count, const double* theta, const double* phi,
                                       double _Complex* e_theta, double _Complex* e_phi) {
  PhaseTimer phase_timer_(PHASE_FAR_FIELD);
  return total_run(k0, wlam, segments, x, y, z, si, cab, sab, salp, segment_current, patches, patch_points,
                   patch_area, patch_current, cells, centers, sizes, volume_current, ground, count, theta, phi,
                   e_theta, e_phi);
}

extern "C" int vie_hip_total_radiated_power(double k0, double wlam, int segments, const double* x, const double* y,
                                            const double* z, const double* si, const double* cab,
                                            const double* sab, const double* salp,
                                            const double _Complex* segment_current, int patches,
                                            const double* patch_points, const double* patch_area,
                                            const double _Complex* patch_current, size_t cells,
                                            const double* centers, const double* sizes,
                                            const double _Complex* volume_current, int ground, size_t count,
                                            double* power) {
  PhaseTimer phase_timer_(PHASE_FAR_FIELD);
  if (!power || count == 0 || count > ((size_t)1 << 28)) {
    set_error("total radiated power: invalid argument");
    return 0;
  }
  // the golden-spiral rule of dw_radiated_power (dielectric_wire.c:793-840)
  std::vector<double> th(count), ph(count);
  for (size_t i = 0; i < count; i++) {
    const double zc = 1.0 - (ground ? 1.0 : 2.0) * ((double)i + 0.5) / (double)count;
    th[i] = acos(zc);
    ph[i] = kPi * (3.0 - sqrt(5.0)) * (double)i;
  }
  std::vector<c64> et(count), ep(count);
  if (!total_run(k0, wlam, segments, x, y, z, si, cab, sab, salp, segment_current, patches, patch_points,
                 patch_area, patch_current, cells, centers, sizes, volume_current, ground, count, th.data(),
                 ph.data(), (double _Complex*)et.data(), (double _Complex*)ep.data()))
    return 0;
  double sum = 0;
  for (size_t i = 0; i < count; i++)
    sum += et[i].re * et[i].re + et[i].im * et[i].im + ep[i].re * ep[i].re + ep[i].im * ep[i].im;
  *power = sum * ((ground ? 2.0 : 4.0) * kPi / (double)count) / (2.0 * 376.73);  // NEC's eta (nec2c.h)
  return 1;
}

// Frees the device buffers the far-field, conductor-row, conductor-column and
// coupled far-field entry points keep between calls (they are reallocated on
// the next call).
// ------------------------------------------------------------------ mesh relationship checks
namespace {
struct MeshWork {
  std::mutex mu;
  DevBuf boxes, cond, part, first, res;
  hipStream_t stream = nullptr;
};
MeshWork& mesh_work() {
  static MeshWork* w = new MeshWork();  // lives for the process (device buffers)
  return *w;
}
}  // namespace

// The overlap scan of validate_box_relationships (dielectric_mesh.c:81-123)
// over resident corners; repeats > 1 times the scan (average ms).
static int overlap_run(const double* minimum, const double* maximum, size_t count, long long* left, long long* right,
                       int repeats, double* ms) {
  g_last_error.clear();
  size_t bytes;
  if ((count && (!minimum || !maximum)) || !left || !right || repeats < 1 ||
      !mul_ok(count, 6 * sizeof(double), &bytes) || count > ((size_t)1 << 31)) {
    set_error("box overlap: invalid argument");
    return 0;
  }
  *left = *right = -1;
  if (!ensure_device()) return 0;
  MeshWork& w = mesh_work();
  std::lock_guard<std::mutex> lock(w.mu);
  if (!w.stream) HIP_OK(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking));
  if (!(w.boxes.alloc(bytes) && w.first.alloc(sizeof(unsigned long long)))) return 0;
  hipStream_t st = w.stream;
  double* mn = w.boxes.as<double>();
  double* mx = mn + 3 * count;
  if (count) {
    HIP_OK(hipMemcpyAsync(mn, minimum, 3 * count * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(mx, maximum, 3 * count * sizeof(double), hipMemcpyHostToDevice, st));
  }
  unsigned long long* first = w.first.as<unsigned long long>();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ms) {
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventRecord(e0, st));
  }
  for (int r = 0; r < repeats; r++)
    if (!vie::box_overlap(mn, mx, count, first, st)) {
      set_error("box overlap kernel failed to launch");
      return 0;
    }
  if (ms) HIP_OK(hipEventRecord(e1, st));
  unsigned long long key = ~0ull;
  HIP_OK(hipMemcpyAsync(&key, first, sizeof(key), hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  if (ms) {
    float t = 0;
    HIP_OK(hipEventElapsedTime(&t, e0, e1));
    *ms = t / repeats;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  if (key != ~0ull) {
    *left = (long long)(key / count);
    *right = (long long)(key % count);
  }
  return 1;
}

extern "C" int vie_hip_box_overlap(const double* minimum, const double* maximum, size_t count, long long* left,
                                   long long* right) {
  PhaseTimer phase_timer_(PHASE_MESH);
  return overlap_run(minimum, maximum, count, left, right, 1, nullptr);
}

extern "C" int vie_hip_time_box_overlap(const double* minimum, const double* maximum, size_t count, int repeats,
                                        long long* left, long long* right, double* ms) {
  if (!ms) {
    set_error("box overlap timing: invalid argument");
    return 0;
  }
  return overlap_run(minimum, maximum, count, left, right, repeats, ms);
}

// evaluate_clearance (dielectric_mesh.c:215-316) over boxes and NEC's wire
// segments and patches.
static int clearance_run(const double* minimum, const double* maximum, const double* cell_longest, size_t boxes,
                         int segments, const double* x, const double* y, const double* z, const double* si,
                         const double* cab, const double* sab, const double* salp, const double* bi, int patches,
                         const double* patch_points, const double* patch_area, vie_hip_clearance* out, int repeats,
                         double* ms) {
  g_last_error.clear();
  size_t bytes;
  if (out) {  // the reference's values for "nothing evaluated", so a failed call leaves no garbage
    memset(out, 0, sizeof(*out));
    out->worst_margin = DBL_MAX;
    out->minimum_gap = DBL_MAX;
  }
  if (!out || segments < 0 || patches < 0 || repeats < 1 ||
      (boxes && (!minimum || !maximum || !cell_longest)) ||
      (segments && (!x || !y || !z || !si || !cab || !sab || !salp || !bi)) ||
      (patches && (!patch_points || !patch_area)) || !mul_ok(boxes, 7 * sizeof(double), &bytes) ||
      boxes > ((size_t)1 << 40)) {
    set_error("conductor clearance: invalid argument");
    return 0;
  }
  vie::ClearanceResult r{DBL_MAX, 0.0, 0.0, DBL_MAX, 0.0, 0, 0, 0};
  bool any_cells = false;
  for (size_t b = 0; b < boxes && !any_cells; b++) any_cells = cell_longest[b] > 0.0;
  out->has_conductors = any_cells && (segments > 0 || patches > 0);
  if (out->has_conductors) {
    if (!ensure_device()) return 0;
    MeshWork& w = mesh_work();
    std::lock_guard<std::mutex> lock(w.mu);
    if (!w.stream) HIP_OK(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking));
    const size_t N = (size_t)segments, M = (size_t)patches;
    const size_t pairs = boxes * (N + M);
    if (!(w.boxes.alloc(bytes) && w.cond.alloc((8 * N + 4 * M + 1) * sizeof(double)) &&
          w.part.alloc(vie::clearance_partials(pairs) * 32) && w.res.alloc(sizeof(vie::ClearanceResult))))
      return 0;
    hipStream_t st = w.stream;
    double* bx = w.boxes.as<double>();
    HIP_OK(hipMemcpyAsync(bx, minimum, 3 * boxes * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(bx + 3 * boxes, maximum, 3 * boxes * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(bx + 6 * boxes, cell_longest, boxes * 8, hipMemcpyHostToDevice, st));
    double* g = w.cond.as<double>();
    const double* segs[8] = {x, y, z, si, cab, sab, salp, bi};
    for (int k = 0; k < 8 && N; k++) HIP_OK(hipMemcpyAsync(g + k * N, segs[k], N * 8, hipMemcpyHostToDevice, st));
    // patch centres are given as M x 3 points; the kernel reads them per axis
    std::vector<double> pt(3 * M);
    for (size_t p = 0; p < M; p++)
      for (int a = 0; a < 3; a++) pt[a * M + p] = patch_points[3 * p + a];
    if (M) {
      HIP_OK(hipMemcpyAsync(g + 8 * N, pt.data(), 3 * M * 8, hipMemcpyHostToDevice, st));
      HIP_OK(hipMemcpyAsync(g + 8 * N + 3 * M, patch_area, M * 8, hipMemcpyHostToDevice, st));
    }
    vie::ClearanceArgs A{bx,        bx + 3 * boxes, bx + 6 * boxes, boxes,     (long long)N,  (long long)M,
                         g,         g + N,          g + 2 * N,      g + 3 * N, g + 4 * N,     g + 5 * N,
                         g + 6 * N, g + 7 * N,      g + 8 * N,      g + 8 * N + M, g + 8 * N + 2 * M,
                         g + 8 * N + 3 * M};
    vie::ClearanceResult* dres = w.res.as<vie::ClearanceResult>();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ms) {
      HIP_OK(hipEventCreate(&e0));
      HIP_OK(hipEventCreate(&e1));
      HIP_OK(hipEventRecord(e0, st));
    }
    for (int k = 0; k < repeats; k++)
      if (!vie::conductor_clearance(A, w.part.p, dres, st)) {
        set_error("conductor clearance kernel failed to launch");
        return 0;
      }
    if (ms) HIP_OK(hipEventRecord(e1, st));
    HIP_OK(hipMemcpyAsync(&r, dres, sizeof(r), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (ms) {
      float t = 0;
      HIP_OK(hipEventElapsedTime(&t, e0, e1));
      *ms = t / repeats;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
    }
  } else if (ms) {
    *ms = 0.0;
  }
  out->worst_margin = r.worst_margin;
  out->worst_gap = r.worst_gap;
  out->worst_required = r.worst_required;
  out->minimum_gap = r.minimum_gap;
  out->required_gap = r.required_gap;
  out->worst_box = r.worst_box;
  out->worst_conductor = r.worst_conductor;
  out->worst_is_patch = r.worst_is_patch;
  out->insufficient = r.worst_margin < -1.0e-12;  // dielectric_mesh.c:305
  return 1;
}

extern "C" int vie_hip_conductor_clearance(const double* minimum, const double* maximum, const double* cell_longest,
                                           size_t boxes, int segments, const double* x, const double* y,
                                           const double* z, const double* si, const double* cab, const double* sab,
                                           const double* salp, const double* bi, int patches,
                                           const double* patch_points, const double* patch_area,
                                           vie_hip_clearance* out) {
  PhaseTimer phase_timer_(PHASE_MESH);
  return clearance_run(minimum, maximum, cell_longest, boxes, segments, x, y, z, si, cab, sab, salp, bi, patches,
                       patch_points, patch_area, out, 1, nullptr);
}

extern "C" int vie_hip_time_conductor_clearance(const double* minimum, const double* maximum,
                                                const double* cell_longest, size_t boxes, int segments,
                                                const double* x, const double* y, const double* z, const double* si,
                                                const double* cab, const double* sab, const double* salp,
                                                const double* bi, int patches, const double* patch_points,
                                                const double* patch_area, int repeats, vie_hip_clearance* out,
                                                double* ms) {
  if (!ms) {
    set_error("conductor clearance timing: invalid argument");
    return 0;
  }
  return clearance_run(minimum, maximum, cell_longest, boxes, segments, x, y, z, si, cab, sab, salp, bi, patches,
                       patch_points, patch_area, out, repeats, ms);
}

extern "C" int vie_hip_release_workspaces(void) {
  {
    FarfieldWork& w = farfield_work();
    std::lock_guard<std::mutex> lock(w.mu);
    if (w.stream) HIP_OK(hipStreamSynchronize(w.stream));
    for (DevBuf* b : {&w.centers, &w.sizes, &w.current, &w.dirs, &w.rs, &w.rp, &w.partial, &w.amp}) b->release();
  }
  {
    RowsWork& w = rows_work();
    std::lock_guard<std::mutex> lock(w.mu);
    if (w.stream) HIP_OK(hipStreamSynchronize(w.stream));
    for (DevBuf* b : {&w.centers, &w.sizes, &w.cond, &w.rows, &w.failed}) b->release();
  }
  {
    ColumnsWork& w = columns_work();
    std::lock_guard<std::mutex> lock(w.mu);
    if (w.stream) HIP_OK(hipStreamSynchronize(w.stream));
    for (DevBuf* b : {&w.geom, &w.cells, &w.chi, &w.active, &w.jbase, &w.term_seg, &w.coef, &w.skip, &w.F, &w.out})
      b->release();
  }
  {
    TotalWork& w = total_work();
    std::lock_guard<std::mutex> lock(w.mu);
    if (w.stream) HIP_OK(hipStreamSynchronize(w.stream));
    for (DevBuf* b : {&w.geom, &w.cur, &w.pcur, &w.ang, &w.e}) b->release();
  }
  {
    MeshWork& w = mesh_work();
    std::lock_guard<std::mutex> lock(w.mu);
    if (w.stream) HIP_OK(hipStreamSynchronize(w.stream));
    for (DevBuf* b : {&w.boxes, &w.cond, &w.part, &w.first, &w.res}) b->release();
  }
  return 1;
}

extern "C" int dielectric_fft_last_solve_recycled(const dielectric_fft_system* s) {
  return s && s->sw.rec_ready ? s->sw.rk : 0;
}

extern "C" int vie_hip_harmonic_ritz(const double _Complex* hbar, int m, int k, double _Complex* p) {
  namespace D = vie::dense;
  if (!hbar || !p || m < 1 || m > vie::kMaxRestart || k < 1 || k > m) return 0;
  const c64* h = (const c64*)hbar;
  D::Mat Hb(m + 1, m), P;
  for (int c = 0; c < m; c++)
    for (int r = 0; r <= m; r++) Hb(r, c) = D::zd(h[(size_t)c * (m + 1) + r].re, h[(size_t)c * (m + 1) + r].im);
  if (!D::harmonic_ritz(Hb, k, P)) return 0;
  c64* out = (c64*)p;
  for (int c = 0; c < k; c++)
    for (int r = 0; r < m; r++) out[(size_t)c * m + r] = c64{P(r, c).re, P(r, c).im};
  return 1;
}

// Mean basis column of the matvec Arnoldi steps the system's own solver
// issued since the last reset (the column sets the reduction's width).
extern "C" int dielectric_fft_solve_stats(dielectric_fft_system* s, double* mean_column, unsigned long long* steps,
                                          int reset) {
  if (!s) return 0;
  if (mean_column) *mean_column = s->sw.col_steps ? s->sw.col_sum / (double)s->sw.col_steps : 0.0;
  if (steps) *steps = s->sw.col_steps;
  if (reset) {
    s->sw.col_sum = 0;
    s->sw.col_steps = 0;
  }
  return 1;
}

// Times the two streaming kernels of a delayed-CGS2 step at column j on the
// system's resident solver workspace (which it overwrites).  Bytes per launch
// with n unknowns and S bytes a value:
//   reduction : columns 0..j-1, u and w read                   (j + 2) n S
//   update    : columns 0..j-1 read, u and w read and written  (j + 4) n S
extern "C" int dielectric_fft_time_arnoldi(dielectric_fft_system* s, vie_hip_precision prec, int j, int repeats,
                                           double* ms_reduce, double* ms_update, double* bytes_reduce,
                                           double* bytes_update) {
  if (!s || j < 1 || j + 1 > vie::kMaxRestart || repeats < 1) return 0;
  const int p = prec == VIE_HIP_MIXED ? VIE_HIP_MIXED : VIE_HIP_FP64;
  SolverWork& w = s->sw;
  if (!ensure_solver(s, w, std::max(w.prec >= 0 ? w.m : 0, j + 1), p)) return 0;
  vie::ReduceScratch rs;
  rs.partial = w.partial.as<c64>();
  float a = 0, b = 0;
  int ok = p == VIE_HIP_MIXED
               ? vie::dcgs2_time<float>(w.V.as<c32>(), w.ldv, w.n, w.state.as<vie::ArnoldiState>(), rs, j, repeats,
                                        s->stream, &a, &b)
               : vie::dcgs2_time<double>(w.V.as<c64>(), w.ldv, w.n, w.state.as<vie::ArnoldiState>(), rs, j, repeats,
                                         s->stream, &a, &b);
  if (!ok) return 0;
  const double S = p == VIE_HIP_MIXED ? 8.0 : 16.0, n = (double)w.n;
  if (ms_reduce) *ms_reduce = a;
  if (ms_update) *ms_update = b;
  if (bytes_reduce) *bytes_reduce = (j + 2) * n * S;
  if (bytes_update) *bytes_update = (j + 4) * n * S;
  return 1;
}

// Algorithmic bytes per fused apply (pruned five-pass pipeline):
//   Z fwd : 3 N x in (+ N eps for the fused Jacobi), 3 nx ny Pz out
//   Y fwd : 3 nx ny Pz in, 3 nx Py Pz out
//   X     : 3 nx Py Pz in, 6 (Px/2+1)(Py/2+1)(Pz/2+1) spectrum, 3 nx Py Pz out
//   Y inv : 3 nx Py Pz in, 3 nx ny Pz out
//   Z inv : 3 nx ny Pz in, 3 N x + N eps re-read, 3 N out
//   cross : c^2 + 2 c (3 N) block values + the 3 N volume vector read for B z_v
//           and the 3 N read-modify-write of y_v for D z_c
// with N the unknown-box cell count and 8 (FP32) or 16 (FP64) bytes a value.
extern "C" int dielectric_fft_apply_bytes(const dielectric_fft_system* s, vie_hip_precision prec, double* pass,
                                          double* total) {
  if (!s) return 0;
  double S = prec == VIE_HIP_MIXED ? 8.0 : 16.0;
  double p[6] = {0, 0, 0, 0, 0, 0};
  if (s->plan) {
    const vie::Geom& g = s->plan->g;
    double N = (double)g.nbox(), a = 3.0 * g.nx * g.ny * g.Pz, b = 3.0 * g.nx * g.Py * g.Pz;
    double oct = 6.0 * (g.Px / 2 + 1) * (g.Py / 2 + 1) * (g.Pz / 2 + 1);
    p[0] = (3 * N + N + a) * S;
    p[1] = (a + b) * S;
    p[2] = (2 * b + oct + (g.ground ? 6.0 * (g.Px / 2 + 1) * (g.Py / 2 + 1) * g.Pz : 0.0)) * S;
    p[3] = (b + a) * S;
    p[4] = (a + 3 * N + N + 3 * N) * S;
    // fused z+y passes move no intermediate workspace: their bytes are the
    // fused kernel's, and the skipped y pass moves nothing
    const unsigned char* var = s->plan->var[prec == VIE_HIP_MIXED ? 1 : 0];
    if (var[1] == vie::kFuseSkip) {
      p[0] = (3 * N + N + b) * S;
      p[1] = 0;
    }
    if (var[3] == vie::kFuseSkip) {
      p[3] = 0;
      p[4] = (b + 3 * N + N + 3 * N) * S;
    }
  }
  double c = (double)s->c, nv = s->plan ? 3.0 * s->plan->g.nbox() : 0.0;
  if (s->c) p[5] = (c * c + 2 * c * nv + (nv ? 3 * nv : 0)) * S;
  double t = 0;
  for (int q = 0; q < 6; q++) {
    if (pass) pass[q] = p[q];
    t += p[q];
  }
  if (total) *total = t;
  return 1;
}

// ------------------------------------------------------------------ batched solves
// Many independent solves in flight at once.  Each job owns a pooled solver
// workspace (Krylov basis, vectors, reduction scratch, pinned status) and is
// advanced by event queries only, so the host interleaves the Arnoldi steps
// of every job: jobs of one system share its stream (they run in submission
// order), jobs of different systems run on different streams and overlap on
// the GPU.  The object is not reentrant; one host thread drives it.
namespace {
enum { JB_RUNNING, JB_DOWNLOAD, JB_READY };
struct BatchJob {
  std::unique_ptr<SolverWork> w;
  int state = JB_RUNNING;
  bool reported = false;
  size_t nu = 0;
  double t_ready = 0;
};
}  // namespace

struct vie_hip_batch {
  std::vector<std::unique_ptr<SolverWork>> pool;
  std::map<long long, BatchJob> jobs;
  std::vector<long long> order;  // submission order of unfinished jobs
  long long next_id = 1;
  // recycled subspaces shared among the jobs of one system (off by default)
  bool share = false;
  unsigned long long shared_published = 0, shared_adopted = 0;
  // multi-job step groups (0: FP64 cycles, 1: FP32 cycles); declared last,
  // so they are destroyed (their streams drained) before the jobs
  std::unique_ptr<StepGroup> groups[2];
};

extern "C" vie_hip_batch* vie_hip_batch_create(void) {
  g_last_error.clear();
  return new (std::nothrow) vie_hip_batch();
}

// The step groups are made with the first job (their streams need the device).
static int batch_groups(vie_hip_batch* b) {
  if (!multijob_enabled()) return 1;
  for (int i = 0; i < 2; i++) {
    if (b->groups[i]) continue;
    std::unique_ptr<StepGroup> g(new (std::nothrow) StepGroup());
    if (!g || !g->init()) {
      set_error("batch: step group setup failed");
      return 0;
    }
    g->fp32 = i == 1;
    b->groups[i] = std::move(g);
  }
  return 1;
}

extern "C" void vie_hip_batch_destroy(vie_hip_batch* b) {
  if (!b) return;
  // queued device work must end before its buffers are freed
  for (auto& kv : b->jobs)
    if (kv.second.w && kv.second.w->sys) (void)hipStreamSynchronize(kv.second.w->sys->stream);
  delete b;
}

static std::unique_ptr<SolverWork> pool_take(vie_hip_batch* b) {
  if (b->pool.empty()) return std::unique_ptr<SolverWork>(new (std::nothrow) SolverWork());
  std::unique_ptr<SolverWork> w = std::move(b->pool.back());
  b->pool.pop_back();
  return w;
}

static int host_io_reserve(SolverWork& w, size_t nu) {
  if (w.host_io_cap >= nu && w.host_io) return 1;
  if (w.host_io) (void)hipHostFree(w.host_io);
  w.host_io = nullptr;
  w.host_io_cap = 0;
  HIP_OK(hipHostMalloc((void**)&w.host_io, std::max<size_t>(nu, 1) * 16, hipHostMallocDefault));
  w.host_io_cap = nu;
  return 1;
}

extern "C" size_t vie_hip_batch_job_bytes(const dielectric_fft_system* s, const vie_gmres_options* options,
                                          const vie_hip_solve_options* so) {
  if (!s) return 0;
  vie_gmres_options defaults = dielectric_gmres_settings;
  if (!options) options = &defaults;
  size_t m = options->restart, n = s->c + s->v;
  if (m > n) m = n;
  if (m > options->maximum_iterations) m = options->maximum_iterations;
  int prec = so && so->precision == VIE_HIP_MIXED ? VIE_HIP_MIXED : VIE_HIP_FP64;
  return solver_bytes(s->nsys(), (int)m, prec, recycle_count(so, prec, (int)m));
}

extern "C" int vie_hip_batch_reserve(vie_hip_batch* b, size_t count, dielectric_fft_system* s,
                                     const vie_gmres_options* options) {
  if (!b || !s) return 0;
  vie_gmres_options defaults = dielectric_gmres_settings;
  if (!options) options = &defaults;
  if (!validate_options(options)) return 0;
  size_t m = device_restart(s->c + s->v, options);
  if (!m) return 0;
  while (b->pool.size() < count) {
    std::unique_ptr<SolverWork> w(new (std::nothrow) SolverWork());
    if (!w) return 0;
    b->pool.push_back(std::move(w));
  }
  // FP64-sized bases serve either precision
  for (size_t i = 0; i < count; i++) {
    SolverWork& w = *b->pool[b->pool.size() - 1 - i];
    if (!ensure_solver(s, w, (int)m, VIE_HIP_FP64) || !ensure_recycle(w) || !host_io_reserve(w, s->c + s->v))
      return 0;
    w.sys = nullptr;
  }
  HIP_OK(hipDeviceSynchronize());
  return 1;
}

extern "C" long long vie_hip_batch_submit(vie_hip_batch* b, dielectric_fft_system* s, const double _Complex* rhs,
                                          const vie_gmres_options* options, const vie_hip_solve_options* so) {
  g_last_error.clear();
  if (!b || !s || !rhs) {
    set_error("batch submit: null argument");
    return 0;
  }
  vie_gmres_options defaults = dielectric_gmres_settings;
  if (!options) options = &defaults;
  if (!validate_options(options)) {
    set_error("batch submit: invalid GMRES options");
    return 0;
  }
  if (!so) {
    init_solve_defaults();
    so = &g_solve;
  }
  if (!batch_groups(b)) return 0;
  BatchJob job;
  job.w = pool_take(b);
  if (!job.w) return 0;
  SolverWork& w = *job.w;
  vie_gmres_report rep;
  report_init(&rep);
  size_t nu = s->c + s->v, n = s->nsys();
  if (!job_prepare(s, w, options, so, &rep) || !host_io_reserve(w, nu)) {
    b->pool.push_back(std::move(job.w));
    return 0;
  }
  for (int i = 0; i < 2; i++) w.groups[i] = b->groups[i].get();
  w.share = b->share;
  // b: user layout -> pinned -> device -> system layout; x0 = 0
  memcpy(w.host_io, rhs, nu * 16);
  hipStream_t st = s->stream;
  if (hipMemcpyAsync(w.ax.p, w.host_io, nu * 16, hipMemcpyHostToDevice, st) != hipSuccess) {
    set_error("batch submit: right-hand side copy failed");
    b->pool.push_back(std::move(job.w));
    return 0;
  }
  vie::vec_zero<double>(w.b.as<c64>(), n, st);
  vie::to_box<double>(w.ax.as<c64>(), s->c, s->cells, s->plan ? s->plan->map_dev.as<size_t>() : nullptr, s->nbox(),
                      w.b.as<c64>(), st);
  vie::vec_zero<double>(w.x.as<c64>(), n, st);
  if (!job_start(w, options, so, true)) {
    b->pool.push_back(std::move(job.w));
    return 0;
  }
  job.nu = nu;
  long long id = b->next_id++;
  b->jobs.emplace(id, std::move(job));
  b->order.push_back(id);
  return id;
}

// Advances one job; returns 1 when it changed state.
static int batch_step(vie_hip_batch* b, BatchJob& j, bool blocking) {
  SolverWork& w = *j.w;
  if (j.state == JB_RUNNING) {
    if (!job_advance_any(w, blocking)) return 0;
    b->shared_published += w.published;
    b->shared_adopted += w.adopted;
    // solution: system layout -> user layout -> pinned
    dielectric_fft_system* s = w.sys;
    hipStream_t st = s->stream;
    vie::from_box<double>(w.x.as<c64>(), s->c, s->cells, s->plan ? s->plan->map_dev.as<size_t>() : nullptr, s->nbox(),
                          w.r.as<c64>(), st);
    if (hipMemcpyAsync(w.host_io, w.r.p, j.nu * 16, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipEventRecord(w.ev_misc, st) != hipSuccess) {
      w.result = 0;
      w.rep.status = VIE_ITERATIVE_CALLBACK_FAILED;
      j.state = JB_READY;
      j.t_ready = now_s();
      return 1;
    }
    j.state = JB_DOWNLOAD;
    return 1;
  }
  if (j.state == JB_DOWNLOAD) {
    int r = job_ready(w.ev_misc, blocking);
    if (r == 0) return 0;
    if (r < 0) {
      w.result = 0;
      w.rep.status = VIE_ITERATIVE_CALLBACK_FAILED;
    }
    j.state = JB_READY;
    j.t_ready = now_s();
    return 1;
  }
  return 0;
}

static int groups_progress(vie_hip_batch* b) {
  int changed = 0;
  if (b->groups[0]) changed += group_progress<double>(*b->groups[0]);
  if (b->groups[1]) changed += group_progress<float>(*b->groups[1]);
  return changed;
}

// One non-blocking sweep over the unfinished jobs; returns how many changed.
static int batch_sweep(vie_hip_batch* b) {
  int changed = groups_progress(b);
  size_t keep = 0;
  for (size_t i = 0; i < b->order.size(); i++) {
    long long id = b->order[i];
    BatchJob& j = b->jobs[id];
    while (j.state != JB_READY && batch_step(b, j, false)) changed++;
    if (j.state != JB_READY) b->order[keep++] = id;
  }
  b->order.resize(keep);
  return changed + groups_progress(b);
}

extern "C" int vie_hip_batch_share_subspaces(vie_hip_batch* b, int on) {
  if (!b) return 0;
  b->share = on != 0;
  return 1;
}

extern "C" int vie_hip_batch_shared_stats(const vie_hip_batch* b, unsigned long long* published,
                                          unsigned long long* adopted) {
  if (!b) return 0;
  if (published) *published = b->shared_published;
  if (adopted) *adopted = b->shared_adopted;
  return 1;
}

extern "C" int vie_hip_batch_group_stats(const vie_hip_batch* b, unsigned long long* group_steps,
                                         unsigned long long* job_steps) {
  if (!b) return 0;
  unsigned long long gs = 0, js = 0;
  for (auto& g : b->groups)
    if (g) {
      gs += g->launches;
      js += g->steps;
    }
  if (group_steps) *group_steps = gs;
  if (job_steps) *job_steps = js;
  return 1;
}

extern "C" size_t vie_hip_batch_progress(vie_hip_batch* b, long long* ready, size_t max_ready, size_t* running) {
  if (!b) return 0;
  batch_sweep(b);
  size_t count = 0;
  for (auto& kv : b->jobs) {
    if (count >= max_ready) break;
    if (kv.second.state == JB_READY && !kv.second.reported) {
      kv.second.reported = true;
      if (ready) ready[count] = kv.first;
      count++;
    }
  }
  if (running) *running = b->order.size();
  return count;
}

extern "C" size_t vie_hip_batch_running(const vie_hip_batch* b) { return b ? b->order.size() : 0; }

extern "C" int vie_hip_batch_wait(vie_hip_batch* b, long long id) {
  if (!b) return 0;
  auto it = b->jobs.find(id);
  if (it == b->jobs.end()) {
    set_error("batch wait: unknown job %lld", id);
    return 0;
  }
  while (it->second.state != JB_READY) {
    if (!batch_sweep(b)) std::this_thread::yield();
  }
  return 1;
}

extern "C" int vie_hip_batch_wait_all(vie_hip_batch* b) {
  if (!b) return 0;
  while (!b->order.empty())
    if (!batch_sweep(b)) std::this_thread::yield();
  return 1;
}

// 1 when the job has converged (the vie_gmres return value); x (may be NULL)
// receives the solution in the user layout, report (may be NULL) the report,
// seconds (may be NULL) the time from submission to completion.
extern "C" int vie_hip_batch_result(vie_hip_batch* b, long long id, double _Complex* x, vie_gmres_report* report,
                                    double* seconds) {
  if (!b) return 0;
  auto it = b->jobs.find(id);
  if (it == b->jobs.end() || it->second.state != JB_READY) {
    set_error("batch result: job %lld is not finished", id);
    if (report) report_init(report);
    return 0;
  }
  BatchJob& j = it->second;
  if (x && j.w->rep.status != VIE_ITERATIVE_INVALID_INPUT) memcpy(x, j.w->host_io, j.nu * 16);
  if (report) *report = j.w->rep;
  if (seconds) *seconds = j.t_ready - j.w->t_submit;
  j.reported = true;
  return j.w->result;
}

extern "C" int vie_hip_batch_release(vie_hip_batch* b, long long id) {
  if (!b) return 0;
  auto it = b->jobs.find(id);
  if (it == b->jobs.end()) return 0;
  if (it->second.state != JB_READY) {
    set_error("batch release: job %lld is still running", id);
    return 0;
  }
  it->second.w->sys = nullptr;
  b->pool.push_back(std::move(it->second.w));
  b->jobs.erase(it);
  return 1;
}
