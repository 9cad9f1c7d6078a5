/*
 * Extensions of libvie_hip.so beyond the reference's interface.  Each entry
 * point below is new: the reference has no device-resident solve (its GMRES
 * calls the operator with host pointers, vie_iterative.c:146-148), no
 * precision choice and no batching.
 */
#ifndef VIE_HIP_H
#define VIE_HIP_H
#include "dielectric_fft.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  VIE_HIP_FP64 = 0,  /* every matvec, basis vector and reduction in FP64 */
  VIE_HIP_MIXED = 1  /* FP32 matvecs and Krylov basis inside each cycle; FP64
                        true residual b - A x, solution update, dot-product
                        accumulation and Givens least squares.  A cycle that
                        cannot lower the true residual in FP32 hands the rest
                        of the solve to FP64 cycles. */
} vie_hip_precision;

typedef struct {
  vie_hip_precision precision;
  /* MIXED: an FP32 cycle ends once |g_{j+1}| <= max(target, inner_reduction * beta);
     0 selects the default (1e-4 with a recycled subspace, else 1e-5). */
  double inner_reduction;
  /* Recycled Krylov subspace (GCRO with a subspace fixed after the first
     cycle): after the first cycle the k harmonic Ritz vectors of smallest
     modulus span U, with C = A M^{-1} U orthonormal, and every later cycle
     minimises over span(U) + its own Krylov space of m - k columns.  k > 0
     sets k (at most 32 and m / 2), -1 turns it off, 0 selects the default:
     off for VIE_HIP_FP64 (the reference's restarted GMRES), on for
     VIE_HIP_MIXED, where the restarts of the FP32 cycles lose most. */
  int recycle;
} vie_hip_solve_options;

/* Hybrid system over a real (lossy dielectric) ground z < 0 of relative
   permittivity ground_re + i ground_im (new: the reference's FFT path rejects
   a real ground, dielectric_ground.c:28-29; its dense and GMRES paths add the
   reflected tensor vie_halfspace_cuboids to every pair, dielectric_ground.c:
   116-131).  The volume operator adds the reflected tensor, a convolution in
   x, y and a correlation in z, through a second spectrum.  Every cell must
   lie strictly above z = 0 (dielectric_ground.c:41-45); the Jacobi
   preconditioner keeps the free-space self term, as the PEC variant does.
   Pass nec2c's dielectric_ground_epsilon(frequency) (dielectric_ground.c:97-100).
   Otherwise as dielectric_fft_prepare. */
int dielectric_fft_prepare_ground(dielectric_fft_system **out, const dielectric_model *model,
                                  const dielectric_mesh_result *mesh, size_t conductors, double k0,
                                  size_t extra_bytes, double ground_re, double ground_im,
                                  dielectric_diagnostic *diagnostic);
/* Device-resident solve of a system prepared by dielectric_fft_prepare[_pec]
   (new; the drop-in for vie_gmres at dielectric_wire.c:1303 and
   dielectric_solver.c:526).  Same contract and report as vie_gmres. */
int dielectric_fft_solve(dielectric_fft_system *system, const VIE_CDOUBLE *rhs, VIE_CDOUBLE *x,
                         const vie_gmres_options *options, vie_gmres_report *report);
/* Same for a bare volume plan of vie_fft_create. */
int vie_fft_solve(vie_fft_plan *plan, const VIE_CDOUBLE *rhs, VIE_CDOUBLE *x,
                  const vie_gmres_options *options, vie_gmres_report *report);
/* Precision used by dielectric_fft_solve / vie_fft_solve / vie_gmres dispatch
   (default VIE_HIP_FP64, or VIE_HIP_PRECISION=mixed in the environment). */
void vie_hip_set_solve_options(const vie_hip_solve_options *options);
void vie_hip_get_solve_options(vie_hip_solve_options *options);

/* HIP backend storage cap for dielectric_fft_prepare (default 64 GiB or
   VIE_HIP_BUDGET_BYTES); the reference's is the 512 MiB macro. */
size_t dielectric_fft_budget(void);
void dielectric_fft_set_budget(size_t bytes);
/* Forces the next apply/solve to upload C/B/D again after entry writes. */
void dielectric_fft_blocks_changed(dielectric_fft_system *system);
/* Underlying volume plan of a system (NULL when it has no cells). */
vie_fft_plan *dielectric_fft_plan(dielectric_fft_system *system);

/* Last error text of this thread ("" when none). */
const char *vie_hip_last_error(void);
/* Number of HIP devices visible (0 without a GPU; never initialises one
   when the runtime reports none). */
int vie_hip_device_count(void);
/* Free and total memory of the current device (for admission control);
   0 without a device. */
int vie_hip_memory_info(size_t *free_bytes, size_t *total_bytes);

/* Bulk per-cell permittivity of a plan (count == cells); each value passes
   the same checks as vie_fft_set_permittivity.  Returns 0 and changes
   nothing when any value is rejected. */
int vie_fft_set_permittivities(vie_fft_plan *plan, const VIE_CDOUBLE *epsilon_r, size_t count);
/* Operator application at an explicit precision on host vectors in the
   reference layout: y = A M^{-1} x when `jacobi` is nonzero (the product the
   Krylov cycle forms), else y = A x.  FP32 rounds x to complex float before the
   device work and returns y widened to double. */
int dielectric_fft_apply_ex(dielectric_fft_system *system, vie_hip_precision precision, int jacobi,
                            const VIE_CDOUBLE *x, VIE_CDOUBLE *y);
/* Allocation-free variant of dielectric_fft_apply_ex (new; no counterpart in
   the reference).  For callers that cannot allocate device memory: code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
   Same arguments and result as dielectric_fft_apply_ex, but every temporary
   (staging, box-order vectors, FFT workspace, reduction partials) lives in a
   96 MiB array of device memory that the library holds on each GPU it runs on
   until the process ends; no device memory is allocated or freed.  Systems
   needing more scratch than that (dielectric_fft_apply_noalloc_bytes) are
   refused with an error (FP64 fits grids up to about 48^3, FP32 larger).  The
   system's device copies must be current: call dielectric_fft_reserve, or
   any apply or solve, after the last dielectric_fft_entry or permittivity
   change.  All calls on one GPU share its array, so calls that overlap on one
   GPU (from several threads) race; each call finishes its device work before
   it returns.  1 on success, 0 with vie_hip_last_error(). */
int dielectric_fft_apply_noalloc(dielectric_fft_system *system, vie_hip_precision precision, int jacobi,
                                 const VIE_CDOUBLE *x, VIE_CDOUBLE *y);
/* Scratch bytes one dielectric_fft_apply_noalloc of this system needs at this
   precision (new). */
size_t dielectric_fft_apply_noalloc_bytes(const dielectric_fft_system *system, vie_hip_precision precision);
/* Pass-shape autotuner: times each launch-shape variant (lines per
   workgroup, workgroup size) of the five volume passes on this plan's
   geometry and keeps the fastest per pass and precision.  Choices are cached
   per size class (box, padded extents, PEC) for the life of the process.
   Runs by itself before a plan's first solve or apply; VIE_HIP_TUNE=0 in the
   environment keeps variant 0.  variants (optional, 10 ints) receives the
   FP64 then the FP32 choice of each pass. */
int vie_fft_tune(vie_fft_plan *plan, int *variants);
/* Timing phases of the library's hot-path work (new; the reference times its
   own phases with performance_add, performance.h:8-17): wall seconds and calls,
   cumulative over the process or since vie_hip_phases_reset, of "fft_build"
   (prepare / vie_fft_create with the spectrum), "fft_upload" (C/B/D to the
   device when called outside a solve), "fft_gmres" (device solves),
   "fft_matvec" (operator applications on host vectors), "fft_conductor_rows",
   "fft_conductor_columns", "fft_far_field" and "fft_mesh_checks".  Work done
   inside another timed call counts once, in the outer phase.  Writes up to max
   phases that have calls (names are static strings) and returns how many there
   are. */
typedef struct {
  const char *name;
  double seconds;
  unsigned long long calls;
} vie_hip_phase;
size_t vie_hip_phases(vie_hip_phase *phases, size_t max);
void vie_hip_phases_reset(void);
/* ---- compact per-voxel resource (new: the reference's format writes a
   per-voxel model as one BOX record per voxel, dielectric_io.c:472-512, whose
   mesher then tests every pair of boxes, dielectric_mesh.c:81-123).  A
   resource "NEC2C-DIELECTRIC 4" holds MATERIAL records (ER [EI|SIGMA], or
   TABLE with SAMPLE records, as versions 2 and 3), one
     VOXELS <nx> <ny> <nz> PITCH <m> ORIGIN <x0> <y0> <z0> MATERIALS|PERMITTIVITY <file>
   record and END.  The side file holds nx*ny*nz little-endian uint32 material
   ids (0: no dielectric) or (Re, Im) float64 permittivities ((1, 0): vacuum),
   x slowest and z fastest; voxel (i, j, k) spans ORIGIN + PITCH (i, j, k) to
   ORIGIN + PITCH (i+1, j+1, k+1).  The lattice is one by construction, so no
   overlap check is needed.  Diagnostics name the record, value and line as
   the reference's parser does. */
typedef struct vie_hip_voxels vie_hip_voxels;
int vie_hip_voxels_load(vie_hip_voxels **out, const char *path, dielectric_diagnostic *diagnostic);
void vie_hip_voxels_free(vie_hip_voxels *voxels);
/* Lattice and the number of dielectric voxels (the cells of a system). */
int vie_hip_voxels_info(const vie_hip_voxels *voxels, size_t shape[3], double *pitch, double origin[3],
                        size_t *cells);
/* Per cell (the dielectric voxels in storage order): centre and size (3 each,
   m) and material id (PERMITTIVITY: the cell index + 1); any may be NULL. */
int vie_hip_voxels_cells(const vie_hip_voxels *voxels, double *centers, double *sizes, int *material_ids);
/* dielectric_fft_prepare (pec = 0) or dielectric_fft_prepare_pec (pec = 1) of
   the voxel model: the same system the equivalent cell list would give. */
int dielectric_fft_prepare_voxels(dielectric_fft_system **out, const vie_hip_voxels *voxels, size_t conductors,
                                  double k0, size_t extra_bytes, int pec, dielectric_diagnostic *diagnostic);
/* Sets the launch variant of each pass (10 ints: FP64 then FP32, passes
   z-forward, y-forward, x, y-inverse, z-inverse) instead of the tuner's choice,
   for tests and measurements (new).  Every variant computes the same
   arithmetic.  0 when a variant does not exist on this geometry. */
int vie_fft_set_variants(vie_fft_plan *plan, const int *variants);
/* The reflected tensor of a passive lower half-space z < 0 (permittivity
   eps_re + i eps_im, eps_im <= 0) between the cells of an nx x ny x nz lattice
   of cubes of edge h whose lowest layer is centred at height z0 > h / 2, on
   the device (new; the reference evaluates it one pair at a time,
   vie_halfspace_cuboids, vie_halfspace.c:230, for its dense GN 2 path):
   out[((c nx + dx) ny + dy)(2 nz - 1) + s] for components c = xx, yy, zz, xy,
   xz, yz (zx = -xz, zy = -yz), lateral offsets dx, dy >= 0 (in cells, test
   minus source) and layer sums s = iz_test + iz_source; Gauss-Legendre order
   `order` per transverse-wavenumber segment.  The field it maps a source
   current to is the reference's N (dimensionless).  1 on success. */
int vie_hip_ground_table(int nx, int ny, int nz, double h, double z0, double k0, double eps_re, double eps_im,
                         int order, VIE_CDOUBLE *out);
/* Waits for all work on the current device. */
int vie_hip_synchronize(void);
/* Copies `bytes` (rounded down to 16) between two device buffers `repeats`
   times with a plain streaming kernel; returns average ms per copy.  Used to
   calibrate HBM traffic counters against a kernel of known size. */
double vie_hip_copy_probe(size_t bytes, int repeats);

/* ---- resident work (for benchmarks and servers) ----
   Right-hand side and solution stay in device memory between calls. */
int dielectric_fft_upload_rhs(dielectric_fft_system *system, const VIE_CDOUBLE *rhs);
/* Makes everything a later solve with these options needs resident: uploads
   C/B/D and the permittivity, allocates the Krylov basis for either
   precision and the solver vectors.  Solves then allocate and copy nothing. */
int dielectric_fft_reserve(dielectric_fft_system *system, const vie_gmres_options *options);
int dielectric_fft_solve_resident(dielectric_fft_system *system, const vie_gmres_options *options,
                                  const vie_hip_solve_options *solve, vie_gmres_report *report);
int dielectric_fft_download_solution(dielectric_fft_system *system, VIE_CDOUBLE *x);
/* Times `repeats` device-resident operator applications (input resident,
   Jacobi fused) at the given precision with HIP events on the system's
   stream.  pass_ms (optional, 6 entries) receives the average per launch of
   the five volume passes and the conductor coupling. */
int dielectric_fft_time_apply(dielectric_fft_system *system, vie_hip_precision precision, int repeats,
                              double *ms_per_apply, double *pass_ms);
/* 1 when the system's last device solve in VIE_HIP_MIXED ran out of FP32
   accuracy (an FP32 cycle gained less than 2x on the FP64 true residual) and
   finished with FP64 cycles. */
int dielectric_fft_last_solve_escalated(const dielectric_fft_system *system);
/* Size of the recycled subspace the system's last device solve built and used
   (0 when its cycles were plain: recycling off, one cycle was enough, or the
   first cycle was too short to supply one). */
int dielectric_fft_last_solve_recycled(const dielectric_fft_system *system);
/* Scattered far-field amplitudes of a solved volume current for many
   directions at once on the device (the reference evaluates
   dielectric_radiation_amplitude, dielectric_solver.c:124-158, one direction
   at a time; dielectric.h:233-238).  cells: centers[3 cells] and sizes[3 cells]
   in m; current[3 cells] cell-major, as that function takes it; k0 in 1/m;
   directions[3 count] unit vectors.  ground 0: free space; 1: PEC plane z = 0
   (mirrored cells, reflection -1 / +1 as dielectric_ground.c:133-144 with
   gnd.iperf == 1); 2: a finite ground with the caller's per-direction Fresnel
   coefficients rs[count], rp[count] (vie_halfspace_fresnel there).  Writes
   amplitude[3 count] (E r e^{ikr}, SI).  1 on success, 0 with
   vie_hip_last_error() on invalid input or a device failure. */
int vie_hip_radiation_amplitudes(size_t cells, const double *centers, const double *sizes,
                                 const VIE_CDOUBLE *current, double k0, int ground, const VIE_CDOUBLE *rs,
                                 const VIE_CDOUBLE *rp, size_t count, const double *directions,
                                 VIE_CDOUBLE *amplitude);
/* Scattered power in W by the reference's 974-point golden-spiral rule over
   the sphere (ground 0) or the upper hemisphere over a PEC plane (ground 1)
   (scattered_power, dielectric_solver.c:160-217). */
int vie_hip_scattered_power(size_t cells, const double *centers, const double *sizes, const VIE_CDOUBLE *current,
                            double k0, int ground, double *power);
/* Volume-to-conductor rows of the coupled wire/patch + volume system on the
   device: the two task loops of dielectric_wire.c:1164-1262 (wire rows from
   dw_point_cuboid_n, patch rows from dw_point_cuboid_h, :101-214), every
   (conductor, cell) pair at once.  Cells: centers[3 cells], sizes[3 cells] in
   m; k0 in 1/m, omega = k0 c0.  Wire w: match point wire_points[3 w], unit
   tangent wire_tangents[3 w] (the reference's x, y, z and cab, sab, salp);
   patch p: patch_points[3 p], tangents patch_t1[3 p], patch_t2[3 p], scale
   patch_scale[p] (psalp).  ground 0: free space; 1: PEC plane z = 0 (the
   mirrored cell with column signs -1, -1, +1, gnd.iperf == 1); a finite
   ground is rejected.  Writes row w = t.N/(i omega eps0) for each wire, rows
   wires + 2p and wires + 2p + 1 = -psalp t2.H and -psalp t1.H for each patch,
   entry (row, 3 cell + q) at rows[row ld + 3 cell + q], ld >= 3 cells (the
   reference's system row with rows = &entry(0, conductor_unknowns), ld =
   total).  1 on success; 0 with vie_hip_last_error() on invalid input, a
   device failure, or a quadrature node on a match point (the reference's
   message for its first failing pair: "wire W, volume cell C: ..."). */
int vie_hip_volume_conductor_rows(double k0, size_t cells, const double *centers, const double *sizes, int wires,
                                  const double *wire_points, const double *wire_tangents, int patches,
                                  const double *patch_points, const double *patch_t1, const double *patch_t2,
                                  const double *patch_scale, int ground, VIE_CDOUBLE *rows, size_t ld);
/* Conductor-to-volume coupling on the device: the volume rows of every
   conductor column, coupling[column * 3 cells + 3 cell + q] =
   -i omega eps0 chi[cell] <E_q>, <E> NEC's near field of the column's frozen
   basis coefficients averaged over the cell's 6^3 Gauss points (the task
   loop dw_wire_to_volume_range / dw_average_wire_field,
   dielectric_wire.c:316-420, over nefld, fields.c:1123-1370; the columns of
   dw_assemble_wire_to_volume, :422-508).  omega = k0 c0.
   NEC's classic engine: segment data in wavelengths as NEC holds it (x, y, z,
   si, bi, cab, sab, salp [segments]); patches: patch_points[3 patches],
   patch_area[patches] (pbi), patch_t1/patch_t2[3 patches], in wavelengths;
   wlam NEC's wavelength in m (cells are in m: centers, sizes[3 cells]; chi[cells]
   the contrast eps_r - 1 of each cell).  sources[column * (3 segments + 3
   patches)]: (a, b, c) per segment, then the patch current (3 per patch), as
   dw_prepare_conductor_sources freezes them.  rkh: NEC's lumped-element range
   (wavelengths).  ground 0: free space; 1: PEC plane (gnd.iperf == 1); a
   finite ground is refused.  1 on success; 0 with vie_hip_last_error(). */
int vie_hip_conductor_volume_columns(double k0, double wlam, size_t cells, const double *centers,
                                     const double *sizes, const VIE_CDOUBLE *chi, int segments, const double *x,
                                     const double *y, const double *z, const double *si, const double *bi,
                                     const double *cab, const double *sab, const double *salp, int patches,
                                     const double *patch_points, const double *patch_area, const double *patch_t1,
                                     const double *patch_t2, size_t columns, const VIE_CDOUBLE *sources,
                                     double rkh, int ground, VIE_CDOUBLE *coupling);
/* The same columns on device buffers only, `repeats` times between HIP
   events after one untimed pass; *ms receives the mean time per pass (new;
   for measurement). */
int vie_hip_time_conductor_volume_columns(double k0, double wlam, size_t cells, const double *centers,
                                          const double *sizes, const VIE_CDOUBLE *chi, int segments,
                                          const double *x, const double *y, const double *z, const double *si,
                                          const double *bi, const double *cab, const double *sab,
                                          const double *salp, int patches, const double *patch_points,
                                          const double *patch_area, const double *patch_t1,
                                          const double *patch_t2, size_t columns, const VIE_CDOUBLE *sources,
                                          double rkh, int ground, int repeats, float *ms);
/* The conductor rows (B block) computed straight into a prepared system's
   device storage (new): the volume-to-conductor rows of
   vie_hip_volume_conductor_rows for the system's cells (centers, sizes: 3
   values for each of the system's cells, in the mesh order given to
   dielectric_fft_prepare; wires + 2 patches must equal the system's conductor
   count), placed in the system's device layout without a
   host round trip.  Replaces the whole B block; the host copy behind
   dielectric_fft_entry is refreshed from the device at the next
   dielectric_fft_entry call (write C through dielectric_fft_entry before
   this call, not after, to avoid that download).  1 on success. */
int dielectric_fft_assemble_conductor_rows(dielectric_fft_system *system, double k0, const double *centers,
                                           const double *sizes, int wires, const double *wire_points,
                                           const double *wire_tangents, int patches, const double *patch_points,
                                           const double *patch_t1, const double *patch_t2,
                                           const double *patch_scale, int ground);
/* The conductor columns (D block) computed straight into a prepared system's
   device storage (new), as vie_hip_conductor_volume_columns with one column
   per conductor unknown (sources[c * (3 segments + 3 patches)]); same
   storage rules as dielectric_fft_assemble_conductor_rows. */
int dielectric_fft_assemble_conductor_columns(dielectric_fft_system *system, double k0, double wlam,
                                              const double *centers, const double *sizes, const VIE_CDOUBLE *chi,
                                              int segments, const double *x, const double *y, const double *z,
                                              const double *si, const double *bi, const double *cab,
                                              const double *sab, const double *salp, int patches,
                                              const double *patch_points, const double *patch_area,
                                              const double *patch_t1, const double *patch_t2,
                                              const VIE_CDOUBLE *sources, double rkh, int ground);
/* The same rows computed on device buffers only, `repeats` times between HIP
   events on the library's stream after one untimed pass; *ms receives the mean
   kernel time per pass (new; for measurement). */
int vie_hip_time_volume_conductor_rows(double k0, size_t cells, const double *centers, const double *sizes,
                                       int wires, const double *wire_points, const double *wire_tangents,
                                       int patches, const double *patch_points, const double *patch_t1,
                                       const double *patch_t2, const double *patch_scale, int ground, int repeats,
                                       float *ms);
/* Far field of the coupled solve on the device (dw_far_field,
   dielectric_wire.c:768-791, for many directions at once): the solved
   wires' and patches' field by NEC's ffld / fflds (radiation.c:36-373,
   classic engine) times wlam, plus the volume current's amplitude
   (vie_hip_radiation_amplitudes) projected on theta and phi.  Segments as
   NEC holds them (x, y, z, si in wavelengths; cab, sab, salp), their solved
   current coefficients segment_current[3 segments] (constant, sine, cosine:
   crnt.air + i crnt.aii, bir/bii, cir/cii, in NEC's amperes per wavelength);
   patches patch_points[3 patches], patch_area[patches] (wavelengths) with
   patch_current[3 patches] (crnt.cur[n + 3 j ..]); the volume cells and
   current as vie_hip_radiation_amplitudes takes them; ground 0 free space,
   1 PEC plane.  Writes e_theta[count], e_phi[count] (r E in V) for the
   directions (theta[d], phi[d]).  1 on success. */
int vie_hip_total_far_field(double k0, double wlam, int segments, const double *x, const double *y, const double *z,
                            const double *si, const double *cab, const double *sab, const double *salp,
                            const VIE_CDOUBLE *segment_current, int patches, const double *patch_points,
                            const double *patch_area, const VIE_CDOUBLE *patch_current, size_t cells,
                            const double *centers, const double *sizes, const VIE_CDOUBLE *volume_current,
                            int ground, size_t count, const double *theta, const double *phi,
                            VIE_CDOUBLE *e_theta, VIE_CDOUBLE *e_phi);
/* Radiated power in W of the coupled solve by dw_radiated_power's golden
   spiral of `count` directions (dielectric_wire.c:793-840; the reference
   uses 3896 and 974): the sphere in free space, the upper hemisphere over a
   PEC plane, NEC's eta = 376.73 in the denominator as there. */
int vie_hip_total_radiated_power(double k0, double wlam, int segments, const double *x, const double *y,
                                 const double *z, const double *si, const double *cab, const double *sab,
                                 const double *salp, const VIE_CDOUBLE *segment_current, int patches,
                                 const double *patch_points, const double *patch_area,
                                 const VIE_CDOUBLE *patch_current, size_t cells, const double *centers,
                                 const double *sizes, const VIE_CDOUBLE *volume_current, int ground, size_t count,
                                 double *power);
/* Box-interior overlap scan of the mesher's validate_box_relationships
   (dielectric_mesh.c:81-123): boxes [minimum, maximum] (count x 3 each, the
   model's BOX records in order).  *left, *right <- the 0-based indices of the
   first pair (left < right, left the outer loop as there) whose interiors
   overlap on all three axes, or -1, -1 when none does; the reference's
   diagnostic reports that pair's region ids and the line of box `right`.
   The region lookups that precede the scan stay with the caller.  All
   count (count - 1) / 2 pairs are tested on the GPU.  1 on success. */
int vie_hip_box_overlap(const double *minimum, const double *maximum, size_t count, long long *left,
                        long long *right);
/* vie_hip_box_overlap `repeats` times on resident corners; *ms <- average
   milliseconds of one scan (HIP events).  Used by the benchmark. */
int vie_hip_time_box_overlap(const double *minimum, const double *maximum, size_t count, int repeats,
                             long long *left, long long *right, double *ms);
typedef struct {
  double worst_margin;    /* gap - required of the worst pair (DBL_MAX: none evaluated) */
  double worst_gap;       /* its gap (m) and required clearance (m), as the diagnostic prints them */
  double worst_required;
  double minimum_gap;     /* mesh->minimum_conductor_gap and mesh->required_conductor_gap */
  double required_gap;
  long long worst_box;    /* 1-based box of the worst pair (0: none) */
  long long worst_conductor; /* 1-based wire segment or patch of the worst pair */
  int worst_is_patch;
  int has_conductors;     /* mesh->has_conductors: cells and conductors exist */
  int insufficient;       /* worst_margin < -1e-12: the reference fails with CLEARANCE */
} vie_hip_clearance;
/* Dielectric-to-conductor clearance of the mesher's evaluate_clearance
   (dielectric_mesh.c:215-316): every box (corners as above; cell_longest =
   the longest edge of the box's first cell, 0 for a box without cells, which
   is skipped) against every wire segment (NEC's centre, length, direction
   cosines and radius bi, in metres) and patch (centre patch_points[3 j ..],
   area).  The worst pair and the smallest gap are the first in the
   reference's loop order (boxes outer, then wires, then patches) and carry
   the same values.  1 on success. */
int vie_hip_conductor_clearance(const double *minimum, const double *maximum, const double *cell_longest,
                                size_t boxes, int segments, const double *x, const double *y, const double *z,
                                const double *si, const double *cab, const double *sab, const double *salp,
                                const double *bi, int patches, const double *patch_points, const double *patch_area,
                                vie_hip_clearance *out);
/* vie_hip_conductor_clearance `repeats` times on resident inputs; *ms <-
   average milliseconds of one evaluation.  Used by the benchmark. */
int vie_hip_time_conductor_clearance(const double *minimum, const double *maximum, const double *cell_longest,
                                     size_t boxes, int segments, const double *x, const double *y, const double *z,
                                     const double *si, const double *cab, const double *sab, const double *salp,
                                     const double *bi, int patches, const double *patch_points,
                                     const double *patch_area, int repeats, vie_hip_clearance *out, double *ms);
/* Frees the device buffers that the far-field, conductor-row, conductor-column,
   coupled far-field and mesh-check entry points keep between calls (new; for
   long-running servers; the next call allocates again).  1 on success. */
int vie_hip_release_workspaces(void);
/* Host-side step of the recycled subspace, exported for tests: the k harmonic
   Ritz vectors of smallest modulus (Morgan 2002) of the (m+1) x m Arnoldi
   Hessenberg hbar (column-major), as the m x k column-major p with unit
   columns.  1 on success.  No device work. */
int vie_hip_harmonic_ritz(const VIE_CDOUBLE *hbar, int m, int k, VIE_CDOUBLE *p);
/* Mean basis column j of the matvec Arnoldi steps that the system's own
   resident solver issued since the last call with reset != 0, and their count. */
int dielectric_fft_solve_stats(dielectric_fft_system *system, double *mean_column, unsigned long long *steps,
                               int reset);
/* Times the two streaming kernels of one delayed-CGS2 Arnoldi step at basis
   column j (the reduction and the fused update), averaged over `repeats`
   launches, with HIP events on the system's stream; overwrites the resident
   solver workspace.  bytes_* receive the algorithmic bytes per launch:
   (j + 2) n S and (j + 4) n S for n unknowns of S bytes. */
int dielectric_fft_time_arnoldi(dielectric_fft_system *system, vie_hip_precision precision, int j, int repeats,
                                double *ms_reduce, double *ms_update, double *bytes_reduce, double *bytes_update);
/* Algorithmic HBM bytes of one fused apply at a precision (per pass and total),
   from the pruned five-pass model: see the implementation for the formula. */
int dielectric_fft_apply_bytes(const dielectric_fft_system *system, vie_hip_precision precision,
                               double *pass_bytes, double *total_bytes);

/* ---- batched solves (new: the reference solves one right-hand side at a
   time, each frequency in its own forked process, dielectric_wire.c:1296-1304)
   Many independent solves in flight on one GPU: each submitted job is a
   device-resident GMRES (the vie_gmres contract, x0 = 0) with its own pooled
   workspace; the host interleaves the Arnoldi steps of all jobs, jobs of
   one system run in submission order on its stream and jobs of different
   systems overlap on the device.  A job's result does not depend on what
   else is in flight.  One host thread drives a batch; systems must outlive
   the jobs submitted on them. */
typedef struct vie_hip_batch vie_hip_batch;
vie_hip_batch *vie_hip_batch_create(void);
/* Waits for queued device work, then frees every job and workspace. */
void vie_hip_batch_destroy(vie_hip_batch *batch);
/* Device bytes one job of this system and options holds while in flight
   (Krylov basis, solver vectors, reduction scratch): the admission unit. */
size_t vie_hip_batch_job_bytes(const dielectric_fft_system *system, const vie_gmres_options *options,
                               const vie_hip_solve_options *solve);
/* Pre-allocates `count` pooled workspaces large enough for jobs of this
   system and options, so later submissions allocate nothing. */
int vie_hip_batch_reserve(vie_hip_batch *batch, size_t count, dielectric_fft_system *system,
                          const vie_gmres_options *options);
/* Queues a solve of system x = rhs (user layout, copied before return).
   Returns the job id (> 0), or 0 with vie_hip_last_error set. */
long long vie_hip_batch_submit(vie_hip_batch *batch, dielectric_fft_system *system, const VIE_CDOUBLE *rhs,
                               const vie_gmres_options *options, const vie_hip_solve_options *solve);
/* Advances every job as far as it can without waiting; writes up to
   max_ready ids of jobs finished since they were last reported and returns
   their count; *running (optional) receives the jobs still in flight. */
size_t vie_hip_batch_progress(vie_hip_batch *batch, long long *ready, size_t max_ready, size_t *running);
size_t vie_hip_batch_running(const vie_hip_batch *batch);
/* Drives all jobs until job `id` (or every job) has finished. */
int vie_hip_batch_wait(vie_hip_batch *batch, long long id);
int vie_hip_batch_wait_all(vie_hip_batch *batch);
/* Result of a finished job: returns what vie_gmres would (1 on convergence);
   x (optional) receives the solution, report (optional) the report and
   seconds (optional) the time from submission to completion. */
int vie_hip_batch_result(vie_hip_batch *batch, long long id, VIE_CDOUBLE *x, vie_gmres_report *report,
                         double *seconds);
/* Returns a finished job's workspace to the pool. */
int vie_hip_batch_release(vie_hip_batch *batch, long long id);

/* Recycled subspace shared among the right-hand sides of one system (new;
   the reference solves one right-hand side per vie_gmres call,
   dielectric_wire.c:1296-1304, and restarts from nothing).  With on != 0,
   jobs submitted afterwards take part: the first of them on a system whose
   first cycle ends (converged or not) and that solves with a recycled
   subspace (vie_hip_solve_options.recycle, the MIXED default) publishes its
   U and C (A M^-1 U = C, C orthonormal: they depend on the operator only);
   a job of the same system, precision and recycled size that starts later,
   on an unchanged operator, adopts them, so its first cycle already
   minimises over span(U) + its own Krylov space (GCRO with a fixed subspace,
   Parks et al., SIAM J. Sci. Comput. 28, 2006: recycling across a sequence
   of right-hand sides).  Results differ from solo solves by the solver
   tolerance, not bitwise; acceptance stays the true residual.  The system
   holds the shared subspace (2 k basis vectors per precision) until it is
   freed.  Submit one right-hand side per system first (a seed) and the rest
   after it for the others to benefit. */
int vie_hip_batch_share_subspaces(vie_hip_batch *batch, int on);
/* Subspaces published and adopted by the batch's finished jobs. */
int vie_hip_batch_shared_stats(const vie_hip_batch *batch, unsigned long long *published,
                               unsigned long long *adopted);

/* Multi-job step counters of a batch: group_steps = multi-job Arnoldi steps
 * issued (each one launch per kernel for all members), job_steps = the job
 * steps they covered.  No counterpart in the reference (it solves one
 * right-hand side per vie_gmres call, dielectric_wire.c:1296-1304). */
int vie_hip_batch_group_stats(const vie_hip_batch *batch, unsigned long long *group_steps,
                              unsigned long long *job_steps);

#ifdef __cplusplus
}
#endif
#endif
