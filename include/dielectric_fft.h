/*
 * Hybrid conductor/volume FFT system, C-ABI of libvie_hip.so.
 *
 * Replaces the reference's dielectric_fft.h:1-26 symbol for symbol.  The
 * operator is the block matrix [C B; D V] (conductors first), V the FFT
 * volume operator of vie_fft.h; with the PEC variant V acts on the physical
 * cells of an image-embedded lattice (image currents diag(-1,-1,+1) J).
 *
 * Inside nec2c, include the reference's dielectric.h before this header; the
 * types below are then taken from there.  Standalone users get these
 * layout-identical definitions (dielectric.h:16-179 of the reference), since
 * the library reads model->materials, mesh->cells and writes
 * diagnostic->message through exactly these layouts.
 */
#ifndef DIELECTRIC_FFT_H
#define DIELECTRIC_FFT_H
#include "vie_fft.h"
#ifdef __cplusplus
extern "C" {
#endif

#ifndef DIELECTRIC_H
#define DIELECTRIC_H 1
#define VIE_HIP_DIELECTRIC_TYPES 1
#ifndef NEC2C_PATH_MAX
#define NEC2C_PATH_MAX 255
#endif
typedef struct {
  int material_id, conductivity_kind;
  double frequency_mhz, epsilon_r, loss;
  size_t source_line;
} dielectric_sample;
typedef struct {
  int id;
  double epsilon_r;
  size_t source_line;
  double epsilon_loss; /* positive imaginary-part magnitude, exp(+i omega t) */
  double conductivity; /* S/m */
  int tabulated;
  const dielectric_sample *samples;
  size_t sample_count;
} dielectric_material;
typedef struct { int id; int material_id; size_t source_line; } dielectric_region;
typedef struct { int region_id; double minimum[3]; double maximum[3]; size_t source_line; } dielectric_box;
typedef struct {
  int loaded, format_version, path_is_absolute, mesh_record_present;
  char *resource_path;
  dielectric_material *materials; size_t material_count;
  dielectric_sample *samples; size_t sample_count;
  dielectric_region *regions; size_t region_count;
  dielectric_box *boxes; size_t box_count;
  double max_kh; size_t max_cells; double max_aspect;
} dielectric_model;
typedef struct {
  char resource[NEC2C_PATH_MAX * 4 + 1];
  char record[32];
  char value[128];
  char message[256];
  size_t line;
} dielectric_diagnostic;
typedef struct {
  size_t box_id; int region_id; int material_id;
  double center[3]; double size[3]; double q; double aspect;
} dielectric_cell;
typedef struct {
  int region_id, material_id; size_t divisions[3]; size_t cell_offset, cell_count;
  double physical_volume, numerical_volume;
} dielectric_box_mesh;
typedef struct { int region_id, material_id; double diameter, electrical_diameter; } dielectric_region_mesh;
typedef struct {
  int built, accuracy_warning; double frequency_mhz;
  dielectric_region_mesh *regions; size_t region_count;
  dielectric_box_mesh *boxes; size_t box_count;
  dielectric_cell *cells; size_t cell_count;
  size_t vacuum_box_count, feature_warning_count;
  double declared_volume, active_volume, vacuum_volume, numerical_volume, volume_relative_error;
  double q_min, q_median, q_p95, q_max, aspect_max;
  int has_conductors; double minimum_conductor_gap, required_conductor_gap;
  size_t conductor_unknowns, volume_unknowns, total_unknowns, matrix_bytes, workspace_bytes;
} dielectric_mesh_result;
#endif

/* Compile-time cap of the reference's CPU backend (dielectric_fft.h:5).  Kept
   for source compatibility; the HIP backend's cap is dielectric_fft_budget(). */
#define DIELECTRIC_FFT_BUDGET ((size_t)512*1024*1024)

typedef struct dielectric_fft_system dielectric_fft_system;
/* {30, 300, 1e-12, 0} (dielectric_fft.c:8) */
extern vie_gmres_options dielectric_gmres_settings;

/* dielectric_fft.c:89-180.  extra_bytes reserves caller-owned storage. */
int dielectric_fft_prepare(dielectric_fft_system **out, const dielectric_model *model,
                           const dielectric_mesh_result *mesh, size_t conductors, double k0,
                           size_t extra_bytes, dielectric_diagnostic *diagnostic);
/* dielectric_fft.c:182-185: PEC image embedding; the bottom gap must be an
   integer multiple of half the pitch. */
int dielectric_fft_prepare_pec(dielectric_fft_system **out, const dielectric_model *model,
                               const dielectric_mesh_result *mesh, size_t conductors, double k0,
                               size_t extra_bytes, dielectric_diagnostic *diagnostic);
void dielectric_fft_free(dielectric_fft_system *system);                    /* dielectric_fft.c:27-28 */
/* dielectric_fft.c:34-40: C/B/D entries only (NULL for the volume block).
   Writes land in host storage and reach the device at the next apply/solve
   after dielectric_fft_get_operator (or dielectric_fft_blocks_changed). */
VIE_CDOUBLE *dielectric_fft_entry(dielectric_fft_system *system, size_t row, size_t column);
void dielectric_fft_get_operator(dielectric_fft_system *system, vie_linear_operator *op); /* :84-88 */
size_t dielectric_fft_storage(const dielectric_fft_system *system);        /* :29 */
size_t dielectric_fft_threads(const dielectric_fft_system *system);        /* :30-31 */
const char *dielectric_fft_backend(const dielectric_fft_system *system);   /* :32-33 */

#ifdef __cplusplus
}
#endif
#endif
