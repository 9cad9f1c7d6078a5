/* Pointer-only ctypes entry point around the reference's own mesher
 * (dielectric_mesh.c, built from its sources with shared.c) -- TEST
 * INFRASTRUCTURE ONLY.  Builds a model of one material (epsilon_r 2) and one
 * region per box (region id = box index + 1, source line = box index + 1),
 * runs dielectric_mesh_build and reports its diagnostic or the clearance
 * figures and the longest edge of each box's first cell.
 */
#include "dielectric.h"

#include <stdlib.h>
#include <string.h>

int ref_mesh_check(size_t nbox, const double *mn, const double *mx, double frequency_mhz, double max_kh,
                   size_t max_cells, int n, double *x, double *y, double *z, double *si, double *cab, double *sab,
                   double *salp, double *bi, int m, double *px, double *py, double *pz, double *pbi, char *record,
                   char *value, size_t *line, double *gaps, int *has_conductors, double *cell_longest) {
  dielectric_material material;
  dielectric_region *regions = calloc(nbox ? nbox : 1, sizeof(*regions));
  dielectric_box *boxes = calloc(nbox ? nbox : 1, sizeof(*boxes));
  dielectric_model model;
  dielectric_mesh_result mesh;
  dielectric_diagnostic diag;
  data_t geometry;
  int ok;
  memset(&material, 0, sizeof(material));
  memset(&model, 0, sizeof(model));
  memset(&diag, 0, sizeof(diag));
  memset(&geometry, 0, sizeof(geometry));
  material.id = 1;
  material.epsilon_r = 2.0;
  for (size_t b = 0; b < nbox; b++) {
    regions[b].id = (int)b + 1;
    regions[b].material_id = 1;
    regions[b].source_line = b + 1;
    boxes[b].region_id = (int)b + 1;
    boxes[b].source_line = b + 1;
    for (int a = 0; a < 3; a++) {
      boxes[b].minimum[a] = mn[3 * b + a];
      boxes[b].maximum[a] = mx[3 * b + a];
    }
  }
  model.loaded = 1;
  model.format_version = 1;
  model.materials = &material;
  model.material_count = 1;
  model.regions = regions;
  model.region_count = nbox;
  model.boxes = boxes;
  model.box_count = nbox;
  model.max_kh = max_kh;
  model.max_cells = max_cells;
  model.max_aspect = 2.0;
  geometry.n = n;
  geometry.m = m;
  geometry.x = x;
  geometry.y = y;
  geometry.z = z;
  geometry.si = si;
  geometry.cab = cab;
  geometry.sab = sab;
  geometry.salp = salp;
  geometry.bi = bi;
  geometry.px = px;
  geometry.py = py;
  geometry.pz = pz;
  geometry.pbi = pbi;
  dielectric_mesh_init(&mesh);
  ok = dielectric_mesh_build(&model, frequency_mhz, &geometry, &mesh, &diag);
  memcpy(record, diag.record, sizeof(diag.record));
  memcpy(value, diag.value, sizeof(diag.value));
  *line = diag.line;
  if (ok) {
    gaps[0] = mesh.minimum_conductor_gap;
    gaps[1] = mesh.required_conductor_gap;
    *has_conductors = mesh.has_conductors;
    for (size_t b = 0; b < nbox; b++) {
      const dielectric_box_mesh *bm = &mesh.boxes[b];
      const dielectric_cell *c = &mesh.cells[bm->cell_offset];
      cell_longest[b] = bm->cell_count ? fmax(c->size[0], fmax(c->size[1], c->size[2])) : 0.0;
    }
    dielectric_mesh_free(&mesh);
  }
  free(regions);
  free(boxes);
  return ok;
}

/* The cells the reference's mesher emits for one box per region (as above,
   no conductors): centre and size (3 each) and the box index of each cell, in
   the mesher's own order.  *count <- the cell count (at most max_out are
   written).  Returns 0 with the diagnostic record/value when meshing fails. */
int ref_mesh_cells(size_t nbox, const double *mn, const double *mx, double frequency_mhz, double max_kh,
                   size_t max_cells, double *centers, double *sizes, size_t *box_index, size_t max_out,
                   size_t *count, char *record, char *value) {
  dielectric_material material;
  dielectric_region *regions = calloc(nbox ? nbox : 1, sizeof(*regions));
  dielectric_box *boxes = calloc(nbox ? nbox : 1, sizeof(*boxes));
  dielectric_model model;
  dielectric_mesh_result mesh;
  dielectric_diagnostic diag;
  data_t geometry;
  int ok;
  memset(&material, 0, sizeof(material));
  memset(&model, 0, sizeof(model));
  memset(&diag, 0, sizeof(diag));
  memset(&geometry, 0, sizeof(geometry));
  material.id = 1;
  material.epsilon_r = 2.0;
  for (size_t b = 0; b < nbox; b++) {
    regions[b].id = (int)b + 1;
    regions[b].material_id = 1;
    regions[b].source_line = b + 1;
    boxes[b].region_id = (int)b + 1;
    boxes[b].source_line = b + 1;
    for (int a = 0; a < 3; a++) {
      boxes[b].minimum[a] = mn[3 * b + a];
      boxes[b].maximum[a] = mx[3 * b + a];
    }
  }
  model.loaded = 1;
  model.format_version = 1;
  model.materials = &material;
  model.material_count = 1;
  model.regions = regions;
  model.region_count = nbox;
  model.boxes = boxes;
  model.box_count = nbox;
  model.max_kh = max_kh;
  model.max_cells = max_cells;
  model.max_aspect = 2.0;
  dielectric_mesh_init(&mesh);
  ok = dielectric_mesh_build(&model, frequency_mhz, &geometry, &mesh, &diag);
  memcpy(record, diag.record, sizeof(diag.record));
  memcpy(value, diag.value, sizeof(diag.value));
  *count = 0;
  if (ok) {
    *count = mesh.cell_count;
    for (size_t c = 0; c < mesh.cell_count && c < max_out; c++) {
      memcpy(centers + 3 * c, mesh.cells[c].center, 3 * sizeof(double));
      memcpy(sizes + 3 * c, mesh.cells[c].size, 3 * sizeof(double));
      box_index[c] = mesh.cells[c].box_id;
    }
    dielectric_mesh_free(&mesh);
  }
  free(regions);
  free(boxes);
  return ok;
}
