// Compact per-voxel dielectric resource: one record for a whole voxel lattice.
//
// The reference's resource format describes a model as BOX records
// (dielectric_io.c:472-512), and its mesher checks every pair of boxes for
// overlapping interiors (dielectric_mesh.c:81-123), which is O(B^2) when a
// model with a permittivity per voxel is written as one BOX per voxel
// (3.4e10 pairs at 64^3).  This record states the lattice once and reads the
// voxels from a binary side file, so the model is on one lattice by
// construction and no pair check exists:
//
//   NEC2C-DIELECTRIC 4
//   MATERIAL <id> ER <er> [EI <loss> | SIGMA <S/m>]     as resource version 2
//   MATERIAL <id> TABLE                                  as resource version 3,
//   SAMPLE <id> FREQ_MHZ <f> ER <er> EI|SIGMA <loss>     with its samples
//   VOXELS <nx> <ny> <nz> PITCH <m> ORIGIN <x0> <y0> <z0> MATERIALS <file>
//   VOXELS <nx> <ny> <nz> PITCH <m> ORIGIN <x0> <y0> <z0> PERMITTIVITY <file>
//   END
//
// MATERIALS: nx*ny*nz little-endian uint32 material ids (0: no dielectric);
// PERMITTIVITY: nx*ny*nz pairs of little-endian float64 (Re eps, Im eps), Im
// <= 0 under exp(+i omega t), (1, 0) meaning vacuum (no unknowns).  Voxel
// (i, j, k) is the cube of edge PITCH with its lowest corner at ORIGIN +
// PITCH (i, j, k), stored with x slowest and z fastest.  A relative side-file
// path is taken relative to the resource.  Keywords are case-insensitive, '#'
// starts a comment, paths may be double-quoted.  The cells of a prepared
// system are the dielectric voxels in storage order.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "vie_hip.h"

struct vie_hip_voxels {
  size_t shape[3] = {0, 0, 0};
  double pitch = 0, origin[3] = {0, 0, 0};
  bool permittivity = false;
  std::vector<uint32_t> ids;            // MATERIALS: id per voxel
  std::vector<double> eps;              // PERMITTIVITY: (re, im) per voxel
  std::vector<dielectric_material> materials;
  std::vector<dielectric_sample> samples;
  std::vector<size_t> active;           // voxel index of each cell
  std::string resource;
};

static int fail(dielectric_diagnostic* d, const std::string& resource, size_t line, const char* record,
                const std::string& value, const char* message) {
  if (d) {
    memset(d, 0, sizeof(*d));
    snprintf(d->resource, sizeof(d->resource), "%s", resource.c_str());
    snprintf(d->record, sizeof(d->record), "%s", record);
    snprintf(d->value, sizeof(d->value), "%s", value.c_str());
    snprintf(d->message, sizeof(d->message), "%s", message);
    d->line = line;
  }
  return 0;
}

static bool keyword(const std::string& a, const char* b) {
  if (a.size() != strlen(b)) return false;
  for (size_t i = 0; i < a.size(); i++)
    if (toupper((unsigned char)a[i]) != b[i]) return false;
  return true;
}

// Whitespace-separated tokens of one line; '#' outside quotes ends it; a
// double-quoted token may hold spaces.  false on an unterminated quote.
static bool tokenize(const std::string& line, std::vector<std::string>& out) {
  out.clear();
  size_t i = 0, n = line.size();
  while (i < n) {
    while (i < n && isspace((unsigned char)line[i])) i++;
    if (i >= n || line[i] == '#') break;
    std::string t;
    if (line[i] == '"') {
      size_t j = line.find('"', i + 1);
      if (j == std::string::npos) return false;
      t = line.substr(i + 1, j - i - 1);
      i = j + 1;
    } else {
      while (i < n && !isspace((unsigned char)line[i]) && line[i] != '#') t += line[i++];
    }
    out.push_back(t);
  }
  return true;
}

static bool number(const std::string& s, double* v) {
  if (s.empty()) return false;
  char* end = nullptr;
  errno = 0;
  double x = strtod(s.c_str(), &end);
  if (*end || !std::isfinite(x)) return false;
  *v = x;
  return true;
}

static bool positive_int(const std::string& s, size_t* v) {
  if (s.empty() || s.size() > 18) return false;
  size_t x = 0;
  for (char ch : s) {
    if (ch < '0' || ch > '9') return false;
    x = 10 * x + (size_t)(ch - '0');
  }
  if (x == 0) return false;
  *v = x;
  return true;
}

static bool read_file(const std::string& path, std::vector<unsigned char>& buf, size_t expect) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return false;
  buf.resize(expect + 1);
  size_t got = fread(buf.data(), 1, expect + 1, f);
  fclose(f);
  if (got != expect) return false;
  buf.resize(expect);
  return true;
}

static inline uint32_t le32(const unsigned char* p) {
  return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}
static inline double le64f(const unsigned char* p) {
  uint64_t u = 0;
  for (int i = 7; i >= 0; i--) u = u << 8 | p[i];
  double d;
  memcpy(&d, &u, 8);
  return d;
}

extern "C" int vie_hip_voxels_load(vie_hip_voxels** out, const char* path, dielectric_diagnostic* diag) {
  if (!out) return 0;
  *out = nullptr;
  std::string res = path ? path : "";
  FILE* f = path ? fopen(path, "r") : nullptr;
  if (!f) return fail(diag, res, 0, "DF", res, "cannot open the dielectric resource");
  vie_hip_voxels* v = new (std::nothrow) vie_hip_voxels();
  if (!v) {
    fclose(f);
    return fail(diag, res, 0, "DF", res, "allocation failed");
  }
  v->resource = res;
  std::vector<std::string> tok;
  std::string line, side;
  size_t lineno = 0, voxels_line = 0;
  bool header = false, ended = false, have_voxels = false;
  char buf[4096];
  auto bad = [&](const char* rec, const std::string& val, const char* msg) {
    fclose(f);
    delete v;
    return fail(diag, res, lineno, rec, val, msg);
  };
  while (fgets(buf, sizeof(buf), f)) {
    lineno++;
    line = buf;
    if (line.size() == sizeof(buf) - 1 && line.back() != '\n') return bad("LINE", "", "line too long");
    if (!tokenize(line, tok)) return bad("LINE", "", "unterminated quoted token");
    if (tok.empty()) continue;
    if (ended) return bad(tok[0].c_str(), tok[0], "only blank lines or comments may follow END");
    if (!header) {
      if (tok.size() != 2 || !keyword(tok[0], "NEC2C-DIELECTRIC") || tok[1] != "4")
        return bad("NEC2C-DIELECTRIC", tok[0], "expected the version header NEC2C-DIELECTRIC 4");
      header = true;
      continue;
    }
    if (keyword(tok[0], "MATERIAL")) {
      dielectric_material m;
      memset(&m, 0, sizeof(m));
      size_t id;
      if (tok.size() < 3 || !positive_int(tok[1], &id) || id > 0x7fffffff)
        return bad("MATERIAL", tok.size() > 1 ? tok[1] : tok[0], "material ID must be a positive decimal integer");
      m.id = (int)id;
      for (auto& q : v->materials)
        if (q.id == m.id) return bad("MATERIAL", tok[1], "duplicate material ID");
      m.source_line = lineno;
      if (tok.size() == 3 && keyword(tok[2], "TABLE")) {
        m.tabulated = 1;
        m.epsilon_r = 1;
      } else {
        if ((tok.size() != 4 && tok.size() != 6) || !keyword(tok[2], "ER"))
          return bad("MATERIAL", tok[0], "expected MATERIAL <id> ER <epsilon-r> [EI <loss> | SIGMA <S/m>] or TABLE");
        if (!number(tok[3], &m.epsilon_r) || m.epsilon_r < 1)
          return bad("MATERIAL", tok[3], "ER must be finite and greater than or equal to 1");
        if (tok.size() == 6) {
          double loss;
          if (!number(tok[5], &loss) || loss < 0 || (!keyword(tok[4], "EI") && !keyword(tok[4], "SIGMA")))
            return bad("MATERIAL", tok[4], "expected EI or SIGMA with a finite nonnegative value");
          if (keyword(tok[4], "EI")) m.epsilon_loss = loss;
          else m.conductivity = loss;
        }
      }
      v->materials.push_back(m);
    } else if (keyword(tok[0], "SAMPLE")) {
      dielectric_sample a;
      memset(&a, 0, sizeof(a));
      size_t id;
      if (tok.size() != 8 || !positive_int(tok[1], &id) || id > 0x7fffffff || !keyword(tok[2], "FREQ_MHZ") ||
          !number(tok[3], &a.frequency_mhz) || a.frequency_mhz <= 0 || !keyword(tok[4], "ER") ||
          !number(tok[5], &a.epsilon_r) || a.epsilon_r < 1 || (!keyword(tok[6], "EI") && !keyword(tok[6], "SIGMA")) ||
          !number(tok[7], &a.loss) || a.loss < 0 || (a.epsilon_r == 1 && a.loss == 0))
        return bad("SAMPLE", tok[0],
                   "expected SAMPLE <id> FREQ_MHZ <positive MHz> ER <ER>=1> EI|SIGMA <nonnegative loss>; sample "
                   "must differ from vacuum");
      a.material_id = (int)id;
      a.conductivity_kind = keyword(tok[6], "SIGMA");
      a.source_line = lineno;
      v->samples.push_back(a);
    } else if (keyword(tok[0], "VOXELS")) {
      if (have_voxels) return bad("VOXELS", tok[0], "only one VOXELS record is allowed");
      double o[3], h;
      if (tok.size() != 12 || !keyword(tok[4], "PITCH") || !keyword(tok[6], "ORIGIN") ||
          !(keyword(tok[10], "MATERIALS") || keyword(tok[10], "PERMITTIVITY")))
        return bad("VOXELS", tok[0],
                   "expected VOXELS <nx> <ny> <nz> PITCH <m> ORIGIN <x> <y> <z> MATERIALS|PERMITTIVITY <file>");
      for (int a = 0; a < 3; a++)
        if (!positive_int(tok[1 + a], &v->shape[a]) || v->shape[a] > (1u << 20))
          return bad("VOXELS", tok[1 + a], "voxel counts must be positive decimal integers");
      if (!number(tok[5], &h) || !(h > 0)) return bad("VOXELS", tok[5], "PITCH must be finite and positive");
      for (int a = 0; a < 3; a++)
        if (!number(tok[7 + a], &o[a])) return bad("VOXELS", tok[7 + a], "ORIGIN must be finite");
      v->pitch = h;
      memcpy(v->origin, o, sizeof(o));
      v->permittivity = keyword(tok[10], "PERMITTIVITY");
      side = tok[11];
      voxels_line = lineno;
      have_voxels = true;
    } else if (keyword(tok[0], "END")) {
      if (tok.size() != 1) return bad("END", tok[0], "END takes no fields");
      ended = true;
    } else {
      return bad(tok[0].c_str(), tok[0], "unknown record (voxel resources hold MATERIAL, SAMPLE, VOXELS, END)");
    }
  }
  fclose(f);
  auto bad_after = [&](size_t ln, const char* rec, const std::string& val, const char* msg) {
    delete v;
    return fail(diag, res, ln, rec, val, msg);
  };
  if (!header) return bad_after(lineno, "NEC2C-DIELECTRIC", "", "missing version header");
  if (!ended) return bad_after(lineno, "END", "", "missing END");
  if (!have_voxels) return bad_after(lineno, "VOXELS", "", "missing VOXELS record");
  // tables: samples in ascending frequency per tabulated material, at least two (dielectric_io.c:630-655)
  std::stable_sort(v->samples.begin(), v->samples.end(), [](const dielectric_sample& a, const dielectric_sample& b) {
    return a.material_id != b.material_id ? a.material_id < b.material_id : a.frequency_mhz < b.frequency_mhz;
  });
  for (auto& m : v->materials) {
    size_t first = v->samples.size(), count = 0;
    for (size_t i = 0; i < v->samples.size(); i++)
      if (v->samples[i].material_id == m.id) {
        if (first == v->samples.size()) first = i;
        count++;
      }
    if (m.tabulated) {
      if (count < 2) return bad_after(m.source_line, "SAMPLE", "table", "a TABLE material needs at least two samples");
      for (size_t i = first + 1; i < first + count; i++)
        if (!(v->samples[i].frequency_mhz > v->samples[i - 1].frequency_mhz))
          return bad_after(v->samples[i].source_line, "SAMPLE", "table", "sample frequencies must increase");
      m.sample_count = count;
      m.samples = nullptr;  // pointed into v->samples below, after the vector stops moving
    } else if (count) {
      return bad_after(v->samples[first].source_line, "SAMPLE", "table", "SAMPLE for a material that is not a TABLE");
    }
  }
  for (auto& sm : v->samples) {
    bool found = false;
    for (auto& m : v->materials) found = found || (m.id == sm.material_id && m.tabulated);
    if (!found) return bad_after(sm.source_line, "SAMPLE", "table", "SAMPLE for an undeclared material");
  }
  for (auto& m : v->materials)
    if (m.tabulated)
      for (size_t i = 0; i < v->samples.size(); i++)
        if (v->samples[i].material_id == m.id) {
          m.samples = &v->samples[i];
          break;
        }
  // side file
  const size_t nvox = v->shape[0] * v->shape[1] * v->shape[2];
  std::string sp = side;
  if (!sp.empty() && sp[0] != '/') {
    size_t slash = res.rfind('/');
    if (slash != std::string::npos) sp = res.substr(0, slash + 1) + sp;
  }
  std::vector<unsigned char> raw;
  const size_t each = v->permittivity ? 16 : 4;
  if (nvox > ((size_t)1 << 34) / each || !read_file(sp, raw, nvox * each))
    return bad_after(voxels_line, "VOXELS", side, "voxel file missing, unreadable or not nx*ny*nz entries long");
  if (v->permittivity) {
    v->eps.resize(2 * nvox);
    for (size_t i = 0; i < nvox; i++) {
      double re = le64f(&raw[16 * i]), im = le64f(&raw[16 * i + 8]);
      if (!std::isfinite(re) || !std::isfinite(im) || re < 1 || im > 0) {
        char val[64];
        snprintf(val, sizeof(val), "voxel %zu", i);
        return bad_after(voxels_line, "VOXELS", val, "voxel permittivity must be finite and passive (Re >= 1, Im <= 0)");
      }
      v->eps[2 * i] = re;
      v->eps[2 * i + 1] = im;
      if (!(re == 1 && im == 0)) v->active.push_back(i);
    }
  } else {
    v->ids.resize(nvox);
    for (size_t i = 0; i < nvox; i++) {
      uint32_t id = le32(&raw[4 * i]);
      v->ids[i] = id;
      if (!id) continue;
      bool found = false;
      for (auto& m : v->materials) found = found || (uint32_t)m.id == id;
      if (!found) {
        char val[64];
        snprintf(val, sizeof(val), "voxel %zu id %u", i, id);
        return bad_after(voxels_line, "VOXELS", val, "voxel references an undeclared material");
      }
      v->active.push_back(i);
    }
  }
  *out = v;
  return 1;
}

extern "C" void vie_hip_voxels_free(vie_hip_voxels* v) { delete v; }

extern "C" int vie_hip_voxels_info(const vie_hip_voxels* v, size_t shape[3], double* pitch, double origin[3],
                                   size_t* cells) {
  if (!v) return 0;
  for (int a = 0; a < 3; a++) {
    if (shape) shape[a] = v->shape[a];
    if (origin) origin[a] = v->origin[a];
  }
  if (pitch) *pitch = v->pitch;
  if (cells) *cells = v->active.size();
  return 1;
}

// Centre of voxel q (x slowest, z fastest).
static void voxel_center(const vie_hip_voxels* v, size_t q, double c[3]) {
  size_t k = q % v->shape[2], j = (q / v->shape[2]) % v->shape[1], i = q / (v->shape[2] * v->shape[1]);
  const size_t ijk[3] = {i, j, k};
  for (int a = 0; a < 3; a++) c[a] = v->origin[a] + v->pitch * ((double)ijk[a] + 0.5);
}

extern "C" int vie_hip_voxels_cells(const vie_hip_voxels* v, double* centers, double* sizes, int* material_ids) {
  if (!v) return 0;
  for (size_t c = 0; c < v->active.size(); c++) {
    const size_t q = v->active[c];
    if (centers) voxel_center(v, q, centers + 3 * c);
    if (sizes)
      for (int a = 0; a < 3; a++) sizes[3 * c + a] = v->pitch;
    if (material_ids) material_ids[c] = v->permittivity ? (int)(c + 1) : (int)v->ids[q];
  }
  return 1;
}

extern "C" int dielectric_fft_prepare_voxels(dielectric_fft_system** out, const vie_hip_voxels* v, size_t conductors,
                                             double k0, size_t extra_bytes, int pec, dielectric_diagnostic* diag) {
  if (!out) return 0;
  *out = nullptr;
  if (!v) return fail(diag, "", 0, "VOXELS", "", "FFT: no voxel model");
  const size_t cells = v->active.size();
  if (v->permittivity && cells > 0x7ffffffe) return fail(diag, v->resource, 0, "VOXELS", "", "FFT: too many voxels");
  std::vector<dielectric_cell> cl;
  std::vector<dielectric_material> mats;
  try {
    cl.resize(cells);
    if (v->permittivity) mats.resize(cells);
  } catch (...) {
    return fail(diag, v->resource, 0, "VOXELS", "", "FFT: cell list allocation failed");
  }
  for (size_t c = 0; c < cells; c++) {
    const size_t q = v->active[c];
    memset(&cl[c], 0, sizeof(cl[c]));
    voxel_center(v, q, cl[c].center);
    for (int a = 0; a < 3; a++) cl[c].size[a] = v->pitch;
    cl[c].box_id = q;
    if (v->permittivity) {
      // one material per voxel: eps = er - i loss exactly (dielectric.h:36-56 with no conductivity)
      memset(&mats[c], 0, sizeof(mats[c]));
      mats[c].id = (int)(c + 1);
      mats[c].epsilon_r = v->eps[2 * q];
      mats[c].epsilon_loss = -v->eps[2 * q + 1];
      cl[c].material_id = (int)(c + 1);
    } else {
      cl[c].material_id = (int)v->ids[q];
    }
  }
  dielectric_model model;
  dielectric_mesh_result mesh;
  memset(&model, 0, sizeof(model));
  memset(&mesh, 0, sizeof(mesh));
  const std::vector<dielectric_material>& use = v->permittivity ? mats : v->materials;
  model.materials = const_cast<dielectric_material*>(use.data());
  model.material_count = use.size();
  model.loaded = 1;
  model.format_version = 4;
  mesh.cells = cl.data();
  mesh.cell_count = cells;
  mesh.built = 1;
  return (pec ? dielectric_fft_prepare_pec : dielectric_fft_prepare)(out, &model, &mesh, conductors, k0, extra_bytes,
                                                                      diag);
}
