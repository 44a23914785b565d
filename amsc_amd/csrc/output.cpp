// output.cpp -- the reference solver's output files (optim_par.cpp:269,
// 311-337, 655-665) written from a device-resident simulation through the
// public C ABI: particles as CSV, nodal fields as a VTK XML structured grid.
// Host code only; every value is read back with amsc_sim_get_field.
#include "amsc_mpm.h"

#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

namespace {

// Particle fields of the C ABI (amsc_sim_get_field), in alphabetical order
// after x and y.  Output-only ones are skipped when the last step did not
// produce them.
const char *const kParticleFields[] = {
    "Ap",     "F_11",   "F_12",   "F_21",    "F_22",    "Fb_x",   "Fb_y",
    "Fpx",    "Fpy",    "Fric_px", "Fric_py", "H",      "Mp",     "Vp",
    "Zp",     "apx",    "apy",    "dZxp",    "dZyp",    "hp",     "hpZ",
    "mom_px", "mom_py", "vpx",    "vpx_dx",  "vpx_dy",  "vpy",    "vpy_dx",
    "vpy_dy"};

const char *const kNodalFields[] = {
    "Mv",       "mom_vx",   "mom_vy",   "F_ext_vx", "F_ext_vy", "F_int_vx",
    "F_int_vy", "Fric_x",   "Fric_y",   "FPxv",     "FPyv",     "Ftot_vx",
    "Ftot_vy",  "avx",      "avy",      "vvx",      "vvy",      "vvxL",
    "vvyL"};

struct File {
  FILE *f = nullptr;
  explicit File(const char *path) : f(std::fopen(path, "w")) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

}  // namespace

extern "C" int amsc_sim_write_particles_csv(amsc_sim *s, const char *path) {
  if (!s || !path) return AMSC_ERR_INVALID;
  int np = 0;
  if (int rc = amsc_sim_num_particles(s, &np)) return rc;
  const size_t m = np > 0 ? np : 1;  // the API takes no null buffers
  std::vector<std::string> names = {"x", "y"};
  std::vector<std::vector<double>> cols;
  for (const char *n : {"x", "y"}) {
    cols.emplace_back(m);
    if (int rc = amsc_sim_get_field(s, n, cols.back().data(), np)) return rc;
  }
  for (const char *n : kParticleFields) {
    std::vector<double> v(m);
    const int rc = amsc_sim_get_field(s, n, v.data(), np);
    if (rc == AMSC_ERR_STALE) continue;  // not produced by the last step
    if (rc) return rc;
    names.push_back(n);
    cols.push_back(std::move(v));
  }
  std::vector<int> label(m), level(m);
  if (int rc = amsc_sim_get_field_int(s, "label", label.data(), np)) return rc;
  if (int rc = amsc_sim_get_field_int(s, "level", level.data(), np)) return rc;
  File out(path);
  if (!out.f) return AMSC_ERR_INVALID;
  for (size_t k = 0; k < names.size(); ++k)
    std::fprintf(out.f, "\"%s\",", names[k].c_str());
  std::fprintf(out.f, "\"label\",\"level\"\n");
  for (int i = 0; i < np; ++i) {
    for (size_t k = 0; k < cols.size(); ++k)
      std::fprintf(out.f, "%.17g,", cols[k][i]);
    std::fprintf(out.f, "%d,%d\n", label[i], level[i]);
  }
  return std::ferror(out.f) ? AMSC_ERR_INVALID : AMSC_OK;
}

extern "C" int amsc_sim_write_grid_vts(amsc_sim *s, const char *path) {
  if (!s || !path) return AMSC_ERR_INVALID;
  amsc_params p;
  if (int rc = amsc_sim_params(s, &p)) return rc;
  const int nr = p.nrows, nc = p.ncols, nn = (nr + 1) * (nc + 1);
  std::vector<std::string> names = {"Z"};
  std::vector<std::vector<double>> cols(1, std::vector<double>(nn));
  if (int rc = amsc_sim_get_field(s, "Z", cols[0].data(), nn)) return rc;
  for (const char *n : kNodalFields) {
    std::vector<double> v(nn);
    const int rc = amsc_sim_get_field(s, n, v.data(), nn);
    if (rc == AMSC_ERR_STALE) continue;
    if (rc) return rc;
    names.push_back(n);
    cols.push_back(std::move(v));
  }
  File out(path);
  if (!out.f) return AMSC_ERR_INVALID;
  // VTK orders structured points with x (columns) fastest; node ids are
  // r + c*(nrows+1)
  std::fprintf(out.f,
               "<?xml version=\"1.0\"?>\n"
               "<VTKFile type=\"StructuredGrid\" version=\"0.1\" "
               "byte_order=\"LittleEndian\">\n"
               "  <StructuredGrid WholeExtent=\"0 %d 0 %d 0 0\">\n"
               "    <Piece Extent=\"0 %d 0 %d 0 0\">\n"
               "      <PointData>\n",
               nc, nr, nc, nr);
  for (size_t k = 0; k < cols.size(); ++k) {
    std::fprintf(out.f,
                 "        <DataArray type=\"Float64\" Name=\"%s\" "
                 "format=\"ascii\">\n",
                 names[k].c_str());
    for (int r = 0; r <= nr; ++r)
      for (int c = 0; c <= nc; ++c)
        std::fprintf(out.f, "%.17g\n", cols[k][r + c * (nr + 1)]);
    std::fprintf(out.f, "        </DataArray>\n");
  }
  std::fprintf(out.f,
               "      </PointData>\n"
               "      <Points>\n"
               "        <DataArray type=\"Float64\" NumberOfComponents=\"3\" "
               "format=\"ascii\">\n");
  for (int r = 0; r <= nr; ++r)
    for (int c = 0; c <= nc; ++c)
      std::fprintf(out.f, "%.17g %.17g 0\n", c * p.hx, r * p.hy);
  std::fprintf(out.f,
               "        </DataArray>\n"
               "      </Points>\n"
               "    </Piece>\n"
               "  </StructuredGrid>\n"
               "</VTKFile>\n");
  return std::ferror(out.f) ? AMSC_ERR_INVALID : AMSC_OK;
}
