// mfc_output.cpp — legacy VTK files of a handle's particles and tool (vtk_writer.cpp:20-263),
// host code over the library's own C ABI (mfc_get_info / mfc_get_field / mfc_get_tool).
//
// The format is the reference's: ASCII legacy VTK, an unstructured grid of vertex cells with
// density, temperature, fixed and tool markers, accumulated plastic strain, velocity and von
// Mises stress; the tool as triangles of its segments and of its fillet (20 samples). Two
// defects of the reference are not reproduced: it never copies the blanked flags to the host
// before testing them (vtk_writer.cpp:73-78), and its vertex cells index the particle's slot in
// the full array, which points past the written points once particles are skipped. Here the
// blanked flags are the state's, and cell k is written point k.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/mfc.h"

namespace {

struct file_closer {
  FILE *fp;
  ~file_closer() {
    if (fp) std::fclose(fp);
  }
};

int get(mfc_sim *sim, int field, std::vector<double> &v, int N) {
  v.resize((size_t)N * mfc_field_components(field));
  return mfc_get_field(sim, field, v.data(), sizeof(double) * v.size());
}

}  // namespace

extern "C" {

int mfc_write_vtk(mfc_sim *sim, const char *path) {
  if (!sim || !path) return MFC_E_INVALID;
  mfc_info info;
  int rc = mfc_get_info(sim, &info);
  if (rc) return rc;
  const int N = info.n_particles;
  std::vector<double> pos, vel, rho, T, eps, S, p;
  std::vector<int32_t> flags((size_t)N);
  if ((rc = get(sim, MFC_FIELD_POS, pos, N)) || (rc = get(sim, MFC_FIELD_VEL, vel, N)) ||
      (rc = get(sim, MFC_FIELD_RHO, rho, N)) || (rc = get(sim, MFC_FIELD_T, T, N)) ||
      (rc = get(sim, MFC_FIELD_EPS_PL, eps, N)) || (rc = get(sim, MFC_FIELD_S, S, N)) ||
      (rc = get(sim, MFC_FIELD_P, p, N)) ||
      (rc = mfc_get_field(sim, MFC_FIELD_FLAGS, flags.data(), sizeof(int32_t) * flags.size())))
    return rc;
  std::vector<int> keep;
  keep.reserve(N);
  for (int i = 0; i < N; i++)
    if (!(flags[i] & MFC_FLAG_BLANKED)) keep.push_back(i);
  const int n = (int)keep.size();
  file_closer f{std::fopen(path, "w+")};
  if (!f.fp) return MFC_E_INVALID;
  FILE *fp = f.fp;
  std::fprintf(fp, "# vtk DataFile Version 2.0\nmfree iwf\nASCII\n\n");
  std::fprintf(fp, "DATASET UNSTRUCTURED_GRID\n");
  std::fprintf(fp, "POINTS %d float\n", n);
  for (int i : keep) std::fprintf(fp, "%f %f %f\n", pos[2 * i], pos[2 * i + 1], 0.);
  std::fprintf(fp, "\n");
  std::fprintf(fp, "CELLS %d %d\n", n, 2 * n);
  for (int k = 0; k < n; k++) std::fprintf(fp, "%d %d\n", 1, k);
  std::fprintf(fp, "\n");
  std::fprintf(fp, "CELL_TYPES %d\n", n);
  for (int k = 0; k < n; k++) std::fprintf(fp, "%d\n", 1);
  std::fprintf(fp, "\n");
  std::fprintf(fp, "POINT_DATA %d\n", n);
  auto scalars = [&](const char *name, auto value) {
    std::fprintf(fp, "SCALARS %s float 1\nLOOKUP_TABLE default\n", name);
    for (int i : keep) std::fprintf(fp, "%f\n", value(i));
    std::fprintf(fp, "\n");
  };
  scalars("density", [&](int i) { return rho[i]; });
  scalars("Temperature", [&](int i) { return T[i]; });
  scalars("Fixed", [&](int i) { return (flags[i] & MFC_FLAG_FIXED) ? 1. : 0.; });
  scalars("Tool", [&](int i) { return (flags[i] & MFC_FLAG_TOOL) ? 1. : 0.; });
  scalars("EquivAccumPlasticStrain", [&](int i) { return eps[i]; });
  std::fprintf(fp, "VECTORS Velocity float\n");
  for (int i : keep) std::fprintf(fp, "%f %f %f\n", vel[2 * i], vel[2 * i + 1], 0.);
  std::fprintf(fp, "\n");
  // von Mises stress of the Cauchy stress S - p I (vtk_writer.cpp:160-172)
  scalars("SvM", [&](int i) {
    const double cxx = S[4 * i] - p[i], cyy = S[4 * i + 2] - p[i], czz = S[4 * i + 3] - p[i], cxy = S[4 * i + 1];
    const double svm2 = (cxx * cxx + cyy * cyy + czz * czz) - cxx * cyy - cxx * czz - cyy * czz + 3.0 * cxy * cxy;
    return svm2 > 0 ? std::sqrt(svm2) : 0.;
  });
  return std::ferror(fp) ? MFC_E_INVALID : MFC_OK;
}

int mfc_write_tool_vtk(mfc_sim *sim, const char *path) {
  if (!sim || !path) return MFC_E_INVALID;
  mfc_tool_state t;
  int rc = mfc_get_tool(sim, &t);
  if (rc) return rc;
  if (t.num_seg != 4 && t.num_seg != 5) return MFC_E_INVALID;  // the reference's assert
  struct tri {
    double p[3][2];
  };
  std::vector<tri> tris;
  auto add = [&](const double *a, const double *b, const double *c) {
    tri x = {{{a[0], a[1]}, {b[0], b[1]}, {c[0], c[1]}}};
    tris.push_back(x);
  };
  // the tool body (vtk_writer.cpp:192-200)
  if (t.num_seg == 4) {
    add(t.left[0], t.right[0], t.right[1]);
    add(t.left[2], t.right[2], t.right[3]);
  } else {
    add(t.left[0], t.right[0], t.right[2]);
    add(t.left[1], t.right[1], t.right[2]);
    add(t.left[3], t.right[3], t.right[4]);
  }
  // the fillet as a fan of 19 triangles (vtk_writer.cpp:203-225)
  if (t.has_fillet) {
    const int num_discr = 20;
    const double t1 = std::fmin(t.fillet_t1, t.fillet_t2), t2 = std::fmax(t.fillet_t1, t.fillet_t2);
    const double lo = t1 - 0.1 * t1;
    const double d_angle = (t2 - t1) / (num_discr - 1);
    const double r = t.fillet_r;
    for (int i = 0; i < num_discr - 1; i++) {
      const double a1 = lo + (i + 0) * d_angle, a2 = lo + (i + 1) * d_angle;
      const double p1[2] = {t.fillet_p[0], t.fillet_p[1]};
      const double p2[2] = {p1[0] + r * std::sin(a1), p1[1] + r * std::cos(a1)};
      const double p3[2] = {p1[0] + r * std::sin(a2), p1[1] + r * std::cos(a2)};
      add(p1, p2, p3);
    }
  }
  const int nt = (int)tris.size();
  file_closer f{std::fopen(path, "w+")};
  if (!f.fp) return MFC_E_INVALID;
  FILE *fp = f.fp;
  std::fprintf(fp, "# vtk DataFile Version 2.0\nmfree iwf\nASCII\n\n");
  std::fprintf(fp, "DATASET UNSTRUCTURED_GRID\n");
  std::fprintf(fp, "POINTS %d float\n", 3 * nt);
  for (const tri &x : tris)
    for (int v = 0; v < 3; v++) std::fprintf(fp, "%f %f %f\n", x.p[v][0], x.p[v][1], 0.);
  std::fprintf(fp, "\n");
  std::fprintf(fp, "CELLS %d %d\n", nt, 3 * nt + nt);
  for (int i = 0; i < nt; i++) std::fprintf(fp, "3 %d %d %d\n", 3 * i + 0, 3 * i + 1, 3 * i + 2);
  std::fprintf(fp, "\n");
  std::fprintf(fp, "CELL_TYPES %d\n", nt);
  for (int i = 0; i < nt; i++) std::fprintf(fp, "5\n");
  return std::ferror(fp) ? MFC_E_INVALID : MFC_OK;
}

}  // extern "C"
