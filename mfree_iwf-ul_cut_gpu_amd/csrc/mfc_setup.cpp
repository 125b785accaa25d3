// mfc_setup.cpp — scenario generation for the orthogonal cut: workpiece lattice, boundary
// flags, rigid tool with optional edge fillet placed at the depth of cut, tool particles
// for thermal contact, cell grid and time step.
//
// Follows setup_ref_cut (benchmarks.cpp:255-537), the tool constructors (tool.cpp:36-102,
// 342-375), tool sampling (tool.cpp:104-144), tool bounding box (tool.cpp:156-197) and the
// line/segment primitives (geometric_primitives.cpp:20-122). Host code: transcendental
// calls (cos, sin, atan2, sqrt) use the host C library as the reference's host code does.
// Compiled with -ffp-contract=off so every expression rounds as written.
#include "mfc_setup.h"

#include <cfloat>
#include <cmath>
#include <cstring>

namespace mfc {
namespace {

struct pt { double x, y; };

// line through two points: y = a x + b, or vertical x = b (geometric_primitives.cpp:70-85)
struct line2 { double a, b; bool vertical; };

line2 line_through(pt p1, pt p2) {
  line2 l;
  double det = p1.x * 1. - 1. * p2.x;
  if (std::fabs(det) < 1e-12) { l.vertical = true; l.a = DBL_MAX; l.b = p1.x; return l; }
  l.vertical = false;
  l.a = (p1.y * 1. - 1. * p2.y) / det;
  l.b = (p2.y * p1.x - p2.x * p1.y) / det;
  return l;
}

// geometric_primitives.cpp:20-47
pt line_cross(const line2 &m, const line2 &l) {
  if (m.vertical && l.vertical) return {DBL_MAX, DBL_MAX};
  if (std::fabs(m.a - l.a) < 1e-12) return {DBL_MAX, DBL_MAX};
  if (m.vertical) return {m.b, l.a * m.b + l.b};
  if (l.vertical) return {l.b, m.a * l.b + m.b};
  double x = (l.b - m.b) / (m.a - l.a);
  return {x, m.a * x + m.b};
}

// foot of the perpendicular from q (geometric_primitives.cpp:49-64)
pt line_foot(const line2 &l, pt q) {
  if (l.vertical) return {l.b, q.y};
  double den = l.a * l.a + (-1.) * (-1.);
  double fx = ((-1.) * ((-1.) * q.x - l.a * q.y) - l.a * l.b) / den;
  double fy = (l.a * (-(-1.) * q.x + l.a * q.y) - (-1.) * l.b) / den;
  return {fx, fy};
}

pt unit(pt v) {  // v * (1/sqrt(v.v))
  double is = 1.0 / std::sqrt(v.x * v.x + v.y * v.y);
  return {v.x * is, v.y * is};
}

double angle_pos(double y, double x) {  // tool.cpp:27-34
  double t = std::atan2(y, x);
  return (t > 0.) ? t : t + 2 * M_PI;
}

void set_segment(mfc_tool_state &t, int i, pt a, pt b) {  // geometric_primitives.cpp:112-122
  t.left[i][0] = a.x; t.left[i][1] = a.y;
  t.right[i][0] = b.x; t.right[i][1] = b.y;
  pt nrm = unit({b.y - a.y, -(b.x - a.x)});
  t.n[i][0] = nrm.x; t.n[i][1] = nrm.y;
  line2 l = line_through(a, b);
  t.line_a[i] = l.a; t.line_b[i] = l.b; t.vertical[i] = l.vertical ? 1 : 0;
}

void set_polygon(mfc_tool_state &t, const pt *p, int n) {  // tool.cpp:64-71 (closed polygon)
  t.num_seg = n;
  for (int i = 0; i < n; i++) set_segment(t, i, p[i], p[(i + 1 > n - 1) ? 0 : i + 1]);
}

// rigid tool from its top-left corner, length, height, rake and clearance angles and edge
// radius (tool.cpp:342-375 with construct_points_and_fillet :73-102, fit_fillet :36-62)
double build_tool(mfc_tool_state &t, pt tl, double length, double height, double rake, double clear, double r) {
  std::memset(&t, 0, sizeof(t));
  pt tr = {tl.x + length, tl.y};
  pt bl = {tl.x, tl.y - height};
  double ar = rake * M_PI / 180.;
  double af = (180 - 90 - clear) * M_PI / 180.;
  // (0,-1) * mat2(cos a, -sin a, sin a, cos a)
  pt rr = {0. * std::cos(ar) + (-1.) * (-std::sin(ar)), 0. * std::sin(ar) + (-1.) * std::cos(ar)};
  pt rf = {0. * std::cos(af) + (-1.) * (-std::sin(af)), 0. * std::sin(af) + (-1.) * std::cos(af)};
  line2 l1 = line_through(tr, {tr.x + rr.x, tr.y + rr.y});
  line2 l2 = line_through(bl, {bl.x + rf.x, bl.y + rf.y});
  pt br = line_cross(l1, l2);
  if (r == 0.) {
    pt poly[4] = {tl, tr, br, bl};
    set_polygon(t, poly, 4);
    return br.x;
  }
  // bisector of the wedge at br
  pt nt = unit({tr.x - br.x, tr.y - br.y});
  pt nl = unit({bl.x - br.x, bl.y - br.y});
  pt nm = unit({0.5 * (nt.x + nl.x), 0.5 * (nt.y + nl.y)});
  line2 lm = line_through(br, {br.x + nm.x, br.y + nm.y});
  line2 e1 = line_through(tr, br), e2 = line_through(bl, br);
  pt c;  // fillet centre: the point of lm at distance r from e1
  if (e1.vertical) {
    line2 par = {DBL_MAX, e1.b - r, true};
    c = line_cross(par, lm);
  } else {
    double A = e1.a - lm.a, B = e1.b - lm.b, C = r * std::sqrt(e1.a * e1.a + 1.);
    double qa = A * A, qb = 2 * A * B, qc = B * B - C * C;
    double disc = std::sqrt(qb * qb - 4 * qa * qc);
    double x1 = (-qb + disc) / (2 * qa), x2 = (-qb - disc) / (2 * qa);
    double xm = std::fmin(x1, x2);
    c = {xm, lm.a * xm + lm.b};
  }
  pt trc = line_foot(e1, c), blc = line_foot(e2, c);
  t.has_fillet = 1;
  t.fillet_r = r;
  t.fillet_t1 = angle_pos(c.y - trc.y, c.x - trc.x);
  t.fillet_t2 = angle_pos(c.y - blc.y, c.x - blc.x);
  t.fillet_p[0] = c.x; t.fillet_p[1] = c.y;
  pt poly[5] = {tl, tr, trc, blc, bl};
  set_polygon(t, poly, 5);
  return br.x;
}

double tool_low(const mfc_tool_state &t) {  // tool.cpp:187-197
  if (t.has_fillet) return t.fillet_p[1] - t.fillet_r;
  double y = DBL_MAX;
  for (int i = 0; i < t.num_seg; i++) y = std::fmin(y, t.left[i][1]);
  return y;
}

void tool_box(const mfc_tool_state &t, double safety, double &x0, double &x1, double &y0, double &y1) {
  x0 = DBL_MAX; x1 = -DBL_MAX; y0 = DBL_MAX; y1 = -DBL_MAX;  // tool.cpp:156-185
  for (int i = 0; i < t.num_seg; i++) {
    x0 = std::fmin(t.left[i][0], x0); x1 = std::fmax(t.left[i][0], x1);
    y0 = std::fmin(t.left[i][1], y0); y1 = std::fmax(t.left[i][1], y1);
  }
  if (t.has_fillet) { y0 = tool_low(t); x1 = t.fillet_p[0] + t.fillet_r; }
  x0 -= safety; y0 -= safety; x1 += safety; y1 += safety;
}

bool tool_contains(const mfc_tool_state &t, double x, double y) {  // tool.cpp:104-119
  bool in = true;
  for (int i = 0; i < t.num_seg; i++) {
    double dx = t.left[i][0] - x, dy = t.left[i][1] - y;
    in = in && (dx * t.n[i][0] + dy * t.n[i][1]) < 0.;
  }
  if (t.has_fillet) {
    double dx = t.fillet_p[0] - x, dy = t.fillet_p[1] - y;
    in = in || std::sqrt(dx * dx + dy * dy) < t.fillet_r;
  }
  return in;
}

uint64_t splitmix64(uint64_t &s) {
  s += 0x9E3779B97F4A7C15ULL;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

}  // namespace

void make_dev_consts(const sim_constants &c, dev_consts &k) {
  std::memset(&k, 0, sizeof(k));
  k.N = c.N; k.gnx = c.gnx; k.gny = c.gny; k.num_cell = c.num_cell;
  k.bbmin_x = c.bbmin_x; k.bbmin_y = c.bbmin_y; k.cell = c.cell;
  k.dt = c.dt; k.mass = c.mass; k.K = c.K; k.G = c.G; k.rho0 = c.rho0; k.c0 = c.c0;
  k.wdeltap = c.wdeltap; k.av_alpha = c.av_alpha; k.av_beta = c.av_beta; k.av_eta = c.av_eta;
  k.xsph_eps = c.xsph_eps; k.stress_eps = c.stress_eps;
  k.alpha_wp = c.alpha_wp; k.alpha_tool = c.alpha_tool; k.cp = c.cp; k.tq = c.tq; k.eta_heat = c.eta_heat;
  k.T_init = c.T_init;
  k.jc_A = c.jc_A; k.jc_B = c.jc_B; k.jc_C = c.jc_C; k.jc_n = c.jc_n; k.jc_m = c.jc_m;
  k.jc_Tmelt = c.jc_Tmelt; k.jc_Tref = c.jc_Tref; k.jc_eps_dot_ref = c.jc_eps_dot_ref;
  k.contact_alpha = c.contact_alpha; k.mu = c.mu; k.tool_vx = c.tool_vx; k.tool_vy = c.tool_vy;
  k.thermal_bc = c.thermal_bc; k.fric_heat = c.fric_heat; k.thermal_wp = c.thermal_wp; k.pse = c.pse;
  k.pse_quirk = c.pse_quirk;
  k.promote = c.promote;
  k.secant_tol = c.secant_tol;
  k.jc_tanh = c.jc_tanh;
  k.tanh_a = c.tanh_a; k.tanh_b = c.tanh_b; k.tanh_c = c.tanh_c; k.tanh_d = c.tanh_d;
  k.plastic = c.plastic;
  k.stage_limit = 0;
  k.adaptive = c.adaptive;
  k.ncell_table = c.ncell_table;
  k.rothlin = c.rothlin;
}

void adaptive_geometry(const double *px, const double *py, const double *h, int N, double &bbmin_x, double &bbmin_y,
                       double &cell, int64_t &gnx, int64_t &gny) {
  double x0 = px[0], x1 = px[0], y0 = py[0], y1 = py[0], hm = h[0];
  for (int i = 1; i < N; i++) {
    x0 = std::fmin(x0, px[i]); x1 = std::fmax(x1, px[i]);
    y0 = std::fmin(y0, py[i]); y1 = std::fmax(y1, py[i]);
    hm = std::fmax(hm, h[i]);
  }
  bbmin_x = x0 - 1e-6; bbmin_y = y0 - 1e-6;
  const double bx1 = x1 + 1e-6, by1 = y1 + 1e-6;
  cell = hm * 2. / 2.;
  gnx = (int64_t)std::ceil((bx1 - bbmin_x) / cell);
  gny = (int64_t)std::ceil((by1 - bbmin_y) / cell);
}

// the adaptive grid of a scenario: the geometry of its initial particles and a binning table of
// max_cell_factor * N cells (grid_gpu_green(10*n, n) in the reference, benchmarks.cpp:135)
static int set_adaptive(const mfc_config &c, scenario &S, std::string &err) {
  sim_constants &k = S.k;
  if (c.grid_kind != MFC_GRID_GREEN && c.grid_kind != MFC_GRID_ROTHLIN) { err = "unknown grid_kind"; return MFC_E_INVALID; }
  k.rothlin = c.grid_kind == MFC_GRID_ROTHLIN;
  k.adaptive = 0;
  k.ncell_table = k.num_cell;
  if (!c.grid_adaptive) return MFC_OK;
  const int64_t factor = c.max_cell_factor > 0 ? c.max_cell_factor : 10;
  const int64_t table = factor * (int64_t)k.N;
  if (table > 1073741824LL) { err = "adaptive cell table too large"; return MFC_E_RANGE; }
  int64_t gnx, gny;
  adaptive_geometry(S.px.data(), S.py.data(), S.h.data(), k.N, k.bbmin_x, k.bbmin_y, k.cell, gnx, gny);
  if (!(gnx >= 1 && gny >= 1) || gnx * gny >= table) {
    err = "the initial adaptive grid needs more cells than max_cell_factor * N";
    return MFC_E_RANGE;
  }
  k.gnx = (int)gnx; k.gny = (int)gny; k.num_cell = (int)(gnx * gny);
  k.adaptive = 1;
  k.ncell_table = (int)table;
  return MFC_OK;
}

void tool_advance(mfc_tool_state &t, double dt) {
  double sx = dt * t.vel[0], sy = dt * t.vel[1];
  for (int i = 0; i < t.num_seg; i++) {
    t.left[i][0] = t.left[i][0] + sx; t.left[i][1] = t.left[i][1] + sy;
    t.right[i][0] = t.right[i][0] + sx; t.right[i][1] = t.right[i][1] + sy;
    line2 l = line_through({t.left[i][0], t.left[i][1]}, {t.right[i][0], t.right[i][1]});
    t.line_a[i] = l.a; t.line_b[i] = l.b; t.vertical[i] = l.vertical ? 1 : 0;
  }
  if (t.has_fillet) { t.fillet_p[0] = t.fillet_p[0] + sx; t.fillet_p[1] = t.fillet_p[1] + sy; }
}

void config_ref_cut(mfc_config &c, int ny, double rake, double clear, double chamfer, double speed, double feed) {
  std::memset(&c, 0, sizeof(c));
  c.lattice = MFC_LATTICE_REF;
  c.ny = ny;
  c.hi_x = 0.200; c.hi_y = 0.060; c.lo_x = 0.005;
  c.hdx = 1.5;
  c.fixed_bottom_rows = 1; c.fix_left = 1; c.fix_right = 1;
  c.E = 1.1; c.nu = 0.35; c.rho0 = 4.429998;
  c.jc_A = 0.0086200; c.jc_B = 0.0033100; c.jc_C = 0.0100000; c.jc_m = 0.8; c.jc_n = 0.34;
  c.jc_Tref = 300.; c.jc_Tmelt = 1836.0000; c.jc_eps_dot_ref = 1e-6;
  c.cp = 553 * 1e-8; c.tq = 0.9; c.k_wp = 7.1 * 1e-13; c.eta = 0.9; c.T_init = c.jc_Tref;
  c.rho_tool = 15.25; c.cp_tool = 292 * 1e-08; c.k_tool = 88 * 1e-13;
  c.av_alpha = 1.; c.av_beta = 1.; c.av_eta = 0.1; c.xsph_eps = 0.5; c.stress_eps = 0.3;
  c.tool_tl_x = -0.05 + 0.004; c.tool_tl_y = 0.0986074;
  c.tool_length = -0.0086824 - -0.05;
  c.tool_height = 0.0986074 - 0.0555074;
  c.rake_deg = rake; c.clear_deg = clear; c.edge_radius = chamfer;
  c.feed = feed; c.speed = speed; c.mu = 0.35; c.contact_alpha = 0.1;
  c.sample_tool = 1;
  c.thermal_solver = MFC_THERMAL_BROOKSHAW;
  c.precision = MFC_PREC_FP64;
  c.cfl = 0.5;
}

// material, Johnson-Cook, thermal and corrector constants of every scenario
// (benchmarks.cpp:279-332, :31-62, :160-186)
void set_constants(const mfc_config &c, double dx, sim_constants &k) {
  const double hdx = c.hdx;
  // material constants (benchmarks.cpp:279-284)
  k.rho0 = c.rho0;
  k.G = c.E / (2. * (1. + c.nu));
  k.K = 2.0 * k.G * (1 + c.nu) / (3 * (1 - 2 * c.nu));
  k.mass = dx * dx * c.rho0;
  k.c0 = c.promote_fp64 ? std::sqrt(k.K / k.rho0) : (double)sqrtf((float)(k.K / k.rho0));
  k.promote = c.promote_fp64 != 0;
  k.secant_tol = (c.secant_tol > 0.) ? c.secant_tol : 1e-6;
  k.jc_tanh = c.jc_model == MFC_JC_TANH;
  k.tanh_a = c.jc_tanh_a; k.tanh_b = c.jc_tanh_b; k.tanh_c = c.jc_tanh_c; k.tanh_d = c.jc_tanh_d;
  k.jc_A = c.jc_A; k.jc_B = c.jc_B; k.jc_C = c.jc_C; k.jc_n = c.jc_n; k.jc_m = c.jc_m;
  k.jc_Tmelt = c.jc_Tmelt; k.jc_Tref = c.jc_Tref; k.jc_eps_dot_ref = c.jc_eps_dot_ref;
  k.cp = c.cp; k.tq = c.tq; k.eta_heat = c.eta; k.T_init = c.T_init;
  k.alpha_wp = c.k_wp / (c.rho0 * c.cp);
  // the tool's diffusivity; 0 without tool material (the body scenarios zero it), which also
  // keeps the thermal boundary condition off as the reference leaves it without a tool
  k.alpha_tool = (c.rho_tool * c.cp_tool > 0.) ? c.k_tool / (c.rho_tool * c.cp_tool) : 0.;
  k.av_alpha = c.av_alpha; k.av_beta = c.av_beta; k.av_eta = c.av_eta;
  k.xsph_eps = c.xsph_eps; k.stress_eps = c.stress_eps;
  {  // W(dx) of the cubic spline (benchmarks.cpp:327-332)
    double h1 = 1. / (hdx * dx);
    double q = dx * h1;
    double fac = 10 * (M_1_PI) / 7.0 * h1 * h1;
    k.wdeltap = fac * (1 - 1.5 * q * q * (1 - 0.5 * q));
  }
  k.plastic = c.elastic == 0;
}

int build_bodies(const mfc_config &c, scenario &S, std::string &err);

// the corrector constants and solver settings of both verification scenarios
// (benchmarks.cpp:50-62, :178-186)
static void config_bodies(mfc_config &c, int lattice, int nbox) {
  std::memset(&c, 0, sizeof(c));
  c.lattice = lattice;
  c.ny = nbox;
  c.hdx = 1.7;
  c.av_alpha = 1.; c.av_beta = 1.; c.av_eta = 0.1; c.xsph_eps = 0.5; c.stress_eps = 0.3;
  c.cp = 1.;  // unused: no plastic heat (tq = 0) and no conduction (k_wp = 0)
  c.thermal_solver = MFC_THERMAL_BROOKSHAW;
  c.precision = MFC_PREC_FP64;
}

void config_rings(mfc_config &c, int nbox) {  // benchmarks.cpp:155-253
  config_bodies(c, MFC_LATTICE_RINGS, nbox);
  c.E = 1e7; c.nu = 0.4; c.rho0 = 1.;
  c.elastic = 1;
  c.T_init = 0.;
  c.cfl = 0.5; c.speed = 170.;
}

void config_impact(mfc_config &c, int nbox) {  // benchmarks.cpp:23-153
  config_bodies(c, MFC_LATTICE_IMPACT, nbox);
  c.grid_adaptive = 1;  // grid_gpu_green(10*n, n), benchmarks.cpp:135
  c.max_cell_factor = 10;
  c.E = 200e9; c.nu = 0.29; c.rho0 = 7830.0;
  c.jc_A = 792.0e6; c.jc_B = 510.0e6; c.jc_C = 0.014; c.jc_m = 1.03; c.jc_n = 0.26;
  c.jc_Tref = 273.0; c.jc_Tmelt = 1793.0; c.jc_eps_dot_ref = 1;
  c.T_init = c.jc_Tref;
  c.cfl = 0.3; c.speed = 1e3;
}

int build_scenario(const mfc_config &c, scenario &S, std::string &err) {
  sim_constants &k = S.k;
  std::memset(&k, 0, sizeof(k));
  if (!(c.hdx > 0) || !(c.rho0 > 0) || !(c.cfl > 0) || !(c.cp > 0)) { err = "hdx, rho0, cfl and cp must be positive"; return MFC_E_INVALID; }
  if (c.thermal_solver != MFC_THERMAL_BROOKSHAW && c.thermal_solver != MFC_THERMAL_PSE) { err = "unknown thermal solver"; return MFC_E_INVALID; }
  if (c.precision != MFC_PREC_FP64 && c.precision != MFC_PREC_FP32C) { err = "unknown precision"; return MFC_E_INVALID; }
  if (c.jc_model != MFC_JC_PLAIN && c.jc_model != MFC_JC_TANH) { err = "unknown jc_model"; return MFC_E_INVALID; }
  // lattice (benchmarks.cpp:266-273)
  double hi_y = c.hi_y, lo_x = c.lo_x, hi_x, lo_y, dx;
  int nx, ny = c.ny;
  if (c.lattice == MFC_LATTICE_REF) {
    if (ny < 2) { err = "ny must be >= 2"; return MFC_E_INVALID; }
    hi_x = c.hi_x;
    lo_y = hi_y - 3 * c.feed;
    dx = (hi_y - lo_y) / (ny - 1);
    nx = (int)((hi_x - lo_x) / dx);
  } else if (c.lattice == MFC_LATTICE_EXPLICIT) {
    if (c.nx < 1 || c.ny < 1 || !(c.dx > 0)) { err = "explicit lattice needs nx, ny >= 1 and dx > 0"; return MFC_E_INVALID; }
    dx = c.dx; nx = c.nx;
    lo_y = hi_y - (ny - 1) * dx;
    hi_x = lo_x + nx * dx;
  } else if (c.lattice == MFC_LATTICE_RINGS || c.lattice == MFC_LATTICE_IMPACT) {
    const int rc = build_bodies(c, S, err);
    return rc ? rc : set_adaptive(c, S, err);
  } else {
    err = "unknown lattice rule";
    return MFC_E_INVALID;
  }
  if ((int64_t)nx * ny < 1 || (int64_t)nx * ny > 100000000) { err = "particle count out of range"; return MFC_E_RANGE; }
  const double hdx = c.hdx;
  set_constants(c, dx, k);
  // workpiece particles, x-major (benchmarks.cpp:345-357) with optional seeded jitter
  const int n_wp = nx * ny;
  S.px.resize(n_wp); S.py.resize(n_wp);
  double min_x = FLT_MAX, max_x = -FLT_MAX, min_y = FLT_MAX, max_y = -FLT_MAX;
  uint64_t rng = c.seed;
  const double inv53 = 1.0 / 9007199254740992.0;
  for (int i = 0, q = 0; i < nx; i++)
    for (int j = 0; j < ny; j++, q++) {
      double x = i * dx + lo_x, y = j * dx + lo_y;
      if (c.jitter != 0.) {
        x = x + (2.0 * ((double)(splitmix64(rng) >> 11) * inv53) - 1.0) * c.jitter * dx;
        y = y + (2.0 * ((double)(splitmix64(rng) >> 11) * inv53) - 1.0) * c.jitter * dx;
      }
      S.px[q] = x; S.py[q] = y;
      max_x = std::fmax(max_x, x); min_x = std::fmin(min_x, x);
      max_y = std::fmax(max_y, y); min_y = std::fmin(min_y, y);
    }
  S.rho.assign(n_wp, c.rho0);
  S.h.assign(n_wp, hdx * dx);
  S.T.assign(n_wp, c.T_init);
  S.flags.assign(n_wp, 0);
  const double y_bottom = min_y + (c.fixed_bottom_rows - 0.5) * dx;
  for (int q = 0; q < n_wp; q++) {  // benchmarks.cpp:373-378
    bool fx = S.py[q] < y_bottom;
    if (c.fix_left) fx = fx || (S.px[q] < min_x + dx / 2 && S.py[q] < min_y + 0.5 * (hi_y - lo_y));
    if (c.fix_right) fx = fx || (S.px[q] > max_x - dx);
    S.flags[q] = fx ? MFC_FLAG_FIXED : 0;
  }
  // tool placed with its front at the workpiece edge and its lowest point at the depth of
  // cut (benchmarks.cpp:384-418)
  const double vel_tool = (c.speed == 0.f) ? 83.333328 * 1e-5 : c.speed;
  mfc_tool_state &t = S.tool;
  double front = build_tool(t, {c.tool_tl_x, c.tool_tl_y}, c.tool_length, c.tool_height, c.rake_deg, c.clear_deg, c.edge_radius);
  if (front > lo_x) {
    t.vel[0] = vel_tool; t.vel[1] = 0.;
    tool_advance(t, -((front - lo_x) / vel_tool));
  }
  {
    double current = hi_y - tool_low(t);
    double corr_time = std::fabs(current - c.feed) / vel_tool;
    double sign = (current > c.feed) ? 1 : -1.;
    t.vel[0] = 0.; t.vel[1] = vel_tool;
    tool_advance(t, corr_time * sign);
  }
  t.vel[0] = vel_tool; t.vel[1] = 0.;
  k.contact_alpha = c.contact_alpha; k.mu = c.mu;
  k.tool_vx = t.vel[0]; k.tool_vy = t.vel[1];
  // tool particles on a dx lattice inside the tool (tool.cpp:121-144, benchmarks.cpp:424-459)
  int n_tool = 0;
  if (c.sample_tool) {
    double x0, x1, y0, y1;
    tool_box(t, 0.1 * dx, x0, x1, y0, y1);
    int tnx = (int)((x1 - x0) / dx + 1), tny = (int)((y1 - y0) / dx + 1);
    std::vector<double> tx, ty;
    for (int i = 0; i < tnx; i++)
      for (int j = 0; j < tny; j++) {
        double x = i * dx + x0, y = j * dx + y0;
        if (tool_contains(t, x, y)) { tx.push_back(x); ty.push_back(y); }
      }
    if (tx.empty()) { err = "tool sampling produced no particles"; return MFC_E_INVALID; }
    double bx = FLT_MAX, by = -FLT_MAX;
    for (size_t q = 0; q < tx.size(); q++) { bx = std::fmin(bx, tx[q]); by = std::fmax(by, ty[q]); }
    n_tool = (int)tx.size();
    for (int q = 0; q < n_tool; q++) {
      S.px.push_back(tx[q]); S.py.push_back(ty[q]);
      S.rho.push_back(c.rho_tool); S.h.push_back(hdx * dx); S.T.push_back(c.T_init);
      bool fx = tx[q] < bx + dx / 2 || ty[q] > by - dx / 2;
      S.flags.push_back(MFC_FLAG_TOOL | (fx ? MFC_FLAG_FIXED : 0));
    }
  }
  const int N = n_wp + n_tool;
  k.N = N; k.n_tool = n_tool;
  // domain box: particles, tool box, room for the chip (benchmarks.cpp:464-487)
  double bx0 = FLT_MAX, by0 = FLT_MAX, bx1 = -FLT_MAX, by1 = -FLT_MAX;
  for (int q = 0; q < N; q++) {
    bx0 = std::fmin(S.px[q], bx0); by0 = std::fmin(S.py[q], by0);
    bx1 = std::fmax(S.px[q], bx1); by1 = std::fmax(S.py[q], by1);
  }
  bx0 -= 1e-8; by0 -= 1e-8; bx1 += 1e-8; by1 += 1e-8;
  {
    double tx0, tx1, ty0, ty1;
    tool_box(t, 1e-3, tx0, tx1, ty0, ty1);
    bx0 = std::fmin(bx0, tx0);
    by0 = std::fmin(by0, ty0);
    bx1 = std::fmax(bx1, tx1) + 10 * dx;
    by1 = std::fmax(by1, ty1);
    by1 = std::fmax(by1, hi_y + hi_x - lo_x);
  }
  // locked cell grid with cell size h (grid.cu:139-158)
  k.bbmin_x = bx0 - 1e-6; k.bbmin_y = by0 - 1e-6;
  k.cell = hdx * dx * 2. / 2.;
  double gnx = std::ceil(((bx1 + 1e-6) - k.bbmin_x) / k.cell);
  double gny = std::ceil(((by1 + 1e-6) - k.bbmin_y) / k.cell);
  if (!(gnx >= 1 && gny >= 1 && gnx * gny <= 1073741824.0)) { err = "cell grid too large"; return MFC_E_RANGE; }
  k.gnx = (int)gnx; k.gny = (int)gny; k.num_cell = k.gnx * k.gny;
  // switches (actions_gpu.cu:448-468, interactions_gpu.cu:443)
  k.thermal_bc = k.alpha_tool != 0.;
  k.fric_heat = k.eta_heat != 0.;
  k.thermal_wp = k.alpha_wp != 0.;
  k.pse = c.thermal_solver == MFC_THERMAL_PSE;
  k.pse_quirk = c.pse_flag_quirk != 0;
  // CFL time step (benchmarks.cpp:530-531)
  k.dt = c.cfl * hdx * dx / (std::sqrt(k.K / k.rho0) + std::sqrt(vel_tool));
  k.t_final = 0.1 / vel_tool;
  return set_adaptive(c, S, err);
}

// the reference's verification scenarios: two colliding elastic rings (setup_rings,
// benchmarks.cpp:155-253) and a disk flying into a wall (setup_impact, benchmarks.cpp:23-153).
// Neither has a tool: the contact test of an empty tool finds no penetration and the tool force
// stays zero. The locked cell grid is the box of the reference's fixed-grid constructor
// (benchmarks.cpp:218 for the rings, :136 for the impact); with grid_adaptive (the impact's
// active line, :135) the geometry follows the particles instead.
int build_bodies(const mfc_config &c, scenario &S, std::string &err) {
  sim_constants &k = S.k;
  const bool rings = c.lattice == MFC_LATTICE_RINGS;
  const int nbox = c.ny;
  if (nbox < 2 || nbox > 20000) { err = "ny (lattice points across a body) must be in [2, 20000]"; return MFC_E_INVALID; }
  if (!(c.speed > 0)) { err = "speed (the velocity of the CFL rule) must be positive"; return MFC_E_INVALID; }
  const double hdx = c.hdx;
  const double ro = rings ? 0.04 : 0.2, ri = 0.03;
  const double dx = 2 * ro / (nbox - 1);
  set_constants(c, dx, k);
  S.px.clear(); S.py.clear(); S.vx.clear(); S.vy.clear();
  auto push = [&](double x, double y, double vx) {
    S.px.push_back(x); S.py.push_back(y); S.vx.push_back(vx); S.vy.push_back(0.);
  };
  if (rings) {  // benchmarks.cpp:189-205, 238-240: interleaved left/right ring points
    const double spacing = ro + 0.001;
    for (int i = 0; i < nbox; i++)
      for (int j = 0; j < nbox; j++) {
        double px = -ro + i * dx, py = -ro + j * dx;
        double dist = std::sqrt(px * px + py * py);
        if (dist < ro && dist >= ri) {
          push(px - spacing, py, 0.);
          push(px + spacing, py, 0.);
        }
      }
    for (size_t q = 0; q < S.px.size(); q++) S.vx[q] = (S.px[q] < 0) ? 200 : -200;
  } else {  // benchmarks.cpp:64-95, 111-118: the flyer, then the wall column by column
    const double spacing = ro + 0.05;
    for (int i = 0; i < nbox; i++)
      for (int j = 0; j < nbox; j++) {
        double px = -ro + i * dx, py = -ro + j * dx;
        if (std::sqrt(px * px + py * py) < ro) push(px - spacing, py, 1e3);
      }
    for (int i = 0; i < nbox / 2; i++)
      for (int j = 0; j < 10 * nbox; j++) {
        if (j * dx - 1. > 1.) continue;
        push(i * dx, j * dx - 1., 0.);
      }
  }
  const size_t N = S.px.size();
  if (N < 1 || N > 100000000) { err = "particle count out of range"; return MFC_E_RANGE; }
  S.rho.assign(N, c.rho0);
  S.h.assign(N, hdx * dx);
  S.T.assign(N, c.T_init);
  S.flags.assign(N, 0);
  if (!rings)  // top and bottom of the wall (benchmarks.cpp:120-125)
    for (size_t q = 0; q < N; q++)
      if (S.py[q] < -1 + dx / 2 || S.py[q] > 1 - dx / 2) S.flags[q] = MFC_FLAG_FIXED;
  std::memset(&S.tool, 0, sizeof(S.tool));
  k.contact_alpha = 0.; k.mu = 0.; k.tool_vx = 0.; k.tool_vy = 0.;
  k.N = (int)N; k.n_tool = 0;
  const double bx0 = rings ? -5 * ro : -.5, by0 = rings ? -2 * ro : -1.2;
  const double bx1 = rings ? 5 * ro : 1., by1 = rings ? 2 * ro : 1.2;
  k.bbmin_x = bx0 - 1e-6; k.bbmin_y = by0 - 1e-6;
  k.cell = hdx * dx * 2. / 2.;
  double gnx = std::ceil(((bx1 + 1e-6) - k.bbmin_x) / k.cell);
  double gny = std::ceil(((by1 + 1e-6) - k.bbmin_y) / k.cell);
  if (!(gnx >= 1 && gny >= 1 && gnx * gny <= 1073741824.0)) { err = "cell grid too large"; return MFC_E_RANGE; }
  k.gnx = (int)gnx; k.gny = (int)gny; k.num_cell = k.gnx * k.gny;
  // no tool: no thermal contact and no friction heat; conduction in the body if k_wp != 0
  k.thermal_bc = 0;
  k.fric_heat = 0;
  k.thermal_wp = k.alpha_wp != 0.;
  k.pse = c.thermal_solver == MFC_THERMAL_PSE;
  k.pse_quirk = c.pse_flag_quirk != 0;
  // CFL step and end time (benchmarks.cpp:210-211, :98-99)
  k.dt = c.cfl * hdx * dx / (std::sqrt(k.K / k.rho0) + std::sqrt(c.speed));
  k.t_final = (rings ? 2.5 : 2.0) * ro / c.speed;
  return MFC_OK;
}

}  // namespace mfc
