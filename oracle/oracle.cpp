// oracle.cpp — FP64 host restatement of the reference's 2D SPH cutting step.
//
// TEST INFRASTRUCTURE ONLY. This file is the correctness oracle and the CPU baseline of the
// MI355X solver; only the tests, the smoke check and the benchmark's CPU-baseline leg load it.
// The product library never links, loads or calls it.
//
// It restates, one loop per reference kernel and in the reference's order, the arithmetic of
//   leap_frog::step            leap_frog.cu:94-155  (init/predict/correct :30-92)
//   compute_hashes             grid.cu:34-51
//   sort + cell start/end      grid_gpu_green.cu:22-116, 177-201 (thrust::sort_by_key is stable)
//   EOS, artificial stress     actions_gpu.cu:35-97
//   do_interactions_monaghan   interactions_gpu.cu:146-345, cubic_spline kernels.cuh:30-70
//   do_interactions_heat_pse   interactions_gpu.cu:88-144, lapl_pse kernels.cuh:72-89
//   Jaumann rate               actions_gpu.cu:99-133
//   fric heat, continuity, momentum, advection   actions_gpu.cu:135-224
//   Johnson-Cook radial return actions_gpu.cu:226-285, plasticity.cuh:30-119
//   tool move + contact        tool.cpp:211-221, geometric_primitives.cpp:70-85,
//                              tool_gpu.cu:29-221
//   boundary conditions, NaN blanking            actions_gpu.cu:287-349
// and the setup of setup_ref_cut (benchmarks.cpp:255-537, tool.cpp:20-144,342-375,
// geometric_primitives.cpp:20-146).
//
// Every floating-point expression keeps the reference's operand order (C++ left-to-right
// evaluation) and its float truncations (sqrtf, fmaxf, atan2f/cosf/sinf in FP64 mode). The
// build compiles this file with -ffp-contract=off. Device transcendentals come from
// mfc_math.h (pinned against libm by the tests); host-only setup code calls libm as the
// reference's host code does.
//
// Parallelism: OpenMP over particles. Every loop writes only its own particle and every
// neighbour sum is a gather in a fixed order, so results are independent of the thread count.
#include "../include/mfc.h"
#include "../mfree_iwf-ul_cut_gpu_amd/csrc/mfc_math.h"

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct V2 { double x, y; };
struct V4 { double x, y, z, w; };

inline V2 v2(double x, double y) { V2 r; r.x = x; r.y = y; return r; }
inline V2 add(V2 a, V2 b) { return v2(a.x + b.x, a.y + b.y); }
inline V2 sub(V2 a, V2 b) { return v2(a.x - b.x, a.y - b.y); }
inline V2 mul(double s, V2 a) { return v2(s * a.x, s * a.y); }
inline V2 mulv(V2 a, double s) { return v2(a.x * s, a.y * s); }
// glm::dot(vec2) = (a*b).x + (a*b).y
inline double dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
inline double length(V2 a) { return std::sqrt(dot(a, a)); }
// glm::normalize = v * inversesqrt(dot(v,v)), inversesqrt(x) = 1/sqrt(x)
inline V2 normalize(V2 a) { double is = 1.0 / std::sqrt(dot(a, a)); return v2(a.x * is, a.y * is); }

inline double f32sqrt(double x) { return (double)sqrtf((float)x); }
// sqrtf of the reference's FP64 build, or the FP64 square root when its truncations are promoted
inline double tsqrt(double x, bool promote) { return promote ? std::sqrt(x) : f32sqrt(x); }

// ---------------------------------------------------------------- geometric primitives
// geometric_primitives.h:26-63, geometric_primitives.cpp:20-85,112-122
struct line_t {
  double a = 0., b = 0.;
  bool vertical = false;
  line_t() {}
  line_t(double a_, double b_, bool v) : a(a_), b(b_), vertical(v) {}
  line_t(V2 p1, V2 p2) {
    double Mxx = p1.x, Mxy = 1.;
    double Myx = p2.x, Myy = 1.;
    double detM = Mxx * Myy - Mxy * Myx;
    if (std::fabs(detM) < 1e-12) { vertical = true; a = DBL_MAX; b = p1.x; return; }
    a = (p1.y * Myy - Mxy * p2.y) / detM;
    b = (p2.y * Mxx - Myx * p1.y) / detM;
  }
  V2 intersect(const line_t &l) const {
    if (vertical && l.vertical) return v2(DBL_MAX, DBL_MAX);
    if (std::fabs(a - l.a) < 1e-12) return v2(DBL_MAX, DBL_MAX);
    if (vertical) { double vx = b; return v2(vx, l.a * vx + l.b); }
    if (l.vertical) { double vx = l.b; return v2(vx, a * vx + b); }
    double x = (l.b - b) / (a - l.a);
    return v2(x, a * x + b);
  }
  V2 closest_point(V2 xq) const {
    if (vertical) return v2(b, xq.y);
    double bb = -1, cc = b, aa = a;
    double px = (bb * (bb * xq.x - aa * xq.y) - aa * cc) / (aa * aa + bb * bb);
    double py = (aa * (-bb * xq.x + aa * xq.y) - bb * cc) / (aa * aa + bb * bb);
    return v2(px, py);
  }
};

struct segment_t {
  V2 left, right, n;
  line_t l;
  segment_t(V2 lft, V2 rgt) {
    left = lft; right = rgt;
    V2 dist = sub(right, left);
    n = normalize(v2(dist.y, -dist.x));
    l = line_t(left, right);
  }
};

struct circle_t { double r = 0., t1 = 0., t2 = 0.; V2 p = {0., 0.}; };

// tool.cpp host class: construction, feed placement, sampling, per-step move
struct tool_t {
  std::vector<segment_t> seg;
  bool has_fillet = false;
  circle_t fillet;
  V2 vel = {0., 0.};
  double mu = 0., front = 0.;

  static double host_atan2_pos(double y, double x) {  // tool.cpp:27-34
    double t = std::atan2(y, x);
    return (t > 0.) ? t : t + 2 * M_PI;
  }
  void construct_segments(const std::vector<V2> &p) {  // tool.cpp:64-71
    unsigned n = (unsigned)p.size();
    for (unsigned i = 0; i < n; i++) {
      unsigned next = (i + 1 > n - 1) ? 0 : i + 1;
      seg.push_back(segment_t(p[i], p[next]));
    }
  }
  V2 fit_fillet(double r, line_t lm, line_t l1) const {  // tool.cpp:36-62
    if (l1.vertical) {
      line_t lpar(DBL_MAX, l1.b - r, true);
      return lpar.intersect(lm);
    }
    double A0 = lm.a, B0 = lm.b, a = l1.a, b = l1.b;
    double A = a - A0, B = b - B0, C = r * std::sqrt(a * a + 1.);
    double qa = A * A, qb = 2 * A * B, qc = B * B - C * C;
    double x1 = (-qb + std::sqrt(qb * qb - 4 * qa * qc)) / (2 * qa);
    double x2 = (-qb - std::sqrt(qb * qb - 4 * qa * qc)) / (2 * qa);
    double xm = std::fmin(x1, x2);
    return v2(xm, lm.a * xm + lm.b);
  }
  std::vector<V2> points_and_fillet(V2 tl, V2 tr, V2 br, V2 bl, double r) {  // tool.cpp:73-102
    V2 pm = br;
    V2 nt = normalize(sub(tr, br));
    V2 nl = normalize(sub(bl, br));
    V2 nm = normalize(mul(0.5, add(nt, nl)));
    line_t lm(pm, add(pm, nm));
    line_t l1(tr, br), l2(bl, br);
    V2 p = fit_fillet(r, lm, l1);
    V2 trc = l1.closest_point(p);
    V2 blc = l2.closest_point(p);
    fillet.r = r;
    fillet.t1 = host_atan2_pos(p.y - trc.y, p.x - trc.x);
    fillet.t2 = host_atan2_pos(p.y - blc.y, p.x - blc.x);
    fillet.p = p;
    has_fillet = true;
    return {tl, tr, trc, blc, bl};
  }
  // tool(tl, length, height, rake, clearance, r, mu)  tool.cpp:342-375
  void build(V2 tl, double length, double height, double rake, double clear, double r, double mu_) {
    mu = mu_;
    V2 tr = v2(tl.x + length, tl.y);
    V2 bl = v2(tl.x, tl.y - height);
    double ar = rake * M_PI / 180.;
    double af = (180 - 90 - clear) * M_PI / 180.;
    // down * mat2(cos,-sin,sin,cos): row vector times column-major matrix
    V2 down = v2(0., -1.);
    V2 rr = v2(down.x * std::cos(ar) + down.y * (-std::sin(ar)), down.x * std::sin(ar) + down.y * std::cos(ar));
    V2 rf = v2(down.x * std::cos(af) + down.y * (-std::sin(af)), down.x * std::sin(af) + down.y * std::cos(af));
    V2 trc = add(tr, rr);
    V2 blc = add(bl, rf);
    line_t l1(tr, trc), l2(bl, blc);
    V2 br = l1.intersect(l2);
    if (r == 0.) {
      construct_segments({tl, tr, br, bl});
    } else {
      construct_segments(points_and_fillet(tl, tr, br, bl, r));
    }
    front = br.x;
  }
  double low() const {  // tool.cpp:187-197
    if (has_fillet) return fillet.p.y - fillet.r;
    double y = DBL_MAX;
    for (auto &s : seg) y = std::fmin(y, s.left.y);
    return y;
  }
  void safe_bb(double safety, double &x0, double &x1, double &y0, double &y1) const {  // tool.cpp:156-185
    x0 = DBL_MAX; x1 = -DBL_MAX; y0 = DBL_MAX; y1 = -DBL_MAX;
    for (auto &s : seg) {
      x0 = std::fmin(s.left.x, x0); x1 = std::fmax(s.left.x, x1);
      y0 = std::fmin(s.left.y, y0); y1 = std::fmax(s.left.y, y1);
    }
    if (has_fillet) { y0 = low(); x1 = fillet.p.x + fillet.r; }
    x0 -= safety; y0 -= safety; x1 += safety; y1 += safety;
  }
  bool inside(V2 q) const {  // tool.cpp:104-119
    bool in = true;
    for (auto &s : seg) in = in && dot(sub(s.left, q), s.n) < 0.;
    if (has_fillet) in = in || length(sub(fillet.p, q)) < fillet.r;
    return in;
  }
  void update(double dt) {  // tool.cpp:211-221
    for (auto &s : seg) {
      s.left = add(s.left, mul(dt, vel));
      s.right = add(s.right, mul(dt, vel));
      s.l = line_t(s.left, s.right);
    }
    if (has_fillet) fillet.p = add(fillet.p, mul(dt, vel));
  }
  std::vector<V2> sample(double dx) const {  // tool.cpp:121-144
    std::vector<V2> out;
    double x0, x1, y0, y1;
    safe_bb(0.1 * dx, x0, x1, y0, y1);
    int nx = (int)((x1 - x0) / dx + 1);
    int ny = (int)((y1 - y0) / dx + 1);
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++) {
        V2 p = v2(i * dx + x0, j * dx + y0);
        if (inside(p)) out.push_back(p);
      }
    return out;
  }
};

// ---------------------------------------------------------------- device-side tool math
// tool_gpu.cu:29-139 (contact geometry on the GPU, transcendental from mfc_math.h)
double dev_myatan2(double y, double x) {
  double t = mfcm::atan2(y, x);
  return (t > 0.) ? t : t + 2 * M_PI;
}

double circle_segment_distance(const circle_t &f, V2 qp) {
  double Ax = f.p.x, Ay = f.p.y, Bx = qp.x, By = qp.y, r = f.r;
  double Cx = Ax + r * (Bx - Ax) / std::sqrt((Bx - Ax) * (Bx - Ax) + (By - Ay) * (By - Ay));
  double Cy = Ay + r * (By - Ay) / std::sqrt((Bx - Ax) * (Bx - Ax) + (By - Ay) * (By - Ay));
  V2 cp = v2(Cx, Cy);
  double t = dev_myatan2(f.p.y - cp.y, f.p.x - cp.x);
  bool valid = t > std::fmin(f.t1, f.t2) && t < std::fmax(f.t1, f.t2);
  return valid ? length(sub(cp, qp)) : (double)FLT_MAX;
}

V2 segment_closest_point(const line_t &l, V2 xq) {
  if (l.vertical) return v2(l.b, xq.y);
  double b = l.b, a = l.a, bb = -1, cc = b, aa = a;
  double px = (bb * (bb * xq.x - aa * xq.y) - aa * cc) / (aa * aa + bb * bb);
  double py = (aa * (-bb * xq.x + aa * xq.y) - bb * cc) / (aa * aa + bb * bb);
  return v2(px, py);
}

bool establish_contact(const tool_t &T, V2 qp, V2 &cp, V2 &n) {
  int num_seg = (int)T.seg.size();
  // an empty tool (the reference's default tool(), mfree_gpu.cpp:214-216): its contact constants
  // are zero, so whatever the test returns the forces are zero; here it is no contact
  if (num_seg == 0) return false;
  bool has_fillet = num_seg == 5;
  bool in = true;
  for (int i = 0; i < num_seg; i++) in = in && dot(sub(T.seg[i].left, qp), T.seg[i].n) < 0.;
  if (has_fillet) in = in || length(sub(T.fillet.p, qp)) < T.fillet.r;
  if (!in) return false;
  double depth = FLT_MAX;
  for (int i = 0; i < num_seg; i++) {
    double d;
    const segment_t &s = T.seg[i];
    if (s.l.vertical) {
      d = std::fabs(s.left.x - qp.x);
      if (d < depth) { depth = d; n = normalize(v2(s.left.x - qp.x, 0.)); cp = v2(s.left.x, qp.y); }
    } else if (has_fillet && i == 2) {
      d = circle_segment_distance(T.fillet, qp);
      if (d < depth) {
        depth = d;
        n = normalize(sub(qp, T.fillet.p));
        cp = add(T.fillet.p, mulv(n, T.fillet.r));
      }
    } else {
      V2 p = segment_closest_point(s.l, qp);
      d = length(sub(p, qp));
      if (d < depth) { depth = d; n = v2(-s.n.x, -s.n.y); cp = p; }
    }
  }
  return true;
}

// ---------------------------------------------------------------- kernels.cuh
constexpr double PI_FAC = 0.454728408833987;

inline void cubic_spline(V2 pi, V2 pj, double hi, double w[3], bool promote = false) {  // kernels.cuh:30-70
  w[0] = w[1] = w[2] = 0.;
  double xij = pi.x - pj.x, yij = pi.y - pj.y;
  double rr2 = xij * xij + yij * yij;
  double h1 = 1 / hi;
  double fourh2 = 4 * hi * hi;
  if (rr2 >= fourh2 || rr2 < 1e-8) return;
  double der, val;
  double rad = tsqrt(rr2, promote);
  double q = rad * h1;
  double fac = PI_FAC * h1 * h1;
  if (q > 1) {
    double m = 2 - q, m2 = m * m;
    val = 0.25 * m2 * m2 * m2;
    der = -0.75 * m2 * h1 / rad;
  } else {
    val = 1 - 1.5 * q * q * (1 - 0.5 * q);
    der = -3.0 * q * (1 - 0.75 * q) * h1 / rad;
  }
  w[0] = val * fac;
  w[1] = der * xij * fac;
  w[2] = der * yij * fac;
}

inline double lapl_pse(V2 pi, V2 pj, double hi) {  // kernels.cuh:72-89
  double xij = pi.x - pj.x, yij = pi.y - pj.y;
  double xx = std::sqrt(xij * xij + yij * yij);
  double h2 = hi * hi, h4 = h2 * h2;
  return +4. / (h4 * M_PI) * mfcm::exp(-xx * xx / (h2));
}

// ---------------------------------------------------------------- plasticity.cuh
struct joco_t {
  double A, B, C, n, m, Tmelt, Tref, eps_dot_ref;
  bool tanh = false;                    // m_s_yield = sigma_yield_tanh (plasticity.cuh:72)
  double ta = 0., tb = 0., tc = 0., td = 0.;  // joco_constants tanh_a..d
};

inline double sigma_yield(const joco_t &jc, double eps_pl, double eps_pl_dot, double t) {
  double theta = (t - jc.Tref) / (jc.Tmelt - jc.Tref);
  double TA = jc.A + jc.B * mfcm::pow(eps_pl, jc.n);
  double TB = 1.0;
  double ed = eps_pl_dot / jc.eps_dot_ref;
  if (ed > 1.0) TB = 1.0 + jc.C * mfcm::log(ed);
  else TB = mfcm::pow((1.0 + ed), jc.C);
  double TC = 1.0 - mfcm::pow(theta, jc.m);
  return TA * TB * TC;
}

// sigma_yield_tanh (plasticity.cuh:48-70), the tanh-modified flow stress of Calamaz 2008
inline double sigma_yield_tanh(const joco_t &jc, double eps_pl, double eps_dot_pl, double t) {
  double theta = (t - jc.Tref) / (jc.Tmelt - jc.Tref);
  double TA = jc.A + jc.B * mfcm::pow(eps_pl, jc.n) * 1.0 / mfcm::exp(mfcm::pow(eps_pl, jc.ta));
  double TB = 1.0;
  double ed = eps_dot_pl / jc.eps_dot_ref;
  if (ed > 1.0) TB = 1.0 + jc.C * mfcm::log(ed);
  else TB = mfcm::pow((1.0 + ed), jc.C);
  double TC = 1.0 - mfcm::pow(theta, jc.m);
  double tD = 1.0 - mfcm::pow((t / jc.Tmelt), jc.td);
  double tS = mfcm::pow(t / jc.Tmelt, jc.tb);
  double TD = tD + (1.0 - tD) * mfcm::tanh(1.0 / (mfcm::pow(eps_pl + tS, jc.tc)));
  return TA * TB * TC * TD;
}

// opt_fun (plasticity.cuh:74-83) through the function pointer m_s_yield (:72); the yield test
// and the plastic heat of the step call sigma_yield itself (actions_gpu.cu:253,281)
inline double opt_fun(const joco_t &jc, double dl, double norm_Strial, double eps_pl_init, double t,
                      double delta_t, double mu) {
  double e = eps_pl_init + std::sqrt(2.0 / 3.0) * std::fmax(dl, 0.);
  double ed = std::sqrt(2.0 / 3.0) * std::fmax(dl, 0.) / delta_t;
  double sY = jc.tanh ? sigma_yield_tanh(jc, e, ed, t) : sigma_yield(jc, e, ed, t);
  return norm_Strial - 2 * mu * dl - std::sqrt(2.0 / 3.0) * sY;
}

inline double solve_secant(const joco_t &jc, double init, double tol, double norm_Strial, double eps_pl_init,
                           double t, double delta_t, double mu, int *iters) {
  double dl = init, dl_old = dl;
  int iter = 0, max_iter = 100;
  while (true) {
    double g = opt_fun(jc, dl, norm_Strial, eps_pl_init, t, delta_t, mu);
    double dl1 = dl * 1.01;
    double g1 = opt_fun(jc, dl1, norm_Strial, eps_pl_init, t, delta_t, mu);
    double slope = (g - g1) / (dl - dl1);
    dl_old = dl;
    dl = dl - g / slope;
    if (dl < 0.) dl = -0.1 * dl;
    if (std::fabs(dl - dl_old) < tol) { if (iters) *iters = iter; return dl; }
    if (iter > max_iter) { if (iters) *iters = iter; return dl; }
    iter++;
  }
}

// ---------------------------------------------------------------- simulation
struct sim_t {
  mfc_config cfg;
  int N = 0, n_tool = 0;
  // physical constants (constants_structs.h)
  double E, nu, rho0, K, G, mass;
  double cp, tq, alpha_wp, eta_heat, T_init;
  double alpha_tool;
  double wdeltap, c_alpha, c_beta, c_eta, xspheps, stresseps;
  joco_t jc;
  bool thermal_bc = false, fric_heat = false, thermal_wp = false, pse = false;
  bool promote = false;     // reference float truncations evaluated in FP64
  bool plastic = true;      // m_plastic (actions_gpu.cu:25, set by actions_setup_johnson_cook_constants)
  double secant_tol = 1e-6; // actions_gpu.cu:261
  // grid (grid.cu:139-158); an adaptive grid recomputes it every step (grid.cu:97-136) and
  // sorts over a table of ncell_table = max_cell cells (num_cell for a locked grid)
  int gnx = 0, gny = 0, num_cell = 0;
  double bbmin_x = 0, bbmin_y = 0, gdx = 0;
  bool adaptive = false;
  int ncell_table = 0;
  int64_t grid_overflow = 0;
  bool rothlin = false;  // grid_gpu_rothlin: ranges by unique/scatter/scan, ft not sorted
  // tool
  tool_t tool;
  double contact_alpha = 0.;
  double dt = 0., t_final = 0.;
  int64_t step = 0;
  int64_t out_of_grid = 0;
  // particle_gpu (particle_gpu.h:32-87); flags as doubles
  std::vector<V2> pos, vel, fc, ft, n, pos_t, vel_t;
  std::vector<double> h, rho, p, fixed, blanked, tool_p, eps_pl, eps_pl_dot, T, rho_t, T_t;
  std::vector<V4> S, R, S_t, S_der, v_der;
  std::vector<int> idx, hash, id;
  // leap_frog (leap_frog.h:36-50)
  std::vector<V2> pos_init, vel_init;
  std::vector<V4> S_init;
  std::vector<double> rho_init, T_init_arr;
  std::vector<int> cell_start, cell_end;
  std::vector<double> force_hist;
};

thread_local std::string g_err;

uint64_t splitmix64(uint64_t &s) {
  s += 0x9E3779B97F4A7C15ULL;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
double u01(uint64_t &s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

int setup_bodies(sim_t &s, const mfc_config &c);

// grid_base::update_geometry (grid.cu:97-136) with the kernel width 2 of leap_frog::step: the
// box of all positions (thrust::min_element / max_element over the whole array, blanked
// particles included) widened by 1e-6, cell size max h. NaN positions are passed over. A
// geometry that needs at least max_cell cells (the reference asserts) is not taken: the
// previous one stays and the step is counted.
void update_geometry(sim_t &s) {
  double x0 = s.pos[0].x, x1 = x0, y0 = s.pos[0].y, y1 = y0, hm = s.h[0];
  for (int i = 1; i < s.N; i++) {
    x0 = std::fmin(x0, s.pos[i].x); x1 = std::fmax(x1, s.pos[i].x);
    y0 = std::fmin(y0, s.pos[i].y); y1 = std::fmax(y1, s.pos[i].y);
    hm = std::fmax(hm, s.h[i]);
  }
  double bx = x0 - 1e-6, by = y0 - 1e-6;
  double ex = x1 + 1e-6, ey = y1 + 1e-6;
  double cell = hm * 2. / 2.;
  double fnx = std::ceil((ex - bx) / cell), fny = std::ceil((ey - by) / cell);
  if (!(fnx >= 1. && fny >= 1. && fnx * fny < (double)s.ncell_table)) { s.grid_overflow++; return; }
  s.bbmin_x = bx; s.bbmin_y = by; s.gdx = cell;
  s.gnx = (int)fnx; s.gny = (int)fny; s.num_cell = s.gnx * s.gny;
}

// the cell table of a scenario: num_cell cells for a locked grid; max_cell_factor * N (10 N, as
// grid_gpu_green(10*n, n) of benchmarks.cpp:135) for an adaptive one, whose first geometry is
// that of the initial particles
int adaptive_setup(sim_t &s, const mfc_config &c) {
  if (c.grid_kind != MFC_GRID_GREEN && c.grid_kind != MFC_GRID_ROTHLIN) { g_err = "unknown grid_kind"; return MFC_E_INVALID; }
  s.rothlin = c.grid_kind == MFC_GRID_ROTHLIN;
  s.adaptive = c.grid_adaptive != 0;
  s.ncell_table = s.num_cell;
  if (s.adaptive) {
    int64_t factor = c.max_cell_factor > 0 ? c.max_cell_factor : 10;
    int64_t table = factor * (int64_t)s.N;
    if (table > (1LL << 30)) { g_err = "adaptive cell table too large"; return MFC_E_RANGE; }
    s.ncell_table = (int)table;
    update_geometry(s);
    if (s.grid_overflow) { g_err = "the initial adaptive grid needs more cells than max_cell_factor * N"; return MFC_E_RANGE; }
  }
  s.cell_start.assign(s.ncell_table, -1);
  s.cell_end.assign(s.ncell_table, 0);
  return MFC_OK;
}

// setup_ref_cut, generalised through mfc_config (benchmarks.cpp:255-537)
int setup(sim_t &s, const mfc_config &c) {
  s.cfg = c;
  if (c.lattice == MFC_LATTICE_RINGS || c.lattice == MFC_LATTICE_IMPACT) return setup_bodies(s, c);
  double hi_y = c.hi_y, lo_x = c.lo_x, hi_x, lo_y, dx;
  int nx, ny = c.ny;
  if (c.lattice == MFC_LATTICE_REF) {
    if (ny < 2) { g_err = "ny < 2"; return MFC_E_INVALID; }
    hi_x = c.hi_x;
    lo_y = hi_y - 3 * c.feed;
    double dy = (hi_y - lo_y) / (ny - 1);
    dx = dy;
    nx = (int)((hi_x - lo_x) / dx);
  } else {
    if (c.nx < 1 || c.ny < 1 || !(c.dx > 0)) { g_err = "bad explicit lattice"; return MFC_E_INVALID; }
    dx = c.dx; nx = c.nx;
    lo_y = hi_y - (ny - 1) * dx;
    hi_x = lo_x + nx * dx;
  }
  double hdx = c.hdx;
  s.E = c.E; s.nu = c.nu; s.rho0 = c.rho0;
  s.G = s.E / (2. * (1. + s.nu));
  s.K = 2.0 * s.G * (1 + s.nu) / (3 * (1 - 2 * s.nu));
  s.mass = dx * dx * s.rho0;
  s.jc.A = c.jc_A; s.jc.B = c.jc_B; s.jc.C = c.jc_C; s.jc.m = c.jc_m; s.jc.n = c.jc_n;
  s.jc.Tref = c.jc_Tref; s.jc.Tmelt = c.jc_Tmelt; s.jc.eps_dot_ref = c.jc_eps_dot_ref;
  s.jc.tanh = c.jc_model == MFC_JC_TANH;
  s.jc.ta = c.jc_tanh_a; s.jc.tb = c.jc_tanh_b; s.jc.tc = c.jc_tanh_c; s.jc.td = c.jc_tanh_d;
  s.cp = c.cp; s.tq = c.tq; s.alpha_wp = c.k_wp / (s.rho0 * c.cp); s.eta_heat = c.eta; s.T_init = c.T_init;
  s.alpha_tool = (c.rho_tool * c.cp_tool > 0.) ? c.k_tool / (c.rho_tool * c.cp_tool) : 0.;
  s.c_alpha = c.av_alpha; s.c_beta = c.av_beta; s.c_eta = c.av_eta; s.xspheps = c.xsph_eps; s.stresseps = c.stress_eps;
  {
    double h1 = 1. / (hdx * dx);
    double q = dx * h1;
    double fac = 10 * (M_1_PI) / 7.0 * h1 * h1;
    s.wdeltap = fac * (1 - 1.5 * q * q * (1 - 0.5 * q));
  }
  // lattice (benchmarks.cpp:334-357)
  int n = nx * ny;
  std::vector<V2> P(n);
  double min_x = FLT_MAX, max_x = -FLT_MAX, min_y = FLT_MAX, max_y = -FLT_MAX;
  uint64_t rng = c.seed;
  int it = 0;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      P[it].x = i * dx + lo_x;
      P[it].y = j * dx + lo_y;
      if (c.jitter != 0.) {
        P[it].x = P[it].x + (2.0 * u01(rng) - 1.0) * c.jitter * dx;
        P[it].y = P[it].y + (2.0 * u01(rng) - 1.0) * c.jitter * dx;
      }
      max_x = std::fmax(max_x, P[it].x); min_x = std::fmin(min_x, P[it].x);
      max_y = std::fmax(max_y, P[it].y); min_y = std::fmin(min_y, P[it].y);
      it++;
    }
  std::vector<V2> vel(n, v2(0., 0.));
  std::vector<double> rho(n), h(n), T(n), fixed(n), toolp(n);
  for (int i = 0; i < n; i++) {
    rho[i] = s.rho0;
    h[i] = hdx * dx;
    T[i] = c.T_init;
    double f = (P[i].y < min_y + (c.fixed_bottom_rows - 0.5) * dx) ? 1. : 0.;
    if (c.fix_left) f = (f || (P[i].x < min_x + dx / 2 && P[i].y < min_y + 0.5 * (hi_y - lo_y))) ? 1. : 0.;
    if (c.fix_right) f = (f || (P[i].x > max_x - dx)) ? 1. : 0.;
    fixed[i] = f;
    toolp[i] = 0.;
  }
  double vel_tool = (c.speed == 0.f) ? 83.333328 * 1e-5 : c.speed;
  // tool (benchmarks.cpp:389-421)
  s.tool.build(v2(c.tool_tl_x, c.tool_tl_y), c.tool_length, c.tool_height, c.rake_deg, c.clear_deg, c.edge_radius, c.mu);
  {
    if (s.tool.front > lo_x) {
      double distance = s.tool.front - lo_x;
      double time = distance / vel_tool;
      s.tool.vel = v2(vel_tool, 0.);
      s.tool.update(-time);
    }
    double target = c.feed;
    double current = hi_y - s.tool.low();
    double dist = std::fabs(current - target);
    double corr_time = dist / vel_tool;
    double sign = (current > target) ? 1 : -1.;
    s.tool.vel = v2(0., vel_tool);
    s.tool.update(corr_time * sign);
  }
  s.tool.vel = v2(vel_tool, 0.);
  s.contact_alpha = c.contact_alpha;
  // tool particles (benchmarks.cpp:424-459)
  int n_tool = 0;
  if (c.sample_tool) {
    std::vector<V2> smp = s.tool.sample(dx);
    double bx = FLT_MAX, by = -FLT_MAX;
    for (auto &q : smp) { bx = std::fmin(bx, q.x); by = std::fmax(by, q.y); }
    n_tool = (int)smp.size();
    for (int k = 0; k < n_tool; k++) {
      P.push_back(smp[k]);
      rho.push_back(c.rho_tool);
      h.push_back(hdx * dx);
      vel.push_back(v2(0., 0.));
      double f = 0.;
      if (smp[k].x < bx + dx / 2 || smp[k].y > by - dx / 2) f = 1.;
      fixed.push_back(f);
      T.push_back(c.T_init);
      toolp.push_back(1.);
    }
  }
  int N = n + n_tool;
  // bounding box + grid (benchmarks.cpp:464-495, grid.cu:139-158)
  V2 bbmin = v2(FLT_MAX, FLT_MAX), bbmax = v2(-FLT_MAX, -FLT_MAX);
  for (int i = 0; i < N; i++) {
    bbmin.x = std::fmin(P[i].x, bbmin.x); bbmin.y = std::fmin(P[i].y, bbmin.y);
    bbmax.x = std::fmax(P[i].x, bbmax.x); bbmax.y = std::fmax(P[i].y, bbmax.y);
  }
  bbmin.x -= 1e-8; bbmin.y -= 1e-8; bbmax.x += 1e-8; bbmax.y += 1e-8;
  double tx0, tx1, ty0, ty1;
  s.tool.safe_bb(1e-3, tx0, tx1, ty0, ty1);
  bbmin.x = std::fmin(bbmin.x, tx0);
  bbmin.y = std::fmin(bbmin.y, ty0);
  bbmax.x = std::fmax(bbmax.x, tx1) + 10 * dx;
  bbmax.y = std::fmax(bbmax.y, ty1);
  double max_height = hi_y + hi_x - lo_x;
  bbmax.y = std::fmax(bbmax.y, max_height);
  {
    double hh = hdx * dx;
    s.bbmin_x = bbmin.x - 1e-6; s.bbmin_y = bbmin.y - 1e-6;
    double bx1 = bbmax.x + 1e-6, by1 = bbmax.y + 1e-6;
    double kernel_width = 2.;
    s.gdx = hh * kernel_width / 2.;
    double lx = bx1 - s.bbmin_x, ly = by1 - s.bbmin_y;
    s.gnx = (int)(unsigned)std::ceil(lx / s.gdx);
    s.gny = (int)(unsigned)std::ceil(ly / s.gdx);
    int64_t nc = (int64_t)s.gnx * s.gny;
    if (nc <= 0 || nc > (1LL << 30)) { g_err = "grid too large"; return MFC_E_RANGE; }
    s.num_cell = (int)nc;
  }
  s.thermal_bc = s.alpha_tool != 0.;
  s.fric_heat = s.eta_heat != 0.;
  s.thermal_wp = s.alpha_wp != 0.;
  s.pse = c.thermal_solver == MFC_THERMAL_PSE;
  s.promote = c.promote_fp64 != 0;
  s.secant_tol = (c.secant_tol > 0.) ? c.secant_tol : 1e-6;
  s.dt = c.cfl * hdx * dx / (std::sqrt(s.K / s.rho0) + std::sqrt(vel_tool));
  s.t_final = 0.1 / vel_tool;
  // particle_gpu(pos, vel, rho, T, h, fixed, tool_p, N) (particle_gpu.cu:20-125)
  s.N = N; s.n_tool = n_tool;
  s.pos = P; s.vel = vel; s.rho = rho; s.h = h; s.T = T; s.fixed = fixed; s.tool_p = toolp;
  V2 z2 = v2(0., 0.); V4 z4 = {0., 0., 0., 0.};
  s.fc.assign(N, z2); s.ft.assign(N, z2); s.n.assign(N, z2); s.pos_t.assign(N, z2); s.vel_t.assign(N, z2);
  s.p.assign(N, 0.); s.blanked.assign(N, 0.); s.eps_pl.assign(N, 0.); s.eps_pl_dot.assign(N, 0.);
  s.rho_t.assign(N, 0.); s.T_t.assign(N, 0.);
  s.S.assign(N, z4); s.R.assign(N, z4); s.S_t.assign(N, z4); s.S_der.assign(N, z4); s.v_der.assign(N, z4);
  s.idx.resize(N); s.hash.assign(N, 0); s.id.resize(N);
  for (int i = 0; i < N; i++) { s.idx[i] = i; s.id[i] = i; }
  s.pos_init.assign(N, z2); s.vel_init.assign(N, z2); s.S_init.assign(N, z4);
  s.rho_init.assign(N, 0.); s.T_init_arr.assign(N, 0.);
  return adaptive_setup(s, c);
}

// setup_rings (benchmarks.cpp:155-253) and setup_impact (benchmarks.cpp:23-153): bodies with an
// initial velocity, no tool, the grid locked to the box of the reference's fixed-grid constructor
// (set_geometry, grid.cu:139-158, with the boxes of benchmarks.cpp:218 and :136)
int setup_bodies(sim_t &s, const mfc_config &c) {
  const bool rings = c.lattice == MFC_LATTICE_RINGS;
  const int nbox = c.ny;
  if (nbox < 2 || nbox > 20000) { g_err = "bad ny"; return MFC_E_INVALID; }
  if (!(c.speed > 0)) { g_err = "speed must be positive"; return MFC_E_INVALID; }
  const double hdx = c.hdx;
  double ro, dx;
  std::vector<V2> P, vel;
  if (rings) {
    double ri = 0.03;
    ro = 0.04;
    double spacing = ro + 0.001;
    dx = 2 * ro / (nbox - 1);
    for (int i = 0; i < nbox; i++)
      for (int j = 0; j < nbox; j++) {
        double px = -ro + i * dx, py = -ro + j * dx;
        double dist = std::sqrt(px * px + py * py);
        if (dist < ro && dist >= ri) {
          P.push_back(v2(px - spacing, py));
          P.push_back(v2(px + spacing, py));
        }
      }
    for (auto &q : P) vel.push_back(v2((q.x < 0) ? 200 : -200, 0.));
  } else {
    ro = 0.2;
    dx = 2 * ro / (nbox - 1);
    double spacing = ro + 0.05;
    for (int i = 0; i < nbox; i++)
      for (int j = 0; j < nbox; j++) {
        double px = -ro + i * dx, py = -ro + j * dx;
        double dist = std::sqrt(px * px + py * py);
        if (dist < ro) { P.push_back(v2(px - spacing, py)); vel.push_back(v2(1e3, 0.)); }
      }
    for (int i = 0; i < nbox / 2; i++)
      for (int j = 0; j < 10 * nbox; j++) {
        if (j * dx - 1. > 1.) continue;
        P.push_back(v2(i * dx, j * dx - 1.));
        vel.push_back(v2(0., 0.));
      }
  }
  const int N = (int)P.size();
  s.E = c.E; s.nu = c.nu; s.rho0 = c.rho0;
  s.G = s.E / (2. * (1. + s.nu));
  s.K = 2.0 * s.G * (1 + s.nu) / (3 * (1 - 2 * s.nu));
  s.mass = dx * dx * s.rho0;
  s.jc.A = c.jc_A; s.jc.B = c.jc_B; s.jc.C = c.jc_C; s.jc.m = c.jc_m; s.jc.n = c.jc_n;
  s.jc.Tref = c.jc_Tref; s.jc.Tmelt = c.jc_Tmelt; s.jc.eps_dot_ref = c.jc_eps_dot_ref;
  s.jc.tanh = c.jc_model == MFC_JC_TANH;
  s.jc.ta = c.jc_tanh_a; s.jc.tb = c.jc_tanh_b; s.jc.tc = c.jc_tanh_c; s.jc.td = c.jc_tanh_d;
  s.cp = c.cp; s.tq = c.tq; s.alpha_wp = c.k_wp / (s.rho0 * c.cp); s.eta_heat = c.eta; s.T_init = c.T_init;
  s.alpha_tool = 0.;
  s.c_alpha = c.av_alpha; s.c_beta = c.av_beta; s.c_eta = c.av_eta; s.xspheps = c.xsph_eps; s.stresseps = c.stress_eps;
  {
    double h1 = 1. / (hdx * dx);
    double q = dx * h1;
    double fac = 10 * (M_1_PI) / 7.0 * h1 * h1;
    s.wdeltap = fac * (1 - 1.5 * q * q * (1 - 0.5 * q));
  }
  s.plastic = c.elastic == 0;
  std::vector<double> fixed(N, 0.);
  if (!rings)
    for (int i = 0; i < N; i++)
      if (P[i].y < -1 + dx / 2 || P[i].y > 1 - dx / 2) fixed[i] = 1.;
  s.tool = tool_t();  // empty: no segments, no fillet, at rest
  s.contact_alpha = 0.;
  {
    double bx0 = rings ? -5 * ro : -.5, by0 = rings ? -2 * ro : -1.2;
    double bx1 = rings ? 5 * ro : 1., by1 = rings ? 2 * ro : 1.2;
    s.bbmin_x = bx0 - 1e-6; s.bbmin_y = by0 - 1e-6;
    double mx = bx1 + 1e-6, my = by1 + 1e-6;
    s.gdx = hdx * dx * 2. / 2.;
    s.gnx = (int)(unsigned)std::ceil((mx - s.bbmin_x) / s.gdx);
    s.gny = (int)(unsigned)std::ceil((my - s.bbmin_y) / s.gdx);
    int64_t nc = (int64_t)s.gnx * s.gny;
    if (nc <= 0 || nc > (1LL << 30)) { g_err = "grid too large"; return MFC_E_RANGE; }
    s.num_cell = (int)nc;
  }
  s.thermal_bc = false;
  s.fric_heat = false;
  s.thermal_wp = s.alpha_wp != 0.;
  s.pse = c.thermal_solver == MFC_THERMAL_PSE;
  s.promote = c.promote_fp64 != 0;
  s.secant_tol = (c.secant_tol > 0.) ? c.secant_tol : 1e-6;
  s.dt = c.cfl * hdx * dx / (std::sqrt(s.K / s.rho0) + std::sqrt(c.speed));
  s.t_final = (rings ? 2.5 : 2.0) * ro / c.speed;
  s.N = N; s.n_tool = 0;
  s.pos = P; s.vel = vel;
  s.rho.assign(N, s.rho0); s.h.assign(N, hdx * dx); s.T.assign(N, c.T_init); s.fixed = fixed; s.tool_p.assign(N, 0.);
  V2 z2 = v2(0., 0.); V4 z4 = {0., 0., 0., 0.};
  s.fc.assign(N, z2); s.ft.assign(N, z2); s.n.assign(N, z2); s.pos_t.assign(N, z2); s.vel_t.assign(N, z2);
  s.p.assign(N, 0.); s.blanked.assign(N, 0.); s.eps_pl.assign(N, 0.); s.eps_pl_dot.assign(N, 0.);
  s.rho_t.assign(N, 0.); s.T_t.assign(N, 0.);
  s.S.assign(N, z4); s.R.assign(N, z4); s.S_t.assign(N, z4); s.S_der.assign(N, z4); s.v_der.assign(N, z4);
  s.idx.resize(N); s.hash.assign(N, 0); s.id.resize(N);
  for (int i = 0; i < N; i++) { s.idx[i] = i; s.id[i] = i; }
  s.pos_init.assign(N, z2); s.vel_init.assign(N, z2); s.S_init.assign(N, z4);
  s.rho_init.assign(N, 0.); s.T_init_arr.assign(N, 0.);
  return adaptive_setup(s, c);
}

// cell coordinate of a particle: C++ truncation of (x - x0)/dx; values outside the int range
// (including NaN) saturate so the hash below can be clamped instead of indexing out of bounds
inline int64_t cell_coord(double f) {
  if (!(f == f)) return -1;
  if (f >= 1073741824.0) return 1073741824;
  if (f <= -1073741824.0) return -1073741824;
  return (int64_t)(int)f;
}

void assign_hashes(sim_t &s) {  // grid.cu:34-51
  int64_t oog = 0;
#pragma omp parallel for schedule(static) reduction(+ : oog)
  for (int i = 0; i < s.N; i++) {
    if (s.blanked[i] == 1.) continue;
    int64_t ix = cell_coord((s.pos[i].x - s.bbmin_x) / s.gdx);
    int64_t iy = cell_coord((s.pos[i].y - s.bbmin_y) / s.gdx);
    int64_t hsh = ix * s.gny + iy;
    // outside the grid on an axis: counted (the id itself is the reference's, clamped only
    // where it would leave the cell table)
    if (ix < 0 || ix >= s.gnx || iy < 0 || iy >= s.gny) oog++;
    if (hsh < 0 || hsh >= s.num_cell) hsh = hsh < 0 ? 0 : s.num_cell - 1;
    s.hash[i] = (int)hsh;
  }
  s.out_of_grid += oog;
}

template <class T>
void permute(std::vector<T> &a, const std::vector<int> &perm) {
  std::vector<T> b(a.size());
  for (size_t i = 0; i < a.size(); i++) b[i] = a[perm[i]];
  a.swap(b);
}

// grid_gpu_rothlin::do_get_cells (grid_gpu_rothlin.cu:28-127) on the sorted ids: the indices
// where the id changes (unique_by_key_copy) and the ids there (the map); the particles of each
// occupied cell as the difference of consecutive indices (the stencil, 0 for the last occupied
// cell); the stencil scattered to its cells and inclusively scanned into offsets[1..]; the
// offsets from the last occupied cell on set to N, offsets[0] to 0. Cell c is then
// [offsets[c], offsets[c+1]) and start/end are those (empty cells are empty ranges).
void rothlin_cells(sim_t &s) {
  const int N = s.N, ncell = s.ncell_table;
  std::vector<int> indices, map;
  for (int i = 0; i < N; i++)
    if (i == 0 || s.hash[i] != s.hash[i - 1]) { indices.push_back(i); map.push_back(s.hash[i]); }
  const int occupied = (int)indices.size();
  std::vector<int> stencil(occupied, 0), scatter(ncell, 0), offsets(ncell + 1, 0);
  for (int q = 0; q + 1 < occupied; q++) stencil[q] = std::abs(indices[q] - indices[q + 1]);
  for (int q = 0; q < occupied; q++) scatter[map[q]] = stencil[q];
  int run = 0;
  for (int c = 0; c < ncell; c++) { run += scatter[c]; offsets[c + 1] = run; }
  const int last_hash = s.hash[N - 1];
  for (int c = last_hash + 1; c <= ncell; c++) offsets[c] = N;
  offsets[0] = 0;
  for (int c = 0; c < ncell; c++) { s.cell_start[c] = offsets[c]; s.cell_end[c] = offsets[c + 1]; }
}

void sort_and_cells(sim_t &s) {  // grid_gpu_green.cu:22-116,177-201 with a stable key sort
  int N = s.N;
  // stable counting sort of (hash, current position)
  std::vector<int> cnt(s.ncell_table + 1, 0);
  for (int i = 0; i < N; i++) cnt[s.hash[i] + 1]++;
  for (int c = 0; c < s.ncell_table; c++) cnt[c + 1] += cnt[c];
  std::vector<int> perm(N);
  for (int i = 0; i < N; i++) perm[cnt[s.hash[i]]++] = i;
  s.idx = perm;
  permute(s.hash, perm);
  permute(s.pos, perm); permute(s.vel, perm); permute(s.fc, perm);
  if (!s.rothlin) permute(s.ft, perm);  // grid_gpu_rothlin's sort leaves ft in place (:183-247)
  permute(s.h, perm); permute(s.rho, perm); permute(s.S, perm); permute(s.eps_pl, perm);
  permute(s.eps_pl_dot, perm); permute(s.T, perm); permute(s.fixed, perm); permute(s.blanked, perm);
  permute(s.tool_p, perm); permute(s.pos_t, perm); permute(s.vel_t, perm); permute(s.rho_t, perm);
  permute(s.S_t, perm); permute(s.T_t, perm); permute(s.id, perm);
  if (s.rothlin) {
    rothlin_cells(s);
    return;
  }
  std::fill(s.cell_start.begin(), s.cell_start.end(), -1);
  for (int i = 0; i < N; i++) {
    int hsh = s.hash[i];
    if (i == 0 || hsh != s.hash[i - 1]) {
      s.cell_start[hsh] = i;
      if (i > 0) s.cell_end[s.hash[i - 1]] = i;
    }
    if (i == N - 1) s.cell_end[hsh] = i + 1;
  }
}

void lf_init(sim_t &s) {  // leap_frog.cu:30-40
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.N; i++) {
    if (s.blanked[i] == 1.) continue;
    s.pos_init[i] = s.pos[i]; s.vel_init[i] = s.vel[i]; s.S_init[i] = s.S[i];
    s.rho_init[i] = s.rho[i]; s.T_init_arr[i] = s.T[i];
  }
}

void lf_predict(sim_t &s) {  // leap_frog.cu:42-66
  double dt = s.dt;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.N; i++) {
    if (s.blanked[i] == 1.) continue;
    s.pos[i].x = s.pos_init[i].x + 0.5 * dt * s.pos_t[i].x;
    s.pos[i].y = s.pos_init[i].y + 0.5 * dt * s.pos_t[i].y;
    s.vel[i].x = s.vel_init[i].x + 0.5 * dt * s.vel_t[i].x;
    s.vel[i].y = s.vel_init[i].y + 0.5 * dt * s.vel_t[i].y;
    s.S[i].x = s.S_init[i].x + 0.5 * dt * s.S_t[i].x;
    s.S[i].y = s.S_init[i].y + 0.5 * dt * s.S_t[i].y;
    s.S[i].z = s.S_init[i].z + 0.5 * dt * s.S_t[i].z;
    s.S[i].w = s.S_init[i].w + 0.5 * dt * s.S_t[i].w;
    s.rho[i] = s.rho_init[i] + 0.5 * dt * s.rho_t[i];
    s.T[i] = s.T_init_arr[i] + 0.5 * dt * s.T_t[i];
  }
}

void lf_correct(sim_t &s) {  // leap_frog.cu:68-92
  double dt = s.dt;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.N; i++) {
    if (s.blanked[i] == 1.) continue;
    s.pos[i].x = s.pos_init[i].x + dt * s.pos_t[i].x;
    s.pos[i].y = s.pos_init[i].y + dt * s.pos_t[i].y;
    s.vel[i].x = s.vel_init[i].x + dt * s.vel_t[i].x;
    s.vel[i].y = s.vel_init[i].y + dt * s.vel_t[i].y;
    s.S[i].x = s.S_init[i].x + dt * s.S_t[i].x;
    s.S[i].y = s.S_init[i].y + dt * s.S_t[i].y;
    s.S[i].z = s.S_init[i].z + dt * s.S_t[i].z;
    s.S[i].w = s.S_init[i].w + dt * s.S_t[i].w;
    s.rho[i] = s.rho_init[i] + dt * s.rho_t[i];
    s.T[i] = s.T_init_arr[i] + dt * s.T_t[i];
  }
}

void material_eos(sim_t &s) {  // actions_gpu.cu:41-51
  double c0 = tsqrt(s.K / s.rho0, s.promote);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.N; i++) {
    if (s.tool_p[i] != 0.) continue;
    s.p[i] = c0 * c0 * (s.rho[i] - s.rho0);
  }
}

inline V4 artificial_stress(double rhoi, double pi, V4 Si, double eps, bool promote = false) {  // actions_gpu.cu:53-97
  double sxx = Si.x, sxy = Si.y, syy = Si.z;
  sxx -= pi; syy -= pi;
  double rhoi21 = 1. / (rhoi * rhoi);
  double numer = 2. * sxy, denom = sxx - syy + 0.;
  double theta, ct, st;
  if (promote) {
    theta = 0.5 * mfcm::atan2(numer, denom);
    ct = mfcm::cos(theta); st = mfcm::sin(theta);
  } else {
    theta = 0.5 * (double)mfcm::atan2f(numer, denom);
    ct = (double)mfcm::cosf(theta); st = (double)mfcm::sinf(theta);
  }
  double ct2 = ct * ct, st2 = st * st;
  double rsxx = ct2 * sxx + 2.0 * ct * st * sxy + st2 * syy;
  double rsyy = st2 * sxx - 2.0 * ct * st * sxy + ct2 * syy;
  double rrxx = 0., rryy = 0.;
  if (rsxx > 0) rrxx = -eps * rsxx * rhoi21;
  if (rsyy > 0) rryy = -eps * rsyy * rhoi21;
  V4 R = {ct2 * rrxx + st2 * rryy, ct * st * (rrxx - rryy), st2 * rrxx + ct2 * rryy, 0.};
  return R;
}

void corrector_artificial_stress(sim_t &s) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.N; i++) {
    if (s.tool_p[i] != 0.) continue;
    s.R[i] = artificial_stress(s.rho[i], s.p[i], s.S[i], s.stresseps, s.promote);
  }
}

inline void window(const sim_t &s, int hsh, int &li, int &lj, int &hi_, int &hj) {  // interactions_gpu.cu:79-86,200-207
  int gi = hsh / s.gny;
  int gj = hsh - gi * s.gny;
  li = gi - 2 < 0 ? 0 : gi - 2;
  lj = gj - 2 < 0 ? 0 : gj - 2;
  hi_ = gi + 3 > s.gnx ? s.gnx : gi + 3;
  hj = gj + 3 > s.gny ? s.gny : gj + 3;
}

void interactions_monaghan(sim_t &s) {  // interactions_gpu.cu:146-345
  const bool brookshaw = !s.pse;
  const double mass = s.mass, K = s.K;
  const double wdeltap = s.wdeltap, alpha = s.c_alpha, beta = s.c_beta, eta = s.c_eta, eps = s.xspheps;
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < s.N; i++) {
    if (s.blanked[i] == 1.) continue;
    double tool_i = s.tool_p[i];
    if (!brookshaw && tool_i == 1.) continue;
    double thermal_alpha = 0.;
    if (brookshaw) thermal_alpha = (tool_i == 0.) ? s.alpha_wp : s.alpha_tool;
    V2 pi = s.pos[i], vi = s.vel[i];
    V4 Si = s.S[i], Ri = s.R[i];
    double hi = s.h[i], rhoi = s.rho[i], prsi = s.p[i];
    double Ti = 0.;
    if (brookshaw && thermal_alpha != 0.) Ti = s.T[i];
    double rhoi21 = 1. / (rhoi * rhoi);
    int li, lj, hi_, hj;
    window(s, s.hash[i], li, lj, hi_, hj);
    V4 vd = {0., 0., 0., 0.}, Sd = {0., 0., 0., 0.};
    V2 vt = v2(0., 0.), xt = v2(0., 0.);
    double T_lapl = 0.;
    for (int ii = li; ii < hi_; ii++)
      for (int jj = lj; jj < hj; jj++) {
        int cell = ii * s.gny + jj;
        int c_start = s.cell_start[cell], c_end = s.cell_end[cell];
        if (c_start == -1) continue;
        for (int j = c_start; j < c_end; j++) {
          if (s.blanked[j] == 1.) continue;
          V2 pj = s.pos[j], vj = s.vel[j];
          V4 Sj = s.S[j], Rj = s.R[j];
          double hj_ = s.h[j], rhoj = s.rho[j], prsj = s.p[j];
          double Tj = 0.;
          if (brookshaw && thermal_alpha != 0.) Tj = s.T[j];
          double tool_j = s.tool_p[j];
          double volj = mass / rhoj;
          double rhoj21 = 1. / (rhoj * rhoj);
          double ww[3];
          cubic_spline(pi, pj, hi, ww, s.promote);
          double w = ww[0], w_x = ww[1], w_y = ww[2];
          if (!(tool_i == 1.) && !(tool_j == 1.)) {
            vd.x += (vj.x - vi.x) * w_x * volj;
            vd.y += (vj.x - vi.x) * w_y * volj;
            vd.z += (vj.y - vi.y) * w_x * volj;
            vd.w += (vj.y - vi.y) * w_y * volj;
            double fab = w / wdeltap;
            fab *= fab;
            fab *= fab;
            double Rx = fab * (Ri.x + Rj.x), Ry = fab * (Ri.y + Rj.y), Rz = fab * (Ri.z + Rj.z);
            Sd.x += mass * ((Si.x - prsi) * rhoi21 + (Sj.x - prsj) * rhoj21 + Rx) * w_x;
            Sd.y += mass * (Si.y * rhoi21 + Sj.y * rhoj21 + Ry) * w_y;
            Sd.z += mass * (Si.y * rhoi21 + Sj.y * rhoj21 + Ry) * w_x;
            Sd.w += mass * ((Si.z - prsi) * rhoi21 + (Sj.z - prsj) * rhoj21 + Rz) * w_y;
            double xij = pi.x - pj.x, yij = pi.y - pj.y;
            double vijx = vi.x - vj.x, vijy = vi.y - vj.y;
            double vijposij = vijx * xij + vijy * yij;
            double rhoij = 0.5 * (rhoi + rhoj);
            if (vijposij < 0.) {
              double ci = tsqrt(K / rhoi, s.promote), cj = tsqrt(K / rhoj, s.promote);
              double cij = 0.5 * (ci + cj);
              double hij = 0.5 * (hi + hj_);
              double r2ij = xij * xij + yij * yij;
              double muij = (hij * vijposij) / (r2ij + eta * eta * hij * hij);
              double piij = (-alpha * cij * muij + beta * muij * muij) / rhoij;
              vt.x += -mass * piij * w_x;
              vt.y += -mass * piij * w_y;
            }
            xt.x += -eps * w * mass / rhoij * vijx;
            xt.y += -eps * w * mass / rhoij * vijy;
          }
          if (brookshaw && thermal_alpha != 0.) {
            double xij = pi.x - pj.x, yij = pi.y - pj.y;
            double rij = std::sqrt(xij * xij + yij * yij);
            if (rij > 1e-8) {
              double eijx = xij / rij, eijy = yij / rij;
              double rij1 = 1. / rij;
              T_lapl += 2.0 * (mass / rhoj) * (Ti - Tj) * rij1 * (eijx * w_x + eijy * w_y);
            }
          }
        }
      }
    s.S_der[i] = Sd; s.v_der[i] = vd; s.pos_t[i] = xt; s.vel_t[i] = vt;
    if (brookshaw && thermal_alpha != 0.) s.T_t[i] = thermal_alpha * T_lapl;
  }
}

void interactions_heat_pse(sim_t &s) {  // interactions_gpu.cu:88-144,391-421
  if (!s.thermal_wp) return;
  const double mass = s.mass;
  const int half = s.N / 2;
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < s.N; i++) {
    if (s.blanked[i] != 0.) continue;
    V2 pi = s.pos[i];
    double hi = s.h[i], Ti = s.T[i];
    int li, lj, hi_, hj;
    window(s, s.hash[i], li, lj, hi_, hj);
    double T_ti = 0.;
    double tflag = s.tool_p[i];
    if (s.cfg.pse_flag_quirk && i >= half) tflag = 0.;
    double a = (tflag == 1.) ? s.alpha_tool : s.alpha_wp;
    for (int ii = li; ii < hi_; ii++)
      for (int jj = lj; jj < hj; jj++) {
        int cell = ii * s.gny + jj;
        int c_start = s.cell_start[cell], c_end = s.cell_end[cell];
        if (c_start == -1) continue;
        for (int j = c_start; j < c_end; j++) {
          double w2 = lapl_pse(pi, s.pos[j], hi);
          T_ti += (s.T[j] - Ti) * w2 * mass / s.rho[j];
        }
      }
    s.T_t[i] = a * T_ti;
  }
}

inline V4 jaumann(double G, V4 vder, V4 Sv) {  // actions_gpu.cu:99-133 with glm::dmat3 products
  double vx_x = vder.x, vx_y = vder.y, vy_x = vder.z, vy_y = vder.w;
  // column-major matrices m[col][row]
  double ep[3][3] = {{vx_x, 0.5 * (vx_y + vy_x), 0.}, {0.5 * (vx_y + vy_x), vy_y, 0.}, {0., 0., 0.}};
  double om[3][3] = {{0., 0.5 * (vy_x - vx_y), 0.}, {0.5 * (vx_y - vy_x), 0., 0.}, {0., 0., 0.}};
  double Sm[3][3] = {{Sv.x, Sv.y, 0.}, {Sv.y, Sv.z, 0.}, {0., 0., Sv.w}};
  double I[3][3] = {{1., 0., 0.}, {0., 1., 0.}, {0., 0., 1.}};
  double tr = ep[0][0] + ep[1][1] + ep[2][2];
  double s13 = double(1. / 3.) * tr;
  double twoG = double(2.) * G;
  double omT[3][3], A[3][3], B[3][3], out[3][3];
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) omT[c][r] = om[r][c];
  // glm mat3 * mat3: Result[c][r] = A[0][r]*B[c][0] + A[1][r]*B[c][1] + A[2][r]*B[c][2]
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) {
      A[c][r] = om[0][r] * Sm[c][0] + om[1][r] * Sm[c][1] + om[2][r] * Sm[c][2];
      B[c][r] = Sm[0][r] * omT[c][0] + Sm[1][r] * omT[c][1] + Sm[2][r] * omT[c][2];
    }
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) out[c][r] = twoG * (ep[c][r] - s13 * I[c][r]) + A[c][r] + B[c][r];
  V4 o = {out[0][0], out[0][1], out[1][1], out[2][2]};
  return o;
}

void stress_rate_jaumann(sim_t &s) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.N; i++) {
    if (s.tool_p[i] == 1.) continue;
    s.S_t[i] = jaumann(s.G, s.v_der[i], s.S[i]);
  }
}

void contmech(sim_t &s) {  // actions_gpu.cu:164-224 (three kernels, each a per-particle map)
  const double mass = s.mass;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.N; i++) {
    if (s.tool_p[i] != 0.) continue;
    s.rho_t[i] = -s.rho[i] * (s.v_der[i].x + s.v_der[i].w);
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.N; i++) {
    if (s.tool_p[i] != 0.) continue;
    V4 d = s.S_der[i];
    s.vel_t[i].x += d.x + d.y + s.fc[i].x / mass + s.ft[i].x / mass;
    s.vel_t[i].y += d.z + d.w + s.fc[i].y / mass + s.ft[i].y / mass;
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.N; i++) {
    if (s.tool_p[i] != 0.) continue;
    s.pos_t[i].x = s.pos_t[i].x + s.vel[i].x;
    s.pos_t[i].y = s.pos_t[i].y + s.vel[i].y;
  }
}

void fric_heat_gen(sim_t &s) {  // actions_gpu.cu:135-162
  if (!s.fric_heat) return;
  V2 vt = s.tool.vel;
  double dt = s.dt;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.N; i++) {
    if (s.tool_p[i] != 0.) continue;
    V2 fT = s.ft[i];
    double fmag = tsqrt(fT.x * fT.x + fT.y * fT.y, s.promote);
    if (fmag == 0.) continue;
    V2 nrm = s.n[i], vp = s.vel[i];
    V2 vd = v2(vp.x - vt.x, vp.y - vt.y);
    double vdn = vd.x * nrm.x + vd.y * nrm.y;
    V2 vr = v2(vd.x - vdn, vd.y - vdn);
    double vmag = tsqrt(vr.x * vr.x + vr.y * vr.y, s.promote);
    s.T[i] += s.eta_heat * dt * fmag * vmag / (s.cp * s.mass);
  }
}

void plasticity_johnson_cook(sim_t &s, int64_t *iter_total) {  // actions_gpu.cu:226-285
  const double dt = s.dt, mu = s.G;
  int64_t itot = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : itot)
  for (int i = 0; i < s.N; i++) {
    if (s.tool_p[i] != 0.) continue;
    if (!s.plastic) continue;  // plasticity_johnson_cook returns at once (actions_gpu.cu:400)
    V4 Sv = s.S[i];
    double Sxx = Sv.x, Syy = Sv.z, Szz = Sv.w, Sxy = Sv.y;
    double norm_Strial = tsqrt(Sxx * Sxx + Syy * Syy + Szz * Szz + Sxy * Sxy, s.promote);
    double p = s.p[i];
    double cxx = Sxx - p, cyy = Syy - p, czz = Szz - p, cxy = Sxy;
    double eps_pl = s.eps_pl[i], eps_pl_dot = s.eps_pl_dot[i], T = s.T[i];
    double svm = tsqrt((cxx * cxx + cyy * cyy + czz * czz) - cxx * cyy - cxx * czz - cyy * czz + 3.0 * cxy * cxy,
                       s.promote);
    double sigma_Y = sigma_yield(s.jc, eps_pl, eps_pl_dot, T);
    if (svm < sigma_Y) { s.eps_pl_dot[i] = 0.; continue; }
    int its = 0;
    double dl = solve_secant(s.jc, std::fmax(eps_pl_dot * dt * std::sqrt(2. / 3.), 1e-8), s.secant_tol, norm_Strial,
                             eps_pl, T, dt, s.G, &its);
    itot += its + 1;
    double inc;
    if (s.promote) {
      inc = std::sqrt(2.0 / 3.0) * std::fmax(dl, 0.);
    } else {
      float sq23 = sqrtf((float)(2.0 / 3.0));
      inc = (double)(sq23 * fmaxf((float)dl, 0.f));
    }
    double eps_pl_new = eps_pl + inc;
    double eps_pl_dot_new = inc / dt;
    s.eps_pl[i] = eps_pl_new;
    s.eps_pl_dot[i] = eps_pl_dot_new;
    V4 Sn;
    Sn.x = Sxx - Sxx / norm_Strial * dl * 2. * mu;
    Sn.z = Syy - Syy / norm_Strial * dl * 2. * mu;
    Sn.w = Szz - Szz / norm_Strial * dl * 2. * mu;
    Sn.y = Sxy - Sxy / norm_Strial * dl * 2. * mu;
    s.S[i] = Sn;
    if (s.tq != 0.) {
      double deps = eps_pl_new - eps_pl;
      double sY = sigma_yield(s.jc, eps_pl_new, eps_pl_dot_new, T);
      double rho = s.rho[i];
      s.T[i] += s.tq / (s.cp * rho) * deps * sY;
    }
  }
  if (iter_total) *iter_total += itot;
}

void tool_update(sim_t &s) {  // tool_gpu.cu:275-316, tool.cpp:211-221, update_tool_particles :210-221
  s.tool.update(s.dt);
  V2 vt = s.tool.vel;
  double dt = s.dt;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.N; i++) {
    if (s.tool_p[i] == 0.) continue;
    s.pos[i].x += dt * vt.x;
    s.pos[i].y += dt * vt.y;
  }
}

void contact_forces(sim_t &s) {  // tool_gpu.cu:151-207
  const double dt = s.dt, ms = s.mass, ca = s.contact_alpha, mu = s.tool.mu;
  V2 tv = s.tool.vel;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.N; i++) {
    if (s.tool_p[i] != 0.) continue;
    V2 xs = s.pos[i], vs = s.vel[i];
    V2 xm = v2(0., 0.), n = v2(0., 0.);
    bool inside = establish_contact(s.tool, xs, xm, n);
    if (!inside) {
      s.fc[i] = v2(0., 0.); s.ft[i] = v2(0., 0.); s.n[i] = v2(0., 0.);
      continue;
    }
    double dt2 = dt * dt;
    double gN = dot(sub(xs, xm), n);
    double sc = -ms * gN;
    V2 fN = mulv(v2((sc * n.x) / dt2, (sc * n.y) / dt2), ca);
    V2 fT = v2(0., 0.);
    if (mu != 0.) {
      V2 v = sub(vs, tv);
      V2 vr = v2(v.x - v.x * n.x, v.y - v.y * n.y);
      V2 fold = s.ft[i];
      double km = ca * ms;
      V2 kde = v2((km * vr.x) / dt, (km * vr.y) / dt);
      double fy = mu * length(fN);
      V2 fstar = sub(fold, kde);
      if (length(fstar) > fy) fT = v2((fy * fstar.x) / length(fstar), (fy * fstar.y) / length(fstar));
      else fT = fstar;
    }
    s.fc[i] = fN; s.ft[i] = fT; s.n[i] = n;
  }
}

void boundary_conditions(sim_t &s) {  // actions_gpu.cu:296-311
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.N; i++) {
    if (s.tool_p[i] != 0.) continue;
    if (s.fixed[i] != 0.) {
      s.vel[i] = v2(0., 0.); s.fc[i] = v2(0., 0.); s.pos_t[i] = v2(0., 0.); s.vel_t[i] = v2(0., 0.);
    }
  }
}

void boundary_conditions_thermal(sim_t &s) {  // actions_gpu.cu:287-294
  if (!s.thermal_bc) return;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.N; i++)
    if (s.fixed[i] == 1.) s.T[i] = s.T_init;
}

inline bool naninf(double v) { return std::isnan(v) || std::isinf(v); }

void invalidate(sim_t &s) {  // actions_gpu.cu:317-349
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.N; i++) {
    if (s.blanked[i] == 1.) continue;
    bool bad = naninf(s.pos_t[i].x) || naninf(s.pos_t[i].y) || naninf(s.vel_t[i].x) || naninf(s.vel_t[i].y) ||
               naninf(s.S_t[i].x) || naninf(s.S_t[i].y) || naninf(s.S_t[i].z) || naninf(s.S_t[i].w) ||
               naninf(s.rho_t[i]) || naninf(s.T_t[i]);
    if (bad) s.blanked[i] = 1.;
  }
}

// report_force (output.cu:30-39) with the solver's fixed summation order (the reference reduces
// with thrust::reduce, whose order is unspecified): tiles of 256 particles in storage order, each reduced by halving
// strides; then lane l of 256 adds the tile results l, l + 256, l + 512, ... in ascending order
// starting from 0, and the 256 lane sums are reduced by halving strides; sum(fc) + sum(ft)
// the build's fixed summation order of report_force (thrust's reduce order is unspecified):
// groups of 64 particles in storage order, each reduced by halving strides; lane l of 256 adds
// the group results l, l + 256, ... in ascending order (groups counted up to 4 per 256 particles
// started, the missing ones adding zeros); the 256 lane sums reduced by halving strides
double tree_sum(const std::vector<V2> &a, bool y) {
  int N = (int)a.size();
  double v[64], lane[256];
  for (int k = 0; k < 256; k++) lane[k] = 0.;
  const int ngroups = 4 * ((N + 255) / 256);
  for (int g = 0; g < ngroups; g++) {
    const int g0 = 64 * g;
    for (int k = 0; k < 64; k++) v[k] = (g0 + k < N) ? (y ? a[g0 + k].y : a[g0 + k].x) : 0.;
    for (int st = 32; st > 0; st >>= 1)
      for (int k = 0; k < st; k++) v[k] = v[k] + v[k + st];
    lane[g % 256] = lane[g % 256] + v[0];
  }
  for (int st = 128; st > 0; st >>= 1)
    for (int k = 0; k < st; k++) lane[k] = lane[k] + lane[k + st];
  return lane[0];
}

void report_force(const sim_t &s, double F[2]) {
  double fcx = tree_sum(s.fc, false), fcy = tree_sum(s.fc, true);
  double ftx = tree_sum(s.ft, false), fty = tree_sum(s.ft, true);
  F[0] = fcx + ftx;
  F[1] = fcy + fty;
}

void step(sim_t &s, int64_t *plastic_iters) {  // leap_frog.cu:94-155
  if (s.adaptive) update_geometry(s);
  assign_hashes(s);
  sort_and_cells(s);
  lf_init(s);
  lf_predict(s);
  material_eos(s);
  corrector_artificial_stress(s);
  interactions_monaghan(s);
  if (s.pse) interactions_heat_pse(s);
  stress_rate_jaumann(s);
  contmech(s);
  lf_correct(s);
  fric_heat_gen(s);
  plasticity_johnson_cook(s, plastic_iters);
  tool_update(s);
  contact_forces(s);
  boundary_conditions(s);
  boundary_conditions_thermal(s);
  invalidate(s);
  double F[2];
  report_force(s, F);
  s.force_hist.push_back(F[0]);
  s.force_hist.push_back(F[1]);
  s.step++;
}

int field_comp(int f) {
  switch (f) {
    case MFC_FIELD_POS: case MFC_FIELD_VEL: case MFC_FIELD_FC: case MFC_FIELD_FT: case MFC_FIELD_N:
    case MFC_FIELD_POS_T: case MFC_FIELD_VEL_T: return 2;
    case MFC_FIELD_S: case MFC_FIELD_S_T: return 4;
    case MFC_FIELD_R: return 3;
    case MFC_FIELD_RHO: case MFC_FIELD_P: case MFC_FIELD_T: case MFC_FIELD_EPS_PL: case MFC_FIELD_EPS_PL_DOT:
    case MFC_FIELD_H: case MFC_FIELD_FLAGS: case MFC_FIELD_ID: case MFC_FIELD_HASH: case MFC_FIELD_RHO_T:
    case MFC_FIELD_T_T: return 1;
    default: return 0;
  }
}
bool field_is_int(int f) { return f == MFC_FIELD_FLAGS || f == MFC_FIELD_ID || f == MFC_FIELD_HASH; }

}  // namespace

// ====================================================================== C interface (tests)
struct oracle_sim { sim_t s; int64_t plastic_iters = 0; };

extern "C" {

const char *oracle_last_error(void) { return g_err.c_str(); }

int oracle_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

// the literal constants of setup_ref_cut (benchmarks.cpp:255-332,384-421)
int oracle_config_ref_cut(mfc_config *c, int ny, double rake, double clear, double chamfer, double speed, double feed) {
  if (!c) return MFC_E_INVALID;
  std::memset(c, 0, sizeof(*c));
  c->lattice = MFC_LATTICE_REF;
  c->ny = ny; c->nx = 0; c->dx = 0.;
  c->hi_x = 0.200; c->hi_y = 0.060; c->lo_x = 0.005;
  c->hdx = 1.5;
  c->fixed_bottom_rows = 1; c->fix_left = 1; c->fix_right = 1;
  c->E = 1.1; c->nu = 0.35; c->rho0 = 4.429998;
  c->jc_A = 0.0086200; c->jc_B = 0.0033100; c->jc_C = 0.0100000; c->jc_m = 0.8; c->jc_n = 0.34;
  c->jc_Tref = 300.; c->jc_Tmelt = 1836.0000; c->jc_eps_dot_ref = 1e-6;
  c->cp = 553 * 1e-8; c->tq = 0.9; c->k_wp = 7.1 * 1e-13; c->eta = 0.9; c->T_init = c->jc_Tref;
  c->rho_tool = 15.25; c->cp_tool = 292 * 1e-08; c->k_tool = 88 * 1e-13;
  c->av_alpha = 1.; c->av_beta = 1.; c->av_eta = 0.1; c->xsph_eps = 0.5; c->stress_eps = 0.3;
  double nudge = 0.004;
  c->tool_tl_x = -0.05 + nudge; c->tool_tl_y = 0.0986074;
  c->tool_length = -0.0086824 - -0.05;
  c->tool_height = 0.0986074 - 0.0555074;
  c->rake_deg = rake; c->clear_deg = clear; c->edge_radius = chamfer;
  c->feed = feed; c->speed = speed; c->mu = 0.35; c->contact_alpha = 0.1;
  c->sample_tool = 1;
  c->thermal_solver = MFC_THERMAL_BROOKSHAW;
  c->precision = MFC_PREC_FP64;
  c->cfl = 0.5;
  return MFC_OK;
}

int oracle_create(const mfc_config *cfg, oracle_sim **out) {
  if (!cfg || !out) { g_err = "null argument"; return MFC_E_INVALID; }
  oracle_sim *o = new oracle_sim();
  int rc = setup(o->s, *cfg);
  if (rc != MFC_OK) { delete o; return rc; }
  *out = o;
  return MFC_OK;
}

void oracle_destroy(oracle_sim *o) { delete o; }

int oracle_step(oracle_sim *o, int nsteps) {
  if (!o || nsteps < 0) return MFC_E_INVALID;
  for (int k = 0; k < nsteps; k++) step(o->s, &o->plastic_iters);
  return MFC_OK;
}

int oracle_get_info(const oracle_sim *o, mfc_info *info) {
  if (!o || !info) return MFC_E_INVALID;
  const sim_t &s = o->s;
  std::memset(info, 0, sizeof(*info));
  info->n_particles = s.N; info->n_tool = s.n_tool;
  info->grid_nx = s.gnx; info->grid_ny = s.gny; info->num_cell = s.num_cell;
  info->precision = MFC_PREC_FP64;
  info->dt = s.dt; info->t_final = s.t_final; info->mass = s.mass;
  info->bbmin_x = s.bbmin_x; info->bbmin_y = s.bbmin_y; info->cell_size = s.gdx;
  info->step = s.step;
  info->out_of_grid = s.out_of_grid;
  info->grid_overflow = s.grid_overflow;
  info->grid_adaptive = s.adaptive;
  info->cell_table = s.ncell_table;
  return MFC_OK;
}

int oracle_tool_force(const oracle_sim *o, double F[2]) {
  if (!o || !F) return MFC_E_INVALID;
  report_force(o->s, F);
  return MFC_OK;
}

int oracle_force_history(const oracle_sim *o, double *F, int64_t cap, int64_t *n) {
  if (!o || !n) return MFC_E_INVALID;
  int64_t m = (int64_t)o->s.force_hist.size() / 2;
  if (!F) { *n = m; return MFC_OK; }
  if (m > cap) m = cap;
  std::memcpy(F, o->s.force_hist.data(), sizeof(double) * 2 * m);
  *n = m;
  return MFC_OK;
}

int64_t oracle_plastic_iterations(const oracle_sim *o) { return o ? o->plastic_iters : -1; }
int64_t oracle_out_of_grid(const oracle_sim *o) { return o ? o->s.out_of_grid : -1; }

int oracle_get_field(const oracle_sim *o, int field, void *host, size_t bytes) {
  if (!o || !host) return MFC_E_INVALID;
  const sim_t &s = o->s;
  int nc = field_comp(field);
  if (nc == 0) return MFC_E_INVALID;
  size_t need = (size_t)s.N * nc * (field_is_int(field) ? 4 : 8);
  if (bytes != need) { g_err = "size mismatch"; return MFC_E_RANGE; }
  double *d = (double *)host;
  int32_t *q = (int32_t *)host;
  for (int i = 0; i < s.N; i++) {
    switch (field) {
      case MFC_FIELD_POS: d[2 * i] = s.pos[i].x; d[2 * i + 1] = s.pos[i].y; break;
      case MFC_FIELD_VEL: d[2 * i] = s.vel[i].x; d[2 * i + 1] = s.vel[i].y; break;
      case MFC_FIELD_FC: d[2 * i] = s.fc[i].x; d[2 * i + 1] = s.fc[i].y; break;
      case MFC_FIELD_FT: d[2 * i] = s.ft[i].x; d[2 * i + 1] = s.ft[i].y; break;
      case MFC_FIELD_N: d[2 * i] = s.n[i].x; d[2 * i + 1] = s.n[i].y; break;
      case MFC_FIELD_POS_T: d[2 * i] = s.pos_t[i].x; d[2 * i + 1] = s.pos_t[i].y; break;
      case MFC_FIELD_VEL_T: d[2 * i] = s.vel_t[i].x; d[2 * i + 1] = s.vel_t[i].y; break;
      case MFC_FIELD_S: d[4 * i] = s.S[i].x; d[4 * i + 1] = s.S[i].y; d[4 * i + 2] = s.S[i].z; d[4 * i + 3] = s.S[i].w; break;
      case MFC_FIELD_S_T: d[4 * i] = s.S_t[i].x; d[4 * i + 1] = s.S_t[i].y; d[4 * i + 2] = s.S_t[i].z; d[4 * i + 3] = s.S_t[i].w; break;
      case MFC_FIELD_R: d[3 * i] = s.R[i].x; d[3 * i + 1] = s.R[i].y; d[3 * i + 2] = s.R[i].z; break;
      case MFC_FIELD_RHO: d[i] = s.rho[i]; break;
      case MFC_FIELD_P: d[i] = s.p[i]; break;
      case MFC_FIELD_T: d[i] = s.T[i]; break;
      case MFC_FIELD_EPS_PL: d[i] = s.eps_pl[i]; break;
      case MFC_FIELD_EPS_PL_DOT: d[i] = s.eps_pl_dot[i]; break;
      case MFC_FIELD_H: d[i] = s.h[i]; break;
      case MFC_FIELD_RHO_T: d[i] = s.rho_t[i]; break;
      case MFC_FIELD_T_T: d[i] = s.T_t[i]; break;
      case MFC_FIELD_FLAGS:
        q[i] = (s.fixed[i] != 0. ? MFC_FLAG_FIXED : 0) | (s.blanked[i] != 0. ? MFC_FLAG_BLANKED : 0) |
               (s.tool_p[i] != 0. ? MFC_FLAG_TOOL : 0);
        break;
      case MFC_FIELD_ID: q[i] = s.id[i]; break;
      case MFC_FIELD_HASH: q[i] = s.hash[i]; break;
    }
  }
  return MFC_OK;
}

int oracle_set_field(oracle_sim *o, int field, const void *host, size_t bytes) {
  if (!o || !host) return MFC_E_INVALID;
  sim_t &s = o->s;
  int nc = field_comp(field);
  if (nc == 0 || field == MFC_FIELD_ID || field == MFC_FIELD_HASH) return MFC_E_INVALID;
  size_t need = (size_t)s.N * nc * (field_is_int(field) ? 4 : 8);
  if (bytes != need) return MFC_E_RANGE;
  const double *d = (const double *)host;
  const int32_t *q = (const int32_t *)host;
  for (int i = 0; i < s.N; i++) {
    switch (field) {
      case MFC_FIELD_POS: s.pos[i] = v2(d[2 * i], d[2 * i + 1]); break;
      case MFC_FIELD_VEL: s.vel[i] = v2(d[2 * i], d[2 * i + 1]); break;
      case MFC_FIELD_FC: s.fc[i] = v2(d[2 * i], d[2 * i + 1]); break;
      case MFC_FIELD_FT: s.ft[i] = v2(d[2 * i], d[2 * i + 1]); break;
      case MFC_FIELD_N: s.n[i] = v2(d[2 * i], d[2 * i + 1]); break;
      case MFC_FIELD_POS_T: s.pos_t[i] = v2(d[2 * i], d[2 * i + 1]); break;
      case MFC_FIELD_VEL_T: s.vel_t[i] = v2(d[2 * i], d[2 * i + 1]); break;
      case MFC_FIELD_S: s.S[i] = V4{d[4 * i], d[4 * i + 1], d[4 * i + 2], d[4 * i + 3]}; break;
      case MFC_FIELD_S_T: s.S_t[i] = V4{d[4 * i], d[4 * i + 1], d[4 * i + 2], d[4 * i + 3]}; break;
      case MFC_FIELD_R: s.R[i] = V4{d[3 * i], d[3 * i + 1], d[3 * i + 2], 0.}; break;
      case MFC_FIELD_RHO: s.rho[i] = d[i]; break;
      case MFC_FIELD_P: s.p[i] = d[i]; break;
      case MFC_FIELD_T: s.T[i] = d[i]; break;
      case MFC_FIELD_EPS_PL: s.eps_pl[i] = d[i]; break;
      case MFC_FIELD_EPS_PL_DOT: s.eps_pl_dot[i] = d[i]; break;
      case MFC_FIELD_H: s.h[i] = d[i]; break;
      case MFC_FIELD_RHO_T: s.rho_t[i] = d[i]; break;
      case MFC_FIELD_T_T: s.T_t[i] = d[i]; break;
      case MFC_FIELD_FLAGS:
        s.fixed[i] = (q[i] & MFC_FLAG_FIXED) ? 1. : 0.;
        s.blanked[i] = (q[i] & MFC_FLAG_BLANKED) ? 1. : 0.;
        s.tool_p[i] = (q[i] & MFC_FLAG_TOOL) ? 1. : 0.;
        break;
    }
  }
  return MFC_OK;
}

// cell ranges as offsets (empty cells have offsets[c] == offsets[c+1])
int oracle_get_cells(const oracle_sim *o, int32_t *off, size_t count) {
  if (!o || !off) return MFC_E_INVALID;
  const sim_t &s = o->s;
  if (count != (size_t)s.num_cell + 1) return MFC_E_RANGE;
  // offsets from start/end of occupied cells; empty cells take the next occupied start (cells
  // of the table past num_cell hold only blanked particles of an adaptive grid, sorted last)
  int next = s.N;
  for (int c = s.ncell_table - 1; c >= 0; c--) {
    if (s.cell_start[c] != -1) next = s.cell_start[c];
    if (c <= s.num_cell) off[c] = next;
  }
  if (s.ncell_table == s.num_cell) off[s.num_cell] = s.N;
  return MFC_OK;
}

// tool_wear::eval_usui (tool_wear.cu:20-60): the reference's sequential host loop over the
// particles in storage order; out = {min, max, mean, count} of the wear rate over the particles
// in contact (min and mean 0 when none). The accumulation dt * mean is the caller's.
int oracle_tool_wear(const oracle_sim *o, double K, double alpha, const double tool_vel[2], double out[4]) {
  if (!o || !tool_vel || !out) return MFC_E_INVALID;
  const sim_t &s = o->s;
  double umin = FLT_MAX, umax = 0., total = 0.;
  int n = 0;
  for (int i = 0; i < s.N; i++) {
    const bool in_contact = fabs(s.fc[i].x) > 0. || fabs(s.fc[i].y) > 0. || fabs(s.ft[i].x) > 0. ||
                            fabs(s.fc[i].y) > 0.;  // tool_wear.cu:36, fc.y tested twice
    if (!in_contact) continue;
    n++;
    const double vx = s.vel[i].x - tool_vel[0], vy = s.vel[i].y - tool_vel[1];
    const double rx = vx - vx * s.n[i].x, ry = vy - vy * s.n[i].y;  // v - v*n, component-wise
    const double vmag = sqrt(rx * rx + ry * ry);
    const double fmag = sqrt(s.fc[i].x * s.fc[i].x + s.fc[i].y * s.fc[i].y);
    const double A = 2 * s.h[i] * M_PI;
    const double p = fmag / A;
    const double w = K * mfcm::exp(alpha / s.T[i]) * p * vmag;
    total += w;
    umin = fmin(umin, w);
    umax = fmax(umax, w);
  }
  out[0] = n == 0 ? 0. : umin;
  out[1] = umax;
  out[2] = n == 0 ? 0. : total / n;
  out[3] = n;
  return MFC_OK;
}

int oracle_get_tool(const oracle_sim *o, mfc_tool_state *t) {
  if (!o || !t) return MFC_E_INVALID;
  const tool_t &T = o->s.tool;
  std::memset(t, 0, sizeof(*t));
  t->num_seg = (int)T.seg.size();
  for (int i = 0; i < t->num_seg; i++) {
    t->left[i][0] = T.seg[i].left.x; t->left[i][1] = T.seg[i].left.y;
    t->right[i][0] = T.seg[i].right.x; t->right[i][1] = T.seg[i].right.y;
    t->n[i][0] = T.seg[i].n.x; t->n[i][1] = T.seg[i].n.y;
    t->line_a[i] = T.seg[i].l.a; t->line_b[i] = T.seg[i].l.b; t->vertical[i] = T.seg[i].l.vertical ? 1 : 0;
  }
  t->has_fillet = T.has_fillet ? 1 : 0;
  t->fillet_p[0] = T.fillet.p.x; t->fillet_p[1] = T.fillet.p.y;
  t->fillet_r = T.fillet.r; t->fillet_t1 = T.fillet.t1; t->fillet_t2 = T.fillet.t2;
  t->vel[0] = T.vel.x; t->vel[1] = T.vel.y;
  return MFC_OK;
}

// ------------------------------------------------------------------ unit-level entry points
void oracle_cubic_spline(double xi, double yi, double xj, double yj, double h, double out[3]) {
  cubic_spline(v2(xi, yi), v2(xj, yj), h, out);
}
double oracle_lapl_pse(double xi, double yi, double xj, double yj, double h) { return lapl_pse(v2(xi, yi), v2(xj, yj), h); }

static joco_t mkjc(const double jc[8]) {
  joco_t j;
  j.A = jc[0]; j.B = jc[1]; j.C = jc[2]; j.n = jc[3]; j.m = jc[4]; j.Tmelt = jc[5]; j.Tref = jc[6]; j.eps_dot_ref = jc[7];
  return j;
}
// jc = {A, B, C, n, m, Tmelt, Tref, eps_dot_ref}
// jc: A, B, C, n, m, Tmelt, Tref, eps_dot_ref, tanh_a, tanh_b, tanh_c, tanh_d
double oracle_sigma_yield_tanh(const double jc[12], double eps_pl, double eps_pl_dot, double T) {
  joco_t j = mkjc(jc);
  j.ta = jc[8]; j.tb = jc[9]; j.tc = jc[10]; j.td = jc[11];
  return sigma_yield_tanh(j, eps_pl, eps_pl_dot, T);
}
double oracle_sigma_yield(const double jc[8], double eps_pl, double eps_pl_dot, double T) {
  return sigma_yield(mkjc(jc), eps_pl, eps_pl_dot, T);
}
double oracle_opt_fun(const double jc[8], double dl, double norm, double eps_pl, double T, double dt, double G) {
  return opt_fun(mkjc(jc), dl, norm, eps_pl, T, dt, G);
}
double oracle_solve_secant(const double jc[8], double init, double tol, double norm, double eps_pl, double T, double dt,
                           double G, int *iters) {
  return solve_secant(mkjc(jc), init, tol, norm, eps_pl, T, dt, G, iters);
}
// one radial return on a single particle: state = {Sxx, Sxy, Syy, Szz, p, eps_pl, eps_pl_dot, T, rho}
void oracle_radial_return(const double jc[8], double G, double dt, double tq, double cp, double state[9]) {
  sim_t s;
  s.N = 1; s.jc = mkjc(jc); s.G = G; s.dt = dt; s.tq = tq; s.cp = cp;
  s.tool_p.assign(1, 0.);
  s.S.assign(1, V4{state[0], state[1], state[2], state[3]});
  s.p.assign(1, state[4]); s.eps_pl.assign(1, state[5]); s.eps_pl_dot.assign(1, state[6]);
  s.T.assign(1, state[7]); s.rho.assign(1, state[8]);
  plasticity_johnson_cook(s, nullptr);
  state[0] = s.S[0].x; state[1] = s.S[0].y; state[2] = s.S[0].z; state[3] = s.S[0].w;
  state[5] = s.eps_pl[0]; state[6] = s.eps_pl_dot[0]; state[7] = s.T[0];
}
void oracle_artificial_stress(double rho, double p, const double S[4], double eps, double R[3]) {
  V4 r = artificial_stress(rho, p, V4{S[0], S[1], S[2], S[3]}, eps);
  R[0] = r.x; R[1] = r.y; R[2] = r.z;
}
void oracle_jaumann(double G, const double vder[4], const double S[4], double out[4]) {
  V4 o = jaumann(G, V4{vder[0], vder[1], vder[2], vder[3]}, V4{S[0], S[1], S[2], S[3]});
  out[0] = o.x; out[1] = o.y; out[2] = o.z; out[3] = o.w;
}
// contact of one point against the sim's current tool: out = {inside, fN.x, fN.y, fT.x, fT.y, n.x, n.y, cp.x, cp.y}
void oracle_contact_point(const oracle_sim *o, double x, double y, double vx, double vy, double ftx, double fty,
                          double out[9]) {
  sim_t s;
  const sim_t &r = o->s;
  s.N = 1; s.tool = r.tool; s.dt = r.dt; s.mass = r.mass; s.contact_alpha = r.contact_alpha;
  s.tool_p.assign(1, 0.);
  s.pos.assign(1, v2(x, y)); s.vel.assign(1, v2(vx, vy)); s.ft.assign(1, v2(ftx, fty));
  s.fc.assign(1, v2(0., 0.)); s.n.assign(1, v2(0., 0.));
  V2 xm = v2(0., 0.), n = v2(0., 0.);
  out[0] = establish_contact(r.tool, v2(x, y), xm, n) ? 1. : 0.;
  contact_forces(s);
  out[1] = s.fc[0].x; out[2] = s.fc[0].y; out[3] = s.ft[0].x; out[4] = s.ft[0].y;
  out[5] = s.n[0].x; out[6] = s.n[0].y; out[7] = xm.x; out[8] = xm.y;
}
// host transcendental library (same function table as mfc_device_math)
int oracle_math(int fn, const double *x, const double *y, double *out, int64_t n) {
  for (int64_t i = 0; i < n; i++) {
    double a = x[i], b = y ? y[i] : 0.;
    double r = 0.;
    switch (fn) {
      case 0: r = mfcm::exp(a); break;
      case 1: r = mfcm::log(a); break;
      case 2: r = mfcm::pow(a, b); break;
      case 3: r = mfcm::atan2(a, b); break;
      case 4: r = mfcm::sin(a); break;
      case 5: r = mfcm::cos(a); break;
      case 6: r = (double)mfcm::atan2f(a, b); break;
      case 7: r = (double)mfcm::cosf(a); break;
      case 8: r = (double)mfcm::sinf(a); break;
      case 9: r = std::sqrt(a); break;
      case 10: r = (double)sqrtf((float)a); break;
      case 11: r = a / b; break;
      case 12: r = mfcm::tanh(a); break;
      default: return MFC_E_INVALID;
    }
    out[i] = r;
  }
  return MFC_OK;
}

}  // extern "C"
