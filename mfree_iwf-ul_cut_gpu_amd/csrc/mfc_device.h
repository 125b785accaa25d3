// mfc_device.h — per-particle physics of the SPH cutting step as gfx950 device functions.
//
// Every function keeps the operand order of the reference expression it implements (cited
// per function) so that, compiled with -ffp-contract=off, the FP64 kernels reproduce the
// host oracle bit for bit. Single-precision truncations of the reference's FP64 build
// (sqrtf, fmaxf, atan2f/cosf/sinf with double arguments) are kept as such.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mfc.h"
#include "mfc_consts.h"
#include "mfc_math.h"

namespace mfc {


// constants of one simulation, passed by value to every kernel
__device__ inline double f32sqrt(double x) { return (double)sqrtf((float)x); }
// sqrtf of the reference's FP64 build, or the FP64 square root when its truncations are promoted
__device__ inline double tsqrt(const dev_consts &k, double x) { return k.promote ? sqrt(x) : f32sqrt(x); }

// Work-groups are dispatched round-robin over the 8 XCDs of MI355X; map block b so that
// each XCD owns one contiguous range of tiles (its neighbour data then stays in that XCD's
// L2). Returns -1 for padding blocks.
__device__ inline int xcd_tile(int b, int ntiles) {
  int per = (ntiles + 7) >> 3;
  int t = (b & 7) * per + (b >> 3);
  return t < ntiles ? t : -1;
}

// The interaction kernels' map: the tiles in chunks of `chunk` consecutive tiles, dealt to
// the 8 XCDs in turn (XCD x takes chunks x, x + 8, ...). The cost of a tile varies along the
// sorted order (tool regions, free surfaces, the cut zone); contiguous ranges per XCD leave the
// XCDs unevenly loaded, while chunks keep each chunk's neighbour data together in one L2 and give
// every XCD, and every CU, a mix of the domain. A chunk's window reaches two cell columns past
// its ends, whose particles are fetched into two L2s: longer chunks re-read less, shorter ones
// balance better (a launch setting, XCD_CHUNK by default). The grid must be a multiple of
// 8 * chunk blocks (chunked_blocks).
constexpr int XCD_CHUNK = 8;
constexpr int XCD_CHUNK_MAX = 64;  // MFC_XCD_CHUNK_MAX
__device__ inline int xcd_chunk_tile(int b, int ntiles, int chunk = XCD_CHUNK) {
  const int x = b & 7, k = b >> 3;
  const int t = (k / chunk) * (8 * chunk) + x * chunk + k % chunk;
  return t < ntiles ? t : -1;
}
inline int chunked_blocks(int ntiles, int chunk = XCD_CHUNK) {
  return 8 * chunk * ((ntiles + 8 * chunk - 1) / (8 * chunk));
}

// cell coordinate: C++ truncation of a double, saturated outside the int range
__device__ inline int64_t cell_coord(double f) {
  if (!(f == f)) return -1;
  if (f >= 1073741824.0) return 1073741824;
  if (f <= -1073741824.0) return -1073741824;
  return (int64_t)(int)f;
}

// cell id of a position (grid.cu:34-51). A position outside the grid on either axis is
// reported in `outside` (the reference's id then points into another column or past the cell
// table); an id outside the table is clamped into it, where the reference would index out of
// bounds.
__device__ inline int cell_of(const dev_consts &k, double x, double y, bool &outside) {
  int64_t ix = cell_coord((x - k.bbmin_x) / k.cell);
  int64_t iy = cell_coord((y - k.bbmin_y) / k.cell);
  int64_t h = ix * k.gny + iy;
  outside = ix < 0 || ix >= k.gnx || iy < 0 || iy >= k.gny;
  if (h < 0) h = 0;
  if (h >= k.num_cell) h = k.num_cell - 1;
  return (int)h;
}

// artificial stress R of Gray & Monaghan in the principal frame (actions_gpu.cu:35-97)
__device__ inline void artificial_stress(double rhoi, double pi, double Sxx, double Sxy, double Syy, double eps,
                                         double &Rxx, double &Rxy, double &Ryy, bool promote) {
  double sxx = Sxx - pi, sxy = Sxy, syy = Syy - pi;
  double rhoi21 = 1. / (rhoi * rhoi);
  double theta, ct, st;
  if (promote) {
    theta = 0.5 * mfcm::atan2(2. * sxy, sxx - syy + 0.);
    ct = mfcm::cos(theta); st = mfcm::sin(theta);
  } else {
    theta = 0.5 * (double)mfcm::atan2f(2. * sxy, sxx - syy + 0.);
    ct = (double)mfcm::cosf(theta); st = (double)mfcm::sinf(theta);
  }
  double ct2 = ct * ct, st2 = st * st;
  double rsxx = ct2 * sxx + 2.0 * ct * st * sxy + st2 * syy;
  double rsyy = st2 * sxx - 2.0 * ct * st * sxy + ct2 * syy;
  double rrxx = 0., rryy = 0.;
  if (rsxx > 0) rrxx = -eps * rsxx * rhoi21;
  if (rsyy > 0) rryy = -eps * rsyy * rhoi21;
  Rxx = ct2 * rrxx + st2 * rryy;
  Rxy = ct * st * (rrxx - rryy);
  Ryy = st2 * rrxx + ct2 * rryy;
}

// Jaumann stress rate (actions_gpu.cu:99-133). The glm::dmat3 expression
// 2G(epsdot - tr/3 I) + Omega S + S Omega^T, expanded: the products with zero entries of
// Omega, S and I only add or subtract exact zeros, which leaves every non-zero sum unchanged.
__device__ inline void jaumann(double G, double vx_x, double vx_y, double vy_x, double vy_y, double Sxx, double Sxy,
                               double Syy, double Szz, double &txx, double &txy, double &tyy, double &tzz) {
  double exy = 0.5 * (vx_y + vy_x);
  double w = 0.5 * (vy_x - vx_y);   // Omega[0][1]
  double wt = 0.5 * (vx_y - vy_x);  // Omega[1][0]
  double tr = vx_x + vy_y + 0.;
  double s13 = double(1. / 3.) * tr;
  double twoG = double(2.) * G;
  // (Omega S)[c][r] = sum_k Omega[k][r] S[c][k];  (S Omega^T)[c][r] = sum_k S[k][r] Omega[r... ]
  double a00 = 0. * Sxx + wt * Sxy + 0. * 0.;
  double b00 = Sxx * 0. + Sxy * wt + 0. * 0.;
  double a01 = w * Sxx + 0. * Sxy + 0. * 0.;
  double b01 = Sxy * 0. + Syy * wt + 0. * 0.;
  double a11 = w * Sxy + 0. * Syy + 0. * 0.;
  double b11 = Sxy * w + Syy * 0. + 0. * 0.;
  double a22 = 0. * 0. + 0. * 0. + 0. * Szz;
  double b22 = 0. * 0. + 0. * 0. + Szz * 0.;
  txx = twoG * (vx_x - s13 * 1.) + a00 + b00;
  txy = twoG * (exy - s13 * 0.) + a01 + b01;
  tyy = twoG * (vy_y - s13 * 1.) + a11 + b11;
  tzz = twoG * (0. - s13 * 1.) + a22 + b22;
}

// The temperature factor 1 - theta^m of the Johnson-Cook flow stress (plasticity.cuh:44). A
// particle's return evaluates the flow stress at one temperature many times (the yield test, two
// residuals per secant iteration, the plastic heating), so the factor is evaluated once per
// particle; it is the same expression of the same values, so the results are unchanged.
__device__ inline double jc_temperature_factor(const dev_consts &k, double t) {
  double theta = (t - k.jc_Tref) / (k.jc_Tmelt - k.jc_Tref);
  return 1.0 - mfcm::pow(theta, k.jc_m);
}

// Johnson-Cook flow stress (plasticity.cuh:30-46), TC = jc_temperature_factor(k, t)
__device__ inline double sigma_yield(const dev_consts &k, double eps_pl, double eps_pl_dot, double TC) {
  double TA = k.jc_A + k.jc_B * mfcm::pow(eps_pl, k.jc_n);
  double TB = 1.0;
  double ed = eps_pl_dot / k.jc_eps_dot_ref;
  if (ed > 1.0) TB = 1.0 + k.jc_C * mfcm::log(ed);
  else TB = mfcm::pow((1.0 + ed), k.jc_C);
  return TA * TB * TC;
}

// tanh-modified Johnson-Cook flow stress after Calamaz 2008 (plasticity.cuh:48-70),
// TC = jc_temperature_factor(k, t)
__device__ inline double sigma_yield_tanh(const dev_consts &k, double eps_pl, double eps_dot_pl, double t, double TC) {
  double TA = k.jc_A + k.jc_B * mfcm::pow(eps_pl, k.jc_n) * 1.0 / mfcm::exp(mfcm::pow(eps_pl, k.tanh_a));
  double TB = 1.0;
  double ed = eps_dot_pl / k.jc_eps_dot_ref;
  if (ed > 1.0) TB = 1.0 + k.jc_C * mfcm::log(ed);
  else TB = mfcm::pow((1.0 + ed), k.jc_C);
  double tD = 1.0 - mfcm::pow((t / k.jc_Tmelt), k.tanh_d);
  double tS = mfcm::pow(t / k.jc_Tmelt, k.tanh_b);
  double TD = tD + (1.0 - tD) * mfcm::tanh(1.0 / (mfcm::pow(eps_pl + tS, k.tanh_c)));
  return TA * TB * TC * TD;
}

// residual of the radial return (plasticity.cuh:74-83); the flow stress is the model of the
// function pointer m_s_yield (plasticity.cuh:72), which only this residual uses. The same
// expressions as sigma_yield / sigma_yield_tanh, with the constants passed in registers: the
// secant below calls it twice per iteration, and as an outlined function it keeps the step
// kernels' register budget (inlined, two interleaved residuals would cost them occupancy), while
// taking its constants from the caller instead of reloading them from the instance descriptor in
// memory at every call, which made a wave with a long return the last of the interaction launch
__device__ __attribute__((noinline)) double rr_residual_jc(double dl, double norm, double eps0, double TC, double A,
                                                          double B, double n, double C, double edr, double dt,
                                                          double G) {
  const double s23 = sqrt(2.0 / 3.0);
  double e = eps0 + s23 * fmax(dl, 0.);
  double ed = s23 * fmax(dl, 0.) / dt;
  double TA = A + B * mfcm::pow(e, n);
  double TB = 1.0;
  double es = ed / edr;
  if (es > 1.0) TB = 1.0 + C * mfcm::log(es);
  else TB = mfcm::pow((1.0 + es), C);
  double sY = TA * TB * TC;
  return norm - 2 * G * dl - s23 * sY;
}

// The same residual for the FP32C path, with the device library's pow and log: FP64 throughout
// (the return stays in double precision), but not the oracle's own pow and log, whose bits the
// FP64 path must reproduce; the library's are about half the instructions and branch-free, and
// the secant's chain of residuals sets the end of the interaction launch late in a cut.
__device__ __attribute__((noinline)) double rr_residual_jc_lib(double dl, double norm, double eps0, double TC,
                                                              double A, double B, double n, double C, double edr,
                                                              double dt, double G) {
  const double s23 = sqrt(2.0 / 3.0);
  double e = eps0 + s23 * fmax(dl, 0.);
  double ed = s23 * fmax(dl, 0.) / dt;
  double TA = A + B * pow(e, n);
  double TB = 1.0;
  double es = ed / edr;
  if (es > 1.0) TB = 1.0 + C * log(es);
  else TB = pow((1.0 + es), C);
  double sY = TA * TB * TC;
  return norm - 2 * G * dl - s23 * sY;
}

__device__ __attribute__((noinline)) double rr_residual_tanh(const dev_consts &k, double dl, double norm, double eps0,
                                                            double t, double TC) {
  const double s23 = sqrt(2.0 / 3.0);
  double e = eps0 + s23 * fmax(dl, 0.);
  double ed = s23 * fmax(dl, 0.) / k.dt;
  double sY = sigma_yield_tanh(k, e, ed, t, TC);
  return norm - 2 * k.G * dl - s23 * sY;
}

// Johnson-Cook plasticity by radial return on one particle (actions_gpu.cu:226-285), in three
// parts: the yield test (plasticity_test), the secant for dl (rr_secant_wave, for all the
// returning particles of a wave at once), and the return itself (plasticity_finish).
struct rr_input {
  double norm, TC, init;  // |S| (S_xy counted once), the temperature factor, the secant's start
};

// the yield test; false (and eps_pl_dot = 0) for an elastic step
__device__ inline bool plasticity_test(const dev_consts &k, double p, double Sxx, double Sxy, double Syy, double Szz,
                                       double eps_pl, double &eps_pl_dot, double T, rr_input &in) {
  in.norm = tsqrt(k, Sxx * Sxx + Syy * Syy + Szz * Szz + Sxy * Sxy);
  double cxx = Sxx - p, cyy = Syy - p, czz = Szz - p, cxy = Sxy;
  double svm = tsqrt(k, (cxx * cxx + cyy * cyy + czz * czz) - cxx * cyy - cxx * czz - cyy * czz + 3.0 * cxy * cxy);
  in.TC = jc_temperature_factor(k, T);
  double sY = sigma_yield(k, eps_pl, eps_pl_dot, in.TC);
  if (svm < sY) { eps_pl_dot = 0.; return false; }
  const double s23 = sqrt(2.0 / 3.0);
  in.init = fmax(eps_pl_dot * k.dt * s23, 1e-8);
  return true;
}

// the return with the secant's dl: stress, plastic strain and rate, plastic heating
__device__ inline void plasticity_finish(const dev_consts &k, double rho, double dl, const rr_input &in, double &Sxx,
                                         double &Sxy, double &Syy, double &Szz, double &eps_pl, double &eps_pl_dot,
                                         double &T) {
  const double s23 = sqrt(2.0 / 3.0);
  const double norm = in.norm;
  double inc;
  if (k.promote) inc = s23 * fmax(dl, 0.);
  else inc = (double)(sqrtf((float)(2.0 / 3.0)) * fmaxf((float)dl, 0.f));
  double e_new = eps_pl + inc;
  double ed_new = inc / k.dt;
  double mu = k.G;
  double nxx = Sxx - Sxx / norm * dl * 2. * mu;
  double nyy = Syy - Syy / norm * dl * 2. * mu;
  double nzz = Szz - Szz / norm * dl * 2. * mu;
  double nxy = Sxy - Sxy / norm * dl * 2. * mu;
  Sxx = nxx; Syy = nyy; Szz = nzz; Sxy = nxy;
  double e_old = eps_pl;
  eps_pl = e_new; eps_pl_dot = ed_new;
  if (k.tq != 0.) {
    double deps = e_new - e_old;
    double sY2 = sigma_yield(k, e_new, ed_new, in.TC);  // T is still the temperature of the return
    T += k.tq / (k.cp * rho) * deps * sY2;
  }
}

// position of the n-th set bit (from 0) of m, which has more than n set bits
__device__ inline int nth_set_bit(unsigned long long m, int n) {
  int b = 0;
#pragma unroll
  for (int s = 32; s > 0; s >>= 1)
    if (__popcll(m & ((1ull << (b + s)) - 1)) <= n) b += s;
  return b;
}

// The secant iteration of the radial return (plasticity.cuh:85-119) for every returning particle
// of a wave, two lanes per particle: the lanes 2r and 2r + 1 take the r-th returning particle of
// the pass (up to 32 per pass) and evaluate the two residuals of an iteration at once, g(dl) and
// g(1.01 dl), exchanging them by a shuffle; both then take the same secant update (dl - g/slope,
// negative iterates folded to -0.1 dl, stop when the step is below the tolerance or after 101
// iterations), so each particle gets the iterates of a one-lane secant, bit for bit, in half the
// dependent chain. Called by all 64 lanes of the wave together; need marks the lanes whose
// particle returns, with their inputs. Returns dl on those lanes, and in folded whether the last
// update took the fold branch (the reverse mode differentiates the fold's fixed point there).
// LIB (the FP32C path): plain Johnson-Cook residuals with the device library's pow and log.
template <bool LIB>
__device__ inline double rr_secant_wave(const dev_consts &k, bool need, const rr_input &in, double eps0, double t,
                                        bool &folded) {
  const int lane = (int)(threadIdx.x & 63);
  unsigned long long pending = __ballot(need);
  double out = 0.;
  folded = false;
  const double A = k.jc_A, B = k.jc_B, n = k.jc_n, C = k.jc_C, edr = k.jc_eps_dot_ref, dt = k.dt, G = k.G;
  const double tol = k.secant_tol;
  const bool tanh_model = k.jc_tanh != 0;
  while (pending) {  // the same in every lane
    const int cnt = __popcll(pending);
    const int take = cnt < 32 ? cnt : 32;
    const unsigned long long pass = take == cnt ? pending : pending & ((1ull << nth_set_bit(pending, take)) - 1);
    const int slot = lane >> 1;
    const bool active = slot < take;
    const int owner = active ? nth_set_bit(pass, slot) : lane;
    const double norm = __shfl(in.norm, owner, 64), TC = __shfl(in.TC, owner, 64);
    const double e0 = __shfl(eps0, owner, 64), tt = __shfl(t, owner, 64);
    double dl = __shfl(in.init, owner, 64);
    int fold = 0;
    if (active) {
      const bool second = lane & 1;
      int iter = 0;
      while (true) {
        const double dl1 = dl * 1.01;
        const double x = second ? dl1 : dl;
        const double gx = tanh_model ? rr_residual_tanh(k, x, norm, e0, tt, TC)
                          : LIB      ? rr_residual_jc_lib(x, norm, e0, TC, A, B, n, C, edr, dt, G)
                                     : rr_residual_jc(x, norm, e0, TC, A, B, n, C, edr, dt, G);
        const double gy = __shfl_xor(gx, 1, 64);
        const double g = second ? gy : gx, g1 = second ? gx : gy;
        const double slope = (g - g1) / (dl - dl1);
        const double dl_old = dl;
        dl = dl - g / slope;
        fold = dl < 0.;
        if (fold) dl = -0.1 * dl;
        if (fabs(dl - dl_old) < tol) break;
        if (iter > 100) break;
        iter++;
      }
    }
    // the owner of slot r reads the result of lane 2r
    const bool mine = (pass >> lane) & 1ull;
    const int src = mine ? 2 * __popcll(pass & ((1ull << lane) - 1)) : lane;
    const double got = __shfl(dl, src, 64);
    const int got_fold = __shfl(fold, src, 64);
    if (mine) {
      out = got;
      folded = got_fold != 0;
    }
    pending &= ~pass;
  }
  return out;
}

// ---------------------------------------------------------------- tool contact (tool_gpu.cu)
__device__ inline double fillet_angle(double y, double x) {  // myatan2, tool_gpu.cu:29-36
  double t = mfcm::atan2(y, x);
  return (t > 0.) ? t : t + 2 * M_PI;
}

// distance to the fillet arc, FLT_MAX outside its angular range (tool_gpu.cu:38-63)
__device__ inline double fillet_distance(const mfc_tool_state &t, double qx, double qy) {
  double Ax = t.fillet_p[0], Ay = t.fillet_p[1], r = t.fillet_r;
  double d = sqrt((qx - Ax) * (qx - Ax) + (qy - Ay) * (qy - Ay));
  double Cx = Ax + r * (qx - Ax) / d;
  double Cy = Ay + r * (qy - Ay) / d;
  double a = fillet_angle(Ay - Cy, Ax - Cx);
  bool valid = a > fmin(t.fillet_t1, t.fillet_t2) && a < fmax(t.fillet_t1, t.fillet_t2);
  if (!valid) return FLT_MAX_D;
  double ex = Cx - qx, ey = Cy - qy;
  return sqrt(ex * ex + ey * ey);
}

// penetration test and closest tool point / normal (tool_gpu.cu:65-139)
__device__ inline bool tool_contact(const mfc_tool_state &t, double qx, double qy, double &cx, double &cy, double &nx,
                                    double &ny) {
  if (t.num_seg == 0) return false;  // no tool (the scenarios without one)
  const bool fil = t.num_seg == 5;
  bool in = true;
  for (int i = 0; i < t.num_seg; i++) {
    double dx = t.left[i][0] - qx, dy = t.left[i][1] - qy;
    in = in && (dx * t.n[i][0] + dy * t.n[i][1]) < 0.;
  }
  if (fil) {
    double dx = t.fillet_p[0] - qx, dy = t.fillet_p[1] - qy;
    in = in || sqrt(dx * dx + dy * dy) < t.fillet_r;
  }
  if (!in) return false;
  double depth = FLT_MAX_D;
  for (int i = 0; i < t.num_seg; i++) {
    double d;
    if (t.vertical[i]) {
      double ddx = t.left[i][0] - qx;
      d = fabs(ddx);
      if (d < depth) {
        depth = d;
        double is = 1.0 / sqrt(ddx * ddx + 0. * 0.);
        nx = ddx * is; ny = 0. * is;
        cx = t.left[i][0]; cy = qy;
      }
    } else if (fil && i == 2) {
      d = fillet_distance(t, qx, qy);
      if (d < depth) {
        depth = d;
        double vx = qx - t.fillet_p[0], vy = qy - t.fillet_p[1];
        double is = 1.0 / sqrt(vx * vx + vy * vy);
        nx = vx * is; ny = vy * is;
        cx = t.fillet_p[0] + nx * t.fillet_r;
        cy = t.fillet_p[1] + ny * t.fillet_r;
      }
    } else {
      double a = t.line_a[i], b = t.line_b[i];
      double den = a * a + (-1.) * (-1.);
      double px = ((-1.) * ((-1.) * qx - a * qy) - a * b) / den;
      double py = (a * (-(-1.) * qx + a * qy) - (-1.) * b) / den;
      double ex = px - qx, ey = py - qy;
      d = sqrt(ex * ex + ey * ey);
      if (d < depth) { depth = d; nx = -t.n[i][0]; ny = -t.n[i][1]; cx = px; cy = py; }
    }
  }
  return true;
}

// penalty normal force and LS-DYNA style Coulomb friction with history (tool_gpu.cu:151-207);
// returns false (and zero forces) outside the tool
__device__ inline bool contact_force(const dev_consts &k, const mfc_tool_state &t, double x, double y, double vx,
                                     double vy, double &fcx, double &fcy, double &ftx, double &fty, double &nx,
                                     double &ny) {
  double cx = 0., cy = 0.;
  nx = 0.; ny = 0.;
  if (!tool_contact(t, x, y, cx, cy, nx, ny)) {
    fcx = fcy = ftx = fty = 0.; nx = ny = 0.;
    return false;
  }
  double dt2 = k.dt * k.dt;
  double gN = (x - cx) * nx + (y - cy) * ny;
  double sc = -k.mass * gN;
  fcx = ((sc * nx) / dt2) * k.contact_alpha;
  fcy = ((sc * ny) / dt2) * k.contact_alpha;
  double tx = 0., ty = 0.;
  if (k.mu != 0.) {
    double rvx = vx - k.tool_vx, rvy = vy - k.tool_vy;
    double vrx = rvx - rvx * nx, vry = rvy - rvy * ny;
    double km = k.contact_alpha * k.mass;
    double kx = (km * vrx) / k.dt, ky = (km * vry) / k.dt;
    double fy = k.mu * sqrt(fcx * fcx + fcy * fcy);
    double sx = ftx - kx, sy = fty - ky;
    double sl = sqrt(sx * sx + sy * sy);
    if (sl > fy) { tx = (fy * sx) / sl; ty = (fy * sy) / sl; }
    else { tx = sx; ty = sy; }
  }
  ftx = tx; fty = ty;
  return true;
}

// one incremental tool move by dt: end points, line coefficients, fillet centre
// (tool.cpp:211-221, geometric_primitives.cpp:70-85)
__device__ inline void tool_move(mfc_tool_state &t, double dt) {
  double sx = dt * t.vel[0], sy = dt * t.vel[1];
#pragma unroll
  for (int i = 0; i < 5; i++) {
    if (i >= t.num_seg) break;
    double lx = t.left[i][0] + sx, ly = t.left[i][1] + sy;
    double rx = t.right[i][0] + sx, ry = t.right[i][1] + sy;
    t.left[i][0] = lx; t.left[i][1] = ly; t.right[i][0] = rx; t.right[i][1] = ry;
    double det = lx * 1. - 1. * rx;
    if (fabs(det) < 1e-12) { t.vertical[i] = 1; t.line_a[i] = 1.7976931348623157e308; t.line_b[i] = lx; }
    else {
      t.vertical[i] = 0;
      t.line_a[i] = (ly * 1. - 1. * ry) / det;
      t.line_b[i] = (ry * lx - rx * ly) / det;
    }
  }
  if (t.has_fillet) { t.fillet_p[0] = t.fillet_p[0] + sx; t.fillet_p[1] = t.fillet_p[1] + sy; }
}

}  // namespace mfc
