// mfc_adjoint.h — reverse-mode derivatives (vector-Jacobian products) of the SPH cutting step.
//
// The adjoint differentiates the step with the reference's single-precision truncations
// promoted to FP64 (a smooth map; the truncated map is a staircase at the float resolution) and
// with the radial return solved to a tight tolerance, so that the converged root can be
// differentiated by the implicit-function theorem. Branches (elastic/plastic, in/out of contact,
// stick/slip, inside/outside the kernel support, approaching/receding pairs) are held fixed at
// the branch the forward step took.
//
// Host and device code share these functions: the kernels of mfc_adjoint.hip call them on the
// GPU; the tests compile them for the host and check every VJP against central differences of
// the forward companions defined next to it.
//
// Forward expressions followed (reference file:line):
//   pair terms          interactions_gpu.cu:249-326, cubic spline kernels.cuh:30-70,
//                       PSE kernels.cuh:72-89 with interactions_gpu.cu:130-143
//   record / EOS / R    actions_gpu.cu:41-97 (R of Gray & Monaghan, principal frame)
//   per-particle tail   actions_gpu.cu:99-311 (Jaumann, fric heat, continuity, momentum,
//                       advection, JC return, BCs), leap_frog.cu:42-92, tool_gpu.cu:87-207
#pragma once

#include <stdint.h>

#include "../../include/mfc.h"
#include "mfc_consts.h"
#include "mfc_math.h"

namespace mfc {
namespace adj {

MFC_HD double sq(double x) { return x * x; }

// ------------------------------------------------------------------ neighbour record
// fields of the FP64 neighbour record (16 doubles per particle, written by the predict kernel)
enum { R_X = 0, R_Y, R_FLAGS, R_H, R_U, R_V, R_RHO, R_VOL, R_AXX, R_AXY, R_AYY, R_RXX, R_RXY, R_RYY, R_C, R_T };

// neighbour sums of one particle
enum { L_VD0 = 0, L_VD1, L_VD2, L_VD3, L_SD0, L_SD1, L_SD2, L_SD3, L_VT0, L_VT1, L_XT0, L_XT1, L_TL, L_TPSE, L_N };

// cubic spline value w and F = (1/r) dW/dr with their r-derivatives, promoted (kernels.cuh:30-70,
// including the reference's (2-q)^6 value for 1 < q < 2)
MFC_HD void spline(double r, double h, double &w, double &dw, double &F, double &dF) {
  const double h1 = 1.0 / h, q = r * h1, fac = PI_FAC * h1 * h1;
  if (q > 1) {
    const double m = 2 - q, m2 = m * m;
    w = fac * 0.25 * m2 * m2 * m2;
    dw = fac * (-1.5) * m2 * m2 * m * h1;
    F = fac * (-0.75) * m2 * h1 / r;
    dF = fac * (-0.75) * h1 * (-2.0 * m * h1 / r - m2 / (r * r));
  } else {
    w = fac * (1 - 1.5 * q * q * (1 - 0.5 * q));
    dw = fac * h1 * (-3.0 * q + 2.25 * q * q);
    F = fac * (-3.0) * h1 * h1 * (1 - 0.75 * q);
    dF = fac * 2.25 * h1 * h1 * h1;
  }
}

// contribution of neighbour j to the sums of particle i (promoted pair_terms); adds to out[L_*]
MFC_HD void pair_fwd(const dev_consts &k, const double *ri, const double *rj, bool tool_i, bool tool_j, double ta,
                     double *out) {
  const double xij = ri[R_X] - rj[R_X], yij = ri[R_Y] - rj[R_Y];
  const double rr2 = xij * xij + yij * yij;
  const double hi = ri[R_H];
  if (rr2 >= 4 * hi * hi || rr2 < 1e-8) return;
  const double r = sqrt(rr2);
  double w, dw, F, dF;
  spline(r, hi, w, dw, F, dF);
  const double gx = F * xij, gy = F * yij, m = k.mass;
  if (!tool_i && !tool_j) {
    const double du = rj[R_U] - ri[R_U], dv = rj[R_V] - ri[R_V], V = rj[R_VOL];
    out[L_VD0] += du * gx * V; out[L_VD1] += du * gy * V; out[L_VD2] += dv * gx * V; out[L_VD3] += dv * gy * V;
    const double f1 = w / k.wdeltap, fab = f1 * f1 * f1 * f1;
    const double sx = ri[R_AXX] + rj[R_AXX] + fab * (ri[R_RXX] + rj[R_RXX]);
    const double sy = ri[R_AXY] + rj[R_AXY] + fab * (ri[R_RXY] + rj[R_RXY]);
    const double sz = ri[R_AYY] + rj[R_AYY] + fab * (ri[R_RYY] + rj[R_RYY]);
    out[L_SD0] += m * sx * gx; out[L_SD1] += m * sy * gy; out[L_SD2] += m * sy * gx; out[L_SD3] += m * sz * gy;
    const double vijx = -du, vijy = -dv, vdx = vijx * xij + vijy * yij;
    const double rhoij = 0.5 * (ri[R_RHO] + rj[R_RHO]);
    if (vdx < 0.) {
      const double cij = 0.5 * (ri[R_C] + rj[R_C]), hij = 0.5 * (hi + rj[R_H]);
      const double mu = hij * vdx / (rr2 + k.av_eta * k.av_eta * hij * hij);
      const double Pi = (-k.av_alpha * cij * mu + k.av_beta * mu * mu) / rhoij;
      out[L_VT0] += -m * Pi * gx; out[L_VT1] += -m * Pi * gy;
    }
    const double xs = -k.xsph_eps * w * m / rhoij;
    out[L_XT0] += xs * vijx; out[L_XT1] += xs * vijy;
  }
  if (ta != 0.) out[L_TL] += 2.0 * rj[R_VOL] * (ri[R_T] - rj[R_T]) * F;
}

// VJP of pair_fwd. lam: adjoints of i's sums. Adds d/d(record of i) to gi and d/d(record of j)
// to gj (want_i / want_j select which are needed).
MFC_HD void pair_vjp(const dev_consts &k, const double *ri, const double *rj, bool tool_i, bool tool_j, double ta,
                     const double *lam, bool want_i, bool want_j, double *gi, double *gj) {
  const double xij = ri[R_X] - rj[R_X], yij = ri[R_Y] - rj[R_Y];
  const double rr2 = xij * xij + yij * yij;
  const double hi = ri[R_H];
  if (rr2 >= 4 * hi * hi || rr2 < 1e-8) return;
  const double r = sqrt(rr2);
  double w, dw, F, dF;
  spline(r, hi, w, dw, F, dF);
  const double gx = F * xij, gy = F * yij;
  const double m = k.mass;
  double b_w = 0, b_F = 0, b_gx = 0, b_gy = 0, b_xij = 0, b_yij = 0, b_rr2 = 0;
  if (!tool_i && !tool_j) {
    const double du = rj[R_U] - ri[R_U], dv = rj[R_V] - ri[R_V];
    const double V = rj[R_VOL];
    const double L0 = lam[L_VD0], L1 = lam[L_VD1], L2 = lam[L_VD2], L3 = lam[L_VD3];
    double b_du = (L0 * gx + L1 * gy) * V;
    double b_dv = (L2 * gx + L3 * gy) * V;
    b_gx += (L0 * du + L2 * dv) * V;
    b_gy += (L1 * du + L3 * dv) * V;
    const double b_V = L0 * du * gx + L1 * du * gy + L2 * dv * gx + L3 * dv * gy;
    const double wd = k.wdeltap, f1 = w / wd, fab = f1 * f1 * f1 * f1;
    const double RX = ri[R_RXX] + rj[R_RXX], RY = ri[R_RXY] + rj[R_RXY], RZ = ri[R_RYY] + rj[R_RYY];
    const double sx = ri[R_AXX] + rj[R_AXX] + fab * RX;
    const double sy = ri[R_AXY] + rj[R_AXY] + fab * RY;
    const double sz = ri[R_AYY] + rj[R_AYY] + fab * RZ;
    const double S0 = lam[L_SD0], S1 = lam[L_SD1], S2 = lam[L_SD2], S3 = lam[L_SD3];
    const double b_sx = m * S0 * gx, b_sy = m * (S1 * gy + S2 * gx), b_sz = m * S3 * gy;
    b_gx += m * (S0 * sx + S2 * sy);
    b_gy += m * (S1 * sy + S3 * sz);
    const double b_fab = b_sx * RX + b_sy * RY + b_sz * RZ;
    b_w += b_fab * 4.0 * f1 * f1 * f1 / wd;
    if (want_i) {
      gi[R_AXX] += b_sx; gi[R_AXY] += b_sy; gi[R_AYY] += b_sz;
      gi[R_RXX] += b_sx * fab; gi[R_RXY] += b_sy * fab; gi[R_RYY] += b_sz * fab;
    }
    if (want_j) {
      gj[R_AXX] += b_sx; gj[R_AXY] += b_sy; gj[R_AYY] += b_sz;
      gj[R_RXX] += b_sx * fab; gj[R_RXY] += b_sy * fab; gj[R_RYY] += b_sz * fab;
    }
    const double vijx = -du, vijy = -dv;
    const double vdx = vijx * xij + vijy * yij;
    const double rhoij = 0.5 * (ri[R_RHO] + rj[R_RHO]);
    double b_rhoij = 0, b_vijx = 0, b_vijy = 0;
    if (vdx < 0.) {
      const double cij = 0.5 * (ri[R_C] + rj[R_C]);
      const double hij = 0.5 * (hi + rj[R_H]);
      const double D = rr2 + k.av_eta * k.av_eta * hij * hij;
      const double mu = hij * vdx / D;
      const double Pi = (-k.av_alpha * cij * mu + k.av_beta * mu * mu) / rhoij;
      const double P = -m * Pi;
      const double T0 = lam[L_VT0], T1 = lam[L_VT1];
      const double b_P = T0 * gx + T1 * gy;
      b_gx += T0 * P;
      b_gy += T1 * P;
      const double b_Pi = -m * b_P;
      const double b_cij = b_Pi * (-k.av_alpha * mu) / rhoij;
      const double b_mu = b_Pi * (-k.av_alpha * cij + 2.0 * k.av_beta * mu) / rhoij;
      b_rhoij += -b_Pi * Pi / rhoij;
      const double b_vdx = b_mu * hij / D;
      b_rr2 += -b_mu * mu / D;
      b_vijx += b_vdx * xij; b_vijy += b_vdx * yij;
      b_xij += b_vdx * vijx; b_yij += b_vdx * vijy;
      if (want_i) gi[R_C] += 0.5 * b_cij;
      if (want_j) gj[R_C] += 0.5 * b_cij;
    }
    const double xs = -k.xsph_eps * w * m / rhoij;
    const double X0 = lam[L_XT0], X1 = lam[L_XT1];
    const double b_xs = X0 * vijx + X1 * vijy;
    b_vijx += X0 * xs; b_vijy += X1 * xs;
    b_w += b_xs * (-k.xsph_eps * m / rhoij);
    b_rhoij += -b_xs * xs / rhoij;
    b_du -= b_vijx;
    b_dv -= b_vijy;
    if (want_i) { gi[R_U] -= b_du; gi[R_V] -= b_dv; gi[R_RHO] += 0.5 * b_rhoij; }
    if (want_j) { gj[R_U] += b_du; gj[R_V] += b_dv; gj[R_RHO] += 0.5 * b_rhoij; gj[R_VOL] += b_V; }
  }
  if (ta != 0.) {  // Brookshaw: 2 V_j (T_i - T_j) (e . grad W)/r = 2 V_j (T_i - T_j) F
    const double V = rj[R_VOL], dT = ri[R_T] - rj[R_T];
    const double Lt = lam[L_TL];
    b_F += Lt * 2.0 * V * dT;
    if (want_i) gi[R_T] += Lt * 2.0 * V * F;
    if (want_j) { gj[R_T] -= Lt * 2.0 * V * F; gj[R_VOL] += Lt * 2.0 * dT * F; }
  }
  b_F += b_gx * xij + b_gy * yij;
  b_xij += b_gx * F;
  b_yij += b_gy * F;
  const double b_r = b_w * dw + b_F * dF;
  b_rr2 += b_r / (2.0 * r);
  b_xij += 2.0 * xij * b_rr2;
  b_yij += 2.0 * yij * b_rr2;
  if (want_i) { gi[R_X] += b_xij; gi[R_Y] += b_yij; }
  if (want_j) { gj[R_X] -= b_xij; gj[R_Y] -= b_yij; }
}

// one term of the PSE sum of particle i (interactions_gpu.cu:130-139, no cutoff, promoted):
// (T_j - T_i) 4/(pi h^4) exp(-r^2/h^2) m / rho_j
MFC_HD double pse_fwd(const dev_consts &k, const double *ri, const double *rj) {
  const double rr2 = sq(ri[R_X] - rj[R_X]) + sq(ri[R_Y] - rj[R_Y]);
  const double h2 = ri[R_H] * ri[R_H];
  return (rj[R_T] - ri[R_T]) * (4. / (h2 * h2 * M_PI)) * mfcm::exp(-rr2 / h2) * k.mass / rj[R_RHO];
}

MFC_HD void pse_vjp(const dev_consts &k, const double *ri, const double *rj, double lam_tpse, bool want_i,
                    bool want_j, double *gi, double *gj) {
  const double xij = ri[R_X] - rj[R_X], yij = ri[R_Y] - rj[R_Y];
  const double rr2 = xij * xij + yij * yij;
  const double h2 = ri[R_H] * ri[R_H];
  const double w2 = (4. / (h2 * h2 * M_PI)) * mfcm::exp(-rr2 / h2);
  const double dT = rj[R_T] - ri[R_T];
  const double rhoj = rj[R_RHO];
  const double term = dT * w2 * k.mass / rhoj;
  const double b_dT = lam_tpse * w2 * k.mass / rhoj;
  const double b_rr2 = lam_tpse * term * (-1.0 / h2);
  if (want_i) { gi[R_T] -= b_dT; gi[R_X] += 2.0 * xij * b_rr2; gi[R_Y] += 2.0 * yij * b_rr2; }
  if (want_j) {
    gj[R_T] += b_dT; gj[R_X] -= 2.0 * xij * b_rr2; gj[R_Y] -= 2.0 * yij * b_rr2;
    gj[R_RHO] += -lam_tpse * term / rhoj;
  }
}

// ------------------------------------------------------------------ artificial stress
// R(rho, p, S) of actions_gpu.cu:53-97, promoted
MFC_HD void art_stress_fwd(double rho, double p, double Sxx, double Sxy, double Syy, double eps, double *R) {
  const double sxx = Sxx - p, syy = Syy - p, sxy = Sxy;
  const double r21 = 1. / (rho * rho);
  const double th = 0.5 * mfcm::atan2(2. * sxy, sxx - syy + 0.);
  const double ct = mfcm::cos(th), st = mfcm::sin(th);
  const double c2 = ct * ct, s2 = st * st, cs = ct * st;
  const double rsxx = c2 * sxx + 2.0 * cs * sxy + s2 * syy;
  const double rsyy = s2 * sxx - 2.0 * cs * sxy + c2 * syy;
  const double rrxx = rsxx > 0 ? -eps * rsxx * r21 : 0.0, rryy = rsyy > 0 ? -eps * rsyy * r21 : 0.0;
  R[0] = c2 * rrxx + s2 * rryy;
  R[1] = cs * (rrxx - rryy);
  R[2] = s2 * rrxx + c2 * rryy;
}

// VJP of art_stress_fwd: b_R -> adds to b_rho, b_p, b_S[3] (Sxx, Sxy, Syy)
MFC_HD void art_stress_vjp(double rho, double p, double Sxx, double Sxy, double Syy, double eps, const double *b_R,
                           double &b_rho, double &b_p, double *b_S) {
  const double sxx = Sxx - p, syy = Syy - p, sxy = Sxy;
  const double r21 = 1. / (rho * rho);
  const double a = sxx - syy + 0., b = 2. * sxy;
  const double th = 0.5 * mfcm::atan2(b, a);
  const double ct = mfcm::cos(th), st = mfcm::sin(th);
  const double c2 = ct * ct, s2 = st * st, cs = ct * st;
  const double rsxx = c2 * sxx + 2.0 * cs * sxy + s2 * syy;
  const double rsyy = s2 * sxx - 2.0 * cs * sxy + c2 * syy;
  const bool px = rsxx > 0, py = rsyy > 0;
  const double rrxx = px ? -eps * rsxx * r21 : 0.0;
  const double rryy = py ? -eps * rsyy * r21 : 0.0;
  const double bRxx = b_R[0], bRxy = b_R[1], bRyy = b_R[2];
  double b_c2 = bRxx * rrxx + bRyy * rryy;
  double b_s2 = bRxx * rryy + bRyy * rrxx;
  double b_cs = bRxy * (rrxx - rryy);
  const double b_rrxx = bRxx * c2 + bRxy * cs + bRyy * s2;
  const double b_rryy = bRxx * s2 - bRxy * cs + bRyy * c2;
  double b_rsxx = 0, b_rsyy = 0, b_r21 = 0;
  if (px) { b_rsxx += b_rrxx * (-eps * r21); b_r21 += b_rrxx * (-eps * rsxx); }
  if (py) { b_rsyy += b_rryy * (-eps * r21); b_r21 += b_rryy * (-eps * rsyy); }
  double b_sxx = b_rsxx * c2 + b_rsyy * s2;
  double b_syy = b_rsxx * s2 + b_rsyy * c2;
  double b_sxy = (b_rsxx - b_rsyy) * 2.0 * cs;
  b_c2 += b_rsxx * sxx + b_rsyy * syy;
  b_s2 += b_rsxx * syy + b_rsyy * sxx;
  b_cs += (b_rsxx - b_rsyy) * 2.0 * sxy;
  const double b_th = b_c2 * (-2.0 * cs) + b_s2 * (2.0 * cs) + b_cs * (c2 - s2);
  const double den = a * a + b * b;
  if (den > 0) {
    const double b_a = b_th * 0.5 * (-b / den), b_b = b_th * 0.5 * (a / den);
    b_sxx += b_a; b_syy -= b_a; b_sxy += 2.0 * b_b;
  }
  b_S[0] += b_sxx; b_S[1] += b_sxy; b_S[2] += b_syy;
  b_p -= b_sxx + b_syy;
  b_rho += b_r21 * (-2.0 * r21 / rho);
}

// ------------------------------------------------------------------ Johnson-Cook
struct jc_grad { double s, e, ed, T, A, B, n, C, m; };

// sigma_Y = (A + B e^n)(1 + C ln ed* or (1 + ed*)^C)(1 - theta^m) and its partials
// (plasticity.cuh:30-46)
MFC_HD jc_grad sigma_yield_grad(const dev_consts &k, double e, double ed, double t) {
  jc_grad g;
  const double th = (t - k.jc_Tref) / (k.jc_Tmelt - k.jc_Tref);
  const double en = mfcm::pow(e, k.jc_n);
  const double TA = k.jc_A + k.jc_B * en;
  const double es = ed / k.jc_eps_dot_ref;
  double TB, dTB, dTB_C;
  if (es > 1.0) { TB = 1.0 + k.jc_C * mfcm::log(es); dTB = k.jc_C / es; dTB_C = mfcm::log(es); }
  else {
    TB = mfcm::pow(1.0 + es, k.jc_C);
    dTB = k.jc_C * mfcm::pow(1.0 + es, k.jc_C - 1.0);
    dTB_C = TB * mfcm::log(1.0 + es);
  }
  const double thm = mfcm::pow(th, k.jc_m);
  const double TC = 1.0 - thm;
  g.s = TA * TB * TC;
  g.e = (e > 0) ? k.jc_B * k.jc_n * mfcm::pow(e, k.jc_n - 1.0) * TB * TC : 0.0;
  g.ed = TA * dTB / k.jc_eps_dot_ref * TC;
  g.T = (th > 0) ? TA * TB * (-k.jc_m * mfcm::pow(th, k.jc_m - 1.0)) / (k.jc_Tmelt - k.jc_Tref) : 0.0;
  g.A = TB * TC;
  g.B = en * TB * TC;
  g.n = (e > 0) ? k.jc_B * en * mfcm::log(e) * TB * TC : 0.0;
  g.C = TA * dTB_C * TC;
  g.m = (th > 0) ? -TA * TB * thm * mfcm::log(th) : 0.0;
  return g;
}

// the temperature factor 1 - theta^m of sigma_yield_grad, and its value sigma_Y alone with that
// factor given: the same expressions, for the secant's residuals, which evaluate the flow stress
// at one temperature many times
MFC_HD double jc_temperature_factor(const dev_consts &k, double t) {
  const double th = (t - k.jc_Tref) / (k.jc_Tmelt - k.jc_Tref);
  return 1.0 - mfcm::pow(th, k.jc_m);
}
MFC_HD double sigma_yield_value(const dev_consts &k, double e, double ed, double TC) {
  const double TA = k.jc_A + k.jc_B * mfcm::pow(e, k.jc_n);
  const double es = ed / k.jc_eps_dot_ref;
  double TB;
  if (es > 1.0) TB = 1.0 + k.jc_C * mfcm::log(es);
  else TB = mfcm::pow(1.0 + es, k.jc_C);
  return TA * TB * TC;
}

// ------------------------------------------------------------------ contact
// penalty contact and Coulomb friction of one particle (tool_gpu.cu:87-207, promoted) with its
// VJP: forward outputs fc, ft, n; adjoints b_fc, b_ft, b_n in; adjoints of the position, the
// velocity, the previous friction force and mu out (added).
struct contact_io {
  double fcx, fcy, ftx, fty, nx, ny;  // forward outputs
  bool inside;
};

MFC_HD contact_io contact_fwd_vjp(const dev_consts &k, const mfc_tool_state &tool, double X, double Y, double U,
                                  double V, double ftx0, double fty0, double bfcx, double bfcy, double bftx,
                                  double bfty, double bnx, double bny, double &bX, double &bY, double &bU, double &bV,
                                  double &bftx0, double &bfty0, double &bmu) {
  contact_io o = {0, 0, 0, 0, 0, 0, false};
  const bool fil = tool.num_seg == 5;
  bool in = true;
  for (int i = 0; i < 5; i++) {
    if (i >= tool.num_seg) break;
    in = in && ((tool.left[i][0] - X) * tool.n[i][0] + (tool.left[i][1] - Y) * tool.n[i][1]) < 0.;
  }
  if (fil) in = in || sqrt(sq(tool.fillet_p[0] - X) + sq(tool.fillet_p[1] - Y)) < tool.fillet_r;
  if (!in) return o;
  o.inside = true;
  double depth = FLT_MAX_D;
  int kind = -1, seg = -1;  // 0 vertical, 1 fillet, 2 line
  for (int i = 0; i < 5; i++) {
    if (i >= tool.num_seg) break;
    double d;
    if (tool.vertical[i]) {
      d = fabs(tool.left[i][0] - X);
      if (d < depth) { depth = d; kind = 0; seg = i; }
    } else if (fil && i == 2) {
      const double Ax = tool.fillet_p[0], Ay = tool.fillet_p[1], r = tool.fillet_r;
      const double dd = sqrt((X - Ax) * (X - Ax) + (Y - Ay) * (Y - Ay));
      const double Cx = Ax + r * (X - Ax) / dd, Cy = Ay + r * (Y - Ay) / dd;
      double t = mfcm::atan2(Ay - Cy, Ax - Cx);
      t = (t > 0.) ? t : t + 2 * M_PI;
      const bool valid = t > fmin(tool.fillet_t1, tool.fillet_t2) && t < fmax(tool.fillet_t1, tool.fillet_t2);
      d = valid ? sqrt(sq(Cx - X) + sq(Cy - Y)) : FLT_MAX_D;
      if (d < depth) { depth = d; kind = 1; seg = i; }
    } else {
      const double la = tool.line_a[i], lb = tool.line_b[i];
      const double den = la * la + 1.0;
      const double px = (-1.0 * (-1.0 * X - la * Y) - la * lb) / den;
      const double py = (la * (X + la * Y) + lb) / den;
      d = sqrt(sq(px - X) + sq(py - Y));
      if (d < depth) { depth = d; kind = 2; seg = i; }
    }
  }
  // penetration gN and normal n: dgN/dq = n for every primitive, dn/dq != 0 only on the fillet
  double nx = 0, ny = 0, gN = 0, dist = 1.0;
  if (kind == 0) {
    const double ddx = tool.left[seg][0] - X;
    nx = ddx / fabs(ddx); ny = 0.;
    gN = (X - tool.left[seg][0]) * nx;
  } else if (kind == 1) {
    const double vx = X - tool.fillet_p[0], vy = Y - tool.fillet_p[1];
    dist = sqrt(vx * vx + vy * vy);
    nx = vx / dist; ny = vy / dist;
    gN = dist - tool.fillet_r;
  } else if (kind == 2) {
    nx = -tool.n[seg][0]; ny = -tool.n[seg][1];
    gN = (X - tool.left[seg][0]) * nx + (Y - tool.left[seg][1]) * ny;
  }
  const double m = k.mass, dt = k.dt;
  const double Kc = -m * k.contact_alpha / (dt * dt);
  const double fNx = Kc * gN * nx, fNy = Kc * gN * ny;
  o.fcx = fNx; o.fcy = fNy; o.nx = nx; o.ny = ny;
  double bfNx = bfcx, bfNy = bfcy, bnxc = bnx, bnyc = bny;
  if (k.mu != 0.) {
    const double rvx = U - k.tool_vx, rvy = V - k.tool_vy;
    const double vrx = rvx - rvx * nx, vry = rvy - rvy * ny;
    const double km = k.contact_alpha * m / dt;
    const double fsx = ftx0 - km * vrx, fsy = fty0 - km * vry;
    const double fnm = sqrt(fNx * fNx + fNy * fNy);
    const double fy = k.mu * fnm;
    const double sl = sqrt(fsx * fsx + fsy * fsy);
    double bfsx, bfsy, bfy = 0;
    if (sl > fy) {
      o.ftx = fy * fsx / sl; o.fty = fy * fsy / sl;
      const double dot = bftx * fsx + bfty * fsy;
      bfy = dot / sl;
      bfsx = fy * (bftx / sl - dot * fsx / (sl * sl * sl));
      bfsy = fy * (bfty / sl - dot * fsy / (sl * sl * sl));
    } else {
      o.ftx = fsx; o.fty = fsy;
      bfsx = bftx; bfsy = bfty;
    }
    bmu += bfy * fnm;
    if (fnm > 0) { bfNx += bfy * k.mu * fNx / fnm; bfNy += bfy * k.mu * fNy / fnm; }
    bftx0 += bfsx; bfty0 += bfsy;
    const double bvrx = -km * bfsx, bvry = -km * bfsy;
    bU += bvrx * (1.0 - nx); bV += bvry * (1.0 - ny);
    bnxc += -bvrx * rvx; bnyc += -bvry * rvy;
  }
  const double bgN = Kc * (bfNx * nx + bfNy * ny);
  bnxc += Kc * gN * bfNx; bnyc += Kc * gN * bfNy;
  bX += bgN * nx; bY += bgN * ny;
  if (kind == 1) {
    const double dot = bnxc * nx + bnyc * ny;
    bX += (bnxc - dot * nx) / dist; bY += (bnyc - dot * ny) / dist;
  }
  return o;
}

// ------------------------------------------------------------------ per-particle tail
// state field order (= the structure-of-arrays of the step kernels)
enum { TS_X = 0, TS_Y, TS_U, TS_V, TS_SXX, TS_SXY, TS_SYY, TS_SZZ, TS_RHO, TS_T, TS_EPS, TS_EPSD, TS_H,
       TS_FCX, TS_FCY, TS_FTX, TS_FTY, TS_PTX, TS_PTY, TS_VTX, TS_VTY, TS_RHOT, TS_STXX, TS_STXY, TS_STYY,
       TS_STZZ, TS_TT, TS_N };

// inputs of the part of the step that follows the neighbour sums, for one particle in sorted
// order (leap_frog.cu:120-155 from material_stress_rate_jaumann on)
struct tail_in {
  double s[TS_N];    // sorted state at the start of the step (TS_*; TS_H unused)
  double p;          // pressure of this step
  double nx, ny;     // positional contact normal written by the previous step
  int flags;
  bool pse_flag_zero;  // PSE tool-flag quirk applies to this index
  double acc[L_N];   // neighbour sums of this step
  double dl_rec = 0.;  // the forward's secant result for this step (-dl: its last update
                       // folded); 0: run the secant here
};

struct tail_grad {
  double s[TS_N];    // adjoint of the sorted start-of-step state
  double nx, ny;     // adjoint of the positional normal of the previous step
  double acc[L_N];   // adjoint of the neighbour sums
  double theta[6];   // A, B, n, C, m, mu
};

// forward recomputation and VJP of the tail of one particle. out_bar: adjoint of the state at the
// end of the step (TS_*) and of the normal written by this step; Fbar: adjoint of this step's
// tool force (F = sum fc + sum ft). fwd (optional): the forward outputs, TS_* then nx, ny.
MFC_HD void tail_vjp(const dev_consts &k, const mfc_tool_state &tool, const tail_in &in, const double *out_bar,
                     double nbar_x, double nbar_y, double Fbar_x, double Fbar_y, tail_grad &g, double *fwd) {
  for (int q = 0; q < TS_N; q++) g.s[q] = 0;
  g.nx = 0; g.ny = 0;
  for (int q = 0; q < L_N; q++) g.acc[q] = 0;
  for (int q = 0; q < 6; q++) g.theta[q] = 0;
  const double *s = in.s;
  const int fl = in.flags;
  const bool blank = fl & MFC_FLAG_BLANKED, tool_p = fl & MFC_FLAG_TOOL, fixed = fl & MFC_FLAG_FIXED;
  const bool pse = k.pse != 0;
  const double dt = k.dt, hdt = 0.5 * dt, m = k.mass, G = k.G;
  // ---------------- forward
  double pu = s[TS_U], pv = s[TS_V], psxx = s[TS_SXX], psxy = s[TS_SXY], psyy = s[TS_SYY];
  double prho = s[TS_RHO];
  if (!blank) {
    pu = s[TS_U] + hdt * s[TS_VTX]; pv = s[TS_V] + hdt * s[TS_VTY];
    psxx = s[TS_SXX] + hdt * s[TS_STXX]; psxy = s[TS_SXY] + hdt * s[TS_STXY]; psyy = s[TS_SYY] + hdt * s[TS_STYY];
    prho = s[TS_RHO] + hdt * s[TS_RHOT];
  }
  const double ta = pse ? 0. : (tool_p ? k.alpha_tool : k.alpha_wp);
  const bool run_mech = !blank && (!pse || !tool_p);
  const bool run_pse = !blank && pse && k.thermal_wp;
  const double apse = (tool_p && !in.pse_flag_zero) ? k.alpha_tool : k.alpha_wp;
  const double *a = in.acc;
  double PTX = s[TS_PTX], PTY = s[TS_PTY], VTX = s[TS_VTX], VTY = s[TS_VTY], RHOT = s[TS_RHOT];
  double STXX = s[TS_STXX], STXY = s[TS_STXY], STYY = s[TS_STYY], STZZ = s[TS_STZZ], TT = s[TS_TT];
  if (run_mech) {
    PTX = a[L_XT0]; PTY = a[L_XT1]; VTX = a[L_VT0]; VTY = a[L_VT1];
    if (ta != 0.) TT = ta * a[L_TL];
  }
  if (run_pse) TT = apse * a[L_TPSE];
  const double vxx = a[L_VD0], vxy = a[L_VD1], vyx = a[L_VD2], vyy = a[L_VD3];
  if (!tool_p) {
    const double exy = 0.5 * (vxy + vyx), w = 0.5 * (vyx - vxy), tr = vxx + vyy;
    STXX = 2 * G * (vxx - tr / 3.) - 2.0 * w * psxy;
    STXY = 2 * G * exy + w * (psxx - psyy);
    STYY = 2 * G * (vyy - tr / 3.) + 2.0 * w * psxy;
    STZZ = 2 * G * (-tr / 3.);
    RHOT = -prho * (vxx + vyy);
    VTX = VTX + a[L_SD0] + a[L_SD1] + s[TS_FCX] / m + s[TS_FTX] / m;
    VTY = VTY + a[L_SD2] + a[L_SD3] + s[TS_FCY] / m + s[TS_FTY] / m;
    PTX = PTX + pu;
    PTY = PTY + pv;
  }
  double X = s[TS_X], Y = s[TS_Y], U = s[TS_U], V = s[TS_V], SXX = s[TS_SXX], SXY = s[TS_SXY], SYY = s[TS_SYY];
  double SZZ = s[TS_SZZ], RHO = s[TS_RHO], T1 = s[TS_T];
  if (!blank) {
    X = s[TS_X] + dt * PTX; Y = s[TS_Y] + dt * PTY; U = s[TS_U] + dt * VTX; V = s[TS_V] + dt * VTY;
    SXX = s[TS_SXX] + dt * STXX; SXY = s[TS_SXY] + dt * STXY; SYY = s[TS_SYY] + dt * STYY;
    SZZ = s[TS_SZZ] + dt * STZZ; RHO = s[TS_RHO] + dt * RHOT; T1 = s[TS_T] + dt * TT;
  }
  // friction heat with the positional normal of the previous step
  const double ftx0 = s[TS_FTX], fty0 = s[TS_FTY];
  bool heat = false;
  double fm = 0, vm = 0, rx = 0, ry = 0, Tf = T1;
  if (k.fric_heat && !tool_p) {
    fm = sqrt(ftx0 * ftx0 + fty0 * fty0);
    if (fm != 0.) {
      heat = true;
      const double dvx = U - k.tool_vx, dvy = V - k.tool_vy;
      const double dvn = dvx * in.nx + dvy * in.ny;
      rx = dvx - dvn; ry = dvy - dvn;
      vm = sqrt(rx * rx + ry * ry);
      Tf = T1 + k.eta_heat * dt * fm * vm / (k.cp * m);
    }
  }
  // radial return: root of g(dl) = norm - 2 G dl - s23 sigma_Y(e0 + s23 dl, s23 dl / dt, T)
  const double s23 = sqrt(2.0 / 3.0);
  const double e0 = s[TS_EPS];
  bool plastic = false;
  bool folded = false;  // the last secant update took the fold branch (x < 0 -> -0.1 x)
  double norm = 0, dl = 0, ratio = 1.0;
  double S2xx = SXX, S2xy = SXY, S2yy = SYY, S2zz = SZZ, E2 = e0, ED2 = s[TS_EPSD], T2 = Tf;
  jc_grad gr = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (!tool_p) {
    norm = sqrt(SXX * SXX + SYY * SYY + SZZ * SZZ + SXY * SXY);
    const double p = in.p;
    const double cxx = SXX - p, cyy = SYY - p, czz = SZZ - p, cxy = SXY;
    const double svm = sqrt((cxx * cxx + cyy * cyy + czz * czz) - cxx * cyy - cxx * czz - cyy * czz + 3.0 * cxy * cxy);
    const double TCf = jc_temperature_factor(k, Tf);
    const double sY0 = sigma_yield_value(k, e0, s[TS_EPSD], TCf);
    if (svm < sY0) {
      ED2 = 0.;
    } else {
      plastic = true;
      double x = fmax(s[TS_EPSD] * dt * s23, 1e-8), xo;
      int it = 0;
      if (in.dl_rec != 0.) {  // the iterates of the forward's own secant, ended as recorded
        x = fabs(in.dl_rec);
        folded = in.dl_rec < 0.;
      } else {
        while (true) {
          const double g0 = norm - 2 * G * x - s23 * sigma_yield_value(k, e0 + s23 * fmax(x, 0.), s23 * fmax(x, 0.) / dt, TCf);
          const double x1 = x * 1.01;
          const double g1 = norm - 2 * G * x1 - s23 * sigma_yield_value(k, e0 + s23 * fmax(x1, 0.), s23 * fmax(x1, 0.) / dt, TCf);
          const double slope = (g0 - g1) / (x - x1);
          xo = x;
          x = x - g0 / slope;
          folded = false;
          if (x < 0.) { x = -0.1 * x; folded = true; }
          if (fabs(x - xo) < k.secant_tol || it > 100) break;
          it++;
        }
      }
      dl = x;
      const double inc = s23 * fmax(dl, 0.);
      E2 = e0 + inc;
      ED2 = inc / dt;
      ratio = 1.0 - dl * 2. * G / norm;
      S2xx = SXX * ratio; S2xy = SXY * ratio; S2yy = SYY * ratio; S2zz = SZZ * ratio;
      gr = sigma_yield_grad(k, E2, ED2, Tf);
      if (k.tq != 0.) T2 = Tf + k.tq / (k.cp * RHO) * inc * gr.s;
    }
  }
  double Xo = X, Yo = Y;
  if (tool_p) { Xo += dt * k.tool_vx; Yo += dt * k.tool_vy; }
  // ---------------- reverse
  double bX = out_bar[TS_X], bY = out_bar[TS_Y], bU = out_bar[TS_U], bV = out_bar[TS_V];
  double bS2xx = out_bar[TS_SXX], bS2xy = out_bar[TS_SXY], bS2yy = out_bar[TS_SYY], bS2zz = out_bar[TS_SZZ];
  double bRHO = out_bar[TS_RHO], bT2 = out_bar[TS_T], bE2 = out_bar[TS_EPS], bED2 = out_bar[TS_EPSD];
  double bFCX = out_bar[TS_FCX] + Fbar_x, bFCY = out_bar[TS_FCY] + Fbar_y;
  double bFTX = out_bar[TS_FTX] + Fbar_x, bFTY = out_bar[TS_FTY] + Fbar_y;
  double bPTX = out_bar[TS_PTX], bPTY = out_bar[TS_PTY], bVTX = out_bar[TS_VTX], bVTY = out_bar[TS_VTY];
  double bRHOT = out_bar[TS_RHOT], bSTXX = out_bar[TS_STXX], bSTXY = out_bar[TS_STXY];
  double bSTYY = out_bar[TS_STYY], bSTZZ = out_bar[TS_STZZ], bTT = out_bar[TS_TT];
  // thermal and mechanical boundary conditions overwrite their targets
  if (k.thermal_bc && fixed) bT2 = 0.;
  if (!tool_p && fixed) { bU = 0; bV = 0; bFCX = 0; bFCY = 0; bPTX = 0; bPTY = 0; bVTX = 0; bVTY = 0; }
  double bftx0 = 0, bfty0 = 0, bfcx0 = 0, bfcy0 = 0;
  double FCX = s[TS_FCX], FCY = s[TS_FCY], FTX = s[TS_FTX], FTY = s[TS_FTY], NX = 0, NY = 0;
  if (tool_p) {  // tool particles carry their forces and the positional normal through unchanged
    bfcx0 += bFCX; bfcy0 += bFCY; bftx0 += bFTX; bfty0 += bFTY;
    g.nx += nbar_x; g.ny += nbar_y;
  } else {
    contact_io c = contact_fwd_vjp(k, tool, X, Y, U, V, ftx0, fty0, bFCX, bFCY, bFTX, bFTY, nbar_x, nbar_y, bX, bY,
                                   bU, bV, bftx0, bfty0, g.theta[5]);
    FCX = c.fcx; FCY = c.fcy; FTX = c.ftx; FTY = c.fty; NX = c.nx; NY = c.ny;
  }
  // radial return
  double bSXX, bSXY, bSYY, bSZZ, bTf = bT2, be0 = 0, bed0 = 0, bRHO1 = bRHO;
  if (tool_p) {
    bSXX = bS2xx; bSXY = bS2xy; bSYY = bS2yy; bSZZ = bS2zz; be0 = bE2; bed0 = bED2;
  } else if (!plastic) {
    bSXX = bS2xx; bSXY = bS2xy; bSYY = bS2yy; bSZZ = bS2zz; be0 = bE2;
  } else {
    double bdl = 0, bnorm = 0;
    const double bratio = bS2xx * SXX + bS2xy * SXY + bS2yy * SYY + bS2zz * SZZ;
    bSXX = bS2xx * ratio; bSXY = bS2xy * ratio; bSYY = bS2yy * ratio; bSZZ = bS2zz * ratio;
    bdl += bratio * (-2. * G / norm);
    bnorm += bratio * (dl * 2. * G / (norm * norm));
    bdl += bE2 * s23 + bED2 * s23 / dt;
    be0 += bE2;
    if (k.tq != 0.) {
      const double cT = k.tq / (k.cp * RHO), inc = s23 * dl;
      bdl += bT2 * cT * s23 * gr.s;
      bRHO1 += bT2 * (-cT / RHO) * inc * gr.s;
      const double bs = bT2 * cT * inc;  // adjoint of sigma_Y(E2, ED2, Tf)
      be0 += bs * gr.e;
      bdl += bs * (gr.e * s23 + gr.ed * s23 / dt);
      bTf += bs * gr.T;
      g.theta[0] += bs * gr.A; g.theta[1] += bs * gr.B; g.theta[2] += bs * gr.n;
      g.theta[3] += bs * gr.C; g.theta[4] += bs * gr.m;
    }
    const double dgd = -2. * G - s23 * (gr.e * s23 + gr.ed * s23 / dt);
    if (!folded) {
      // implicit function theorem on g(dl; norm, e0, Tf, theta) = 0
      const double mlt = -bdl / dgd;
      bnorm += mlt;
      be0 += mlt * (-s23 * gr.e);
      bTf += mlt * (-s23 * gr.T);
      g.theta[0] += mlt * (-s23 * gr.A); g.theta[1] += mlt * (-s23 * gr.B); g.theta[2] += mlt * (-s23 * gr.n);
      g.theta[3] += mlt * (-s23 * gr.C); g.theta[4] += mlt * (-s23 * gr.m);
    } else {
      // g has no root (g(0) < 0 can happen: the return's norm counts S_xy once while the yield
      // test uses the von Mises stress), and the secant stops at the fixed point of its fold:
      // x = -0.1 (x - g(x)/s(x)) with s(x) = (g(x) - g(1.01 x)) / (-0.01 x), that is
      // H(x) = 1101 g(x) - 1100 g(1.01 x) = 0; implicit function theorem on H
      const double x1 = 1.01 * dl;
      const jc_grad g1 = sigma_yield_grad(k, e0 + s23 * x1, s23 * x1 / dt, Tf);
      const double dgd1 = -2. * G - s23 * (g1.e * s23 + g1.ed * s23 / dt);
      const double dH = 1101. * dgd - 1100. * 1.01 * dgd1;
      const double mlt = -bdl / dH;
      bnorm += mlt * (1101. - 1100.);
      be0 += mlt * (-s23) * (1101. * gr.e - 1100. * g1.e);
      bTf += mlt * (-s23) * (1101. * gr.T - 1100. * g1.T);
      g.theta[0] += mlt * (-s23) * (1101. * gr.A - 1100. * g1.A);
      g.theta[1] += mlt * (-s23) * (1101. * gr.B - 1100. * g1.B);
      g.theta[2] += mlt * (-s23) * (1101. * gr.n - 1100. * g1.n);
      g.theta[3] += mlt * (-s23) * (1101. * gr.C - 1100. * g1.C);
      g.theta[4] += mlt * (-s23) * (1101. * gr.m - 1100. * g1.m);
    }
    bSXX += bnorm * SXX / norm; bSXY += bnorm * SXY / norm; bSYY += bnorm * SYY / norm; bSZZ += bnorm * SZZ / norm;
  }
  // friction heat
  double bT1 = bTf;
  if (heat) {
    const double c = k.eta_heat * dt / (k.cp * m);
    const double bfm = bTf * c * vm, bvm = bTf * c * fm;
    bftx0 += bfm * ftx0 / fm; bfty0 += bfm * fty0 / fm;
    if (vm > 0) {
      const double brx = bvm * rx / vm, bry = bvm * ry / vm;
      const double bdvn = -(brx + bry);  // r = dv - (dv . n) (1, 1)
      bU += brx + bdvn * in.nx;
      bV += bry + bdvn * in.ny;
      g.nx += bdvn * (U - k.tool_vx);
      g.ny += bdvn * (V - k.tool_vy);
    }
  }
  // leap-frog corrector
  g.s[TS_X] += bX; g.s[TS_Y] += bY; g.s[TS_U] += bU; g.s[TS_V] += bV;
  g.s[TS_SXX] += bSXX; g.s[TS_SXY] += bSXY; g.s[TS_SYY] += bSYY; g.s[TS_SZZ] += bSZZ;
  g.s[TS_RHO] += bRHO1; g.s[TS_T] += bT1;
  if (!blank) {
    bPTX += dt * bX; bPTY += dt * bY; bVTX += dt * bU; bVTY += dt * bV;
    bSTXX += dt * bSXX; bSTXY += dt * bSXY; bSTYY += dt * bSYY; bSTZZ += dt * bSZZ;
    bRHOT += dt * bRHO1; bTT += dt * bT1;
  }
  g.s[TS_EPS] += be0;
  g.s[TS_EPSD] += bed0;
  // Jaumann, continuity, momentum, advection on the predicted state
  double bpu = 0, bpv = 0, bpsxx = 0, bpsxy = 0, bpsyy = 0, bprho = 0;
  double bRHOTi = 0, bSTXXi = 0, bSTXYi = 0, bSTYYi = 0, bSTZZi = 0;
  if (!tool_p) {
    bpu += bPTX; bpv += bPTY;
    g.acc[L_SD0] += bVTX; g.acc[L_SD1] += bVTX; g.acc[L_SD2] += bVTY; g.acc[L_SD3] += bVTY;
    bfcx0 += bVTX / m; bftx0 += bVTX / m; bfcy0 += bVTY / m; bfty0 += bVTY / m;
    bprho += -bRHOT * (vxx + vyy);
    double bvxx = -bRHOT * prho, bvyy = -bRHOT * prho;
    const double w = 0.5 * (vyx - vxy);
    double bw = 0, bexy = 0, btr = 0;
    bvxx += bSTXX * 2 * G; btr += bSTXX * (-2 * G / 3.); bw += bSTXX * (-2.0 * psxy); bpsxy += bSTXX * (-2.0 * w);
    bexy += bSTXY * 2 * G; bw += bSTXY * (psxx - psyy); bpsxx += bSTXY * w; bpsyy -= bSTXY * w;
    bvyy += bSTYY * 2 * G; btr += bSTYY * (-2 * G / 3.); bw += bSTYY * (2.0 * psxy); bpsxy += bSTYY * (2.0 * w);
    btr += bSTZZ * (-2 * G / 3.);
    bvxx += btr; bvyy += btr;
    g.acc[L_VD0] += bvxx; g.acc[L_VD3] += bvyy;
    g.acc[L_VD1] += 0.5 * bexy - 0.5 * bw;
    g.acc[L_VD2] += 0.5 * bexy + 0.5 * bw;
  } else {
    bRHOTi = bRHOT; bSTXXi = bSTXX; bSTXYi = bSTXY; bSTYYi = bSTYY; bSTZZi = bSTZZ;
  }
  // values before the stage above: the interaction outputs or the previous derivatives
  double bPTXi = bPTX, bPTYi = bPTY, bVTXi = bVTX, bVTYi = bVTY, bTTi = bTT;
  if (run_pse) { g.acc[L_TPSE] += bTT * apse; bTTi = 0; }
  if (run_mech) {
    g.acc[L_XT0] += bPTXi; g.acc[L_XT1] += bPTYi; g.acc[L_VT0] += bVTXi; g.acc[L_VT1] += bVTYi;
    bPTXi = 0; bPTYi = 0; bVTXi = 0; bVTYi = 0;
    if (ta != 0.) { g.acc[L_TL] += bTTi * ta; bTTi = 0; }
  }
  g.s[TS_PTX] += bPTXi; g.s[TS_PTY] += bPTYi; g.s[TS_VTX] += bVTXi; g.s[TS_VTY] += bVTYi;
  g.s[TS_RHOT] += bRHOTi; g.s[TS_STXX] += bSTXXi; g.s[TS_STXY] += bSTXYi; g.s[TS_STYY] += bSTYYi;
  g.s[TS_STZZ] += bSTZZi; g.s[TS_TT] += bTTi;
  g.s[TS_FCX] += bfcx0; g.s[TS_FCY] += bfcy0; g.s[TS_FTX] += bftx0; g.s[TS_FTY] += bfty0;
  // the predicted state used directly above
  g.s[TS_U] += bpu; g.s[TS_V] += bpv; g.s[TS_SXX] += bpsxx; g.s[TS_SXY] += bpsxy; g.s[TS_SYY] += bpsyy;
  g.s[TS_RHO] += bprho;
  if (!blank) {
    g.s[TS_VTX] += hdt * bpu; g.s[TS_VTY] += hdt * bpv; g.s[TS_STXX] += hdt * bpsxx; g.s[TS_STXY] += hdt * bpsxy;
    g.s[TS_STYY] += hdt * bpsyy; g.s[TS_RHOT] += hdt * bprho;
  }
  if (fwd) {
    // outputs after the boundary conditions
    double Uo = U, Vo = V, FCXo = FCX, FCYo = FCY, PTXo = PTX, PTYo = PTY, VTXo = VTX, VTYo = VTY, T2o = T2;
    if (!tool_p && fixed) { Uo = 0; Vo = 0; FCXo = 0; FCYo = 0; PTXo = 0; PTYo = 0; VTXo = 0; VTYo = 0; }
    if (k.thermal_bc && fixed) T2o = k.T_init;
    fwd[TS_X] = Xo; fwd[TS_Y] = Yo; fwd[TS_U] = Uo; fwd[TS_V] = Vo;
    fwd[TS_SXX] = S2xx; fwd[TS_SXY] = S2xy; fwd[TS_SYY] = S2yy; fwd[TS_SZZ] = S2zz;
    fwd[TS_RHO] = RHO; fwd[TS_T] = T2o; fwd[TS_EPS] = E2; fwd[TS_EPSD] = ED2; fwd[TS_H] = s[TS_H];
    fwd[TS_FCX] = FCXo; fwd[TS_FCY] = FCYo; fwd[TS_FTX] = FTX; fwd[TS_FTY] = FTY;
    fwd[TS_PTX] = PTXo; fwd[TS_PTY] = PTYo; fwd[TS_VTX] = VTXo; fwd[TS_VTY] = VTYo; fwd[TS_RHOT] = RHOT;
    fwd[TS_STXX] = STXX; fwd[TS_STXY] = STXY; fwd[TS_STYY] = STYY; fwd[TS_STZZ] = STZZ; fwd[TS_TT] = TT;
    fwd[TS_N] = tool_p ? in.nx : NX; fwd[TS_N + 1] = tool_p ? in.ny : NY;
  }
}

// ------------------------------------------------------------------ record
// the neighbour record of one particle from its sorted start-of-step state (predict, promoted);
// p: in/out positional pressure (recomputed unless tool particle), R3: in/out artificial stress
MFC_HD void record_fwd(const dev_consts &k, const double *s, int flags, double &p, double *R3, double *rec) {
  const bool blank = flags & MFC_FLAG_BLANKED, tool_p = flags & MFC_FLAG_TOOL;
  const double hdt = 0.5 * k.dt;
  double x = s[TS_X], y = s[TS_Y], u = s[TS_U], v = s[TS_V], sxx = s[TS_SXX], sxy = s[TS_SXY], syy = s[TS_SYY];
  double rho = s[TS_RHO], T = s[TS_T];
  if (!blank) {
    x = s[TS_X] + hdt * s[TS_PTX]; y = s[TS_Y] + hdt * s[TS_PTY];
    u = s[TS_U] + hdt * s[TS_VTX]; v = s[TS_V] + hdt * s[TS_VTY];
    sxx = s[TS_SXX] + hdt * s[TS_STXX]; sxy = s[TS_SXY] + hdt * s[TS_STXY]; syy = s[TS_SYY] + hdt * s[TS_STYY];
    rho = s[TS_RHO] + hdt * s[TS_RHOT]; T = s[TS_T] + hdt * s[TS_TT];
  }
  if (!tool_p) {
    p = k.c0 * k.c0 * (rho - k.rho0);
    art_stress_fwd(rho, p, sxx, sxy, syy, k.stress_eps, R3);
  }
  const double r21 = 1. / (rho * rho);
  rec[R_X] = x; rec[R_Y] = y; rec[R_FLAGS] = 0; rec[R_H] = s[TS_H]; rec[R_U] = u; rec[R_V] = v; rec[R_RHO] = rho;
  rec[R_VOL] = k.mass / rho; rec[R_AXX] = (sxx - p) * r21; rec[R_AXY] = sxy * r21; rec[R_AYY] = (syy - p) * r21;
  rec[R_RXX] = R3[0]; rec[R_RXY] = R3[1]; rec[R_RYY] = R3[2]; rec[R_C] = sqrt(k.K / rho); rec[R_T] = T;
}

// VJP of record_fwd: b_rec (adjoint of the record) -> adjoint of the sorted start-of-step state
// (added to gs, TS_*)
MFC_HD void record_vjp(const dev_consts &k, const double *s, int flags, double p, const double *b_rec, double *gs) {
  const bool blank = flags & MFC_FLAG_BLANKED, tool_p = flags & MFC_FLAG_TOOL;
  const double hdt = 0.5 * k.dt;
  double sxx = s[TS_SXX], sxy = s[TS_SXY], syy = s[TS_SYY], rho = s[TS_RHO];
  if (!blank) {
    sxx = s[TS_SXX] + hdt * s[TS_STXX]; sxy = s[TS_SXY] + hdt * s[TS_STXY]; syy = s[TS_SYY] + hdt * s[TS_STYY];
    rho = s[TS_RHO] + hdt * s[TS_RHOT];
  }
  double brho = b_rec[R_RHO], bsxx = 0, bsxy = 0, bsyy = 0, bp = 0;
  const double r21 = 1. / (rho * rho);
  brho += b_rec[R_VOL] * (-k.mass / (rho * rho));
  bsxx += b_rec[R_AXX] * r21; bsxy += b_rec[R_AXY] * r21; bsyy += b_rec[R_AYY] * r21;
  bp -= (b_rec[R_AXX] + b_rec[R_AYY]) * r21;
  brho += (b_rec[R_AXX] * (sxx - p) + b_rec[R_AXY] * sxy + b_rec[R_AYY] * (syy - p)) * (-2.0 * r21 / rho);
  brho += b_rec[R_C] * (-0.5 * sqrt(k.K / rho) / rho);
  if (!tool_p) {
    double bS[3] = {0, 0, 0};
    art_stress_vjp(rho, p, sxx, sxy, syy, k.stress_eps, b_rec + R_RXX, brho, bp, bS);
    bsxx += bS[0]; bsxy += bS[1]; bsyy += bS[2];
    brho += bp * k.c0 * k.c0;  // p = c0^2 (rho - rho0)
  }
  const double bx = b_rec[R_X], by = b_rec[R_Y], bu = b_rec[R_U], bv = b_rec[R_V], bT = b_rec[R_T];
  gs[TS_X] += bx; gs[TS_Y] += by; gs[TS_U] += bu; gs[TS_V] += bv;
  gs[TS_SXX] += bsxx; gs[TS_SXY] += bsxy; gs[TS_SYY] += bsyy; gs[TS_RHO] += brho; gs[TS_T] += bT;
  if (!blank) {
    gs[TS_PTX] += hdt * bx; gs[TS_PTY] += hdt * by; gs[TS_VTX] += hdt * bu; gs[TS_VTY] += hdt * bv;
    gs[TS_STXX] += hdt * bsxx; gs[TS_STXY] += hdt * bsxy; gs[TS_STYY] += hdt * bsyy; gs[TS_RHOT] += hdt * brho;
    gs[TS_TT] += hdt * bT;
  }
}

}  // namespace adj
}  // namespace mfc
