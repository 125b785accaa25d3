// adj_harness.cpp — host build of the adjoint building blocks (mfc_adjoint.h) for the CPU tests.
//
// Every VJP of mfc_adjoint.h is exported next to its forward companion, so that the tests can
// compare the VJP with central differences of the forward function. The constants come from
// the product's own scenario builder (build_scenario + make_dev_consts), so the checks run at
// the magnitudes of a real cutting simulation. Test infrastructure only.
#include <cstring>
#include <string>

#include "../../include/mfc.h"
#include "../../mfree_iwf-ul_cut_gpu_amd/csrc/mfc_adjoint.h"
#include "../../mfree_iwf-ul_cut_gpu_amd/csrc/mfc_setup.h"

using namespace mfc;
using namespace mfc::adj;

extern "C" {

int adjh_consts_size() { return (int)sizeof(dev_consts); }

// constants and initial tool of a configuration; returns an MFC_E_* code
int adjh_make_consts(const mfc_config *cfg, void *k_out, mfc_tool_state *tool_out) {
  scenario sc;
  std::string err;
  int rc = build_scenario(*cfg, sc, err);
  if (rc) return rc;
  make_dev_consts(sc.k, *(dev_consts *)k_out);
  *tool_out = sc.tool;
  return 0;
}

// theta = (A, B, n, C, m, mu)
void adjh_set_theta(void *kp, const double *th) {
  dev_consts &k = *(dev_consts *)kp;
  k.jc_A = th[0]; k.jc_B = th[1]; k.jc_n = th[2]; k.jc_C = th[3]; k.jc_m = th[4]; k.mu = th[5];
}

void adjh_set_secant_tol(void *kp, double tol) { ((dev_consts *)kp)->secant_tol = tol; }

void adjh_pair_fwd(const void *kp, const double *ri, const double *rj, int ti, int tj, double ta, double *out) {
  pair_fwd(*(const dev_consts *)kp, ri, rj, ti != 0, tj != 0, ta, out);
}

void adjh_pair_vjp(const void *kp, const double *ri, const double *rj, int ti, int tj, double ta, const double *lam,
                   double *gi, double *gj) {
  pair_vjp(*(const dev_consts *)kp, ri, rj, ti != 0, tj != 0, ta, lam, true, true, gi, gj);
}

double adjh_pse_fwd(const void *kp, const double *ri, const double *rj) {
  return pse_fwd(*(const dev_consts *)kp, ri, rj);
}

void adjh_pse_vjp(const void *kp, const double *ri, const double *rj, double lam, double *gi, double *gj) {
  pse_vjp(*(const dev_consts *)kp, ri, rj, lam, true, true, gi, gj);
}

// x = (rho, p, Sxx, Sxy, Syy)
void adjh_art_fwd(const double *x, double eps, double *R) { art_stress_fwd(x[0], x[1], x[2], x[3], x[4], eps, R); }

void adjh_art_vjp(const double *x, double eps, const double *bR, double *gx) {
  double brho = 0, bp = 0, bS[3] = {0, 0, 0};
  art_stress_vjp(x[0], x[1], x[2], x[3], x[4], eps, bR, brho, bp, bS);
  gx[0] = brho; gx[1] = bp; gx[2] = bS[0]; gx[3] = bS[1]; gx[4] = bS[2];
}

// out = (s, d/de, d/ded, d/dT, d/dA, d/dB, d/dn, d/dC, d/dm)
void adjh_jc(const void *kp, double e, double ed, double T, double *out) {
  jc_grad g = sigma_yield_grad(*(const dev_consts *)kp, e, ed, T);
  out[0] = g.s; out[1] = g.e; out[2] = g.ed; out[3] = g.T; out[4] = g.A;
  out[5] = g.B; out[6] = g.n; out[7] = g.C; out[8] = g.m;
}

// in = (X, Y, U, V, ftx0, fty0); bar = adjoints of (fcx, fcy, ftx, fty, nx, ny);
// fwd = (fcx, fcy, ftx, fty, nx, ny); grad = adjoints of (X, Y, U, V, ftx0, fty0, mu)
int adjh_contact(const void *kp, const mfc_tool_state *tool, const double *in, const double *bar, double *fwd,
                 double *grad) {
  for (int q = 0; q < 7; q++) grad[q] = 0;
  contact_io c = contact_fwd_vjp(*(const dev_consts *)kp, *tool, in[0], in[1], in[2], in[3], in[4], in[5], bar[0],
                                 bar[1], bar[2], bar[3], bar[4], bar[5], grad[0], grad[1], grad[2], grad[3], grad[4],
                                 grad[5], grad[6]);
  fwd[0] = c.fcx; fwd[1] = c.fcy; fwd[2] = c.ftx; fwd[3] = c.fty; fwd[4] = c.nx; fwd[5] = c.ny;
  return c.inside ? 1 : 0;
}

// the tail of one particle: s[TS_N] start state, n_prev[2], acc[L_N]; adjoints of the outputs
// out_bar[TS_N], nbar[2], Fbar[2]; returns the adjoints g_s, g_n, g_acc, g_theta and the
// forward outputs fwd[TS_N + 2]
void adjh_tail(const void *kp, const mfc_tool_state *tool, const double *s, double p, const double *n_prev,
               int flags, int quirk, const double *acc, const double *out_bar, const double *nbar,
               const double *Fbar, double *g_s, double *g_n, double *g_acc, double *g_theta, double *fwd) {
  tail_in in;
  std::memcpy(in.s, s, sizeof(in.s));
  in.p = p; in.nx = n_prev[0]; in.ny = n_prev[1];
  in.flags = flags; in.pse_flag_zero = quirk != 0;
  std::memcpy(in.acc, acc, sizeof(in.acc));
  tail_grad g;
  tail_vjp(*(const dev_consts *)kp, *tool, in, out_bar, nbar[0], nbar[1], Fbar[0], Fbar[1], g, fwd);
  std::memcpy(g_s, g.s, sizeof(g.s));
  g_n[0] = g.nx; g_n[1] = g.ny;
  std::memcpy(g_acc, g.acc, sizeof(g.acc));
  std::memcpy(g_theta, g.theta, sizeof(g.theta));
}

void adjh_record_fwd(const void *kp, const double *s, int flags, double *p, double *R3, double *rec) {
  record_fwd(*(const dev_consts *)kp, s, flags, *p, R3, rec);
}

void adjh_record_vjp(const void *kp, const double *s, int flags, double p, const double *brec, double *gs) {
  for (int q = 0; q < TS_N; q++) gs[q] = 0;
  record_vjp(*(const dev_consts *)kp, s, flags, p, brec, gs);
}

}  // extern "C"
