// adj_chain.cpp — host restatement of the whole reverse-mode chain, built from the blocks of
// mfc_adjoint.h (test infrastructure).
//
// adjh_chain runs the cut forward with the blocks' forward companions (record_fwd, pair_fwd,
// pse_fwd, tail_vjp's forward outputs) and a plain stable sort, keeping every step's start
// state, then reverses every step: tail VJP per particle, pair and PSE VJPs pushed to both
// particles of each pair, record VJP, and the scatter through the sort permutation. The tests
// check (1) that the forward equals the oracle's step (so the blocks' forward is the step that
// is differentiated), (2) the host gradient against central differences of the host forward,
// and (3) the GPU gradient against the host gradient.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/mfc.h"
#include "../../mfree_iwf-ul_cut_gpu_amd/csrc/mfc_adjoint.h"
#include "../../mfree_iwf-ul_cut_gpu_amd/csrc/mfc_setup.h"

using namespace mfc;
using namespace mfc::adj;

namespace {

constexpr int FLAG_MASK = MFC_FLAG_FIXED | MFC_FLAG_BLANKED | MFC_FLAG_TOOL;

int64_t cell_coord(double f) {
  if (!(f == f)) return -1;
  if (f >= 1073741824.0) return 1073741824;
  if (f <= -1073741824.0) return -1073741824;
  return (int64_t)(int)f;
}

int cell_of(const dev_consts &k, double x, double y) {
  int64_t ix = cell_coord((x - k.bbmin_x) / k.cell);
  int64_t iy = cell_coord((y - k.bbmin_y) / k.cell);
  int64_t h = ix * k.gny + iy;
  if (h < 0) h = 0;
  if (h >= k.num_cell) h = k.num_cell - 1;
  return (int)h;
}

struct tape_step {
  std::vector<double> S;        // TS_N x N, storage order of the previous step
  std::vector<double> pos;      // 6 x N positional (nx, ny, p, rxx, rxy, ryy)
  std::vector<int> flags, keys;  // storage order; keys: the previous step's sorted ids
  mfc_tool_state tool;
};

struct stepper {
  dev_consts k;
  int N;
  // per step scratch
  std::vector<int> key, perm, offsets, sflags;
  std::vector<double> rec, acc;
  std::vector<double> s_sorted;  // TS_N x N

  // binning, sorted start state, records (positional p, R updated), sums
  void prepare(const tape_step &in, std::vector<double> &pos) {
    key.assign(N, 0);
    for (int i = 0; i < N; i++)
      key[i] = (in.flags[i] & MFC_FLAG_BLANKED) ? in.keys[i] : cell_of(k, in.S[TS_X * N + i], in.S[TS_Y * N + i]);
    perm.resize(N);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return key[a] < key[b]; });
    offsets.assign(k.num_cell + 1, 0);
    for (int i = 0; i < N; i++) offsets[key[i] + 1]++;
    for (int c = 0; c < k.num_cell; c++) offsets[c + 1] += offsets[c];
    s_sorted.assign((size_t)TS_N * N, 0.);
    sflags.assign(N, 0);
    rec.assign((size_t)REC_N * N, 0.);
    for (int a = 0; a < N; a++) {
      const int s = perm[a];
      double sa[TS_N];
      for (int q = 0; q < TS_N; q++) sa[q] = s_sorted[(size_t)q * N + a] = in.S[(size_t)q * N + s];
      sflags[a] = in.flags[s] & FLAG_MASK;
      double p = pos[2 * N + a], R3[3] = {pos[3 * N + a], pos[4 * N + a], pos[5 * N + a]};
      record_fwd(k, sa, sflags[a], p, R3, &rec[(size_t)REC_N * a]);
      pos[2 * N + a] = p;
      pos[3 * N + a] = R3[0]; pos[4 * N + a] = R3[1]; pos[5 * N + a] = R3[2];
    }
    acc.assign((size_t)L_N * N, 0.);
    for (int a = 0; a < N; a++) window(a, [&](int j) {
      const int fa = sflags[a], fj = sflags[j];
      const bool pse = k.pse != 0, tool_a = fa & MFC_FLAG_TOOL, blank_a = fa & MFC_FLAG_BLANKED;
      const double ta = pse ? 0. : (tool_a ? k.alpha_tool : k.alpha_wp);
      const bool mech = !blank_a && (!pse || !tool_a);
      const bool run_pse = !blank_a && pse && k.thermal_wp;
      double *o = &acc[(size_t)L_N * a];
      if (run_pse) o[L_TPSE] += pse_fwd(k, &rec[(size_t)REC_N * a], &rec[(size_t)REC_N * j]);
      if (mech && !(fj & MFC_FLAG_BLANKED))
        pair_fwd(k, &rec[(size_t)REC_N * a], &rec[(size_t)REC_N * j], tool_a, (fj & MFC_FLAG_TOOL) != 0, ta, o);
    });
  }

  template <class F>
  void window(int a, F f) {
    const int hsh = key[perm[a]];
    const int gi = hsh / k.gny, gj = hsh - gi * k.gny;
    const int li = gi - 2 < 0 ? 0 : gi - 2, lj = gj - 2 < 0 ? 0 : gj - 2;
    const int hi = gi + 3 > k.gnx ? k.gnx : gi + 3, hj = gj + 3 > k.gny ? k.gny : gj + 3;
    for (int ii = li; ii < hi; ii++)
      for (int j = offsets[ii * k.gny + lj]; j < offsets[ii * k.gny + hj]; j++) f(j);
  }

  tail_in tin(int a, const std::vector<double> &pos, const std::vector<double> &n_prev) const {
    tail_in in;
    for (int q = 0; q < TS_N; q++) in.s[q] = s_sorted[(size_t)q * N + a];
    in.p = pos[2 * N + a];
    in.nx = n_prev[a];
    in.ny = n_prev[N + a];
    in.flags = sflags[a];
    in.pse_flag_zero = k.pse_quirk && a >= N / 2;
    for (int q = 0; q < L_N; q++) in.acc[q] = acc[(size_t)L_N * a + q];
    return in;
  }
  static constexpr int REC_N = 16;
};

}  // namespace

extern "C" {

// nsteps of the cut from its initial state; Fmean over steps [w0, w1); grad = cot . dFmean/dtheta.
// state_out (optional, TS_N x N): the final state in sorted order; Fhist (optional, 2 x nsteps)
// cap (optional): the tail inputs of every particle at step nsteps - 1, per particle TS_N state
// values, p, n_prev (2), flags, quirk, L_N sums (TS_N + 5 + L_N doubles), plus the tool after
// the move of that step in cap_tool
// pre / theta_run (adjh_chain_sub): the first pre steps run with the configuration's theta and
// are not differentiated; the steps from pre on run with theta_run, and grad is with respect to
// theta_run for the mean force over [w0, w1) (pre <= w0)
int adjh_chain_cap(const mfc_config *cfg, int nsteps, int w0, int w1, const double *cot, double *Fmean,
                   double *grad, double *state_out, double *Fhist, double *cap, mfc_tool_state *cap_tool, int pre,
                   const double *theta_run);

int adjh_chain(const mfc_config *cfg, int nsteps, int w0, int w1, const double *cot, double *Fmean, double *grad,
               double *state_out, double *Fhist) {
  return adjh_chain_cap(cfg, nsteps, w0, w1, cot, Fmean, grad, state_out, Fhist, nullptr, nullptr, 0, nullptr);
}

static const double *g_state_bar = nullptr;  // optional seed: adjoint of the final state

int adjh_chain_sub(const mfc_config *cfg, int pre, int nsteps, int w0, int w1, const double *theta_run,
                   const double *cot, double *Fmean, double *grad) {
  return adjh_chain_cap(cfg, nsteps, w0, w1, cot, Fmean, grad, nullptr, nullptr, nullptr, nullptr, pre, theta_run);
}

// as adjh_chain_sub, with J = cot . Fmean + sum(state_bar * final state) (state_bar: TS_N x N,
// sorted order of the last step); state_out receives the final state
int adjh_chain_state(const mfc_config *cfg, int pre, int nsteps, int w0, int w1, const double *theta_run,
                     const double *cot, const double *state_bar, double *Fmean, double *grad, double *state_out) {
  g_state_bar = state_bar;
  int rc = adjh_chain_cap(cfg, nsteps, w0, w1, cot, Fmean, grad, state_out, nullptr, nullptr, nullptr, pre, theta_run);
  g_state_bar = nullptr;
  return rc;
}

int adjh_chain_cap(const mfc_config *cfg, int nsteps, int w0, int w1, const double *cot, double *Fmean,
                   double *grad, double *state_out, double *Fhist, double *cap, mfc_tool_state *cap_tool, int pre,
                   const double *theta_run) {
  scenario sc;
  std::string err;
  if (build_scenario(*cfg, sc, err)) return -1;
  stepper st;
  make_dev_consts(sc.k, st.k);
  dev_consts k_run = st.k;
  if (theta_run) {
    k_run.jc_A = theta_run[0]; k_run.jc_B = theta_run[1]; k_run.jc_n = theta_run[2];
    k_run.jc_C = theta_run[3]; k_run.jc_m = theta_run[4]; k_run.mu = theta_run[5];
  }
  const dev_consts k_pre = st.k;
  dev_consts &k = st.k;
  const int N = st.N = k.N;
  tape_step cur;
  cur.S.assign((size_t)TS_N * N, 0.);
  for (int i = 0; i < N; i++) {
    cur.S[TS_X * N + i] = sc.px[i]; cur.S[TS_Y * N + i] = sc.py[i]; cur.S[TS_RHO * N + i] = sc.rho[i];
    cur.S[TS_H * N + i] = sc.h[i]; cur.S[TS_T * N + i] = sc.T[i];
  }
  cur.pos.assign((size_t)6 * N, 0.);
  cur.flags = sc.flags;
  cur.keys.assign(N, 0);
  cur.tool = sc.tool;
  std::vector<tape_step> tape;
  std::vector<double> F((size_t)2 * nsteps, 0.);
  double zero_bar[TS_N] = {0};
  for (int t = 0; t < nsteps; t++) {
    k = t < pre ? k_pre : k_run;
    if (t >= pre) tape.push_back(cur);
    std::vector<double> pos = cur.pos;
    st.prepare(cur, pos);
    mfc_tool_state tool_now = cur.tool;
    tool_advance(tool_now, k.dt);
    tape_step nx;
    nx.S.assign((size_t)TS_N * N, 0.);
    nx.pos = pos;
    nx.flags = st.sflags;
    nx.keys.resize(N);
    for (int a = 0; a < N; a++) nx.keys[a] = st.key[st.perm[a]];
    nx.tool = tool_now;
    std::vector<double> n_prev(2 * (size_t)N);
    for (int a = 0; a < N; a++) { n_prev[a] = cur.pos[0 * (size_t)N + a]; n_prev[N + a] = cur.pos[1 * (size_t)N + a]; }
    double fx = 0., fy = 0.;
    if (cap && t == nsteps - 1) *cap_tool = tool_now;
    for (int a = 0; a < N; a++) {
      tail_in in = st.tin(a, pos, n_prev);
      if (cap && t == nsteps - 1) {
        double *c = cap + (size_t)(TS_N + 5 + L_N) * a;
        for (int q = 0; q < TS_N; q++) c[q] = in.s[q];
        c[TS_N] = in.p; c[TS_N + 1] = in.nx; c[TS_N + 2] = in.ny; c[TS_N + 3] = in.flags;
        c[TS_N + 4] = in.pse_flag_zero ? 1 : 0;
        for (int q = 0; q < L_N; q++) c[TS_N + 5 + q] = in.acc[q];
      }
      tail_grad g;
      double fwd[TS_N + 2];
      tail_vjp(k, tool_now, in, zero_bar, 0., 0., 0., 0., g, fwd);
      for (int q = 0; q < TS_N; q++) nx.S[(size_t)q * N + a] = fwd[q];
      nx.pos[0 * (size_t)N + a] = fwd[TS_N];
      nx.pos[1 * (size_t)N + a] = fwd[TS_N + 1];
      fx += fwd[TS_FCX] + fwd[TS_FTX];
      fy += fwd[TS_FCY] + fwd[TS_FTY];
    }
    F[2 * t] = fx;
    F[2 * t + 1] = fy;
    cur = nx;
  }
  if (state_out) std::memcpy(state_out, cur.S.data(), sizeof(double) * cur.S.size());
  if (Fhist) std::memcpy(Fhist, F.data(), sizeof(double) * F.size());
  double mx = 0., my = 0.;
  for (int t = w0; t < w1; t++) { mx += F[2 * t]; my += F[2 * t + 1]; }
  Fmean[0] = mx / (w1 - w0);
  Fmean[1] = my / (w1 - w0);
  if (!cot || !grad) return 0;
  // reverse
  for (int q = 0; q < 6; q++) grad[q] = 0.;
  std::vector<double> lam((size_t)TS_N * N, 0.), lam_n((size_t)2 * N, 0.);
  if (g_state_bar) lam.assign(g_state_bar, g_state_bar + (size_t)TS_N * N);
  const double Fb[2] = {cot[0] / (w1 - w0), cot[1] / (w1 - w0)};
  k = k_run;
  for (int t = nsteps - 1; t >= pre; t--) {
    const tape_step &in_t = tape[t - pre];
    std::vector<double> pos = in_t.pos;
    st.prepare(in_t, pos);
    mfc_tool_state tool_now = in_t.tool;
    tool_advance(tool_now, k.dt);
    std::vector<double> n_prev(2 * (size_t)N);
    for (int a = 0; a < N; a++) { n_prev[a] = in_t.pos[a]; n_prev[N + a] = in_t.pos[(size_t)N + a]; }
    const bool win = t >= w0 && t < w1;
    std::vector<double> g_s((size_t)TS_N * N), lam_acc((size_t)L_N * N), g_rec((size_t)16 * N, 0.);
    std::vector<double> lam_n_prev((size_t)2 * N);
    for (int a = 0; a < N; a++) {
      tail_in in = st.tin(a, pos, n_prev);
      double ob[TS_N];
      for (int q = 0; q < TS_N; q++) ob[q] = lam[(size_t)q * N + a];
      tail_grad g;
      tail_vjp(k, tool_now, in, ob, lam_n[a], lam_n[N + a], win ? Fb[0] : 0., win ? Fb[1] : 0., g, nullptr);
      for (int q = 0; q < TS_N; q++) g_s[(size_t)q * N + a] = g.s[q];
      for (int q = 0; q < L_N; q++) lam_acc[(size_t)L_N * a + q] = g.acc[q];
      lam_n_prev[a] = g.nx;
      lam_n_prev[N + a] = g.ny;
      for (int q = 0; q < 6; q++) grad[q] += g.theta[q];
    }
    // pair and PSE VJPs pushed to both particles
    for (int a = 0; a < N; a++) st.window(a, [&](int j) {
      const int fa = st.sflags[a], fj = st.sflags[j];
      const bool pse = k.pse != 0, tool_a = fa & MFC_FLAG_TOOL, blank_a = fa & MFC_FLAG_BLANKED;
      const double ta = pse ? 0. : (tool_a ? k.alpha_tool : k.alpha_wp);
      const bool mech = !blank_a && (!pse || !tool_a);
      const bool run_pse = !blank_a && pse && k.thermal_wp;
      const double *ra = &st.rec[(size_t)16 * a], *rj = &st.rec[(size_t)16 * j];
      if (run_pse) pse_vjp(k, ra, rj, lam_acc[(size_t)L_N * a + L_TPSE], true, true, &g_rec[16 * (size_t)a],
                           &g_rec[16 * (size_t)j]);
      if (mech && !(fj & MFC_FLAG_BLANKED))
        pair_vjp(k, ra, rj, tool_a, (fj & MFC_FLAG_TOOL) != 0, ta, &lam_acc[(size_t)L_N * a], true, true,
                 &g_rec[16 * (size_t)a], &g_rec[16 * (size_t)j]);
    });
    std::vector<double> lam_prev((size_t)TS_N * N, 0.);
    for (int a = 0; a < N; a++) {
      double s[TS_N], gs[TS_N];
      for (int q = 0; q < TS_N; q++) { s[q] = st.s_sorted[(size_t)q * N + a]; gs[q] = g_s[(size_t)q * N + a]; }
      record_vjp(k, s, st.sflags[a], pos[2 * (size_t)N + a], &g_rec[16 * (size_t)a], gs);
      for (int q = 0; q < TS_N; q++) lam_prev[(size_t)q * N + st.perm[a]] = gs[q];
    }
    lam.swap(lam_prev);
    lam_n.swap(lam_n_prev);
  }
  return 0;
}

}  // extern "C"
