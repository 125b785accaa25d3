// mfc_pairs.h — the FP64 neighbour sums of one particle, shared by the step kernels and the
// adjoint's forward recomputation.
#pragma once

#include "mfc_kernels.h"

namespace mfc {

constexpr int REC_DOUBLES = 16;  // doubles per FP64 neighbour record

// candidates of the window walk whose loads are issued together (FP32C direct walk: WALK_U32,
// with their whole records)
constexpr int WALK_U = 4;
constexpr int WALK_U32 = 2;

// where the window walk reads candidate positions and flags: candidate j at [j - base] of pos
// and fl (the records in memory with base 0, or a tile's copy in LDS)
struct cand_src {
  const double2 *pos;
  const int *fl;
  int base;
};

// The 5 x 5 cell window around the cell of a particle (interactions_gpu.cu:219-229): up to 5
// columns of the cell grid (clipped to the grid), each a contiguous range of the sorted order.
// The candidates are numbered 0..total-1 column by column in ascending sorted index, which is
// the reference's visiting order; at(t) is the sorted index of candidate t. The ten range
// bounds are loaded together, before any candidate.
struct window {
  int total;
  int c1, c2, c3, c4;      // number of the first candidate of columns 1..4
  int d0;                  // sorted index minus candidate number in column 0
  int e1, e2, e3, e4;      // its change from column r - 1 to column r
  __device__ inline int at(int t) const {
    // sums of selected increments (not a select among the five offsets, which the compiler
    // may turn into an indexed load from private memory)
    return t + d0 + (t >= c1 ? e1 : 0) + (t >= c2 ? e2 : 0) + (t >= c3 ? e3 : 0) + (t >= c4 ? e4 : 0);
  }
};

__device__ inline window make_window(const dev_consts &k, const int *__restrict__ offsets, int hsh) {
  const int gi = hsh / k.gny, gj = hsh - gi * k.gny;
  const int li = gi - 2 < 0 ? 0 : gi - 2;
  const int lj = gj - 2 < 0 ? 0 : gj - 2;
  const int hi = gi + 3 > k.gnx ? k.gnx : gi + 3;
  const int hj = gj + 3 > k.gny ? k.gny : gj + 3;
  int b[5], e[5];
#pragma unroll
  for (int r = 0; r < 5; r++) {
    const bool in = li + r < hi;
    const int ii = in ? li + r : hi - 1;
    b[r] = offsets[ii * k.gny + lj];
    e[r] = offsets[ii * k.gny + hj];
    if (!in) e[r] = b[r];
  }
  window w;
  // d_r = b[r] - c_r with c_r the candidates before column r; e_r = d_r - d_(r-1)
  int c = e[0] - b[0];
  w.d0 = b[0];
  w.c1 = c; w.e1 = b[1] - c - w.d0; c += e[1] - b[1];
  w.c2 = c; w.e2 = b[2] - c - (b[1] - w.c1); c += e[2] - b[2];
  w.c3 = c; w.e3 = b[3] - c - (b[2] - w.c2); c += e[3] - b[3];
  w.c4 = c; w.e4 = b[4] - c - (b[3] - w.c3); c += e[4] - b[4];
  w.total = c;
  return w;
}

__device__ inline double bits_to_double(int f) { return __longlong_as_double((long long)f); }
__device__ inline int double_to_bits(double d) { return (int)__double_as_longlong(d); }
__device__ inline bool finite_d(double v) { return !mfcm::is_nan(v) && !mfcm::is_inf(v); }

struct pair_acc {
  double vd0, vd1, vd2, vd3;   // v_der (vx,x vx,y vy,x vy,y)
  double sd0, sd1, sd2, sd3;   // S_der
  double vt0, vt1, xt0, xt1;   // vel_t, pos_t (XSPH)
  double tl;                   // Brookshaw sum
};

// neighbour sums stored per particle as ACC_N arrays of N doubles (structure of arrays, in
// the order vd0..vd3, sd0..sd3, vt0, vt1, xt0, xt1, tl, PSE sum)

__device__ inline void store_acc(double *__restrict__ acc, size_t N, int i, const pair_acc &a, double tpse) {
  acc[0 * N + i] = a.vd0; acc[1 * N + i] = a.vd1; acc[2 * N + i] = a.vd2; acc[3 * N + i] = a.vd3;
  acc[4 * N + i] = a.sd0; acc[5 * N + i] = a.sd1; acc[6 * N + i] = a.sd2; acc[7 * N + i] = a.sd3;
  acc[8 * N + i] = a.vt0; acc[9 * N + i] = a.vt1; acc[10 * N + i] = a.xt0; acc[11 * N + i] = a.xt1;
  acc[12 * N + i] = a.tl; acc[13 * N + i] = tpse;
}

__device__ inline void load_acc(const double *__restrict__ acc, size_t N, int i, pair_acc &a, double &tpse) {
  a.vd0 = acc[0 * N + i]; a.vd1 = acc[1 * N + i]; a.vd2 = acc[2 * N + i]; a.vd3 = acc[3 * N + i];
  a.sd0 = acc[4 * N + i]; a.sd1 = acc[5 * N + i]; a.sd2 = acc[6 * N + i]; a.sd3 = acc[7 * N + i];
  a.vt0 = acc[8 * N + i]; a.vt1 = acc[9 * N + i]; a.xt0 = acc[10 * N + i]; a.xt1 = acc[11 * N + i];
  a.tl = acc[12 * N + i]; tpse = acc[13 * N + i];
}

// the body of the neighbour loop of do_interactions_monaghan for one pair
// (interactions_gpu.cu:249-326 with cubic_spline kernels.cuh:30-70)
__device__ inline void pair_terms(const dev_consts &k, pair_acc &a, double xi, double yi, double ui, double vi_,
                                  double hi, double rhoi, double Axx_i, double Axy_i, double Ayy_i, double Rxx_i,
                                  double Rxy_i, double Ryy_i, double ci, double Ti, bool tool_i, double ta,
                                  double h1, double fourh2, double fac, const double *rj, double xj, double yj,
                                  bool tool_j) {
  const double2 *r2 = reinterpret_cast<const double2 *>(rj);
  double2 q1 = r2[1], q2 = r2[2], q3 = r2[3], q4 = r2[4], q5 = r2[5], q6 = r2[6], q7 = r2[7];
  const double hj = q1.y, uj = q2.x, vj = q2.y, rhoj = q3.x, volj = q3.y;
  const double Axx_j = q4.x, Axy_j = q4.y, Ayy_j = q5.x, Rxx_j = q5.y, Rxy_j = q6.x, Ryy_j = q6.y;
  const double cj = q7.x, Tj = (ta != 0.) ? q7.y : 0.;
  const double mass = k.mass;
  // cubic spline with h_i
  double w = 0., w_x = 0., w_y = 0.;
  const double xij = xi - xj, yij = yi - yj;
  const double rr2 = xij * xij + yij * yij;
  if (!(rr2 >= fourh2 || rr2 < 1e-8)) {
    double der, val;
    double rad = tsqrt(k, rr2);
    double q = rad * h1;
    if (q > 1) {
      double m = 2 - q, m2 = m * m;
      val = 0.25 * m2 * m2 * m2;
      der = -0.75 * m2 * h1 / rad;
    } else {
      val = 1 - 1.5 * q * q * (1 - 0.5 * q);
      der = -3.0 * q * (1 - 0.75 * q) * h1 / rad;
    }
    w = val * fac;
    w_x = der * xij * fac;
    w_y = der * yij * fac;
  }
  if (!tool_i && !tool_j) {
    a.vd0 += (uj - ui) * w_x * volj;
    a.vd1 += (uj - ui) * w_y * volj;
    a.vd2 += (vj - vi_) * w_x * volj;
    a.vd3 += (vj - vi_) * w_y * volj;
    double fab = w / k.wdeltap;
    fab *= fab;
    fab *= fab;
    const double Rx = fab * (Rxx_i + Rxx_j), Ry = fab * (Rxy_i + Rxy_j), Rz = fab * (Ryy_i + Ryy_j);
    a.sd0 += mass * (Axx_i + Axx_j + Rx) * w_x;
    a.sd1 += mass * (Axy_i + Axy_j + Ry) * w_y;
    a.sd2 += mass * (Axy_i + Axy_j + Ry) * w_x;
    a.sd3 += mass * (Ayy_i + Ayy_j + Rz) * w_y;
    const double vijx = ui - uj, vijy = vi_ - vj;
    const double vijposij = vijx * xij + vijy * yij;
    const double rhoij = 0.5 * (rhoi + rhoj);
    if (vijposij < 0.) {
      const double cij = 0.5 * (ci + cj);
      const double hij = 0.5 * (hi + hj);
      const double muij = (hij * vijposij) / (rr2 + k.av_eta * k.av_eta * hij * hij);
      const double piij = (-k.av_alpha * cij * muij + k.av_beta * muij * muij) / rhoij;
      a.vt0 += -mass * piij * w_x;
      a.vt1 += -mass * piij * w_y;
    }
    a.xt0 += -k.xsph_eps * w * mass / rhoij * vijx;
    a.xt1 += -k.xsph_eps * w * mass / rhoij * vijy;
  }
  if (ta != 0.) {  // Brookshaw
    const double rij = sqrt(rr2);
    if (rij > 1e-8) {
      const double eijx = xij / rij, eijy = yij / rij;
      const double rij1 = 1. / rij;
      a.tl += 2.0 * volj * (Ti - Tj) * rij1 * (eijx * w_x + eijy * w_y);
    }
  }
}

// FP64 neighbour sums of particle i: the reference's loop over the 5 x 5 cell window
// (interactions_gpu.cu:219-331), walked as 5 contiguous column ranges of the sorted order.
// COMPACT: the window is walked first, marking the neighbours that contribute in a per-lane
// bit mask, and the marked ones are evaluated afterwards in the same order, so that a wave runs
// the pair body about as often as its busiest lane has pairs instead of once per window
// candidate.
template <bool COMPACT>
__device__ inline void sums_f64(const dev_consts &k, int i, const double *__restrict__ rec,
                                const int *__restrict__ key_sorted, const int *__restrict__ offsets,
                                const cand_src &cs, bool tool, double ta, bool run_mech, bool run_pse, pair_acc &acc,
                                double &tpse) {
  const double2 *r2 = reinterpret_cast<const double2 *>(rec + (size_t)REC_DOUBLES * i);
  double2 a0 = r2[0], a1 = r2[1], a2 = r2[2], a3 = r2[3], a4 = r2[4], a5 = r2[5], a6 = r2[6], a7 = r2[7];
  const double xi = a0.x, yi = a0.y;
  const int rfl = double_to_bits(a1.x);
  const double hi = a1.y, ui = a2.x, vi_ = a2.y, rhoi = a3.x;
  const double Axx_i = a4.x, Axy_i = a4.y, Ayy_i = a5.x, Rxx_i = a5.y, Rxy_i = a6.x, Ryy_i = a6.y;
  const double ci = a7.x, Tpred = a7.y;
  const bool nonfin_i = rfl & FLAG_NONFINITE;
  const double Ti = (ta != 0.) ? Tpred : 0.;
  const double h1 = 1 / hi;
  const double fourh2 = 4 * hi * hi;
  const double fac = PI_FAC * h1 * h1;
  const double h2 = hi * hi, h4 = h2 * h2;
  const double pse_c = 4. / (h4 * M_PI);
  const window w = make_window(k, offsets, key_sorted[i]);
  // what the walk loads per candidate ahead of its evaluation: position and flags (density and
  // temperature for PSE are loaded when the candidate is evaluated, which keeps the registers
  // of the group at 5 per candidate)
  struct cand {
    double2 p;
    int f;
  };
  auto load = [&](int j, cand &c) {
    c.p = cs.pos[j - cs.base];
    c.f = cs.fl[j - cs.base];
  };
  // the PSE term of every window member and the test whether j adds to the mechanical sums
  auto visit = [&](int j, const cand &c) -> bool {
    const double xj = c.p.x, yj = c.p.y;
    if (run_pse) {  // do_interactions_heat_pse (interactions_gpu.cu:130-139), all cell members
      const double *rj = rec + (size_t)REC_DOUBLES * j;
      const double dx = xi - xj, dy = yi - yj;
      const double xx = sqrt(dx * dx + dy * dy);
      const double w2 = pse_c * mfcm::exp(-xx * xx / (h2));
      tpse += (rj[15] - Tpred) * w2 * k.mass / rj[6];
    }
    if (!run_mech) return false;
    if (c.f & MFC_FLAG_BLANKED) return false;
    const double dx = xi - xj, dy = yi - yj;
    const double rr2 = dx * dx + dy * dy;
    // outside the support the reference adds exact zeros, unless a value is non-finite
    return !((rr2 >= fourh2 || rr2 < 1e-8) && !nonfin_i && !(c.f & FLAG_NONFINITE));
  };
  auto pair = [&](int j, double xj, double yj, int fj) {
    pair_terms(k, acc, xi, yi, ui, vi_, hi, rhoi, Axx_i, Axy_i, Ayy_i, Rxx_i, Rxy_i, Ryy_i, ci, Ti, tool, ta, h1,
               fourh2, fac, rec + (size_t)REC_DOUBLES * j, xj, yj, (fj & MFC_FLAG_TOOL) != 0);
  };
  // the window in groups of WALK_U candidates, the loads of a group issued before its first
  // candidate is evaluated; the order of evaluation is the candidates' order
  if constexpr (!COMPACT) {
    for (int t0 = 0; t0 < w.total; t0 += WALK_U) {
      int j[WALK_U];
      cand c[WALK_U];
#pragma unroll
      for (int u = 0; u < WALK_U; u++) {
        j[u] = w.at(t0 + u < w.total ? t0 + u : w.total - 1);
        load(j[u], c[u]);
      }
#pragma unroll
      for (int u = 0; u < WALK_U; u++)
        if (t0 + u < w.total && visit(j[u], c[u])) pair(j[u], c[u].p.x, c[u].p.y, c[u].f);
    }
  } else {
    // masked walk: the tests of up to 64 candidates set the bits of a per-lane mask, then the
    // candidates of the set bits are evaluated in candidate order (the sorted index of a bit is
    // recomputed by window::at, so no list is kept in memory)
    for (int c0 = 0; c0 < w.total; c0 += 64) {
      const int n = w.total - c0 < 64 ? w.total - c0 : 64;
      unsigned long long m = 0ull;
      for (int u0 = 0; u0 < n; u0 += WALK_U) {
        int j[WALK_U];
        cand c[WALK_U];
#pragma unroll
        for (int u = 0; u < WALK_U; u++) {
          j[u] = w.at(c0 + (u0 + u < n ? u0 + u : n - 1));
          load(j[u], c[u]);
        }
#pragma unroll
        for (int u = 0; u < WALK_U; u++)
          if (u0 + u < n && visit(j[u], c[u])) m |= 1ull << (u0 + u);
      }
      while (m) {
        const int jj = w.at(c0 + __builtin_ctzll(m));
        m &= m - 1ull;
        const double2 p = cs.pos[jj - cs.base];
        pair(jj, p.x, p.y, cs.fl[jj - cs.base]);
      }
    }
  }
}

}  // namespace mfc
