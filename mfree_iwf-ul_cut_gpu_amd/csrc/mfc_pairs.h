// mfc_pairs.h — the FP64 neighbour sums of one particle, shared by the step kernels and the
// adjoint's forward recomputation.
#pragma once

#include "mfc_kernels.h"

namespace mfc {

constexpr int REC_DOUBLES = 16;  // doubles per FP64 neighbour record

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
// COMPACT: the window is walked first, collecting the neighbours that contribute in a per-lane
// list (in the same order), and the list is evaluated afterwards, so that a wave runs the pair
// body about as often as its busiest lane has pairs instead of once per window candidate.
template <bool COMPACT>
__device__ inline void sums_f64(const dev_consts &k, int i, const double *__restrict__ rec,
                                const int *__restrict__ key_sorted, const int *__restrict__ offsets, bool tool,
                                double ta, bool run_mech, bool run_pse, pair_acc &acc, double &tpse) {
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
  const int hsh = key_sorted[i];
  const int gi = hsh / k.gny, gj = hsh - gi * k.gny;
  const int li = gi - 2 < 0 ? 0 : gi - 2;
  const int lj = gj - 2 < 0 ? 0 : gj - 2;
  const int hi_ = gi + 3 > k.gnx ? k.gnx : gi + 3;
  const int hj = gj + 3 > k.gny ? k.gny : gj + 3;
  // the PSE term of every window member and the test whether j adds to the mechanical sums
  auto visit = [&](int j) -> bool {
    const double *rj = rec + (size_t)REC_DOUBLES * j;
    const double2 b0 = reinterpret_cast<const double2 *>(rj)[0];
    const double2 b1 = reinterpret_cast<const double2 *>(rj)[1];
    const double xj = b0.x, yj = b0.y;
    const int fj = double_to_bits(b1.x);
    if (run_pse) {  // do_interactions_heat_pse (interactions_gpu.cu:130-139), all cell members
      const double2 b7 = reinterpret_cast<const double2 *>(rj)[7];
      const double2 b3 = reinterpret_cast<const double2 *>(rj)[3];
      const double dx = xi - xj, dy = yi - yj;
      const double xx = sqrt(dx * dx + dy * dy);
      const double w2 = pse_c * mfcm::exp(-xx * xx / (h2));
      tpse += (b7.y - Tpred) * w2 * k.mass / b3.x;
    }
    if (!run_mech) return false;
    if (fj & MFC_FLAG_BLANKED) return false;
    const double dx = xi - xj, dy = yi - yj;
    const double rr2 = dx * dx + dy * dy;
    // outside the support the reference adds exact zeros, unless a value is non-finite
    return !((rr2 >= fourh2 || rr2 < 1e-8) && !nonfin_i && !(fj & FLAG_NONFINITE));
  };
  auto pair = [&](int j) {
    const double *rj = rec + (size_t)REC_DOUBLES * j;
    const double2 b0 = reinterpret_cast<const double2 *>(rj)[0];
    const double2 b1 = reinterpret_cast<const double2 *>(rj)[1];
    pair_terms(k, acc, xi, yi, ui, vi_, hi, rhoi, Axx_i, Axy_i, Ayy_i, Rxx_i, Rxy_i, Ryy_i, ci, Ti, tool, ta, h1,
               fourh2, fac, rj, b0.x, b0.y, (double_to_bits(b1.x) & MFC_FLAG_TOOL) != 0);
  };
  if constexpr (!COMPACT) {
    for (int ii = li; ii < hi_; ii++) {
      const int jb = offsets[ii * k.gny + lj], je = offsets[ii * k.gny + hj];
      for (int j = jb; j < je; j++)
        if (visit(j)) pair(j);
    }
  } else {
    __shared__ int list[PAIR_LIST][TILE];
    int ii = li, j = 0, je = 0;
    bool more = ii < hi_;
    if (more) { j = offsets[ii * k.gny + lj]; je = offsets[ii * k.gny + hj]; }
    while (more) {
      int cnt = 0;
      while (cnt < PAIR_LIST) {
        if (j >= je) {
          if (++ii >= hi_) { more = false; break; }
          j = offsets[ii * k.gny + lj];
          je = offsets[ii * k.gny + hj];
          continue;
        }
        if (visit(j)) list[cnt++][threadIdx.x] = j;
        j++;
      }
      for (int q = 0; q < cnt; q++) pair(list[q][threadIdx.x]);
    }
  }
}

}  // namespace mfc
