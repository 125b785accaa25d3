// mfc_adjoint.hip — reverse mode of the SPH cutting step on gfx950: gradients of the mean tool
// force over a window of steps with respect to the Johnson-Cook constants and the friction
// coefficient, theta = (A, B, n, C, m, mu).
//
// Tape: the state at the start of every K-th step (state block, positional block, stale cell
// ids of blanked particles, tool) is copied into a checkpoint. The backward pass walks the
// checkpoint segments from last to first; each segment is replayed forward with the step
// kernels to recover the start state of each of its steps, and the steps are then reversed
// one by one:
//   F1  binning          launch_binning (cell ids, stable counting sort)
//   F2  predict          launch_predict with the sorted copy of the start state (gather,
//                        predictor, EOS, artificial stress, records)
//   F3  launch_sums      neighbour sums of every particle (the sums kernel of the step)
//   B1  k_adj_tail       VJP of everything after the sums, per particle (tail_vjp), and the
//                        per-tile theta partials of the step (k_adj_theta adds them up)
//   B2  k_adj_pull       VJP of the neighbour sums with respect to the records: each particle
//                        pulls the contributions of the pairs it takes part in over the same
//                        symmetric 5 x 5 cell window, in a fixed order (no atomics); two passes,
//                        pairs of the neighbours' sums, then of its own sums
//   B3  k_adj_record     VJP of the record and predictor, then the scatter back to the
//                        storage order of the previous step (inverse of the sort permutation)
// The positional contact normal (read by the friction heat of the next step at the same sorted
// index) carries its adjoint from step to step without permutation, like its forward value.
//
// Derivatives follow the smooth map of promote_fp64 = 1 with a tight secant tolerance; the
// handle must use FP64 records.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "mfc_adjoint.h"
#include "mfc_pairs.h"
#include "mfc_sim.h"

namespace mfc {

namespace {

using namespace adj;

// ------------------------------------------------------------------ B1: per-particle tail
// st: sorted start state (DSTATE_DOUBLES arrays of stride N, TS_* order); lam_s: adjoint of the
// end state of the step (same layout); lam_n: adjoint of the positional normal written by this
// step (in) and of the one read by it (out)
__global__ __launch_bounds__(TILE) void k_adj_tail(dev_consts k, int ntiles, const double *__restrict__ st,
                                                   const int *__restrict__ flags, dpositional pos,
                                                   const double *__restrict__ acc, const double *__restrict__ lam_s,
                                                   double *__restrict__ lam_n, double Fbx, double Fby,
                                                   const mfc_tool_state *__restrict__ tool_now,
                                                   double *__restrict__ g_s, double *__restrict__ lam_acc,
                                                   double *__restrict__ theta_part) {
  const int t = xcd_tile(blockIdx.x, ntiles);
  if (t < 0) return;
  const int i = t * TILE + threadIdx.x;
  const size_t N = k.N;
  double th[6] = {0., 0., 0., 0., 0., 0.};
  if (i < k.N) {
    tail_in in;
#pragma unroll
    for (int q = 0; q < TS_N; q++) in.s[q] = st[q * N + i];
    in.p = pos.p[i];
    in.nx = pos.nx[i];
    in.ny = pos.ny[i];
    in.flags = flags[i];
    in.pse_flag_zero = k.pse_quirk && i >= k.N / 2;
#pragma unroll
    for (int q = 0; q < L_N; q++) in.acc[q] = acc[q * N + i];
    double ob[TS_N];
#pragma unroll
    for (int q = 0; q < TS_N; q++) ob[q] = lam_s[q * N + i];
    const double nbx = lam_n[i], nby = lam_n[N + i];
    tail_grad g;
    tail_vjp(k, *tool_now, in, ob, nbx, nby, Fbx, Fby, g, nullptr);
#pragma unroll
    for (int q = 0; q < TS_N; q++) g_s[q * N + i] = g.s[q];
#pragma unroll
    for (int q = 0; q < L_N; q++) lam_acc[q * N + i] = g.acc[q];
    lam_n[i] = g.nx;
    lam_n[N + i] = g.ny;
#pragma unroll
    for (int q = 0; q < 6; q++) th[q] = g.theta[q];
  }
  // theta partials: halving-stride tree per tile
  __shared__ double red[6][TILE];
#pragma unroll
  for (int q = 0; q < 6; q++) red[q][threadIdx.x] = th[q];
  for (int s = TILE / 2; s > 0; s >>= 1) {
    __syncthreads();
    if (threadIdx.x < s) {
#pragma unroll
      for (int q = 0; q < 6; q++) red[q][threadIdx.x] = red[q][threadIdx.x] + red[q][threadIdx.x + s];
    }
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 6; q++) theta_part[6 * t + q] = red[q][0];
  }
}

// adds the step's theta partials to the running totals: lane l sums the tiles l, l + TILE, ...
// in ascending order, lane sums reduced by halving strides
__global__ __launch_bounds__(TILE) void k_adj_theta(int ntiles, const double *__restrict__ theta_part,
                                                    double *__restrict__ theta_acc) {
  __shared__ double red[6][TILE];
  double a[6] = {0., 0., 0., 0., 0., 0.};
  for (int u = threadIdx.x; u < ntiles; u += TILE)
#pragma unroll
    for (int q = 0; q < 6; q++) a[q] = a[q] + theta_part[6 * u + q];
#pragma unroll
  for (int q = 0; q < 6; q++) red[q][threadIdx.x] = a[q];
  for (int s = TILE / 2; s > 0; s >>= 1) {
    __syncthreads();
    if (threadIdx.x < s) {
#pragma unroll
      for (int q = 0; q < 6; q++) red[q][threadIdx.x] = red[q][threadIdx.x] + red[q][threadIdx.x + s];
    }
  }
  if (threadIdx.x < 6) theta_acc[threadIdx.x] = theta_acc[threadIdx.x] + red[threadIdx.x][0];
}

// ------------------------------------------------------------------ B2: pull over pairs
// OWN = false: the pairs (a, b) of the sums of a window member a that include b, with the
// adjoint of a's sums; g_rec[b] is written. OWN = true: the pairs (b, a) of b's own sums, with
// b's adjoint; added to g_rec[b]. Two passes keep fewer records live per lane than one.
template <bool OWN>
__global__ __launch_bounds__(TILE) void k_adj_pull(dev_consts k, int ntiles, const double *__restrict__ rec,
                                                   const int *__restrict__ key_sorted,
                                                   const int *__restrict__ offsets, const int *__restrict__ flags,
                                                   const double *__restrict__ lam_acc, double *__restrict__ g_rec) {
  const int t = xcd_tile(blockIdx.x, ntiles);
  if (t < 0) return;
  const int b = t * TILE + threadIdx.x;
  if (b >= k.N) return;
  const size_t N = k.N;
  const bool pse = k.pse != 0;
  double rb[REC_DOUBLES], gb[REC_DOUBLES], lb[L_N], dummy[REC_DOUBLES];
#pragma unroll
  for (int q = 0; q < REC_DOUBLES; q++) {
    rb[q] = rec[REC_DOUBLES * (size_t)b + q];
    gb[q] = OWN ? g_rec[q * N + b] : 0.;
  }
  if constexpr (OWN) {
#pragma unroll
    for (int q = 0; q < L_N; q++) lb[q] = lam_acc[q * N + b];
  }
  const int fb = flags[b];
  const bool blank_b = fb & MFC_FLAG_BLANKED, tool_b = fb & MFC_FLAG_TOOL;
  const double ta_b = pse ? 0. : (tool_b ? k.alpha_tool : k.alpha_wp);
  const bool mech_b = !blank_b && (!pse || !tool_b);
  const bool pse_b = !blank_b && pse && k.thermal_wp;
  const int hsh = key_sorted[b];
  const int gi = hsh / k.gny, gj = hsh - gi * k.gny;
  const int li = gi - 2 < 0 ? 0 : gi - 2;
  const int lj = gj - 2 < 0 ? 0 : gj - 2;
  const int hi_ = gi + 3 > k.gnx ? k.gnx : gi + 3;
  const int hj = gj + 3 > k.gny ? k.gny : gj + 3;
  for (int ii = li; ii < hi_; ii++) {
    const int jb = offsets[ii * k.gny + lj], je = offsets[ii * k.gny + hj];
    for (int a = jb; a < je; a++) {
      if (a == b) continue;
      const int fa = flags[a];
      const bool blank_a = fa & MFC_FLAG_BLANKED, tool_a = fa & MFC_FLAG_TOOL;
      const bool mech_a = !blank_a && (!pse || !tool_a);
      const bool pse_a = !blank_a && pse && k.thermal_wp;
      const double *pa = rec + REC_DOUBLES * (size_t)a;
      const double dx = pa[R_X] - rb[R_X], dy = pa[R_Y] - rb[R_Y];
      const double rr2 = dx * dx + dy * dy;
      if constexpr (!OWN) {
        const double ha = pa[R_H];
        const bool mech = mech_a && !blank_b && rr2 < 4 * ha * ha && rr2 >= 1e-8;  // a's sums include b
        if (!mech && !pse_a) continue;
        double ra[REC_DOUBLES], la[L_N];
#pragma unroll
        for (int q = 0; q < REC_DOUBLES; q++) ra[q] = pa[q];
#pragma unroll
        for (int q = 0; q < L_N; q++) la[q] = lam_acc[q * N + a];
        const double ta_a = pse ? 0. : (tool_a ? k.alpha_tool : k.alpha_wp);
        if (mech) pair_vjp(k, ra, rb, tool_a, tool_b, ta_a, la, false, true, dummy, gb);
        if (pse_a) pse_vjp(k, ra, rb, la[L_TPSE], false, true, dummy, gb);
      } else {
        const double hb = rb[R_H];
        const bool mech = mech_b && !blank_a && rr2 < 4 * hb * hb && rr2 >= 1e-8;  // b's sums include a
        if (!mech && !pse_b) continue;
        double ra[REC_DOUBLES];
#pragma unroll
        for (int q = 0; q < REC_DOUBLES; q++) ra[q] = pa[q];
        if (mech) pair_vjp(k, rb, ra, tool_b, tool_a, ta_b, lb, true, false, gb, dummy);
        if (pse_b) pse_vjp(k, rb, ra, lb[L_TPSE], true, false, gb, dummy);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < REC_DOUBLES; q++) g_rec[q * N + b] = gb[q];
}

// ------------------------------------------------------------------ B3: record VJP + unsort
__global__ __launch_bounds__(TILE) void k_adj_record(dev_consts k, int ntiles, const double *__restrict__ st,
                                                     const int *__restrict__ flags, const double *__restrict__ p,
                                                     const double *__restrict__ g_rec,
                                                     const double *__restrict__ g_s, const int *__restrict__ perm,
                                                     double *__restrict__ lam_prev) {
  const int t = xcd_tile(blockIdx.x, ntiles);
  if (t < 0) return;
  const int a = t * TILE + threadIdx.x;
  if (a >= k.N) return;
  const size_t N = k.N;
  double s[TS_N], gs[TS_N], br[REC_DOUBLES];
#pragma unroll
  for (int q = 0; q < TS_N; q++) { s[q] = st[q * N + a]; gs[q] = g_s[q * N + a]; }
#pragma unroll
  for (int q = 0; q < REC_DOUBLES; q++) br[q] = g_rec[q * N + a];
  record_vjp(k, s, flags[a], p[a], br, gs);
  const int dst = perm[a];
#pragma unroll
  for (int q = 0; q < TS_N; q++) lam_prev[q * N + dst] = gs[q];
}

int padded(int ntiles) { return ((ntiles + 7) / 8) * 8; }

}  // namespace

// ------------------------------------------------------------------ workspace
struct snapshot {
  char *state;            // state block (state_block_bytes)
  char *pos;              // positional block
  int *keys;              // cell ids of the previous step (kept by blanked particles)
  mfc_tool_state *tool;
};

struct adj_work {
  int N = 0, nsteps = 0, K = 1, w0 = 0, w1 = 0;
  bool valid = false;
  int64_t hist0 = 0;  // force-history index of the first step of the run
  double Fmean[2] = {0., 0.};
  std::vector<void *> allocs;
  std::vector<snapshot> ckpt, seg;
  char *work[2] = {nullptr, nullptr};   // replay state blocks
  char *posw = nullptr;                 // replay positional block
  mfc_tool_state *toolw = nullptr;      // 2
  double *partials = nullptr, *force_last = nullptr, *hist = nullptr;
  int64_t *ctr = nullptr;
  int *pending = nullptr;
  int *keys_keep = nullptr;  // the handle's cell ids, kept across the replay of the backward pass
  double *acc = nullptr, *lam_acc = nullptr, *lam_s[2] = {nullptr, nullptr}, *lam_n = nullptr, *g_s = nullptr,
         *g_rec = nullptr, *theta_part = nullptr, *theta_acc = nullptr;

  template <class T>
  int alloc(T **p, size_t n) {
    void *q = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&q, n * sizeof(T));
    if (e != hipSuccess) return fail(MFC_E_NOMEM, std::string("hipMalloc (adjoint): ") + hipGetErrorString(e));
    allocs.push_back(q);
    *p = (T *)q;
    return MFC_OK;
  }
  int alloc_snapshot(snapshot &sn) {
    int rc;
    if ((rc = alloc(&sn.state, state_block_bytes(N))) || (rc = alloc(&sn.pos, positional_block_bytes(N))) ||
        (rc = alloc(&sn.keys, (size_t)N)) || (rc = alloc(&sn.tool, 1)))
      return rc;
    return MFC_OK;
  }
  ~adj_work() {
    for (void *q : allocs) (void)hipFree(q);
  }
};

void adj_free(adj_work *w) { delete w; }

namespace {

int save(hipStream_t st, const snapshot &d, const void *state, const void *pos, const int *keys,
         const mfc_tool_state *tool, int N) {
  HIP_TRY(hipMemcpyAsync(d.state, state, state_block_bytes(N), hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(d.pos, pos, positional_block_bytes(N), hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(d.keys, keys, sizeof(int) * (size_t)N, hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(d.tool, tool, sizeof(mfc_tool_state), hipMemcpyDeviceToDevice, st));
  return MFC_OK;
}

int restore(hipStream_t st, const snapshot &s, void *state, void *pos, int *keys, mfc_tool_state *tool, int N) {
  HIP_TRY(hipMemcpyAsync(state, s.state, state_block_bytes(N), hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(pos, s.pos, positional_block_bytes(N), hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(keys, s.keys, sizeof(int) * (size_t)N, hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(tool, s.tool, sizeof(mfc_tool_state), hipMemcpyDeviceToDevice, st));
  return MFC_OK;
}

int setup_work(mfc_sim *s, int nsteps, int K, int w0, int w1) {
  const int N = s->k.N;
  adj_work *w = s->adj;
  const int nck = (nsteps + K - 1) / K;
  const bool reuse = w && w->N == N && w->K == K && (int)w->ckpt.size() >= nck;
  if (!reuse) {
    adj_free(w);
    w = s->adj = new adj_work();
    w->N = N;
    w->K = K;
    int rc;
    w->ckpt.resize(nck);
    for (auto &c : w->ckpt)
      if ((rc = w->alloc_snapshot(c))) return rc;
    w->seg.resize(K);
    for (auto &c : w->seg)
      if ((rc = w->alloc_snapshot(c))) return rc;
    const size_t n = (size_t)N;
    if ((rc = w->alloc(&w->work[0], state_block_bytes(N))) || (rc = w->alloc(&w->work[1], state_block_bytes(N))) ||
        (rc = w->alloc(&w->posw, positional_block_bytes(N))) || (rc = w->alloc(&w->toolw, 2)) ||
        (rc = w->alloc(&w->partials, 16 * (size_t)s->ntiles)) || (rc = w->alloc(&w->force_last, 2)) ||
        (rc = w->alloc(&w->hist, 2)) || (rc = w->alloc(&w->ctr, 1)) || (rc = w->alloc(&w->pending, 1)) ||
        (rc = w->alloc(&w->keys_keep, n)) ||
        (rc = w->alloc(&w->acc, L_N * n)) || (rc = w->alloc(&w->lam_acc, L_N * n)) ||
        (rc = w->alloc(&w->lam_s[0], TS_N * n)) || (rc = w->alloc(&w->lam_s[1], TS_N * n)) ||
        (rc = w->alloc(&w->lam_n, 2 * n)) || (rc = w->alloc(&w->g_s, TS_N * n)) ||
        (rc = w->alloc(&w->g_rec, REC_DOUBLES * n)) || (rc = w->alloc(&w->theta_part, 6 * (size_t)s->ntiles)) ||
        (rc = w->alloc(&w->theta_acc, 6)))
      return rc;
    // the replay's force reduction keeps no history (cap 0) but counts steps
    HIP_TRY(hipMemsetAsync(w->ctr, 0, sizeof(int64_t), s->stream));
    HIP_TRY(hipMemsetAsync(w->pending, 0, sizeof(int), s->stream));
  }
  w->nsteps = nsteps;
  w->w0 = w0;
  w->w1 = w1;
  w->valid = false;
  return MFC_OK;
}

}  // namespace
}  // namespace mfc

namespace mfc {

int adj_begin(mfc_sim *s, int nsteps, int checkpoint_every, int w0, int w1) {
  if (nsteps < 1 || checkpoint_every < 1 || w0 < 0 || w1 <= w0 || w1 > nsteps)
    return fail(MFC_E_INVALID, "need nsteps >= 1, checkpoint_every >= 1 and 0 <= w0 < w1 <= nsteps");
  if (s->precision != MFC_PREC_FP64 || !s->k.promote)
    return fail(MFC_E_INVALID, "the adjoint needs precision FP64 and promote_fp64 = 1");
  if (s->sc.tool.num_seg == 0 || !s->k.plastic)
    return fail(MFC_E_INVALID, "the adjoint differentiates the cutting scenario (tool and plasticity present)");
  if (s->k.jc_tanh) return fail(MFC_E_INVALID, "the adjoint differentiates the plain Johnson-Cook model (jc_model = 0)");
  if (s->k.adaptive) return fail(MFC_E_INVALID, "the adjoint runs on a locked grid (grid_adaptive = 0)");
  if (s->k.rothlin) return fail(MFC_E_INVALID, "the adjoint runs on the green grid (grid_kind = 0)");
  int64_t base = 0;
  int rc = mfc_force_history(s, nullptr, 0, &base);
  if (rc) return rc;
  if (base + nsteps > HIST_CAP) return fail(MFC_E_INVALID, "force history full: reset it before the run");
  if ((rc = setup_work(s, nsteps, checkpoint_every, w0, w1))) return rc;
  s->adj->hist0 = base;
  return MFC_OK;
}

int adj_checkpoint(mfc_sim *s, int t) {
  adj_work *w = s->adj;
  if (t % w->K) return MFC_OK;
  const int c = s->cur;
  return save(s->stream, w->ckpt[t / w->K], s->set[c].px, s->pos.nx, s->bins.key_sorted, s->tool + c, s->k.N);
}

int adj_finish(mfc_sim *s, double Fmean[2]) {
  adj_work *w = s->adj;
  const int n = w->w1 - w->w0;
  std::vector<double> F((size_t)2 * n);
  HIP_TRY(hipMemcpyAsync(F.data(), s->hist + 2 * (w->hist0 + w->w0), sizeof(double) * F.size(),
                         hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  double sx = 0., sy = 0.;
  for (int t = 0; t < n; t++) {
    sx += F[2 * t];
    sy += F[2 * t + 1];
  }
  w->Fmean[0] = Fmean[0] = sx / n;
  w->Fmean[1] = Fmean[1] = sy / n;
  w->valid = true;
  return MFC_OK;
}

}  // namespace mfc

using namespace mfc;

extern "C" {

int mfc_vjp_forward(mfc_sim *s, int nsteps, int checkpoint_every, int w0, int w1, double Fmean[2]) {
  if (!s || !Fmean) return fail(MFC_E_INVALID, "null argument");
  if (s->batched) return fail(MFC_E_INVALID, "instance of a batch: use mfc_batch_vjp_forward");
  int rc = adj_begin(s, nsteps, checkpoint_every, w0, w1);
  if (rc || (rc = ensure_keys(s))) return rc;
  for (int t = 0; t < nsteps; t++) {
    if ((rc = adj_checkpoint(s, t))) return rc;
    enqueue_step(s, s->cur);
    s->cur = 1 - s->cur;
    s->step++;
  }
  finish_steps(s);
  HIP_TRY(hipGetLastError());
  return adj_finish(s, Fmean);
}

int mfc_vjp_backward(mfc_sim *s, const double cot[2], double grad[6]) {
  if (!s || !cot || !grad) return fail(MFC_E_INVALID, "null argument");
  adj_work *w = s->adj;
  if (!w || !w->valid) return fail(MFC_E_INVALID, "no recorded run: call mfc_vjp_forward first");
  const int N = s->k.N, nt = s->ntiles, K = w->K;
  const dev_consts &k = s->k;
  hipStream_t st = s->stream;
  const size_t n = (size_t)N;
  HIP_TRY(hipMemsetAsync(w->lam_s[0], 0, sizeof(double) * TS_N * n, st));
  HIP_TRY(hipMemsetAsync(w->lam_n, 0, sizeof(double) * 2 * n, st));
  HIP_TRY(hipMemsetAsync(w->theta_acc, 0, sizeof(double) * 6, st));
  HIP_TRY(hipMemcpyAsync(w->keys_keep, s->bins.key_sorted, sizeof(int) * n, hipMemcpyDeviceToDevice, st));
  // the replays bin with the handle's cell table, which holds the next step's counts: clear them
  // (each replayed binning leaves them cleared), and bin the handle's state again at its next step
  HIP_TRY(hipMemsetAsync(s->bins.count, 0, sizeof(int) * (size_t)k.ncell_table, st));
  s->keys_fresh = false;
  int cur = 0;  // lam_s[cur] holds the adjoint of the end state of the step being reversed
  const double Fb[2] = {cot[0] / (w->w1 - w->w0), cot[1] / (w->w1 - w->w0)};
  // the step kernels run on the replay buffers; cell table and records are the handle's
  // (recomputed every step, so the handle's next step is unaffected)
  dstate A, B;
  dpositional P;
  carve_positional(w->posw, N, P);
  dforce fo = {w->partials, w->force_last, w->hist, 0, w->ctr, w->pending};
  const int nseg = (w->nsteps + K - 1) / K;
  int rc;
  for (int sg = nseg - 1; sg >= 0; sg--) {
    const int t0 = sg * K, t1 = (t0 + K < w->nsteps) ? t0 + K : w->nsteps;
    // replay the segment, keeping the start of each of its steps
    int wa = 0;
    if ((rc = restore(st, w->ckpt[sg], w->work[wa], w->posw, s->bins.key_sorted, w->toolw, N))) return rc;
    for (int t = t0; t < t1; t++) {
      if ((rc = save(st, w->seg[t - t0], w->work[wa], w->posw, s->bins.key_sorted, w->toolw, N))) return rc;
      if (t + 1 == t1) break;
      carve_state(w->work[wa], N, A);
      carve_state(w->work[1 - wa], N, B);
      launch_binning(k, A, s->bins, st);
      launch_predict(k, A, B, s->bins, P, s->recs, w->toolw, w->toolw + 1, false, st);
      launch_interact(k, A, B, s->bins, P, s->recs, w->toolw + 1, fo, s->acc, s->variant(), st);
      HIP_TRY(hipMemcpyAsync(w->toolw, w->toolw + 1, sizeof(mfc_tool_state), hipMemcpyDeviceToDevice, st));
      wa = 1 - wa;
    }
    // reverse the steps of the segment
    for (int t = t1 - 1; t >= t0; t--) {
      if ((rc = restore(st, w->seg[t - t0], w->work[0], w->posw, s->bins.key_sorted, w->toolw, N))) return rc;
      carve_state(w->work[0], N, A);
      carve_state(w->work[1], N, B);
      launch_binning(k, A, s->bins, st);
      launch_predict(k, A, B, s->bins, P, s->recs, w->toolw, w->toolw + 1, true, st);
      const double *rec = s->recs.rec;
      const double *sb = B.px;  // sorted start state, TS_* order with stride N
      const bool in_win = t >= w->w0 && t < w->w1;
      launch_sums(k, A.flags, s->bins, s->recs, w->acc, s->variant(), st);
      hipLaunchKernelGGL(k_adj_tail, dim3(padded(nt)), dim3(TILE), 0, st, k, nt, sb, B.flags, P, w->acc,
                         w->lam_s[cur], w->lam_n, in_win ? Fb[0] : 0., in_win ? Fb[1] : 0., w->toolw + 1, w->g_s,
                         w->lam_acc, w->theta_part);
      hipLaunchKernelGGL(k_adj_theta, dim3(1), dim3(TILE), 0, st, nt, w->theta_part, w->theta_acc);
      hipLaunchKernelGGL(k_adj_pull<false>, dim3(padded(nt)), dim3(TILE), 0, st, k, nt, rec, s->bins.key_sorted,
                         s->bins.offsets, B.flags, w->lam_acc, w->g_rec);
      hipLaunchKernelGGL(k_adj_pull<true>, dim3(padded(nt)), dim3(TILE), 0, st, k, nt, rec, s->bins.key_sorted,
                         s->bins.offsets, B.flags, w->lam_acc, w->g_rec);
      hipLaunchKernelGGL(k_adj_record, dim3(padded(nt)), dim3(TILE), 0, st, k, nt, sb, B.flags, P.p, w->g_rec,
                         w->g_s, s->bins.perm, w->lam_s[1 - cur]);
      cur = 1 - cur;
    }
  }
  HIP_TRY(hipGetLastError());
  // the replay overwrote the handle's cell ids, which blanked particles keep into the next step
  HIP_TRY(hipMemcpyAsync(s->bins.key_sorted, w->keys_keep, sizeof(int) * n, hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(grad, w->theta_acc, sizeof(double) * 6, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return MFC_OK;
}

int mfc_vjp(mfc_sim *s, mfc_vjp_req *req, double grad[6]) {
  if (!s || !req || !grad) return fail(MFC_E_INVALID, "null argument");
  int rc = mfc_vjp_forward(s, req->nsteps, req->checkpoint_every, req->w0, req->w1, req->Fmean);
  if (rc) return rc;
  return mfc_vjp_backward(s, req->cot, grad);
}

}  // extern "C"
