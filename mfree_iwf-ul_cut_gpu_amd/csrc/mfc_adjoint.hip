// mfc_adjoint.hip — reverse mode of the SPH cutting step on gfx950: gradients of the mean tool
// force over a window of steps with respect to the Johnson-Cook constants and the friction
// coefficient, theta = (A, B, n, C, m, mu).
//
// Tape: the state at the start of every K-th step (state block, positional block, stale cell
// ids of blanked particles, tool) is copied into a checkpoint. The backward pass walks the
// checkpoint segments from last to first; each segment is replayed forward with the step
// kernels to recover the start state of each of its steps, and the steps are then reversed
// one by one:
//   F1  binning          cell ids and the stable counting sort of the restored state
//   F2  predict          with the sorted copy of the start state (gather, predictor, EOS,
//                        artificial stress, records)
//   F3  sums             neighbour sums of every particle, kept from the segment's replay (the
//                        replayed interaction stores them, one slot per step of the segment)
//   B1  kb_adj_tail      VJP of everything after the sums, per particle (tail_vjp), and the
//                        per-tile theta partials of the step (kb_adj_theta adds them up)
//   B2  kb_adj_pull      VJP of the neighbour sums with respect to the records: each particle
//                        pulls the contributions of the pairs it takes part in over the same
//                        symmetric 5 x 5 cell window, in a fixed order (no atomics), the pairs
//                        of the neighbours' sums and of its own sums in one walk
//   B3  kb_adj_record    VJP of the record and predictor, then the scatter back to the
//                        storage order of the previous step (inverse of the sort permutation)
// The positional contact normal (read by the friction heat of the next step at the same sorted
// index) carries its adjoint from step to step without permutation, like its forward value.
//
// A group of instances that ran the same taped forward (one handle, or the instances of a
// batch) is reversed together: every phase, the checkpoint copies included, is one launch over
// the group's (instance, tile) work table, with per-instance descriptors in device memory, as
// the batched forward step runs. Each instance computes exactly what it computes alone.
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

}  // namespace

// a checkpoint of one instance: state block, positional block, cell ids, tool
struct snapshot {
  char *state;            // state block (state_block_bytes)
  char *pos;              // positional block
  int *keys;              // cell ids of the previous step (kept by blanked particles)
  mfc_tool_state *tool;
};

namespace {

int padded(int ntiles) { return ((ntiles + 7) / 8) * 8; }

static_assert(L_N == ACC_N, "a replay slot holds the interaction's neighbour sums in the tail's order");

// one instance of the backward pass as its kernels see it (device memory, one per instance)
struct adj_args {
  dev_consts k;
  int ntiles;
  const double *st;                 // sorted start state of the reversed step (TS_* order, stride N)
  const int *flags;                 // its flags
  dpositional pos;                  // replay positional block
  const double *rec;                // FP64 neighbour records
  const int *key_sorted, *offsets, *perm;
  const mfc_tool_state *tool_now;   // the tool after the reversed step's move
  double *acc_seg;                  // the replay's neighbour sums and secant results, one slot of
                                    // (ACC_N + 1) x N per step
  size_t acc_stride;                // doubles per slot
  double *lam_acc, *lam_s[2], *lam_n, *g_s, *g_rec, *theta_part, *theta_acc;
  double Fb[2];                     // cotangent over the window length
  // checkpoint copies: the live replay buffers and the instance's snapshots
  size_t state_words, pos_words;    // 8-byte words of a state block and of a positional block
  uint64_t *live_state[2];
  uint64_t *live_pos;
  int *live_keys;                   // the handle's sorted cell ids
  mfc_tool_state *live_tool;        // 2 replay tool states
  const snapshot *ckpt, *seg;       // checkpoints of the tape, and the start states of a segment
};

// an instance's descriptor read through the constant address space: the pointers loaded from it
// are then known to point to global memory, and the kernels use global instead of flat loads
// (a flat load also holds up every wait for an LDS or scalar load)
#define MFC_ADJ_ARGS(d, q) const __attribute__((address_space(4))) adj_args &d = \
    *(const __attribute__((address_space(4))) adj_args *)(D + (q))

// ------------------------------------------------------------------ checkpoint copies
// dir 0: a snapshot (ckpt[idx] or seg[idx]) into replay set 0, the positional block, the cell
// ids and replay tool 0; dir 1: replay set wa, the positional block, the cell ids and replay
// tool wa into seg[idx]. Grid (copy_blocks, instances).
__global__ __launch_bounds__(256) void kb_adj_copy(const adj_args *__restrict__ D, int dir, int from_ckpt, int idx,
                                                   int wa) {
  MFC_ADJ_ARGS(d, blockIdx.y);
  const snapshot sn = from_ckpt ? d.ckpt[idx] : d.seg[idx];
  const int lw = dir ? wa : 0;
  uint64_t *ls = d.live_state[lw];
  const uint64_t *src_s = dir ? ls : (const uint64_t *)sn.state;
  uint64_t *dst_s = dir ? (uint64_t *)sn.state : ls;
  const uint64_t *src_p = dir ? d.live_pos : (const uint64_t *)sn.pos;
  uint64_t *dst_p = dir ? (uint64_t *)sn.pos : d.live_pos;
  const int *src_k = dir ? d.live_keys : sn.keys;
  int *dst_k = dir ? sn.keys : d.live_keys;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (size_t q = i0; q < d.state_words; q += stride) dst_s[q] = src_s[q];
  for (size_t q = i0; q < d.pos_words; q += stride) dst_p[q] = src_p[q];
  for (size_t q = i0; q < (size_t)d.k.N; q += stride) dst_k[q] = src_k[q];
  if (i0 == 0) {
    if (dir) *sn.tool = d.live_tool[lw];
    else d.live_tool[0] = *sn.tool;
  }
}

// points the replay descriptors of both parities at the neighbour-sum slot of a step (one
// thread per instance; the next launches read the descriptors after this one completes)
__global__ __launch_bounds__(64) void kb_adj_acc_slot(inst_args *__restrict__ rep, const adj_args *__restrict__ D,
                                                      int n, int slot) {
  const int q = blockIdx.x * 64 + threadIdx.x;
  if (q >= n) return;
  double *a = D[q].acc_seg + (size_t)slot * D[q].acc_stride;
  rep[q].acc = a; rep[q].acc_out = a;
  rep[n + q].acc = a; rep[n + q].acc_out = a;
}

// ------------------------------------------------------------------ B1: per-particle tail
// st: sorted start state; acc: the step's neighbour sums (slot of the segment's replay);
// lam_s[cur]: adjoint of the end state of the step (same layout); lam_n: adjoint of the positional
// normal written by this step (in) and of the one read by it (out); in_win: the step's force is
// in the loss window
__global__ __launch_bounds__(TILE) void kb_adj_tail(const adj_args *__restrict__ D, const int2 *__restrict__ work,
                                                    int nwork, int cur, int in_win, int slot) {
  const int w = xcd_tile(blockIdx.x, nwork);
  if (w < 0) return;
  const int2 it = work[w];
  MFC_ADJ_ARGS(d, it.x);
  const dev_consts &k = D[it.x].k;
  const int t = it.y;
  const int i = t * TILE + threadIdx.x;
  const size_t N = k.N;
  const double *__restrict__ acc = d.acc_seg + (size_t)slot * d.acc_stride;
  double th[6] = {0., 0., 0., 0., 0., 0.};
  if (i < k.N) {
    tail_in in;
#pragma unroll
    for (int q = 0; q < TS_N; q++) in.s[q] = d.st[q * N + i];
    in.p = d.pos.p[i];
    in.nx = d.pos.nx[i];
    in.ny = d.pos.ny[i];
    in.flags = d.flags[i];
    in.pse_flag_zero = k.pse_quirk && i >= k.N / 2;
#pragma unroll
    for (int q = 0; q < L_N; q++) in.acc[q] = acc[q * N + i];
    in.dl_rec = acc[L_N * N + i];
    const double *lam_s = d.lam_s[cur];
    double ob[TS_N];
#pragma unroll
    for (int q = 0; q < TS_N; q++) ob[q] = lam_s[q * N + i];
    const double nbx = d.lam_n[i], nby = d.lam_n[N + i];
    tail_grad g;
    tail_vjp(k, *d.tool_now, in, ob, nbx, nby, in_win ? d.Fb[0] : 0., in_win ? d.Fb[1] : 0., g, nullptr);
#pragma unroll
    for (int q = 0; q < TS_N; q++) d.g_s[q * N + i] = g.s[q];
#pragma unroll
    for (int q = 0; q < L_N; q++) d.lam_acc[q * N + i] = g.acc[q];
    d.lam_n[i] = g.nx;
    d.lam_n[N + i] = g.ny;
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
    for (int q = 0; q < 6; q++) d.theta_part[6 * t + q] = red[q][0];
  }
}

// adds the step's theta partials of instance blockIdx.x to its running totals: lane l sums the
// tiles l, l + TILE, ... in ascending order, lane sums reduced by halving strides
__global__ __launch_bounds__(TILE) void kb_adj_theta(const adj_args *__restrict__ D) {
  MFC_ADJ_ARGS(d, blockIdx.x);
  __shared__ double red[6][TILE];
  double a[6] = {0., 0., 0., 0., 0., 0.};
  for (int u = threadIdx.x; u < d.ntiles; u += TILE)
#pragma unroll
    for (int q = 0; q < 6; q++) a[q] = a[q] + d.theta_part[6 * u + q];
#pragma unroll
  for (int q = 0; q < 6; q++) red[q][threadIdx.x] = a[q];
  for (int s = TILE / 2; s > 0; s >>= 1) {
    __syncthreads();
    if (threadIdx.x < s) {
#pragma unroll
      for (int q = 0; q < 6; q++) red[q][threadIdx.x] = red[q][threadIdx.x] + red[q][threadIdx.x + s];
    }
  }
  if (threadIdx.x < 6) d.theta_acc[threadIdx.x] = d.theta_acc[threadIdx.x] + red[threadIdx.x][0];
}

// ------------------------------------------------------------------ B2: pull over pairs
// For each window member a of particle b, both kinds of pair that link them: the pairs (a, b) of
// a's sums that include b, with the adjoint of a's sums, and the pairs (b, a) of b's own sums,
// with b's adjoint; a's record is loaded once for both, and g_rec[b] is written. (Two passes,
// one per kind, kept fewer records live per lane at the same 2 waves per SIMD and walked every
// window twice: 71 + 62 us per step against 113 us for the one pass on the 37.5k cut.)
// PULL_LANES consecutive lanes share a particle: lane r takes the window candidates r,
// r + PULL_LANES, ... (numbered column by column, as the forward walk numbers them), and the
// lanes' partial sums are added by a butterfly of shuffles, which gives every lane the same bits
// (((g0 + g1) + (g2 + g3)) + ...). The pass is latency-bound (a long dependent chain of record
// loads and pair adjoints per particle, 2 waves per SIMD at its 256 VGPRs; forcing 3 waves
// spills 336 bytes per lane and takes 40 % longer), so more lanes per particle shorten the
// chain. The count is fixed, so that an instance's gradient has the same bits alone and in a
// batch. Measured backward pass, ms per step with 2 / 4 / 8 lanes: 6k particles 0.357 / 0.317 /
// 0.286, 37.5k 0.340 / 0.313 / 0.295, the 32-instance batch of 577k 1.80 / 1.88 / 1.92 (the
// calibration cuts are the ones the reverse mode is for). A block takes the PULL_LANES-th part
// of a tile: grid padded(ntiles * PULL_LANES).
constexpr int PULL_LANES = 8;

__global__ __launch_bounds__(TILE) void kb_adj_pull(const adj_args *__restrict__ D, const int2 *__restrict__ work,
                                                    int nwork) {
  const int u = xcd_tile(blockIdx.x, nwork * PULL_LANES);
  if (u < 0) return;
  const int2 it = work[u / PULL_LANES];
  MFC_ADJ_ARGS(d, it.x);
  const dev_consts &k = D[it.x].k;
  const int r = threadIdx.x % PULL_LANES;
  const int b = it.y * TILE + (u % PULL_LANES) * (TILE / PULL_LANES) + threadIdx.x / PULL_LANES;
  const bool live = b < k.N;  // the lanes of a particle stay together for the shuffles
  const size_t N = k.N;
  const double *__restrict__ rec = d.rec;
  const int *__restrict__ flags = d.flags;
  const double *__restrict__ lam_acc = d.lam_acc;
  double *__restrict__ g_rec = d.g_rec;
  const bool pse = k.pse != 0;
  double rb[REC_DOUBLES], gb[REC_DOUBLES], lb[L_N], dummy[REC_DOUBLES];
  const int bb = live ? b : 0;
#pragma unroll
  for (int q = 0; q < REC_DOUBLES; q++) {
    rb[q] = rec[REC_DOUBLES * (size_t)bb + q];
    gb[q] = 0.;
  }
#pragma unroll
  for (int q = 0; q < L_N; q++) lb[q] = lam_acc[q * N + bb];
  const int fb = flags[bb];
  const bool blank_b = fb & MFC_FLAG_BLANKED, tool_b = fb & MFC_FLAG_TOOL;
  const double ta_b = pse ? 0. : (tool_b ? k.alpha_tool : k.alpha_wp);
  const bool mech_b = !blank_b && (!pse || !tool_b);
  const bool pse_b = !blank_b && pse && k.thermal_wp;
  const window w = make_window(k, d.offsets, d.key_sorted[bb]);
  const int total = live ? w.total : 0;
  for (int t = r; t < total; t += PULL_LANES) {
    const int a = w.at(t);
    if (a == b) continue;
    const int fa = flags[a];
    const bool blank_a = fa & MFC_FLAG_BLANKED, tool_a = fa & MFC_FLAG_TOOL;
    const bool mech_a = !blank_a && (!pse || !tool_a);
    const bool pse_a = !blank_a && pse && k.thermal_wp;
    const double *pa = rec + REC_DOUBLES * (size_t)a;
    const double dx = pa[R_X] - rb[R_X], dy = pa[R_Y] - rb[R_Y];
    const double rr2 = dx * dx + dy * dy;
    const double ha = pa[R_H], hb = rb[R_H];
    const bool mech1 = mech_a && !blank_b && rr2 < 4 * ha * ha && rr2 >= 1e-8;  // a's sums include b
    const bool mech2 = mech_b && !blank_a && rr2 < 4 * hb * hb && rr2 >= 1e-8;  // b's sums include a
    if (!mech1 && !pse_a && !mech2 && !pse_b) continue;
    double ra[REC_DOUBLES];
#pragma unroll
    for (int q = 0; q < REC_DOUBLES; q++) ra[q] = pa[q];
    if (mech1 || pse_a) {
      double la[L_N];
#pragma unroll
      for (int q = 0; q < L_N; q++) la[q] = lam_acc[q * N + a];
      const double ta_a = pse ? 0. : (tool_a ? k.alpha_tool : k.alpha_wp);
      if (mech1) pair_vjp(k, ra, rb, tool_a, tool_b, ta_a, la, false, true, dummy, gb);
      if (pse_a) pse_vjp(k, ra, rb, la[L_TPSE], false, true, dummy, gb);
    }
    if (mech2) pair_vjp(k, rb, ra, tool_b, tool_a, ta_b, lb, true, false, gb, dummy);
    if (pse_b) pse_vjp(k, rb, ra, lb[L_TPSE], true, false, gb, dummy);
  }
#pragma unroll
  for (int st = 1; st < PULL_LANES; st <<= 1)
#pragma unroll
    for (int q = 0; q < REC_DOUBLES; q++) gb[q] = gb[q] + __shfl_xor(gb[q], st, 64);
  if (live && r == 0) {
#pragma unroll
    for (int q = 0; q < REC_DOUBLES; q++) g_rec[q * N + b] = gb[q];
  }
}

// ------------------------------------------------------------------ B3: record VJP + unsort
__global__ __launch_bounds__(TILE) void kb_adj_record(const adj_args *__restrict__ D, const int2 *__restrict__ work,
                                                      int nwork, int cur) {
  const int w = xcd_tile(blockIdx.x, nwork);
  if (w < 0) return;
  const int2 it = work[w];
  MFC_ADJ_ARGS(d, it.x);
  const dev_consts &k = D[it.x].k;
  const int a = it.y * TILE + threadIdx.x;
  if (a >= k.N) return;
  const size_t N = k.N;
  double s[TS_N], gs[TS_N], br[REC_DOUBLES];
#pragma unroll
  for (int q = 0; q < TS_N; q++) { s[q] = d.st[q * N + a]; gs[q] = d.g_s[q * N + a]; }
#pragma unroll
  for (int q = 0; q < REC_DOUBLES; q++) br[q] = d.g_rec[q * N + a];
  record_vjp(k, s, d.flags[a], d.pos.p[a], br, gs);
  const int dst = d.perm[a];
  double *lam_prev = d.lam_s[1 - cur];
#pragma unroll
  for (int q = 0; q < TS_N; q++) lam_prev[q * N + dst] = gs[q];
}


}  // namespace

// ------------------------------------------------------------------ workspace
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
  double *acc_seg = nullptr, *lam_acc = nullptr, *lam_s[2] = {nullptr, nullptr}, *lam_n = nullptr, *g_s = nullptr,
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
        (rc = w->alloc(&w->acc_seg, (size_t)K * (L_N + 1) * n)) || (rc = w->alloc(&w->lam_acc, L_N * n)) ||
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

int adj_backward(mfc_sim *const *sims, int n, const batch_work &fw, hipStream_t st, const double *cot, double *grad) {
  if (n < 1) return fail(MFC_E_INVALID, "empty group");
  const adj_work *w0 = sims[0]->adj;
  for (int i = 0; i < n; i++) {
    const adj_work *w = sims[i]->adj;
    if (!w || !w->valid) return fail(MFC_E_INVALID, "no recorded run: call mfc_vjp_forward first");
    if (w->nsteps != w0->nsteps || w->K != w0->K || w->w0 != w0->w0 || w->w1 != w0->w1)
      return fail(MFC_E_INVALID, "the instances recorded different runs");
  }
  const int nsteps = w0->nsteps, K = w0->K;
  const int nck = (nsteps + K - 1) / K;
  // descriptors: replay steps of start parity 0 and 1 (inst_args, the step kernels' view, with
  // the replay buffers as state and no binning of the next step), the backward kernels' view
  // (adj_args), and each instance's checkpoints and segment snapshots
  std::vector<inst_args> rep(2 * (size_t)n);
  std::vector<adj_args> da(n);
  std::vector<snapshot> snaps((size_t)n * (nck + K));
  for (int i = 0; i < n; i++) {
    mfc_sim *s = sims[i];
    adj_work *w = s->adj;
    const int N = s->k.N;
    dstate A, B;
    dpositional P;
    carve_state(w->work[0], N, A);
    carve_state(w->work[1], N, B);
    carve_positional(w->posw, N, P);
    for (int p = 0; p < 2; p++) {
      inst_args &a = rep[(size_t)p * n + i];
      std::memset(&a, 0, sizeof(a));
      a.k = s->k;
      a.cur = p ? B : A;
      a.nxt = p ? A : B;
      a.b = s->bins;
      a.pos = P;
      a.r = s->recs;
      a.tool_in = w->toolw + p;
      a.tool_out = w->toolw + 1 - p;
      a.f = {w->partials, w->force_last, w->hist, 0, w->ctr, w->pending};  // no history, counted
      a.acc = a.acc_out = w->acc_seg;  // pointed at each step's slot by kb_adj_acc_slot
      a.ntiles = s->ntiles;
      a.nscan = scan_blocks(s->k.ncell_table);
      a.bin_next = 0;
    }
    adj_args &d = da[i];
    std::memset(&d, 0, sizeof(d));
    d.k = s->k;
    d.ntiles = s->ntiles;
    d.st = B.px;  // the sorted start state, TS_* order with stride N
    d.flags = B.flags;
    d.pos = P;
    d.rec = s->recs.rec;
    d.key_sorted = s->bins.key_sorted;
    d.offsets = s->bins.offsets;
    d.perm = s->bins.perm;
    d.tool_now = w->toolw + 1;
    d.acc_seg = w->acc_seg;
    d.acc_stride = (size_t)(L_N + 1) * N;
    d.lam_acc = w->lam_acc;
    d.lam_s[0] = w->lam_s[0];
    d.lam_s[1] = w->lam_s[1];
    d.lam_n = w->lam_n;
    d.g_s = w->g_s;
    d.g_rec = w->g_rec;
    d.theta_part = w->theta_part;
    d.theta_acc = w->theta_acc;
    d.Fb[0] = cot[2 * i] / (w->w1 - w->w0);
    d.Fb[1] = cot[2 * i + 1] / (w->w1 - w->w0);
    d.state_words = state_block_bytes(N) / sizeof(uint64_t);
    d.pos_words = positional_block_bytes(N) / sizeof(uint64_t);
    d.live_state[0] = (uint64_t *)w->work[0];
    d.live_state[1] = (uint64_t *)w->work[1];
    d.live_pos = (uint64_t *)w->posw;
    d.live_keys = s->bins.key_sorted;
    d.live_tool = w->toolw;
    for (int c = 0; c < nck; c++) snaps[(size_t)i * (nck + K) + c] = w->ckpt[c];
    for (int c = 0; c < K; c++) snaps[(size_t)i * (nck + K) + nck + c] = w->seg[c];
  }
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t b_rep = up(sizeof(inst_args) * rep.size()), b_da = up(sizeof(adj_args) * da.size());
  const size_t b_sn = up(sizeof(snapshot) * snaps.size());
  char *dmem = nullptr;
  HIP_TRY(hipMalloc(&dmem, b_rep + b_da + b_sn));
  inst_args *d_rep = (inst_args *)dmem;
  adj_args *d_da = (adj_args *)(dmem + b_rep);
  snapshot *d_sn = (snapshot *)(dmem + b_rep + b_da);
  for (int i = 0; i < n; i++) {
    da[i].ckpt = d_sn + (size_t)i * (nck + K);
    da[i].seg = da[i].ckpt + nck;
  }
  // copy grid: about 8 words of the largest state block per thread
  int maxN = 0;
  for (int i = 0; i < n; i++) maxN = sims[i]->k.N > maxN ? sims[i]->k.N : maxN;
  const size_t want = (state_block_bytes(maxN) / sizeof(uint64_t) + 2047) / 2048;
  const int copy_blocks = want < 16 ? 16 : (want > 1024 ? 1024 : (int)want);
  int rc = MFC_OK;
  auto run = [&]() -> int {
    HIP_TRY(hipMemcpyAsync(d_rep, rep.data(), sizeof(inst_args) * rep.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d_da, da.data(), sizeof(adj_args) * da.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d_sn, snaps.data(), sizeof(snapshot) * snaps.size(), hipMemcpyHostToDevice, st));
    for (int i = 0; i < n; i++) {
      mfc_sim *s = sims[i];
      adj_work *w = s->adj;
      const size_t N = (size_t)s->k.N;
      HIP_TRY(hipMemsetAsync(w->lam_s[0], 0, sizeof(double) * TS_N * N, st));
      HIP_TRY(hipMemsetAsync(w->lam_n, 0, sizeof(double) * 2 * N, st));
      HIP_TRY(hipMemsetAsync(w->theta_acc, 0, sizeof(double) * 6, st));
      HIP_TRY(hipMemcpyAsync(w->keys_keep, s->bins.key_sorted, sizeof(int) * N, hipMemcpyDeviceToDevice, st));
      // every binning of the pass counts into cleared chunk sums and leaves them cleared; the
      // handle's chunk sums hold the next step's counts: clear them, and bin the handle's state
      // again at its next step
      HIP_TRY(hipMemsetAsync(s->bins.bsum, 0, sizeof(int) * (size_t)scan_blocks(s->k.ncell_table), st));
      s->keys_fresh = false;
    }
    const batch_work rw[2] = {{d_rep, n, fw.tiles, fw.ntiles, fw.chunks, fw.nchunks, false, fw.v, false, false},
                              {d_rep + n, n, fw.tiles, fw.ntiles, fw.chunks, fw.nchunks, false, fw.v, false, false}};
    const dim3 gcopy(copy_blocks, n), gt(padded(fw.ntiles)), gpull(padded(fw.ntiles * PULL_LANES));
    const dim3 gslot((n + 63) / 64);
    int cur = 0;  // lam_s[cur] holds the adjoint of the end state of the step being reversed
    const int nseg = nck;
    for (int sg = nseg - 1; sg >= 0; sg--) {
      const int t0 = sg * K, t1 = (t0 + K < nsteps) ? t0 + K : nsteps;
      // replay the segment, keeping the start of each of its steps
      int wa = 0;
      hipLaunchKernelGGL(kb_adj_copy, gcopy, dim3(256), 0, st, d_da, 0, 1, sg, 0);
      for (int t = t0; t < t1; t++) {
        hipLaunchKernelGGL(kb_adj_copy, gcopy, dim3(256), 0, st, d_da, 1, 0, t - t0, wa);
        hipLaunchKernelGGL(kb_adj_acc_slot, gslot, dim3(64), 0, st, d_rep, d_da, n, t - t0);
        launch_batch_hash(rw[wa], st);
        launch_batch_step(rw[wa], st);  // the interaction keeps its sums in the step's slot
        wa = 1 - wa;
      }
      // reverse the steps of the segment
      for (int t = t1 - 1; t >= t0; t--) {
        hipLaunchKernelGGL(kb_adj_copy, gcopy, dim3(256), 0, st, d_da, 0, 0, t - t0, 0);
        launch_batch_hash(rw[0], st);
        launch_batch_binning(rw[0], st);
        launch_batch_predict_gather(rw[0], st);
        const int in_win = t >= w0->w0 && t < w0->w1;
        hipLaunchKernelGGL(kb_adj_tail, gt, dim3(TILE), 0, st, d_da, fw.tiles, fw.ntiles, cur, in_win, t - t0);
        hipLaunchKernelGGL(kb_adj_theta, dim3(n), dim3(TILE), 0, st, d_da);
        hipLaunchKernelGGL(kb_adj_pull, gpull, dim3(TILE), 0, st, d_da, fw.tiles, fw.ntiles);
        hipLaunchKernelGGL(kb_adj_record, gt, dim3(TILE), 0, st, d_da, fw.tiles, fw.ntiles, cur);
        cur = 1 - cur;
      }
    }
    HIP_TRY(hipGetLastError());
    for (int i = 0; i < n; i++) {
      mfc_sim *s = sims[i];
      adj_work *w = s->adj;
      // the replay overwrote the handle's cell ids, which blanked particles keep into the next step
      HIP_TRY(hipMemcpyAsync(s->bins.key_sorted, w->keys_keep, sizeof(int) * (size_t)s->k.N, hipMemcpyDeviceToDevice,
                             st));
      HIP_TRY(hipMemcpyAsync(grad + 6 * i, w->theta_acc, sizeof(double) * 6, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    return MFC_OK;
  };
  rc = run();
  if (rc) (void)hipStreamSynchronize(st);  // nothing queued may still read the descriptors
  (void)hipFree(dmem);
  return rc;
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
  if (s->batched) return fail(MFC_E_INVALID, "instance of a batch: use mfc_batch_vjp_backward");
  return adj_backward(&s, 1, s->work(s->cur), s->stream, cot, grad);
}

int mfc_vjp(mfc_sim *s, mfc_vjp_req *req, double grad[6]) {
  if (!s || !req || !grad) return fail(MFC_E_INVALID, "null argument");
  int rc = mfc_vjp_forward(s, req->nsteps, req->checkpoint_every, req->w0, req->w1, req->Fmean);
  if (rc) return rc;
  return mfc_vjp_backward(s, req->cot, grad);
}

}  // extern "C"
