// mfc_api.hip — C ABI of the solver (declared in include/mfc.h).
//
// A handle owns every device buffer of one simulation, its own HIP stream (default priority)
// and the hipGraphs of GRAPH_STEPS consecutive steps. Steps are queued asynchronously; calls that
// return data synchronise the handle's stream.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mfc_sim.h"

namespace {

thread_local std::string g_err;

}  // namespace

int mfc::fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

namespace {
using mfc::fail;

int field_components(int f) {
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

static double **state_fields(mfc::dstate &d, int q) {
  double **dp[mfc::DSTATE_DOUBLES] = {&d.px, &d.py, &d.vx, &d.vy, &d.sxx, &d.sxy, &d.syy, &d.szz, &d.rho,
                                      &d.T, &d.eps, &d.epsd, &d.h, &d.fcx, &d.fcy, &d.ftx, &d.fty, &d.ptx,
                                      &d.pty, &d.vtx, &d.vty, &d.rhot, &d.stxx, &d.stxy, &d.styy, &d.stzz, &d.Tt};
  return dp[q];
}

size_t state_block_bytes(int N) { return (size_t)N * (mfc::DSTATE_DOUBLES * sizeof(double) + 2 * sizeof(int)); }

void carve_state(void *base, int N, mfc::dstate &d) {
  double *b = (double *)base;
  for (int q = 0; q < mfc::DSTATE_DOUBLES; q++) *state_fields(d, q) = b + (size_t)q * N;
  d.flags = (int *)(b + (size_t)mfc::DSTATE_DOUBLES * N);
  d.id = d.flags + N;
}

size_t positional_block_bytes(int N) { return (size_t)N * 6 * sizeof(double); }

void carve_positional(void *base, int N, mfc::dpositional &p) {
  double *b = (double *)base;
  p.nx = b; p.ny = b + N; p.p = b + 2 * (size_t)N;
  p.rxx = b + 3 * (size_t)N; p.rxy = b + 4 * (size_t)N; p.ryy = b + 5 * (size_t)N;
}

void enqueue_step(mfc_sim *s, int c) { mfc::launch_batch_step(s->work(c), s->stream); }

int ensure_keys(mfc_sim *s) {
  if (s->keys_fresh) return MFC_OK;
  HIP_TRY(hipMemsetAsync(s->bins.bsum, 0, sizeof(int) * (size_t)mfc::scan_blocks(s->k.ncell_table), s->stream));
  mfc::launch_batch_hash(s->work(s->cur), s->stream);
  s->keys_fresh = true;
  return MFC_OK;
}

static_assert(mfc::XCD_CHUNK_MAX == MFC_XCD_CHUNK_MAX, "XCD chunk bound of the header and the kernels");

int check_tuning(const mfc_tuning *t) {
  if ((t->split_interact != 0 && t->split_interact != 1) || (t->compact_pairs != 0 && t->compact_pairs != 1))
    return fail(MFC_E_INVALID, "split_interact and compact_pairs must be 0 or 1");
  if (t->stage_limit < 0 || t->stage_limit > MFC_STAGE_CAP)
    return fail(MFC_E_INVALID, "stage_limit must be in 0.." + std::to_string(MFC_STAGE_CAP));
  if (t->xcd_chunk < 0 || t->xcd_chunk > mfc::XCD_CHUNK_MAX)
    return fail(MFC_E_INVALID, "xcd_chunk must be in 0.." + std::to_string(mfc::XCD_CHUNK_MAX));
  if (t->sort_fast < 0 || t->sort_fast > MFC_SORT_FAST_MAX)
    return fail(MFC_E_INVALID, "sort_fast must be in 0.." + std::to_string(MFC_SORT_FAST_MAX));
  if (t->sort_list < 0) return fail(MFC_E_INVALID, "sort_list must be >= 0");
  return MFC_OK;
}

// each step's tool force is reduced by the next step's first kernel; this reduces the last one
void finish_steps(mfc_sim *s) { mfc::launch_batch_force(s->work(s->cur), s->stream); }

mfc::inst_args instance_args(const mfc_sim *s, int c) {
  mfc::inst_args a;
  std::memset(&a, 0, sizeof(a));
  a.k = s->k;
  a.cur = s->set[c];
  a.nxt = s->set[1 - c];
  a.b = s->bins;
  a.pos = s->pos;
  a.r = s->recs;
  a.tool_in = s->tool + c;
  a.tool_out = s->tool + (1 - c);
  a.f = s->force();
  a.acc = s->acc;
  a.geo = s->geo;
  a.ntiles = s->ntiles;
  a.nscan = mfc::scan_blocks(s->k.ncell_table);
  a.bin_next = 1;
  return a;
}

int current_geometry(const mfc_sim *s, mfc::grid_geo &g) {
  const mfc::sim_constants &c = s->sc.k;
  g = {c.bbmin_x, c.bbmin_y, c.cell, c.gnx, c.gny, c.num_cell, 0};
  if (!s->k.adaptive) return MFC_OK;
  HIP_TRY(hipMemcpyAsync(&g, s->geo, sizeof(g), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return MFC_OK;
}

namespace {

int alloc_state(mfc_sim *s, mfc::dstate &d, int N) {
  char *base = nullptr;
  int rc = s->alloc(&base, state_block_bytes(N));
  if (rc) return rc;
  carve_state(base, N, d);
  return MFC_OK;
}

// captures GRAPH_STEPS steps from buffer parity `parity`. The capture is always ended, and the
// handle keeps the executable graph only when capture and instantiation both succeeded (a
// failed instantiation leaves no half-made graph behind for a later launch).
int build_graph(mfc_sim *s, int parity) {
  HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  for (int q = 0; q < GRAPH_STEPS; q++) enqueue_step(s, (parity + q) & 1);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s->stream, &g);
  if (e != hipSuccess || g == nullptr) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return fail(MFC_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  }
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(MFC_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  s->graph[parity] = x;
  return MFC_OK;
}

int copy_out(mfc_sim *s, void *dst, const void *src, size_t bytes) {
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return MFC_OK;
}

}  // namespace

extern "C" {

const char *mfc_last_error(void) { return g_err.c_str(); }

int mfc_field_components(int field) { return field_components(field); }

int mfc_config_ref_cut(mfc_config *cfg, int ny, double rake, double clear, double chamfer, double speed, double feed) {
  if (!cfg) return fail(MFC_E_INVALID, "null config");
  mfc::config_ref_cut(*cfg, ny, rake, clear, chamfer, speed, feed);
  return MFC_OK;
}

int mfc_config_rings(mfc_config *cfg, int nbox) {
  if (!cfg) return fail(MFC_E_INVALID, "null config");
  mfc::config_rings(*cfg, nbox);
  return MFC_OK;
}

int mfc_config_impact(mfc_config *cfg, int nbox) {
  if (!cfg) return fail(MFC_E_INVALID, "null config");
  mfc::config_impact(*cfg, nbox);
  return MFC_OK;
}

void mfc_destroy(mfc_sim *s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (int p = 0; p < 2; p++)
    if (s->graph[p]) (void)hipGraphExecDestroy(s->graph[p]);
  mfc::adj_free(s->adj);
  for (void *q : s->allocs) (void)hipFree(q);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

int mfc_create(const mfc_config *cfg, mfc_sim **out) {
  if (!cfg || !out) return fail(MFC_E_INVALID, "null argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(MFC_E_HIP, "no HIP device");
  mfc_sim *s = new mfc_sim();
  s->cfg = *cfg;
  std::string err;
  int rc = mfc::build_scenario(*cfg, s->sc, err);
  if (rc) { delete s; return fail(rc, err); }
  const mfc::sim_constants &c = s->sc.k;
  mfc::dev_consts &k = s->k;
  mfc::make_dev_consts(c, k);
  const int N = c.N;
  s->ntiles = (N + mfc::TILE - 1) / mfc::TILE;
  s->precision = cfg->precision;
  s->recs.f32 = cfg->precision == MFC_PREC_FP32C;

  auto bail = [&](int code) { mfc_destroy(s); return code; };
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) return bail(fail(MFC_E_HIP, "stream"));
  for (int q = 0; q < 2; q++)
    if ((rc = alloc_state(s, s->set[q], N))) return bail(rc);
  mfc::dbins &b = s->bins;
  b.ccap = mfc::chunk_list_cap(N);
  if ((rc = s->alloc(&b.key, N)) || (rc = s->alloc(&b.key_sorted, N)) ||
      (rc = s->alloc(&b.offsets, (size_t)c.ncell_table + 1)) ||
      (rc = s->alloc(&b.bucket, N)) || (rc = s->alloc(&b.perm, N)) || (rc = s->alloc(&b.spos, N)) ||
      (rc = s->alloc(&b.bsum, (mfc::scan_blocks(c.ncell_table) + 3) & ~3)) || (rc = s->alloc(&b.clamped, 1)) ||
      (rc = s->alloc(&b.clist, (size_t)mfc::scan_blocks(c.ncell_table) * mfc::chunk_list_cap(N))) ||
      (rc = s->alloc(&s->geo, 1)) || (c.rothlin && (rc = s->alloc(&b.offsets_r, (size_t)c.ncell_table + 1))))
    return bail(rc);
  mfc::dpositional &p = s->pos;
  {
    char *pb = nullptr;
    if ((rc = s->alloc(&pb, positional_block_bytes(N)))) return bail(rc);
    carve_positional(pb, N, p);
  }
  if ((rc = s->alloc(&s->recs.posd, N)) || (rc = s->alloc(&s->recs.vfl, N))) return bail(rc);
  if (s->recs.f32) {
    if ((rc = s->alloc(&s->recs.rec32, (size_t)16 * N))) return bail(rc);
  } else if ((rc = s->alloc(&s->recs.rec, (size_t)mfc::REC * N))) {
    return bail(rc);
  }
  if ((rc = s->alloc(&s->tool, 2)) ||
      (rc = s->alloc(&s->partials, (size_t)16 * s->ntiles)) || (rc = s->alloc(&s->force_last, 2)) ||
      (rc = s->alloc(&s->hist, (size_t)2 * HIST_CAP)) || (rc = s->alloc(&s->hist_ctr, 1)) ||
      (rc = s->alloc(&s->force_pending, 1)) ||
      (rc = s->alloc(&s->acc, (size_t)mfc::ACC_N * N)))
    return bail(rc);
  // zero everything, then upload the initial state (particle_gpu.cu:20-125)
  {
    std::vector<double> zero(N, 0.);
    std::vector<int> ids(N);
    for (int i = 0; i < N; i++) ids[i] = i;
    auto up = [&](void *dst, const void *src, size_t bytes) {
      return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->stream);
    };
    for (int q = 0; q < 2; q++) {
      mfc::dstate &d = s->set[q];
      double *all[mfc::DSTATE_DOUBLES] = {d.px, d.py, d.vx, d.vy, d.sxx, d.sxy, d.syy, d.szz, d.rho,
                                          d.T, d.eps, d.epsd, d.h, d.fcx, d.fcy, d.ftx, d.fty, d.ptx,
                                          d.pty, d.vtx, d.vty, d.rhot, d.stxx, d.stxy, d.styy, d.stzz, d.Tt};
      for (double *a : all)
        if (up(a, zero.data(), sizeof(double) * N) != hipSuccess) return bail(fail(MFC_E_HIP, "upload"));
      if (up(d.flags, zero.data(), sizeof(int) * N) != hipSuccess) return bail(fail(MFC_E_HIP, "upload"));
      if (up(d.id, ids.data(), sizeof(int) * N) != hipSuccess) return bail(fail(MFC_E_HIP, "upload"));
    }
    mfc::dstate &d = s->set[0];
    const mfc::scenario &S = s->sc;
    if (up(d.px, S.px.data(), sizeof(double) * N) || up(d.py, S.py.data(), sizeof(double) * N) ||
        up(d.rho, S.rho.data(), sizeof(double) * N) || up(d.h, S.h.data(), sizeof(double) * N) ||
        up(d.T, S.T.data(), sizeof(double) * N) || up(d.flags, S.flags.data(), sizeof(int) * N))
      return bail(fail(MFC_E_HIP, "upload"));
    if (!S.vx.empty() && (up(d.vx, S.vx.data(), sizeof(double) * N) || up(d.vy, S.vy.data(), sizeof(double) * N)))
      return bail(fail(MFC_E_HIP, "upload"));
    double *pz[6] = {p.nx, p.ny, p.p, p.rxx, p.rxy, p.ryy};
    for (double *a : pz)
      if (up(a, zero.data(), sizeof(double) * N) != hipSuccess) return bail(fail(MFC_E_HIP, "upload"));
    if (hipMemsetAsync(b.key_sorted, 0, sizeof(int) * N, s->stream) != hipSuccess ||
        hipMemsetAsync(b.offsets, 0, sizeof(int) * ((size_t)c.ncell_table + 1), s->stream) != hipSuccess ||
        (b.offsets_r && hipMemsetAsync(b.offsets_r, 0, sizeof(int) * ((size_t)c.ncell_table + 1), s->stream) != hipSuccess) ||
        hipMemsetAsync(b.clamped, 0, sizeof(unsigned long long), s->stream) != hipSuccess ||
        hipMemsetAsync(s->hist_ctr, 0, sizeof(int64_t), s->stream) != hipSuccess ||
        hipMemsetAsync(s->force_pending, 0, sizeof(int), s->stream) != hipSuccess ||
        hipMemsetAsync(b.bsum, 0, sizeof(int) * (size_t)mfc::scan_blocks(c.ncell_table), s->stream) != hipSuccess ||
        hipMemsetAsync(s->force_last, 0, 2 * sizeof(double), s->stream) != hipSuccess ||
        (s->recs.rec && hipMemsetAsync(s->recs.rec, 0, sizeof(double) * mfc::REC * (size_t)N, s->stream) != hipSuccess) ||
        hipMemsetAsync(s->recs.posd, 0, sizeof(double2) * (size_t)N, s->stream) != hipSuccess ||
        hipMemsetAsync(s->recs.vfl, 0, sizeof(int) * (size_t)N, s->stream) != hipSuccess ||
        (s->recs.rec32 && hipMemsetAsync(s->recs.rec32, 0, sizeof(float) * 16 * (size_t)N, s->stream) != hipSuccess))
      return bail(fail(MFC_E_HIP, "memset"));
    mfc_tool_state t2[2] = {S.tool, S.tool};
    if (up(s->tool, t2, sizeof(t2)) != hipSuccess) return bail(fail(MFC_E_HIP, "upload tool"));
    const mfc::grid_geo g0 = {c.bbmin_x, c.bbmin_y, c.cell, c.gnx, c.gny, c.num_cell, 0};
    if (up(s->geo, &g0, sizeof(g0)) != hipSuccess) return bail(fail(MFC_E_HIP, "upload grid"));
    // descriptors and work tables of the step kernels (a batch of one)
    const int nscan = mfc::scan_blocks(c.ncell_table);
    std::vector<int2> tiles(s->ntiles), chunks(nscan);
    for (int t = 0; t < s->ntiles; t++) tiles[t] = make_int2(0, t);
    for (int q = 0; q < nscan; q++) chunks[q] = make_int2(0, q);
    s->nchunks = nscan;
    mfc::inst_args a2[2] = {instance_args(s, 0), instance_args(s, 1)};
    if ((rc = s->alloc(&s->dargs, 2)) || (rc = s->alloc(&s->dtiles, (size_t)s->ntiles)) ||
        (rc = s->alloc(&s->dchunks, (size_t)nscan)))
      return bail(rc);
    if (up(s->dargs, a2, sizeof(a2)) != hipSuccess ||
        up(s->dtiles, tiles.data(), sizeof(int2) * tiles.size()) != hipSuccess ||
        up(s->dchunks, chunks.data(), sizeof(int2) * chunks.size()) != hipSuccess)
      return bail(fail(MFC_E_HIP, "upload of the step tables"));
    if (hipStreamSynchronize(s->stream) != hipSuccess) return bail(fail(MFC_E_HIP, "sync"));
  }
  *out = s;
  return MFC_OK;
}

int mfc_set_tuning(mfc_sim *s, const mfc_tuning *t) {
  if (!s || !t) return fail(MFC_E_INVALID, "null argument");
  if (s->batched) return fail(MFC_E_INVALID, "instance of a batch: use mfc_batch_set_tuning");
  int rc = check_tuning(t);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(s->stream));
  s->tuning = *t;
  s->k.stage_limit = t->stage_limit;
  s->k.sort_fast = t->sort_fast;
  s->k.sort_list = t->sort_list;
  mfc::inst_args a2[2] = {instance_args(s, 0), instance_args(s, 1)};
  HIP_TRY(hipMemcpy(s->dargs, a2, sizeof(a2), hipMemcpyHostToDevice));
  for (int p = 0; p < 2; p++)
    if (s->graph[p]) {
      (void)hipGraphExecDestroy(s->graph[p]);
      s->graph[p] = nullptr;
    }
  return MFC_OK;
}

int mfc_get_tuning(const mfc_sim *s, mfc_tuning *t) {
  if (!s || !t) return fail(MFC_E_INVALID, "null argument");
  *t = s->tuning;
  return MFC_OK;
}

int mfc_step(mfc_sim *s, int nsteps) {
  if (!s || nsteps < 0) return fail(MFC_E_INVALID, "bad argument");
  if (s->batched) return fail(MFC_E_INVALID, "instance of a batch: step it with mfc_batch_step");
  int rc = ensure_keys(s);
  if (rc) return rc;
  int done = 0;
  if (s->use_graphs && nsteps >= GRAPH_STEPS) {
    while (nsteps - done >= GRAPH_STEPS) {
      int par = s->cur;
      if (!s->graph[par] && (rc = build_graph(s, par))) return rc;
      HIP_TRY(hipGraphLaunch(s->graph[par], s->stream));
      done += GRAPH_STEPS;
      s->step += GRAPH_STEPS;
    }
  }
  for (; done < nsteps; done++) {
    enqueue_step(s, s->cur);
    s->cur = 1 - s->cur;
    s->step++;
  }
  finish_steps(s);
  HIP_TRY(hipGetLastError());
  return MFC_OK;
}

int mfc_profile(mfc_sim *s, int nsteps, double *ms) {
  if (!s || !ms || nsteps < 1) return fail(MFC_E_INVALID, "bad argument");
  if (s->batched) return fail(MFC_E_INVALID, "instance of a batch: step it with mfc_batch_step");
  // ev: the phase boundaries on the stream; kt: the interaction kernels' own start and end
  constexpr int NE = MFC_PROFILE_PHASES + 1;
  hipEvent_t ev[NE], kt[4];
  for (int q = 0; q < NE; q++) HIP_TRY(hipEventCreate(&ev[q]));
  for (int q = 0; q < 4; q++) HIP_TRY(hipEventCreate(&kt[q]));
  double acc[MFC_PROFILE_PHASES] = {0., 0., 0., 0., 0.};
  int rc = ensure_keys(s);
  for (int it = 0; it < nsteps && rc == MFC_OK; it++) {
    const int c = s->cur, n = 1 - c;
    const mfc::batch_work w = s->work(c);
    (void)hipEventRecord(ev[0], s->stream);
    mfc::launch_batch_binning(w, s->stream);
    (void)hipEventRecord(ev[1], s->stream);
    mfc::launch_batch_predict(w, s->stream);
    (void)hipEventRecord(ev[2], s->stream);
    if (w.v.split) {
      mfc::launch_batch_sums(w, s->stream, kt[0], kt[1]);
      (void)hipEventRecord(ev[3], s->stream);
      mfc::launch_batch_update(w, s->stream, kt[2], kt[3]);
    } else {
      mfc::launch_batch_interact_fused(w, s->stream, kt[0], kt[1]);
      (void)hipEventRecord(ev[3], s->stream);
    }
    (void)hipEventRecord(ev[4], s->stream);
    mfc::launch_batch_force(w, s->stream);
    (void)hipEventRecord(ev[5], s->stream);
    s->cur = n;
    s->step++;
    if (hipEventSynchronize(ev[NE - 1]) != hipSuccess) { rc = fail(MFC_E_HIP, "event sync"); break; }
    float t[MFC_PROFILE_PHASES] = {0.f, 0.f, 0.f, 0.f, 0.f};
    if (hipEventElapsedTime(&t[0], ev[0], ev[1]) != hipSuccess ||
        hipEventElapsedTime(&t[1], ev[1], ev[2]) != hipSuccess ||
        hipEventElapsedTime(&t[2], kt[0], kt[1]) != hipSuccess ||
        (w.v.split && hipEventElapsedTime(&t[3], kt[2], kt[3]) != hipSuccess) ||
        hipEventElapsedTime(&t[4], ev[4], ev[5]) != hipSuccess) {
      rc = fail(MFC_E_HIP, "event time");
      break;
    }
    for (int q = 0; q < MFC_PROFILE_PHASES; q++) acc[q] += t[q];
  }
  for (int q = 0; q < NE; q++) (void)hipEventDestroy(ev[q]);
  for (int q = 0; q < 4; q++) (void)hipEventDestroy(kt[q]);
  if (rc) return rc;
  for (int q = 0; q < MFC_PROFILE_PHASES; q++) ms[q] = acc[q] / nsteps;
  return MFC_OK;
}

int mfc_sync(mfc_sim *s) {
  if (!s) return fail(MFC_E_INVALID, "null handle");
  HIP_TRY(hipStreamSynchronize(s->stream));
  return MFC_OK;
}

int mfc_tool_force(mfc_sim *s, double F[2]) {
  if (!s || !F) return fail(MFC_E_INVALID, "null argument");
  if (s->step == 0) { F[0] = 0.; F[1] = 0.; return MFC_OK; }
  return copy_out(s, F, s->force_last, 2 * sizeof(double));
}

int mfc_force_history(mfc_sim *s, double *F, int64_t cap, int64_t *n) {
  if (!s || !n || cap < 0) return fail(MFC_E_INVALID, "bad argument");
  int64_t ctr = 0;
  int rc = copy_out(s, &ctr, s->hist_ctr, sizeof(int64_t));
  if (rc) return rc;
  int64_t m = ctr < HIST_CAP ? ctr : HIST_CAP;
  if (!F) { *n = m; return MFC_OK; }
  if (m > cap) m = cap;
  if (m > 0) {
    rc = copy_out(s, F, s->hist, sizeof(double) * 2 * (size_t)m);
    if (rc) return rc;
  }
  *n = m;
  return MFC_OK;
}

int mfc_force_history_reset(mfc_sim *s) {
  if (!s) return fail(MFC_E_INVALID, "null handle");
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemsetAsync(s->hist_ctr, 0, sizeof(int64_t), s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return MFC_OK;
}

int mfc_get_info(const mfc_sim *s, mfc_info *info) {
  if (!s || !info) return fail(MFC_E_INVALID, "null argument");
  const mfc::sim_constants &c = s->sc.k;
  std::memset(info, 0, sizeof(*info));
  info->n_particles = c.N; info->n_tool = c.n_tool;
  info->grid_nx = c.gnx; info->grid_ny = c.gny; info->num_cell = c.num_cell;
  info->precision = s->precision;
  info->dt = c.dt; info->t_final = c.t_final; info->mass = c.mass;
  info->bbmin_x = c.bbmin_x; info->bbmin_y = c.bbmin_y; info->cell_size = c.cell;
  info->step = s->step;
  info->grid_adaptive = s->k.adaptive;
  info->cell_table = s->k.ncell_table;
  if (s->k.adaptive) {
    mfc::grid_geo g;
    int rc = current_geometry(s, g);
    if (rc) return rc;
    info->grid_nx = g.gnx; info->grid_ny = g.gny; info->num_cell = g.num_cell;
    info->bbmin_x = g.bbmin_x; info->bbmin_y = g.bbmin_y; info->cell_size = g.cell;
    info->grid_overflow = g.overflow;
  }
  unsigned long long oog = 0;
  HIP_TRY(hipMemcpyAsync(&oog, s->bins.clamped, sizeof(oog), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  info->out_of_grid = (int64_t)oog;
  return MFC_OK;
}

static int field_ptrs(mfc_sim *s, int field, void **ptr, int &nc) {
  const mfc::dstate &d = s->set[s->cur];
  nc = field_components(field);
  for (int q = 0; q < 4; q++) ptr[q] = nullptr;
  switch (field) {
    case MFC_FIELD_POS: ptr[0] = d.px; ptr[1] = d.py; break;
    case MFC_FIELD_VEL: ptr[0] = d.vx; ptr[1] = d.vy; break;
    case MFC_FIELD_S: ptr[0] = d.sxx; ptr[1] = d.sxy; ptr[2] = d.syy; ptr[3] = d.szz; break;
    case MFC_FIELD_RHO: ptr[0] = d.rho; break;
    case MFC_FIELD_P: ptr[0] = s->pos.p; break;
    case MFC_FIELD_T: ptr[0] = d.T; break;
    case MFC_FIELD_EPS_PL: ptr[0] = d.eps; break;
    case MFC_FIELD_EPS_PL_DOT: ptr[0] = d.epsd; break;
    case MFC_FIELD_FC: ptr[0] = d.fcx; ptr[1] = d.fcy; break;
    case MFC_FIELD_FT: ptr[0] = d.ftx; ptr[1] = d.fty; break;
    case MFC_FIELD_N: ptr[0] = s->pos.nx; ptr[1] = s->pos.ny; break;
    case MFC_FIELD_H: ptr[0] = d.h; break;
    case MFC_FIELD_FLAGS: ptr[0] = d.flags; break;
    case MFC_FIELD_ID: ptr[0] = d.id; break;
    case MFC_FIELD_HASH: ptr[0] = s->bins.key_sorted; break;
    case MFC_FIELD_POS_T: ptr[0] = d.ptx; ptr[1] = d.pty; break;
    case MFC_FIELD_VEL_T: ptr[0] = d.vtx; ptr[1] = d.vty; break;
    case MFC_FIELD_RHO_T: ptr[0] = d.rhot; break;
    case MFC_FIELD_S_T: ptr[0] = d.stxx; ptr[1] = d.stxy; ptr[2] = d.styy; ptr[3] = d.stzz; break;
    case MFC_FIELD_T_T: ptr[0] = d.Tt; break;
    case MFC_FIELD_R: ptr[0] = s->pos.rxx; ptr[1] = s->pos.rxy; ptr[2] = s->pos.ryy; break;
    default: return fail(MFC_E_INVALID, "unknown field");
  }
  return MFC_OK;
}

int mfc_get_field(mfc_sim *s, int field, void *host, size_t bytes) {
  if (!s || !host) return fail(MFC_E_INVALID, "null argument");
  void *ptr[4];
  int nc;
  int rc = field_ptrs(s, field, ptr, nc);
  if (rc) return rc;
  const int N = s->k.N;
  const size_t es = field_is_int(field) ? sizeof(int) : sizeof(double);
  if (bytes != (size_t)N * nc * es) return fail(MFC_E_RANGE, "buffer size does not match the field");
  if (nc == 1) {
    rc = copy_out(s, host, ptr[0], es * N);
    if (rc) return rc;
    if (field == MFC_FIELD_FLAGS) {
      int *f = (int *)host;
      for (int i = 0; i < N; i++) f[i] &= (MFC_FLAG_FIXED | MFC_FLAG_BLANKED | MFC_FLAG_TOOL);
    }
    return MFC_OK;
  }
  std::vector<double> tmp((size_t)N);
  double *out = (double *)host;
  for (int c = 0; c < nc; c++) {
    rc = copy_out(s, tmp.data(), ptr[c], sizeof(double) * N);
    if (rc) return rc;
    for (int i = 0; i < N; i++) out[(size_t)nc * i + c] = tmp[i];
  }
  return MFC_OK;
}

int mfc_set_field(mfc_sim *s, int field, const void *host, size_t bytes) {
  if (!s || !host) return fail(MFC_E_INVALID, "null argument");
  if (field == MFC_FIELD_ID || field == MFC_FIELD_HASH) return fail(MFC_E_INVALID, "field is read-only");
  void *ptr[4];
  int nc;
  int rc = field_ptrs(s, field, ptr, nc);
  if (rc) return rc;
  const int N = s->k.N;
  const size_t es = field_is_int(field) ? sizeof(int) : sizeof(double);
  if (bytes != (size_t)N * nc * es) return fail(MFC_E_RANGE, "buffer size does not match the field");
  HIP_TRY(hipStreamSynchronize(s->stream));
  s->keys_fresh = false;  // positions and flags decide the binning
  if (nc == 1) {
    HIP_TRY(hipMemcpy(ptr[0], host, es * N, hipMemcpyHostToDevice));
    return MFC_OK;
  }
  std::vector<double> tmp((size_t)N);
  const double *in = (const double *)host;
  for (int c = 0; c < nc; c++) {
    for (int i = 0; i < N; i++) tmp[i] = in[(size_t)nc * i + c];
    HIP_TRY(hipMemcpy(ptr[c], tmp.data(), sizeof(double) * N, hipMemcpyHostToDevice));
  }
  return MFC_OK;
}

int mfc_get_cells(mfc_sim *s, int32_t *offsets, size_t count) {
  if (!s || !offsets) return fail(MFC_E_INVALID, "null argument");
  mfc::grid_geo g;
  int rc = current_geometry(s, g);
  if (rc) return rc;
  if (count != (size_t)g.num_cell + 1) return fail(MFC_E_RANGE, "count must be num_cell + 1");
  return copy_out(s, offsets, s->k.rothlin ? s->bins.offsets_r : s->bins.offsets, sizeof(int) * count);
}

int mfc_tool_wear(mfc_sim *s, double K, double alpha, const double tool_vel[2], double dt, double out[4]) {
  if (!s || !tool_vel || !out) return fail(MFC_E_INVALID, "null argument");
  if (!s->wear_part) {
    int rc = s->alloc(&s->wear_part, 4 * (size_t)s->ntiles + 4);
    if (rc) return rc;
  }
  double *res = s->wear_part + 4 * (size_t)s->ntiles;
  mfc::launch_tool_wear(s->k, s->set[s->cur], s->pos, K, alpha, tool_vel[0], tool_vel[1], s->wear_part, res,
                        s->stream);
  HIP_TRY(hipGetLastError());
  double r[4];
  HIP_TRY(hipMemcpyAsync(r, res, sizeof(r), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  const double n = r[3];
  const double avg = n == 0. ? 0. : r[2] / n;
  s->wear_accum += dt * avg;
  out[0] = n == 0. ? 0. : r[0];
  out[1] = r[1];
  out[2] = avg;
  out[3] = s->wear_accum;
  return MFC_OK;
}

int mfc_get_tool(mfc_sim *s, mfc_tool_state *tool) {
  if (!s || !tool) return fail(MFC_E_INVALID, "null argument");
  return copy_out(s, tool, s->tool + s->cur, sizeof(mfc_tool_state));
}

int mfc_scenario(const mfc_config *cfg, mfc_info *info, double *pos, double *rho, double *h, double *T,
                 int32_t *flags, mfc_tool_state *tool) {
  if (!cfg || !info) return fail(MFC_E_INVALID, "null argument");
  mfc::scenario S;
  std::string err;
  int rc = mfc::build_scenario(*cfg, S, err);
  if (rc) return fail(rc, err);
  const mfc::sim_constants &c = S.k;
  std::memset(info, 0, sizeof(*info));
  info->n_particles = c.N; info->n_tool = c.n_tool;
  info->grid_nx = c.gnx; info->grid_ny = c.gny; info->num_cell = c.num_cell;
  info->precision = cfg->precision;
  info->dt = c.dt; info->t_final = c.t_final; info->mass = c.mass;
  info->bbmin_x = c.bbmin_x; info->bbmin_y = c.bbmin_y; info->cell_size = c.cell;
  info->grid_adaptive = c.adaptive;
  info->cell_table = c.ncell_table;
  for (int i = 0; i < c.N; i++) {
    if (pos) { pos[2 * i] = S.px[i]; pos[2 * i + 1] = S.py[i]; }
    if (rho) rho[i] = S.rho[i];
    if (h) h[i] = S.h[i];
    if (T) T[i] = S.T[i];
    if (flags) flags[i] = S.flags[i];
  }
  if (tool) *tool = S.tool;
  return MFC_OK;
}

int mfc_device_math(int fn, const double *x, const double *y, double *out, int64_t n) {
  if (!x || !out || n < 0 || fn < 0 || fn > 12) return fail(MFC_E_INVALID, "bad argument");
  if (n == 0) return MFC_OK;
  double *dx = nullptr, *dy = nullptr, *dout = nullptr;
  size_t bytes = sizeof(double) * (size_t)n;
  int rc = MFC_OK;
  if (hipMalloc(&dx, bytes) != hipSuccess || hipMalloc(&dout, bytes) != hipSuccess ||
      (y && hipMalloc(&dy, bytes) != hipSuccess)) {
    rc = fail(MFC_E_NOMEM, "hipMalloc");
  } else if (hipMemcpy(dx, x, bytes, hipMemcpyHostToDevice) != hipSuccess ||
             (y && hipMemcpy(dy, y, bytes, hipMemcpyHostToDevice) != hipSuccess)) {
    rc = fail(MFC_E_HIP, "upload");
  } else {
    mfc::launch_device_math(fn, dx, dy, dout, n, nullptr);
    if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail(MFC_E_HIP, "device math");
  }
  if (dx) (void)hipFree(dx);
  if (dy) (void)hipFree(dy);
  if (dout) (void)hipFree(dout);
  return rc;
}

}  // extern "C"
