// mfc_batch.hip — many independent cutting instances stepped together (C ABI in include/mfc.h).
//
// A batch owns one stream and n instance handles created by mfc_create (each with its own
// buffers and constants). One batched step is one launch per phase for all instances
// (launch_batch_step): every block takes an (instance, tile) or (instance, scan chunk) item from
// a work table and runs the per-tile device body on that instance's descriptor, so each
// instance computes exactly the bits of its own single run while small instances share the
// chip. Two consecutive batched steps are captured in a hipGraph per step parity.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "mfc_sim.h"

struct mfc_batch {
  std::vector<mfc_sim *> sims;
  hipStream_t stream = nullptr;
  bool f32 = false;
  mfc_tuning tuning = {0, 0, 0, 0, 0, 0, {0, 0}};
  mfc::inst_args *args[2] = {nullptr, nullptr};  // descriptors per start parity (device)
  int2 *tiles = nullptr, *chunks = nullptr;
  int ntiles = 0, nchunks = 0;
  int cur = 0;
  int64_t step = 0;
  hipGraphExec_t graph[2] = {nullptr, nullptr};

  bool adaptive = false;  // some instance has an adaptive grid
  bool rothlin = false;   // some instance uses grid_gpu_rothlin's ranges
  mfc::batch_work work(int parity) const {
    return {args[parity], (int)sims.size(), tiles, ntiles, chunks, nchunks, f32,
            {tuning.split_interact != 0, tuning.compact_pairs != 0, tuning.xcd_chunk ? tuning.xcd_chunk : mfc::XCD_CHUNK},
            adaptive, rothlin};
  }
  void drop_graphs() {
    for (int p = 0; p < 2; p++)
      if (graph[p]) {
        (void)hipGraphExecDestroy(graph[p]);
        graph[p] = nullptr;
      }
  }
};

using mfc::fail;

namespace {

void batch_release(mfc_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  b->drop_graphs();
  for (mfc_sim *s : b->sims) {
    if (!s) continue;
    s->stream = nullptr;  // the batch's stream, destroyed below
    mfc_destroy(s);
  }
  for (void *p : {(void *)b->args[0], (void *)b->args[1], (void *)b->tiles, (void *)b->chunks})
    if (p) (void)hipFree(p);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

// as build_graph of a single handle: the capture is always ended, and the batch keeps the
// executable graph only when capture and instantiation both succeeded
int build_batch_graph(mfc_batch *b, int parity) {
  HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  for (int q = 0; q < GRAPH_STEPS; q++) mfc::launch_batch_step(b->work((parity + q) & 1), b->stream);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess || g == nullptr) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return fail(MFC_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  }
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(MFC_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  b->graph[parity] = x;
  return MFC_OK;
}

// the binning of every instance's current state (computed by the last step's interaction),
// recomputed for all instances when any of them lacks it
int batch_keys(mfc_batch *b) {
  bool fresh = true;
  for (mfc_sim *s : b->sims) fresh = fresh && s->keys_fresh;
  if (fresh) return MFC_OK;
  for (mfc_sim *s : b->sims) {
    HIP_TRY(hipMemsetAsync(s->bins.bsum, 0, sizeof(int) * (size_t)mfc::scan_blocks(s->k.ncell_table), b->stream));
  }
  mfc::launch_batch_hash(b->work(b->cur), b->stream);
  for (mfc_sim *s : b->sims) s->keys_fresh = true;
  return MFC_OK;
}

void advance(mfc_batch *b) {
  b->cur = 1 - b->cur;
  b->step++;
  for (mfc_sim *s : b->sims) {
    s->cur = b->cur;
    s->step++;
  }
}

}  // namespace

extern "C" {

int mfc_batch_create(const mfc_config *cfgs, int n, mfc_batch **out) {
  if (!cfgs || !out || n < 1) return fail(MFC_E_INVALID, "need n >= 1 configurations");
  *out = nullptr;
  for (int i = 1; i < n; i++)
    if (cfgs[i].precision != cfgs[0].precision) return fail(MFC_E_INVALID, "all instances need the same precision");
  mfc_batch *b = new mfc_batch();
  b->f32 = cfgs[0].precision == MFC_PREC_FP32C;
  auto bail = [&](int code) {
    batch_release(b);
    return code;
  };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return bail(fail(MFC_E_HIP, "stream"));
  b->sims.assign(n, nullptr);
  for (int i = 0; i < n; i++) {
    mfc_sim *s = nullptr;
    int rc = mfc_create(&cfgs[i], &s);
    if (rc) {
      std::string msg = "instance " + std::to_string(i) + ": " + mfc_last_error();
      return bail(fail(rc, msg));
    }
    b->sims[i] = s;
    if (hipStreamSynchronize(s->stream) != hipSuccess) return bail(fail(MFC_E_HIP, "instance setup"));
    (void)hipStreamDestroy(s->stream);
    s->stream = b->stream;
    s->batched = true;
    s->use_graphs = false;
    b->adaptive = b->adaptive || s->k.adaptive;
    b->rothlin = b->rothlin || s->k.rothlin;
  }
  std::vector<int2> tiles, chunks;
  std::vector<mfc::inst_args> a0(n), a1(n);
  for (int i = 0; i < n; i++) {
    const mfc_sim *s = b->sims[i];
    a0[i] = instance_args(s, 0);
    a1[i] = instance_args(s, 1);
    for (int t = 0; t < a0[i].ntiles; t++) tiles.push_back(make_int2(i, t));
    for (int c = 0; c < a0[i].nscan; c++) chunks.push_back(make_int2(i, c));
  }
  b->ntiles = (int)tiles.size();
  b->nchunks = (int)chunks.size();
  if (hipMalloc(&b->args[0], sizeof(mfc::inst_args) * n) != hipSuccess ||
      hipMalloc(&b->args[1], sizeof(mfc::inst_args) * n) != hipSuccess ||
      hipMalloc(&b->tiles, sizeof(int2) * tiles.size()) != hipSuccess ||
      hipMalloc(&b->chunks, sizeof(int2) * chunks.size()) != hipSuccess)
    return bail(fail(MFC_E_NOMEM, "hipMalloc (batch tables)"));
  if (hipMemcpyAsync(b->args[0], a0.data(), sizeof(mfc::inst_args) * n, hipMemcpyHostToDevice, b->stream) ||
      hipMemcpyAsync(b->args[1], a1.data(), sizeof(mfc::inst_args) * n, hipMemcpyHostToDevice, b->stream) ||
      hipMemcpyAsync(b->tiles, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice, b->stream) ||
      hipMemcpyAsync(b->chunks, chunks.data(), sizeof(int2) * chunks.size(), hipMemcpyHostToDevice, b->stream) ||
      hipStreamSynchronize(b->stream))
    return bail(fail(MFC_E_HIP, "upload of the batch tables"));
  *out = b;
  return MFC_OK;
}

void mfc_batch_destroy(mfc_batch *b) { batch_release(b); }

int mfc_batch_size(const mfc_batch *b) { return b ? (int)b->sims.size() : 0; }

int mfc_batch_instance(mfc_batch *b, int i, mfc_sim **sim) {
  if (!b || !sim || i < 0 || i >= (int)b->sims.size()) return fail(MFC_E_INVALID, "bad instance");
  *sim = b->sims[i];
  return MFC_OK;
}

int mfc_batch_set_tuning(mfc_batch *b, const mfc_tuning *t) {
  if (!b || !t) return fail(MFC_E_INVALID, "null argument");
  int rc = check_tuning(t);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->tuning = *t;
  const size_t n = b->sims.size();
  std::vector<mfc::inst_args> a0(n), a1(n);
  for (size_t i = 0; i < n; i++) {
    mfc_sim *s = b->sims[i];
    s->tuning = *t;
    s->k.stage_limit = t->stage_limit;
    s->k.sort_fast = t->sort_fast;
    s->k.sort_list = t->sort_list;
    a0[i] = instance_args(s, 0);
    a1[i] = instance_args(s, 1);
  }
  HIP_TRY(hipMemcpy(b->args[0], a0.data(), sizeof(mfc::inst_args) * n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b->args[1], a1.data(), sizeof(mfc::inst_args) * n, hipMemcpyHostToDevice));
  b->drop_graphs();
  return MFC_OK;
}

int mfc_batch_step(mfc_batch *b, int nsteps) {
  if (!b || nsteps < 0) return fail(MFC_E_INVALID, "bad argument");
  int rc = batch_keys(b);
  if (rc) return rc;
  int done = 0;
  if (nsteps >= GRAPH_STEPS) {
    while (nsteps - done >= GRAPH_STEPS) {
      const int par = b->cur;
      if (!b->graph[par] && (rc = build_batch_graph(b, par))) return rc;
      HIP_TRY(hipGraphLaunch(b->graph[par], b->stream));
      for (int q = 0; q < GRAPH_STEPS; q++) advance(b);
      done += GRAPH_STEPS;
    }
  }
  for (; done < nsteps; done++) {
    mfc::launch_batch_step(b->work(b->cur), b->stream);
    advance(b);
  }
  mfc::launch_batch_force(b->work(b->cur), b->stream);  // the last step's tool force
  HIP_TRY(hipGetLastError());
  return MFC_OK;
}

int mfc_batch_sync(mfc_batch *b) {
  if (!b) return fail(MFC_E_INVALID, "null handle");
  HIP_TRY(hipStreamSynchronize(b->stream));
  return MFC_OK;
}

int mfc_batch_vjp_forward(mfc_batch *b, int nsteps, int checkpoint_every, int w0, int w1, double *Fmean) {
  if (!b || !Fmean) return fail(MFC_E_INVALID, "null argument");
  int rc;
  for (mfc_sim *s : b->sims)
    if ((rc = mfc::adj_begin(s, nsteps, checkpoint_every, w0, w1))) return rc;
  if ((rc = batch_keys(b))) return rc;
  for (int t = 0; t < nsteps; t++) {
    for (mfc_sim *s : b->sims)
      if ((rc = mfc::adj_checkpoint(s, t))) return rc;
    mfc::launch_batch_step(b->work(b->cur), b->stream);
    advance(b);
  }
  mfc::launch_batch_force(b->work(b->cur), b->stream);
  HIP_TRY(hipGetLastError());
  for (size_t i = 0; i < b->sims.size(); i++)
    if ((rc = mfc::adj_finish(b->sims[i], Fmean + 2 * i))) return rc;
  return MFC_OK;
}

int mfc_batch_vjp_backward(mfc_batch *b, const double *cot, double *grad) {
  if (!b || !cot || !grad) return fail(MFC_E_INVALID, "null argument");
  if (b->f32) return fail(MFC_E_INVALID, "the adjoint needs precision FP64");
  return mfc::adj_backward(b->sims.data(), (int)b->sims.size(), b->work(b->cur), b->stream, cot, grad);
}

}  // extern "C"
