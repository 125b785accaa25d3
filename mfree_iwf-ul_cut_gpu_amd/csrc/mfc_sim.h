// mfc_sim.h — the simulation handle behind the C ABI (internal to the library).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/mfc.h"
#include "mfc_kernels.h"
#include "mfc_setup.h"

namespace mfc {

// records the message of the last failing call of this thread; returns code
int fail(int code, const std::string &msg);

}  // namespace mfc

#define HIP_TRY(expr)                                                                         \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return mfc::fail(MFC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int64_t HIST_CAP = 1 << 20;

struct mfc_sim;

namespace mfc {
struct adj_work;  // reverse-mode workspace (mfc_adjoint.hip)
void adj_free(adj_work *w);
// taped forward in three parts, so that a batch can interleave its instances: validate and
// allocate, keep the state at the start of step t when t is a checkpoint, and mean the window
int adj_begin(mfc_sim *s, int nsteps, int checkpoint_every, int w0, int w1);
int adj_checkpoint(mfc_sim *s, int t);
int adj_finish(mfc_sim *s, double Fmean[2]);
// the backward pass of a group of instances that recorded the same run (steps, checkpoint
// interval, window): one handle, or the instances of a batch, reversed together with one launch
// per phase over the group's work tables (fw: tile and chunk tables, launch variant);
// cot: 2 per instance, grad: 6 per instance
int adj_backward(mfc_sim *const *sims, int n, const batch_work &fw, hipStream_t st, const double *cot, double *grad);
}  // namespace mfc

struct mfc_sim {
  mfc_config cfg;      // the configuration the handle was created from (checkpoints keep it)
  mfc::scenario sc;
  mfc::dev_consts k;
  int ntiles = 0;
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
  mfc::dstate set[2];
  mfc::dbins bins = {};
  mfc::dpositional pos;
  mfc::drecords recs = {false, nullptr, nullptr, nullptr};
  int precision = MFC_PREC_FP64;
  mfc_tool_state *tool = nullptr;  // two states, alternating each step
  double *partials = nullptr, *force_last = nullptr, *hist = nullptr;
  int64_t *hist_ctr = nullptr;
  int *force_pending = nullptr;  // the last step's tool force is not yet reduced
  mfc::dforce force() const { return {partials, force_last, hist, HIST_CAP, hist_ctr, force_pending}; }
  int64_t hist_base = 0;           // steps recorded before the last history reset
  int cur = 0;                     // set holding the current state, and index of the current tool state
  int64_t step = 0;
  hipGraphExec_t graph[2] = {nullptr, nullptr};  // GRAPH_STEPS steps starting from parity 0 / 1
  bool use_graphs = true;
  mfc::adj_work *adj = nullptr;    // reverse-mode tape of the last mfc_vjp_forward
  bool batched = false;            // instance of a batch: stepped by mfc_batch_step only
  // bins.key/bsum/clist hold the binning of the current state (computed by the last step's
  // interaction); false after creation and after anything else wrote the state or the bins
  bool keys_fresh = false;
  double *acc = nullptr;           // neighbour sums between the split interaction kernels
  double *wear_part = nullptr;     // tool wear: 4 doubles per tile, then the 4 results
  double wear_accum = 0.;          // accumulated wear, sum of dt * mean rate (tool_wear.cu:59)
  mfc_tuning tuning = {0, 0, 0, 0, 0, 0, {0, 0}};
  mfc::variant variant() const {
    return {tuning.split_interact != 0, tuning.compact_pairs != 0, tuning.xcd_chunk ? tuning.xcd_chunk : mfc::XCD_CHUNK};
  }
  // the step runs the batched kernels on a batch of one: descriptors of this handle for start
  // parity 0 and 1, and its (instance, block) work tables
  mfc::inst_args *dargs = nullptr;
  int2 *dtiles = nullptr, *dchunks = nullptr;
  int nchunks = 0;
  mfc::grid_geo *geo = nullptr;    // adaptive grid: the geometry of the last step (device)
  mfc::batch_work work(int c) const {
    return {dargs + c, 1, dtiles, ntiles, dchunks, nchunks, recs.f32, variant(), k.adaptive != 0, k.rothlin != 0};
  }

  template <class T>
  int alloc(T **p, size_t n) {
    void *q = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&q, n * sizeof(T));
    if (e != hipSuccess) return mfc::fail(MFC_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    allocs.push_back(q);
    *p = (T *)q;
    return MFC_OK;
  }
};

// the state arrays of one set, carved from one allocation in the order of dstate
// (DSTATE_DOUBLES doubles, then flags and id), so that a set is one contiguous block
size_t state_block_bytes(int N);
void carve_state(void *base, int N, mfc::dstate &d);
// the positional arrays (nx, ny, p, rxx, rxy, ryy), one contiguous block
size_t positional_block_bytes(int N);
void carve_positional(void *base, int N, mfc::dpositional &p);
// queue one step: set[c] -> set[1-c]
// steps per captured hipGraph (even: a graph starts and ends on the same buffer parity)
constexpr int GRAPH_STEPS = 8;
void enqueue_step(mfc_sim *s, int c);
// bins.key/bsum/clist of the current state, recomputed when keys_fresh is false
int ensure_keys(mfc_sim *s);
void finish_steps(mfc_sim *s);
int check_tuning(const mfc_tuning *t);
// the batched kernels' view of a handle for start parity c
mfc::inst_args instance_args(const mfc_sim *s, int c);
// the grid of the last step: the creation geometry for a locked grid, the device's record for
// an adaptive one (synchronises the handle's stream)
int current_geometry(const mfc_sim *s, mfc::grid_geo &g);
