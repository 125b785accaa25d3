// mfc_kernels.h — launch interface of the gfx950 kernels of one SPH cutting step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mfc_device.h"

namespace mfc {

// one set of the permuted persistent particle fields, structure of arrays in storage order
struct dstate {
  double *px, *py, *vx, *vy;
  double *sxx, *sxy, *syy, *szz;
  double *rho, *T, *eps, *epsd, *h;
  double *fcx, *fcy, *ftx, *fty;
  double *ptx, *pty, *vtx, *vty, *rhot, *stxx, *stxy, *styy, *stzz, *Tt;
  int *flags, *id;
};
constexpr int DSTATE_DOUBLES = 27;

// binning scratch and cell table
struct dbins {
  int *key;          // cell id per particle in the current (unsorted) order
  int *key_sorted;   // cell id per particle in sorted order (= reference's sorted hash)
  int *offsets;      // num_cell + 1 exclusive prefix sums: cell c = [offsets[c], offsets[c+1])
  int *bucket;       // particle indices grouped by cell, unordered inside a cell
  int *perm;         // sorted position -> current position
  int *spos;         // current position -> sorted position
  int *bsum;         // particles counted in each scan chunk of cells, zero between steps
  int2 *clist;       // per scan chunk, (current position, cell id) of the particles counted in
                     // its cells, in arrival order (ccap entries per chunk; a chunk that counts
                     // more finds them by a pass over all particles)
  int ccap;
  int *offsets_r;    // MFC_GRID_ROTHLIN: the ranges rebuilt from the sorted ids (the step reads these)
  unsigned long long *clamped;  // binned particles whose position lay outside the grid on an axis
};

// positional (not permuted) arrays, as particle_gpu::{n,p,R} (particle_gpu.h:39-47)
struct dpositional {
  double *nx, *ny, *p, *rxx, *rxy, *ryy;
};

constexpr int REC = 16;    // doubles per neighbour record
constexpr int ACC_N = 14;  // neighbour sums per particle

// neighbour records written by the predict kernel and read by the interaction kernel, in
// sorted order: posd (FP64 predicted position) and vfl (flags of the record) for the window
// walk, and the pair record, rec (16 doubles per particle, FP64 path) or rec32 (16 floats,
// FP32C path)
struct drecords {
  bool f32;
  double *rec;
  double2 *posd;
  int *vfl;
  float *rec32;
};

// window candidates of a tile that the interaction kernels stage in LDS (positions and flags,
// 20 bytes each); a tile whose candidates span more sorted indices reads them from memory
constexpr int STAGE_CAP = MFC_STAGE_CAP;

// tool-force reduction: per-tile partials written by the interaction kernel, reduced by the
// force-final pass (the next step's first kernel, or a kernel of its own after the last step),
// which appends F to the history
struct dforce {
  double *partials;      // 4 per wave of 64 particles: sum fc.x, fc.y, ft.x, ft.y
  double *force_last;    // F of the last step
  double *hist;          // force history, 2 per step, cap entries
  int64_t cap;
  int64_t *ctr;          // steps recorded
  int *pending;          // 1: the partials of a step are not yet reduced
};

// the particle flags of the state (FIXED, BLANKED, TOOL)
constexpr int FLAG_MASK = MFC_FLAG_FIXED | MFC_FLAG_BLANKED | MFC_FLAG_TOOL;

// internal flag: some value of the particle's neighbour record is NaN or Inf
constexpr int FLAG_NONFINITE = 8;

// launch variant of the interaction (results are identical for every variant)
struct variant {
  bool split;    // sums kernel + update kernel instead of one fused kernel
  bool compact;  // per-lane mask of in-support neighbours before the pair loop
  int chunk;     // tiles per XCD chunk of the interaction kernels' tile map (1..XCD_CHUNK_MAX)
};

// one instance of a batch as the batched kernels see it (an array of these lives in device
// memory, one per instance and step parity)
struct inst_args {
  dev_consts k;
  dstate cur, nxt;  // state at the start of the step, and the sorted copy updated by the step
  dbins b;
  dpositional pos;
  drecords r;
  const mfc_tool_state *tool_in;
  mfc_tool_state *tool_out;
  dforce f;
  double *acc;  // split interaction only
  double *acc_out;  // when set, the interaction also stores its neighbour sums there (ACC_N
                    // arrays of N doubles), then each particle's secant result (N doubles,
                    // wave_returns); the reverse mode's replay keeps them
  grid_geo *geo;  // adaptive grid: the geometry of the current step (written by the device)
  int ntiles;   // particle tiles of TILE
  int nscan;    // scan chunks of the cell table
  int bin_next;  // 1: the interaction bins the next step's start state (0: the reverse mode's
                 // replay, which bins every restored state itself)
};

// the work of one batched step: instance descriptors and (instance, block) work tables
struct batch_work {
  const inst_args *args;  // n descriptors, device memory
  int n;
  const int2 *tiles;      // (instance, particle tile) for every tile of every instance
  int ntiles;
  const int2 *chunks;     // (instance, scan chunk) for every chunk of every cell table
  int nchunks;
  bool f32;
  variant v;
  bool adaptive;          // some instance has an adaptive grid (a geometry pass starts the step)
  bool rothlin;           // some instance uses grid_gpu_rothlin's cell ranges (a pass after predict)
};

// one step of every instance of a batch: 3 (fused) or 4 (split) launches, and its phases. The
// cell ids, chunk lists and chunk sums of a step's start state are computed by the previous
// step's interaction; launch_batch_hash computes them from the state itself (chunk sums zeroed
// before)
void launch_batch_step(const batch_work &w, hipStream_t s);
void launch_batch_hash(const batch_work &w, hipStream_t s);
void launch_batch_binning(const batch_work &w, hipStream_t s);
void launch_batch_predict(const batch_work &w, hipStream_t s);
// predict that also writes the sorted copy of the start state into nxt (FP64 records; the
// reverse mode reads it)
void launch_batch_predict_gather(const batch_work &w, hipStream_t s);
// with events t0 and t1 the launch records the kernel's own start and end in them
// (hipExtLaunchKernelGGL), which times the kernel without the gap before it
void launch_batch_interact_fused(const batch_work &w, hipStream_t s, hipEvent_t t0 = nullptr,
                                 hipEvent_t t1 = nullptr);
void launch_batch_sums(const batch_work &w, hipStream_t s, hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);
void launch_batch_update(const batch_work &w, hipStream_t s, hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);
void launch_batch_force(const batch_work &w, hipStream_t s);
int scan_chunk_cells();
// entries of a scan chunk's particle list: 4 particles per cell of the chunk (a cell holds
// about 2.25 at the lattice spacing), at most N
int chunk_list_cap(int N);

// Usui wear rate of state st: out = {min, max, sum, count} over the particles in contact
// (part: 4 doubles per tile of scratch)
void launch_tool_wear(const dev_consts &k, const dstate &st, const dpositional &pos, double K, double alpha,
                      double tvx, double tvy, double *part, double *out, hipStream_t s);
int scan_blocks(int num_cell);
void launch_device_math(int fn, const double *x, const double *y, double *out, int64_t n, hipStream_t s);

}  // namespace mfc
