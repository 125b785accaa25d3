// mfc_kernels.hip — gfx950 kernels of one SPH cutting step (FP64 and FP32C).
//
// Per step (the reference's ~20 launches of leap_frog::step, leap_frog.cu:94-155, become 3):
//   sort         one work-group per chunk of cells: the cells' offsets, buckets and the rank of
//                every particle in its cell, a stable counting sort by cell
//                (thrust::sort_by_key + reorder_data_and_find_cell_start,
//                grid_gpu_green.cu:22-116,177-201) from the cell ids (grid.cu:34-51), per-cell
//                and per-chunk counts and the chunks' particle lists that the previous step's
//                interaction computed for the updated state (hash_count computes them when the
//                state was written from outside, and for every state the reverse mode's replay
//                restores)
//   predict      leap-frog predict (leap_frog.cu:30-66) + EOS
//                (actions_gpu.cu:41-51) + artificial stress (:53-97) + neighbour record + tool
//                move, in the start state's storage order; the work-group of the first tile
//                also reduces the previous step's tool force (output.cu:30-39) from its tile
//                partials
//   interact     neighbour sums (interactions_gpu.cu:88-345) fused with every per-particle action
//                that follows them in the step: Jaumann, continuity, momentum, advection, correct,
//                friction heat, Johnson-Cook return, tool-particle move, contact, boundary
//                conditions, NaN blanking, the gather into sorted order, and the per-wave partial
//                sums of the tool force (or split into sums + update), and the cell ids and
//                counts of the next step
// After the last step of a call, force_final reduces the last step's tool force.
// Each body is a device function of its work index; the kb_* kernels run it over an
// (instance, tile) work table for any number of instances (a single handle is a batch of one).
//
// Memory layout: the persistent particle state is a structure of arrays in cell-sorted order,
// double-buffered across the step. Neighbour data are packed per particle into one 128-byte
// record (two 64-byte lines; FP32C: 64 bytes plus the FP64 position) so that the divergent
// per-lane neighbour reads are 16-byte vector loads from one cache line. Work-group to tile
// mapping is XCD-aware.
#include <hip/hip_ext.h>

#include <cfloat>

#include "mfc_kernels.h"
#include "mfc_pairs.h"

namespace mfc {

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 2;                         // cells per thread of the sort
constexpr int SCAN_CHUNK = SCAN_THREADS * SCAN_ITEMS;  // cells per chunk of the sort
constexpr int SORT_BATCH = 6;            // particles per thread the sort holds in registers
static_assert(SCAN_THREADS * SORT_BATCH == MFC_SORT_FAST_MAX, "the sort's register path bound of the header");
constexpr int CARRY_BATCH = 4;            // 16-byte loads of chunk sums per thread the sort
                                          // has in flight at once


// exclusive scan of one int per thread over a 256-thread block; total receives the block sum
__device__ inline int block_exclusive_scan(int v, int &total) {
  __shared__ int wsum[SCAN_THREADS / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  int pre = 0, tot = 0;
#pragma unroll
  for (int q = 0; q < SCAN_THREADS / 64; q++) {
    if (q < w) pre += wsum[q];
    tot += wsum[q];
  }
  total = tot;
  __syncthreads();
  return pre + x - v;
}

// ------------------------------------------------------------------------------- binning
// Each kernel body is a device function of its work index, called by the batched kernel with
// the index of an (instance, block) item of its work table.

// cell id of a particle for the binning (grid.cu:34-51; blanked particles keep their last
// sorted id, grid.cu:39-41); positions outside the grid are counted
__device__ inline int particle_cell(const dev_consts &k, double x, double y, int fl, int sorted_id,
                                    unsigned long long *__restrict__ clamped) {
  if (fl & MFC_FLAG_BLANKED) return sorted_id;
  bool cl;
  const int h = cell_of(k, x, y, cl);
  if (cl) atomicAdd(clamped, 1ull);
  return h;
}

// records cell id h of particle i (storage order) and appends (i, h) to the particle list of
// its chunk of SCAN_CHUNK cells (bsum counts the chunk's particles; the list keeps the first ccap
// of them). All 64 lanes of the wave take part (lanes past N carry h = -1). The storage order is
// the previous step's cell order, so neighbouring lanes mostly share a chunk: each run of equal
// chunks in the wave takes its list places with one atomic.
__device__ inline void count_cell(int i, int h, int *__restrict__ key, int *__restrict__ bsum,
                                  int2 *__restrict__ clist, int ccap) {
  if (h >= 0) key[i] = h;
  const int lane = threadIdx.x & 63;
  const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  const int c = h >= 0 ? h / SCAN_CHUNK : -1;
  const int cprev = __shfl_up(c, 1, 64);
  const bool chead = lane == 0 || cprev != c;
  const unsigned long long cheads = __ballot(chead);
  const int my_chead = 63 - __clzll(cheads & upto);
  const unsigned long long cafter = (lane == 63) ? 0ull : (cheads & ~upto);
  const int cnext = cafter ? __ffsll((long long)cafter) - 1 : 64;
  int cbase = 0;
  if (chead && c >= 0) cbase = atomicAdd(&bsum[c], cnext - lane);
  cbase = __shfl(cbase, my_chead, 64);
  const int place = cbase + (lane - my_chead);
  if (c >= 0 && place < ccap) clist[(size_t)c * ccap + place] = make_int2(i, h);
}

// cell id of every particle of the start state, counted in its chunk
__device__ inline void hash_count(const dev_consts &k, int i, const double *__restrict__ px,
                                  const double *__restrict__ py, const int *__restrict__ flags,
                                  const dbins &b) {
  const int h = i < k.N ? particle_cell(k, px[i], py[i], flags[i], b.key_sorted[i], b.clamped) : -1;
  count_cell(i, h, b.key, b.bsum, b.clist, b.ccap);
}

// The counting sort of one chunk of SCAN_CHUNK cells (SCAN_ITEMS per thread), by one work-group,
// from the chunk's particle list:
// 1. the particles counted per cell (LDS atomics, which also give each its arrival slot);
// 2. the cells' offsets (the chunk's offset is the sum of the chunk sums before it, added up
//    here, so no separate scan of the chunk sums);
// 3. every particle into its cell's bucket at its arrival slot;
// 4. each one ranked in its cell by current position, the order thrust::sort_by_key gives on
//    (cell id, 0..N-1): its sorted position (spos), perm and the sorted cell id there.
// The chunks' cells, bucket ranges and particles are disjoint, so the work-groups share nothing
// but the chunk sums they read. A chunk of at most SCAN_THREADS * SORT_BATCH particles (a dense
// chunk holds about 2.25 per cell at the lattice spacing) keeps them in registers and their
// buckets in LDS, and its list, chunk sum and preceding chunk sums are loaded at once: the
// only global loads of the work-group. A larger chunk walks its list (or, when the list
// overflowed, every particle's cell id) three times and uses the global bucket. Two cells per
// thread keep most work-groups of a large table resident at once (a 500k-particle cut has 1.18
// million cells: 2,310 work-groups). sort_fast and sort_list (launch settings, 0: none) lower
// the two bounds, so that the other paths can be run on any cut.
__device__ inline void sort_chunk(int blk, int N, int n, const dbins &b, int sort_fast, int sort_list) {
  __shared__ int off_l[SCAN_CHUNK], cnt_l[SCAN_CHUNK], fill_l[SCAN_CHUNK];
  __shared__ int bk_l[SCAN_THREADS * SORT_BATCH];
  const int c0 = blk * SCAN_CHUNK;
  const int2 *__restrict__ list = b.clist + (size_t)blk * b.ccap;
  const int list_cap = sort_list > 0 && sort_list < b.ccap ? sort_list : b.ccap;
  int fast_cap = list_cap < SCAN_THREADS * SORT_BATCH ? list_cap : SCAN_THREADS * SORT_BATCH;
  if (sort_fast > 0 && sort_fast < fast_cap) fast_cap = sort_fast;
  // the list's first entries, loaded before the chunk sum says how many are valid
  int2 e[SORT_BATCH];
#pragma unroll
  for (int u = 0; u < SORT_BATCH; u++) {
    const int p = u * SCAN_THREADS + threadIdx.x;
    e[u] = p < fast_cap ? list[p] : make_int2(-1, -1);
  }
  const int members = b.bsum[blk];
  // the chunk sums before this chunk, four per load (the sums are allocated in whole
  // 16-byte words), CARRY_BATCH loads per thread in flight at once
  const int4 *__restrict__ bsum4 = reinterpret_cast<const int4 *>(b.bsum);
  int carry_part = 0;
  for (int q0 = 0; q0 < blk; q0 += 4 * SCAN_THREADS * CARRY_BATCH) {
    int4 t[CARRY_BATCH];
#pragma unroll
    for (int u = 0; u < CARRY_BATCH; u++) {
      const int q = q0 + 4 * (u * SCAN_THREADS + (int)threadIdx.x);
      t[u] = q < blk ? bsum4[q >> 2] : make_int4(0, 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < CARRY_BATCH; u++) {
      const int q = q0 + 4 * (u * SCAN_THREADS + (int)threadIdx.x);
      carry_part += (q < blk ? t[u].x : 0) + (q + 1 < blk ? t[u].y : 0) + (q + 2 < blk ? t[u].z : 0) +
                    (q + 3 < blk ? t[u].w : 0);
    }
  }
#pragma unroll
  for (int q = 0; q < SCAN_ITEMS; q++) {
    cnt_l[q * SCAN_THREADS + threadIdx.x] = 0;
    fill_l[q * SCAN_THREADS + threadIdx.x] = 0;
  }
  const bool fast = members <= fast_cap;  // the same in every thread
  const bool listed = members <= list_cap;
  const int lim = listed ? members : N;
  __syncthreads();
  int ls[SORT_BATCH];
  if (fast) {
#pragma unroll
    for (int u = 0; u < SORT_BATCH; u++) {
      const int p = u * SCAN_THREADS + threadIdx.x;
      if (p >= members) e[u].x = -1;
      ls[u] = e[u].x >= 0 ? atomicAdd(&cnt_l[e[u].y - c0], 1) : 0;
    }
  } else {
    for (int p = threadIdx.x; p < lim; p += SCAN_THREADS) {
      const int h = (listed ? list[p].y : b.key[p]) - c0;
      if (h >= 0 && h < SCAN_CHUNK) atomicAdd(&cnt_l[h], 1);
    }
  }
  int carry;
  block_exclusive_scan(carry_part, carry);  // its barriers also complete the counts
  // the thread's SCAN_ITEMS consecutive cells
  const int l0 = threadIdx.x * SCAN_ITEMS;
  int v[SCAN_ITEMS], sv = 0;
#pragma unroll
  for (int q = 0; q < SCAN_ITEMS; q++) {
    v[q] = cnt_l[l0 + q];
    sv += v[q];
  }
  int total;
  int o = block_exclusive_scan(sv, total) + carry;
#pragma unroll
  for (int q = 0; q < SCAN_ITEMS; q++) {
    const int cell = c0 + l0 + q;
    off_l[l0 + q] = o;
    if (cell < n) {
      b.offsets[cell] = o;
      if (cell == n - 1) b.offsets[n] = o + v[q];
    }
    o += v[q];
  }
  __syncthreads();  // off_l
  const int base = off_l[0];
  // the sorted position of particle i (cell h of the chunk): its rank among the entries of the
  // cell's bucket (bk, indexed from offset bbase)
  auto rank = [&](int i, int h, const int *bk, int bbase) {
    const int cb = off_l[h], ce = cb + cnt_l[h];
    int rnk = 0;
    for (int q = cb; q < ce; q++) rnk += bk[q - bbase] < i;
    const int j = cb + rnk;
    b.spos[i] = j;
    b.perm[j] = i;
    b.key_sorted[j] = h + c0;
  };
  if (fast) {
#pragma unroll
    for (int u = 0; u < SORT_BATCH; u++)
      if (e[u].x >= 0) bk_l[off_l[e[u].y - c0] - base + ls[u]] = e[u].x;
    __syncthreads();  // the chunk's buckets
#pragma unroll
    for (int u = 0; u < SORT_BATCH; u++)
      if (e[u].x >= 0) rank(e[u].x, e[u].y - c0, bk_l, base);
  } else {
    int *__restrict__ bucket = b.bucket;
    for (int p = threadIdx.x; p < lim; p += SCAN_THREADS) {
      const int2 q = listed ? list[p] : make_int2(p, b.key[p]);
      const int h = q.y - c0;
      if (h >= 0 && h < SCAN_CHUNK) bucket[off_l[h] + atomicAdd(&fill_l[h], 1)] = q.x;
    }
    __syncthreads();  // the chunk's buckets
    for (int p = threadIdx.x; p < lim; p += SCAN_THREADS) {
      const int2 q = listed ? list[p] : make_int2(p, b.key[p]);
      const int h = q.y - c0;
      if (h >= 0 && h < SCAN_CHUNK) rank(q.x, h, bucket, 0);
    }
  }
}


// ------------------------------------------------------------------------------- predict
// artificial stress in single precision without trigonometry: with d = sqrt((sxx-syy)^2 + 4 sxy^2),
// cos 2theta = (sxx-syy)/d and sin 2theta = 2 sxy/d for theta = atan2(2 sxy, sxx-syy)/2, and the
// rotated normal stresses are the principal stresses (sxx+syy)/2 +- d/2 (actions_gpu.cu:53-97)
__device__ inline void artificial_stress_f32(float rho, float p, float Sxx, float Sxy, float Syy, float eps, float &Rxx,
                                             float &Rxy, float &Ryy) {
  const float sxx = Sxx - p, syy = Syy - p, sxy = Sxy;
  const float r21 = 1.0f / (rho * rho);
  const float a = sxx - syy, b = 2.0f * sxy;
  const float d = sqrtf(a * a + b * b);
  float c2 = 1.0f, s2 = 0.0f;
  if (d > 0.0f) { c2 = a / d; s2 = b / d; }
  const float ct2 = 0.5f * (1.0f + c2), st2 = 0.5f * (1.0f - c2), cs = 0.5f * s2;
  const float mid = 0.5f * (sxx + syy);
  const float r1 = mid + 0.5f * d, r2 = mid - 0.5f * d;
  const float q1 = r1 > 0.0f ? -eps * r1 * r21 : 0.0f;
  const float q2 = r2 > 0.0f ? -eps * r2 * r21 : 0.0f;
  Rxx = ct2 * q1 + st2 * q2;
  Rxy = cs * (q1 - q2);
  Ryy = st2 * q1 + ct2 * q2;
}

// leap-frog init + predict, EOS, artificial stress and the neighbour record written at the
// particle's sorted position (from the sort). Runs in the storage order of the start state (its
// reads are coalesced). F32: the record is 16 floats (64 bytes) plus the FP64 predicted
// position.
template <bool F32, bool GATHER>
__device__ inline void predict(const dev_consts &k, int t, const dbins &bins, const dstate &cur, const dstate &nxt,
                               const dpositional &pos, const drecords &recs,
                               const mfc_tool_state *__restrict__ tool_in, mfc_tool_state *__restrict__ tool_out) {
  double *__restrict__ rec = recs.rec;
  double2 *__restrict__ posd = recs.posd;
  int *__restrict__ vfl = recs.vfl;
  float *__restrict__ rec32 = recs.rec32;
  if (t == 0 && threadIdx.x == 0) {
    mfc_tool_state ts = *tool_in;
    tool_move(ts, k.dt);
    *tool_out = ts;
  }
  const int s = t * TILE + threadIdx.x;
  if (s >= k.N) return;
  const int i = bins.spos[s];
  const double x0 = cur.px[s], y0 = cur.py[s], u0 = cur.vx[s], v0 = cur.vy[s];
  const double sxx0 = cur.sxx[s], sxy0 = cur.sxy[s], syy0 = cur.syy[s];
  const double rho0 = cur.rho[s], T0 = cur.T[s];
  const double ptx = cur.ptx[s], pty = cur.pty[s], vtx = cur.vtx[s], vty = cur.vty[s];
  const double rhot = cur.rhot[s], stxx = cur.stxx[s], stxy = cur.stxy[s], styy = cur.styy[s];
  const double Tt = cur.Tt[s];
  const int fl = cur.flags[s] & FLAG_MASK;
  const double h = cur.h[s];
  if constexpr (GATHER) {  // the sorted copy of the start state (the reverse mode reads it)
    nxt.px[i] = x0; nxt.py[i] = y0; nxt.vx[i] = u0; nxt.vy[i] = v0;
    nxt.sxx[i] = sxx0; nxt.sxy[i] = sxy0; nxt.syy[i] = syy0; nxt.szz[i] = cur.szz[s];
    nxt.rho[i] = rho0; nxt.T[i] = T0; nxt.h[i] = h;
    nxt.eps[i] = cur.eps[s]; nxt.epsd[i] = cur.epsd[s];
    nxt.fcx[i] = cur.fcx[s]; nxt.fcy[i] = cur.fcy[s]; nxt.ftx[i] = cur.ftx[s]; nxt.fty[i] = cur.fty[s];
    nxt.ptx[i] = ptx; nxt.pty[i] = pty; nxt.vtx[i] = vtx; nxt.vty[i] = vty;
    nxt.rhot[i] = rhot; nxt.stxx[i] = stxx; nxt.stxy[i] = stxy; nxt.styy[i] = styy; nxt.stzz[i] = cur.stzz[s];
    nxt.Tt[i] = Tt;
    nxt.flags[i] = fl;
    nxt.id[i] = cur.id[s];
  }

  double x = x0, y = y0, u = u0, v = v0, sxx = sxx0, sxy = sxy0, syy = syy0, rho = rho0, T = T0;
  if (!(fl & MFC_FLAG_BLANKED)) {  // leap_frog.cu:42-66
    const double hdt = 0.5 * k.dt;
    x = x0 + hdt * ptx; y = y0 + hdt * pty;
    u = u0 + hdt * vtx; v = v0 + hdt * vty;
    sxx = sxx0 + hdt * stxx; sxy = sxy0 + hdt * stxy; syy = syy0 + hdt * styy;
    rho = rho0 + hdt * rhot;
    T = T0 + hdt * Tt;
  }
  // the positional values at i are recomputed below for every particle but the tool's, which
  // keep what the slot holds (the reference's positional arrays are not permuted)
  double p = 0., rxx = 0., rxy = 0., ryy = 0.;
  if (fl & MFC_FLAG_TOOL) {
    p = pos.p[i]; rxx = pos.rxx[i]; rxy = pos.rxy[i]; ryy = pos.ryy[i];
  }
  if constexpr (F32) {
    float fxx = (float)rxx, fxy = (float)rxy, fyy = (float)ryy;
    if (!(fl & MFC_FLAG_TOOL)) {
      p = k.c0 * k.c0 * (rho - k.rho0);
      artificial_stress_f32((float)rho, (float)p, (float)sxx, (float)sxy, (float)syy, (float)k.stress_eps, fxx, fxy, fyy);
      pos.p[i] = p; pos.rxx[i] = fxx; pos.rxy[i] = fxy; pos.ryy[i] = fyy;
    }
    const float rf = (float)rho, pf = (float)p;
    const float r21 = 1.0f / (rf * rf);
    float4 q0 = make_float4((float)u, (float)v, (float)h, rf);
    float4 q1 = make_float4((float)(k.mass / rho), ((float)sxx - pf) * r21, (float)sxy * r21, ((float)syy - pf) * r21);
    float4 q2 = make_float4(fxx, fxy, fyy, sqrtf((float)(k.K / rho)));
    float4 q3 = make_float4((float)T, __int_as_float(fl), 0.0f, 0.0f);
    float4 *dst = reinterpret_cast<float4 *>(rec32 + (size_t)16 * i);
    dst[0] = q0; dst[1] = q1; dst[2] = q2; dst[3] = q3;
    posd[i] = make_double2(x, y);
    vfl[i] = fl;
    return;
  } else {
    if (!(fl & MFC_FLAG_TOOL)) {
      p = k.c0 * k.c0 * (rho - k.rho0);
      artificial_stress(rho, p, sxx, sxy, syy, k.stress_eps, rxx, rxy, ryy, k.promote != 0);
      pos.p[i] = p; pos.rxx[i] = rxx; pos.rxy[i] = rxy; pos.ryy[i] = ryy;
    }
  }
  const double r21 = 1. / (rho * rho);
  double r[REC];
  r[0] = x; r[1] = y;
  r[3] = h;
  r[4] = u; r[5] = v;
  r[6] = rho;
  r[7] = k.mass / rho;
  r[8] = (sxx - p) * r21;
  r[9] = sxy * r21;
  r[10] = (syy - p) * r21;
  r[11] = rxx; r[12] = rxy; r[13] = ryy;
  r[14] = tsqrt(k, k.K / rho);
  r[15] = T;
  bool fin = true;
#pragma unroll
  for (int q = 0; q < REC; q++)
    if (q != 2) fin = fin && finite_d(r[q]);
  const int rfl = fl | (fin ? 0 : FLAG_NONFINITE);
  r[2] = bits_to_double(rfl);
  double2 *dst = reinterpret_cast<double2 *>(rec + (size_t)REC * i);
#pragma unroll
  for (int q = 0; q < REC / 2; q++) dst[q] = make_double2(r[2 * q], r[2 * q + 1]);
  posd[i] = make_double2(x, y);
  vfl[i] = rfl;
}

// ------------------------------------------------------------------------------- interact
// the sums kernels keep at least 3 waves per SIMD (at most 168 VGPRs): a 150k-particle step
// has about 3 waves per SIMD in all, so all of them are resident at once
#define SUMS_OCCUPANCY __attribute__((amdgpu_waves_per_eu(3)))

// compensated (Kahan) accumulation in single precision
struct kahan {
  float s = 0.0f, c = 0.0f;
  __device__ inline void add(float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ inline double value() const { return (double)s - (double)c; }
};

// FP32 neighbour sums with Kahan compensation (the same pair terms as pair_terms, rearranged
// for single precision: hardware reciprocals and reciprocal square roots instead of divisions,
// the particle mass and the factor 2 of Brookshaw taken out of the sums, e.gradW/r of Brookshaw
// written as der*fac, which is its exact value, and the stress and viscosity terms of each
// acceleration component added before their one compensated sum). Pair distances are formed in
// FP64 and then rounded.
template <bool COMPACT>
__device__ inline void sums_f32(const dev_consts &k, int i, const double2 *__restrict__ posd,
                                const float *__restrict__ rec32, const int *__restrict__ key_sorted,
                                const int *__restrict__ offsets, const cand_src &cs, bool tool, double ta_d,
                                bool run_mech, bool run_pse, pair_acc &acc, double &tpse) {
  const float4 *ri = reinterpret_cast<const float4 *>(rec32 + (size_t)16 * i);
  const float4 a0 = ri[0], a1 = ri[1], a2 = ri[2], a3 = ri[3];
  const double2 pi = posd[i];
  const float ui = a0.x, vi_ = a0.y, hi = a0.z, rhoi = a0.w;
  const float Axx_i = a1.y, Axy_i = a1.z, Ayy_i = a1.w, Rxx_i = a2.x, Rxy_i = a2.y, Ryy_i = a2.z, ci = a2.w;
  const float Tpred = a3.x;
  const float ta = (float)ta_d;
  const float Ti = ta != 0.0f ? Tpred : 0.0f;
  const float h1 = 1.0f / hi;
  const float fourh2 = 4.0f * hi * hi;
  const float fac = (float)PI_FAC * h1 * h1;
  const float inv_wd = (float)(1.0 / k.wdeltap);
  const float eta2 = (float)(k.av_eta * k.av_eta);
  const float alpha = (float)k.av_alpha, beta = (float)k.av_beta, xeps = (float)k.xsph_eps;
  const float inv_h2 = h1 * h1;
  const float pse_c = (float)(4. / M_PI) * inv_h2 * inv_h2;
  const float massf = (float)k.mass;
  // the stress and viscosity sums enter the acceleration only as sd0 + sd1 + vt0 and
  // sd2 + sd3 + vt1 (actions_gpu.cu momentum), so they are accumulated as those two sums
  kahan vd0, vd1, vd2, vd3, ax, ay, xt0, xt1, tl, tp;
  const window w = make_window(k, offsets, key_sorted[i]);
  // the mechanical and Brookshaw terms of one in-support neighbour
  struct rec4 {
    float4 b0, b1, b2, b3;
  };
  auto load_rec = [&](int j, rec4 &r) {
    const float4 *rj = reinterpret_cast<const float4 *>(rec32 + (size_t)16 * j);
    r.b0 = rj[0]; r.b1 = rj[1]; r.b2 = rj[2]; r.b3 = rj[3];
  };
  auto pair = [&](float dx, float dy, float rr2, const rec4 &r) {
    const float4 b0 = r.b0, b1 = r.b1, b2 = r.b2, b3 = r.b3;
    const int fj = __float_as_int(b3.y);
    const float uj = b0.x, vj = b0.y, hj_ = b0.z, rhoj = b0.w, volj = b1.x;
    const float inv_r = __builtin_amdgcn_rsqf(rr2);
    const float q = rr2 * inv_r * h1;
    float val, der;  // der = dW/dr / r without the factor fac
    if (q > 1.0f) {
      const float m = 2.0f - q, m2 = m * m;
      val = 0.25f * m2 * m2 * m2;
      der = -0.75f * m2 * h1 * inv_r;
    } else {
      val = 1.0f - 1.5f * q * q * (1.0f - 0.5f * q);
      der = -3.0f * q * (1.0f - 0.75f * q) * h1 * inv_r;
    }
    const float w = val * fac, gr = der * fac;
    const float w_x = gr * dx, w_y = gr * dy;
    if (!tool && !(fj & MFC_FLAG_TOOL)) {
      const float tx = w_x * volj, ty = w_y * volj;
      const float du = uj - ui, dv = vj - vi_;
      vd0.add(du * tx); vd1.add(du * ty); vd2.add(dv * tx); vd3.add(dv * ty);
      float fab = w * inv_wd;
      fab *= fab;
      fab *= fab;
      const float sx = Axx_i + b1.y + fab * (Rxx_i + b2.x);
      const float sy = Axy_i + b1.z + fab * (Rxy_i + b2.y);
      const float sz = Ayy_i + b1.w + fab * (Ryy_i + b2.z);
      float fx = __builtin_fmaf(sx, w_x, sy * w_y), fy = __builtin_fmaf(sy, w_x, sz * w_y);
      const float vijx = -du, vijy = -dv;
      const float vdx = __builtin_fmaf(vijx, dx, vijy * dy);
      const float inv_rhoij = 2.0f * __builtin_amdgcn_rcpf(rhoi + rhoj);
      if (vdx < 0.0f) {
        const float cij = 0.5f * (ci + b2.w);
        const float hij = 0.5f * (hi + hj_);
        const float muij = hij * vdx * __builtin_amdgcn_rcpf(__builtin_fmaf(eta2 * hij, hij, rr2));
        const float piij = (-alpha * cij * muij + beta * muij * muij) * inv_rhoij;
        fx = __builtin_fmaf(-piij, w_x, fx);
        fy = __builtin_fmaf(-piij, w_y, fy);
      }
      ax.add(fx); ay.add(fy);
      const float xs = -xeps * w * inv_rhoij;
      xt0.add(xs * vijx); xt1.add(xs * vijy);
    }
    if (ta != 0.0f) tl.add(volj * (Ti - b3.x) * gr);
  };
  // what the walk loads per candidate ahead of its evaluation: FP64 position and flags (the
  // temperature and density for PSE are read when the candidate is evaluated)
  struct cand {
    double2 p;
    int f;
  };
  auto load = [&](int j, cand &c) {
    c.p = cs.pos[j - cs.base];
    c.f = cs.fl[j - cs.base];
  };
  // the PSE term of every window member, and the support test; true: j is a mechanical pair
  auto visit = [&](int j, const cand &c, float &dx, float &dy, float &rr2) -> bool {
    dx = (float)(pi.x - c.p.x);
    dy = (float)(pi.y - c.p.y);
    rr2 = __builtin_fmaf(dx, dx, dy * dy);
    if (run_pse) {
      const float Tj = rec32[(size_t)16 * j + 12], rhoj = rec32[(size_t)16 * j + 3];
      const float w2 = pse_c * __expf(-rr2 * inv_h2);
      tp.add((Tj - Tpred) * w2 * (massf * __builtin_amdgcn_rcpf(rhoj)));
    }
    if (!run_mech) return false;
    if (rr2 >= fourh2 || rr2 < 1e-8f) return false;
    return !(c.f & MFC_FLAG_BLANKED);
  };
  // the window in groups of candidates, the loads of a group issued before its first candidate
  // is evaluated; the order of evaluation is the candidates' order
  if constexpr (!COMPACT) {
    // direct walk: the support tests of a group come first, then the records of the group's
    // in-support candidates are loaded together, then their pairs are evaluated in order
    for (int t0 = 0; t0 < w.total; t0 += WALK_U32) {
      int j[WALK_U32];
      cand c[WALK_U32];
      bool in[WALK_U32];
      float dx[WALK_U32], dy[WALK_U32], rr2[WALK_U32];
      rec4 r[WALK_U32];
#pragma unroll
      for (int u = 0; u < WALK_U32; u++) {
        j[u] = w.at(t0 + u < w.total ? t0 + u : w.total - 1);
        load(j[u], c[u]);
      }
#pragma unroll
      for (int u = 0; u < WALK_U32; u++) in[u] = t0 + u < w.total && visit(j[u], c[u], dx[u], dy[u], rr2[u]);
#pragma unroll
      for (int u = 0; u < WALK_U32; u++)
        if (in[u]) load_rec(j[u], r[u]);
#pragma unroll
      for (int u = 0; u < WALK_U32; u++)
        if (in[u]) pair(dx[u], dy[u], rr2[u], r[u]);
    }
  } else {
    // masked walk: the support tests of up to 64 candidates set the bits of a per-lane mask,
    // then the candidates of the set bits are evaluated in candidate order, the loads of the
    // next one issued before the current one is evaluated. The wave runs the pair body about as
    // often as its busiest lane has pairs instead of once per window candidate, with no list in
    // memory (the sorted index of a bit is recomputed by window::at).
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
        for (int u = 0; u < WALK_U; u++) {
          float dx, dy, rr2;
          if (u0 + u < n && visit(j[u], c[u], dx, dy, rr2)) m |= 1ull << (u0 + u);
        }
      }
      double2 pn;
      rec4 rn;
      if (m) {
        const int jj = w.at(c0 + __builtin_ctzll(m));
        pn = cs.pos[jj - cs.base];
        load_rec(jj, rn);
      }
      while (m) {
        m &= m - 1ull;
        const double2 pj = pn;
        const rec4 r = rn;
        if (m) {
          const int jj = w.at(c0 + __builtin_ctzll(m));
          pn = cs.pos[jj - cs.base];
          load_rec(jj, rn);
        }
        const float dx = (float)(pi.x - pj.x), dy = (float)(pi.y - pj.y);
        pair(dx, dy, __builtin_fmaf(dx, dx, dy * dy), r);
      }
    }
  }
  acc.vd0 = vd0.value(); acc.vd1 = vd1.value(); acc.vd2 = vd2.value(); acc.vd3 = vd3.value();
  acc.sd0 = k.mass * ax.value(); acc.sd1 = 0.;
  acc.sd2 = k.mass * ay.value(); acc.sd3 = 0.;
  acc.vt0 = 0.; acc.vt1 = 0.;
  acc.xt0 = k.mass * xt0.value(); acc.xt1 = k.mass * xt1.value();
  acc.tl = 2.0 * tl.value();
  tpse = tp.value();
}

// a particle's Johnson-Cook return, left for the whole wave's secant after update_particle: the
// yield test's inputs and the state the return starts from (the stress, plastic strain and rate
// and temperature that update_particle stored are replaced once the return is done)
struct deferred_return {
  bool need;   // the particle returns this step
  bool fixed;  // the thermal boundary condition resets its temperature after the return
  rr_input in;
  double rho, sxx, sxy, syy, szz, eps, epsd, T;
};

// everything after the neighbour sums for particle i (sorted index), in the reference's order;
// returns the particle's fc and ft for the tool force. The secant of a Johnson-Cook return runs
// for the whole wave after this call (finish_return); the other results do not depend on it
// reads particle s of the unsorted start state src (s = perm[i]) and writes the updated
// particle to position i of dst: the gather into sorted order happens here, once
__device__ inline void update_particle(const dev_consts &k, int i, int s, const pair_acc &acc, double tpse,
                                       const dstate &src, const dstate &dst, const dpositional &pos,
                                       const mfc_tool_state &tool_now, double fsum[4], double &x_out, double &y_out,
                                       int &fl_out, deferred_return &dr) {
  const bool pse = k.pse != 0;
  int fl = src.flags[s] & FLAG_MASK;
  const bool blank = fl & MFC_FLAG_BLANKED, tool = fl & MFC_FLAG_TOOL, fixed = fl & MFC_FLAG_FIXED;
  const double ta = pse ? 0. : (tool ? k.alpha_tool : k.alpha_wp);
  const bool run_mech = !blank && (!pse || !tool);
  const bool run_pse = !blank && pse && k.thermal_wp;
  // state at the start of the step (sorted) and previous derivatives
  const double x0 = src.px[s], y0 = src.py[s], u0 = src.vx[s], v0 = src.vy[s];
  const double sxx0 = src.sxx[s], sxy0 = src.sxy[s], syy0 = src.syy[s], szz0 = src.szz[s];
  const double rho0 = src.rho[s], T0 = src.T[s];
  double ptx = src.ptx[s], pty = src.pty[s], vtx = src.vtx[s], vty = src.vty[s];
  double rhot = src.rhot[s], stxx = src.stxx[s], stxy = src.stxy[s], styy = src.styy[s], stzz = src.stzz[s];
  double Tt = src.Tt[s];
  // grid_gpu_rothlin's sort leaves ft in place: the contact reads the history at the sorted index
  const int sft = k.rothlin ? i : s;
  double fcx = src.fcx[s], fcy = src.fcy[s], ftx = src.ftx[sft], fty = src.fty[sft];
  double eps = src.eps[s], epsd = src.epsd[s];
  // predicted state (same expressions as predict)
  double pu = u0, pv = v0, psxx = sxx0, psxy = sxy0, psyy = syy0, pszz = szz0, prho = rho0;
  if (!blank) {
    const double hdt = 0.5 * k.dt;
    pu = u0 + hdt * vtx; pv = v0 + hdt * vty;
    psxx = sxx0 + hdt * stxx; psxy = sxy0 + hdt * stxy; psyy = syy0 + hdt * styy; pszz = szz0 + hdt * stzz;
    prho = rho0 + hdt * rhot;
  }
  // interaction outputs (interactions_gpu.cu:334-344, :143)
  if (run_mech) {
    ptx = acc.xt0; pty = acc.xt1; vtx = acc.vt0; vty = acc.vt1;
    if (ta != 0.) Tt = ta * acc.tl;
  }
  if (run_pse) {
    double tflag = tool ? 1. : 0.;
    if (k.pse_quirk && i >= k.N / 2) tflag = 0.;
    Tt = ((tflag == 1.) ? k.alpha_tool : k.alpha_wp) * tpse;
  }
  if (!tool) {
    // Jaumann rate, continuity, momentum, advection (actions_gpu.cu:99-224)
    jaumann(k.G, acc.vd0, acc.vd1, acc.vd2, acc.vd3, psxx, psxy, psyy, pszz, stxx, stxy, styy, stzz);
    rhot = -prho * (acc.vd0 + acc.vd3);
    vtx += acc.sd0 + acc.sd1 + fcx / k.mass + ftx / k.mass;
    vty += acc.sd2 + acc.sd3 + fcy / k.mass + fty / k.mass;
    ptx = ptx + pu;
    pty = pty + pv;
  }
  // leap-frog corrector (leap_frog.cu:68-92)
  double x = x0, y = y0, u = u0, v = v0, sxx = sxx0, sxy = sxy0, syy = syy0, szz = szz0, rho = rho0, T = T0;
  if (!blank) {
    x = x0 + k.dt * ptx; y = y0 + k.dt * pty;
    u = u0 + k.dt * vtx; v = v0 + k.dt * vty;
    sxx = sxx0 + k.dt * stxx; sxy = sxy0 + k.dt * stxy; syy = syy0 + k.dt * styy; szz = szz0 + k.dt * stzz;
    rho = rho0 + k.dt * rhot;
    T = T0 + k.dt * Tt;
  }
  // frictional heating with the positional normal of the previous step (actions_gpu.cu:135-162)
  if (k.fric_heat && !tool) {
    const double fmag = tsqrt(k, ftx * ftx + fty * fty);
    if (fmag != 0.) {
      const double nx = pos.nx[i], ny = pos.ny[i];
      const double dvx = u - k.tool_vx, dvy = v - k.tool_vy;
      const double dvn = dvx * nx + dvy * ny;
      const double rx = dvx - dvn, ry = dvy - dvn;
      const double vmag = tsqrt(k, rx * rx + ry * ry);
      T += k.eta_heat * k.dt * fmag * vmag / (k.cp * k.mass);
    }
  }
  if (!tool && k.plastic) {
    dr.need = plasticity_test(k, pos.p[i], sxx, sxy, syy, szz, eps, epsd, T, dr.in);
    if (dr.need) {
      dr.fixed = fixed;
      dr.rho = rho; dr.sxx = sxx; dr.sxy = sxy; dr.syy = syy; dr.szz = szz;
      dr.eps = eps; dr.epsd = epsd; dr.T = T;
    }
  }
  if (tool) {  // update_tool_particles (tool_gpu.cu:210-221)
    x += k.dt * k.tool_vx;
    y += k.dt * k.tool_vy;
  }
  if (!tool) {
    double nx, ny;
    contact_force(k, tool_now, x, y, u, v, fcx, fcy, ftx, fty, nx, ny);
    pos.nx[i] = nx; pos.ny[i] = ny;
    if (fixed) {  // actions_gpu.cu:296-311
      u = 0.; v = 0.; fcx = 0.; fcy = 0.; ptx = 0.; pty = 0.; vtx = 0.; vty = 0.;
    }
  }
  if (k.thermal_bc && fixed) T = k.T_init;  // actions_gpu.cu:287-294
  if (!blank) {  // actions_gpu.cu:317-349
    bool bad = !finite_d(ptx) || !finite_d(pty) || !finite_d(vtx) || !finite_d(vty) || !finite_d(stxx) ||
               !finite_d(stxy) || !finite_d(styy) || !finite_d(stzz) || !finite_d(rhot) || !finite_d(Tt);
    if (bad) fl |= MFC_FLAG_BLANKED;
  }
  dst.px[i] = x; dst.py[i] = y; dst.vx[i] = u; dst.vy[i] = v;
  dst.sxx[i] = sxx; dst.sxy[i] = sxy; dst.syy[i] = syy; dst.szz[i] = szz;
  dst.rho[i] = rho; dst.T[i] = T; dst.eps[i] = eps; dst.epsd[i] = epsd; dst.h[i] = src.h[s];
  dst.fcx[i] = fcx; dst.fcy[i] = fcy; dst.ftx[i] = ftx; dst.fty[i] = fty;
  dst.ptx[i] = ptx; dst.pty[i] = pty; dst.vtx[i] = vtx; dst.vty[i] = vty;
  dst.rhot[i] = rhot; dst.stxx[i] = stxx; dst.stxy[i] = stxy; dst.styy[i] = styy; dst.stzz[i] = stzz;
  dst.Tt[i] = Tt;
  dst.flags[i] = fl;
  dst.id[i] = src.id[s];
  fsum[0] = fcx; fsum[1] = fcy; fsum[2] = ftx; fsum[3] = fty;
  x_out = x; y_out = y; fl_out = fl;
}

// the return of a particle that needed one, with the secant's dl: the stress, plastic strain and
// rate and temperature at sorted index i (the thermal boundary condition after the return, as in
// update_particle)
__device__ inline void finish_return(const dev_consts &k, int i, double dl, deferred_return &dr, const dstate &dst) {
  plasticity_finish(k, dr.rho, dl, dr.in, dr.sxx, dr.sxy, dr.syy, dr.szz, dr.eps, dr.epsd, dr.T);
  if (k.thermal_bc && dr.fixed) dr.T = k.T_init;  // actions_gpu.cu:287-294
  dst.sxx[i] = dr.sxx; dst.sxy[i] = dr.sxy; dst.syy[i] = dr.syy; dst.szz[i] = dr.szz;
  dst.eps[i] = dr.eps; dst.epsd[i] = dr.epsd; dst.T[i] = dr.T;
}

// the deferred returns of the wave (all 64 lanes together); dl_out (the reverse mode's replay,
// the same in every thread): each particle's secant result, -dl where its last update folded,
// 0 for no return. F32: the FP32C path's secant (rr_secant_wave<true>)
template <bool F32>
__device__ inline void wave_returns(const dev_consts &k, int i, deferred_return &dr, const dstate &dst,
                                    double *__restrict__ dl_out) {
  bool folded;
  const double dl = rr_secant_wave<F32>(k, dr.need, dr.in, dr.eps, dr.T, folded);
  if (dr.need) finish_return(k, i, dl, dr, dst);
  if (dl_out && i < k.N) dl_out[i] = dr.need ? (folded ? -dl : dl) : 0.;
}

// the neighbour sums of particle i as pair_acc + PSE sum
template <bool F32, bool COMPACT>
__device__ inline void particle_sums(const dev_consts &k, int i, const double *__restrict__ rec,
                                     const double2 *__restrict__ posd, const float *__restrict__ rec32,
                                     const int *__restrict__ key_sorted, const int *__restrict__ offsets,
                                     const cand_src &cs, int fl, pair_acc &acc, double &tpse) {
  const bool pse = k.pse != 0;
  const bool blank = fl & MFC_FLAG_BLANKED, tool = fl & MFC_FLAG_TOOL;
  const double ta = pse ? 0. : (tool ? k.alpha_tool : k.alpha_wp);
  const bool run_mech = !blank && (!pse || !tool);
  const bool run_pse = !blank && pse && k.thermal_wp;
  acc = {0., 0., 0., 0., 0., 0., 0., 0., 0., 0., 0., 0., 0.};
  tpse = 0.;
  if (run_mech || run_pse) {
    if constexpr (F32)
      sums_f32<COMPACT>(k, i, posd, rec32, key_sorted, offsets, cs, tool, ta, run_mech, run_pse, acc, tpse);
    else
      sums_f64<COMPACT>(k, i, rec, key_sorted, offsets, cs, tool, ta, run_mech, run_pse, acc, tpse);
  }
}

// partial sums of the tool force per wave of 64 particles (storage order): the halving-stride
// tree over the wave's lanes (lane k adds lane k + st for st = 32, 16, ..., 1, so lane 0 ends with
// the tree's root), then lane 0 writes the wave's four sums at partials[4 * wave]. Every lane of
// the wave calls it (lanes past N carry zeros).
__device__ inline void force_partials(int wave, const double fsum[4], double *__restrict__ partials) {
  double v0 = fsum[0], v1 = fsum[1], v2 = fsum[2], v3 = fsum[3];
#pragma unroll
  for (int st = 32; st > 0; st >>= 1) {
    v0 = v0 + __shfl_down(v0, st, 64);
    v1 = v1 + __shfl_down(v1, st, 64);
    v2 = v2 + __shfl_down(v2, st, 64);
    v3 = v3 + __shfl_down(v3, st, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    partials[4 * wave + 0] = v0; partials[4 * wave + 1] = v1;
    partials[4 * wave + 2] = v2; partials[4 * wave + 3] = v3;
  }
}

// LDS copy of the window candidates of a tile (positions and flags), when they fit
struct stage_buf {
  double2 pos[STAGE_CAP];
  int fl[STAGE_CAP];
};

// [lo, hi): the sorted indices of every window candidate of the particles of tile t. The first
// and the last cell of a particle's window (column-major cell ids) do not decrease along the
// sorted order, so the first particle's window start and the last particle's window end bound
// the tile; where a window is clipped at the first or last two columns of the grid the bound
// is the start or the end of the table instead.
__device__ inline int2 tile_span(const dev_consts &k, int t, const int *__restrict__ key_sorted,
                                 const int *__restrict__ offsets) {
  const int first = t * TILE;
  const int last = (first + TILE < k.N ? first + TILE : k.N) - 1;
  const int hf = key_sorted[first], hl = key_sorted[last];
  const int gif = hf / k.gny, gjf = hf - gif * k.gny;
  const int gil = hl / k.gny, gjl = hl - gil * k.gny;
  const int lo = gif - 2 < 0 ? 0 : offsets[(gif - 2) * k.gny + (gjf - 2 < 0 ? 0 : gjf - 2)];
  const int hi = gil + 3 > k.gnx ? k.N : offsets[(gil + 2) * k.gny + (gjl + 3 > k.gny ? k.gny : gjl + 3)];
  return make_int2(lo, hi);
}

// copies the tile's candidates into st when they fit (the same decision in every thread) and
// returns the sorted index of st's first entry, or -1 when they do not fit; ends with a
// work-group barrier
__device__ inline int stage_tile(const dev_consts &k, int t, const drecords &r, const int *__restrict__ key_sorted,
                                 const int *__restrict__ offsets, stage_buf &st) {
  const int2 sp = tile_span(k, t, key_sorted, offsets);
  const int n = sp.y - sp.x;
  const int cap = k.stage_limit > 0 && k.stage_limit < STAGE_CAP ? k.stage_limit : STAGE_CAP;
  int lo = -1;
  if (n <= cap) {
    for (int q = threadIdx.x; q < n; q += TILE) {
      st.pos[q] = r.posd[sp.x + q];
      st.fl[q] = r.vfl[sp.x + q];
    }
    lo = sp.x;
  }
  __syncthreads();
  return lo;
}

// the next step's first binning pass for the updated particle i (its storage position in the
// next step): cell id and arrival slot, as hash_count would take them from the updated state;
// bin == false: the next step bins its start state itself (the bins are passed by value so that
// they stay in registers)
__device__ inline void next_cell(const dev_consts &k, int i, bool live, double x, double y, int fl, const dbins next,
                                 bool bin) {
  if (!bin) return;  // the same in every thread
  const int h = live ? particle_cell(k, x, y, fl, next.key_sorted[i], next.clamped) : -1;
  count_cell(i, h, next.key, next.bsum, next.clist, next.ccap);
}

// fused: sums + update in one pass (registers of both live in one kernel)
template <bool F32, bool COMPACT>
__device__ inline void interact(const dev_consts &k, int t, const drecords &r, const int *__restrict__ key_sorted,
                                const int *__restrict__ offsets, const int *__restrict__ perm, const dstate &src,
                                const dstate &dst, const dpositional &pos,
                                const mfc_tool_state *__restrict__ tool_now, const dforce &fo, const dbins next,
                                bool bin_next, double *__restrict__ acc_out) {
  __shared__ stage_buf st;
  const int lo = stage_tile(k, t, r, key_sorted, offsets, st);
  const int i = t * TILE + threadIdx.x;
  double fsum[4] = {0., 0., 0., 0.};
  double xn = 0., yn = 0.;
  int fln = 0;
  deferred_return dr;
  dr.need = false;
  if (i < k.N) {
    const int s = perm[i];
    const int fl = src.flags[s] & FLAG_MASK;
    pair_acc acc;
    double tpse;
    // two call sites, so that each walk reads one kind of memory (LDS or global)
    if (lo >= 0)
      particle_sums<F32, COMPACT>(k, i, r.rec, r.posd, r.rec32, key_sorted, offsets, cand_src{st.pos, st.fl, lo}, fl,
                                  acc, tpse);
    else
      particle_sums<F32, COMPACT>(k, i, r.rec, r.posd, r.rec32, key_sorted, offsets, cand_src{r.posd, r.vfl, 0}, fl,
                                  acc, tpse);
    if (acc_out) store_acc(acc_out, k.N, i, acc, tpse);  // the same in every thread
    update_particle(k, i, s, acc, tpse, src, dst, pos, *tool_now, fsum, xn, yn, fln, dr);
  }
  wave_returns<F32>(k, i, dr, dst, acc_out ? acc_out + (size_t)ACC_N * k.N : nullptr);
  next_cell(k, i, i < k.N, xn, yn, fln, next, bin_next);
  force_partials(i >> 6, fsum, fo.partials);
  if (t == 0 && threadIdx.x == 0) *fo.pending = 1;
}

// split, first half: the neighbour sums of every particle into acc (ACC_N arrays of N doubles)
template <bool F32, bool COMPACT>
__device__ inline void sums(const dev_consts &k, int t, const drecords &r, const int *__restrict__ key_sorted,
                            const int *__restrict__ offsets, const int *__restrict__ perm,
                            const int *__restrict__ src_flags, double *__restrict__ acc) {
  __shared__ stage_buf st;
  const int lo = stage_tile(k, t, r, key_sorted, offsets, st);
  const int i = t * TILE + threadIdx.x;
  if (i >= k.N) return;
  const int fl = src_flags[perm[i]] & FLAG_MASK;
  pair_acc a;
  double tpse;
  if (lo >= 0)
    particle_sums<F32, COMPACT>(k, i, r.rec, r.posd, r.rec32, key_sorted, offsets, cand_src{st.pos, st.fl, lo}, fl, a,
                                tpse);
  else
    particle_sums<F32, COMPACT>(k, i, r.rec, r.posd, r.rec32, key_sorted, offsets, cand_src{r.posd, r.vfl, 0}, fl, a,
                                tpse);
  store_acc(acc, k.N, i, a, tpse);
}

// split, second half: the per-particle update from the stored sums (dl_out: as wave_returns)
template <bool F32>
__device__ inline void update(const dev_consts &k, int t, const double *__restrict__ acc,
                              const int *__restrict__ perm, const dstate &src, const dstate &dst,
                              const dpositional &pos, const mfc_tool_state *__restrict__ tool_now, const dforce &fo,
                              const dbins next, bool bin_next, double *__restrict__ dl_out) {
  const int i = t * TILE + threadIdx.x;
  double fsum[4] = {0., 0., 0., 0.};
  double xn = 0., yn = 0.;
  int fln = 0;
  deferred_return dr;
  dr.need = false;
  if (i < k.N) {
    pair_acc a;
    double tpse;
    load_acc(acc, k.N, i, a, tpse);
    update_particle(k, i, perm[i], a, tpse, src, dst, pos, *tool_now, fsum, xn, yn, fln, dr);
  }
  wave_returns<F32>(k, i, dr, dst, dl_out);
  next_cell(k, i, i < k.N, xn, yn, fln, next, bin_next);
  force_partials(i >> 6, fsum, fo.partials);
  if (t == 0 && threadIdx.x == 0) *fo.pending = 1;
}

constexpr int FF_BATCH = 8;  // wave partials a thread of force_final loads at a time

// tool force of the step from the wave partials, when the interaction of a step has left them
// (pending): lane l adds the results of waves l, l + TILE, ... in ascending order, then the lane
// sums are reduced by halving strides; F = sum(fc) + sum(ft). All threads of the work-group
// call it.
__device__ inline void force_final(int ntiles, const dforce &fo) {
  if (*fo.pending == 0) return;  // the same value in every thread
  __shared__ double red[4][TILE];
  double a0 = 0., a1 = 0., a2 = 0., a3 = 0.;
  // each thread adds the wave partials q = threadIdx.x, + TILE, + 2 TILE, ... in that order (4
  // per tile), eight loads in flight at a time: this work-group is the first of the step's
  // predict launch, and one load at a time made it that launch's last
  const double4 *__restrict__ p4 = reinterpret_cast<const double4 *>(fo.partials);
  const int nq = 4 * ntiles;
  for (int q0 = threadIdx.x; q0 < nq; q0 += TILE * FF_BATCH) {
    double4 v[FF_BATCH];
#pragma unroll
    for (int u = 0; u < FF_BATCH; u++) v[u] = q0 + u * TILE < nq ? p4[q0 + u * TILE] : make_double4(0., 0., 0., 0.);
#pragma unroll
    for (int u = 0; u < FF_BATCH; u++)
      if (q0 + u * TILE < nq) {
        a0 = a0 + v[u].x; a1 = a1 + v[u].y; a2 = a2 + v[u].z; a3 = a3 + v[u].w;
      }
  }
  red[0][threadIdx.x] = a0; red[1][threadIdx.x] = a1;
  red[2][threadIdx.x] = a2; red[3][threadIdx.x] = a3;
  for (int s = TILE / 2; s > 0; s >>= 1) {
    __syncthreads();
    if (threadIdx.x < s) {
#pragma unroll
      for (int c = 0; c < 4; c++) red[c][threadIdx.x] = red[c][threadIdx.x] + red[c][threadIdx.x + s];
    }
  }
  if (threadIdx.x == 0) {
    const double F0 = red[0][0] + red[2][0], F1 = red[1][0] + red[3][0];
    fo.force_last[0] = F0;
    fo.force_last[1] = F1;
    const int64_t n = *fo.ctr;
    if (n >= 0 && n < fo.cap) { fo.hist[2 * n] = F0; fo.hist[2 * n + 1] = F1; }
    *fo.ctr = n + 1;
    *fo.pending = 0;
  }
}

// ------------------------------------------------------------------------------- tool wear
// Usui wear rate (tool_wear.cu:20-60) of the current state: particles count as in contact when
// |fc.x|, |fc.y| or |ft.x| is non-zero (the reference tests fc.y twice and never ft.y); the
// relative velocity is v - v*n with the component-wise product of its glm code, the pressure
// |fc| / (2 pi h), the rate K exp(alpha / T) p |v_r|. Per tile: min (from FLT_MAX, as the
// reference), max (from 0), sum and count over the tile's particles in contact, then one
// work-group folds the tiles in a fixed order. Min, max and count do not depend on the order;
// the sum is the fixed-order tree instead of the reference's sequential loop.
__device__ inline void wear_tree(double (*red)[TILE]) {
  for (int st = TILE / 2; st > 0; st >>= 1) {
    __syncthreads();
    if (threadIdx.x < st) {
      red[0][threadIdx.x] = fmin(red[0][threadIdx.x], red[0][threadIdx.x + st]);
      red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + st]);
      red[2][threadIdx.x] = red[2][threadIdx.x] + red[2][threadIdx.x + st];
      red[3][threadIdx.x] = red[3][threadIdx.x] + red[3][threadIdx.x + st];
    }
  }
}

__global__ __launch_bounds__(TILE) void k_wear_tiles(int N, int ntiles, dstate st, const double *__restrict__ nx,
                                                     const double *__restrict__ ny, double K, double alpha,
                                                     double tvx, double tvy, double *__restrict__ part) {
  const int t = xcd_tile(blockIdx.x, ntiles);
  if (t < 0) return;
  __shared__ double red[4][TILE];
  const int i = t * TILE + threadIdx.x;
  double wmin = (double)FLT_MAX, wmax = 0., wsum = 0., cnt = 0.;
  if (i < N) {
    const double fcx = st.fcx[i], fcy = st.fcy[i], ftx = st.ftx[i];
    if (fabs(fcx) > 0. || fabs(fcy) > 0. || fabs(ftx) > 0. || fabs(fcy) > 0.) {
      const double vx = st.vx[i] - tvx, vy = st.vy[i] - tvy;
      const double rx = vx - vx * nx[i], ry = vy - vy * ny[i];
      const double vmag = sqrt(rx * rx + ry * ry);
      const double fmag = sqrt(fcx * fcx + fcy * fcy);
      const double A = 2 * st.h[i] * M_PI;
      const double p = fmag / A;
      const double w = K * mfcm::exp(alpha / st.T[i]) * p * vmag;
      wmin = w; wmax = w; wsum = w; cnt = 1.;
    }
  }
  red[0][threadIdx.x] = wmin; red[1][threadIdx.x] = wmax; red[2][threadIdx.x] = wsum; red[3][threadIdx.x] = cnt;
  wear_tree(red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < 4; c++) part[4 * t + c] = red[c][0];
  }
}

__global__ __launch_bounds__(TILE) void k_wear_final(int ntiles, const double *__restrict__ part,
                                                     double *__restrict__ out) {
  __shared__ double red[4][TILE];
  double wmin = (double)FLT_MAX, wmax = 0., wsum = 0., cnt = 0.;
  for (int q = threadIdx.x; q < ntiles; q += TILE) {
    wmin = fmin(wmin, part[4 * q]); wmax = fmax(wmax, part[4 * q + 1]);
    wsum = wsum + part[4 * q + 2]; cnt = cnt + part[4 * q + 3];
  }
  red[0][threadIdx.x] = wmin; red[1][threadIdx.x] = wmax; red[2][threadIdx.x] = wsum; red[3][threadIdx.x] = cnt;
  wear_tree(red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < 4; c++) out[c] = red[c][0];
  }
}

// ------------------------------------------------------------------------------- batched
// Fields of the instance descriptors read through the constant address space: the compiler
// then takes the pointers loaded from them for global pointers and emits global loads, which
// count only against the vector-memory wait counter, instead of flat loads, which also hold up
// every wait for an LDS or scalar load. (A whole-struct copy loses that, hence the field lists.)
// The descriptors do not change while a kernel runs.
#define MFC_CONST(T) const __attribute__((address_space(4))) T
#define MFC_F(f) o.f = c->f
__device__ inline dstate load_state(MFC_CONST(dstate) * c) {
  dstate o;
  MFC_F(px); MFC_F(py); MFC_F(vx); MFC_F(vy); MFC_F(sxx); MFC_F(sxy); MFC_F(syy); MFC_F(szz); MFC_F(rho);
  MFC_F(T); MFC_F(eps); MFC_F(epsd); MFC_F(h); MFC_F(fcx); MFC_F(fcy); MFC_F(ftx); MFC_F(fty); MFC_F(ptx);
  MFC_F(pty); MFC_F(vtx); MFC_F(vty); MFC_F(rhot); MFC_F(stxx); MFC_F(stxy); MFC_F(styy); MFC_F(stzz); MFC_F(Tt);
  MFC_F(flags); MFC_F(id);
  return o;
}
__device__ inline dbins load_bins(MFC_CONST(dbins) * c) {
  dbins o;
  MFC_F(key); MFC_F(key_sorted); MFC_F(offsets); MFC_F(bucket); MFC_F(perm);
  MFC_F(spos); MFC_F(bsum); MFC_F(clist); MFC_F(ccap); MFC_F(offsets_r); MFC_F(clamped);
  return o;
}
__device__ inline dpositional load_pos(MFC_CONST(dpositional) * c) {
  dpositional o;
  MFC_F(nx); MFC_F(ny); MFC_F(p); MFC_F(rxx); MFC_F(rxy); MFC_F(ryy);
  return o;
}
__device__ inline drecords load_recs(MFC_CONST(drecords) * c) {
  drecords o;
  MFC_F(f32); MFC_F(rec); MFC_F(posd); MFC_F(vfl); MFC_F(rec32);
  return o;
}
__device__ inline dforce load_force(MFC_CONST(dforce) * c) {
  dforce o;
  MFC_F(partials); MFC_F(force_last); MFC_F(hist); MFC_F(cap); MFC_F(ctr); MFC_F(pending);
  return o;
}
__device__ inline MFC_CONST(inst_args) * const_args(const inst_args *A, int q) {
  return (MFC_CONST(inst_args) *)(A + q);
}
__device__ inline inst_args load_args(const inst_args *A, int q) {
  MFC_CONST(inst_args) *c = const_args(A, q);
  inst_args o;
  o.k = A[q].k;  // no pointers
  o.cur = load_state(&c->cur);
  o.nxt = load_state(&c->nxt);
  o.b = load_bins(&c->b);
  o.pos = load_pos(&c->pos);
  o.r = load_recs(&c->r);
  MFC_F(tool_in); MFC_F(tool_out);
  o.f = load_force(&c->f);
  MFC_F(acc); MFC_F(geo); MFC_F(ntiles); MFC_F(nscan);
  return o;
}
#undef MFC_F

// One launch per phase for all instances of a batch: block b takes work item xcd_tile(b) (the
// interaction kernels: xcd_chunk_tile(b)) of the (instance, block) table, and runs the device
// body with that instance's constants and buffers, so every instance computes exactly its own
// run.
__global__ __launch_bounds__(256) void kb_hash_count(const inst_args *__restrict__ A, const int2 *__restrict__ work,
                                                     int nwork) {
  const int w = xcd_tile(blockIdx.x, nwork);
  if (w < 0) return;
  const int2 it = work[w];
  const inst_args a = load_args(A, it.x);
  if (a.k.adaptive) return;  // its step starts with the geometry pass, which bins the state
  hash_count(a.k, it.y * 256 + threadIdx.x, a.cur.px, a.cur.py, a.cur.flags, a.b);
}

// Adaptive grid (grid_base::update_geometry, grid.cu:97-136), one work-group per instance: the
// box of all positions (blanked particles included, as thrust::min_element / max_element over
// the whole array) and the largest h give this step's geometry, which the work-group writes
// into the instance's descriptor for the step's later kernels, then it bins the start state with
// it (cell ids, arrival slots, cell and chunk counts). Minimum and maximum do not depend on the
// order they are taken in; NaN positions are passed over. A geometry that needs as many cells as
// the binning table holds is not taken (the reference asserts): the previous step's (kept in
// geo, which every step writes) stays and the step is counted in geo->overflow.
constexpr int GEO_THREADS = 1024;
__global__ __launch_bounds__(GEO_THREADS) void kb_geometry(inst_args *__restrict__ A) {
  // plain loads: this kernel writes the descriptor it reads (not through the constant space)
  const inst_args a = A[blockIdx.x];
  if (!a.k.adaptive) return;  // the same in every thread of the work-group
  __shared__ double red[5][GEO_THREADS / 64];
  __shared__ double geo_d[3];
  __shared__ int geo_i[3];
  const int N = a.k.N;
  const double inf = __builtin_huge_val();
  double x0 = inf, x1 = -inf, y0 = inf, y1 = -inf, hm = -inf;
  for (int i = threadIdx.x; i < N; i += GEO_THREADS) {
    const double x = a.cur.px[i], y = a.cur.py[i], h = a.cur.h[i];
    x0 = fmin(x0, x); x1 = fmax(x1, x);
    y0 = fmin(y0, y); y1 = fmax(y1, y);
    hm = fmax(hm, h);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    x0 = fmin(x0, __shfl_xor(x0, o, 64)); x1 = fmax(x1, __shfl_xor(x1, o, 64));
    y0 = fmin(y0, __shfl_xor(y0, o, 64)); y1 = fmax(y1, __shfl_xor(y1, o, 64));
    hm = fmax(hm, __shfl_xor(hm, o, 64));
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) { red[0][wv] = x0; red[1][wv] = x1; red[2][wv] = y0; red[3][wv] = y1; red[4][wv] = hm; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < GEO_THREADS / 64; q++) {
      x0 = fmin(x0, red[0][q]); x1 = fmax(x1, red[1][q]);
      y0 = fmin(y0, red[2][q]); y1 = fmax(y1, red[3][q]);
      hm = fmax(hm, red[4][q]);
    }
    double bx = x0 - 1e-6, by = y0 - 1e-6;
    const double ex = x1 + 1e-6, ey = y1 + 1e-6;
    double cell = hm * 2. / 2.;
    const double fnx = ceil((ex - bx) / cell), fny = ceil((ey - by) / cell);
    int gnx, gny;
    if (fnx >= 1. && fny >= 1. && fnx * fny < (double)a.k.ncell_table) {
      gnx = (int)fnx; gny = (int)fny;
    } else {  // keep the previous step's geometry (the instance's record: the descriptor of
              // this buffer parity holds the one of two steps before)
      bx = a.geo->bbmin_x; by = a.geo->bbmin_y; cell = a.geo->cell; gnx = a.geo->gnx; gny = a.geo->gny;
      a.geo->overflow = a.geo->overflow + 1;
    }
    geo_d[0] = bx; geo_d[1] = by; geo_d[2] = cell;
    geo_i[0] = gnx; geo_i[1] = gny; geo_i[2] = gnx * gny;
    dev_consts *kd = &A[blockIdx.x].k;
    kd->bbmin_x = bx; kd->bbmin_y = by; kd->cell = cell;
    kd->gnx = gnx; kd->gny = gny; kd->num_cell = gnx * gny;
    a.geo->bbmin_x = bx; a.geo->bbmin_y = by; a.geo->cell = cell;
    a.geo->gnx = gnx; a.geo->gny = gny; a.geo->num_cell = gnx * gny;
  }
  __syncthreads();
  dev_consts k = a.k;
  k.bbmin_x = geo_d[0]; k.bbmin_y = geo_d[1]; k.cell = geo_d[2];
  k.gnx = geo_i[0]; k.gny = geo_i[1]; k.num_cell = geo_i[2];
  for (int base = 0; base < N; base += GEO_THREADS)  // every wave runs every round (count_cell)
    hash_count(k, base + threadIdx.x, a.cur.px, a.cur.py, a.cur.flags, a.b);
}

__global__ __launch_bounds__(SCAN_THREADS) void kb_sort(const inst_args *__restrict__ A, const int2 *__restrict__ work,
                                                         int nwork) {
  const int w = xcd_tile(blockIdx.x, nwork);
  if (w < 0) return;
  const int2 it = work[w];
  MFC_CONST(inst_args) *c = const_args(A, it.x);
  const dev_consts &k = A[it.x].k;
  sort_chunk(it.y, k.N, k.ncell_table, load_bins(&c->b), k.sort_fast, k.sort_list);
}

template <bool F32, bool GATHER>
__global__ __launch_bounds__(TILE) void kb_predict(const inst_args *__restrict__ A, const int2 *__restrict__ work,
                                                   int nwork) {
  const int w = xcd_tile(blockIdx.x, nwork);
  if (w < 0) return;
  const int2 it = work[w];
  // records, bins and the start state through the constant address space (global loads and
  // stores); the other pointers through generic loads, where they are used (holding all of
  // them would take more scalar registers than there are)
  const inst_args &a = A[it.x];
  MFC_CONST(inst_args) *c = const_args(A, it.x);
  const dbins b = load_bins(&c->b);
  predict<F32, GATHER>(a.k, it.y, b, load_state(&c->cur), a.nxt, a.pos, load_recs(&c->r), a.tool_in, a.tool_out);
  // the chunk sums are consumed (the sort): cleared for the counts of the next step
  for (int q = it.y * TILE + threadIdx.x; q < a.nscan; q += a.ntiles * TILE) b.bsum[q] = 0;
  // the tool force of the previous step, if its partials are still pending (the first tile's
  // work-group is dispatched first, so the longer kernel hides the reduction)
  if (it.y == 0) force_final(a.ntiles, load_force(&c->f));
}

// The cell ranges of grid_gpu_rothlin (grid_gpu_rothlin.cu:28-127) rebuilt from the sorted cell
// ids: the lane at the end of each run of equal ids (where unique_by_key_copy starts the next run)
// writes the run's end into the offsets of its cell's successor and of every empty cell up to the
// next run's cell, which is the scatter of the run lengths and their scan over the empty cells
// between in one pass; the first lane writes 0 up to the first id, the last run fills the table's
// end with N. Every entry of offsets_r is written by exactly one lane; the same ranges as the
// counting sort's offsets.
__global__ __launch_bounds__(256) void kb_rothlin_cells(const inst_args *__restrict__ A, const int2 *__restrict__ work,
                                                        int nwork) {
  const int w = xcd_tile(blockIdx.x, nwork);
  if (w < 0) return;
  const int2 it = work[w];
  MFC_CONST(inst_args) *c = const_args(A, it.x);
  const int rothlin = A[it.x].k.rothlin, N = A[it.x].k.N, nc = A[it.x].k.ncell_table;
  if (!rothlin) return;
  const int i = it.y * 256 + threadIdx.x;
  if (i >= N) return;
  const dbins b = load_bins(&c->b);
  const int h = b.key_sorted[i];
  if (i == 0)
    for (int q = 0; q <= h; q++) b.offsets_r[q] = 0;
  const int hn = i + 1 < N ? b.key_sorted[i + 1] : nc;
  if (hn != h)
    for (int q = h + 1; q <= hn; q++) b.offsets_r[q] = i + 1;
}

template <bool F32, bool COMPACT>
__global__ __launch_bounds__(TILE) SUMS_OCCUPANCY void kb_interact(const inst_args *__restrict__ A, const int2 *__restrict__ work,
                                                    int nwork, int chunk) {
  const int w = xcd_chunk_tile(blockIdx.x, nwork, chunk);
  if (w < 0) return;
  const int2 it = work[w];
  const inst_args &a = A[it.x];
  MFC_CONST(inst_args) *c = const_args(A, it.x);
  const dbins b = load_bins(&c->b);
  // an adaptive grid bins the next step's start state after its geometry pass instead
  interact<F32, COMPACT>(a.k, it.y, load_recs(&c->r), b.key_sorted, a.k.rothlin ? b.offsets_r : b.offsets, b.perm,
                         a.cur, a.nxt, a.pos, a.tool_out, a.f, b, a.bin_next && !a.k.adaptive, a.acc_out);
}

template <bool F32, bool COMPACT>
__global__ __launch_bounds__(TILE) SUMS_OCCUPANCY void kb_sums(const inst_args *__restrict__ A, const int2 *__restrict__ work,
                                                int nwork, int chunk) {
  const int w = xcd_chunk_tile(blockIdx.x, nwork, chunk);
  if (w < 0) return;
  const int2 it = work[w];
  const inst_args &a = A[it.x];
  MFC_CONST(inst_args) *c = const_args(A, it.x);
  const dbins b = load_bins(&c->b);
  sums<F32, COMPACT>(a.k, it.y, load_recs(&c->r), b.key_sorted, a.k.rothlin ? b.offsets_r : b.offsets, b.perm,
                     a.cur.flags, c->acc);
}

template <bool F32>
__global__ __launch_bounds__(TILE) void kb_update(const inst_args *__restrict__ A, const int2 *__restrict__ work,
                                                  int nwork) {
  const int w = xcd_tile(blockIdx.x, nwork);
  if (w < 0) return;
  const int2 it = work[w];
  const inst_args &a = A[it.x];
  MFC_CONST(inst_args) *c = const_args(A, it.x);
  const dbins b = load_bins(&c->b);
  update<F32>(a.k, it.y, c->acc, b.perm, a.cur, a.nxt, a.pos, a.tool_out, a.f, b, a.bin_next && !a.k.adaptive,
              a.acc_out ? a.acc_out + (size_t)ACC_N * a.k.N : nullptr);
}

__global__ __launch_bounds__(TILE) void kb_force_final(const inst_args *__restrict__ A) {
  const inst_args a = load_args(A, blockIdx.x);
  force_final(a.ntiles, a.f);
}

__global__ void k_device_math(int fn, const double *__restrict__ x, const double *__restrict__ y,
                              double *__restrict__ out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a = x[i], b = y ? y[i] : 0., r = 0.;
  switch (fn) {
    case 0: r = mfcm::exp(a); break;
    case 1: r = mfcm::log(a); break;
    case 2: r = mfcm::pow(a, b); break;
    case 3: r = mfcm::atan2(a, b); break;
    case 4: r = mfcm::sin(a); break;
    case 5: r = mfcm::cos(a); break;
    case 6: r = (double)mfcm::atan2f(a, b); break;
    case 7: r = (double)mfcm::cosf(a); break;
    case 8: r = (double)mfcm::sinf(a); break;
    case 9: r = sqrt(a); break;
    case 10: r = f32sqrt(a); break;
    case 11: r = a / b; break;
    case 12: r = mfcm::tanh(a); break;
    default: r = 0.;
  }
  out[i] = r;
}

}  // namespace

int scan_blocks(int num_cell) { return (num_cell + SCAN_CHUNK - 1) / SCAN_CHUNK; }

static int padded_tiles(int ntiles) { return ((ntiles + 7) / 8) * 8; }

// the four instantiations of a kernel template <F32, COMPACT>
#define MFC_PICK(kern, f32, compact) \
  ((f32) ? ((compact) ? kern<true, true> : kern<true, false>) : ((compact) ? kern<false, true> : kern<false, false>))

void launch_tool_wear(const dev_consts &k, const dstate &st, const dpositional &pos, double K, double alpha,
                      double tvx, double tvy, double *part, double *out, hipStream_t s) {
  const int nt = (k.N + TILE - 1) / TILE;
  hipLaunchKernelGGL(k_wear_tiles, dim3(padded_tiles(nt)), dim3(TILE), 0, s, k.N, nt, st, pos.nx, pos.ny, K, alpha,
                     tvx, tvy, part);
  hipLaunchKernelGGL(k_wear_final, dim3(1), dim3(TILE), 0, s, nt, part, out);
}

void launch_device_math(int fn, const double *x, const double *y, double *out, int64_t n, hipStream_t s) {
  const int64_t g = (n + 255) / 256;
  hipLaunchKernelGGL(k_device_math, dim3((unsigned)g), dim3(256), 0, s, fn, x, y, out, n);
}

int scan_chunk_cells() { return SCAN_CHUNK; }
int chunk_list_cap(int N) { return N < 4 * SCAN_CHUNK ? N : 4 * SCAN_CHUNK; }

void launch_batch_hash(const batch_work &w, hipStream_t s) {
  hipLaunchKernelGGL(kb_hash_count, dim3(padded_tiles(w.ntiles)), dim3(256), 0, s, w.args, w.tiles, w.ntiles);
}

void launch_batch_binning(const batch_work &w, hipStream_t s) {
  const dim3 gc(padded_tiles(w.nchunks));
  if (w.adaptive)  // the descriptors are device memory of the batch; the geometry pass updates them
    hipLaunchKernelGGL(kb_geometry, dim3(w.n), dim3(GEO_THREADS), 0, s, const_cast<inst_args *>(w.args));
  hipLaunchKernelGGL(kb_sort, gc, dim3(SCAN_THREADS), 0, s, w.args, w.chunks, w.nchunks);
}

void launch_batch_predict(const batch_work &w, hipStream_t s) {
  const dim3 gt(padded_tiles(w.ntiles));
  if (w.f32) hipLaunchKernelGGL(HIP_KERNEL_NAME(kb_predict<true, false>), gt, dim3(TILE), 0, s, w.args, w.tiles, w.ntiles);
  else hipLaunchKernelGGL(HIP_KERNEL_NAME(kb_predict<false, false>), gt, dim3(TILE), 0, s, w.args, w.tiles, w.ntiles);
  if (w.rothlin) hipLaunchKernelGGL(kb_rothlin_cells, gt, dim3(256), 0, s, w.args, w.tiles, w.ntiles);
}

void launch_batch_predict_gather(const batch_work &w, hipStream_t s) {
  hipLaunchKernelGGL(HIP_KERNEL_NAME(kb_predict<false, true>), dim3(padded_tiles(w.ntiles)), dim3(TILE), 0, s, w.args,
                     w.tiles, w.ntiles);
}

// a batched launch over the work table, with the kernel's own start and end recorded in t0 and
// t1 when they are given
template <class K>
static void launch_timed(K kern, dim3 grid, const batch_work &w, hipStream_t s, hipEvent_t t0, hipEvent_t t1) {
  if (t0)
    hipExtLaunchKernelGGL(kern, grid, dim3(TILE), 0, s, t0, t1, 0, w.args, w.tiles, w.ntiles);
  else
    hipLaunchKernelGGL(kern, grid, dim3(TILE), 0, s, w.args, w.tiles, w.ntiles);
}

// the same for the interaction kernels, whose tile map takes the XCD chunk length
template <class K>
static void launch_chunked(K kern, const batch_work &w, hipStream_t s, hipEvent_t t0, hipEvent_t t1) {
  const dim3 grid(chunked_blocks(w.ntiles, w.v.chunk));
  if (t0)
    hipExtLaunchKernelGGL(kern, grid, dim3(TILE), 0, s, t0, t1, 0, w.args, w.tiles, w.ntiles, w.v.chunk);
  else
    hipLaunchKernelGGL(kern, grid, dim3(TILE), 0, s, w.args, w.tiles, w.ntiles, w.v.chunk);
}

void launch_batch_sums(const batch_work &w, hipStream_t s, hipEvent_t t0, hipEvent_t t1) {
  launch_chunked(MFC_PICK(kb_sums, w.f32, w.v.compact), w, s, t0, t1);
}

void launch_batch_update(const batch_work &w, hipStream_t s, hipEvent_t t0, hipEvent_t t1) {
  if (w.f32) launch_timed(kb_update<true>, dim3(padded_tiles(w.ntiles)), w, s, t0, t1);
  else launch_timed(kb_update<false>, dim3(padded_tiles(w.ntiles)), w, s, t0, t1);
}

void launch_batch_interact_fused(const batch_work &w, hipStream_t s, hipEvent_t t0, hipEvent_t t1) {
  launch_chunked(MFC_PICK(kb_interact, w.f32, w.v.compact), w, s, t0, t1);
}

void launch_batch_force(const batch_work &w, hipStream_t s) {
  hipLaunchKernelGGL(kb_force_final, dim3(w.n), dim3(TILE), 0, s, w.args);
}

// one step; the cell ids and counts of its start state come from the previous step's interaction
// (or launch_batch_hash), and its own interaction computes those of the next step
void launch_batch_step(const batch_work &w, hipStream_t s) {
  launch_batch_binning(w, s);  // also reduces the previous step's tool force
  launch_batch_predict(w, s);
  if (w.v.split) {
    launch_batch_sums(w, s);
    launch_batch_update(w, s);
  } else {
    launch_batch_interact_fused(w, s);
  }
}

}  // namespace mfc
