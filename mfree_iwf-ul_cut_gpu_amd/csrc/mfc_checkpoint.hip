// mfc_checkpoint.hip — binary checkpoint of a simulation handle (mfc_save / mfc_load, include/mfc.h).
//
// The reference has no restart format: its only state dump is the text dump_state of debug.cpp
// (debug.cpp:20-64), and a run always starts from setup. A checkpoint here holds everything a
// step reads that is not recomputed at the start of the next step, so that a loaded handle
// continues with the bits of the run it was saved from:
//   header       magic, format version, sizes, the creating configuration, launch tuning,
//                completed steps, accumulated tool wear, payload length and FNV-1a checksum
//   state        the current set of persistent particle fields (27 FP64 arrays, flags, ids),
//                in the handle's storage order
//   positional   contact normal, pressure and artificial stress (not permuted, as the
//                reference's particle_gpu::{n,p,R}); tool particles keep p and R across steps
//   sorted ids   the cell id each stored particle was last sorted by (blanked particles keep
//                theirs)
//   tool         the current tool state (moved incrementally every step)
//   forces       the last tool force and the force history
// Cell tables, neighbour records and the binning of the current state are recomputed by the
// next step.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mfc_sim.h"

using mfc::fail;

namespace {

constexpr char CKPT_MAGIC[8] = {'M', 'F', 'C', 'C', 'K', 'P', 'T', '1'};
constexpr uint32_t CKPT_VERSION = 1;

struct ckpt_header {
  char magic[8];
  uint32_t version;
  uint32_t header_bytes;
  int64_t n_particles;
  int64_t num_cell;
  int64_t step;
  int64_t hist_n;
  int32_t precision;
  int32_t pad;
  mfc_config cfg;
  mfc_tuning tuning;
  double wear_accum;
  uint64_t payload_bytes;
  uint64_t checksum;
};

uint64_t fnv1a(const unsigned char *p, size_t n, uint64_t h = 1469598103934665603ull) {
  for (size_t i = 0; i < n; i++) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

struct payload_layout {
  size_t state, positional, keys, tool, force_last, hist, total;
};

payload_layout layout(int64_t N, int64_t hist_n) {
  payload_layout L;
  L.state = state_block_bytes((int)N);
  L.positional = positional_block_bytes((int)N);
  L.keys = sizeof(int) * (size_t)N;
  L.tool = sizeof(mfc_tool_state);
  L.force_last = 2 * sizeof(double);
  L.hist = 2 * sizeof(double) * (size_t)hist_n;
  L.total = L.state + L.positional + L.keys + L.tool + L.force_last + L.hist;
  return L;
}

}  // namespace

extern "C" {

int mfc_save(mfc_sim *s, const char *path) {
  if (!s || !path) return fail(MFC_E_INVALID, "null argument");
  HIP_TRY(hipStreamSynchronize(s->stream));
  const int N = s->k.N;
  int64_t ctr = 0;
  HIP_TRY(hipMemcpy(&ctr, s->hist_ctr, sizeof(ctr), hipMemcpyDeviceToHost));
  const int64_t hist_n = ctr < HIST_CAP ? ctr : HIST_CAP;
  const payload_layout L = layout(N, hist_n);
  std::vector<unsigned char> buf(L.total);
  unsigned char *p = buf.data();
  // the current set is one contiguous block (carve_state), as is the positional block
  HIP_TRY(hipMemcpy(p, s->set[s->cur].px, L.state, hipMemcpyDeviceToHost));
  p += L.state;
  HIP_TRY(hipMemcpy(p, s->pos.nx, L.positional, hipMemcpyDeviceToHost));
  p += L.positional;
  HIP_TRY(hipMemcpy(p, s->bins.key_sorted, L.keys, hipMemcpyDeviceToHost));
  p += L.keys;
  HIP_TRY(hipMemcpy(p, s->tool + s->cur, L.tool, hipMemcpyDeviceToHost));
  p += L.tool;
  HIP_TRY(hipMemcpy(p, s->force_last, L.force_last, hipMemcpyDeviceToHost));
  p += L.force_last;
  if (hist_n > 0) HIP_TRY(hipMemcpy(p, s->hist, L.hist, hipMemcpyDeviceToHost));
  ckpt_header h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, CKPT_MAGIC, sizeof(h.magic));
  h.version = CKPT_VERSION;
  h.header_bytes = sizeof(h);
  h.n_particles = N;
  h.num_cell = s->k.num_cell;
  h.step = s->step;
  h.hist_n = hist_n;
  h.precision = s->precision;
  h.cfg = s->cfg;
  h.tuning = s->tuning;
  h.wear_accum = s->wear_accum;
  h.payload_bytes = L.total;
  h.checksum = fnv1a(buf.data(), buf.size());
  FILE *f = std::fopen(path, "wb");
  if (!f) return fail(MFC_E_INVALID, std::string("cannot open ") + path + " for writing");
  bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 && std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(MFC_E_INVALID, std::string("write of ") + path + " failed");
  return MFC_OK;
}

int mfc_load(const char *path, mfc_sim **out) {
  if (!path || !out) return fail(MFC_E_INVALID, "null argument");
  *out = nullptr;
  FILE *f = std::fopen(path, "rb");
  if (!f) return fail(MFC_E_INVALID, std::string("cannot open ") + path);
  ckpt_header h;
  if (std::fread(&h, sizeof(h), 1, f) != 1 || std::memcmp(h.magic, CKPT_MAGIC, sizeof(h.magic)) != 0 ||
      h.version != CKPT_VERSION || h.header_bytes != sizeof(h)) {
    std::fclose(f);
    return fail(MFC_E_INVALID, std::string(path) + " is not a checkpoint of this format version");
  }
  if (h.n_particles < 1 || h.hist_n < 0 || h.hist_n > HIST_CAP) {
    std::fclose(f);
    return fail(MFC_E_INVALID, std::string(path) + ": bad sizes in the header");
  }
  const payload_layout L = layout(h.n_particles, h.hist_n);
  if (L.total != h.payload_bytes) {
    std::fclose(f);
    return fail(MFC_E_INVALID, std::string(path) + ": payload size does not match the header");
  }
  std::vector<unsigned char> buf(L.total);
  const bool got = std::fread(buf.data(), 1, buf.size(), f) == buf.size();
  std::fclose(f);
  if (!got) return fail(MFC_E_INVALID, std::string(path) + " is truncated");
  if (fnv1a(buf.data(), buf.size()) != h.checksum) return fail(MFC_E_INVALID, std::string(path) + ": checksum mismatch");
  // a handle of the saved configuration, whose state is then replaced
  mfc_sim *s = nullptr;
  int rc = mfc_create(&h.cfg, &s);
  if (rc) return rc;
  if (s->k.N != h.n_particles || s->k.num_cell != h.num_cell || s->precision != h.precision) {
    mfc_destroy(s);
    return fail(MFC_E_INVALID, std::string(path) + ": the configuration no longer builds the saved sizes");
  }
  auto bail = [&](int code) {
    mfc_destroy(s);
    return code;
  };
  if ((rc = mfc_set_tuning(s, &h.tuning))) return bail(rc);
  const unsigned char *p = buf.data();
  s->cur = 0;
  if (hipMemcpy(s->set[0].px, p, L.state, hipMemcpyHostToDevice) != hipSuccess) return bail(fail(MFC_E_HIP, "upload"));
  p += L.state;
  if (hipMemcpy(s->pos.nx, p, L.positional, hipMemcpyHostToDevice) != hipSuccess) return bail(fail(MFC_E_HIP, "upload"));
  p += L.positional;
  if (hipMemcpy(s->bins.key_sorted, p, L.keys, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(MFC_E_HIP, "upload"));
  p += L.keys;
  mfc_tool_state t2[2];
  std::memcpy(&t2[0], p, sizeof(mfc_tool_state));
  t2[1] = t2[0];
  if (hipMemcpy(s->tool, t2, sizeof(t2), hipMemcpyHostToDevice) != hipSuccess) return bail(fail(MFC_E_HIP, "upload"));
  p += L.tool;
  if (hipMemcpy(s->force_last, p, L.force_last, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(MFC_E_HIP, "upload"));
  p += L.force_last;
  if (h.hist_n > 0 && hipMemcpy(s->hist, p, L.hist, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(MFC_E_HIP, "upload"));
  const int64_t ctr = h.hist_n;
  if (hipMemcpy(s->hist_ctr, &ctr, sizeof(ctr), hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(MFC_E_HIP, "upload"));
  s->step = h.step;
  s->wear_accum = h.wear_accum;
  s->keys_fresh = false;  // the next step bins the loaded state
  *out = s;
  return MFC_OK;
}

int mfc_get_config(const mfc_sim *s, mfc_config *cfg) {
  if (!s || !cfg) return fail(MFC_E_INVALID, "null argument");
  *cfg = s->cfg;
  return MFC_OK;
}

}  // extern "C"
