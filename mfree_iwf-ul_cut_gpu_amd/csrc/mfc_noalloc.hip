// mfc_noalloc.hip — mfc_tool_wear_noalloc: the Usui tool-wear evaluation of mfc_tool_wear for
// callers that cannot allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid allocation). mfc_tool_wear allocates its per-tile partial sums on its
// first call; this entry point keeps them in a file-scope device array of fixed size instead,
// one copy per GPU, held until the process ends. Every call on a GPU shares that GPU's array, so
// calls that overlap on one GPU (several host threads, or handles on different streams of one
// device) race; calls on one stream in sequence are safe. Nothing in the array has to start at
// zero: every call writes all the partials it reads.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "mfc_sim.h"

using mfc::fail;

namespace {

constexpr size_t WEAR_SCRATCH_DOUBLES = 4 * (size_t)MFC_NOALLOC_WEAR_MAX_TILES + 4;

// per-tile (min, max, sum, count) of the wear rate, then the 4 folded results
__device__ __align__(256) double scratch[WEAR_SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[MAX_DEVICES] = {};  // device address of `scratch` on each GPU

// the address of this GPU's copy of `scratch`, looked up once per device with the device current
int scratch_of_current_device(double **out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return fail(MFC_E_RANGE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return fail(MFC_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_dev[dev];
  return MFC_OK;
}

}  // namespace

extern "C" int mfc_tool_wear_noalloc(mfc_sim *s, double K, double alpha, const double tool_vel[2], double dt,
                                     double out[4]) {
  if (!s || !tool_vel || !out) return fail(MFC_E_INVALID, "null argument");
  if (s->ntiles > MFC_NOALLOC_WEAR_MAX_TILES)
    return fail(MFC_E_RANGE, "mfc_tool_wear_noalloc handles at most " + std::to_string(MFC_NOALLOC_WEAR_MAX_TILES) +
                                 " tiles of 256 particles (" + std::to_string(256 * MFC_NOALLOC_WEAR_MAX_TILES) +
                                 " particles); this handle has " + std::to_string(s->k.N) + ": use mfc_tool_wear");
  double *ws = nullptr;
  int rc = scratch_of_current_device(&ws);
  if (rc) return rc;
  double *res = ws + 4 * (size_t)MFC_NOALLOC_WEAR_MAX_TILES;
  mfc::launch_tool_wear(s->k, s->set[s->cur], s->pos, K, alpha, tool_vel[0], tool_vel[1], ws, res, s->stream);
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
