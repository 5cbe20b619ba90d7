// The allocation-free variant of lbm_reduce_moments (include/lbm_capi.h: lbm_reduce_moments_noalloc).
//
// lbm_reduce_moments sums per-workgroup partials into a device buffer that lbm_create allocates
// with the context.  Callers that may not allocate device memory at all (code between graph
// captures, serving loops, harnesses that forbid hipMalloc) get the same reduction with its
// partials and results in a fixed device array, `scratch`, that the library image holds on every
// device it runs on until the process ends.  The kernels and the launch geometry are those of
// lbm_reduce_moments, so the results are the same bits.
//
// The host never passes `scratch` itself to a kernel: in host code the name is only a stand-in.
// Its device address is looked up once per device with hipGetSymbolAddress (at lbm_create, with
// the context's device current) and that address is what the kernels receive.
#include <cmath>
#include <mutex>
#include <string>

#include "lbm_ctx.h"
#include "lbm_internal.h"

// 2^21 + 512 doubles = 16 MiB + 4 KiB: 4 partial sums per reduction workgroup plus 8 result slots.
static constexpr int64_t kScratchDoubles = ((int64_t)1 << 21) + 512;
static constexpr int64_t kResultSlots = 8;
static constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

std::mutex g_lock;                          // guards the table and serialises calls per process
double* g_scratch_ptr[kMaxDevices] = {};   // device address of `scratch` on each device

int nfail(lbm_ctx* c, int code, const std::string& msg) {
  c->err = msg;
  return code;
}

}  // namespace

// Called by lbm_create with `device` current.  Looks up the device address of `scratch` on that
// device once; later contexts on the device reuse it.
hipError_t lbm_noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_lock);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    return e != hipSuccess ? e : hipErrorInvalidSymbol;
  }
  g_scratch_ptr[device] = (double*)p;
  return hipSuccess;
}

// Largest number of reduction workgroups the array holds (4 partials each, after the result slots).
int64_t lbm_noalloc_max_partials() { return (kScratchDoubles - kResultSlots) / 4; }

extern "C" int lbm_reduce_moments_noalloc(lbm_ctx* c, double out[4]) {
  if (!c) return LBM_ERR_ARG;
  if (!out) return nfail(c, LBM_ERR_ARG, "output is null");
  if (!c->flags_set) return nfail(c, LBM_ERR_STATE, "boundary flags not set (lbm_set_flags)");
  if (!c->state_set) return nfail(c, LBM_ERR_STATE, "state not initialised (lbm_init_equilibrium / lbm_set_populations)");
  const int64_t np = lbm::dense_moment_partials(c->g);
  if (np > lbm_noalloc_max_partials())
    return nfail(c, LBM_ERR_ARG,
                 "lattice too large for lbm_reduce_moments_noalloc: " + std::to_string(np) +
                     " reduction workgroups, the fixed device array holds " + std::to_string(lbm_noalloc_max_partials()) +
                     " (about 2^27 cells); use lbm_reduce_moments");
  if (hipSetDevice(c->cfg.device) != hipSuccess) return nfail(c, LBM_ERR_HIP, "hipSetDevice failed");
  // One call at a time per process: every call on a device shares its array, and a call returns only
  // after its results are copied out, so holding the lock to the end keeps calls from overlapping.
  std::lock_guard<std::mutex> lk(g_lock);
  double* base = (c->cfg.device >= 0 && c->cfg.device < kMaxDevices) ? g_scratch_ptr[c->cfg.device] : nullptr;
  if (!base) return nfail(c, LBM_ERR_STATE, "the device array of lbm_reduce_moments_noalloc was not set up");
  double* res = base;                   // 8 result slots
  double* partials = base + kResultSlots;  // every slot the reduction reads is written by it first
  double h[4];
  hipError_t e = lbm::launch_dense_moments(lbm_state_geom(c), c->fp32, c->buf[c->cur], c->words, partials, res, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return nfail(c, LBM_ERR_HIP, std::string("no-alloc moment reduction: ") + hipGetErrorString(e));
  out[0] = (double)c->nfluid + h[0];
  out[1] = h[1];
  out[2] = h[2];
  out[3] = h[3];
  for (int k = 0; k < 4; k++)
    if (!std::isfinite(out[k])) return nfail(c, LBM_ERR_NUMERIC, "non-finite population detected");
  return LBM_OK;
}
