// ptc_noalloc.hip -- the allocation-free form of clcPrSrc.
//
// ptc_clc_prsrc keeps the per-workgroup maxima of |div| in the context's
// reduction buffer, which it allocates (or grows) on first use.  Callers that
// cannot call hipMalloc -- code under HIP graph capture, tight serving loops,
// harnesses that forbid it -- use ptc_clc_prsrc_noalloc instead: the same
// kernels, with the partials in the file-scope device array `scratch` below.
//
// Each GPU holds its own copy of `scratch` (16 MiB) from the library's load
// until the process ends.  In host code the array's name is not a device
// address; ptc_create resolves the address once per device with
// hipGetSymbolAddress, and the entry point passes that address to the
// kernels.  Every call on one GPU shares that GPU's array: calls that overlap
// on one GPU (contexts on different streams) race.  Nothing in it needs to
// start at zero: k_prsrc writes every partial that k_final_max reads.
#include <mutex>
#include <string>
#include <vector>

#include "ptc_internal.hpp"

namespace ptc {

// workgroup partials of one call: ceil(nx/64) * ceil(ny/4) * nz at most this
constexpr int kNoAllocPartials = 1 << 21;

static __device__ __align__(256) double scratch[kNoAllocPartials];

namespace {
std::mutex g_scratch_mu;
std::vector<double *> g_scratch_ptr;  // per device: the device address of scratch
}  // namespace

int noalloc_init_device(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  if (g_scratch_ptr[(size_t)device]) return PTC_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr) {
    (void)hipGetLastError();
    set_error(std::string("hipGetSymbolAddress of the allocation-free scratch failed: ") + hipGetErrorString(e));
    return PTC_E_HIP;
  }
  g_scratch_ptr[(size_t)device] = static_cast<double *>(p);
  return PTC_OK;
}

static double *scratch_ptr(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return device >= 0 && device < (int)g_scratch_ptr.size() ? g_scratch_ptr[(size_t)device] : nullptr;
}

}  // namespace ptc

using namespace ptc;

extern "C" {

int ptc_clc_prsrc_noalloc(ptc_ctx *ctx, int ux, int uy, int uz, int prsrc, double *divmax) {
  PTC_CHECK(ctx != nullptr, PTC_E_ARG, "null context");
  const int nb = prsrc_partials(ctx);
  PTC_CHECK(nb <= kNoAllocPartials, PTC_E_ARG,
            "grid too large for ptc_clc_prsrc_noalloc: ceil(nx/64)*ceil(ny/4)*nz = " + std::to_string(nb) +
                " workgroup partials, at most " + std::to_string(kNoAllocPartials));
  double *part = scratch_ptr(ctx->device);
  PTC_CHECK(part != nullptr, PTC_E_STATE, "allocation-free scratch not resolved on this device");
  double *u, *v, *w, *f;
  PTC_TRY(field_existing(ctx, ux, &u));
  PTC_TRY(field_existing(ctx, uy, &v));
  PTC_TRY(field_existing(ctx, uz, &w));
  PTC_TRY(field_existing(ctx, prsrc, &f));
  return launch_prsrc_part(ctx, u, v, w, f, divmax, true, part);
}

}  // extern "C"
