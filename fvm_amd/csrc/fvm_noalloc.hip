// An entry point that allocates no device memory: CALCCOEF with its partial sums in a
// file-scope device array instead of the context's reduction buffer.
//
// The array is this library's static device memory: each GPU the library runs on holds its own
// copy from module load to process end.  Its device address on the current GPU is looked up
// once per device (hipGetSymbolAddress) and kept; host code uses only that address, never the
// array's name.  Every call on one GPU shares the array: calls that overlap on one GPU (from
// two host threads, or on two streams) race.  The kernels write every partial they later read
// within the call, so nothing needs clearing between calls.
#include <mutex>

#include "fvm_ctx.h"

// 1 MiB: the CFL maxima (512) and five force sums per 256 wall triangles, enough for
// (NJ-1)(NK-1) up to about 3.3 million wall cells.
#define FVM_NOALLOC_DOUBLES (1L << 17)

static __device__ __align__(256) double scratch[FVM_NOALLOC_DOUBLES];

namespace {
constexpr int kMaxDev = 64;
std::mutex g_mu;
double* g_ptr[kMaxDev] = {};

// Device address of scratch on device dev (looked up with dev current, once).
int scratch_on(int dev, double** out) {
  if (dev < 0 || dev >= kMaxDev) return FVM_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    int cur = 0;
    HIP_CHECK(hipGetDevice(&cur));
    HIP_CHECK(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) {
      fprintf(stderr, "fvm_amd: no device address of the static scratch array on device %d (%s)\n",
              dev, hipGetErrorString(e));
      return FVM_ERR_HIP;
    }
    g_ptr[dev] = (double*)p;
  }
  *out = g_ptr[dev];
  return FVM_OK;
}
}  // namespace

extern "C" int fvm_calccoef_noalloc(fvm_ctx* c, fvm_step_info* info) {
  if (!c) return FVM_ERR_ARG;
  const long need = fvm_calccoef_scratch(c, false);
  if (need > FVM_NOALLOC_DOUBLES) {
    fprintf(stderr, "fvm_amd: fvm_calccoef_noalloc needs %ld scratch doubles, the static array "
                    "holds %ld (at most about 3.3 million wall cells)\n", need, FVM_NOALLOC_DOUBLES);
    return FVM_ERR_SIZE;
  }
  double* part = nullptr;
  int r = scratch_on(c->dev, &part);
  if (r) return r;
  double o[8];
  r = c->two_d ? fvm_k_calccoef2d(c, o, false, nullptr, part, FVM_NOALLOC_DOUBLES)
               : fvm_k_calccoef(c, o, false, nullptr, part, FVM_NOALLOC_DOUBLES);
  if (r) return r;
  return fvm_calccoef_info(c, o, info);
}
