// OC bisection (TOP3D_XL.m:145-155) on caller-owned device arrays without
// any device allocation, for callers that cannot allocate (code under HIP
// graph capture, serving loops, harnesses that forbid hipMalloc).
//
// The kernels are the ones behind t3x_oc (k-ary bisection, fixed-order
// reductions); their scratch (reduction partials, the bisection state and the
// volume sum) lives in `scratch`, a file-scope device array of fixed size.
// Each GPU holds its own copy at its own address: the address is looked up
// with hipGetSymbolAddress once per device, with that device current, and
// only that address reaches kernels.  Every call on one GPU shares that GPU's
// array, so calls that overlap on one GPU (two host threads, two streams)
// race; each call sets the state it reads before reading it.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>

#include "../../include/t3x.h"
#include "t3x_kernels.h"
#include "t3x_reduce.h"

using namespace t3x;

namespace {

// Layout (doubles): partials of the widest reduction (15 bisection
// candidates, at most kMaxRedBlocks blocks, rounded up to 16 values per
// block), then the 8 bisection slots, then the volume sum.
constexpr long long kPartialDoubles = 16LL * kMaxRedBlocks;
constexpr long long kOcSlot = kPartialDoubles;
constexpr long long kSumSlot = kOcSlot + 8;
constexpr long long kScratchDoubles = kSumSlot + 8;
static_assert(kScratchDoubles * 8 == T3X_OC_NOALLOC_SCRATCH_BYTES, "scratch size in t3x.h");

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex gMu;
double* gScratch[kMaxDevices] = {};

void set_err(char* err, int errLen, const char* msg) {
  if (err && errLen > 0) {
    std::strncpy(err, msg, (size_t)errLen - 1);
    err[errLen - 1] = '\0';
  }
}

// The current device's copy of `scratch` (looked up once per device).
hipError_t device_scratch(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(gMu);
  if (!gScratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    gScratch[dev] = (double*)p;
  }
  *out = gScratch[dev];
  return hipSuccess;
}

}  // namespace

extern "C" int t3x_oc_noalloc(long long nCells, double* x, const double* xPhys, const double* dc,
                              const uint8_t* passive, double V0, void* stream, double* change, int* nBisect,
                              char* err, int errLen) {
  set_err(err, errLen, "");
  if (nCells < 1 || nCells > T3X_OC_NOALLOC_MAX_CELLS) {
    char m[160];
    std::snprintf(m, sizeof m, "t3x_oc_noalloc: nCells = %lld outside 1..%lld", nCells,
                  (long long)T3X_OC_NOALLOC_MAX_CELLS);
    set_err(err, errLen, m);
    return T3X_ERR_INPUT;
  }
  if (!x || !xPhys || !dc || !change) {
    set_err(err, errLen, "t3x_oc_noalloc: x, xPhys, dc and change must not be null");
    return T3X_ERR_INPUT;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_err(err, errLen, "t3x_oc_noalloc: no GPU visible");
    return T3X_ERR_NO_DEVICE;
  }
  double* sp = nullptr;
  hipError_t e = device_scratch(&sp);
  auto fail = [&](hipError_t code, const char* what) {
    char m[200];
    std::snprintf(m, sizeof m, "t3x_oc_noalloc: %s: %s", what, hipGetErrorString(code));
    set_err(err, errLen, m);
    return (int)T3X_ERR_HIP;
  };
  if (e != hipSuccess) return fail(e, "hipGetSymbolAddress(scratch)");
  hipStream_t s = (hipStream_t)stream;
  const RedWs ws{sp, nullptr};
  double* oc = sp + kOcSlot;
  double* vsum = sp + kSumSlot;
  double hb[8];
  // fval = mean(xPhys) - V0 (145); l1 = 0, l2 = 1e9 (146)
  launch_masked_sum(s, nCells, nullptr, xPhys, nullptr, 0, ws, vsum);
  if ((e = hipMemcpyAsync(hb, vsum, sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (e = hipStreamSynchronize(s)) != hipSuccess)
    return fail(e, "volume sum");
  const double fval = hb[0] / (double)nCells - V0;
  const double init[8] = {0.0, 1e9, fval, 0.0, 0.0, 0.0, 0.0, 0.0};
  launch_set_vals(s, oc, 8, init);
  for (int pass = 0; pass < 64; ++pass) {
    launch_oc_pass(s, nCells, nullptr, passive, x, dc, ws, oc, 3);
    if (pass % 8 == 7) {
      if ((e = hipMemcpyAsync(hb, oc, 8 * sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess ||
          (e = hipStreamSynchronize(s)) != hipSuccess)
        return fail(e, "bisection");
      if (hb[3] != 0.0) break;
    }
  }
  launch_oc_final(s, nCells, nullptr, passive, x, dc, ws, oc);
  if ((e = hipGetLastError()) != hipSuccess) return fail(e, "kernel launch");
  if ((e = hipMemcpyAsync(hb, oc, 8 * sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (e = hipStreamSynchronize(s)) != hipSuccess)
    return fail(e, "update");
  *change = hb[6];
  if (nBisect) *nBisect = (int)hb[5];
  return T3X_OK;
}
