// zrk_integrate_device_noalloc: zrk_integrate_device for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  The integrator's scratch -- the
// persistent grid's work counter and one saved-Jacobian / Newton-matrix slot
// per resident team -- comes from a file-scope device array instead of the
// handle's hipMalloc'd workspace; the kernels are the same.
//
// In host code the name `scratch` is only a host-side stand-in for the array:
// its device address is looked up once per device with hipGetSymbolAddress
// (with that device current) and only that address is ever passed on.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/zrk_device.h"
#include "zrk_internal.h"

namespace {

// 64 MiB per device, held until the process ends: the first 256 bytes are
// the work counter, the rest the team slots (2,047 slots of 32 KiB: the
// one-wavefront integrator's full residency at up to 63 species)
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr size_t kCounterBytes = 256;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// device address of this device's copy of scratch (looked up on first use)
int scratch_address(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices) return -1;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return e != hipSuccess ? (int)e : -1;
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return 0;
}

}  // namespace

extern "C" int zrk_integrate_device_noalloc(zrk_mech m, const zrk_bdf_options* o, int n,
                                            double dt, double* T, double* P, double* Y,
                                            int stride, int* nsteps, int* stats, void* stream) {
  if (!m || !o) return -1;
  auto* h = reinterpret_cast<zrk::MechHandle*>(m);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (h->dev_blob && dev != h->device) return -1;  // the image lives on another device
  void* base = nullptr;
  if (int rc = scratch_address(dev, &base)) return rc;
  zrk::ExtWorkspace ws;
  ws.counter = static_cast<unsigned int*>(base);
  ws.jac = reinterpret_cast<double*>(static_cast<unsigned char*>(base) + kCounterBytes);
  ws.doubles = (kScratchBytes - kCounterBytes) / sizeof(double);
  return zrk::integrate_device(h, *o, n, dt, nullptr, T, P, Y, stride, nsteps, stats, nullptr,
                               (hipStream_t)stream, nullptr, nullptr, nullptr, nullptr, nullptr,
                               1, &ws);
}
