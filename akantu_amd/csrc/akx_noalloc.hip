// akx_noalloc.hip -- akx_energy_noalloc: the kinetic energy and the external
// work of akx_energy for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops), on caller device arrays.
//
// The block partial sums of the reduction live in `scratch`, a file-scope
// device array of AKX_NOALLOC_PARTIALS + 1 doubles (32 KiB + 8 bytes) that
// every GPU the library runs on holds until the process ends.  Host code
// never passes the array's name to a kernel: on the first call on a device,
// with that device current, its address is taken once with
// hipGetSymbolAddress and kept per device.  Every call on one GPU shares that
// GPU's array; calls that overlap on one GPU (on different streams or host
// threads) race.  The kernels are the ones of akx_energy, launched with the
// same grid, so the result is bit-identical to akx_energy on the same arrays.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/akx.h"
#include "akx_internal.hpp"

namespace akx {
namespace {

constexpr int AKX_NOALLOC_PARTIALS = 4096;  // block partial sums, then the total
constexpr int AKX_NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[AKX_NOALLOC_PARTIALS + 1];

std::mutex g_mutex;
double* g_scratch[AKX_NOALLOC_MAX_DEVICES] = {};

// out = sum * scale, as akx_energy scales the sum on the host (one IEEE
// multiplication either way)
__global__ void k_scale_total(const double* sum, double scale, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *out = *sum * scale;
}

// The scratch address of the current device (taken on the first call).
hipError_t device_scratch(double** ptr) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= AKX_NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[dev] = static_cast<double*>(p);
  }
  *ptr = g_scratch[dev];
  return hipSuccess;
}

}  // namespace
}  // namespace akx

extern "C" int akx_energy_noalloc(int kind, int dim, int64_t nb_nodes, const double* velocity,
                                  const double* mass, const double* external_force,
                                  const double* internal_force, const uint8_t* blocked,
                                  double time_step, double* energy, void* hip_stream) {
  using namespace akx;
  if (dim != 2 && dim != 3) return AKX_ERR_ARGUMENT;
  if (nb_nodes < 0 || !energy || !velocity) return AKX_ERR_ARGUMENT;
  if (kind == AKX_ENERGY_KINETIC) {
    if (!mass) return AKX_ERR_ARGUMENT;
  } else if (kind == AKX_ENERGY_EXTERNAL_WORK) {
    if (!external_force || !internal_force || !blocked) return AKX_ERR_ARGUMENT;
  } else {
    return AKX_ERR_UNSUPPORTED;
  }
  double* ptr = nullptr;
  if (device_scratch(&ptr) != hipSuccess) return AKX_ERR_HIP;
  hipStream_t s = static_cast<hipStream_t>(hip_stream);
  // any node count fits: the kernels stride over the nodes with at most
  // AKX_NOALLOC_PARTIALS blocks, one partial sum each
  const int et = dim == 2 ? TRI3 : TET4;
  hipError_t e = launch_energy(kind, et, 0, s, nb_nodes, 0, nullptr, velocity, mass,
                               external_force, internal_force, blocked, nullptr, nullptr,
                               nullptr, nullptr, nullptr, nullptr, nullptr, 0, ptr,
                               AKX_NOALLOC_PARTIALS, ptr + AKX_NOALLOC_PARTIALS);
  if (e != hipSuccess) return AKX_ERR_HIP;
  const double scale = kind == AKX_ENERGY_KINETIC ? .5 : time_step;
  k_scale_total<<<1, 64, 0, s>>>(ptr + AKX_NOALLOC_PARTIALS, scale, energy);
  return hipGetLastError() == hipSuccess ? AKX_OK : AKX_ERR_HIP;
}
