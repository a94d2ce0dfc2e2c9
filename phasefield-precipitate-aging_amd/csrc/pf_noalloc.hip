// pf_noalloc.hip -- pf_summarize_f64_noalloc: the field summary of
// pf_summarize_f64 for callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
// It needs no context: the reduction partials live in a file-scope device
// array, `scratch`, of fixed size (REDUCE_BLOCKS * 10 + 8 doubles, 80 KiB),
// which every GPU the library runs on holds until the process ends.  The call
// allocates nothing and calls no allocating routine.
//
// The array's address is looked up once per device with hipGetSymbolAddress
// (the array's name in host code is not a device address) and kept per
// device.  Every call on a GPU shares that GPU's array: the partial kernel
// overwrites every slot the final kernel reads, so nothing needs clearing
// between calls, but calls that overlap on one GPU (two streams, two threads)
// race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "pf_internal.h"
#include "pfpa.h"

namespace pf {

constexpr size_t NOALLOC_SCRATCH_DOUBLES = (size_t)REDUCE_BLOCKS * 10 + 8;

static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mutex;
double* g_scratch[MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up on the
// first call there.
hipError_t device_scratch(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[dev] = (double*)p;
  }
  *out = g_scratch[dev];
  return hipSuccess;
}

int nfail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

}  // namespace
}  // namespace pf

extern "C" int pf_summarize_f64_noalloc(int nx, int ny, const double* conc_Cr,
                                        const double* conc_Nb, const double* phi_del,
                                        const double* phi_lav, double* out5, void* stream) {
  if (!out5) return pf::nfail(PF_ERR_ARG, "null output");
  if (!conc_Cr || !conc_Nb || !phi_del || !phi_lav)
    return pf::nfail(PF_ERR_ARG, "null field pointer");
  if (nx < 3 || ny < 3)
    return pf::nfail(PF_ERR_ARG, "grid must have at least one interior cell (nx, ny >= 3)");
  // the scratch holds the partials of at most REDUCE_BLOCKS blocks, which the
  // reduction never exceeds (it strides over the grid); the grid itself is
  // limited by the 32-bit cell indices of the field layout
  if ((long long)nx * (long long)ny >= (1LL << 31))
    return pf::nfail(PF_ERR_ARG, "grid too large for 32-bit cell indices (nx * ny < 2^31)");
  double* ptr = nullptr;
  hipError_t e = pf::device_scratch(&ptr);
  if (e != hipSuccess)
    return pf::nfail(PF_ERR_HIP, std::string("summary scratch lookup: ") + hipGetErrorString(e));
  pf::Scratch S{};
  S.partials = ptr;
  e = pf::launch_summarize<double>(nx, ny, conc_Cr, conc_Nb, phi_del, phi_lav, nullptr, out5, S,
                                   (hipStream_t)stream);
  if (e != hipSuccess)
    return pf::nfail(PF_ERR_HIP, std::string("summarize kernels: ") + hipGetErrorString(e));
  return PF_OK;
}
