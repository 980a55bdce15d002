// Allocation-free form of cvf_vdw_energy (include/cvf_api.h).
//
// cvf_vdw_energy stages the caller's shell table and radii in device memory
// it allocates per call.  cvf_vdw_energy_noalloc stages them instead in
// `scratch`, a file-scope device array of fixed size that every GPU the
// library runs on holds until the process ends, so the call makes no
// hipMalloc / hipFree.  In host code `scratch` names a host-side stand-in,
// not a device address: the device address is taken once per device with
// hipGetSymbolAddress (first call on that device, with it current) and only
// that address reaches copies and kernel arguments.
//
// Every call on one GPU shares that GPU's array and leaves its contents
// behind; calls that overlap on one GPU (two host threads) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "cvf_api.h"
#include "cvf_kernels.h"

namespace cvf {
int set_error(int code, const char* fmt, ...);
}

namespace {

// scratch layout: shell distances, shell counts, radii, energies
constexpr size_t kDistOff = 0;
constexpr size_t kFreqOff = kDistOff + sizeof(float) * CVF_NOALLOC_MAX_SHELLS;
constexpr size_t kROff = kFreqOff + sizeof(int32_t) * CVF_NOALLOC_MAX_SHELLS;
constexpr size_t kEOff = kROff + sizeof(double) * CVF_NOALLOC_MAX_R;
constexpr size_t kScratchBytes = kEOff + sizeof(double) * CVF_NOALLOC_MAX_R;
static_assert(kROff % 8 == 0 && kEOff % 8 == 0, "FP64 staging must be 8-byte aligned");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_ptr[kMaxDevices] = {};  // device address of `scratch`, per device

// Device address of the current device's `scratch`, looked up on the first
// call on that device.
int scratch_address(int dev, unsigned char** out) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return cvf::set_error(CVF_EDEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                            hipGetErrorString(e));
    g_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_ptr[dev];
  return CVF_OK;
}

#define HIPC(expr)                                                                                    \
  do {                                                                                                \
    const hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess)                                                                             \
      return cvf::set_error(CVF_EDEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                            __LINE__);                                                                \
  } while (0)

}  // namespace

extern "C" int cvf_vdw_energy_noalloc(int32_t device_id, int32_t L, const float* dist, const int32_t* freq,
                                      int32_t nshell, float r_cutoff, const double* r, int32_t n_r,
                                      int32_t precision, double* energy) {
  if (L < 2 || nshell <= 0 || n_r <= 0 || !dist || !freq || !r || !energy)
    return cvf::set_error(CVF_EINVAL, "bad arguments");
  if (nshell > CVF_NOALLOC_MAX_SHELLS || n_r > CVF_NOALLOC_MAX_R)
    return cvf::set_error(CVF_EINVAL,
                          "cvf_vdw_energy_noalloc: %d shells and %d radii exceed its fixed scratch "
                          "(at most %d shells and %d radii per call)",
                          nshell, n_r, CVF_NOALLOC_MAX_SHELLS, CVF_NOALLOC_MAX_R);
  if (precision != CVF_FP64 && precision != CVF_FP32_COMPENSATED)
    return cvf::set_error(CVF_EINVAL, "unknown precision");
  int ndev = 0;
  HIPC(hipGetDeviceCount(&ndev));
  if (device_id < 0 || device_id >= ndev || device_id >= kMaxDevices)
    return cvf::set_error(CVF_EINVAL, "device %d out of range (%d devices)", device_id, ndev);
  HIPC(hipSetDevice(device_id));
  unsigned char* base = nullptr;
  const int rc = scratch_address(device_id, &base);
  if (rc) return rc;
  float* d = reinterpret_cast<float*>(base + kDistOff);
  int32_t* f = reinterpret_cast<int32_t*>(base + kFreqOff);
  double* dr = reinterpret_cast<double*>(base + kROff);
  double* de = reinterpret_cast<double*>(base + kEOff);
  // the null stream orders the copies and the kernel
  HIPC(hipMemcpy(d, dist, (size_t)nshell * sizeof(float), hipMemcpyHostToDevice));
  HIPC(hipMemcpy(f, freq, (size_t)nshell * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPC(hipMemcpy(dr, r, (size_t)n_r * sizeof(double), hipMemcpyHostToDevice));
  cvf::launch_vdw_eval(nullptr, d, f, nshell, r_cutoff, L * L * L, dr, n_r, precision, de);
  HIPC(hipGetLastError());
  HIPC(hipMemcpy(energy, de, (size_t)n_r * sizeof(double), hipMemcpyDeviceToHost));
  return CVF_OK;
}
