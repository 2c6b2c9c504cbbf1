// Host-side shared pieces: error state, device selection, libm image tables, math diagnostics.
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "fbh_internal.h"

namespace fbh {

static thread_local int t_code = FBH_OK;
static thread_local std::string t_msg;

void set_error(int code, const std::string& msg) {
  t_code = code;
  t_msg = msg;
}
int last_code() { return t_code; }

int ensure_device(int device) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
    set_error(FBH_ENODEV, "no HIP device visible");
    return FBH_ENODEV;
  }
  if (device < 0 || device >= count) {
    set_error(FBH_ENODEV, "device ordinal out of range");
    return FBH_ENODEV;
  }
  FBH_HIP_CHECK(hipSetDevice(device));
  return FBH_OK;
}

// ---------------------------------------------------------------- libm image tables
namespace {
constexpr size_t kGrid = size_t(1) << 24;
std::once_flag g_host_once;
std::vector<float> g_logf, g_cosf;
std::vector<double> g_logd;
std::mutex g_dev_mu;
std::map<int, LibmTables> g_dev;

// The argument expressions are those of the generator (src/bart_kernel.cpp:59-60): u is exact in float and the cosine
// argument is the float product 6.28318531f * u. Built with -ffp-contract=off, so nothing is fused.
void build_host_tables() {
  g_logf.resize(kGrid);
  g_cosf.resize(kGrid);
  g_logd.resize(kGrid);
  unsigned nt = std::thread::hardware_concurrency();
  if (nt < 1) nt = 1;
  if (nt > 16) nt = 16;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([t, nt]() {
      const size_t lo = kGrid * t / nt, hi = kGrid * (t + 1) / nt;
      for (size_t k = lo; k < hi; k++) {
        volatile float u = ((float)(uint32_t)k + 0.5f) * (1.0f / 16777216.0f);
        const float uu = u;
        volatile float x = 6.28318531f * uu;
        g_logf[k] = logf(uu);
        g_cosf[k] = cosf(x);
        g_logd[k] = log((double)uu);
      }
    });
  for (auto& x : th) x.join();
}
}  // namespace

int libm_tables(int device, LibmTables* out) {
  std::call_once(g_host_once, build_host_tables);
  std::lock_guard<std::mutex> lk(g_dev_mu);
  auto it = g_dev.find(device);
  if (it != g_dev.end()) {
    *out = it->second;
    return FBH_OK;
  }
  float *lf = nullptr, *cf = nullptr;
  double* ld = nullptr;
  FBH_HIP_CHECK(hipMalloc(&lf, kGrid * sizeof(float)));
  FBH_HIP_CHECK(hipMalloc(&cf, kGrid * sizeof(float)));
  FBH_HIP_CHECK(hipMalloc(&ld, kGrid * sizeof(double)));
  FBH_HIP_CHECK(hipMemcpy(lf, g_logf.data(), kGrid * sizeof(float), hipMemcpyHostToDevice));
  FBH_HIP_CHECK(hipMemcpy(cf, g_cosf.data(), kGrid * sizeof(float), hipMemcpyHostToDevice));
  FBH_HIP_CHECK(hipMemcpy(ld, g_logd.data(), kGrid * sizeof(double), hipMemcpyHostToDevice));
  LibmTables t{lf, cf, ld};
  g_dev[device] = t;
  *out = t;
  return FBH_OK;
}

// ---------------------------------------------------------------- math diagnostics
__global__ void k_debug_math(int kind, const uint32_t* in_u, const double* in_d, long count, float* out_f,
                             double* out_d, LibmTables tb) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  switch (kind) {
    case 0: {
      const uint32_t a = in_u[2 * i] >> 8, b = in_u[2 * i + 1] >> 8;
      out_f[i] = __builtin_sqrtf(-2.f * tb.logf_t[a]) * tb.cosf_t[b];
      break;
    }
    case 1: out_f[i] = __builtin_sqrtf(-2.f * tb.logf_t[in_u[i] >> 8]); break;
    case 2: out_d[i] = log(in_d[i]); break;
    case 3: out_d[i] = exp(in_d[i]); break;
    case 4: out_d[i] = __builtin_sqrt(in_d[i]); break;
    case 5: out_d[i] = in_d[2 * i] / in_d[2 * i + 1]; break;
    default: break;
  }
}

// plain copy of 16-byte words, grid-stride (counter calibration: reads and writes exactly `words` * 16 bytes)
__global__ __launch_bounds__(256) void k_stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                     long words) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) dst[i] = src[i];
}

}  // namespace fbh

using namespace fbh;

extern "C" const char* fbh_last_error(void) { return t_msg.c_str(); }
extern "C" int fbh_last_status(void) { return t_code; }
extern "C" const char* fbh_version(void) { return "fastbart_hip 0.1 gfx950"; }

// host libm image entries [k0, k0+count) of table kind 0 (logf), 1 (cosf of 2 pi u) or 2 (log double), for tests
extern "C" int fbh_debug_libm_host(int kind, long k0, long count, void* out) {
  if (kind < 0 || kind > 2 || k0 < 0 || count < 0 || k0 + count > (long)kGrid || (count && !out)) {
    set_error(FBH_EINVAL, "fbh_debug_libm_host: bad argument");
    return FBH_EINVAL;
  }
  std::call_once(g_host_once, build_host_tables);
  if (kind == 0) std::memcpy(out, g_logf.data() + k0, sizeof(float) * count);
  if (kind == 1) std::memcpy(out, g_cosf.data() + k0, sizeof(float) * count);
  if (kind == 2) std::memcpy(out, g_logd.data() + k0, sizeof(double) * count);
  return FBH_OK;
}

extern "C" int fbh_debug_math(int kind, const uint32_t* in_u, const double* in_d, long count, float* out_f,
                              double* out_d, int device) {
  if (count < 0 || kind < 0 || kind > 5) {
    set_error(FBH_EINVAL, "fbh_debug_math: bad kind/count");
    return FBH_EINVAL;
  }
  if (count == 0) return FBH_OK;
  int rc = ensure_device(device);
  if (rc) return rc;
  LibmTables tb;
  if ((rc = libm_tables(device, &tb))) return rc;
  const long nu = (kind == 0) ? 2 * count : (kind == 1 ? count : 0);
  const long nd = (kind == 5) ? 2 * count : (kind >= 2 ? count : 0);
  uint32_t* du = nullptr;
  double* dd = nullptr;
  float* of = nullptr;
  double* od = nullptr;
  if (nu) {
    FBH_HIP_CHECK(hipMalloc(&du, nu * sizeof(uint32_t)));
    FBH_HIP_CHECK(hipMemcpy(du, in_u, nu * sizeof(uint32_t), hipMemcpyHostToDevice));
  }
  if (nd) {
    FBH_HIP_CHECK(hipMalloc(&dd, nd * sizeof(double)));
    FBH_HIP_CHECK(hipMemcpy(dd, in_d, nd * sizeof(double), hipMemcpyHostToDevice));
  }
  if (kind <= 1) FBH_HIP_CHECK(hipMalloc(&of, count * sizeof(float)));
  else FBH_HIP_CHECK(hipMalloc(&od, count * sizeof(double)));
  const int bs = 256;
  const long grid = (count + bs - 1) / bs;
  hipLaunchKernelGGL(k_debug_math, dim3((unsigned)grid), dim3(bs), 0, 0, kind, du, dd, count, of, od, tb);
  FBH_HIP_CHECK(hipGetLastError());
  FBH_HIP_CHECK(hipDeviceSynchronize());
  if (of) FBH_HIP_CHECK(hipMemcpy(out_f, of, count * sizeof(float), hipMemcpyDeviceToHost));
  if (od) FBH_HIP_CHECK(hipMemcpy(out_d, od, count * sizeof(double), hipMemcpyDeviceToHost));
  (void)hipFree(du);
  (void)hipFree(dd);
  (void)hipFree(of);
  (void)hipFree(od);
  return FBH_OK;
}

extern "C" int fbh_debug_copy(long mib, int device, double* ms) {
  if (mib < 1 || mib > 16384) {
    set_error(FBH_EINVAL, "fbh_debug_copy: size outside 1..16384 MiB");
    return FBH_EINVAL;
  }
  int rc = ensure_device(device);
  if (rc) return rc;
  const long words = (mib << 20) / 16;
  uint4 *src = nullptr, *dst = nullptr;
  FBH_HIP_CHECK(hipMalloc(&src, words * 16));
  if (hipMalloc(&dst, words * 16) != hipSuccess) {
    (void)hipFree(src);
    set_error(FBH_ENOMEM, "fbh_debug_copy: hipMalloc failed");
    return FBH_ENOMEM;
  }
  hipEvent_t e0, e1;
  FBH_HIP_CHECK(hipEventCreate(&e0));
  FBH_HIP_CHECK(hipEventCreate(&e1));
  FBH_HIP_CHECK(hipEventRecord(e0, 0));
  hipLaunchKernelGGL(k_stream_copy, dim3(8192), dim3(256), 0, 0, src, dst, words);
  FBH_HIP_CHECK(hipGetLastError());
  FBH_HIP_CHECK(hipEventRecord(e1, 0));
  FBH_HIP_CHECK(hipEventSynchronize(e1));
  float t = 0;
  FBH_HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
  if (ms) *ms = t;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(src);
  (void)hipFree(dst);
  return FBH_OK;
}
