// Error channel, version, and small elementwise / reduction utilities.
#include <string>

#include "common.h"

#include <mutex>
#include <vector>

namespace steam {

static thread_local std::string g_last_error;

bool g_profiling = false;
struct ProfRec {
  int fam;
  hipEvent_t a, b;
};
static std::mutex g_prof_mu;
static std::vector<ProfRec> g_prof;
static std::vector<hipEvent_t> g_pending;  // begin events awaiting their end

// Timing-only events: without the system-scope fence a record does not
// write back and invalidate the caches, which would otherwise be counted in
// the span of the launch it brackets and slow the launch after it.
static hipError_t prof_event(hipEvent_t* e) {
  return hipEventCreateWithFlags(e, hipEventDisableSystemFence);
}

void prof_begin(int fam, hipStream_t s) {
  hipEvent_t e;
  if (prof_event(&e) != hipSuccess) return;
  (void)hipEventRecord(e, s);
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_pending.push_back(e);
  (void)fam;
}

void prof_end(int fam, hipStream_t s) {
  hipEvent_t e;
  if (prof_event(&e) != hipSuccess) return;
  (void)hipEventRecord(e, s);
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (g_pending.empty()) return;
  hipEvent_t a = g_pending.back();
  g_pending.pop_back();
  g_prof.push_back({fam, a, e});
}

static void prof_clear() {
  for (auto& r : g_prof) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  g_prof.clear();
  for (auto e : g_pending) (void)hipEventDestroy(e);
  g_pending.clear();
}

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return STEAM_OK;
  int code = (e == hipErrorOutOfMemory) ? STEAM_ENOMEM : STEAM_EDEVICE;
  return set_error(code, std::string(what) + ": " + hipGetErrorString(e));
}

__global__ __launch_bounds__(256) void scale_kernel(float* __restrict__ d,
                                                    int64_t n, float s) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256)
    d[i] = __fmul_rn(d[i], s);
}

// per-block float64 partial sums of a window (x0..x1, y0..y1, all z)
__global__ __launch_bounds__(256) void window_sum_kernel(
    const float* __restrict__ d, int64_t ny, int64_t nz, int64_t x0, int64_t x1,
    int64_t y0, int64_t y1, double* __restrict__ part) {
  __shared__ double red[256];
  const int64_t wy = y1 - y0;
  const int64_t lines = (x1 - x0) * wy;
  double acc = 0.0;
  for (int64_t l = blockIdx.x; l < lines; l += gridDim.x) {
    const int64_t x = x0 + l / wy;
    const int64_t y = y0 + l % wy;
    const float* row = d + (x * ny + y) * nz;
    for (int64_t z = threadIdx.x; z < nz; z += 256) acc += (double)row[z];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void part_sum_kernel(const double* part,
                                                       int n, double* out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += part[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// VALU ceiling probe: every lane runs 8 independent fma chains, so a wave
// issues vector instructions back to back (8 fma per iteration; the loop
// counter is scalar).  The sum of the chains is stored so nothing is dead.
__global__ __launch_bounds__(256) void valu_probe_kernel(int iters, float b, float c,
                                                         float* __restrict__ out) {
  float a[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) a[i] = (float)(threadIdx.x + i);
  for (int k = 0; k < iters; ++k) {
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = fmaf(a[i], b, c);
  }
  float sum = 0.f;
#pragma unroll
  for (int i = 0; i < 8; ++i) sum += a[i];
  out[(int64_t)blockIdx.x * 256 + threadIdx.x] = sum;
}

}  // namespace steam

using namespace steam;

extern "C" int steam_valu_probe(int32_t workgroups, int32_t iters, float* out_dev,
                                int64_t* wave_fmas, steam_stream_t stream) {
  if (workgroups < 1 || iters < 1 || !out_dev || !wave_fmas)
    return set_error(STEAM_EINVAL, "steam_valu_probe: bad arguments");
  hipLaunchKernelGGL(valu_probe_kernel, dim3(workgroups), dim3(256), 0, (hipStream_t)stream,
                     iters, 0.999f, 0.001f, out_dev);
  STEAM_LAUNCH_CHECK("valu_probe_kernel");
  *wave_fmas = (int64_t)workgroups * 4 * (int64_t)iters * 8;
  return STEAM_OK;
}

extern "C" int steam_profile_enable(int on) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (on) prof_clear();
  g_profiling = on != 0;
  return STEAM_OK;
}

extern "C" int steam_profile_read(int family, int64_t* launches,
                                  double* total_ms) {
  if (!launches || !total_ms) return set_error(STEAM_EINVAL, "profile_read: NULL");
  std::lock_guard<std::mutex> lk(g_prof_mu);
  int64_t n = 0;
  double ms = 0.0;
  for (auto& r : g_prof) {
    if (r.fam != family) continue;
    STEAM_HIP(hipEventSynchronize(r.b));
    float t = 0.f;
    STEAM_HIP(hipEventElapsedTime(&t, r.a, r.b));
    ms += t;
    ++n;
  }
  *launches = n;
  *total_ms = ms;
  return STEAM_OK;
}

extern "C" int steam_version(void) { return 1 * 10000 + 0 * 100 + 0; }

extern "C" const char* steam_last_error(void) { return g_last_error.c_str(); }

extern "C" int steam_scale_inplace(float* d, int64_t n, float s,
                                   steam_stream_t stream) {
  if (n < 0) return set_error(STEAM_EINVAL, "scale_inplace: negative size");
  if (n == 0) return STEAM_OK;
  if (!d) return set_error(STEAM_EINVAL, "scale_inplace: NULL pointer");
  int64_t g = (n + 255) / 256;
  if (g > (1 << 20)) g = 1 << 20;
  hipLaunchKernelGGL(scale_kernel, dim3((int)g), dim3(256), 0, (hipStream_t)stream,
                     d, n, s);
  STEAM_LAUNCH_CHECK("scale_kernel");
  return STEAM_OK;
}

extern "C" int steam_sum_f64(const float* d, int64_t nx, int64_t ny, int64_t nz,
                             const int64_t* window, double* total,
                             steam_stream_t stream) {
  if (!d || !total) return set_error(STEAM_EINVAL, "sum_f64: NULL pointer");
  if (nx < 1 || ny < 1 || nz < 1) return set_error(STEAM_EINVAL, "sum_f64: empty");
  int64_t x0 = 0, x1 = nx, y0 = 0, y1 = ny;
  if (window) {
    x0 = window[0]; x1 = window[1]; y0 = window[2]; y1 = window[3];
    if (x0 < 0 || x1 > nx || x0 >= x1 || y0 < 0 || y1 > ny || y0 >= y1)
      return set_error(STEAM_EINVAL, "sum_f64: window out of range");
  }
  hipStream_t s = (hipStream_t)stream;
  const int64_t lines = (x1 - x0) * (y1 - y0);
  const int nblk = (int)(lines < 2048 ? lines : 2048);
  double* part = nullptr;
  STEAM_HIP(hipMallocAsync((void**)&part, (size_t)(nblk + 1) * sizeof(double), s));
  hipLaunchKernelGGL(window_sum_kernel, dim3(nblk), dim3(256), 0, s, d, ny, nz, x0,
                     x1, y0, y1, part);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    hipLaunchKernelGGL(part_sum_kernel, dim3(1), dim3(256), 0, s, part, nblk,
                       part + nblk);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(total, part + nblk, sizeof(double), hipMemcpyDeviceToHost, s);
  hipError_t ef = hipFreeAsync(part, s);
  if (e != hipSuccess) return check_hip(e, "sum_f64");
  if (ef != hipSuccess) return check_hip(ef, "hipFreeAsync(sum_f64)");
  STEAM_HIP(hipStreamSynchronize(s));
  return STEAM_OK;
}
