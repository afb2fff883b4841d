// Turbulon kernel synthesis (steam/simulate.py:2503-2556, shapes from
// steam/turbulons.py:20-48): float32 taps in the reference's operation order,
// the two sums of the Gaussian-weighted zero-sum correction in float64 with a
// fixed-order reduction.
#include <math.h>

#include <mutex>
#include <vector>

#include "common.h"

namespace steam {

struct EnvParams {
  int hx, hy, hz;
  float dx, dy, dz;
  float sig2;   // float32((k/pi)^2)
  float g2;     // float32(2 (k/pi)^2)
  float omega;  // float32(6/k)
  float pexp;   // float32(-p/2): power law (1 + rho)^(-p/2)
  int shape;
};

__device__ __forceinline__ void env_tap(const EnvParams& p, int64_t idx,
                                        float& env, float& gauss) {
  const int ny = 2 * p.hy + 1, nz = 2 * p.hz + 1;
  const int c = (int)(idx % nz);
  const int64_t ab = idx / nz;
  const int b = (int)(ab % ny);
  const int a = (int)(ab / ny);
  const float X = __fmul_rn((float)(a - p.hx), p.dx);
  const float Y = __fmul_rn((float)(b - p.hy), p.dy);
  const float Z = __fmul_rn((float)(c - p.hz), p.dz);
  const float r2 = __fadd_rn(__fadd_rn(__fmul_rn(X, X), __fmul_rn(Y, Y)),
                             __fmul_rn(Z, Z));
  const float ratio = __fdiv_rn(r2, p.sig2);
  const float half = expf(__fdiv_rn(-ratio, 2.0f));
  if (p.shape == STEAM_SHAPE_MEXICAN_HAT)
    env = __fmul_rn(__fsub_rn(3.0f, ratio), half);
  else if (p.shape == STEAM_SHAPE_MORLET_OMEGA0_6)
    env = __fmul_rn(cosf(__fmul_rn(p.omega, sqrtf(r2))), half);
  else
    env = powf(__fadd_rn(1.0f, ratio), p.pexp);
  gauss = expf(__fdiv_rn(-r2, p.g2));
}

__global__ __launch_bounds__(256) void env_taps_kernel(EnvParams p,
                                                       int64_t n,
                                                       float* __restrict__ out,
                                                       double* __restrict__ part) {
  __shared__ double red[2][256];
  double se = 0.0, sg = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    float e, g;
    env_tap(p, i, e, g);
    out[i] = e;
    se += e;
    sg += g;
  }
  red[0][threadIdx.x] = se;
  red[1][threadIdx.x] = sg;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = red[0][0];
    part[2 * blockIdx.x + 1] = red[1][0];
  }
}

__global__ __launch_bounds__(256) void env_reduce_kernel(const double* part,
                                                         int nparts,
                                                         float* ratio_out) {
  __shared__ double red[2][256];
  double se = 0.0, sg = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) {
    se += part[2 * i];
    sg += part[2 * i + 1];
  }
  red[0][threadIdx.x] = se;
  red[1][threadIdx.x] = sg;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *ratio_out = (float)(red[0][0] / red[1][0]);
}

__global__ __launch_bounds__(256) void env_correct_kernel(EnvParams p, int64_t n,
                                                          const float* ratio,
                                                          float* __restrict__ out) {
  const float r = *ratio;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    float e, g;
    env_tap(p, i, e, g);
    out[i] = __fsub_rn(out[i], __fmul_rn(g, r));
  }
}

// Workgroups of the tap and correction passes (each takes a grid stride
// over the taps); their float64 partial sums are the scratch a call needs.
constexpr int kEnvMaxWgs = 4096;
constexpr size_t kEnvScratchBytes = (size_t)kEnvMaxWgs * 2 * sizeof(double) + 256;

// Scratch of steam_turbulon_envelope_noalloc: the partial sums of up to
// kEnvMaxWgs workgroups, then the float32 correction ratio.  One copy per
// GPU, held until the process ends.  Host code never uses this name as an
// address: it reads each device's address once (envelope_scratch).
static __device__ __align__(256) unsigned char scratch[kEnvScratchBytes];

// Device address of scratch on the current device, looked up on the first
// call on that device and kept.
static std::mutex g_scratch_mu;
static std::vector<void*> g_scratch_ptr;

static int envelope_scratch(void** out) {
  int dev = 0;
  STEAM_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return check_hip(e, "hipGetSymbolAddress(envelope scratch)");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return STEAM_OK;
}

// Parameters and tap count of a kernel, with the argument checks of the
// public entries.
static int envelope_setup(double k, double dx, double dy, double dz, double sf,
                          int shape_id, const float* out, EnvParams& p, int64_t& n);

// The three passes: taps and their two sums per workgroup into part, the
// fixed-order reduction into the ratio (after the 2 * nblk partials), the
// correction.  part holds env_scratch_bytes(nblk).
static size_t env_scratch_bytes(int nblk) { return (size_t)nblk * 2 * sizeof(double) + 256; }

static hipError_t envelope_passes(const EnvParams& p, int64_t n, int nblk, float* out,
                                  double* part, hipStream_t s) {
  float* ratio = (float*)(part + 2 * nblk);
  hipLaunchKernelGGL(env_taps_kernel, dim3(nblk), dim3(256), 0, s, p, n, out, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(env_reduce_kernel, dim3(1), dim3(256), 0, s, part, nblk, ratio);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(env_correct_kernel, dim3(nblk), dim3(256), 0, s, p, n, ratio, out);
  return hipGetLastError();
}

}  // namespace steam

using namespace steam;

extern "C" int steam_envelope_halfwidths(double k, double dx, double dy,
                                         double dz, double sf, int64_t half[3]) {
  if (!half) return set_error(STEAM_EINVAL, "envelope_halfwidths: NULL output");
  if (!(k > 0) || !(dx > 0) || !(dy > 0) || !(dz > 0) || !(sf > 0))
    return set_error(STEAM_EINVAL, "envelope: k, spacings and support must be positive");
  half[0] = (int64_t)ceil(sf * k / dx);
  half[1] = (int64_t)ceil(sf * k / dy);
  half[2] = (int64_t)ceil(sf * k / dz);
  return STEAM_OK;
}

int steam::envelope_setup(double k, double dx, double dy, double dz, double sf,
                          int shape_id, const float* out, EnvParams& p, int64_t& n) {
  if (shape_id != STEAM_SHAPE_MEXICAN_HAT && shape_id != STEAM_SHAPE_MORLET_OMEGA0_6 &&
      shape_id != STEAM_SHAPE_POWER_LAW)
    return set_error(STEAM_EINVAL, "envelope: unknown shape");
  int64_t h[3];
  int rc = steam_envelope_halfwidths(k, dx, dy, dz, sf, h);
  if (rc) return rc;
  if (!out) return set_error(STEAM_EINVAL, "envelope: NULL output");
  if (h[0] > 4096 || h[1] > 4096 || h[2] > 4096)
    return set_error(STEAM_EINVAL, "envelope: support too large");
  p.hx = (int)h[0];
  p.hy = (int)h[1];
  p.hz = (int)h[2];
  p.dx = (float)dx;
  p.dy = (float)dy;
  p.dz = (float)dz;
  const double sig = k / M_PI;
  p.sig2 = (float)(sig * sig);
  p.g2 = (float)(2.0 * sig * sig);
  p.omega = (float)(6.0 / k);
  p.pexp = 0.f;
  p.shape = shape_id;
  n = (2 * h[0] + 1) * (2 * h[1] + 1) * (2 * h[2] + 1);
  return STEAM_OK;
}

static int envelope_allocating(const EnvParams& p, int64_t n, float* out,
                               steam_stream_t stream);

extern "C" int steam_turbulon_envelope(double k, double dx, double dy, double dz,
                                       double sf, int shape_id, float* out,
                                       steam_stream_t stream) {
  if (shape_id == STEAM_SHAPE_POWER_LAW)
    return set_error(STEAM_EINVAL, "envelope: the power-law shape takes its exponent "
                                   "(steam_turbulon_envelope_power)");
  EnvParams p;
  int64_t n = 0;
  int rc = envelope_setup(k, dx, dy, dz, sf, shape_id, out, p, n);
  if (rc) return rc;
  return envelope_allocating(p, n, out, stream);
}

extern "C" int steam_turbulon_envelope_power(double k, double dx, double dy, double dz,
                                             double sf, double exponent, float* out,
                                             steam_stream_t stream) {
  if (!(exponent > 0.0) || !(exponent < 16.0))
    return set_error(STEAM_EINVAL, "envelope: power-law exponent must lie in (0, 16)");
  EnvParams p;
  int64_t n = 0;
  int rc = envelope_setup(k, dx, dy, dz, sf, STEAM_SHAPE_POWER_LAW, out, p, n);
  if (rc) return rc;
  p.pexp = (float)(-0.5 * exponent);
  return envelope_allocating(p, n, out, stream);
}

static int envelope_allocating(const EnvParams& p, int64_t n, float* out,
                               steam_stream_t stream) {
  const int nblk = (int)std::min<int64_t>((n + 255) / 256, kEnvMaxWgs);
  hipStream_t s = (hipStream_t)stream;
  double* part = nullptr;
  STEAM_HIP(hipMallocAsync((void**)&part, env_scratch_bytes(nblk), s));
  hipError_t e = envelope_passes(p, n, nblk, out, part, s);
  hipError_t ef = hipFreeAsync(part, s);
  if (e != hipSuccess) return check_hip(e, "envelope kernels");
  if (ef != hipSuccess) return check_hip(ef, "hipFreeAsync(envelope)");
  STEAM_HIP(hipStreamSynchronize(s));
  return STEAM_OK;
}

extern "C" int steam_turbulon_envelope_noalloc(double k, double dx, double dy, double dz,
                                               double sf, int shape_id, float* out,
                                               steam_stream_t stream) {
  if (shape_id == STEAM_SHAPE_POWER_LAW)
    return set_error(STEAM_EINVAL, "envelope_noalloc: the power-law shape takes its "
                                   "exponent (steam_turbulon_envelope_power)");
  EnvParams p;
  int64_t n = 0;
  int rc = envelope_setup(k, dx, dy, dz, sf, shape_id, out, p, n);
  if (rc) return rc;
  // the same grid as steam_turbulon_envelope, so the same summation order
  const int nblk = (int)std::min<int64_t>((n + 255) / 256, kEnvMaxWgs);
  if (env_scratch_bytes(nblk) > kEnvScratchBytes)
    return set_error(STEAM_EINVAL, "envelope_noalloc: the kernel needs more scratch than "
                                   "the library's static array holds");
  void* ptr = nullptr;
  rc = envelope_scratch(&ptr);
  if (rc) return rc;
  // every partial slot and the ratio are written before they are read, in
  // stream order: nothing in the scratch needs clearing between calls
  hipError_t e = envelope_passes(p, n, nblk, out, (double*)ptr, (hipStream_t)stream);
  if (e != hipSuccess) return check_hip(e, "envelope kernels (noalloc)");
  return STEAM_OK;
}
