// World-keyed noise kernels: Philox4x32-10 uniforms, the CMS extremal-Levy
// transform, the sparse keyed Levy field and the unit-mean multiplier noise.
#include "conv.h"

namespace steam {

NoiseDev make_noise_dev(const steam_noise_params* np);
int check_noise_params(const steam_noise_params* np);

// keyed_uniforms (steam/noise.py:103-130): one thread per site, C order.
__global__ __launch_bounds__(256) void keyed_uniforms_kernel(
    const int64_t* __restrict__ ix, int64_t nx, const int64_t* __restrict__ iy,
    int64_t ny, const int64_t* __restrict__ iz, int64_t nz, uint32_t k0,
    uint32_t k1, uint32_t k2, float* __restrict__ lane0,
    float* __restrict__ lane1) {
  const int64_t n = nx * ny * nz;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = i % nz;
    const int64_t ab = i / nz;
    const int64_t b = ab % ny;
    const int64_t a = ab / ny;
    float u0, u1;
    uniform_pair(ix[a], iy[b], iz[c], k0, k1, k2, u0, u1);
    lane0[i] = u0;
    lane1[i] = u1;
  }
}

__global__ __launch_bounds__(256) void philox_kernel(const uint32_t* __restrict__ ctr,
                                                     int64_t n, uint32_t k0,
                                                     uint32_t k1,
                                                     uint32_t* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t c0 = ctr[4 * i], c1 = ctr[4 * i + 1], c2 = ctr[4 * i + 2],
             c3 = ctr[4 * i + 3];
    philox_fast(c0, c1, c2, c3, k0, k1);
    out[4 * i] = c0;
    out[4 * i + 1] = c1;
    out[4 * i + 2] = c2;
    out[4 * i + 3] = c3;
  }
}

__global__ __launch_bounds__(256) void levy_kernel(CmsConst k,
                                                   float* __restrict__ phi,
                                                   const float* __restrict__ R,
                                                   int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    phi[i] = cms_f32(phi[i], R[i], k);
}

// Sparse keyed Levy (steam/simulate.py:554-598) or, with MNOISE, the unit-mean
// multiplier noise (steam/simulate.py:2146-2161).  One workgroup row per
// (x, y) line, threads along z (coalesced stores).
template <bool MNOISE>
__global__ __launch_bounds__(256) void field_noise_kernel(NoiseDev nd,
                                                          int64_t nx, int64_t ny,
                                                          int64_t nz,
                                                          float* __restrict__ out) {
  const int64_t lines = nx * ny;
  for (int64_t l = blockIdx.x; l < lines; l += gridDim.x) {
    const int64_t x = l / ny;
    const int64_t y = l - x * ny;
    LineNoise ln = noise_line(nd, (int)x, (int)y);
    float* o = out + l * nz;
    for (int64_t z = threadIdx.x; z < nz; z += blockDim.x) {
      float v;
      if (MNOISE) {
        v = noise_at(nd, ln, (int)z, (int)nz);
      } else {
        v = 0.f;
        if (ln.xy_center && (nd.sz == 1 || ((int)z + ln.zphase) % nd.sz == 0)) {
          float u0, u1;
          uniform_pair(ln.wx, ln.wy, (int64_t)((int)z + (int)nd.oz), nd.key0,
                       nd.key1, nd.key2, u0, u1);
          v = cms_f32(u0, u1, nd.cms);
        }
      }
      o[z] = v;
    }
  }
}

static int grid_for(int64_t n, int per) {
  int64_t g = (n + per - 1) / per;
  if (g > 1 << 20) g = 1 << 20;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace steam

using namespace steam;

extern "C" int steam_keyed_uniforms(const int64_t* ix, int64_t nx,
                                    const int64_t* iy, int64_t ny,
                                    const int64_t* iz, int64_t nz, uint32_t k0,
                                    uint32_t k1, uint32_t k2, float* lane0,
                                    float* lane1, steam_stream_t stream) {
  if (nx < 0 || ny < 0 || nz < 0)
    return set_error(STEAM_EINVAL, "keyed_uniforms: negative extent");
  const int64_t n = nx * ny * nz;
  if (n == 0) return STEAM_OK;
  if (!ix || !iy || !iz || !lane0 || !lane1)
    return set_error(STEAM_EINVAL, "keyed_uniforms: NULL pointer");
  hipLaunchKernelGGL(keyed_uniforms_kernel, dim3(grid_for(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, ix, nx, iy, ny, iz, nz, k0, k1, k2,
                     lane0, lane1);
  STEAM_LAUNCH_CHECK("keyed_uniforms_kernel");
  return STEAM_OK;
}

extern "C" int steam_philox4x32_10(const uint32_t* ctr, int64_t n, uint32_t k0,
                                   uint32_t k1, uint32_t* out,
                                   steam_stream_t stream) {
  if (n < 0) return set_error(STEAM_EINVAL, "philox: negative count");
  if (n == 0) return STEAM_OK;
  if (!ctr || !out) return set_error(STEAM_EINVAL, "philox: NULL pointer");
  hipLaunchKernelGGL(philox_kernel, dim3(grid_for(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, ctr, n, k0, k1, out);
  STEAM_LAUNCH_CHECK("philox_kernel");
  return STEAM_OK;
}

extern "C" int steam_levy_from_uniforms(double alpha, float* phi, float* R,
                                        int64_t n, steam_stream_t stream) {
  if (n < 0) return set_error(STEAM_EINVAL, "levy_from_uniforms: negative size");
  if (n == 0) return STEAM_OK;
  if (!phi || !R) return set_error(STEAM_EINVAL, "levy_from_uniforms: NULL pointer");
  if (!(alpha > 0.0 && alpha <= 2.0))
    return set_error(STEAM_EINVAL, "levy_from_uniforms: alpha must lie in (0, 2]");
  hipLaunchKernelGGL(levy_kernel, dim3(grid_for(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, make_cms(alpha), phi, R, n);
  STEAM_LAUNCH_CHECK("levy_kernel");
  return STEAM_OK;
}

extern "C" int steam_keyed_sparse_levy(int64_t nx, int64_t ny, int64_t nz,
                                       const steam_region* region, double alpha,
                                       float* out, steam_stream_t stream) {
  if (!region) return set_error(STEAM_EINVAL, "keyed_sparse_levy: NULL region");
  steam_noise_params np;
  np.region = *region;
  np.alpha = alpha;
  np.per_class_scale = 1.0f;
  np.shift = 0.0f;
  np.n_zero = 0;
  np.zero_bottom = 0;
  np.zero_top = 0;
  int rc = check_noise_params(&np);
  if (rc) return rc;
  if (nx < 0 || ny < 0 || nz < 0)
    return set_error(STEAM_EINVAL, "keyed_sparse_levy: negative extent");
  if (nx * ny * nz == 0) return STEAM_OK;
  if (!out) return set_error(STEAM_EINVAL, "keyed_sparse_levy: NULL output");
  hipLaunchKernelGGL((field_noise_kernel<false>), dim3(grid_for(nx * ny, 1)),
                     dim3(nz >= 256 ? 256 : 64), 0, (hipStream_t)stream,
                     make_noise_dev(&np), nx, ny, nz, out);
  STEAM_LAUNCH_CHECK("field_noise_kernel");
  return STEAM_OK;
}

extern "C" int steam_multiplier_noise(int64_t nx, int64_t ny, int64_t nz,
                                      const steam_noise_params* noise,
                                      float* out, steam_stream_t stream) {
  int rc = check_noise_params(noise);
  if (rc) return rc;
  if (nx < 0 || ny < 0 || nz < 0)
    return set_error(STEAM_EINVAL, "multiplier_noise: negative extent");
  if (nx * ny * nz == 0) return STEAM_OK;
  if (!out) return set_error(STEAM_EINVAL, "multiplier_noise: NULL output");
  hipLaunchKernelGGL((field_noise_kernel<true>), dim3(grid_for(nx * ny, 1)),
                     dim3(nz >= 256 ? 256 : 64), 0, (hipStream_t)stream,
                     make_noise_dev(noise), nx, ny, nz, out);
  STEAM_LAUNCH_CHECK("field_noise_kernel(multiplier)");
  return STEAM_OK;
}
