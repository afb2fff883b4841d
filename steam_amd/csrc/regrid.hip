// Periodic trilinear regrid (steam/utils.py:435-534): align_corners=False,
// x/y continued periodically (the wrapped pad + crop of _wrap_plan evaluated
// directly with modular source indices), z clamped.  Index and weight
// arithmetic follows torch's linear upsampling rules in float32
// (scale = float(in)/out, src = scale*(dst+0.5)-0.5 clamped at 0,
// i1 = i0 + (i0 < in-1)).  A target window gives the tiled form of
// steam/streaming.py:132-242 with no slab or pad at all.
#include <math.h>

#include "common.h"

namespace steam {

struct AxisMap {
  int n_in;     // source extent
  int pad;      // wrapped pad (source cells) on each side
  int crop;     // matching crop (target cells)
  int big_in;   // n_in + 2 pad
  float scale;  // float(big_in) / big_out
};

__device__ __forceinline__ void axis_weights(const AxisMap& m, int64_t j,
                                             int& i0, int& i1, float& l0,
                                             float& l1) {
  const float dst = (float)(j + m.crop);
  float src = __fsub_rn(__fmul_rn(m.scale, __fadd_rn(dst, 0.5f)), 0.5f);
  if (src < 0.f) src = 0.f;
  int ib = (int)floorf(src);
  if (ib > m.big_in - 1) ib = m.big_in - 1;
  float lam = __fsub_rn(src, (float)ib);
  lam = fminf(fmaxf(lam, 0.f), 1.f);
  int ib1 = ib + (ib < m.big_in - 1 ? 1 : 0);
  l1 = lam;
  l0 = __fsub_rn(1.f, lam);
  i0 = (int)pmod((int64_t)ib - m.pad, m.n_in);
  i1 = (int)pmod((int64_t)ib1 - m.pad, m.n_in);
}

__global__ __launch_bounds__(256) void regrid_kernel(
    const float* __restrict__ src, int64_t sy, int64_t sz, AxisMap mx,
    AxisMap my, AxisMap mz, int64_t x0, int64_t wx, int64_t y0, int64_t wy,
    int64_t tz, float scale, float* __restrict__ dst) {
  const int64_t lines = wx * wy;
  for (int64_t l = blockIdx.x; l < lines; l += gridDim.x) {
    const int64_t i = l / wy;
    const int64_t j = l - i * wy;
    int ax0, ax1, ay0, ay1;
    float lx0, lx1, ly0, ly1;
    axis_weights(mx, x0 + i, ax0, ax1, lx0, lx1);
    axis_weights(my, y0 + j, ay0, ay1, ly0, ly1);
    const float* r00 = src + ((int64_t)ax0 * sy + ay0) * sz;
    const float* r01 = src + ((int64_t)ax0 * sy + ay1) * sz;
    const float* r10 = src + ((int64_t)ax1 * sy + ay0) * sz;
    const float* r11 = src + ((int64_t)ax1 * sy + ay1) * sz;
    float* o = dst + l * tz;
    for (int64_t k = threadIdx.x; k < tz; k += blockDim.x) {
      int az0, az1;
      float lz0, lz1;
      axis_weights(mz, k, az0, az1, lz0, lz1);
      const float v000 = __fmul_rn(scale, r00[az0]), v001 = __fmul_rn(scale, r00[az1]);
      const float v010 = __fmul_rn(scale, r01[az0]), v011 = __fmul_rn(scale, r01[az1]);
      const float v100 = __fmul_rn(scale, r10[az0]), v101 = __fmul_rn(scale, r10[az1]);
      const float v110 = __fmul_rn(scale, r11[az0]), v111 = __fmul_rn(scale, r11[az1]);
      const float a = ly0 * (lz0 * v000 + lz1 * v001) + ly1 * (lz0 * v010 + lz1 * v011);
      const float b = ly0 * (lz0 * v100 + lz1 * v101) + ly1 * (lz0 * v110 + lz1 * v111);
      o[k] = lx0 * a + lx1 * b;
    }
  }
}

static int gcd64(int64_t a, int64_t b) {
  while (b) {
    int64_t t = a % b;
    a = b;
    b = t;
  }
  return (int)a;
}

static AxisMap make_map(int64_t n_in, int64_t n_out, bool periodic) {
  AxisMap m;
  m.n_in = (int)n_in;
  if (periodic) {
    int g = gcd64(n_in, n_out);
    m.pad = (int)(n_in / g);
    m.crop = (int)(n_out / g);
  } else {
    m.pad = 0;
    m.crop = 0;
  }
  m.big_in = (int)(n_in + 2 * m.pad);
  const int64_t big_out = n_out + 2 * m.crop;
  m.scale = (float)m.big_in / (float)big_out;
  return m;
}

}  // namespace steam

using namespace steam;

extern "C" int steam_regrid_trilinear(const float* src, int64_t sx, int64_t sy,
                                      int64_t sz, int64_t tx, int64_t ty,
                                      int64_t tz, int64_t x0, int64_t wx,
                                      int64_t y0, int64_t wy,
                                      const int32_t* periodic, float scale,
                                      float* dst, steam_stream_t stream) {
  if (sx < 1 || sy < 1 || sz < 1 || tx < 1 || ty < 1 || tz < 1)
    return set_error(STEAM_EINVAL, "regrid: extents must be positive");
  if (!src || !dst) return set_error(STEAM_EINVAL, "regrid: NULL pointer");
  if (wx < 0 || wy < 0 || x0 < 0 || y0 < 0 || x0 + wx > tx || y0 + wy > ty)
    return set_error(STEAM_EINVAL, "regrid: window outside the target grid");
  const int64_t n_in[3] = {sx, sy, sz}, n_out[3] = {tx, ty, tz};
  const bool per[3] = {periodic ? periodic[0] != 0 : true,
                       periodic ? periodic[1] != 0 : true,
                       periodic ? periodic[2] != 0 : false};
  for (int a = 0; a < 3; ++a)
    if (per[a] && n_out[a] < n_in[a])
      return set_error(STEAM_ENOTIMPL,
                       "Periodic resampling downsamples; the regrid only refines");
  if (wx * wy == 0) return STEAM_OK;
  AxisMap mx = make_map(sx, tx, per[0]);
  AxisMap my = make_map(sy, ty, per[1]);
  AxisMap mz = make_map(sz, tz, per[2]);
  int64_t lines = wx * wy;
  int grid = (int)(lines < (1 << 20) ? lines : (1 << 20));
  hipLaunchKernelGGL(regrid_kernel, dim3(grid), dim3(tz >= 256 ? 256 : 64), 0,
                     (hipStream_t)stream, src, sy, sz, mx, my, mz, x0, wx, y0, wy,
                     tz, scale, dst);
  STEAM_LAUNCH_CHECK("regrid_kernel");
  return STEAM_OK;
}
