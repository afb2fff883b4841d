// Scalar-increment path of the cascade (steam/simulate.py:1924-2029):
// the advective weight, bound taper and joint pattern, per-level window
// reductions, and the bounded amplitude-preserving add with its ledger.
//
// Fields are float32 (nx, ny, nz), z fastest; a level is the (x, y) plane at
// one z, and every reduction is per level over the window [x0,x1) x [y0,y1).
// Per-level reductions run in float64 in a fixed order (workgroup partials,
// then a tree over workgroups), so results do not depend on launch
// composition or timing.
#include <math.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "common.h"

// Every operation below rounds separately, as the reference's float32 numpy
// arithmetic does (HIP would otherwise contract a*b+c into one fma).  The
// pragma governs the operators written here; the file is also built with
// -ffp-contract=off, which reaches the __f*_rn header functions.
#pragma clang fp contract(off)

namespace steam {

// z differences of np.gradient (steam/simulate.py:2709-2741), float32
struct ZDiff {
  int uniform;
  float two_dz, dz_first, dz_last;
  const float* a;  // [nz-2] coefficients of f[z-1], f[z], f[z+1]
  const float* b;
  const float* c;
};

// ---------------------------------------------------------------------------
// joint pattern W * S_k * g  (steam/simulate.py:1935-1963 with _advective_weight
// 2839-2868 / 2787-2836 and _bound_taper 2905-2930): per cell, in the
// reference's float32 operation order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void pattern_kernel(
    const float* __restrict__ pert, const float* __restrict__ mean_z,
    const float* __restrict__ sk, const float* __restrict__ b_z,
    const float* __restrict__ aspect_z, ZDiff zd, int nx, int ny, int nz,
    float two_dx, float two_dy, float lo, float span, float* __restrict__ out) {
  const int64_t n = (int64_t)nx * ny * nz;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int z = (int)(i % nz);
    const int64_t line = i / nz;
    const int x = (int)(line / ny), y = (int)(line - (int64_t)x * ny);
    const int xp = x + 1 < nx ? x + 1 : 0, xm = x > 0 ? x - 1 : nx - 1;
    const int yp = y + 1 < ny ? y + 1 : 0, ym = y > 0 ? y - 1 : ny - 1;
    const float mz = mean_z[z];
    auto phi = [&](int xx, int yy, int zz, float m) {
      return __fadd_rn(pert[((int64_t)xx * ny + yy) * nz + zz], m);
    };
    const float gx = __fdiv_rn(__fsub_rn(phi(xp, y, z, mz), phi(xm, y, z, mz)), two_dx);
    const float gy = __fdiv_rn(__fsub_rn(phi(x, yp, z, mz), phi(x, ym, z, mz)), two_dy);
    float w = sqrtf(__fadd_rn(__fmul_rn(gx, gx), __fmul_rn(gy, gy)));
    const float fc = phi(x, y, z, mz);
    float gz;
    if (nz == 1) {
      gz = 0.f;
    } else if (z == 0) {
      gz = fabsf(__fdiv_rn(__fsub_rn(phi(x, y, 1, mean_z[1]), fc), zd.dz_first));
    } else if (z == nz - 1) {
      gz = fabsf(__fdiv_rn(__fsub_rn(fc, phi(x, y, nz - 2, mean_z[nz - 2])), zd.dz_last));
    } else {
      const float fm = phi(x, y, z - 1, mean_z[z - 1]);
      const float fp = phi(x, y, z + 1, mean_z[z + 1]);
      if (zd.uniform) {
        gz = fabsf(__fdiv_rn(__fsub_rn(fp, fm), zd.two_dz));
      } else {
        gz = fabsf(__fadd_rn(__fadd_rn(__fmul_rn(zd.a[z - 1], fm), __fmul_rn(zd.b[z - 1], fc)),
                             __fmul_rn(zd.c[z - 1], fp)));
      }
    }
    if (nz > 1) w = __fadd_rn(__fmul_rn(gz, aspect_z[z]), w);
    if (!sk) {  // the advective weight alone
      out[i] = w;
      continue;
    }
    w = __fmul_rn(w, sk[i]);
    // bound taper of the running field
    float g = __fsub_rn(fc, lo);
    g = fminf(g, __fsub_rn(span, g));
    const float bz = b_z[z];
    g = __fdiv_rn(g, bz > 0.f ? bz : 1e-30f);
    g = fminf(fmaxf(g, 0.f), 1.f);
    out[i] = __fmul_rn(w, g);
  }
}

// W = W / (level_mean > 0 ? level_mean : 1) * C_k[z]  (steam/simulate.py:1986-1990)
__global__ __launch_bounds__(256) void pattern_normalize_kernel(
    float* __restrict__ w, const float* __restrict__ level_mean,
    const float* __restrict__ c_z, int64_t n, int nz) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int z = (int)(i % nz);
    const float lm = level_mean[z];
    w[i] = __fmul_rn(__fdiv_rn(w[i], lm > 0.f ? lm : 1.f), c_z[z]);
  }
}

// ---------------------------------------------------------------------------
// per-level window reductions
// ---------------------------------------------------------------------------
enum LevelOp {
  LV_SUM = 0,        // sum f
  LV_ABS = 1,        // sum |f|
  LV_NONZERO = 2,    // count f != 0
  LV_CLAMP = 3,      // sum clamp(f - f32(mu[z]), lo, hi)
  LV_PHI = 4,        // sum (f + mean[z]) (the running field)
  LV_MIN = 5,        // min (f + mean[z]) over the whole level (not windowed)
  LV_MAX = 6,        // max (f + mean[z]) over the whole level
  // the same field in float64, e = (double)f + (double)mean[z] (exact)
  LV_E_SUM = 7,      // sum e
  LV_E_MIN = 8,      // min e over the whole level
  LV_E_MAX = 9,      // max e over the whole level
  LV_E_CLAMP = 10,   // sum clamp(e - mu[z], lo, hi) on levels with active[z]
};

struct LevelArgs {
  const float* f;
  const float* mean_z;
  const double* mu;
  const double* active;  // LV_E_CLAMP: levels to reduce (others give 0)
  int nx, ny, nz;
  int win[4];
  int op;
  double lo, hi;
};

constexpr int LV_LINES = 32;   // lines per first-stage workgroup
constexpr int LV_ROWS = 64;    // partial rows per combine workgroup

__device__ __forceinline__ bool lv_is_min(int op) { return op == LV_MIN || op == LV_E_MIN; }
__device__ __forceinline__ bool lv_is_max(int op) { return op == LV_MAX || op == LV_E_MAX; }
__device__ __forceinline__ double lv_init(int op) {
  return lv_is_min(op) ? INFINITY : (lv_is_max(op) ? -INFINITY : 0.0);
}
__device__ __forceinline__ double lv_join(int op, double a, double b) {
  return lv_is_min(op) ? fmin(a, b) : (lv_is_max(op) ? fmax(a, b) : a + b);
}

// first stage: workgroup w covers lines [w*LV_LINES, ...); thread t handles
// levels t, t+256, ...; partial[w][z]
__global__ __launch_bounds__(256) void level_partial_kernel(LevelArgs a, double* part) {
  const int64_t nlines = (int64_t)a.nx * a.ny;
  const int64_t l0 = (int64_t)blockIdx.x * LV_LINES;
  const bool whole = lv_is_min(a.op) || lv_is_max(a.op);
  const float flo = (float)a.lo, fhi = (float)a.hi;
  for (int z = threadIdx.x; z < a.nz; z += blockDim.x) {
    double acc = lv_init(a.op);
    const float mz = a.mean_z ? a.mean_z[z] : 0.f;
    const float muz = a.mu ? (float)a.mu[z] : 0.f;
    const double mu64 = a.mu ? a.mu[z] : 0.0;
    if (a.op == LV_E_CLAMP && a.active[z] == 0.0) {
      part[(int64_t)blockIdx.x * a.nz + z] = 0.0;
      continue;
    }
    for (int j = 0; j < LV_LINES; ++j) {
      const int64_t line = l0 + j;
      if (line >= nlines) break;
      const int x = (int)(line / a.ny), y = (int)(line - (int64_t)x * a.ny);
      if (!whole && (x < a.win[0] || x >= a.win[1] || y < a.win[2] || y >= a.win[3]))
        continue;
      const float v = a.f[line * a.nz + z];
      double t;
      switch (a.op) {
        case LV_SUM: t = (double)v; break;
        case LV_ABS: t = (double)fabsf(v); break;
        case LV_NONZERO: t = v != 0.f ? 1.0 : 0.0; break;
        case LV_CLAMP: t = (double)fminf(fmaxf(__fsub_rn(v, muz), flo), fhi); break;
        case LV_E_SUM:
        case LV_E_MIN:
        case LV_E_MAX: t = (double)v + (double)mz; break;
        case LV_E_CLAMP: t = fmin(fmax(((double)v + (double)mz) - mu64, a.lo), a.hi); break;
        default: t = (double)__fadd_rn(v, mz); break;
      }
      acc = lv_join(a.op, acc, t);
    }
    part[(int64_t)blockIdx.x * a.nz + z] = acc;
  }
}

// combine stage: rows [r0, r0 + LV_ROWS) of a partial table into one row;
// block (64 levels x 4 row groups), fixed order
__global__ __launch_bounds__(256) void level_combine_kernel(const double* in, int64_t rows,
                                                            int nz, int op, double* out) {
  __shared__ double red[4][64];
  const int zl = threadIdx.x & 63, grp = threadIdx.x >> 6;
  const int z = blockIdx.x * 64 + zl;
  const int64_t r0 = (int64_t)blockIdx.y * LV_ROWS;
  double acc = lv_init(op);
  if (z < nz) {
    for (int j = grp; j < LV_ROWS; j += 4) {
      const int64_t r = r0 + j;
      if (r >= rows) break;
      acc = lv_join(op, acc, in[r * nz + z]);
    }
  }
  red[grp][zl] = acc;
  __syncthreads();
  if (grp == 0 && z < nz) {
    double s = red[0][zl];
    for (int g = 1; g < 4; ++g) s = lv_join(op, s, red[g][zl]);
    out[(int64_t)blockIdx.y * nz + z] = s;
  }
}

// ---------------------------------------------------------------------------
// bounded amplitude-preserving add, device pieces (the algorithm of
// steam/simulate.py:2985-3111, the reference's GPU path)
// ---------------------------------------------------------------------------
// d = clamp(d - f32(shift[z]) + phi, lo, hi) - phi  (shift may be NULL)
__global__ __launch_bounds__(256) void ba_shift_clip_kernel(
    float* __restrict__ d, const float* __restrict__ pert, const float* __restrict__ mean_z,
    const double* __restrict__ shift, int64_t n, int nz, float lo, float hi) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int z = (int)(i % nz);
    const float phi = __fadd_rn(pert[i], mean_z[z]);
    float v = d[i];
    if (shift) v = __fsub_rn(v, (float)shift[z]);
    const float e = fminf(fmaxf(__fadd_rn(v, phi), lo), hi);
    d[i] = __fsub_rn(e, phi);
  }
}

// d *= f32(scale[z])
__global__ __launch_bounds__(256) void ba_scale_kernel(float* __restrict__ d,
                                                       const double* __restrict__ scale,
                                                       int64_t n, int nz) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = __fmul_rn(d[i], (float)scale[i % nz]);
}

// e = d + phi (in place on d)
__global__ __launch_bounds__(256) void ba_to_field_kernel(float* __restrict__ d,
                                                          const float* __restrict__ pert,
                                                          const float* __restrict__ mean_z,
                                                          int64_t n, int nz) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int z = (int)(i % nz);
    d[i] = __fadd_rn(d[i], __fadd_rn(pert[i], mean_z[z]));
  }
}

// final write: v = clamp(e - f32(mu[z]), lo, hi) - mean[z]; with record,
// inc = v - pert; pert = v
__global__ __launch_bounds__(256) void ba_finish_kernel(const float* __restrict__ e,
                                                        float* __restrict__ pert,
                                                        float* __restrict__ inc,
                                                        const float* __restrict__ mean_z,
                                                        const double* __restrict__ mu,
                                                        int64_t n, int nz, float lo, float hi,
                                                        int record) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int z = (int)(i % nz);
    float v = fminf(fmaxf(__fsub_rn(e[i], (float)mu[z]), lo), hi);
    v = __fsub_rn(v, mean_z[z]);
    if (record) inc[i] = __fsub_rn(v, pert[i]);
    pert[i] = v;
  }
}

// per-level scalar steps (one thread per level)
struct BaVec {
  double* a0;
  double* done;       // 1.0 / 0.0
  double* sum;        // last reduction (per level)
  double* demean;     // step vectors
  double* scale;
  double* mu;
  double* mu_lo;
  double* mu_hi;
  double* mean_phi;
};

// demean = done ? 0 : sum / n
__global__ void ba_demean_kernel(BaVec v, int nz, double npts) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z < nz) v.demean[z] = v.done[z] != 0.0 ? 0.0 : v.sum[z] / npts;
}
// m_abs = sum / n; scale = min(a0 / max(m_abs, 1e-300), 2); done |= ...;
// applied = done ? 1 : scale  (written to scale)
__global__ void ba_rescale_kernel(BaVec v, int nz, double npts) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z < nz) {
    const double m_abs = v.sum[z] / npts;
    const double s = fmin(v.a0[z] / fmax(m_abs, 1e-300), 2.0);
    const bool d = v.done[z] != 0.0 || m_abs <= 0.0 || fabs(s - 1.0) < 1e-4;
    v.done[z] = d ? 1.0 : 0.0;
    v.scale[z] = d ? 1.0 : s;
  }
}
__global__ void ba_bisect_init_kernel(BaVec v, int nz, double width) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z < nz) {
    v.mu_lo[z] = -width;
    v.mu_hi[z] = width;
    v.mu[z] = 0.5 * (v.mu_lo[z] + v.mu_hi[z]);
  }
}
// positive = sum / n - mean_phi > 0; narrow the bracket; next midpoint
__global__ void ba_bisect_step_kernel(BaVec v, int nz, double npts) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z < nz) {
    const bool positive = v.sum[z] / npts - v.mean_phi[z] > 0.0;
    if (positive)
      v.mu_lo[z] = v.mu[z];
    else
      v.mu_hi[z] = v.mu[z];
    v.mu[z] = 0.5 * (v.mu_lo[z] + v.mu_hi[z]);
  }
}
// copy sum / n into dst (a level mean)
__global__ void ba_mean_kernel(const double* sum, double* dst, int nz, double npts) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z < nz) dst[z] = sum[z] / npts;
}
// done = a0 <= 0
__global__ void ba_init_done_kernel(BaVec v, int nz) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z < nz) v.done[z] = v.a0[z] <= 0.0 ? 1.0 : 0.0;
}

// ---------------------------------------------------------------------------
// compensation deficits and output composition (steam/simulate.py:1893-1919,
// 1999-2003, 2015-2016, 3558-3649)
// ---------------------------------------------------------------------------
// deficit += f32(coef[z] * inc), inc = after - before (or after alone)
__global__ __launch_bounds__(256) void deficit_kernel(float* __restrict__ deficit,
                                                      const float* __restrict__ after,
                                                      const float* __restrict__ before,
                                                      const float* __restrict__ coef_z,
                                                      int64_t n, int nz) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int z = (int)(i % nz);
    float inc = after[i];
    if (before) inc = __fsub_rn(inc, before[i]);
    deficit[i] = __fadd_rn(deficit[i], __fmul_rn(coef_z[z], inc));
  }
}

// d += s
__global__ __launch_bounds__(256) void add_state_kernel(float* __restrict__ d,
                                                        const float* __restrict__ state,
                                                        int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = __fadd_rn(d[i], state[i]);
}

// out = clamp(src + mean[z], lo, hi) (float32, the written scalar field)
// (out may be src)
__global__ __launch_bounds__(256) void compose_kernel(const float* src,
                                                      const float* __restrict__ mean_z,
                                                      float* out, int64_t n,
                                                      int nz, float lo, float hi) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int z = (int)(i % nz);
    out[i] = fminf(fmaxf(__fadd_rn(src[i], mean_z[z]), lo), hi);
  }
}

// projection onto the bounds: active[z] = level violates; bracket set up
__global__ void proj_init_kernel(BaVec v, const double* vmin, const double* vmax, double* active,
                                 int nz, double lo, double hi, double width) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z < nz) {
    active[z] = (vmin[z] < lo || vmax[z] > hi) ? 1.0 : 0.0;
    v.mu_lo[z] = -width;
    v.mu_hi[z] = width;
    v.mu[z] = 0.5 * (v.mu_lo[z] + v.mu_hi[z]);
  }
}
// recorded mu: NaN on levels left alone
__global__ void proj_record_kernel(const double* mu, const double* active, double* out, int nz) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z < nz) out[z] = active[z] != 0.0 ? mu[z] : __builtin_nan("");
}
// replay: active where the recorded mu is a number
__global__ void proj_active_kernel(const double* mu, double* active, int nz) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z < nz) active[z] = (mu[z] == mu[z]) ? 1.0 : 0.0;
}
// f = f32(clamp(e - mu, lo, hi) - mean) on active levels; e = f + mean in float64
__global__ __launch_bounds__(256) void proj_apply_kernel(float* __restrict__ f,
                                                         const float* __restrict__ mean_z,
                                                         const double* __restrict__ mu,
                                                         const double* __restrict__ active,
                                                         int64_t n, int nz, double lo,
                                                         double hi) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int z = (int)(i % nz);
    if (active[z] == 0.0) continue;
    const double m = (double)mean_z[z];
    const double e = (double)f[i] + m;
    f[i] = (float)(fmin(fmax(e - mu[z], lo), hi) - m);
  }
}

// flux composition: d = max(d + state, 0)
__global__ __launch_bounds__(256) void flux_compose_kernel(float* __restrict__ d,
                                                           const float* __restrict__ state,
                                                           int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = fmaxf(__fadd_rn(d[i], state[i]), 0.f);
}
// total of per-level sums, in level order, and its mean over count cells
// (one thread)
__global__ void sum_levels_kernel(const double* lv, int nz, double count, double* total,
                                  double* mean) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double s = 0.0;
    for (int z = 0; z < nz; ++z) s += lv[z];
    *total = s;
    *mean = s / count;
  }
}
// rescale to the entering volume mean: d *= f32(em / rm), or d = f32(em)
// when nothing is left (means[0] entering, means[1] realized)
__global__ __launch_bounds__(256) void flux_restore_kernel(float* __restrict__ d,
                                                           const double* __restrict__ means,
                                                           int64_t n) {
  const double em = means[0], rm = means[1];
  const bool fill = !(rm > 0.0);
  const float r = fill ? (float)em : (float)(em / rm);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = fill ? r : __fmul_rn(d[i], r);
}

}  // namespace steam

using namespace steam;

namespace {

int grid_1d(int64_t n) {
  int64_t g = (n + 255) / 256;
  return (int)std::min<int64_t>(std::max<int64_t>(g, 1), 1 << 16);
}

// bytes of reduction scratch for a field of nlines lines and nz levels
int64_t level_scratch_bytes(int64_t nlines, int nz) {
  const int64_t rows = (nlines + LV_LINES - 1) / LV_LINES;
  const int64_t rows2 = (rows + LV_ROWS - 1) / LV_ROWS;
  return (rows + rows2 + 1) * (int64_t)nz * 8;
}

// per-level reduction of a into out[nz] (device); scratch from the caller
int level_reduce(const LevelArgs& a, double* out, double* scratch, hipStream_t s) {
  const int64_t nlines = (int64_t)a.nx * a.ny;
  int64_t rows = (nlines + LV_LINES - 1) / LV_LINES;
  double* p0 = scratch;
  double* p1 = scratch + rows * a.nz;
  hipLaunchKernelGGL(level_partial_kernel, dim3((unsigned)rows), dim3(256), 0, s, a, p0);
  STEAM_LAUNCH_CHECK("level_partial_kernel");
  double* src = p0;
  double* dst = p1;
  while (rows > 1) {
    const int64_t next = (rows + LV_ROWS - 1) / LV_ROWS;
    double* target = (next == 1) ? out : dst;
    hipLaunchKernelGGL(level_combine_kernel, dim3((a.nz + 63) / 64, (unsigned)next), dim3(256),
                       0, s, src, rows, a.nz, a.op, target);
    STEAM_LAUNCH_CHECK("level_combine_kernel");
    rows = next;
    std::swap(src, dst);
    src = target;
    dst = (target == p0) ? p1 : p0;
  }
  if (src != out)
    STEAM_HIP(hipMemcpyAsync(out, src, (size_t)a.nz * 8, hipMemcpyDeviceToDevice, s));
  return STEAM_OK;
}

int check_window(const int64_t* w, int nx, int ny, int out[4]) {
  if (!w) {
    out[0] = 0; out[1] = nx; out[2] = 0; out[3] = ny;
    return STEAM_OK;
  }
  if (w[0] < 0 || w[1] > nx || w[0] >= w[1] || w[2] < 0 || w[3] > ny || w[2] >= w[3])
    return set_error(STEAM_EINVAL, "window out of range");
  for (int i = 0; i < 4; ++i) out[i] = (int)w[i];
  return STEAM_OK;
}

}  // namespace

extern "C" int steam_scalar_pattern(const float* pert_dev, const float* mean_z_dev,
                                    const float* sk_dev, const float* b_z_dev,
                                    const float* aspect_z_dev, const steam_zdiff* zdiff,
                                    int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                                    double phi_min, double phi_max, float* out_dev,
                                    steam_stream_t stream) {
  if (!pert_dev || !mean_z_dev || (sk_dev && !b_z_dev) || !aspect_z_dev || !zdiff || !out_dev)
    return set_error(STEAM_EINVAL, "steam_scalar_pattern: NULL argument");
  if (nx < 1 || ny < 1 || nz < 1 || nx * ny * nz >= (1ll << 40))
    return set_error(STEAM_EINVAL, "steam_scalar_pattern: bad shape");
  if (!zdiff->uniform && nz > 2 && (!zdiff->a_dev || !zdiff->b_dev || !zdiff->c_dev))
    return set_error(STEAM_EINVAL, "steam_scalar_pattern: non-uniform z needs coefficients");
  ZDiff zd;
  zd.uniform = zdiff->uniform;
  zd.two_dz = zdiff->two_dz;
  zd.dz_first = zdiff->dz_first;
  zd.dz_last = zdiff->dz_last;
  zd.a = zdiff->a_dev;
  zd.b = zdiff->b_dev;
  zd.c = zdiff->c_dev;
  const float lo = (float)phi_min;
  const float span = (float)(phi_max - phi_min);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(pattern_kernel, dim3(grid_1d(nx * ny * nz)), dim3(256), 0, s, pert_dev,
                     mean_z_dev, sk_dev, b_z_dev, aspect_z_dev, zd, (int)nx, (int)ny, (int)nz,
                     (float)(2.0 * dx), (float)(2.0 * dy), lo, span, out_dev);
  STEAM_LAUNCH_CHECK("pattern_kernel");
  return STEAM_OK;
}

extern "C" int steam_pattern_normalize(float* w_dev, const float* level_mean_dev,
                                       const float* c_z_dev, int64_t n, int64_t nz,
                                       steam_stream_t stream) {
  if (!w_dev || !level_mean_dev || !c_z_dev || nz < 1 || n % nz)
    return set_error(STEAM_EINVAL, "steam_pattern_normalize: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(pattern_normalize_kernel, dim3(grid_1d(n)), dim3(256), 0, s, w_dev,
                     level_mean_dev, c_z_dev, n, (int)nz);
  STEAM_LAUNCH_CHECK("pattern_normalize_kernel");
  return STEAM_OK;
}

extern "C" int64_t steam_level_scratch_bytes(int64_t nx, int64_t ny, int64_t nz) {
  return level_scratch_bytes(nx * ny, (int)nz);
}

extern "C" int steam_level_reduce(const float* f_dev, const float* mean_z_dev,
                                  const double* mu_dev, int64_t nx, int64_t ny, int64_t nz,
                                  const int64_t* window, int op, double lo, double hi,
                                  double* out_dev, void* scratch_dev, int64_t scratch_bytes,
                                  steam_stream_t stream) {
  if (!f_dev || !out_dev || !scratch_dev || nx < 1 || ny < 1 || nz < 1)
    return set_error(STEAM_EINVAL, "steam_level_reduce: bad arguments");
  if (op < LV_SUM || op > LV_E_MAX) return set_error(STEAM_EINVAL, "steam_level_reduce: unknown op");
  if ((op == LV_PHI || op == LV_MIN || op == LV_MAX || op >= LV_E_SUM) && !mean_z_dev)
    return set_error(STEAM_EINVAL, "steam_level_reduce: op needs mean_z");
  if (op == LV_CLAMP && !mu_dev) return set_error(STEAM_EINVAL, "steam_level_reduce: op needs mu");
  if (scratch_bytes < level_scratch_bytes(nx * ny, (int)nz))
    return set_error(STEAM_EINVAL, "steam_level_reduce: scratch too small");
  LevelArgs a;
  memset(&a, 0, sizeof(a));
  a.f = f_dev;
  a.mean_z = mean_z_dev;
  a.mu = mu_dev;
  a.nx = (int)nx;
  a.ny = (int)ny;
  a.nz = (int)nz;
  a.op = op;
  a.lo = (double)(float)lo;
  a.hi = (double)(float)hi;
  int rc = check_window(window, (int)nx, (int)ny, a.win);
  if (rc) return rc;
  return level_reduce(a, out_dev, (double*)scratch_dev, (hipStream_t)stream);
}

extern "C" int64_t steam_bounded_add_workspace_bytes(int64_t nx, int64_t ny, int64_t nz) {
  const int64_t n = nx * ny * nz;
  (void)n;
  return 16 * nz * 8 + level_scratch_bytes(nx * ny, (int)nz) + 256;
}

extern "C" int steam_bounded_add(float* pert_dev, const float* mean_z_dev, float* inc_dev,
                                 int64_t nx, int64_t ny, int64_t nz, const int64_t* window,
                                 double phi_min, double phi_max, int n_rescale,
                                 int record_applied, steam_bounded_add_ledger* ledger,
                                 int replay, void* workspace_dev, int64_t workspace_bytes,
                                 steam_stream_t stream) {
  if (!pert_dev || !mean_z_dev || !inc_dev || !ledger || !workspace_dev)
    return set_error(STEAM_EINVAL, "steam_bounded_add: NULL argument");
  if (nx < 1 || ny < 1 || nz < 1 || n_rescale < 0 || n_rescale > 64)
    return set_error(STEAM_EINVAL, "steam_bounded_add: bad shape or n_rescale");
  if (!ledger->a0 || !ledger->demean || !ledger->scale || !ledger->n_loop ||
      !ledger->n_scale || !ledger->final_demean || !ledger->mu)
    return set_error(STEAM_EINVAL, "steam_bounded_add: ledger arrays missing");
  if (workspace_bytes < steam_bounded_add_workspace_bytes(nx, ny, nz))
    return set_error(STEAM_EINVAL, "steam_bounded_add: workspace too small");
  hipStream_t s = (hipStream_t)stream;
  int win[4];
  int rc = check_window(window, (int)nx, (int)ny, win);
  if (rc) return rc;
  const int64_t n = nx * ny * nz;
  const double npts = (double)(win[1] - win[0]) * (double)(win[3] - win[2]);
  const float lo = (float)phi_min, hi = (float)phi_max;
  const double width = phi_max - phi_min;
  // the solve runs in place on inc; the workspace holds per-level vectors
  // and the reduction scratch only
  double* vec = (double*)workspace_dev;
  BaVec v;
  v.a0 = vec;
  v.done = vec + nz;
  v.sum = vec + 2 * nz;
  v.demean = vec + 3 * nz;
  v.scale = vec + 4 * nz;
  v.mu = vec + 5 * nz;
  v.mu_lo = vec + 6 * nz;
  v.mu_hi = vec + 7 * nz;
  v.mean_phi = vec + 8 * nz;
  double* red = vec + 16 * nz;
  const int gz = (int)((nz + 255) / 256);
  const int gn = grid_1d(n);
  LevelArgs a;
  memset(&a, 0, sizeof(a));
  a.nx = (int)nx;
  a.ny = (int)ny;
  a.nz = (int)nz;
  memcpy(a.win, win, sizeof(win));
  a.mean_z = mean_z_dev;
  a.lo = (double)lo;
  a.hi = (double)hi;
  auto reduce = [&](const float* f, int op, double* out, const double* mu) {
    LevelArgs b = a;
    b.f = f;
    b.op = op;
    b.mu = mu;
    return level_reduce(b, out, red, s);
  };
  std::vector<double> host(nz);
  auto down = [&](const double* src, double* dst) -> int {
    STEAM_HIP(hipMemcpyAsync(dst, src, (size_t)nz * 8, hipMemcpyDeviceToHost, s));
    return STEAM_OK;
  };
  auto up = [&](const double* src, double* dst) -> int {
    STEAM_HIP(hipMemcpyAsync(dst, src, (size_t)nz * 8, hipMemcpyHostToDevice, s));
    return STEAM_OK;
  };

  if (replay) {
    // apply-only: the recorded per-level scalars, no reductions
    int nloop = 0;
    for (int z = 0; z < nz; ++z) nloop = std::max(nloop, (int)ledger->n_loop[z]);
    std::vector<double> col(nz);
    for (int step = 0; step < nloop; ++step) {
      for (int z = 0; z < nz; ++z) col[z] = ledger->demean[(int64_t)z * n_rescale + step];
      if ((rc = up(col.data(), v.demean))) return rc;
      hipLaunchKernelGGL(ba_shift_clip_kernel, dim3(gn), dim3(256), 0, s, inc_dev, pert_dev,
                         mean_z_dev, v.demean, n, (int)nz, lo, hi);
      for (int z = 0; z < nz; ++z) col[z] = ledger->scale[(int64_t)z * n_rescale + step];
      if ((rc = up(col.data(), v.scale))) return rc;
      hipLaunchKernelGGL(ba_scale_kernel, dim3(gn), dim3(256), 0, s, inc_dev, v.scale, n,
                         (int)nz);
      STEAM_HIP(hipStreamSynchronize(s));
    }
    if ((rc = up(ledger->final_demean, v.demean))) return rc;
    hipLaunchKernelGGL(ba_shift_clip_kernel, dim3(gn), dim3(256), 0, s, inc_dev, pert_dev,
                       mean_z_dev, v.demean, n, (int)nz, lo, hi);
    if ((rc = up(ledger->mu, v.mu))) return rc;
  } else {
    // mean of the running field over the window, and the target amplitude
    if ((rc = reduce(pert_dev, LV_PHI, v.sum, nullptr))) return rc;
    hipLaunchKernelGGL(ba_mean_kernel, dim3(gz), dim3(256), 0, s, v.sum, v.mean_phi, (int)nz, npts);
    if ((rc = reduce(inc_dev, LV_ABS, v.sum, nullptr))) return rc;
    hipLaunchKernelGGL(ba_mean_kernel, dim3(gz), dim3(256), 0, s, v.sum, v.a0, (int)nz, npts);
    hipLaunchKernelGGL(ba_init_done_kernel, dim3(gz), dim3(256), 0, s, v, (int)nz);
    if ((rc = down(v.a0, ledger->a0))) return rc;
    for (int64_t i = 0; i < nz * (int64_t)n_rescale; ++i) {
      ledger->demean[i] = 0.0;
      ledger->scale[i] = 1.0;
    }
    int n_loop = 0;
    std::vector<double> dm(nz), sc(nz), dn(nz);
    for (int step = 0; step < n_rescale; ++step) {
      if ((rc = reduce(inc_dev, LV_SUM, v.sum, nullptr))) return rc;
      hipLaunchKernelGGL(ba_demean_kernel, dim3(gz), dim3(256), 0, s, v, (int)nz, npts);
      hipLaunchKernelGGL(ba_shift_clip_kernel, dim3(gn), dim3(256), 0, s, inc_dev, pert_dev,
                         mean_z_dev, v.demean, n, (int)nz, lo, hi);
      n_loop = step + 1;
      if ((rc = reduce(inc_dev, LV_ABS, v.sum, nullptr))) return rc;
      hipLaunchKernelGGL(ba_rescale_kernel, dim3(gz), dim3(256), 0, s, v, (int)nz, npts);
      if ((rc = down(v.demean, dm.data()))) return rc;
      if ((rc = down(v.done, dn.data()))) return rc;
      if ((rc = down(v.scale, sc.data()))) return rc;
      STEAM_HIP(hipStreamSynchronize(s));
      bool all = true;
      for (int z = 0; z < nz; ++z) {
        ledger->demean[(int64_t)z * n_rescale + step] = dm[z];
        all = all && dn[z] != 0.0;
      }
      if (all) break;
      for (int z = 0; z < nz; ++z) ledger->scale[(int64_t)z * n_rescale + step] = sc[z];
      hipLaunchKernelGGL(ba_scale_kernel, dim3(gn), dim3(256), 0, s, inc_dev, v.scale, n,
                         (int)nz);
    }
    for (int z = 0; z < nz; ++z) {
      ledger->n_loop[z] = n_loop;
      ledger->n_scale[z] = n_loop;
    }
    if ((rc = reduce(inc_dev, LV_SUM, v.sum, nullptr))) return rc;
    hipLaunchKernelGGL(ba_mean_kernel, dim3(gz), dim3(256), 0, s, v.sum, v.demean, (int)nz, npts);
    if ((rc = down(v.demean, ledger->final_demean))) return rc;
    hipLaunchKernelGGL(ba_shift_clip_kernel, dim3(gn), dim3(256), 0, s, inc_dev, pert_dev,
                       mean_z_dev, v.demean, n, (int)nz, lo, hi);
  }
  // e = d + phi, in place; bisect mu in the shifted space
  hipLaunchKernelGGL(ba_to_field_kernel, dim3(gn), dim3(256), 0, s, inc_dev, pert_dev,
                     mean_z_dev, n, (int)nz);
  if (!replay) {
    hipLaunchKernelGGL(ba_bisect_init_kernel, dim3(gz), dim3(256), 0, s, v, (int)nz, width);
    for (int it = 0; it < 60; ++it) {
      if ((rc = reduce(inc_dev, LV_CLAMP, v.sum, v.mu))) return rc;
      hipLaunchKernelGGL(ba_bisect_step_kernel, dim3(gz), dim3(256), 0, s, v, (int)nz, npts);
    }
    if ((rc = down(v.mu, ledger->mu))) return rc;
  }
  hipLaunchKernelGGL(ba_finish_kernel, dim3(gn), dim3(256), 0, s, inc_dev, pert_dev, inc_dev,
                     mean_z_dev, v.mu, n, (int)nz, lo, hi, record_applied);
  STEAM_LAUNCH_CHECK("bounded add");
  STEAM_HIP(hipStreamSynchronize(s));
  return STEAM_OK;
}

extern "C" int steam_deficit_accumulate(float* deficit_dev, const float* after_dev,
                                        const float* before_dev, const float* coef_z_dev,
                                        int64_t n, int64_t nz, steam_stream_t stream) {
  if (!deficit_dev || !after_dev || !coef_z_dev || nz < 1 || n % nz)
    return set_error(STEAM_EINVAL, "steam_deficit_accumulate: bad arguments");
  hipLaunchKernelGGL(deficit_kernel, dim3(grid_1d(n)), dim3(256), 0, (hipStream_t)stream,
                     deficit_dev, after_dev, before_dev, coef_z_dev, n, (int)nz);
  STEAM_LAUNCH_CHECK("deficit_kernel");
  return STEAM_OK;
}

extern "C" int64_t steam_project_workspace_bytes(int64_t nx, int64_t ny, int64_t nz) {
  return 16 * nz * 8 + level_scratch_bytes(nx * ny, (int)nz) + 256;
}

namespace {

// projection of f (+ mean) onto [lo, hi] preserving every level's window
// mean, levels with no violation untouched; mu_host[nz] in (replay) or out
int project(float* f, const float* mean_z, int64_t nx, int64_t ny, int64_t nz, const int win[4],
            double lo, double hi, double* mu_host, int replay, void* ws, hipStream_t s) {
  const int64_t n = nx * ny * nz;
  const double npts = (double)(win[1] - win[0]) * (double)(win[3] - win[2]);
  const double width = hi - lo;
  double* vec = (double*)ws;
  BaVec v;
  memset(&v, 0, sizeof(v));
  v.sum = vec;
  v.mu = vec + nz;
  v.mu_lo = vec + 2 * nz;
  v.mu_hi = vec + 3 * nz;
  v.mean_phi = vec + 4 * nz;
  double* active = vec + 5 * nz;
  double* vmin = vec + 6 * nz;
  double* vmax = vec + 7 * nz;
  double* rec = vec + 8 * nz;
  double* red = vec + 16 * nz;
  const int gz = (int)((nz + 255) / 256);
  const int gn = grid_1d(n);
  LevelArgs a;
  memset(&a, 0, sizeof(a));
  a.f = f;
  a.mean_z = mean_z;
  a.nx = (int)nx;
  a.ny = (int)ny;
  a.nz = (int)nz;
  memcpy(a.win, win, sizeof(a.win));
  a.lo = lo;
  a.hi = hi;
  int rc;
  auto reduce = [&](int op, double* out, const double* mu) {
    LevelArgs b = a;
    b.op = op;
    b.mu = mu;
    b.active = active;
    return level_reduce(b, out, red, s);
  };
  if (replay) {
    STEAM_HIP(hipMemcpyAsync(v.mu, mu_host, (size_t)nz * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(proj_active_kernel, dim3(gz), dim3(256), 0, s, v.mu, active, (int)nz);
  } else {
    if ((rc = reduce(LV_E_MIN, vmin, nullptr))) return rc;
    if ((rc = reduce(LV_E_MAX, vmax, nullptr))) return rc;
    hipLaunchKernelGGL(proj_init_kernel, dim3(gz), dim3(256), 0, s, v, vmin, vmax, active,
                       (int)nz, lo, hi, width);
    if ((rc = reduce(LV_E_SUM, v.sum, nullptr))) return rc;
    hipLaunchKernelGGL(ba_mean_kernel, dim3(gz), dim3(256), 0, s, v.sum, v.mean_phi, (int)nz,
                       npts);
    for (int it = 0; it < 60; ++it) {
      if ((rc = reduce(LV_E_CLAMP, v.sum, v.mu))) return rc;
      hipLaunchKernelGGL(ba_bisect_step_kernel, dim3(gz), dim3(256), 0, s, v, (int)nz, npts);
    }
    hipLaunchKernelGGL(proj_record_kernel, dim3(gz), dim3(256), 0, s, v.mu, active, rec,
                       (int)nz);
    STEAM_HIP(hipMemcpyAsync(mu_host, rec, (size_t)nz * 8, hipMemcpyDeviceToHost, s));
  }
  hipLaunchKernelGGL(proj_apply_kernel, dim3(gn), dim3(256), 0, s, f, mean_z, v.mu, active, n,
                     (int)nz, lo, hi);
  STEAM_LAUNCH_CHECK("projection");
  STEAM_HIP(hipStreamSynchronize(s));
  return STEAM_OK;
}

}  // namespace

extern "C" int steam_project_onto_bounds(float* pert_dev, const float* mean_z_dev, int64_t nx,
                                         int64_t ny, int64_t nz, const int64_t* window,
                                         double phi_min, double phi_max, double* mu, int replay,
                                         void* workspace_dev, int64_t workspace_bytes,
                                         steam_stream_t stream) {
  if (!pert_dev || !mean_z_dev || !mu || !workspace_dev || nx < 1 || ny < 1 || nz < 1)
    return set_error(STEAM_EINVAL, "steam_project_onto_bounds: bad arguments");
  if (!(phi_max > phi_min)) return set_error(STEAM_EINVAL, "steam_project_onto_bounds: empty bounds");
  if (workspace_bytes < steam_project_workspace_bytes(nx, ny, nz))
    return set_error(STEAM_EINVAL, "steam_project_onto_bounds: workspace too small");
  int win[4];
  int rc = check_window(window, (int)nx, (int)ny, win);
  if (rc) return rc;
  return project(pert_dev, mean_z_dev, nx, ny, nz, win, phi_min, phi_max, mu, replay,
                 workspace_dev, (hipStream_t)stream);
}

extern "C" int steam_compose_output(const float* state_dev, float* deficit_dev,
                                    const float* mean_z_dev, int64_t nx, int64_t ny, int64_t nz,
                                    const int64_t* window, double phi_min, double phi_max,
                                    double* mu, int replay, float* out_dev, void* workspace_dev,
                                    int64_t workspace_bytes, steam_stream_t stream) {
  if (!state_dev || !mean_z_dev || !out_dev || nx < 1 || ny < 1 || nz < 1)
    return set_error(STEAM_EINVAL, "steam_compose_output: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = nx * ny * nz;
  const float lo = (float)phi_min, hi = (float)phi_max;
  if (!deficit_dev) {
    hipLaunchKernelGGL(compose_kernel, dim3(grid_1d(n)), dim3(256), 0, s, state_dev, mean_z_dev,
                       out_dev, n, (int)nz, lo, hi);
    STEAM_LAUNCH_CHECK("compose_kernel");
    return STEAM_OK;
  }
  if (!mu || !workspace_dev || workspace_bytes < steam_project_workspace_bytes(nx, ny, nz))
    return set_error(STEAM_EINVAL, "steam_compose_output: projection needs mu and a workspace");
  if (!(phi_max > phi_min)) return set_error(STEAM_EINVAL, "steam_compose_output: empty bounds");
  int win[4];
  int rc = check_window(window, (int)nx, (int)ny, win);
  if (rc) return rc;
  hipLaunchKernelGGL(add_state_kernel, dim3(grid_1d(n)), dim3(256), 0, s, deficit_dev, state_dev,
                     n);
  STEAM_LAUNCH_CHECK("add_state_kernel");
  if ((rc = project(deficit_dev, mean_z_dev, nx, ny, nz, win, phi_min, phi_max, mu, replay,
                    workspace_dev, s)))
    return rc;
  hipLaunchKernelGGL(compose_kernel, dim3(grid_1d(n)), dim3(256), 0, s, deficit_dev, mean_z_dev,
                     out_dev, n, (int)nz, lo, hi);
  STEAM_LAUNCH_CHECK("compose_kernel");
  return STEAM_OK;
}

extern "C" int64_t steam_compose_flux_workspace_bytes(int64_t nx, int64_t ny, int64_t nz) {
  return (nz + 8) * 8 + level_scratch_bytes(nx * ny, (int)nz) + 256;
}

extern "C" int steam_compose_flux_output(const float* state_dev, float* deficit_dev, int64_t nx,
                                         int64_t ny, int64_t nz, const int64_t* window,
                                         double* sums, int replay, void* workspace_dev,
                                         int64_t workspace_bytes, steam_stream_t stream) {
  if (!state_dev || !deficit_dev || !sums || !workspace_dev || nx < 1 || ny < 1 || nz < 1)
    return set_error(STEAM_EINVAL, "steam_compose_flux_output: bad arguments");
  if (workspace_bytes < steam_compose_flux_workspace_bytes(nx, ny, nz))
    return set_error(STEAM_EINVAL, "steam_compose_flux_output: workspace too small");
  hipStream_t s = (hipStream_t)stream;
  int win[4];
  int rc = check_window(window, (int)nx, (int)ny, win);
  if (rc) return rc;
  const int64_t n = nx * ny * nz;
  const double count = (double)(win[1] - win[0]) * (double)(win[3] - win[2]) * (double)nz;
  double* vec = (double*)workspace_dev;
  double* lv = vec;            // per-level sums
  double* tot = vec + nz;      // [0] entering total, [1] realized total
  double* means = tot + 2;     // [0] entering mean, [1] realized mean
  double* red = vec + nz + 8;
  LevelArgs a;
  memset(&a, 0, sizeof(a));
  a.nx = (int)nx;
  a.ny = (int)ny;
  a.nz = (int)nz;
  a.op = LV_SUM;
  memcpy(a.win, win, sizeof(a.win));
  if (replay) {
    if (!(sums[1] > 0.0) || !(sums[3] > 0.0))
      return set_error(STEAM_EINVAL, "steam_compose_flux_output: recorded counts must be positive");
    double m[2] = {sums[0] / sums[1], sums[2] / sums[3]};
    STEAM_HIP(hipMemcpyAsync(means, m, 16, hipMemcpyHostToDevice, s));
  } else {
    a.f = state_dev;
    if ((rc = level_reduce(a, lv, red, s))) return rc;
    hipLaunchKernelGGL(sum_levels_kernel, dim3(1), dim3(64), 0, s, lv, (int)nz, count, tot,
                       means);
  }
  hipLaunchKernelGGL(flux_compose_kernel, dim3(grid_1d(n)), dim3(256), 0, s, deficit_dev,
                     state_dev, n);
  if (!replay) {
    a.f = deficit_dev;
    if ((rc = level_reduce(a, lv, red, s))) return rc;
    hipLaunchKernelGGL(sum_levels_kernel, dim3(1), dim3(64), 0, s, lv, (int)nz, count,
                       tot + 1, means + 1);
  }
  hipLaunchKernelGGL(flux_restore_kernel, dim3(grid_1d(n)), dim3(256), 0, s, deficit_dev, means,
                     n);
  STEAM_LAUNCH_CHECK("flux composition");
  if (!replay) {
    double t[2];
    STEAM_HIP(hipMemcpyAsync(t, tot, 16, hipMemcpyDeviceToHost, s));
    STEAM_HIP(hipStreamSynchronize(s));
    sums[0] = t[0];
    sums[1] = count;
    sums[2] = t[1];
    sums[3] = count;
  }
  return STEAM_OK;
}
