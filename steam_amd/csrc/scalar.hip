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

// Position of a grid-stride loop over a (lines, nz) field: the element i,
// its level z and line, and the line's (x, y) when ny is given, advanced by
// additions instead of a 64-bit division per element.
struct ZWalk {
  int64_t i, line, dline;
  int z, dz, nz, x, y, ny, dx_, dy_;
  __device__ ZWalk(int64_t start, int64_t stride, int nz_, int ny_ = 0) : nz(nz_), ny(ny_) {
    i = start;
    line = start / nz;
    z = (int)(start - line * nz);
    dline = stride / nz;
    dz = (int)(stride - dline * nz);
    if (ny > 0) {
      x = (int)(line / ny);
      y = (int)(line - (int64_t)x * ny);
      dx_ = (int)(dline / ny);
      dy_ = (int)(dline - (int64_t)dx_ * ny);
    } else {
      x = y = dx_ = dy_ = 0;
    }
  }
  __device__ __forceinline__ void next(int64_t stride) {
    i += stride;
    int carry = 0;
    z += dz;
    if (z >= nz) {
      z -= nz;
      carry = 1;
    }
    line += dline + carry;
    if (ny > 0) {
      y += dy_ + carry;
      x += dx_;
      if (y >= ny) {
        y -= ny;
        ++x;
      }
    }
  }
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
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (ZWalk zw((int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride, nz, ny); zw.i < n;
       zw.next(stride)) {
    const int64_t i = zw.i;
    const int z = zw.z;
    const int x = zw.x, y = zw.y;
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

// The same joint pattern without the cancellation of the float32 running
// field: phi = pert + mean(z) is never formed.  Horizontal differences are
// differences of pert (the mean is horizontally uniform); the vertical one is
// the difference of pert plus the mean's own z-derivative dmean[z] (formed in
// float64 by the caller); the taper's distances to the bounds are
// pert + (mean - lo) and (hi - mean) - pert with both level terms formed in
// float64.  Used by the precision workload; agrees with the float64
// transform to the rounding of each difference rather than of the field.
__global__ __launch_bounds__(256) void pattern_precise_kernel(
    const float* __restrict__ pert, const float* __restrict__ dmean_z,
    const float* __restrict__ mlo_z, const float* __restrict__ mhi_z,
    const float* __restrict__ sk, const float* __restrict__ b_z,
    const float* __restrict__ aspect_z, ZDiff zd, int nx, int ny, int nz, float two_dx,
    float two_dy, float* __restrict__ out) {
  const int64_t n = (int64_t)nx * ny * nz;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (ZWalk zw((int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride, nz, ny); zw.i < n;
       zw.next(stride)) {
    const int64_t i = zw.i;
    const int z = zw.z;
    const int x = zw.x, y = zw.y;
    const int xp = x + 1 < nx ? x + 1 : 0, xm = x > 0 ? x - 1 : nx - 1;
    const int yp = y + 1 < ny ? y + 1 : 0, ym = y > 0 ? y - 1 : ny - 1;
    auto p = [&](int xx, int yy, int zz) { return pert[((int64_t)xx * ny + yy) * nz + zz]; };
    const float gx = __fdiv_rn(__fsub_rn(p(xp, y, z), p(xm, y, z)), two_dx);
    const float gy = __fdiv_rn(__fsub_rn(p(x, yp, z), p(x, ym, z)), two_dy);
    float w = sqrtf(__fadd_rn(__fmul_rn(gx, gx), __fmul_rn(gy, gy)));
    const float pc = p(x, y, z);
    if (nz > 1) {
      float gz;
      if (z == 0) {
        gz = __fdiv_rn(__fsub_rn(p(x, y, 1), pc), zd.dz_first);
      } else if (z == nz - 1) {
        gz = __fdiv_rn(__fsub_rn(pc, p(x, y, nz - 2)), zd.dz_last);
      } else if (zd.uniform) {
        gz = __fdiv_rn(__fsub_rn(p(x, y, z + 1), p(x, y, z - 1)), zd.two_dz);
      } else {
        gz = __fadd_rn(__fadd_rn(__fmul_rn(zd.a[z - 1], p(x, y, z - 1)),
                                 __fmul_rn(zd.b[z - 1], pc)),
                       __fmul_rn(zd.c[z - 1], p(x, y, z + 1)));
      }
      w = __fadd_rn(__fmul_rn(fabsf(__fadd_rn(gz, dmean_z[z])), aspect_z[z]), w);
    }
    w = __fmul_rn(w, sk[i]);
    float g = fminf(__fadd_rn(pc, mlo_z[z]), __fsub_rn(mhi_z[z], pc));
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
  const int64_t stride_ = (int64_t)gridDim.x * blockDim.x;
  for (ZWalk zw((int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride_, nz); zw.i < n;
       zw.next(stride_)) {
    const int64_t i = zw.i;
    const int z = zw.z;
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
  // the bounded add's rescale step, fused (the field is the increment d)
  LV_SCALED_SUM = 11,  // sum f32(d * f32(scale[z]))
  LV_APPLY_ABS = 12,   // d = clamp(f32(d * f32(scale[z])) - f32(shift[z]) + phi, lo, hi) - phi
                       // over the whole field, written back; sum |d| over the window
};

// The window's lines (x, y) are split into G contiguous runs of
// lines_per_group, G a function of the window alone; workgroup g reduces run
// g into one partial per (candidate, level), and one combine pass joins the
// G partials of each (candidate, level) in a fixed order.  A level's result therefore depends
// only on the field, the window and the shift, never on how many candidate
// shifts share the launch.
#ifndef STEAM_RD_MIN_LINES
#define STEAM_RD_MIN_LINES 128
#endif
constexpr int RD_MAX_GROUPS = 4096;
constexpr int RD_MIN_LINES = STEAM_RD_MIN_LINES;
// lines whose loads a lane issues before it works on them
constexpr int RD_U = 4;
// candidate shifts of one bisection pass: the 2^D - 1 midpoints of a D-deep
// bisection tree, so one pass over the field takes D bisection steps; D = 4
// (15 candidates) where passes are latency-bound, D = 3 (7 candidates) on
// large fields, where the candidates' arithmetic would bound the pass
constexpr int KCAND = 15;
constexpr int KCAND_WIDE = 7;
constexpr int KCAND_NARROW = 3;
// fields above this many cells take the 3-deep tree
#ifndef STEAM_KWIDE_LOG2
#define STEAM_KWIDE_LOG2 20
#endif
constexpr int64_t KWIDE_CELLS = int64_t(1) << STEAM_KWIDE_LOG2;
// above 2^STEAM_KNARROW_LOG2 cells the tree is 2 deep (3 candidates, 30
// passes).  Off by default: at 2048x2048x115 it took the bounded add from
// 18.2 to 19.1 ms and the projection from 20.7 to 22.2 ms (the 20 passes of
// 7 candidates cost less than 30 of 3)
#ifndef STEAM_KNARROW_LOG2
#define STEAM_KNARROW_LOG2 40
#endif
constexpr int64_t KNARROW_CELLS = int64_t(1) << STEAM_KNARROW_LOG2;

struct LevelArgs {
  const float* f;
  const float* mean_z;
  const double* mu;      // [ncand][nz] shifts (clamp ops)
  const double* active;  // LV_E_CLAMP: levels to reduce (others give 0)
  const int* stop;       // when set and nonzero the launch does nothing
  const float* pert;     // LV_APPLY_ABS: the perturbation (phi = pert + mean)
  float* out;            // LV_APPLY_ABS: where d is written (may be f)
  const double* scale;   // LV_SCALED_SUM, LV_APPLY_ABS: pending per-level scale
  const double* shift;   // LV_APPLY_ABS: per-level demean
  int nx, ny, nz;
  int win[4];            // lines visited
  int awin[4];           // LV_APPLY_ABS: lines whose |d| is summed
  int op;
  int ncand;
  double lo, hi;
};

__device__ __forceinline__ bool lv_is_min(int op) { return op == LV_MIN || op == LV_E_MIN; }
__device__ __forceinline__ bool lv_is_max(int op) { return op == LV_MAX || op == LV_E_MAX; }
__device__ __forceinline__ double lv_init(int op) {
  return lv_is_min(op) ? INFINITY : (lv_is_max(op) ? -INFINITY : 0.0);
}
__device__ __forceinline__ double lv_join(int op, double a, double b) {
  return lv_is_min(op) ? fmin(a, b) : (lv_is_max(op) ? fmax(a, b) : a + b);
}

// Workgroup g reduces window lines [g*lpg, (g+1)*lpg).  Levels are taken in
// chunks of up to 256; within a chunk of `sub` levels, lane t owns level
// zc + t % sub and walks every (256/sub)-th line of the run, so a wave reads
// consecutive addresses of consecutive lines.  Lanes sharing a level are then
// summed through LDS in lane order.  OP is the LevelOp; NC > 1 only for the
// two clamp ops (candidate shifts).
template <int NC, int OP>
__global__ __launch_bounds__(256) void level_partial_kernel(LevelArgs a, double* __restrict__ part,
                                                            int64_t lpg) {
  constexpr bool kMin = OP == LV_MIN || OP == LV_E_MIN;
  constexpr bool kMax = OP == LV_MAX || OP == LV_E_MAX;
  constexpr double kInit = kMin ? INFINITY : (kMax ? -INFINITY : 0.0);
  if (a.stop && *a.stop) return;
  __shared__ double red[NC][256];
  const int t = threadIdx.x;
  const int wx0 = a.win[0], wy0 = a.win[2];
  const int wy = a.win[3] - a.win[2];
  const int64_t wlines = (int64_t)(a.win[1] - a.win[0]) * wy;
  const int64_t l0 = (int64_t)blockIdx.x * lpg;
  const int64_t l1 = l0 + lpg < wlines ? l0 + lpg : wlines;
  const int nc = NC == 1 ? 1 : a.ncand;
  const float flo = (float)a.lo, fhi = (float)a.hi;
  for (int zc = 0; zc < a.nz; zc += 256) {
    const int sub = a.nz - zc < 256 ? a.nz - zc : 256;
    const int lpi = 256 / sub;
    const bool on = t < lpi * sub;
    const int z = zc + (on ? t % sub : 0);
    const int first = on ? t / sub : 0;
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = kInit;
    bool skip = false;
    if constexpr (OP == LV_E_CLAMP) skip = a.active && a.active[z] == 0.0;
    if (on && !skip && l0 + first < l1) {
      const float mz = a.mean_z ? a.mean_z[z] : 0.f;
      float muf[NC];
      double mu64[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        mu64[c] = (a.mu && c < nc) ? a.mu[(int64_t)c * a.nz + z] : 0.0;
        muf[c] = (float)mu64[c];
      }
      float sc = 1.f, sh = 0.f;
      if constexpr (OP == LV_SCALED_SUM || OP == LV_APPLY_ABS) sc = (float)a.scale[z];
      if constexpr (OP == LV_APPLY_ABS) sh = (float)a.shift[z];
      // the element work, in line order (loads are issued RD_U lines ahead)
      auto step = [&](int64_t idx, float v, float pv, int x, int y) {
        if constexpr (OP == LV_SCALED_SUM) acc[0] += (double)__fmul_rn(v, sc);
        if constexpr (OP == LV_APPLY_ABS) {
          const float phi = __fadd_rn(pv, mz);
          const float u = __fsub_rn(__fmul_rn(v, sc), sh);
          const float e = fminf(fmaxf(__fadd_rn(u, phi), flo), fhi);
          const float d = __fsub_rn(e, phi);
          a.out[idx] = d;
          if (x >= a.awin[0] && x < a.awin[1] && y >= a.awin[2] && y < a.awin[3])
            acc[0] += (double)fabsf(d);
        }
        if constexpr (OP == LV_SUM) acc[0] += (double)v;
        if constexpr (OP == LV_ABS) acc[0] += (double)fabsf(v);
        if constexpr (OP == LV_NONZERO) acc[0] += v != 0.f ? 1.0 : 0.0;
        if constexpr (OP == LV_PHI) acc[0] += (double)__fadd_rn(v, mz);
        if constexpr (OP == LV_MIN) acc[0] = fmin(acc[0], (double)__fadd_rn(v, mz));
        if constexpr (OP == LV_MAX) acc[0] = fmax(acc[0], (double)__fadd_rn(v, mz));
        if constexpr (OP == LV_E_SUM) acc[0] += (double)v + (double)mz;
        if constexpr (OP == LV_E_MIN) acc[0] = fmin(acc[0], (double)v + (double)mz);
        if constexpr (OP == LV_E_MAX) acc[0] = fmax(acc[0], (double)v + (double)mz);
        // every candidate slot is evaluated (slots past nc hold shift 0 and
        // are never stored), so the loop body is straight-line code
        if constexpr (OP == LV_CLAMP) {
          // med3 is the clamp for finite values and lo <= hi
#pragma unroll
          for (int c = 0; c < NC; ++c)
            acc[c] += (double)__builtin_amdgcn_fmed3f(__fsub_rn(v, muf[c]), flo, fhi);
        }
        if constexpr (OP == LV_E_CLAMP) {
          const double e = (double)v + (double)mz;
#pragma unroll
          for (int c = 0; c < NC; ++c) acc[c] += fmin(fmax(e - mu64[c], a.lo), a.hi);
        }
      };
      int64_t wl = l0 + first;
      int x = wx0 + (int)(wl / wy);
      int y = wy0 + (int)(wl - (int64_t)(x - wx0) * wy);
      const int wy1 = wy0 + wy;
      while (wl < l1) {
        int64_t idx[RD_U];
        int xs[RD_U], ys[RD_U];
        float v[RD_U], pv[RD_U];
#pragma unroll
        for (int u = 0; u < RD_U; ++u) {
          const bool ok = wl + (int64_t)u * lpi < l1;
          xs[u] = x;
          ys[u] = y;
          idx[u] = ((int64_t)x * a.ny + y) * a.nz + z;
          v[u] = ok ? a.f[idx[u]] : 0.f;
          pv[u] = 0.f;
          if constexpr (OP == LV_APPLY_ABS) pv[u] = ok ? a.pert[idx[u]] : 0.f;
          y += lpi;
          if (y >= wy1) {
            const int k = (y - wy0) / wy;
            x += k;
            y -= k * wy;
          }
        }
#pragma unroll
        for (int u = 0; u < RD_U; ++u)
          if (wl + (int64_t)u * lpi < l1) step(idx[u], v[u], pv[u], xs[u], ys[u]);
        wl += (int64_t)RD_U * lpi;
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) red[c][t] = acc[c];
    __syncthreads();
    // one thread per (candidate, level) joins that level's lpi lanes in lane order
    for (int p = t; p < nc * sub; p += 256) {
      const int c = p / sub, zl = p - c * sub;
      double s = red[c][zl];
      for (int j = 1; j < lpi; ++j) {
        const double o = red[c][zl + j * sub];
        s = kMin ? fmin(s, o) : (kMax ? fmax(s, o) : s + o);
      }
      part[((int64_t)c * a.nz + zc + zl) * gridDim.x + blockIdx.x] = s;
    }
    __syncthreads();
  }
}

// out[c][z] = (the G partials of (c, z), part[(c*nz + z)*G + g]) joined in
// a fixed order, divided by div when div > 0: one wave per (c, z), lane l
// takes runs l, l+64, ... in order, then the lanes are joined by a fixed
// butterfly
__global__ __launch_bounds__(256) void level_combine_kernel(const double* __restrict__ part,
                                                            int groups, int nc, int nz, int op,
                                                            double div, double* __restrict__ out,
                                                            const int* stop) {
  if (stop && *stop) return;
  const int lane = threadIdx.x & 63;
  const int64_t pair = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (pair >= (int64_t)nc * nz) return;
  const double* row = part + pair * groups;
  double acc = lv_init(op);
#pragma unroll 8
  for (int g = lane; g < groups; g += 64) acc = lv_join(op, acc, row[g]);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) acc = lv_join(op, acc, __shfl_xor(acc, off, 64));
  if (lane == 0) out[pair] = div > 0.0 ? acc / div : acc;
}

// ---------------------------------------------------------------------------
// bisection on the level means, D steps per pass over the field
//
// A bisection step halves the bracket [lo, hi] at mid = 0.5 * (lo + hi),
// moving lo up to mid when the level mean at mid is "positive" and hi down
// otherwise.  The 2^D - 1 candidates of a pass are the midpoints of the
// complete D-deep tree of brackets below the current one (heap order), each
// computed by that same formula; after the pass the D steps are taken along
// the tree with the measured sums.  Every step therefore sees exactly the
// midpoint and the level sum a one-step-per-pass bisection would, and 60
// steps end on the same bracket whatever D is.
// ---------------------------------------------------------------------------
struct BisectVec {
  double* lo;
  double* hi;
  double* mu;      // 0.5 * (lo + hi), the result after the last pass
  double* cand;    // [2^D - 1][nz] the next pass's candidates
  const double* target;
  const double* sums;  // [2^D - 1][nz] level sums at the candidates
};

template <int D>
__device__ __forceinline__ void bisect_candidates(BisectVec v, int nz, int z, double lo, double hi) {
  constexpr int NC = (1 << D) - 1;
  double a[NC + 1], b[NC + 1];
  a[1] = lo;
  b[1] = hi;
#pragma unroll
  for (int n = 1; n <= NC; ++n) {
    const double m = 0.5 * (a[n] + b[n]);
    v.cand[(int64_t)(n - 1) * nz + z] = m;
    if (2 * n + 1 <= NC) {
      a[2 * n] = a[n];
      b[2 * n] = m;
      a[2 * n + 1] = m;
      b[2 * n + 1] = b[n];
    }
  }
}

template <int D>
__global__ void bisect_init_kernel(BisectVec v, int nz, double width) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z < nz) {
    v.lo[z] = -width;
    v.hi[z] = width;
    v.mu[z] = 0.5 * (v.lo[z] + v.hi[z]);
    bisect_candidates<D>(v, nz, z, -width, width);
  }
}

// The combine of one level's 2^D - 1 candidate sums (wave c joins candidate
// c's run partials exactly as level_combine_kernel does) and the D bisection
// steps of that level, in one launch: one workgroup per level.
// mode 0: positive = sum / npts - target > 0 (the bounded add);
// mode 1: positive = sum / npts > target (the projection)
template <int D>
__global__ __launch_bounds__(1024) void bisect_combine_kernel(const double* __restrict__ part,
                                                              int groups, int nz, int op,
                                                              BisectVec v, double npts,
                                                              int mode) {
  constexpr int NC = (1 << D) - 1;
  __shared__ double sums[NC];
  const int z = blockIdx.x;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  {
    const double* row = part + ((int64_t)w * nz + z) * groups;
    double acc = lv_init(op);
#pragma unroll 8
    for (int g = lane; g < groups; g += 64) acc = lv_join(op, acc, row[g]);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc = lv_join(op, acc, __shfl_xor(acc, off, 64));
    if (lane == 0) sums[w] = acc;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double lo = v.lo[z], hi = v.hi[z];
  const double target = v.target[z];
  int n = 1;
  for (int s = 0; s < D; ++s) {
    const double mid = 0.5 * (lo + hi);
    const double mean = sums[n - 1] / npts;
    const bool positive = mode == 0 ? (mean - target > 0.0) : (mean > target);
    if (positive) {
      lo = mid;
      n = 2 * n + 1;
    } else {
      hi = mid;
      n = 2 * n;
    }
  }
  v.lo[z] = lo;
  v.hi[z] = hi;
  v.mu[z] = 0.5 * (lo + hi);
  bisect_candidates<D>(v, nz, z, lo, hi);
}

// ---------------------------------------------------------------------------
// bounded amplitude-preserving add, device pieces (the algorithm of
// steam/simulate.py:2985-3111, the reference's GPU path).  The rescale loop
// runs without host round trips: a device flag `stop` is raised by the step
// whose levels are all done, and every later launch of the loop returns at
// once, so the loop's launches are the same whatever the data.
// ---------------------------------------------------------------------------
// d = clamp(d - f32(shift[z*stride]) + phi, lo, hi) - phi  (shift may be NULL)
__global__ __launch_bounds__(256) void ba_shift_clip_kernel(
    float* __restrict__ d, const float* __restrict__ pert, const float* __restrict__ mean_z,
    const double* __restrict__ shift, int stride, int64_t n, int nz, float lo, float hi,
    const int* stop) {
  if (stop && *stop) return;
  const int64_t stride_ = (int64_t)gridDim.x * blockDim.x;
  for (ZWalk zw((int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride_, nz); zw.i < n;
       zw.next(stride_)) {
    const int64_t i = zw.i;
    const int z = zw.z;
    const float phi = __fadd_rn(pert[i], mean_z[z]);
    float v = d[i];
    if (shift) v = __fsub_rn(v, (float)shift[(int64_t)z * stride]);
    const float e = fminf(fmaxf(__fadd_rn(v, phi), lo), hi);
    d[i] = __fsub_rn(e, phi);
  }
}

// d *= f32(scale[z*stride])
__global__ __launch_bounds__(256) void ba_scale_kernel(float* __restrict__ d,
                                                       const double* __restrict__ scale,
                                                       int stride, int64_t n, int nz,
                                                       const int* stop) {
  if (stop && *stop) return;
  const int64_t stride_ = (int64_t)gridDim.x * blockDim.x;
  for (ZWalk zw((int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride_, nz); zw.i < n;
       zw.next(stride_))
    d[zw.i] = __fmul_rn(d[zw.i], (float)scale[(int64_t)zw.z * stride]);
}

// e = d + phi (in place on d)
__global__ __launch_bounds__(256) void ba_to_field_kernel(float* __restrict__ d,
                                                          const float* __restrict__ pert,
                                                          const float* __restrict__ mean_z,
                                                          int64_t n, int nz) {
  const int64_t stride_ = (int64_t)gridDim.x * blockDim.x;
  for (ZWalk zw((int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride_, nz); zw.i < n;
       zw.next(stride_)) {
    const int64_t i = zw.i;
    const int z = zw.z;
    d[i] = __fadd_rn(d[i], __fadd_rn(pert[i], mean_z[z]));
  }
}

// final write: v = clamp(e - f32(mu[z]), lo, hi) - mean[z]; with record,
// inc = v - pert; pert = v
__global__ __launch_bounds__(256) void ba_finish_kernel(const float* e, float* __restrict__ pert,
                                                        float* inc,
                                                        const float* __restrict__ mean_z,
                                                        const double* __restrict__ mu,
                                                        int64_t n, int nz, float lo, float hi,
                                                        int record) {
  const int64_t stride_ = (int64_t)gridDim.x * blockDim.x;
  for (ZWalk zw((int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride_, nz); zw.i < n;
       zw.next(stride_)) {
    const int64_t i = zw.i;
    const int z = zw.z;
    float v = fminf(fmaxf(__fsub_rn(e[i], (float)mu[z]), lo), hi);
    v = __fsub_rn(v, mean_z[z]);
    if (record) inc[i] = __fsub_rn(v, pert[i]);
    pert[i] = v;
  }
}

// per-level state of the rescale loop
struct BaVec {
  double* a0;
  double* done;          // 1.0 / 0.0
  double* mean;          // last level mean (of d or |d|)
  double* shift;         // this step's demean
  double* scale;         // this step's applied scale
  double* hist_demean;   // [nz][n_rescale]
  double* hist_scale;    // [nz][n_rescale]
  double* pend;          // scale still to be applied to d (1 when none)
  int* flags;            // [0] stop, [1] n_loop
};

// done = a0 <= 0; history cleared (demean 0, scale 1); stop = 0, n_loop = 0
__global__ void ba_init_kernel(BaVec v, int nz, int n_rescale) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z < nz) {
    v.done[z] = v.a0[z] <= 0.0 ? 1.0 : 0.0;
    v.pend[z] = 1.0;
    for (int s = 0; s < n_rescale; ++s) {
      v.hist_demean[(int64_t)z * n_rescale + s] = 0.0;
      v.hist_scale[(int64_t)z * n_rescale + s] = 1.0;
    }
  }
  if (z == 0) {
    v.flags[0] = 0;
    v.flags[1] = 0;
  }
}
// One level's run partials joined exactly as level_combine_kernel joins
// them (lane l takes runs l, l + 64, ... in order, then a fixed butterfly).
__device__ __forceinline__ double join_level_runs(const double* __restrict__ row, int groups,
                                                  int op, int lane) {
  double acc = lv_init(op);
#pragma unroll 8
  for (int g = lane; g < groups; g += 64) acc = lv_join(op, acc, row[g]);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) acc = lv_join(op, acc, __shfl_xor(acc, off, 64));
  return acc;
}

// A rescale step's first combine (the level mean of the pending-scaled d)
// and its demean: shift = done ? 0 : level mean, recorded for the step.
// One wave per level.
__global__ __launch_bounds__(256) void ba_combine_demean_kernel(const double* __restrict__ part,
                                                                int groups, int nz, double div,
                                                                BaVec v, int step,
                                                                int n_rescale) {
  if (v.flags[0]) return;
  const int lane = threadIdx.x & 63;
  const int z = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (z >= nz) return;
  const double acc = join_level_runs(part + (int64_t)z * groups, groups, LV_SCALED_SUM, lane);
  if (lane == 0) {
    const double mean = acc / div;
    v.mean[z] = mean;
    const double dm = v.done[z] != 0.0 ? 0.0 : mean;
    v.shift[z] = dm;
    v.hist_demean[(int64_t)z * n_rescale + step] = dm;
  }
}

// A rescale step's second combine (m_abs, the level mean of |d| after the
// scale, demean and clip), then in the same workgroup: scale =
// min(a0 / max(m_abs, 1e-300), 2); done |= m_abs <= 0 | |scale - 1| < 1e-4;
// applied = done ? 1 : scale.  When every level is done the stop flag is
// raised (the step's scale is then neither applied nor recorded), else the
// scale is recorded and left pending for the next pass over d; n_loop =
// step + 1 either way.
constexpr int BA_CK_NTH = 1024;
__global__ __launch_bounds__(BA_CK_NTH) void ba_combine_rescale_check_kernel(
    const double* __restrict__ part, int groups, int nz, double div, BaVec v, int step,
    int n_rescale) {
  if (v.flags[0]) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int z = w; z < nz; z += BA_CK_NTH / 64) {
    const double acc = join_level_runs(part + (int64_t)z * groups, groups, LV_APPLY_ABS, lane);
    if (lane == 0) {
      const double m_abs = acc / div;
      v.mean[z] = m_abs;
      const double sc = fmin(v.a0[z] / fmax(m_abs, 1e-300), 2.0);
      const bool d = v.done[z] != 0.0 || m_abs <= 0.0 || fabs(sc - 1.0) < 1e-4;
      v.done[z] = d ? 1.0 : 0.0;
      v.scale[z] = d ? 1.0 : sc;
    }
  }
  __shared__ int all_done;
  if (threadIdx.x == 0) all_done = 1;
  __syncthreads();
  for (int z = threadIdx.x; z < nz; z += BA_CK_NTH)
    if (v.done[z] == 0.0) all_done = 0;
  __syncthreads();
  if (all_done) {
    for (int z = threadIdx.x; z < nz; z += BA_CK_NTH) v.pend[z] = 1.0;
    if (threadIdx.x == 0) {
      v.flags[1] = step + 1;
      v.flags[0] = 1;
    }
    return;
  }
  for (int z = threadIdx.x; z < nz; z += BA_CK_NTH) {
    v.hist_scale[(int64_t)z * n_rescale + step] = v.scale[z];
    v.pend[z] = v.scale[z];
  }
  if (threadIdx.x == 0) v.flags[1] = step + 1;
}

// the final demean and clip, then e = d + phi, in one pass:
// d' = clamp(f32(d * f32(pend)) - f32(shift) + phi, lo, hi) - phi; d = d' + phi
// the loop counts of a device ledger: every level took the same number of
// rescale steps, flags[1]
__global__ __launch_bounds__(256) void ba_counts_kernel(const int* flags, int32_t* n_loop,
                                                         int32_t* n_scale, int nz) {
  for (int z = blockIdx.x * 256 + threadIdx.x; z < nz; z += gridDim.x * 256) {
    n_loop[z] = flags[1];
    n_scale[z] = flags[1];
  }
}

__global__ __launch_bounds__(256) void ba_final_kernel(float* __restrict__ d,
                                                       const float* __restrict__ pert,
                                                       const float* __restrict__ mean_z,
                                                       const double* __restrict__ pend,
                                                       const double* __restrict__ shift,
                                                       int64_t n, int nz, float lo, float hi) {
  const int64_t stride_ = (int64_t)gridDim.x * blockDim.x;
  for (ZWalk zw((int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride_, nz); zw.i < n;
       zw.next(stride_)) {
    const int64_t i = zw.i;
    const int z = zw.z;
    const float phi = __fadd_rn(pert[i], mean_z[z]);
    const float u = __fsub_rn(__fmul_rn(d[i], (float)pend[z]), (float)shift[z]);
    const float e = fminf(fmaxf(__fadd_rn(u, phi), lo), hi);
    d[i] = __fadd_rn(__fsub_rn(e, phi), phi);
  }
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
  const int64_t stride_ = (int64_t)gridDim.x * blockDim.x;
  for (ZWalk zw((int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride_, nz); zw.i < n;
       zw.next(stride_)) {
    const int64_t i = zw.i;
    const int z = zw.z;
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
  const int64_t stride_ = (int64_t)gridDim.x * blockDim.x;
  for (ZWalk zw((int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride_, nz); zw.i < n;
       zw.next(stride_)) {
    const int64_t i = zw.i;
    const int z = zw.z;
    out[i] = fminf(fmaxf(__fadd_rn(src[i], mean_z[z]), lo), hi);
  }
}

// projection onto the bounds: active[z] = the level has a cell out of bounds
__global__ void proj_active_kernel(const double* vmin, const double* vmax, double* active, int nz,
                                   double lo, double hi) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z < nz) active[z] = (vmin[z] < lo || vmax[z] > hi) ? 1.0 : 0.0;
}
// recorded mu: NaN on levels left alone
__global__ void proj_record_kernel(const double* mu, const double* active, double* out, int nz) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z < nz) out[z] = active[z] != 0.0 ? mu[z] : __builtin_nan("");
}
// replay: active where the recorded mu is a number
__global__ void proj_replay_active_kernel(const double* mu, double* active, int nz) {
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
  const int64_t stride_ = (int64_t)gridDim.x * blockDim.x;
  for (ZWalk zw((int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride_, nz); zw.i < n;
       zw.next(stride_)) {
    const int64_t i = zw.i;
    const int z = zw.z;
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

// lines per workgroup run for a window of wlines lines, and the run count
// The runs depend on the window alone, not on the candidate count, so a
// many-candidate pass sums each candidate in the order a one-candidate pass
// does.
int64_t rd_lines_per_group(int64_t wlines) {
  int64_t g = (wlines + RD_MIN_LINES - 1) / RD_MIN_LINES;
  g = std::min<int64_t>(std::max<int64_t>(g, 1), RD_MAX_GROUPS);
  return (wlines + g - 1) / g;
}
int64_t rd_groups(int64_t wlines) {
  const int64_t lpg = rd_lines_per_group(wlines);
  return (wlines + lpg - 1) / lpg;
}

// bytes of reduction scratch for windows of up to nlines lines, nz levels
// and nc candidate shifts per launch (the run count of any window of at
// most nlines lines is at most this bound)
int64_t level_scratch_bytes(int64_t nlines, int64_t nz, int nc) {
  const int64_t g = std::min<int64_t>(
      std::max<int64_t>((nlines + RD_MIN_LINES - 1) / RD_MIN_LINES, 1), RD_MAX_GROUPS);
  return g * nc * nz * 8;
}

// out[c][z] = per-level reduction (divided by div when div > 0) of a's op
// over its window, for each of a.ncand shifts
// the partial pass of a level reduction: per run of lines, one partial per
// (candidate, level) in scratch[(c * nz + z) * groups + run]
int level_partials(const LevelArgs& a, double* scratch, hipStream_t s, int64_t* groups_out) {
  const int64_t wlines = (int64_t)(a.win[1] - a.win[0]) * (a.win[3] - a.win[2]);
  const int64_t lpg = rd_lines_per_group(wlines);
  const int64_t groups = rd_groups(wlines);
  *groups_out = groups;
  const dim3 grid((unsigned)groups), block(256);
  if (a.ncand > 1) {
    if (a.ncand == KCAND && a.op == LV_CLAMP)
      hipLaunchKernelGGL((level_partial_kernel<KCAND, LV_CLAMP>), grid, block, 0, s, a, scratch,
                         lpg);
    else if (a.ncand == KCAND && a.op == LV_E_CLAMP)
      hipLaunchKernelGGL((level_partial_kernel<KCAND, LV_E_CLAMP>), grid, block, 0, s, a, scratch,
                         lpg);
    else if (a.ncand == KCAND_WIDE && a.op == LV_CLAMP)
      hipLaunchKernelGGL((level_partial_kernel<KCAND_WIDE, LV_CLAMP>), grid, block, 0, s, a,
                         scratch, lpg);
    else if (a.ncand == KCAND_WIDE && a.op == LV_E_CLAMP)
      hipLaunchKernelGGL((level_partial_kernel<KCAND_WIDE, LV_E_CLAMP>), grid, block, 0, s, a,
                         scratch, lpg);
    else if (a.ncand == KCAND_NARROW && a.op == LV_CLAMP)
      hipLaunchKernelGGL((level_partial_kernel<KCAND_NARROW, LV_CLAMP>), grid, block, 0, s, a,
                         scratch, lpg);
    else if (a.ncand == KCAND_NARROW && a.op == LV_E_CLAMP)
      hipLaunchKernelGGL((level_partial_kernel<KCAND_NARROW, LV_E_CLAMP>), grid, block, 0, s, a,
                         scratch, lpg);
    else
      return set_error(STEAM_EINVAL, "level reduction: candidates need a clamp op");
  } else {
    switch (a.op) {
#define STEAM_LV_CASE(OPV)                                                                  \
  case OPV:                                                                                \
    hipLaunchKernelGGL((level_partial_kernel<1, OPV>), grid, block, 0, s, a, scratch, lpg); \
    break;
      STEAM_LV_CASE(LV_SUM)
      STEAM_LV_CASE(LV_ABS)
      STEAM_LV_CASE(LV_NONZERO)
      STEAM_LV_CASE(LV_CLAMP)
      STEAM_LV_CASE(LV_PHI)
      STEAM_LV_CASE(LV_MIN)
      STEAM_LV_CASE(LV_MAX)
      STEAM_LV_CASE(LV_E_SUM)
      STEAM_LV_CASE(LV_E_MIN)
      STEAM_LV_CASE(LV_E_MAX)
      STEAM_LV_CASE(LV_E_CLAMP)
      STEAM_LV_CASE(LV_SCALED_SUM)
      STEAM_LV_CASE(LV_APPLY_ABS)
#undef STEAM_LV_CASE
      default:
        return set_error(STEAM_EINVAL, "level reduction: unknown op");
    }
  }
  STEAM_LAUNCH_CHECK("level_partial_kernel");
  return STEAM_OK;
}

int level_reduce(const LevelArgs& a, double* out, double div, double* scratch, hipStream_t s) {
  int64_t groups = 0;
  int rc = level_partials(a, scratch, s, &groups);
  if (rc) return rc;
  const int64_t pairs = (int64_t)a.ncand * a.nz;
  hipLaunchKernelGGL(level_combine_kernel, dim3((unsigned)((pairs + 3) / 4)), dim3(256), 0, s,
                     scratch, (int)groups, a.ncand, a.nz, a.op, div, out, a.stop);
  STEAM_LAUNCH_CHECK("level_combine_kernel");
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

bool shape_ok(int64_t nx, int64_t ny, int64_t nz) {
  return nx >= 1 && ny >= 1 && nz >= 1 && nx < (1 << 30) && ny < (1 << 30) && nz < (1 << 20) &&
         nx * ny < (1ll << 31) && nx * ny * nz < (1ll << 40);
}

// the 60-step bisection of the level means of a's clamp op against target,
// D steps per pass; leaves the result in b.mu
template <int D>
int bisect_depth(LevelArgs a, BisectVec b, int nz, double width, double npts, int mode,
                 double* red, hipStream_t s) {
  const int gz = (nz + 255) / 256;
  hipLaunchKernelGGL((bisect_init_kernel<D>), dim3(gz), dim3(256), 0, s, b, nz, width);
  a.mu = b.cand;
  a.ncand = (1 << D) - 1;
  constexpr int NC = (1 << D) - 1;
  for (int pass = 0; pass < 60 / D; ++pass) {
    int64_t groups = 0;
    int rc = level_partials(a, red, s, &groups);
    if (rc) return rc;
    hipLaunchKernelGGL((bisect_combine_kernel<D>), dim3(nz), dim3(64 * NC), 0, s, red,
                       (int)groups, nz, a.op, b, npts, mode);
  }
  STEAM_LAUNCH_CHECK("bisection");
  return STEAM_OK;
}

int bisect(LevelArgs a, BisectVec b, int nz, double width, double npts, int mode, double* red,
           hipStream_t s) {
  const int64_t cells = (int64_t)a.nx * a.ny * a.nz;
  if (cells > KNARROW_CELLS) return bisect_depth<2>(a, b, nz, width, npts, mode, red, s);
  return cells > KWIDE_CELLS ? bisect_depth<3>(a, b, nz, width, npts, mode, red, s)
                             : bisect_depth<4>(a, b, nz, width, npts, mode, red, s);
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

extern "C" int steam_scalar_pattern_precise(const float* pert_dev, const float* dmean_z_dev,
                                            const float* mlo_z_dev, const float* mhi_z_dev,
                                            const float* sk_dev, const float* b_z_dev,
                                            const float* aspect_z_dev, const steam_zdiff* zdiff,
                                            int64_t nx, int64_t ny, int64_t nz, double dx,
                                            double dy, float* out_dev, steam_stream_t stream) {
  if (!pert_dev || !dmean_z_dev || !mlo_z_dev || !mhi_z_dev || !sk_dev || !b_z_dev ||
      !aspect_z_dev || !zdiff || !out_dev)
    return set_error(STEAM_EINVAL, "steam_scalar_pattern_precise: NULL argument");
  if (nx < 1 || ny < 1 || nz < 1 || nx * ny * nz >= (1ll << 40))
    return set_error(STEAM_EINVAL, "steam_scalar_pattern_precise: bad shape");
  if (!zdiff->uniform && nz > 2 && (!zdiff->a_dev || !zdiff->b_dev || !zdiff->c_dev))
    return set_error(STEAM_EINVAL,
                     "steam_scalar_pattern_precise: non-uniform z needs coefficients");
  ZDiff zd;
  zd.uniform = zdiff->uniform;
  zd.two_dz = zdiff->two_dz;
  zd.dz_first = zdiff->dz_first;
  zd.dz_last = zdiff->dz_last;
  zd.a = zdiff->a_dev;
  zd.b = zdiff->b_dev;
  zd.c = zdiff->c_dev;
  hipLaunchKernelGGL(pattern_precise_kernel, dim3(grid_1d(nx * ny * nz)), dim3(256), 0,
                     (hipStream_t)stream, pert_dev, dmean_z_dev, mlo_z_dev, mhi_z_dev, sk_dev,
                     b_z_dev, aspect_z_dev, zd, (int)nx, (int)ny, (int)nz, (float)(2.0 * dx),
                     (float)(2.0 * dy), out_dev);
  STEAM_LAUNCH_CHECK("pattern_precise_kernel");
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
  return level_scratch_bytes(nx * ny, nz, 1);
}

extern "C" int steam_level_reduce(const float* f_dev, const float* mean_z_dev,
                                  const double* mu_dev, int64_t nx, int64_t ny, int64_t nz,
                                  const int64_t* window, int op, double lo, double hi,
                                  double* out_dev, void* scratch_dev, int64_t scratch_bytes,
                                  steam_stream_t stream) {
  if (!f_dev || !out_dev || !scratch_dev || !shape_ok(nx, ny, nz))
    return set_error(STEAM_EINVAL, "steam_level_reduce: bad arguments");
  if (op < LV_SUM || op > LV_E_CLAMP) return set_error(STEAM_EINVAL, "steam_level_reduce: unknown op");
  if ((op == LV_PHI || op == LV_MIN || op == LV_MAX || op >= LV_E_SUM) && !mean_z_dev)
    return set_error(STEAM_EINVAL, "steam_level_reduce: op needs mean_z");
  if ((op == LV_CLAMP || op == LV_E_CLAMP) && !mu_dev)
    return set_error(STEAM_EINVAL, "steam_level_reduce: op needs mu");
  if (scratch_bytes < level_scratch_bytes(nx * ny, nz, 1))
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
  a.ncand = 1;
  // the float32 clamp takes float32 bounds; the float64 one takes them as given
  a.lo = op == LV_E_CLAMP ? lo : (double)(float)lo;
  a.hi = op == LV_E_CLAMP ? hi : (double)(float)hi;
  const bool whole = op == LV_MIN || op == LV_MAX || op == LV_E_MIN || op == LV_E_MAX;
  int rc = check_window(whole ? nullptr : window, (int)nx, (int)ny, a.win);
  if (rc) return rc;
  return level_reduce(a, out_dev, 0.0, (double*)scratch_dev, (hipStream_t)stream);
}

// bounded-add workspace: per-level vectors (5), the two step histories
// (up to 64 steps), the bisection state (4 + 2 * KCAND), the loop flags,
// then the reduction scratch
constexpr int BA_VEC = 5 + 2 * 64 + 4 + 2 * KCAND + 1;

extern "C" int64_t steam_bounded_add_workspace_bytes(int64_t nx, int64_t ny, int64_t nz) {
  return (BA_VEC * nz + 32) * 8 + level_scratch_bytes(nx * ny, nz, KCAND) + 256;
}

extern "C" int steam_bounded_add(float* pert_dev, const float* mean_z_dev, float* inc_dev,
                                 int64_t nx, int64_t ny, int64_t nz, const int64_t* window,
                                 double phi_min, double phi_max, int n_rescale,
                                 int record_applied, steam_bounded_add_ledger* ledger,
                                 int replay, void* workspace_dev, int64_t workspace_bytes,
                                 steam_stream_t stream) {
  if (!pert_dev || !mean_z_dev || !inc_dev || !ledger || !workspace_dev)
    return set_error(STEAM_EINVAL, "steam_bounded_add: NULL argument");
  if (!shape_ok(nx, ny, nz) || n_rescale < 0 || n_rescale > 64)
    return set_error(STEAM_EINVAL, "steam_bounded_add: bad shape or n_rescale");
  if (replay < 0 || replay > 2)
    return set_error(STEAM_EINVAL, "steam_bounded_add: replay must be 0, 1 or 2");
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
  double* vec = (double*)workspace_dev;
  BaVec v;
  v.a0 = vec;
  v.done = vec + nz;
  v.mean = vec + 2 * nz;
  v.shift = vec + 3 * nz;
  v.scale = vec + 4 * nz;
  v.hist_demean = vec + 5 * nz;
  v.hist_scale = vec + 69 * nz;
  double* mean_phi = vec + 133 * nz;
  BisectVec b;
  b.lo = vec + 134 * nz;
  b.hi = vec + 135 * nz;
  b.mu = vec + 136 * nz;
  b.cand = vec + 137 * nz;
  double* sums = vec + (137 + KCAND) * nz;
  b.sums = sums;
  b.target = mean_phi;
  v.pend = vec + (137 + 2 * KCAND) * nz;
  v.flags = (int*)(vec + BA_VEC * nz);
  double* red = vec + BA_VEC * nz + 32;
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
  a.ncand = 1;
  auto reduce_mean = [&](const float* f, int op, double* out, const int* stop) {
    LevelArgs c = a;
    c.f = f;
    c.op = op;
    c.stop = stop;
    return level_reduce(c, out, npts, red, s);
  };
  const size_t vbytes = (size_t)nz * 8;

  if (replay == 1) {
    // apply-only: the recorded per-level scalars, no reductions
    int nloop = 0;
    for (int z = 0; z < nz; ++z) nloop = std::max(nloop, (int)ledger->n_loop[z]);
    if (nloop > n_rescale) return set_error(STEAM_EINVAL, "steam_bounded_add: ledger n_loop too large");
    STEAM_HIP(hipMemcpyAsync(v.hist_demean, ledger->demean, vbytes * n_rescale,
                             hipMemcpyHostToDevice, s));
    STEAM_HIP(hipMemcpyAsync(v.hist_scale, ledger->scale, vbytes * n_rescale,
                             hipMemcpyHostToDevice, s));
    STEAM_HIP(hipMemcpyAsync(v.shift, ledger->final_demean, vbytes, hipMemcpyHostToDevice, s));
    STEAM_HIP(hipMemcpyAsync(b.mu, ledger->mu, vbytes, hipMemcpyHostToDevice, s));
    for (int step = 0; step < nloop; ++step) {
      hipLaunchKernelGGL(ba_shift_clip_kernel, dim3(gn), dim3(256), 0, s, inc_dev, pert_dev,
                         mean_z_dev, v.hist_demean + step, n_rescale, n, (int)nz, lo, hi,
                         nullptr);
      hipLaunchKernelGGL(ba_scale_kernel, dim3(gn), dim3(256), 0, s, inc_dev,
                         v.hist_scale + step, n_rescale, n, (int)nz, nullptr);
    }
    hipLaunchKernelGGL(ba_shift_clip_kernel, dim3(gn), dim3(256), 0, s, inc_dev, pert_dev,
                       mean_z_dev, v.shift, 1, n, (int)nz, lo, hi, nullptr);
    hipLaunchKernelGGL(ba_to_field_kernel, dim3(gn), dim3(256), 0, s, inc_dev, pert_dev,
                       mean_z_dev, n, (int)nz);
  } else {
    // mean of the running field over the window, and the target amplitude
    if ((rc = reduce_mean(pert_dev, LV_PHI, mean_phi, nullptr))) return rc;
    if ((rc = reduce_mean(inc_dev, LV_ABS, v.a0, nullptr))) return rc;
    hipLaunchKernelGGL(ba_init_kernel, dim3(gz), dim3(256), 0, s, v, (int)nz, n_rescale);
    // each step is two passes over d: the mean of the (pending-scaled) d,
    // then scale, demean and clip written back with the mean |d| taken on
    // the way; the step's own scale stays pending for the next pass
    LevelArgs scaled = a;
    scaled.f = inc_dev;
    scaled.op = LV_SCALED_SUM;
    scaled.scale = v.pend;
    scaled.stop = v.flags;
    LevelArgs apply = scaled;
    apply.op = LV_APPLY_ABS;
    apply.out = inc_dev;
    apply.pert = pert_dev;
    apply.shift = v.shift;
    memcpy(apply.awin, win, sizeof(win));
    apply.win[0] = 0;
    apply.win[1] = (int)nx;
    apply.win[2] = 0;
    apply.win[3] = (int)ny;
    // a step is four launches: each pass's partials, and each combine with
    // the step's per-level work (demean; rescale and the all-done check)
    const unsigned gzw = (unsigned)((nz + 3) / 4);
    for (int step = 0; step < n_rescale; ++step) {
      int64_t groups = 0;
      if ((rc = level_partials(scaled, red, s, &groups))) return rc;
      hipLaunchKernelGGL(ba_combine_demean_kernel, dim3(gzw), dim3(256), 0, s, red, (int)groups,
                         (int)nz, npts, v, step, n_rescale);
      if ((rc = level_partials(apply, red, s, &groups))) return rc;
      hipLaunchKernelGGL(ba_combine_rescale_check_kernel, dim3(1), dim3(BA_CK_NTH), 0, s, red,
                         (int)groups, (int)nz, npts, v, step, n_rescale);
    }
    // the final demean and clip run whatever the loop did; then e = d + phi
    // in place, and mu is bisected in that shifted space
    scaled.stop = nullptr;
    if ((rc = level_reduce(scaled, v.shift, npts, red, s))) return rc;
    hipLaunchKernelGGL(ba_final_kernel, dim3(gn), dim3(256), 0, s, inc_dev, pert_dev,
                       mean_z_dev, v.pend, v.shift, n, (int)nz, lo, hi);
    LevelArgs c = a;
    c.f = inc_dev;
    c.op = LV_CLAMP;
    if ((rc = bisect(c, b, (int)nz, width, npts, 0, red, s))) return rc;
    const hipMemcpyKind kind = replay == 2 ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    STEAM_HIP(hipMemcpyAsync(ledger->a0, v.a0, vbytes, kind, s));
    STEAM_HIP(hipMemcpyAsync(ledger->demean, v.hist_demean, vbytes * n_rescale, kind, s));
    STEAM_HIP(hipMemcpyAsync(ledger->scale, v.hist_scale, vbytes * n_rescale, kind, s));
    STEAM_HIP(hipMemcpyAsync(ledger->final_demean, v.shift, vbytes, kind, s));
    STEAM_HIP(hipMemcpyAsync(ledger->mu, b.mu, vbytes, kind, s));
    if (replay == 2) {
      hipLaunchKernelGGL(ba_counts_kernel, dim3(gz), dim3(256), 0, s, v.flags, ledger->n_loop,
                         ledger->n_scale, (int)nz);
    } else {
      int flags[2];
      STEAM_HIP(hipMemcpyAsync(flags, v.flags, sizeof(flags), hipMemcpyDeviceToHost, s));
      STEAM_HIP(hipStreamSynchronize(s));
      for (int z = 0; z < nz; ++z) {
        ledger->n_loop[z] = flags[1];
        ledger->n_scale[z] = flags[1];
      }
    }
  }
  hipLaunchKernelGGL(ba_finish_kernel, dim3(gn), dim3(256), 0, s, inc_dev, pert_dev, inc_dev,
                     mean_z_dev, b.mu, n, (int)nz, lo, hi, record_applied);
  STEAM_LAUNCH_CHECK("bounded add");
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

// projection workspace: active, vmin, vmax, record, target, the bisection
// state (3 + 2 * KCAND), then the reduction scratch
constexpr int PJ_VEC = 8 + 2 * KCAND;

extern "C" int64_t steam_project_workspace_bytes(int64_t nx, int64_t ny, int64_t nz) {
  return (PJ_VEC * nz + 32) * 8 + level_scratch_bytes(nx * ny, nz, KCAND) + 256;
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
  double* active = vec;
  double* vmin = vec + nz;
  double* vmax = vec + 2 * nz;
  double* rec = vec + 3 * nz;
  double* target = vec + 4 * nz;
  BisectVec b;
  b.lo = vec + 5 * nz;
  b.hi = vec + 6 * nz;
  b.mu = vec + 7 * nz;
  b.cand = vec + 8 * nz;
  double* sums = vec + (8 + KCAND) * nz;
  b.sums = sums;
  b.target = target;
  double* red = vec + PJ_VEC * nz + 32;
  const int gz = (int)((nz + 255) / 256);
  const int gn = grid_1d(n);
  LevelArgs a;
  memset(&a, 0, sizeof(a));
  a.f = f;
  a.mean_z = mean_z;
  a.nx = (int)nx;
  a.ny = (int)ny;
  a.nz = (int)nz;
  a.lo = lo;
  a.hi = hi;
  a.ncand = 1;
  a.active = active;
  int rc;
  if (replay) {
    STEAM_HIP(hipMemcpyAsync(b.mu, mu_host, (size_t)nz * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(proj_replay_active_kernel, dim3(gz), dim3(256), 0, s, b.mu, active,
                       (int)nz);
  } else {
    LevelArgs whole = a;
    whole.win[0] = 0;
    whole.win[1] = (int)nx;
    whole.win[2] = 0;
    whole.win[3] = (int)ny;
    whole.op = LV_E_MIN;
    if ((rc = level_reduce(whole, vmin, 0.0, red, s))) return rc;
    whole.op = LV_E_MAX;
    if ((rc = level_reduce(whole, vmax, 0.0, red, s))) return rc;
    hipLaunchKernelGGL(proj_active_kernel, dim3(gz), dim3(256), 0, s, vmin, vmax, active,
                       (int)nz, lo, hi);
    LevelArgs c = a;
    memcpy(c.win, win, sizeof(c.win));
    c.op = LV_E_SUM;
    if ((rc = level_reduce(c, target, npts, red, s))) return rc;
    c.op = LV_E_CLAMP;
    if ((rc = bisect(c, b, (int)nz, width, npts, 1, red, s))) return rc;
    hipLaunchKernelGGL(proj_record_kernel, dim3(gz), dim3(256), 0, s, b.mu, active, rec,
                       (int)nz);
    STEAM_HIP(hipMemcpyAsync(mu_host, rec, (size_t)nz * 8, hipMemcpyDeviceToHost, s));
  }
  hipLaunchKernelGGL(proj_apply_kernel, dim3(gn), dim3(256), 0, s, f, mean_z, b.mu, active, n,
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
  if (!pert_dev || !mean_z_dev || !mu || !workspace_dev || !shape_ok(nx, ny, nz))
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
  if (!state_dev || !mean_z_dev || !out_dev || !shape_ok(nx, ny, nz))
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
  return (nz + 8) * 8 + level_scratch_bytes(nx * ny, nz, 1) + 256;
}

extern "C" int steam_compose_flux_output(const float* state_dev, float* deficit_dev, int64_t nx,
                                         int64_t ny, int64_t nz, const int64_t* window,
                                         double* sums, int replay, void* workspace_dev,
                                         int64_t workspace_bytes, steam_stream_t stream) {
  if (!state_dev || !deficit_dev || !sums || !workspace_dev || !shape_ok(nx, ny, nz))
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
  a.ncand = 1;
  memcpy(a.win, win, sizeof(a.win));
  if (replay) {
    if (!(sums[1] > 0.0) || !(sums[3] > 0.0))
      return set_error(STEAM_EINVAL, "steam_compose_flux_output: recorded counts must be positive");
    double m[2] = {sums[0] / sums[1], sums[2] / sums[3]};
    STEAM_HIP(hipMemcpyAsync(means, m, 16, hipMemcpyHostToDevice, s));
  } else {
    a.f = state_dev;
    if ((rc = level_reduce(a, lv, 0.0, red, s))) return rc;
    hipLaunchKernelGGL(sum_levels_kernel, dim3(1), dim3(64), 0, s, lv, (int)nz, count, tot,
                       means);
  }
  hipLaunchKernelGGL(flux_compose_kernel, dim3(grid_1d(n)), dim3(256), 0, s, deficit_dev,
                     state_dev, n);
  if (!replay) {
    a.f = deficit_dev;
    if ((rc = level_reduce(a, lv, 0.0, red, s))) return rc;
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
