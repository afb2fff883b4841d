// FFT overlap-add convolution (periodic x/y, zero-padded z) and the fused
// flux-cascade advance, for gfx950.
//
// Operator (steam/utils.py:293-427): out = field (*) kernel with circular
// x/y over the whole extent, linear z with zero padding, kernel centred at
// ((kx-1)/2, (ky-1)/2, (kz-1)/2) and folded onto the field when wider than
// it (steam/utils.py:222-242).  Overlap-add runs along z over blocks of
// L = nfft - kz + 1 input planes; x and y keep their exact periods.
//
// Five passes per operator application, all through LDS:
//   zfwd  real z-lines -> half-length complex FFT -> r2c split -> planes
//         S[block][kz][x][y] (the flux advance generates the keyed noise,
//         multiplies it by the entering flux and takes its partial sums here)
//   ypass contiguous y-lines, forward
//   xpass x-lines (T-wide column tiles): forward, * kernel spectrum, inverse
//   ypass inverse
//   zinv  c2r per block, overlap-add of all blocks in LDS, one coalesced
//         write per output line (the flux advance adds the increment, clips
//         and takes the realized partial sums here)
// Partial sums are per workgroup and reduced in a fixed order by one
// workgroup, so results do not depend on launch composition.
#pragma once

#include <math.h>

#include <algorithm>
#include <type_traits>
#include <cstring>
#include <vector>

#include "conv.h"
#include "fft_reg.h"
#include "fft_zreg.h"

// z-pass FFT stages inlined (false) or called out of line (true)
#ifndef Z_STAGE_CALLS
#define Z_STAGE_CALLS false
#endif

// compile-time tuning knobs (overridable with -D for measurement builds)
#ifndef STEAM_ZR_WAVES
#define STEAM_ZR_WAVES 4
#endif
// register z inverse pass: load the entering flux before the transform (1;
// a 512x512x256 pipelined step 0.910 ms against 0.914 after it over three
// alternating runs) or after it (0)
#ifndef STEAM_ZR_PREFETCH
#define STEAM_ZR_PREFETCH 1
#endif
#ifndef STEAM_XREG_COLS
#define STEAM_XREG_COLS 8
#endif
#ifndef STEAM_ZINV_ATTR
#define STEAM_ZINV_ATTR
#endif
// register x pass: issue the kernel-spectrum loads with the line's loads,
// before the forward transform (1), or after it (0)
#ifndef STEAM_XREG_KPREFETCH
#define STEAM_XREG_KPREFETCH 1
#endif
// register x pass with two transform lanes per thread: y columns per workgroup
#ifndef STEAM_XREG_COLS2
#define STEAM_XREG_COLS2 16
#endif
// register x pass: largest log2(nx) taking two lanes per thread (0: none).
// Off: the forward * spectrum * inverse form of two lanes needs more than
// the 128 registers of two 512-thread workgroups per CU and spills.
#ifndef STEAM_XREG_C2_MAXLOG
#define STEAM_XREG_C2_MAXLOG 0
#endif

namespace steam {

// Lane addressing of the register passes.  An access is a base that is
// uniform over the workgroup (its first line, one plane of its columns, one
// row of a plane) plus a 32-bit lane byte offset plus a compile-time
// constant, so the compiler forms one lane address and takes the constants
// as immediates: no per-access 64-bit vector address arithmetic.
template <class T>
__device__ __forceinline__ T* lane_at(T* base, unsigned lane_bytes, int const_bytes) {
  using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return reinterpret_cast<T*>(reinterpret_cast<B*>(base) + (size_t)lane_bytes + const_bytes);
}

constexpr int XREG_COLS = STEAM_XREG_COLS;  // y columns per register x-pass workgroup
constexpr int XREG_COLS2 = STEAM_XREG_COLS2;  // the same with two lanes per thread
// transform lanes per thread of the register x pass for an nx of 2^logn
__host__ __device__ constexpr int xreg_lanes(int logn) {
  return logn >= 6 && logn <= STEAM_XREG_C2_MAXLOG ? 2 : 1;
}
// y columns per workgroup of the register x pass
__host__ __device__ constexpr int xreg_cols(int logn) {
  const int pt = (1 << logn) / 8 / xreg_lanes(logn);
  const int cap = xreg_lanes(logn) == 2 ? XREG_COLS2 : XREG_COLS;
  return 1024 / pt < cap ? 1024 / pt : cap;
}
// z passes: 512 threads holding 8 values each (same tile as 256 x 16, half
// the registers per thread)
constexpr int Z_NTH = 512;
constexpr int Z_EM = 8;
}

namespace steam {

// ---------------------------------------------------------------------------
// descriptor sources: one instance by value, or a batch in device memory
// ---------------------------------------------------------------------------
struct SingleSrc {
  PassDesc d;
  __device__ __forceinline__ const PassDesc& get(int, int&) const { return d; }
};

struct BatchSrc {
  const PassDesc* d;
  int n;
  __device__ __forceinline__ const PassDesc& get(int pass, int& wg) const {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (d[mid].wg_base[pass] <= wg)
        lo = mid;
      else
        hi = mid - 1;
    }
    wg -= d[lo].wg_base[pass];
    return d[lo];
  }
};

enum { PASS_Z = 0, PASS_Y = 1, PASS_X = 2, PASS_ZINV = 3, PASS_POST = 4,
       PASS_FIN = 5, PASS_NOISE = 6 };

// workgroup sum of NV doubles over NT threads, result valid in thread 0
// (fixed order: independent of launch composition)
template <int NV, int NT>
__device__ void block_sum(double (&v)[NV], double* red) {
  const int tid = threadIdx.x;
  for (int i = 0; i < NV; ++i) red[i * NT + tid] = v[i];
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (tid < s)
      for (int i = 0; i < NV; ++i) red[i * NT + tid] += red[i * NT + tid + s];
    __syncthreads();
  }
  if (tid == 0)
    for (int i = 0; i < NV; ++i) v[i] = red[i * NT];
}

// The entering flux value of a stored cell: the pending rescale of the
// previous advance (sigma_in, or chained from the ledger it left on the
// device) applied on read, or the flat entering mean of an all-clipped class.
// Both cases are one fma with uniform coefficients: sig * raw + 0 rounds as
// the product does, and 0 * raw + em is em (stored flux is finite).
struct Pending {
  float a, b;
};
__device__ __forceinline__ Pending pending_of(const PassDesc& d) {
  float sig = d.sigma_in, em = 0.f;
  int flat = 0;
  if (d.pend) {
    flat = d.pend->flatten;
    sig = d.pend->rescale;
    em = d.pend->entering_mean_f;
  }
  return flat ? Pending{0.f, em} : Pending{sig, 0.f};
}
__device__ __forceinline__ float entering_value(const Pending& p, float raw) {
  return fmaf(p.a, raw, p.b);
}
// The advanced flux as stored: with out_scale (apply mode) the recorded
// rescale applied as the post pass would apply it.
__device__ __forceinline__ float out_value(const PassDesc& d, float v) {
  return d.out_scale ? (d.out_flat ? d.out_em : __fmul_rn(v, d.out_rs)) : v;
}

// kp[x, y, z] of the rolled, folded kernel: taps i = x + offx (mod nx),
// j = y + offy (mod ny), summed x-fold first then y-fold.
__device__ __forceinline__ float kernel_gather(const PassDesc& d, int x, int y,
                                               int z) {
  const ConvGeom& g = d.g;
  int i0 = (int)pmod((int64_t)x + g.offx, g.nx);
  int j0 = (int)pmod((int64_t)y + g.offy, g.ny);
  float acc = 0.f;
  for (int j = j0; j < g.ky; j += g.ny) {
    float s = 0.f;
    for (int i = i0; i < g.kx; i += g.nx) s += d.in[((int64_t)i * g.ky + j) * g.kz + z];
    acc += s;
  }
  return acc;
}

// Kernel-spectrum factor of z bin kz: the normalisation kscale times the
// phase exp(+2 pi i kz cz / nfft), which shifts the circular z output by the
// 'same' trim cz so that output level z is transform sample z.
__device__ __forceinline__ float2 kernel_bin_factor(const PassDesc& d, int kz) {
  const ConvGeom& g = d.g;
  const int64_t r = ((int64_t)kz * g.cz) % g.nfft;
  double sn, cs;
  sincospi(2.0 * (double)r / (double)g.nfft, &sn, &cs);
  return make_float2((float)(cs * d.kscale), (float)(sn * d.kscale));
}

// ---------------------------------------------------------------------------
// noise: the unit-mean multiplier noise field of one advance (VALU-bound:
// Philox + CMS + expm1 per centre).  Independent of the flux, so it may run
// ahead of the advance on another stream.  4 lines per workgroup, threads
// along z (coalesced stores).
// ---------------------------------------------------------------------------
// workgroups of a noise field generated ahead of its advance (0: one per
// four lines), one per CU of the 256.  Generated two classes ahead beside
// the transforms, a 512x512x256 step over 100 steps took 0.852-0.854 ms at
// 256, 0.894 at 320, 0.864-0.870 at 384 and 448, 0.860-0.864 at 512 and
// 0.858-0.862 at 640 (two rounds on one box).  With the noise kernel at 0.41 ms
// alone 512 had been best (0.914 ms; 384 fell behind the transforms on some
// boxes); at 0.33 ms one workgroup per CU keeps up and takes fewer of the
// passes' slots.
#ifndef STEAM_NOISE_WGS
#define STEAM_NOISE_WGS 256
#endif
// STRIDE: a grid smaller than the field's line groups walks them with a
// grid stride (a few workgroups per CU keep the VALU busy beside memory-bound
// passes on another stream without taking their slots).  A workgroup takes
// noise_lines_per_wg(nz) consecutive lines, noise_lines_per_pass(nz) of them
// side by side per pass when lines are short (lane t then holds level
// t mod nz of line t / nz, so a 48-level line keeps 240 of 256 lanes busy
// rather than 48); a line of more than 128 levels is strided over by the
// workgroup, with its x, y and keyed-noise set-up uniform.
// LINES: NOISE_LONG or NOISE_SHORT when every instance of the launch has
// lines of that kind (compiled for it alone), NOISE_ANY otherwise.
enum { NOISE_ANY = 0, NOISE_LONG = 1, NOISE_SHORT = 2 };
template <class Src, bool STRIDE = false, int LINES = NOISE_ANY>
__global__ __launch_bounds__(256) void noise_field_kernel(Src src) {
  int wg = blockIdx.x;
  const PassDesc& d = src.get(PASS_NOISE, wg);
  const ConvGeom& g = d.g;
  const int nz = g.nz;
  const int64_t lines = (int64_t)g.nx * g.ny;
  const int U = LINES == NOISE_LONG ? 1 : noise_lines_per_pass(nz);
  const int lpw = LINES == NOISE_LONG ? 4 : noise_lines_per_wg(nz);
  const int64_t stride =
      STRIDE ? (int64_t)(d.wg_noise_n > 0 ? d.wg_noise_n : (int)gridDim.x) * lpw : lines;
  if (LINES == NOISE_LONG || (LINES == NOISE_ANY && U == 1)) {
    for (int64_t base = (int64_t)wg * lpw; base < lines; base += stride) {
      for (int l = 0; l < lpw; ++l) {
        const int64_t line = base + l;
        if (line >= lines) break;
        const int x = (int)(line / g.ny);
        const int y = (int)(line - (int64_t)x * g.ny);
        const LineNoise ln = noise_line(d.noise, x, y);
        float* o = d.nfield + line * nz;
        for (int z = threadIdx.x; z < nz; z += 256) o[z] = noise_at(d.noise, ln, z, nz);
      }
    }
    return;
  }
  if constexpr (LINES == NOISE_LONG) return;
  // short lines: a fixed (line slot, level) per lane; the lanes past U
  // lines have no cell (the kernel has no barrier, so they leave)
  const int sub = (int)threadIdx.x / nz;
  const int z = (int)threadIdx.x - sub * nz;
  if (sub >= U) return;
  for (int64_t base = (int64_t)wg * lpw; base < lines; base += stride) {
    for (int l0 = 0; l0 < lpw; l0 += U) {
      const int64_t line = base + l0 + sub;
      if (line >= lines) break;
      // lines < 2^24 (x/y extents of at most 4096)
      const unsigned ul = (unsigned)line;
      const int x = (int)(ul / (unsigned)g.ny);
      const int y = (int)(ul - (unsigned)x * (unsigned)g.ny);
      d.nfield[line * nz + z] = noise_at(d.noise, noise_line(d.noise, x, y), z, nz);
    }
  }
}

// ---------------------------------------------------------------------------
// zfwd: real z-lines -> spectral planes
// ---------------------------------------------------------------------------
template <class Src, int MODE, int RS>
__global__ __launch_bounds__(Z_NTH) __attribute__((amdgpu_waves_per_eu(4))) void zfwd_kernel(Src src) {
  int wg = blockIdx.x;
  const PassDesc& d = src.get(PASS_Z, wg);
  const ConvGeom& g = d.g;
  extern __shared__ float2 smem[];
  const int tid = threadIdx.x;
  const int x = wg / g.nyt_z;
  const int y0 = (wg - x * g.nyt_z) * g.Tz;
  const int ldn = g.Tz + 1;  // element (line t, index n) at n*ldn + t
  float* sr = reinterpret_cast<float*>(smem);
  const int nzin = (MODE == MODE_KERNEL) ? g.kz : g.nz;
  const int nblk = (MODE == MODE_KERNEL) ? 1 : g.nblocks;
  const int L = (MODE == MODE_KERNEL) ? g.nfft : g.L;
  double acc[3] = {0.0, 0.0, 0.0};
  bool xin = false;
  Pending pd{1.f, 0.f};
  if (MODE == MODE_FLUX) {
    xin = x >= d.win[0] && x < d.win[1];
    pd = pending_of(d);
  }

  for (int b = 0; b < nblk; ++b) {
    const int z0 = b * L;
    for (int t = 0; t < g.Tz; ++t) {
      const int y = y0 + t;
      const bool yv = y < g.ny;
      const int64_t line = ((int64_t)x * g.ny + y) * nzin;
      bool win = false;
      if (MODE == MODE_FLUX && yv) win = xin && y >= d.win[2] && y < d.win[3];
      for (int m = tid; m < g.nfft; m += Z_NTH) {
        const int z = z0 + m;
        float v = 0.f;
        if (yv && m < L && z < nzin) {
          if (MODE == MODE_PLAIN) {
            v = d.in[line + z];
          } else if (MODE == MODE_FLUX) {
            float f = entering_value(pd, d.unit_in ? 1.f : d.out[line + z]);
            const float nse = d.nfield[line + z];
            float p = __fmul_rn(nse, f);
            if (win) {
              acc[0] += (double)f;
              acc[1] += (double)fabsf(nse);
              acc[2] += (nse != 0.f) ? 1.0 : 0.0;
            }
            if (d.sk) d.sk[line + z] = p;
            v = p;
          } else {
            v = kernel_gather(d, x, y, z);
          }
        }
        sr[((m >> 1) * ldn + t) * 2 + (m & 1)] = v;
      }
    }
    __syncthreads();
    fft_lines<true, RS, Z_STAGE_CALLS, Z_EM, Z_NTH>(smem, g.az, g.Tz, g.logTz, ldn, 1, false, tid);
    // r2c split: X2[k] = (Z[k] + conj Z[n2-k]) - i W^k (Z[k] - conj Z[n2-k])
    const int nh = g.n2 / 2 + 1;
    for (int e = tid; e < g.Tz * nh; e += Z_NTH) {
      const int t = e & (g.Tz - 1);
      const int k = e >> g.logTz;
      float2* Lp = smem + t;
      if (k == 0) {
        float2 z0v = Lp[0];
        Lp[0] = make_float2(2.f * (z0v.x + z0v.y), 0.f);
        Lp[g.n2 * ldn] = make_float2(2.f * (z0v.x - z0v.y), 0.f);
      } else {
        const int km = g.n2 - k;
        float2 zk = Lp[k * ldn], zm = Lp[km * ldn];
        float2 w = g.wr2c[k];
        float2 A = make_float2(zk.x + zm.x, zk.y - zm.y);
        float2 B = make_float2(zk.x - zm.x, zk.y + zm.y);
        float2 wB = cmul(w, B);
        Lp[k * ldn] = make_float2(A.x + wB.y, A.y - wB.x);
        if (km != k) {
          float2 A2 = make_float2(zm.x + zk.x, zm.y - zk.y);
          float2 B2 = make_float2(zm.x - zk.x, zm.y + zk.y);
          float2 wB2 = cmul(make_float2(-w.x, w.y), B2);
          Lp[km * ldn] = make_float2(A2.x + wB2.y, A2.y - wB2.x);
        }
      }
    }
    __syncthreads();
    float2* Sb = d.S + (int64_t)b * g.nzc * g.nx * g.ny;
    for (int e = tid; e < g.nzc * g.Tz; e += Z_NTH) {
      const int kz = e >> g.logTz;
      const int t = e & (g.Tz - 1);
      const int y = y0 + t;
      if (y < g.ny) Sb[((int64_t)kz * g.nx + x) * g.ny + y] = smem[kz * ldn + t];
    }
    __syncthreads();
  }
  if (MODE == MODE_FLUX && !d.apply) {
    __shared__ double red[3 * Z_NTH];
    block_sum<3, Z_NTH>(acc, red);
    if (tid == 0) {
      d.part_z[wg * 3 + 0] = acc[0];
      d.part_z[wg * 3 + 1] = acc[1];
      d.part_z[wg * 3 + 2] = acc[2];
    }
  }
}

// ---------------------------------------------------------------------------
// ypass: contiguous y-lines (Ty per workgroup), in place
// ---------------------------------------------------------------------------
// threads of a generic x/y pass workgroup: the Bluestein set's length-M
// scratch transform (M < 2^13) takes 512 at 16 values each
constexpr int generic_xy_nth(int rs) { return rs == 3 ? 512 : FFT_NTH; }

template <class Src, int RS>
__global__ __launch_bounds__(generic_xy_nth(RS)) __attribute__((amdgpu_waves_per_eu(4))) void ypass_kernel(Src src, int inv) {
  constexpr int NT = generic_xy_nth(RS);
  int wg = blockIdx.x;
  const PassDesc& d = src.get(PASS_Y, wg);
  const ConvGeom& g = d.g;
  extern __shared__ float2 smem[];
  const int tid = threadIdx.x;
  const int P = (d.mode == MODE_KERNEL) ? g.nzc : g.nblocks * g.nzc;
  const int64_t lines = (int64_t)P * g.nx;
  const int64_t l0 = (int64_t)wg * g.Ty;
  const int ny = g.ny;
  const int ldn = g.Ty + 1;
  for (int t = 0; t < g.Ty; ++t) {
    const int64_t l = l0 + t;
    const float2* srcl = d.S + l * ny;
    for (int n = tid; n < ny; n += NT)
      smem[n * ldn + t] = (l < lines) ? srcl[n] : make_float2(0.f, 0.f);
  }
  AxisPlan ap = g.ay;
  if (RS == 2 && g.twl_y) {
    float2* twl = smem + ny * ldn;
    for (int n = tid; n < ny; n += NT) twl[n] = g.ay.tw[n];
    ap.tw = twl;
  }
  __syncthreads();
  fft_lines<true, RS, false, FFT_EMAX, NT>(smem, ap, g.Ty, g.logTy, ldn, 1, inv != 0, tid);
  for (int t = 0; t < g.Ty; ++t) {
    const int64_t l = l0 + t;
    if (l >= lines) break;
    float2* dstl = d.S + l * ny;
    for (int n = tid; n < ny; n += NT) dstl[n] = smem[n * ldn + t];
  }
}

// ---------------------------------------------------------------------------
// xpass: x-lines in column tiles of Tx; XMODE 0: fwd * K inv, 1: fwd * scale
// ---------------------------------------------------------------------------
template <class Src, int XMODE, int RS>
__global__ __launch_bounds__(generic_xy_nth(RS)) __attribute__((amdgpu_waves_per_eu(4))) void xpass_kernel(Src src) {
  constexpr int NT = generic_xy_nth(RS);
  int wg = xcd_remap(blockIdx.x, gridDim.x);
  const PassDesc& d = src.get(PASS_X, wg);
  const ConvGeom& g = d.g;
  extern __shared__ float2 smem[];
  const int tid = threadIdx.x;
  const int p = wg / g.nyt_x;
  const int y0 = (wg - p * g.nyt_x) * g.Tx;
  const int nx = g.nx, ny = g.ny, Tx = g.Tx, logTx = g.logTx;
  float2* Sp = d.S + (int64_t)p * nx * ny;
  const int ne = nx * Tx;
  for (int e = tid; e < ne; e += NT) {
    const int t = e & (Tx - 1);
    const int n = e >> logTx;
    const int y = y0 + t;
    smem[e] = (y < ny) ? Sp[(int64_t)n * ny + y] : make_float2(0.f, 0.f);
  }
  AxisPlan ap = g.ax;
  if (RS == 2 && g.twl_x) {
    float2* twl = smem + ne;
    for (int n = tid; n < nx; n += NT) twl[n] = g.ax.tw[n];
    ap.tw = twl;
  }
  __syncthreads();
  fft_lines<true, RS, false, FFT_EMAX, NT>(smem, ap, Tx, logTx, Tx, 1, false, tid);
  if (XMODE == 0) {
    const int kzp = p % g.nzc;
    const float2* Kp = d.K + (int64_t)kzp * nx * ny;
    for (int e = tid; e < ne; e += NT) {
      const int t = e & (Tx - 1);
      const int n = e >> logTx;
      const int y = y0 + t;
      if (y < ny) smem[e] = cmul(smem[e], Kp[(int64_t)n * ny + y]);
    }
    __syncthreads();
    fft_lines<true, RS, false, FFT_EMAX, NT>(smem, ap, Tx, logTx, Tx, 1, true, tid);
  } else {
    const float2 f = kernel_bin_factor(d, p % g.nzc);
    for (int e = tid; e < ne; e += NT) smem[e] = cmul(smem[e], f);
    __syncthreads();
  }
  for (int e = tid; e < ne; e += NT) {
    const int t = e & (Tx - 1);
    const int n = e >> logTx;
    const int y = y0 + t;
    if (y < ny) Sp[(int64_t)n * ny + y] = smem[e];
  }
}

// ---------------------------------------------------------------------------
// register-FFT y pass (power-of-two ny): P = ny/8 threads per contiguous
// line, 256/P lines per workgroup, HBM -> registers -> HBM
// ---------------------------------------------------------------------------
template <class Src, int LOGN>
__global__ __launch_bounds__(256) void ypass_reg_kernel(Src src, int inv) {
  using F = RegFFT<LOGN>;
  constexpr int P = F::P, LPW = 256 / P;
  __shared__ float2 lds[LPW * F::LDSN];
  int wg = blockIdx.x;
  const PassDesc& d = src.get(PASS_Y, wg);
  const ConvGeom& g = d.g;
  const int tid = threadIdx.x;
  const int q = tid % P, lw = tid / P;
  const int planes = (d.mode == MODE_KERNEL) ? g.nzc : g.nblocks * g.nzc;
  const int64_t lines = (int64_t)planes * g.nx;
  const int64_t line = (int64_t)wg * LPW + lw;
  const bool valid = line < lines;
  float2* Sl = d.S + (valid ? line : 0) * F::N;
  float2 v[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = valid ? Sl[q + m * P] : make_float2(0.f, 0.f);
  RegTw<LOGN> twb;
  reg_tw_load<LOGN>(twb, q, g.ay.tw);
  fft_reg<LOGN>(v, q, lds, 1, lw * F::LDSN, g.ay.tw, twb, inv != 0);
  if (valid) {
#pragma unroll
    for (int m = 0; m < 8; ++m) Sl[q + m * P] = v[m];
  }
}

// register-FFT x pass (power-of-two nx): each x-line is held by P / C
// threads (P = nx/8 lanes of the radix-8 transform, C lanes per thread), Tr
// adjacent y columns per workgroup (thread t % Tr = column, so each row
// access is a Tr*8-byte run); XMODE 0: forward, * kernel spectrum, inverse;
// 1: forward, * scale.
// register x pass: waves per SIMD the compiler is held to (0: its own choice)
#ifndef STEAM_XREG_WAVES
#define STEAM_XREG_WAVES 0
#endif
#if STEAM_XREG_WAVES > 0
#define STEAM_XREG_WAVES_ATTR __attribute__((amdgpu_waves_per_eu(STEAM_XREG_WAVES, STEAM_XREG_WAVES)))
#else
#define STEAM_XREG_WAVES_ATTR
#endif
// One workgroup of the register x pass.  KREAL: the kernel spectrum is real
// and even in x (a kernel with mirror symmetry on every axis), stored as
// floats [kz][y/T][nx/2 + 1][T]; bin x takes the value of min(x, nx - x).
template <int XMODE, int LOGN, int C, bool KREAL>
__device__ __forceinline__ void xpass_reg_body(const PassDesc& d, int wg, float2* lds) {
  using F = RegFFT<LOGN>;
  constexpr int P = F::P, PT = P / C;
  constexpr int T = (1024 / PT < (C == 1 ? XREG_COLS : XREG_COLS2))
                        ? 1024 / PT : (C == 1 ? XREG_COLS : XREG_COLS2);
  const ConvGeom& g = d.g;
  const int tid = threadIdx.x;
  const int t = tid % T, q0 = tid / T;
  const int nyt = (g.ny + T - 1) / T;
  const int p = wg / nyt;
  const int y = (wg - p * nyt) * T + t;
  const bool valid = y < g.ny;
  const int nx = g.nx, ny = g.ny;
  // Value m of lane c is row x = q + m P (q = q0 + c PT) of the plane: a row
  // base uniform over the workgroup (Sb + m P ny) plus the lane's 32-bit byte
  // offset (q ny + y) * 8 < nx ny * 8, so no access re-forms a 64-bit address.
  float2* Sb = d.S + (int64_t)p * nx * ny;
  const int yo = valid ? y : 0;
  auto row = [&](int m) { return Sb + (int64_t)m * P * ny; };
  unsigned lb[C];
#pragma unroll
  for (int c = 0; c < C; ++c) lb[c] = (unsigned)((q0 + c * PT) * ny + yo) * 8u;
  auto off = [&](int c, int m) { return (q0 + c * PT + m * P) * ny + yo; };
  using KV = typename std::conditional<KREAL, float, float2>::type;
  float2 v[C][8];
  KV kv[C][8];
#pragma unroll
  for (int c = 0; c < C; ++c)
#pragma unroll
    for (int m = 0; m < 8; ++m)
      v[c][m] = valid ? *lane_at(row(m), lb[c], 0) : make_float2(0.f, 0.f);
  // the kernel spectrum: in this pass's column tiles [kz][y/T][x][T] when
  // the plan stores it so (the workgroup's block is then one contiguous run)
  // With the real, x-even format, row x takes bin min(x, nx - x): for
  // x = q + m P (q < P, nx = 8 P) that is x itself for m < 4 and
  // (8 - m) P - q for m >= 4 (both give 4 P at x = nx / 2), so both halves
  // are a uniform base plus a lane offset.
  const KV* Kb = nullptr;
  const bool kt = d.ktile != 0;
  auto load_k = [&]() {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int q = q0 + c * PT;
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        if constexpr (KREAL) {
          const unsigned lo = (unsigned)(q * T + t) * 4u;
          const unsigned hi = (unsigned)((P - q) * T + t) * 4u;
          kv[c][m] = !valid ? 0.f
                     : m < 4 ? *lane_at(Kb + m * P * T, lo, 0)
                             : *lane_at(Kb + (7 - m) * P * T, hi, 0);
        } else if (kt) {
          kv[c][m] = valid ? *lane_at(Kb + m * P * T, (unsigned)(q * T + t) * 8u, 0)
                           : make_float2(0.f, 0.f);
        } else {
          kv[c][m] = valid ? Kb[off(c, m)] : make_float2(0.f, 0.f);
        }
      }
    }
  };
  // one lane per thread: the spectrum loads go out with the line's; two
  // lanes: after the forward transform (the registers of both would spill)
  constexpr bool KPRE = STEAM_XREG_KPREFETCH && C == 1;
  if (XMODE == 0) {
    const int64_t tile = (int64_t)(p % g.nzc) * nyt + (wg - p * nyt);
    if constexpr (KREAL)
      Kb = reinterpret_cast<const KV*>(d.K) + tile * (nx / 2 + 1) * T;
    else
      Kb = kt ? d.K + tile * nx * T : d.K + (int64_t)(p % g.nzc) * nx * ny;
    if (KPRE) load_k();
  }
  // the twiddles, formed once for the forward and the inverse transform
  // (two lanes per thread keep only the bases: registers)
  RegTw<LOGN> twb[C];
#pragma unroll
  for (int c = 0; c < C; ++c) reg_tw_load<LOGN>(twb[c], q0 + c * PT, g.ax.tw);
  using TWX = typename std::conditional<C == 1, RegTwFull<LOGN>, RegTw<LOGN>>::type;
  TWX twx[C] = {TWX(twb[0])};
  if constexpr (C == 2) twx[1] = TWX(twb[1]);
  fft_reg_lanes<LOGN, C>(v, q0, PT, lds, T, t, g.ax.tw, twx, false);
  if (XMODE == 0) {
    if (!KPRE) load_k();
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        if constexpr (KREAL)
          v[c][m] = make_float2(v[c][m].x * kv[c][m], v[c][m].y * kv[c][m]);
        else
          v[c][m] = cmul(v[c][m], kv[c][m]);
      }
    fft_reg_lanes<LOGN, C>(v, q0, PT, lds, T, t, g.ax.tw, twx, true);
  } else {
    const float2 f = kernel_bin_factor(d, p % g.nzc);
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
      for (int m = 0; m < 8; ++m) v[c][m] = cmul(v[c][m], f);
  }
  if (valid) {
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
      for (int m = 0; m < 8; ++m) *lane_at(row(m), lb[c], 0) = v[c][m];
  }
}

template <class Src, int XMODE, int LOGN, int C>
__global__ __launch_bounds__(1024) STEAM_XREG_WAVES_ATTR void xpass_reg_kernel(Src src) {
  using F = RegFFT<LOGN>;
  constexpr int P = F::P, PT = P / C;
  constexpr int T = (1024 / PT < (C == 1 ? XREG_COLS : XREG_COLS2))
                        ? 1024 / PT : (C == 1 ? XREG_COLS : XREG_COLS2);
  __shared__ float2 lds[T * F::LDSN];
  int wg = xcd_remap(blockIdx.x, gridDim.x);
  const PassDesc& d = src.get(PASS_X, wg);
  // the spectrum's format word is uniform over the launch's instance
  if (XMODE == 0 && d.kfmt && *d.kfmt == KFMT_REAL_EVEN)
    xpass_reg_body<XMODE, LOGN, C, true>(d, wg, lds);
  else
    xpass_reg_body<XMODE, LOGN, C, false>(d, wg, lds);
}

// ---------------------------------------------------------------------------
// zinv: planes -> c2r per block -> overlap-add in LDS -> epilogue
// ---------------------------------------------------------------------------
template <class Src, int MODE, int RS>
__global__ __launch_bounds__(Z_NTH) __attribute__((amdgpu_waves_per_eu(4))) void zinv_kernel(Src src) {
  int wg = blockIdx.x;
  const PassDesc& d = src.get(PASS_ZINV, wg);
  const ConvGeom& g = d.g;
  extern __shared__ float2 smem[];
  const int tid = threadIdx.x;
  const int x = wg / g.nyt_z;
  const int y0 = (wg - x * g.nyt_z) * g.Tz;
  const int ldn = g.Tz + 1;
  const int nz = g.nz;
  float* sr = reinterpret_cast<float*>(smem);
  float* accz = reinterpret_cast<float*>(smem + (g.n2 + 1) * ldn);
  float* comp = accz + g.Tz * nz;
  const bool kahan = d.kahan != 0;
  for (int e = tid; e < g.Tz * nz; e += Z_NTH) {
    accz[e] = 0.f;
    if (kahan) comp[e] = 0.f;
  }
  for (int b = 0; b < g.nblocks; ++b) {
    const float2* Sb = d.S + (int64_t)b * g.nzc * g.nx * g.ny;
    for (int e = tid; e < g.nzc * g.Tz; e += Z_NTH) {
      const int kz = e >> g.logTz;
      const int t = e & (g.Tz - 1);
      const int y = y0 + t;
      smem[kz * ldn + t] = (y < g.ny) ? Sb[((int64_t)kz * g.nx + x) * g.ny + y]
                                      : make_float2(0.f, 0.f);
    }
    __syncthreads();
    // c2r merge: Z[k] = (Y[k] + conj Y[n2-k]) + i conj(W^k) (Y[k] - conj Y[n2-k])
    const int nh = g.n2 / 2 + 1;
    for (int e = tid; e < g.Tz * nh; e += Z_NTH) {
      const int t = e & (g.Tz - 1);
      const int k = e >> g.logTz;
      float2* Lp = smem + t;
      if (k == 0) {
        float2 Y0 = Lp[0], YN = Lp[g.n2 * ldn];
        float2 A = make_float2(Y0.x + YN.x, Y0.y - YN.y);
        float2 B = make_float2(Y0.x - YN.x, Y0.y + YN.y);
        Lp[0] = make_float2(A.x - B.y, A.y + B.x);
      } else {
        const int km = g.n2 - k;
        float2 Yk = Lp[k * ldn], Ym = Lp[km * ldn];
        float2 w = g.wr2c[k];
        float2 A = make_float2(Yk.x + Ym.x, Yk.y - Ym.y);
        float2 B = make_float2(Yk.x - Ym.x, Yk.y + Ym.y);
        float2 cB = cmul(make_float2(w.x, -w.y), B);  // conj(W^k) B
        Lp[k * ldn] = make_float2(A.x - cB.y, A.y + cB.x);  // A + i cB
        if (km != k) {
          float2 A2 = make_float2(Ym.x + Yk.x, Ym.y - Yk.y);
          float2 B2 = make_float2(Ym.x - Yk.x, Ym.y + Yk.y);
          float2 cB2 = cmul(make_float2(-w.x, -w.y), B2);  // conj(W^{n2-k}) = -W^k
          Lp[km * ldn] = make_float2(A2.x - cB2.y, A2.y + cB2.x);
        }
      }
    }
    __syncthreads();
    fft_lines<true, RS, Z_STAGE_CALLS, Z_EM, Z_NTH>(smem, g.az, g.Tz, g.logTz, ldn, 1, true, tid);
    // sample m is output level b L + m; the last cz samples are the circular
    // wrap, levels b L + m - nfft (the previous block's tail)
    const int wrap = g.nfft - g.cz;
    for (int t = 0; t < g.Tz; ++t) {
      for (int m = tid; m < g.nfft; m += Z_NTH) {
        const int zo = b * g.L + (m < wrap ? m : m - g.nfft);
        if (zo >= 0 && zo < nz) {
          const float v = sr[((m >> 1) * ldn + t) * 2 + (m & 1)];
          const int ia = t * nz + zo;
          if (!kahan) {
            accz[ia] += v;
          } else {
            float s = accz[ia];
            float yv = v - comp[ia];
            float tt = s + yv;
            comp[ia] = (tt - s) - yv;
            accz[ia] = tt;
          }
        }
      }
    }
    __syncthreads();
  }
  double acc[2] = {0.0, 0.0};
  const bool xin = (MODE == MODE_FLUX) && x >= d.win[0] && x < d.win[1];
  Pending pd{1.f, 0.f};
  if (MODE == MODE_FLUX) pd = pending_of(d);
  for (int t = 0; t < g.Tz; ++t) {
    const int y = y0 + t;
    if (y >= g.ny) break;
    const int64_t line = ((int64_t)x * g.ny + y) * nz;
    if (MODE == MODE_PLAIN) {
      for (int z = tid; z < nz; z += Z_NTH) d.out[line + z] = accz[t * nz + z];
    } else {
      const bool win = xin && y >= d.win[2] && y < d.win[3];
      for (int z = tid; z < nz; z += Z_NTH) {
        float f = entering_value(pd, d.unit_in ? 1.f : d.out[line + z]);
        f = __fadd_rn(f, accz[t * nz + z]);
        if (f < 0.f) {
          if (win || !d.clip_win) acc[1] += 1.0;
          f = 0.f;
        }
        if (win) acc[0] += (double)f;
        d.out[line + z] = out_value(d, f);
      }
    }
  }
  if (MODE == MODE_FLUX && !d.apply) {
    __shared__ double red[2 * Z_NTH];
    block_sum<2, Z_NTH>(acc, red);
    if (tid == 0) {
      d.part_o[wg * 2 + 0] = acc[0];
      d.part_o[wg * 2 + 1] = acc[1];
    }
  }
}

// ---------------------------------------------------------------------------
// Single-block z passes (the whole z extent fits one transform, nz % 4 == 0,
// the tile's T*nz values fit one batch of ZT_U float4 per thread).  The tile of
// T adjacent lines is one contiguous run of T*nz floats, read with 16-byte
// loads all issued before the first use; twiddles are staged in LDS; zinv
// prefetches the flux tile before its transforms, and reads the trimmed
// convolution straight out of the transform buffer (one block: no overlap-add).
// ---------------------------------------------------------------------------
constexpr int ZT_U = 4;   // float4 groups per thread (tile <= 16 * Z_NTH floats)
constexpr int ZT_S = 8;   // float2 spectral values per thread per batch

// deterministic workgroup sum (fixed shuffle tree per wave, then the waves in
// order); result valid in thread 0; red holds NV * waves doubles
template <int NV, int NT>
__device__ __forceinline__ void block_sum_w(double (&v)[NV], double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = NT >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = v[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    v[i] = s;
  }
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) red[w * NV + i] = v[i];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double s = 0.0;
      for (int k = 0; k < nw; ++k) s += red[k * NV + i];
      v[i] = s;
    }
  }
}

// LDS layout of the tile kernels: transform buffer [(n2+1) * (T+1)] complex,
// then z twiddles [n2], then r2c twiddles [n2/2 + 1].
__device__ __forceinline__ void stage_twiddles(const ConvGeom& g, float2* twz,
                                               float2* twr, int tid) {
  for (int i = tid; i < g.n2; i += Z_NTH) twz[i] = g.az.tw[i];
  for (int i = tid; i <= g.n2 / 2; i += Z_NTH) twr[i] = g.wr2c[i];
}

template <class Src, int MODE, int RS>
__global__ __launch_bounds__(Z_NTH) __attribute__((amdgpu_waves_per_eu(4))) void zfwd_tile_kernel(Src src) {
  int wg = blockIdx.x;
  const PassDesc& d = src.get(PASS_Z, wg);
  const ConvGeom& g = d.g;
  extern __shared__ float2 smem[];
  const int tid = threadIdx.x;
  const int T = g.Tz, logT = g.logTz, ldn = T + 1;
  const int x = wg / g.nyt_z;
  const int y0 = (wg - x * g.nyt_z) * T;
  const int nz = g.nz, n2 = g.n2;
  float2* twz = smem + (n2 + 1) * ldn;
  float2* twr = twz + n2;
  const int tl = min(T, g.ny - y0);
  const int64_t base = ((int64_t)x * g.ny + y0) * nz;
  const int n4 = (tl * nz) >> 2;
  const float* srcp = (MODE == MODE_FLUX) ? d.out : d.in;
  const float4* in4 = reinterpret_cast<const float4*>(srcp + base);
  const float4* nf4 = reinterpret_cast<const float4*>(d.nfield + base);
  float4 fv[ZT_U], nv[ZT_U];
#pragma unroll
  for (int i = 0; i < ZT_U; ++i) {
    const int e = i * Z_NTH + tid;
    if (e < n4) {
      fv[i] = (MODE == MODE_FLUX && d.unit_in) ? make_float4(1.f, 1.f, 1.f, 1.f) : in4[e];
      if (MODE == MODE_FLUX) nv[i] = nf4[e];
    }
  }
  stage_twiddles(g, twz, twr, tid);
  double acc[3] = {0.0, 0.0, 0.0};
  Pending pd{1.f, 0.f};
  bool xin = false;
  if (MODE == MODE_FLUX) {
    pd = pending_of(d);
    xin = x >= d.win[0] && x < d.win[1];
  }
  float4* sk4 = (MODE == MODE_FLUX && d.sk) ? reinterpret_cast<float4*>(d.sk + base) : nullptr;
#pragma unroll
  for (int i = 0; i < ZT_U; ++i) {
    const int e = i * Z_NTH + tid;
    if (e < n4) {
      const int f = e << 2;
      const int t = (int)((unsigned)f / (unsigned)nz);
      const int z = f - t * nz;
      float4 v = fv[i];
      if (MODE == MODE_FLUX) {
        const float4 nn = nv[i];
        const float f0 = entering_value(pd, v.x), f1 = entering_value(pd, v.y);
        const float f2 = entering_value(pd, v.z), f3 = entering_value(pd, v.w);
        v = make_float4(__fmul_rn(nn.x, f0), __fmul_rn(nn.y, f1), __fmul_rn(nn.z, f2),
                        __fmul_rn(nn.w, f3));
        const int y = y0 + t;
        if (xin && y >= d.win[2] && y < d.win[3]) {
          acc[0] += ((double)f0 + (double)f1) + ((double)f2 + (double)f3);
          acc[1] += ((double)fabsf(nn.x) + (double)fabsf(nn.y)) +
                    ((double)fabsf(nn.z) + (double)fabsf(nn.w));
          acc[2] += (double)((nn.x != 0.f) + (nn.y != 0.f) + (nn.z != 0.f) + (nn.w != 0.f));
        }
        if (sk4) sk4[e] = v;
      }
      const int c = z >> 1;
      smem[c * ldn + t] = make_float2(v.x, v.y);
      smem[(c + 1) * ldn + t] = make_float2(v.z, v.w);
    }
  }
  // zero padding beyond nz (lines past the field edge are never stored)
  const int c0 = nz >> 1;
  for (int e = tid; e < (n2 - c0) * T; e += Z_NTH) {
    const int c = c0 + (e >> logT);
    smem[c * ldn + (e & (T - 1))] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  AxisPlan az = g.az;
  az.tw = twz;
  fft_lines<true, RS, Z_STAGE_CALLS, Z_EM, Z_NTH>(smem, az, T, logT, ldn, 1, false, tid);
  // r2c split: X2[k] = (Z[k] + conj Z[n2-k]) - i W^k (Z[k] - conj Z[n2-k])
  const int nh = n2 / 2 + 1;
  for (int e = tid; e < T * nh; e += Z_NTH) {
    const int t = e & (T - 1);
    const int k = e >> logT;
    float2* Lp = smem + t;
    if (k == 0) {
      float2 z0v = Lp[0];
      Lp[0] = make_float2(2.f * (z0v.x + z0v.y), 0.f);
      Lp[n2 * ldn] = make_float2(2.f * (z0v.x - z0v.y), 0.f);
    } else {
      const int km = n2 - k;
      float2 zk = Lp[k * ldn], zm = Lp[km * ldn];
      float2 w = twr[k];
      float2 A = make_float2(zk.x + zm.x, zk.y - zm.y);
      float2 B = make_float2(zk.x - zm.x, zk.y + zm.y);
      float2 wB = cmul(w, B);
      Lp[k * ldn] = make_float2(A.x + wB.y, A.y - wB.x);
      if (km != k) {
        float2 A2 = make_float2(zm.x + zk.x, zm.y - zk.y);
        float2 B2 = make_float2(zm.x - zk.x, zm.y + zk.y);
        float2 wB2 = cmul(make_float2(-w.x, w.y), B2);
        Lp[km * ldn] = make_float2(A2.x + wB2.y, A2.y - wB2.x);
      }
    }
  }
  __syncthreads();
  const int64_t plane = (int64_t)g.nx * g.ny;
  float2* So = d.S + (int64_t)x * g.ny + y0;
  for (int e = tid; e < g.nzc * T; e += Z_NTH) {
    const int kz = e >> logT;
    const int t = e & (T - 1);
    if (t < tl) So[kz * plane + t] = smem[kz * ldn + t];
  }
  if (MODE == MODE_FLUX && !d.apply) {
    __syncthreads();
    block_sum_w<3, Z_NTH>(acc, reinterpret_cast<double*>(smem));
    if (tid == 0) {
      d.part_z[wg * 3 + 0] = acc[0];
      d.part_z[wg * 3 + 1] = acc[1];
      d.part_z[wg * 3 + 2] = acc[2];
    }
  }
}

template <class Src, int MODE, int RS>
__global__ __launch_bounds__(Z_NTH) __attribute__((amdgpu_waves_per_eu(4))) void zinv_tile_kernel(Src src) {
  int wg = blockIdx.x;
  const PassDesc& d = src.get(PASS_ZINV, wg);
  const ConvGeom& g = d.g;
  extern __shared__ float2 smem[];
  const int tid = threadIdx.x;
  const int T = g.Tz, logT = g.logTz, ldn = T + 1;
  const int x = wg / g.nyt_z;
  const int y0 = (wg - x * g.nyt_z) * T;
  const int nz = g.nz, n2 = g.n2;
  float2* twz = smem + (n2 + 1) * ldn;
  float2* twr = twz + n2;
  const int tl = min(T, g.ny - y0);
  const int64_t base = ((int64_t)x * g.ny + y0) * nz;
  const int n4 = (tl * nz) >> 2;
  float4* io4 = reinterpret_cast<float4*>(d.out + base);
  // flux prefetch: in flight while the tile is transformed
  float4 fv[ZT_U];
  if (MODE == MODE_FLUX) {
#pragma unroll
    for (int i = 0; i < ZT_U; ++i) {
      const int e = i * Z_NTH + tid;
      if (e < n4) fv[i] = d.unit_in ? make_float4(1.f, 1.f, 1.f, 1.f) : io4[e];
    }
  }
  const int64_t plane = (int64_t)g.nx * g.ny;
  const float2* Si = d.S + (int64_t)x * g.ny + y0;
  const int ns = g.nzc * T;
  for (int e0 = 0; e0 < ns; e0 += ZT_S * Z_NTH) {
    float2 sv[ZT_S];
#pragma unroll
    for (int i = 0; i < ZT_S; ++i) {
      const int e = e0 + i * Z_NTH + tid;
      const int t = e & (T - 1);
      if (e < ns && t < tl) sv[i] = Si[(e >> logT) * plane + t];
    }
#pragma unroll
    for (int i = 0; i < ZT_S; ++i) {
      const int e = e0 + i * Z_NTH + tid;
      const int t = e & (T - 1);
      if (e < ns) smem[(e >> logT) * ldn + t] = (t < tl) ? sv[i] : make_float2(0.f, 0.f);
    }
  }
  stage_twiddles(g, twz, twr, tid);
  __syncthreads();
  // c2r merge: Z[k] = (Y[k] + conj Y[n2-k]) + i conj(W^k) (Y[k] - conj Y[n2-k])
  const int nh = n2 / 2 + 1;
  for (int e = tid; e < T * nh; e += Z_NTH) {
    const int t = e & (T - 1);
    const int k = e >> logT;
    float2* Lp = smem + t;
    if (k == 0) {
      float2 Y0 = Lp[0], YN = Lp[n2 * ldn];
      float2 A = make_float2(Y0.x + YN.x, Y0.y - YN.y);
      float2 B = make_float2(Y0.x - YN.x, Y0.y + YN.y);
      Lp[0] = make_float2(A.x - B.y, A.y + B.x);
    } else {
      const int km = n2 - k;
      float2 Yk = Lp[k * ldn], Ym = Lp[km * ldn];
      float2 w = twr[k];
      float2 A = make_float2(Yk.x + Ym.x, Yk.y - Ym.y);
      float2 B = make_float2(Yk.x - Ym.x, Yk.y + Ym.y);
      float2 cB = cmul(make_float2(w.x, -w.y), B);
      Lp[k * ldn] = make_float2(A.x - cB.y, A.y + cB.x);
      if (km != k) {
        float2 A2 = make_float2(Ym.x + Yk.x, Ym.y - Yk.y);
        float2 B2 = make_float2(Ym.x - Yk.x, Ym.y + Yk.y);
        float2 cB2 = cmul(make_float2(-w.x, -w.y), B2);
        Lp[km * ldn] = make_float2(A2.x - cB2.y, A2.y + cB2.x);
      }
    }
  }
  __syncthreads();
  AxisPlan az = g.az;
  az.tw = twz;
  fft_lines<true, RS, Z_STAGE_CALLS, Z_EM, Z_NTH>(smem, az, T, logT, ldn, 1, true, tid);
  // output z of line t is real sample m = z + cz of the block transform
  const float* sr = reinterpret_cast<const float*>(smem);
  double acc[2] = {0.0, 0.0};
  Pending pd{1.f, 0.f};
  bool xin = false;
  if (MODE == MODE_FLUX) {
    pd = pending_of(d);
    xin = x >= d.win[0] && x < d.win[1];
  }
#pragma unroll
  for (int i = 0; i < ZT_U; ++i) {
    const int e = i * Z_NTH + tid;
    if (e < n4) {
      const int f = e << 2;
      const int t = (int)((unsigned)f / (unsigned)nz);
      const int m = f - t * nz;
      float c[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) c[j] = sr[(((m + j) >> 1) * ldn + t) * 2 + ((m + j) & 1)];
      if (MODE == MODE_PLAIN) {
        io4[e] = make_float4(c[0], c[1], c[2], c[3]);
      } else {
        const float4 r = fv[i];
        float o[4] = {r.x, r.y, r.z, r.w};
        const int y = y0 + t;
        const bool win = xin && y >= d.win[2] && y < d.win[3];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          float v = __fadd_rn(entering_value(pd, o[j]), c[j]);
          if (v < 0.f) {
            if (win || !d.clip_win) acc[1] += 1.0;
            v = 0.f;
          }
          o[j] = v;
        }
        if (win) acc[0] += ((double)o[0] + (double)o[1]) + ((double)o[2] + (double)o[3]);
        io4[e] = make_float4(out_value(d, o[0]), out_value(d, o[1]), out_value(d, o[2]),
                             out_value(d, o[3]));
      }
    }
  }
  if (MODE == MODE_FLUX && !d.apply) {
    __syncthreads();
    block_sum_w<2, Z_NTH>(acc, reinterpret_cast<double*>(smem));
    if (tid == 0) {
      d.part_o[wg * 2 + 0] = acc[0];
      d.part_o[wg * 2 + 1] = acc[1];
    }
  }
}

// ---------------------------------------------------------------------------
// Register z passes (single block, nz even, n_fft/2 one of the compiled
// z-line plans).  Lines are numbered line = x*ny + y, so a workgroup takes
// ZR_WAVES * LINES consecutive lines and S[k][x][y] = S[k*plane + line].
// Each lane loads the float2 pairs (z = 2n, 2n+1) of its natural indices
// n = q + m P straight from HBM, which is also the packing of the half-length
// complex transform; the r2c split / c2r merge pair k with n2 - k through
// the wave's LDS line.  Partial sums: one slot per workgroup.
// ---------------------------------------------------------------------------
constexpr int ZR_WAVES = STEAM_ZR_WAVES;  // wavefronts per workgroup
constexpr int ZR_NTH = ZR_WAVES * 64;

using ZP32 = ZPlan<32, 8, 8, 4>;
using ZP64 = ZPlan<64, 8, 8, 8>;
using ZP80 = ZPlan<80, 20, 4, 20>;
using ZP96 = ZPlan<96, 24, 8, 12>;
using ZP128 = ZPlan<128, 16, 8, 16>;
// 160 points: 20 values on 8 lanes per line, or (STEAM_ZP160_V 10) 10 values
// on 16 lanes, radix 10 then four radix-2 stages
#ifndef STEAM_ZP160_V
#define STEAM_ZP160_V 20
#endif
#if STEAM_ZP160_V == 10
using ZP160 = ZPlan<160, 10, 10, 2, 2, 2, 2>;
#else
using ZP160 = ZPlan<160, 20, 4, 4, 10>;
#endif
// 72 and 144 points (z lengths 144 and 288: a 13-tap kernel over 128 and
// 256 levels pads to them instead of 160 and 320): 18 values per lane
using ZP72 = ZPlan<72, 18, 6, 6, 2>;
using ZP144 = ZPlan<144, 18, 6, 6, 2, 2>;
using ZP192 = ZPlan<192, 24, 8, 24>;
using ZP256 = ZPlan<256, 16, 16, 16>;
using ZP320 = ZPlan<320, 20, 4, 4, 20>;
using ZP512 = ZPlan<512, 16, 8, 8, 8>;

// call f(ZPlan{}) for a compiled z-line length; false if there is none
template <class F>
inline bool with_zplan(int n2, F&& f) {
  switch (n2) {
    case 32: f(ZP32{}); return true;
    case 64: f(ZP64{}); return true;
    case 72: f(ZP72{}); return true;
    case 80: f(ZP80{}); return true;
    case 96: f(ZP96{}); return true;
    case 128: f(ZP128{}); return true;
    case 144: f(ZP144{}); return true;
    case 160: f(ZP160{}); return true;
    case 192: f(ZP192{}); return true;
    case 256: f(ZP256{}); return true;
    case 320: f(ZP320{}); return true;
    case 512: f(ZP512{}); return true;
    default: return false;
  }
}
inline bool zplan_lines(int n2, int& lines) {
  return with_zplan(n2, [&](auto zp) { lines = decltype(zp)::LINES * ZR_WAVES; });
}

// ---------------------------------------------------------------------------
// Mixed-radix register x/y passes: the x/y lengths that are not a power of
// two but have a compiled plan (2^a 3^b 5^c lengths such as a batch's 96 and
// 192 or a served tile's padded 80, 144, 288), from HBM to registers and
// back as the radix-8 register passes do.  A line of N points is held by
// P = N / V lanes, lane q holding the V values q + m P.
// ---------------------------------------------------------------------------
using XP288 = ZPlan<288, 18, 6, 6, 2, 2, 2>;
using XP384 = ZPlan<384, 24, 8, 8, 6>;
// 400 points (a 384-cell served tile plus its halo): 50 values on 8 lanes,
// the radix-2 stages first (fewer live registers than 10, 10, 2, 2)
using XP400 = ZPlan<400, 50, 2, 2, 10, 10>;

// call f(plan) for an x/y length with a mixed-radix register plan
template <class F>
inline bool with_xyplan(int n, F&& f) {
  switch (n) {
    case 32: f(ZP32{}); return true;
    case 80: f(ZP80{}); return true;
    case 96: f(ZP96{}); return true;
    case 144: f(ZP144{}); return true;
    case 160: f(ZP160{}); return true;
    case 192: f(ZP192{}); return true;
    case 288: f(XP288{}); return true;
    case 320: f(ZP320{}); return true;
    case 384: f(XP384{}); return true;
    case 400: f(XP400{}); return true;
    default: return false;
  }
}
// threads per workgroup of a plan's x/y passes: 256, or 128 where 256 / P
// wave-private lines would take more than 64 KB of LDS (two workgroups per CU)
template <class XP>
__host__ __device__ constexpr int xy_nth() {
  return (256 / XP::P) * XP::PITCH * 8 > 64 * 1024 ? 128 : 256;
}
// lanes per line and threads per workgroup of an x/y length's plan (0: none)
inline int xyplan_lanes(int n) {
  int p = 0;
  with_xyplan(n, [&](auto xp) { p = decltype(xp)::P; });
  return p;
}
inline int xyplan_threads(int n) {
  int t = 0;
  with_xyplan(n, [&](auto xp) { t = xy_nth<decltype(xp)>(); });
  return t;
}

// y pass: 256 / P contiguous y lines per workgroup, each wave-private
// (a wave holds 64 / P of them), lines in LDS at the z plans' pitch
template <class Src, class XP>
__global__ __launch_bounds__(xy_nth<XP>()) void ypass_mr_kernel(Src src, int inv) {
  constexpr int NTH = xy_nth<XP>();
  constexpr int N = XP::N, V = XP::V, P = XP::P, LPW = NTH / P;
  __shared__ float2 lds[LPW * XP::PITCH + N];
  int wg = blockIdx.x;
  const PassDesc& d = src.get(PASS_Y, wg);
  const ConvGeom& g = d.g;
  const int tid = threadIdx.x;
  const int q = tid % P, lw = tid / P;
  float2* tw = lds + LPW * XP::PITCH;
  for (int i = tid; i < N; i += NTH) tw[i] = g.ay.tw[i];
  __syncthreads();
  const int planes = (d.mode == MODE_KERNEL) ? g.nzc : g.nblocks * g.nzc;
  const int64_t lines = (int64_t)planes * g.nx;
  const int64_t line = (int64_t)wg * LPW + lw;
  const bool valid = line < lines;
  float2* Sl = d.S + (valid ? line : 0) * N;
  const unsigned lo = (unsigned)q * 8u;
  float2 v[V];
#pragma unroll
  for (int m = 0; m < V; ++m)
    v[m] = valid ? *lane_at(Sl, lo, m * P * 8) : make_float2(0.f, 0.f);
  float2* line_lds = lds + lw * XP::PITCH;
  if (inv)
    XP::template fft_xy<true, false>(v, q, line_lds, 1, 0, tw);
  else
    XP::template fft_xy<false, false>(v, q, line_lds, 1, 0, tw);
  if (valid) {
#pragma unroll
    for (int m = 0; m < V; ++m) *lane_at(Sl, lo, m * P * 8) = v[m];
  }
}

// x pass: T = 256 / P adjacent y columns per workgroup (thread t % T is the
// column, so a row access is a T*8-byte run); column lines in LDS at
// zpad(i) * (T + 1) + t.  The kernel spectrum is in the pass's column tiles
// [kz][y/T][x][T], or real and x-even [kz][y/T][nx/2 + 1][T] (KREAL: bin x
// takes the value of min(x, nx - x); for x = q + m P that is x for m < V/2
// and (V - m) P - q from m = V/2 on).  XMODE 0: forward, * spectrum,
// inverse; 1: forward, * the kernel-build scale.
template <class XP>
struct XMrLds {
  static constexpr int T = xy_nth<XP>() / XP::P;
  // column stride: T + 1 (the lane groups of a wave start in different
  // banks) unless that leaves room for only one workgroup per CU
  static constexpr int LS = ((XP::N + XP::N / 8) * (T + 1) + XP::N) * 8 > 80 * 1024 ? T : T + 1;
  static constexpr int LINE = (XP::N + XP::N / 8) * LS;
};

template <int XMODE, class XP, bool KREAL>
__device__ __forceinline__ void xpass_mr_body(const PassDesc& d, int wg, float2* lds,
                                              const float2* tw) {
  constexpr int N = XP::N, V = XP::V, P = XP::P, T = XMrLds<XP>::T, LS = XMrLds<XP>::LS;
  const ConvGeom& g = d.g;
  const int tid = threadIdx.x;
  const int t = tid % T, q = tid / T;
  const int nyt = (g.ny + T - 1) / T;
  const int p = wg / nyt;
  const int y = (wg - p * nyt) * T + t;
  const bool valid = y < g.ny;
  const int ny = g.ny;
  // value m of the lane is row q + m P: the plane base (uniform) plus a
  // 32-bit byte offset (a plane is under 4 GB), so no access needs a 64-bit
  // row pointer of its own
  float2* Sb = d.S + (int64_t)p * N * ny;
  const unsigned lb = (unsigned)(q * ny + (valid ? y : 0)) * 8u;
  const unsigned rs = (unsigned)(P * ny) * 8u;
  float2 v[V];
#pragma unroll
  for (int m = 0; m < V; ++m)
    v[m] = valid ? *lane_at(Sb, lb + (unsigned)m * rs, 0) : make_float2(0.f, 0.f);
  using KV = typename std::conditional<KREAL, float, float2>::type;
  KV kv[V];
  // the spectrum loads go out with the line's (up to 24 values per lane) or
  // after the forward transform (32: both sets would not fit the registers)
  constexpr bool KPRE = V <= 24;
  auto load_k = [&]() {
    const int64_t tile = (int64_t)(p % g.nzc) * nyt + (wg - p * nyt);
    if constexpr (KREAL) {
      const float* Kb = reinterpret_cast<const float*>(d.K) + tile * (N / 2 + 1) * T;
      const unsigned klo = (unsigned)(q * T + t) * 4u;
      const unsigned khi = (unsigned)((P - q) * T + t) * 4u;
#pragma unroll
      for (int m = 0; m < V; ++m)
        kv[m] = !valid ? 0.f
                : m < V / 2 ? *lane_at(Kb + m * P * T, klo, 0)
                            : *lane_at(Kb + (V - 1 - m) * P * T, khi, 0);
    } else {
      const float2* Kb = d.K + tile * N * T;
      const unsigned klo = (unsigned)(q * T + t) * 8u;
#pragma unroll
      for (int m = 0; m < V; ++m)
        kv[m] = valid ? *lane_at(Kb + m * P * T, klo, 0) : make_float2(0.f, 0.f);
    }
  };
  if (XMODE == 0 && KPRE) load_k();
  XP::template fft_xy<false, true>(v, q, lds, LS, t, tw);
  if (XMODE == 0) {
    // the inverse reads its twiddles from a second copy of the table: the
    // compiler cannot then keep the forward transform's twiddles live in
    // registers for it (with 50 values per lane those would not fit)
    constexpr bool REDERIVE = V > 24;
    const float2* twi = REDERIVE ? tw + N : tw;
    if constexpr (!KPRE && !KREAL) {
      // the complex spectrum above 24 values per lane: loaded and applied
      // in chunks of 10 (all of it at once would not fit the registers)
      const int64_t tile = (int64_t)(p % g.nzc) * nyt + (wg - p * nyt);
      const float2* Kb = d.K + tile * N * T;
      const unsigned klo = (unsigned)(q * T + t) * 8u;
#pragma unroll
      for (int m0 = 0; m0 < V; m0 += 10) {
        float2 kc[10];
#pragma unroll
        for (int i = 0; i < 10 && m0 + i < V; ++i)
          kc[i] = valid ? *lane_at(Kb + (m0 + i) * P * T, klo, 0) : make_float2(0.f, 0.f);
#pragma unroll
        for (int i = 0; i < 10 && m0 + i < V; ++i) v[m0 + i] = cmul(v[m0 + i], kc[i]);
        __builtin_amdgcn_sched_barrier(0);
      }
    } else {
      if (!KPRE) load_k();
#pragma unroll
      for (int m = 0; m < V; ++m) {
        if constexpr (KREAL)
          v[m] = make_float2(v[m].x * kv[m], v[m].y * kv[m]);
        else
          v[m] = cmul(v[m], kv[m]);
      }
    }
    // likewise its LDS addresses: the lane's q and t pass through an empty
    // asm, so they are recomputed rather than held from the forward transform
    int qi = q, ti = t;
    if constexpr (REDERIVE) asm volatile("" : "+v"(qi), "+v"(ti));
    XP::template fft_xy<true, true>(v, qi, lds, LS, ti, twi);
  } else {
    const float2 f = kernel_bin_factor(d, p % g.nzc);
#pragma unroll
    for (int m = 0; m < V; ++m) v[m] = cmul(v[m], f);
  }
  if (valid) {
#pragma unroll
    for (int m = 0; m < V; ++m) *lane_at(Sb, lb + (unsigned)m * rs, 0) = v[m];
  }
}


template <class Src, int XMODE, class XP>
__global__ __launch_bounds__(xy_nth<XP>()) __attribute__((amdgpu_waves_per_eu(2)))
void xpass_mr_kernel(Src src) {
  constexpr int N = XP::N;
  constexpr int NT = XP::V > 24 ? 2 * N : N;  // table copies (xpass_mr_body)
  __shared__ float2 lds[XMrLds<XP>::LINE + NT];
  int wg = xcd_remap(blockIdx.x, gridDim.x);
  const PassDesc& d = src.get(PASS_X, wg);
  float2* tw = lds + XMrLds<XP>::LINE;
  for (int i = threadIdx.x; i < NT; i += xy_nth<XP>()) tw[i] = d.g.ax.tw[i < N ? i : i - N];
  __syncthreads();
  // the spectrum's format word is uniform over the launch's instance
  if (XMODE == 0 && d.kfmt && *d.kfmt == KFMT_REAL_EVEN)
    xpass_mr_body<XMODE, XP, true>(d, wg, lds, tw);
  else
    xpass_mr_body<XMODE, XP, false>(d, wg, lds, tw);
}

// The workgroup's plane tile: planes k = 0..N of its LW = ZR_WAVES * LINES
// consecutive lines, tile[k * (LW + 1) + l], in the space of the wave lines.
template <class ZP>
struct ZTile {
  static constexpr int LW = ZR_WAVES * ZP::LINES;
  static constexpr int TP = LW + 1;
  static constexpr int NF = (ZP::N + 1) * LW;
  static constexpr int ITER = (NF + ZR_NTH - 1) / ZR_NTH;
  // LDS of a workgroup: the waves' lines, or the plane tile where larger
  static constexpr int LDS = (ZP::N + 1) * TP > LW * ZP::PITCH ? (ZP::N + 1) * TP
                                                                : LW * ZP::PITCH;
};

// W_2N^k (k < 2N) from the r2c table W_2N^k (k <= N): W_2N^{k+N} = -W_2N^k
__device__ __forceinline__ void stage_tw2(const ConvGeom& g, float2* tw2, int N) {
  for (int i = threadIdx.x; i < 2 * N; i += ZR_NTH) {
    const float2 w = g.wr2c[i < N ? i : i - N];
    tw2[i] = i < N ? w : make_float2(-w.x, -w.y);
  }
}

template <int NV>
__device__ __forceinline__ void zr_block_sum(double (&v)[NV], double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = v[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    v[i] = s;
  }
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) red[w * NV + i] = v[i];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < ZR_WAVES; ++k) s += red[k * NV + i];
      v[i] = s;
    }
}

// The z pair (2n, 2n+1) at lane offset lb plus cb bytes: one 8-byte access
// when the field has an even number of levels (every line start is then
// 8-byte aligned), two 4-byte ones when it is odd.  The _top forms are for
// a pair that may reach past the top level nz: with an odd count its second
// level reads 0 and is not written.
template <bool ODD>
__device__ __forceinline__ float2 ld_pair(const float* base, unsigned lb, int cb) {
  const float* p = lane_at(base, lb, cb);
  if constexpr (!ODD) return *reinterpret_cast<const float2*>(p);
  return make_float2(p[0], p[1]);
}
template <bool ODD>
__device__ __forceinline__ void st_pair(float* base, unsigned lb, int cb, float2 v) {
  float* p = lane_at(base, lb, cb);
  if constexpr (!ODD) {
    *reinterpret_cast<float2*>(p) = v;
  } else {
    p[0] = v.x;
    p[1] = v.y;
  }
}
template <bool ODD>
__device__ __forceinline__ float2 ld_pair_top(const float* base, unsigned lb, int cb, int z0,
                                              int nz) {
  if constexpr (!ODD) return ld_pair<false>(base, lb, cb);
  const float* p = lane_at(base, lb, cb);
  return make_float2(p[0], z0 + 1 < nz ? p[1] : 0.f);
}
template <bool ODD>
__device__ __forceinline__ void st_pair_top(float* base, unsigned lb, int cb, int z0, int nz,
                                            float2 v) {
  if constexpr (!ODD) {
    st_pair<false>(base, lb, cb, v);
  } else {
    float* p = lane_at(base, lb, cb);
    p[0] = v.x;
    if (z0 + 1 < nz) p[1] = v.y;
  }
}

// Lane layout shared by the register z passes.  A workgroup holds LW
// consecutive lines from line0; lane (w, t, q) works on line l = w LINES + t
// of them, natural indices q + m P.  Pair m (levels 2(q + mP), +1) lies
// below nz on every lane when m < m_full, a count uniform over the launch,
// so those pairs need neither a bound check nor a lane mask; lines and level
// offsets inside one workgroup fit 32 bits (x/y extents <= 4096, nz <= 16384).
template <class ZP>
struct ZLanes {
  static constexpr int LW = ZTile<ZP>::LW;
  int t, q, l, line0, nlines, m_full;
  bool valid;
  unsigned lb;  // byte offset of the lane's pair 0 (level 2q of line l) from line0
  __device__ __forceinline__ ZLanes(const ConvGeom& g, int wg, bool full) {
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    t = lane / ZP::P;
    q = lane % ZP::P;
    l = w * ZP::LINES + t;
    line0 = wg * LW;
    nlines = g.nx * g.ny;
    valid = full || line0 + l < nlines;
    m_full = min(ZP::V, g.nz / (2 * ZP::P));
    lb = valid ? (unsigned)(l * g.nz + 2 * q) * 4u : 0u;
  }
  // byte offset of pair m from the lane's pair 0
  static constexpr int pair_bytes(int m) { return 8 * m * ZP::P; }
  // the lane's line is inside the x/y window (32-bit division)
  __device__ __forceinline__ bool in_window(const ConvGeom& g, const int (&win)[4]) const {
    const unsigned li = (unsigned)(line0 + l);
    const int x = (int)(li / (unsigned)g.ny);
    const int y = (int)(li - (unsigned)x * (unsigned)g.ny);
    return valid && x >= win[0] && x < win[1] && y >= win[2] && y < win[3];
  }
};

// XF: the served flux options (a unit entering flux, the apply mode's stored
// rescale) are tested at run time; without it they compile away, so the
// other advances run the kernels without them.
template <int MODE, class ZP, bool ODD, bool FULL, bool XF>
__device__ __forceinline__ void zfwd_reg_body(const PassDesc& d, int wg, float2* lds_lines,
                                              float2* tw2, double* red) {
  constexpr int N = ZP::N, V = ZP::V, P = ZP::P, LPW = ZP::LINES;
  using TL = ZTile<ZP>;
  const ConvGeom& g = d.g;
  const int tid = threadIdx.x, w = tid >> 6;
  const ZLanes<ZP> ln(g, wg, FULL);
  const int q = ln.q, nz = g.nz;
  // uniform bases, 32-bit lane offsets
  const int64_t base = (int64_t)ln.line0 * nz;
  const float* srcp = ((MODE == MODE_FLUX) ? d.out : d.in) + base;
  const float* nfp = d.nfield + base;
  const unsigned lb = ln.lb;
  float2 v[V], nv[V];
  // a unit entering flux (uniform over the launch): the flux is not read
  const bool unit = XF && MODE == MODE_FLUX && d.unit_in;
  if (unit) {
#pragma unroll
    for (int m = 0; m < V; ++m) {
      const int z0 = 2 * (q + m * P);
      const int cb = ZLanes<ZP>::pair_bytes(m);
      // levels past the top stay zero: they are the transform's padding
      v[m] = make_float2(0.f, 0.f);
      nv[m] = make_float2(0.f, 0.f);
      if (m < ln.m_full) {
        if (ln.valid) {
          v[m] = make_float2(1.f, 1.f);
          nv[m] = ld_pair<ODD>(nfp, lb, cb);
        }
      } else if (ln.valid && z0 < nz) {
        v[m] = make_float2(1.f, 1.f);
        nv[m] = ld_pair_top<ODD>(nfp, lb, cb, z0, nz);
      }
    }
  } else {
#pragma unroll
    for (int m = 0; m < V; ++m) {
      const int z0 = 2 * (q + m * P);
      const int cb = ZLanes<ZP>::pair_bytes(m);
      v[m] = make_float2(0.f, 0.f);
      nv[m] = make_float2(0.f, 0.f);
      if (m < ln.m_full) {
        if (ln.valid) {
          v[m] = ld_pair<ODD>(srcp, lb, cb);
          if (MODE == MODE_FLUX) nv[m] = ld_pair<ODD>(nfp, lb, cb);
        }
      } else if (ln.valid && z0 < nz) {
        v[m] = ld_pair_top<ODD>(srcp, lb, cb, z0, nz);
        if (MODE == MODE_FLUX) nv[m] = ld_pair_top<ODD>(nfp, lb, cb, z0, nz);
      }
    }
  }
  stage_tw2(g, tw2, N);
  double acc0 = 0.0, acc1 = 0.0;
  int nonzero = 0;
  if (MODE == MODE_FLUX) {
    const Pending pd = pending_of(d);
    const bool win = ln.in_window(g, d.win);
    float* skp = d.sk ? d.sk + base : nullptr;
#pragma unroll
    for (int m = 0; m < V; ++m) {
      const int z0 = 2 * (q + m * P);
      const bool full = m < ln.m_full;
      if (full ? ln.valid : (ln.valid && z0 < nz)) {
        // with an odd level count the last pair's second level is past the
        // top: it holds no cell and enters no sum
        const bool top = ODD && !full && z0 + 1 >= nz;
        const float f0 = entering_value(pd, v[m].x);
        const float f1 = top ? 0.f : entering_value(pd, v[m].y);
        const float2 pr = make_float2(__fmul_rn(nv[m].x, f0), __fmul_rn(nv[m].y, f1));
        if (win) {
          acc0 += (double)f0 + (double)f1;
          acc1 += (double)fabsf(nv[m].x) + (double)fabsf(nv[m].y);
          nonzero += (nv[m].x != 0.f) + (nv[m].y != 0.f);
        }
        if (skp) {
          const int cb = ZLanes<ZP>::pair_bytes(m);
          if (full)
            st_pair<ODD>(skp, lb, cb, pr);
          else
            st_pair_top<ODD>(skp, lb, cb, z0, nz, pr);
        }
        v[m] = pr;
      }
    }
  }
  __syncthreads();
  float2* L = lds_lines + ln.l * ZP::PITCH;
  ZP::template fft<false>(v, q, L, tw2);
  // r2c split: X2[k] = (Z[k] + conj Z[N-k]) - i W_2N^k (Z[k] - conj Z[N-k])
#pragma unroll
  for (int m = 0; m < V; ++m) L[zpad(q + m * P)] = v[m];
  wave_lds_sync();
  float2 nyq = make_float2(0.f, 0.f);
#pragma unroll
  for (int m = 0; m < V; ++m) {
    const int k = q + m * P;
    const float2 zk = v[m];
    if (m == 0 && q == 0) {
      v[m] = make_float2(2.f * (zk.x + zk.y), 0.f);
      nyq = make_float2(2.f * (zk.x - zk.y), 0.f);
    } else {
      const float2 zm = L[zpad(N - k)];
      const float2 wk = tw2[k];
      const float2 A = make_float2(zk.x + zm.x, zk.y - zm.y);
      const float2 B = make_float2(zk.x - zm.x, zk.y + zm.y);
      const float2 wB = cmul(wk, B);
      v[m] = make_float2(A.x + wB.y, A.y - wB.x);
    }
  }
  // planes k = 0..N of the workgroup's lines, one plane run per wave access
  float2* tile = lds_lines;
  __syncthreads();
#pragma unroll
  for (int m = 0; m < V; ++m) tile[(q + m * P) * TL::TP + ln.l] = v[m];
  if (q == 0) tile[N * TL::TP + ln.l] = nyq;
  __syncthreads();
  const int64_t plane = (int64_t)ln.nlines;
  float2* So = d.S + ln.line0;
  constexpr int KSTEP = ZR_NTH / TL::LW;  // planes per iteration
  const int kl = tid / TL::LW, c = tid % TL::LW;
  // (kl plane + c) complex, below 2^30 bytes for planes of at most 2^24 cells
  const unsigned pb = (unsigned)(kl * ln.nlines + c) * 8u;
  const bool cvalid = FULL || ln.line0 + c < ln.nlines;
#pragma unroll
  for (int i = 0; i < TL::ITER; ++i) {
    const int k = kl + i * KSTEP;
    if ((i + 1) * ZR_NTH <= TL::NF || k <= N) {
      if (cvalid) *lane_at(So + (int64_t)(i * KSTEP) * plane, pb, 0) = tile[k * TL::TP + c];
    }
  }
  if (MODE == MODE_FLUX && !d.apply) {
    double acc[3] = {acc0, acc1, (double)nonzero};
    zr_block_sum<3>(acc, red);
    if (tid == 0) {
      d.part_z[wg * 3 + 0] = acc[0];
      d.part_z[wg * 3 + 1] = acc[1];
      d.part_z[wg * 3 + 2] = acc[2];
    }
  }
  (void)w;
  (void)LPW;
}

template <class Src, int MODE, class ZP, bool ODD, bool XF = false>
__global__ __launch_bounds__(ZR_NTH) void zfwd_reg_kernel(Src src) {
  constexpr int N = ZP::N;
  __shared__ float2 lds_lines[ZTile<ZP>::LDS];
  __shared__ float2 tw2[2 * N];
  __shared__ double red[ZR_WAVES * 3];
  int wg = blockIdx.x;
  const PassDesc& d = src.get(PASS_Z, wg);
  const ConvGeom& g = d.g;
  // a workgroup whose lines all exist takes the unmasked body
  if ((int64_t)(wg + 1) * ZTile<ZP>::LW <= (int64_t)g.nx * g.ny)
    zfwd_reg_body<MODE, ZP, ODD, true, XF>(d, wg, lds_lines, tw2, red);
  else
    zfwd_reg_body<MODE, ZP, ODD, false, XF>(d, wg, lds_lines, tw2, red);
}

template <int MODE, class ZP, bool ODD, bool FULL, bool XF>
__device__ __forceinline__ void zinv_reg_body(const PassDesc& d, int wg, float2* lds_lines,
                                              float2* tw2, double* red) {
  constexpr int N = ZP::N, V = ZP::V, P = ZP::P;
  using TL = ZTile<ZP>;
  const ConvGeom& g = d.g;
  const int tid = threadIdx.x;
  const ZLanes<ZP> ln(g, wg, FULL);
  const int q = ln.q, nz = g.nz;
  // planes k = 0..N of the workgroup's lines, one plane run per wave access
  const int64_t plane = (int64_t)ln.nlines;
  const float2* Si = d.S + ln.line0;
  constexpr int KSTEP = ZR_NTH / TL::LW;
  const int kl = tid / TL::LW, c = tid % TL::LW;
  // (kl plane + c) complex, below 2^30 bytes for planes of at most 2^24 cells
  const unsigned pb = (unsigned)(kl * ln.nlines + c) * 8u;
  const bool cvalid = FULL || ln.line0 + c < ln.nlines;
  float2* tile = lds_lines;
  {
    float2 u[TL::ITER];
#pragma unroll
    for (int i = 0; i < TL::ITER; ++i) {
      const int k = kl + i * KSTEP;
      const bool in = (i + 1) * ZR_NTH <= TL::NF || k <= N;
      u[i] = (in && cvalid) ? *lane_at(Si + (int64_t)(i * KSTEP) * plane, pb, 0)
                            : make_float2(0.f, 0.f);
    }
    stage_tw2(g, tw2, N);
#pragma unroll
    for (int i = 0; i < TL::ITER; ++i) {
      const int k = kl + i * KSTEP;
      if ((i + 1) * ZR_NTH <= TL::NF || k <= N) tile[k * TL::TP + c] = u[i];
    }
  }
  __syncthreads();
  float2 v[V];
  float2 yn = make_float2(0.f, 0.f);
#pragma unroll
  for (int m = 0; m < V; ++m) v[m] = tile[(q + m * P) * TL::TP + ln.l];
  if (q == 0) yn = tile[N * TL::TP + ln.l];
  // output levels z = 2n, 2n+1 of natural index n (the kernel spectrum
  // carries the 'same' trim shift); with an odd nz the last pair's second
  // level is past the top and is neither written nor counted
  const int64_t base = (int64_t)ln.line0 * nz;
  float* outp = d.out + base;
  const unsigned lb = ln.lb;
  float2 fl[V];
  // STEAM_ZR_PREFETCH 1: the flux before the transform, 0: after it
  auto load_flux = [&](int m) {
    const int z0 = 2 * (q + m * P);
    const int cb = ZLanes<ZP>::pair_bytes(m);
    if (m < ln.m_full)
      fl[m] = ln.valid ? ld_pair<ODD>(outp, lb, cb) : make_float2(0.f, 0.f);
    else
      fl[m] = (ln.valid && z0 < nz) ? ld_pair_top<ODD>(outp, lb, cb, z0, nz)
                                    : make_float2(0.f, 0.f);
  };
  // a unit entering flux (uniform over the launch): the flux is not read
  const bool unit = XF && MODE == MODE_FLUX && d.unit_in;
  if (unit) {
#pragma unroll
    for (int m = 0; m < V; ++m) fl[m] = make_float2(1.f, 1.f);
  }
  if (MODE == MODE_FLUX && STEAM_ZR_PREFETCH && !unit) {
#pragma unroll
    for (int m = 0; m < V; ++m) load_flux(m);
  }
  __syncthreads();
  float2* L = lds_lines + ln.l * ZP::PITCH;
#pragma unroll
  for (int m = 0; m < V; ++m) L[zpad(q + m * P)] = v[m];
  wave_lds_sync();
  // c2r merge: Z[k] = (Y[k] + conj Y[N-k]) + i conj(W_2N^k) (Y[k] - conj Y[N-k])
#pragma unroll
  for (int m = 0; m < V; ++m) {
    const int k = q + m * P;
    const float2 yk = v[m];
    const float2 ym = (m == 0 && q == 0) ? yn : L[zpad(N - k)];
    const float2 wk = tw2[k];
    const float2 A = make_float2(yk.x + ym.x, yk.y - ym.y);
    const float2 B = make_float2(yk.x - ym.x, yk.y + ym.y);
    const float2 cB = cmul(make_float2(wk.x, -wk.y), B);
    v[m] = make_float2(A.x - cB.y, A.y + cB.x);
  }
  wave_lds_sync();
  ZP::template fft<true>(v, q, L, tw2);
  if (MODE == MODE_PLAIN) {
#pragma unroll
    for (int m = 0; m < V; ++m) {
      const int z0 = 2 * (q + m * P);
      const int cb = ZLanes<ZP>::pair_bytes(m);
      if (m < ln.m_full) {
        if (ln.valid) st_pair<ODD>(outp, lb, cb, v[m]);
      } else if (ln.valid && z0 < nz) {
        st_pair_top<ODD>(outp, lb, cb, z0, nz, v[m]);
      }
    }
    return;
  }
  if (!STEAM_ZR_PREFETCH && !unit) {
#pragma unroll
    for (int m = 0; m < V; ++m) load_flux(m);
  }
  const Pending pd = pending_of(d);
  if (XF && d.out_scale) {
    // apply mode (uniform over the launch): the recorded rescale, stored as
    // the post pass would store it; no partial sums
    if (d.crop_out) {
      // a crop (uniform over the launch): the lane's line, if inside the
      // crop's x/y extent, stores the levels inside its z extent, each as
      // a float (the crop's lines need not start on an even level)
      const unsigned li = (unsigned)(ln.line0 + ln.l);
      const int cx = (int)(li / (unsigned)g.ny) - d.crop_o[0];
      const int cy = (int)(li % (unsigned)g.ny) - d.crop_o[1];
      const bool cl = ln.valid && cx >= 0 && cx < d.crop_n[0] && cy >= 0 && cy < d.crop_n[1];
      const int zc0 = d.crop_o[2], zc1 = d.crop_o[2] + d.crop_n[2];
      float* co = d.crop_out + ((int64_t)cx * d.crop_n[1] + cy) * d.crop_n[2] - zc0;
#pragma unroll
      for (int m = 0; m < V; ++m) {
        const int z0 = 2 * (q + m * P);
        float o0 = __fadd_rn(entering_value(pd, fl[m].x), v[m].x);
        float o1 = __fadd_rn(entering_value(pd, fl[m].y), v[m].y);
        o0 = (o0 < 0.f) ? 0.f : o0;
        o1 = (o1 < 0.f) ? 0.f : o1;
        if (cl && z0 >= zc0 && z0 < zc1) co[z0] = out_value(d, o0);
        if (cl && z0 + 1 >= zc0 && z0 + 1 < zc1) co[z0 + 1] = out_value(d, o1);
      }
      return;
    }
#pragma unroll
    for (int m = 0; m < V; ++m) {
      const int z0 = 2 * (q + m * P);
      const bool full = m < ln.m_full;
      float o0 = __fadd_rn(entering_value(pd, fl[m].x), v[m].x);
      float o1 = __fadd_rn(entering_value(pd, fl[m].y), v[m].y);
      if (ODD && !full && z0 + 1 >= nz) o1 = 0.f;
      o0 = (o0 < 0.f) ? 0.f : o0;
      o1 = (o1 < 0.f) ? 0.f : o1;
      if (full ? ln.valid : (ln.valid && z0 < nz)) {
        const int cb = ZLanes<ZP>::pair_bytes(m);
        const float2 ov = make_float2(out_value(d, o0), out_value(d, o1));
        if (full)
          st_pair<ODD>(outp, lb, cb, ov);
        else
          st_pair_top<ODD>(outp, lb, cb, z0, nz, ov);
      }
    }
    return;
  }
  const bool win = ln.in_window(g, d.win);
  const bool count_clip = win || !d.clip_win;
  double acc0 = 0.0;
  int nclip = 0;
#pragma unroll
  for (int m = 0; m < V; ++m) {
    const int z0 = 2 * (q + m * P);
    const bool full = m < ln.m_full;
    float o0 = __fadd_rn(entering_value(pd, fl[m].x), v[m].x);
    float o1 = __fadd_rn(entering_value(pd, fl[m].y), v[m].y);
    if (ODD && !full && z0 + 1 >= nz) o1 = 0.f;
    const int clipped = (o0 < 0.f) + (o1 < 0.f);
    o0 = (o0 < 0.f) ? 0.f : o0;
    o1 = (o1 < 0.f) ? 0.f : o1;
    if (full ? ln.valid : (ln.valid && z0 < nz)) {
      if (count_clip) nclip += clipped;
      if (win) acc0 += (double)o0 + (double)o1;
      const int cb = ZLanes<ZP>::pair_bytes(m);
      if (full)
        st_pair<ODD>(outp, lb, cb, make_float2(o0, o1));
      else
        st_pair_top<ODD>(outp, lb, cb, z0, nz, make_float2(o0, o1));
    }
  }
  if (!d.apply) {
    double acc[2] = {acc0, (double)nclip};
    zr_block_sum<2>(acc, red);
    if (tid == 0) {
      d.part_o[wg * 2 + 0] = acc[0];
      d.part_o[wg * 2 + 1] = acc[1];
    }
  }
}

template <class Src, int MODE, class ZP, bool ODD, bool XF = false>
__global__ __launch_bounds__(ZR_NTH) STEAM_ZINV_ATTR void zinv_reg_kernel(Src src) {
  constexpr int N = ZP::N;
  __shared__ float2 lds_lines[ZTile<ZP>::LDS];
  __shared__ float2 tw2[2 * N];
  __shared__ double red[ZR_WAVES * 2];
  int wg = blockIdx.x;
  const PassDesc& d = src.get(PASS_ZINV, wg);
  const ConvGeom& g = d.g;
  if ((int64_t)(wg + 1) * ZTile<ZP>::LW <= (int64_t)g.nx * g.ny)
    zinv_reg_body<MODE, ZP, ODD, true, XF>(d, wg, lds_lines, tw2, red);
  else
    zinv_reg_body<MODE, ZP, ODD, false, XF>(d, wg, lds_lines, tw2, red);
}

// ---------------------------------------------------------------------------
// finalize: fixed-order reduction of the partials into the ledger
// ---------------------------------------------------------------------------
// threads of the finalize workgroup: the partials of a large field (8192
// workgroups at 512 x 512) are read eight per thread, and the workgroup
// sum is a shuffle tree per wave plus one pass over the waves
constexpr int FIN_NTH = 1024;

// The advance's device ledger from its sums: the means, the rescale
// entering / realized (or flatten to the entering mean when nothing is
// left), and 1 / <|noise|> for S_k.
__device__ __forceinline__ void finalize_store(const PassDesc& d, DevLedger L) {
  const double em = L.entering_total / fmax(L.entering_count, 1.0);
  const double rm = L.realized_total / fmax(L.realized_count, 1.0);
  const double ma = L.noise_abs_total / fmax(L.noise_abs_count, 1.0);
  if (rm > 0.0) {
    L.rescale = (float)(em / rm);
    L.flatten = 0;
  } else {
    L.rescale = 1.0f;
    L.flatten = 1;
  }
  L.entering_mean_f = (float)em;
  L.have_mean_abs = ma > 0.0;
  L.inv_mean_abs = ma > 0.0 ? (float)(1.0 / ma) : 1.0f;
  L.pad = 0;
  *d.led = L;
}

template <class Src>
__global__ __launch_bounds__(FIN_NTH) void finalize_kernel(Src src) {
  int wg = blockIdx.x;
  const PassDesc& d = src.get(PASS_FIN, wg);
  const ConvGeom& g = d.g;
  const int tid = threadIdx.x;
  __shared__ double red[5 * (FIN_NTH / 64)];
  DevLedger L;
  if (!d.apply) {
    double v[5] = {0, 0, 0, 0, 0};
    for (int i = tid; i < d.npart; i += FIN_NTH) {
      v[0] += d.part_z[i * 3 + 0];
      v[1] += d.part_z[i * 3 + 1];
      v[2] += d.part_z[i * 3 + 2];
      v[3] += d.part_o[i * 2 + 0];
      v[4] += d.part_o[i * 2 + 1];
    }
    block_sum_w<5, FIN_NTH>(v, red);
    if (tid != 0) return;
    const double cnt = (double)(d.win[1] - d.win[0]) * (double)(d.win[3] - d.win[2]) *
                       (double)g.nz;
    L.entering_total = v[0];
    L.entering_count = cnt;
    L.noise_abs_total = v[1];
    L.noise_abs_count = v[2];
    L.realized_total = v[3];
    L.realized_count = cnt;
    L.n_clipped = v[4];
  } else {
    if (tid != 0) return;
    L = d.led_in;
  }
  finalize_store(d, L);
}

// The apply-mode finalize (the ledger is given, so there is nothing to
// reduce): one wave per instance.  A 1024-thread workgroup would wait for
// a whole CU's wave slots to free up, which beside other streams' kernels
// (the tile server's) held it for hundreds of microseconds.
constexpr int FIN_APPLY_NTH = 64;

template <class Src>
__global__ __launch_bounds__(FIN_APPLY_NTH) void finalize_apply_kernel(Src src) {
  if (threadIdx.x != 0) return;
  int wg = blockIdx.x;
  const PassDesc& d = src.get(PASS_FIN, wg);
  finalize_store(d, d.led_in);
}

// ---------------------------------------------------------------------------
// post: apply the rescale to the flux and the 1/<|noise|> scale to S_k
// ---------------------------------------------------------------------------
template <class Src>
__global__ __launch_bounds__(FFT_NTH) __attribute__((amdgpu_waves_per_eu(4))) void post_kernel(Src src, int do_rescale) {
  int wg = blockIdx.x;
  const PassDesc& d = src.get(PASS_POST, wg);
  const ConvGeom& g = d.g;
  const int64_t n = (int64_t)g.nx * g.ny * g.nz;
  const DevLedger* L = d.led;
  const float rs = L->rescale;
  const int flat = L->flatten;
  const float em = L->entering_mean_f;
  const float inv = L->inv_mean_abs;
  const int have = L->have_mean_abs;
  const int64_t i0 = ((int64_t)wg * FFT_NTH + threadIdx.x) * 4;
  for (int q = 0; q < 4; ++q) {
    const int64_t i = i0 + q;
    if (i >= n) break;
    if (do_rescale) d.out[i] = flat ? em : __fmul_rn(d.out[i], rs);
    if (d.sk && have) d.sk[i] = __fmul_rn(d.sk[i], inv);
  }
}

}  // namespace steam
