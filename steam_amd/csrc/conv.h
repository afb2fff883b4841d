// Descriptors shared by the convolution / flux-advance passes and their
// batched launches.
#pragma once

#include "fft.h"

namespace steam {

// Geometry of one convolution operator instance.
struct ConvGeom {
  int nx, ny, nz;       // field
  int kx, ky, kz;       // kernel after folding (kx <= nx, ky <= ny)
  int offx, offy;       // kernel placement: kp[x] gathers taps i = x+off mod nx
  int cz;               // (kz-1)/2: 'same' trim offset along z
  int nfft, n2, nzc;    // z transform: real length, complex half, kept bins
  int L, nblocks;       // overlap-add block length and count along z
  int Tz, logTz;        // lines per workgroup in the z passes
  int Ty, logTy;        // lines per workgroup in the y pass
  int Tx, logTx;        // lines per workgroup in the x pass
  int nyt_z, nyt_x;     // y-tiles per x line (z passes), per plane (x pass)
  int wg_z, wg_y, wg_x; // workgroups per pass
  int reg_y, reg_x;     // log2 length of the register FFT path (0: generic)
  int mr_y, mr_x;       // length of the mixed-radix register path (0: none)
  int Tr;               // columns per workgroup in the register x pass
  int tile_z;           // single-block z passes with vector tile loads
  int zreg;             // single-block register z passes (compiled n2)
  int wg_zr;            // workgroups of the register z passes
  int twl_y, twl_x;     // generic y / x pass with a direct or radix-7 stage:
                        // the axis twiddle table is staged in LDS after the
                        // lines (its reads are lane-scattered there)
  AxisPlan ax, ay, az;
  const float2* wr2c;   // n2+1 entries exp(-2 pi i k / nfft)
};

// Workgroup index remapped so that consecutive tiles land on one XCD
// (workgroups are dealt round-robin over the 8 XCDs; each has its own L2):
// neighbouring column tiles that share 128-byte lines then share a cache.
__device__ __forceinline__ int xcd_remap(int w, int n) {
  const int q = n >> 3, r = n & 7;
  const int i = w & 7, slot = w >> 3;
  return i * q + (i < r ? i : r) + slot;
}

// Keyed multiplier-noise parameters as the kernels use them.
struct NoiseDev {
  uint32_t key0, key1, key2;
  int64_t ox, oy, oz;       // world origin of local (0,0,0)
  int64_t wx, wy;           // world extent (x/y wrap)
  int sx, sy, sz;           // sparsity
  CmsConst cms;
  float pcs, shift;         // per-class scale, shift (float32)
  int n_zero, zero_bottom, zero_top;
};

// Device-side ledger written by the finalize kernel.
struct DevLedger {
  double entering_total, entering_count;
  double noise_abs_total, noise_abs_count;
  double realized_total, realized_count;
  double n_clipped;
  float rescale;
  int flatten;
  float entering_mean_f;
  float inv_mean_abs;
  int have_mean_abs;
  int pad;
};

enum PassMode { MODE_PLAIN = 0, MODE_FLUX = 1, MODE_KERNEL = 2 };

// Kernel-spectrum formats (the word at the start of a spectrum buffer).
// COMPLEX: complex64 [kz][x][y], or in the register x pass's column tiles.
// REAL_EVEN: the spectrum of a kernel with mirror symmetry on every axis is
// real and even in x; floats [kz][y/T][nx/2 + 1][T] (register x pass only).
enum KernelFormat { KFMT_COMPLEX = 0, KFMT_REAL_EVEN = 1 };
// Layout of a spectrum buffer: the format word, the per-workgroup mismatch
// counts of the symmetry test, then the spectrum itself.
constexpr int KSYM_WGS = 1024;
constexpr int64_t KSPEC_DATA_OFF = 256 + 4 * KSYM_WGS;

struct PassDesc {
  ConvGeom g;
  float2* S;            // spectral workspace [nblocks][nzc][nx][ny]
  const float2* K;      // kernel spectrum [nzc][nx][ny], or with ktile > 0 in the
                        // register x pass's column tiles [nzc][ny/ktile][nx][ktile]
  int ktile;
  const int* kfmt;      // format word of the kernel spectrum (KFMT_*)
  const float* in;      // plain input / kernel taps
  float* out;           // plain output / flux state (in place)
  int mode;
  int kahan;
  // flux advance
  float sigma_in;
  const DevLedger* pend;  // chained: pending rescale of the previous advance
  int apply;            // 1: reductions injected from ledger_in
  NoiseDev noise;
  int win[4];           // x0, x1, y0, y1
  float* sk;            // optional S_k output
  float* nfield;        // multiplier-noise field [nx][ny][nz] (workspace)
  double* part_z;       // [wg_z][3]: entering, |noise|, nonzero count
  double* part_o;       // [wg_z][2]: realized, clipped count
  DevLedger* led;
  DevLedger led_in;     // apply mode values
  float kscale;         // kernel-spectrum normalisation (MODE_KERNEL)
  int npart;            // partial-sum slots written by the z passes
  int clip_win;         // count clipped cells inside the window only
  int wg_noise_n;       // workgroups a strided noise launch gives this instance
                        // (0: the whole grid)
  int unit_in;          // the entering flux is 1 everywhere: the flux buffer
                        // is written, never read
  int out_scale;        // apply mode: the inverse z pass writes the rescaled
  int out_flat;         // flux (flat ? out_em : flux * out_rs), the recorded
  float out_rs, out_em; // ledger's rescale, so no post pass reads it again
  float* crop_out;      // with out_scale: only the box crop_o + [0, crop_n) of
  int crop_o[3];        // the advanced flux is stored, dense in crop_out, and
  int crop_n[3];        // the grid itself is not written (NULL: the grid)
  int wg_base[7];       // batched: first workgroup of z, y, x, zinv, post,
                        // finalize, noise passes
};

}  // namespace steam

namespace steam {

// Work split of the noise-field kernel: lines handled side by side per pass
// (256 / nz of them when a line has at most 128 levels, else one) and lines
// per workgroup (a whole number of passes, at least four lines).
__host__ __device__ inline int noise_lines_per_pass(int nz) {
  return (nz > 0 && nz <= 128) ? 256 / nz : 1;
}
__host__ __device__ inline int noise_lines_per_wg(int nz) {
  const int u = noise_lines_per_pass(nz);
  return u * ((4 + u - 1) / u);
}

// Per-line part of the keyed multiplier noise: the wrapped world x/y of the
// line and whether it lies on the x/y centre sub-lattice.  World coordinates
// are 32-bit here (the host rejects worlds or origins beyond +-2^30).
struct LineNoise {
  int wx, wy;
  int xy_center;
  int zphase;  // world z of local z=0, modulo s_z
};

__device__ __forceinline__ int pmod32(int a, int n) {
  const int r = a % n;
  return r < 0 ? r + n : r;
}

__device__ __forceinline__ LineNoise noise_line(const NoiseDev& nd, int x, int y) {
  LineNoise ln;
  ln.wx = pmod32(x + (int)nd.ox, (int)nd.wx);
  ln.wy = pmod32(y + (int)nd.oy, (int)nd.wy);
  ln.xy_center = (nd.sx == 1 || ln.wx % nd.sx == 0) && (nd.sy == 1 || ln.wy % nd.sy == 0);
  ln.zphase = pmod32((int)nd.oz, nd.sz);
  return ln;
}

// Unit-mean multiplier noise exp(gamma)-1 at local level z of a line of an
// array with nz levels; 0 off-centre and in the zeroed boundary levels.
// Follows steam/simulate.py:2146-2161 (zero levels, scale, off-centre mask,
// shift, expm1) on the keyed draw of steam/simulate.py:554-598 (centres are
// world sites divisible by s after the x/y wrap; z never wraps).  The
// reference finds off-centre cells as those whose scaled draw is exactly 0;
// here they are the cells off the centre lattice.  The two differ only where
// a centre's draw is zero: the measure-zero set where sin(alpha(phi - phi0))
// vanishes (e.g. u0 = 3/8 at alpha = 1.6), on which any evaluation returns
// rounding noise and the FP64 transform, like the reference's own float32
// one, keeps the cell as a centre (noise expm1(-shift)).
__device__ __forceinline__ float noise_at(const NoiseDev& nd,
                                         const LineNoise& ln, int z, int nz) {
  if (!ln.xy_center) return 0.f;
  if (nd.n_zero > 0) {
    if (nd.zero_bottom && z < nd.n_zero) return 0.f;
    if (nd.zero_top && z >= nz - nd.n_zero) return 0.f;
  }
  if (nd.sz > 1 && (z + ln.zphase) % nd.sz != 0) return 0.f;
  float u0, u1;
  uniform_pair(ln.wx, ln.wy, (int64_t)(z + (int)nd.oz), nd.key0, nd.key1,
               nd.key2, u0, u1);
  float gam = cms_f32(u0, u1, nd.cms);
  const float g = __fmul_rn(gam, nd.pcs);
  return expm1_f32(__fsub_rn(g, nd.shift));
}

}  // namespace steam
