// FFT overlap-add convolution and fused flux advance: planning and the
// single-instance C entry points (kernels in conv_kernels.h).
#include <math.h>

#include <algorithm>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "conv_plan.h"

namespace steam {

// ---------------------------------------------------------------------------
// host side: planning
// ---------------------------------------------------------------------------
static bool is_smooth7(int64_t n) {
  for (int p : {2, 3, 5, 7})
    while (n % p == 0) n /= p;
  return n == 1;
}

// Stages of a length: radix 8 first, then 4, 2, 3, 5, 7, then each prime
// factor above 7 as a direct stage (cost: that prime per element; the last
// free slot takes the product of what remains); the whole line as one direct
// stage when the slots run out.
static void factor_axis(int n, AxisPlan& ap) {
  ap.n = n;
  ap.nst = 0;
  if (n <= 1) return;
  int m = n;
  while (m % 8 == 0 && ap.nst < FFT_MAXST) { ap.radix[ap.nst++] = 8; m /= 8; }
  while (m % 4 == 0 && ap.nst < FFT_MAXST) { ap.radix[ap.nst++] = 4; m /= 4; }
  while (m % 2 == 0 && ap.nst < FFT_MAXST) { ap.radix[ap.nst++] = 2; m /= 2; }
  for (int p : {3, 5, 7})
    while (m % p == 0 && ap.nst < FFT_MAXST) { ap.radix[ap.nst++] = p; m /= p; }
  if (m != 1 && (ap.nst == FFT_MAXST || m <= 8)) {
    ap.nst = 1;
    ap.radix[0] = 0;
    return;
  }
  for (int p = 11; m != 1 && p * p <= m && ap.nst < FFT_MAXST - 1; p += 2)
    while (m % p == 0 && m != p && ap.nst < FFT_MAXST - 1) { ap.radix[ap.nst++] = p; m /= p; }
  if (m != 1) ap.radix[ap.nst++] = m;  // a prime, or the product of the rest
}

// A direct stage of radix R costs R complex products per element; the
// Bluestein form costs two transforms of a smooth length M >= 2n - 1 per
// transform, in the LDS-staged passes.  Measured x passes (140 x 97 planes,
// twiddles read from LDS): 2060 = 4 5 103 took 3.04 ms direct and 4.77 ms
// Bluestein, 1036 = 4 7 37 0.65 ms direct and 3.26 ms Bluestein, a 2053-point
// (prime) line 117.8 ms direct and 5.1 ms Bluestein; so it takes over from
// radix 128 on, where a direct stage costs hundreds of products per element
// (a prime line: n of them).  STEAM_BLUESTEIN=0 in the environment keeps the
// direct stages (read once; for measurement).
constexpr int BLUE_MIN_RADIX = 128;

static bool bluestein_on() {
  static const bool on = [] {
    const char* e = getenv("STEAM_BLUESTEIN");
    return !(e && *e == '0');
  }();
  return on;
}

// The x/y form of a length: its stages, and, when they hold a long direct
// stage, the Bluestein length and its stages (the smallest M >= 2n - 1 that
// factors into radices 2..8 within the stage slots).  Tables come later.
static void factor_xy_axis(int n, AxisPlan& ap) {
  factor_axis(n, ap);
  ap.bm = 0;
  ap.bnst = 0;
  int worst = 0;
  for (int s = 0; s < ap.nst; ++s)
    worst = std::max(worst, ap.radix[s] == 0 ? n : ap.radix[s]);
  if (worst < BLUE_MIN_RADIX || !bluestein_on()) return;
  for (int M = 2 * n - 1; M <= FFT_LDSE * 2; ++M) {
    AxisPlan mp;
    factor_axis(M, mp);
    bool smooth = mp.nst > 0;
    for (int s = 0; s < mp.nst; ++s) smooth = smooth && mp.radix[s] >= 2 && mp.radix[s] <= 8;
    if (!smooth) continue;
    ap.bm = M;
    ap.bnst = mp.nst;
    for (int s = 0; s < FFT_MAXST; ++s) ap.bradix[s] = s < mp.nst ? mp.radix[s] : 0;
    return;
  }
}

// Forward DFT in float64 of a 7-smooth length (recursive decimation in time
// over the smallest prime factor), for the Bluestein tables.
static void dft64(std::vector<std::complex<double>>& x) {
  const int n = (int)x.size();
  if (n <= 1) return;
  int p = 2;
  while (n % p) ++p;
  const int q = n / p;
  std::vector<std::vector<std::complex<double>>> sub(p, std::vector<std::complex<double>>(q));
  for (int r = 0; r < p; ++r)
    for (int j = 0; j < q; ++j) sub[r][j] = x[j * p + r];
  for (int r = 0; r < p; ++r) dft64(sub[r]);
  for (int k = 0; k < n; ++k) {
    std::complex<double> acc = 0.0;
    for (int r = 0; r < p; ++r) {
      const double a = -2.0 * M_PI * (double)((int64_t)r * k % n) / (double)n;
      acc += sub[r][k % q] * std::complex<double>(cos(a), sin(a));
    }
    x[k] = acc;
  }
}

// The Bluestein tables of an axis: chirp[m] = exp(-pi i m^2 / n) (m^2 taken
// mod 2n exactly), the transform of the wrapped conjugate chirp over M
// divided by M, and the M-point twiddles; appended to tw.
static void bluestein_tables(const AxisPlan& ap, std::vector<float2>& tw) {
  const int n = ap.n, M = ap.bm;
  std::vector<std::complex<double>> c(n), b(M, 0.0);
  for (int m = 0; m < n; ++m) {
    const int64_t e = ((int64_t)m * m) % (2 * (int64_t)n);
    const double a = -M_PI * (double)e / (double)n;
    c[m] = std::complex<double>(cos(a), sin(a));
  }
  for (int m = 0; m < n; ++m) {
    b[m] = std::conj(c[m]);
    if (m > 0) b[M - m] = std::conj(c[m]);
  }
  dft64(b);
  for (int m = 0; m < n; ++m) tw.push_back(make_float2((float)c[m].real(), (float)c[m].imag()));
  for (int m = 0; m < M; ++m)
    tw.push_back(make_float2((float)(b[m].real() / M), (float)(b[m].imag() / M)));
  for (int m = 0; m < M; ++m) {
    const double a = -2.0 * M_PI * (double)m / (double)M;
    tw.push_back(make_float2((float)cos(a), (float)sin(a)));
  }
}

// the mixed-radix register x/y passes: on unless STEAM_XY_MIXED_RADIX=0 in
// the environment (read once; the generic LDS passes then serve those
// lengths, for measurement)
static bool xy_mixed_radix() {
  static const bool on = [] {
    const char* e = getenv("STEAM_XY_MIXED_RADIX");
    return !(e && *e == '0');
  }();
  return on;
}

// log2 of a power of two in [64, 2048], else 0 (no register FFT path)
static int reg_log2(int n) {
  for (int l = 6; l <= 11; ++l)
    if (n == (1 << l)) return l;
  return 0;
}

static int floor_pow2(int64_t v) {
  int p = 1;
  while ((int64_t)p * 2 <= v) p *= 2;
  return p;
}
static int ilog2(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

}  // namespace steam


using namespace steam;


static const int64_t kMaxAxis = FFT_LDSE;  // longest x/y line one workgroup holds


static void set_tiles(ConvGeom& g, int nz_acc, int kahan, size_t& lds_z,
                      size_t& lds_zinv, size_t& lds_y, size_t& lds_x) {
  const int nzc = g.n2 + 1;
  int Tz = std::min(16, floor_pow2(std::max<int64_t>(1, FFT_LDSE / nzc)));
  Tz = std::min(Tz, 1 << ilog2(g.ny));
  while (Tz > 1 && (size_t)(Tz + 1) * nzc * 8 + (size_t)Tz * nz_acc * 4 * (1 + kahan) >
                       64 * 1024)
    Tz /= 2;
  g.Tz = Tz;
  g.logTz = ilog2(Tz);
  g.Ty = std::max(1, std::min(16, floor_pow2(FFT_LDSE / g.ny)));
  int Tx = std::min(16, floor_pow2(std::max<int64_t>(1, FFT_LDSE / g.nx)));
  Tx = std::min(Tx, 1 << ilog2(g.ny));
  g.Tx = std::max(1, Tx);
  // a Bluestein axis takes one scratch line of bm values after the lines:
  // fewer lines per workgroup until it fits 64 KiB, else the direct stage
  constexpr size_t kLds = 64 * 1024;
  if (g.ay.bm) {
    while (g.Ty > 1 && (size_t)((g.Ty + 1) * g.ny + g.ay.bm) * 8 > kLds) g.Ty /= 2;
    if ((size_t)((g.Ty + 1) * g.ny + g.ay.bm) * 8 > kLds) g.ay.bm = 0;
  }
  if (g.ax.bm) {
    while (g.Tx > 1 && (size_t)(g.nx * g.Tx + g.ax.bm) * 8 > kLds) g.Tx /= 2;
    if ((size_t)(g.nx * g.Tx + g.ax.bm) * 8 > kLds) g.ax.bm = 0;
  }
  g.logTy = ilog2(g.Ty);
  g.logTx = ilog2(g.Tx);
  g.nyt_z = (g.ny + g.Tz - 1) / g.Tz;
  g.nyt_x = (g.ny + g.Tx - 1) / g.Tx;
  const int64_t P = (int64_t)g.nblocks * g.nzc;
  g.reg_y = reg_log2(g.ny);
  g.reg_x = reg_log2(g.nx);
  g.mr_y = (!g.reg_y && xy_mixed_radix() && xyplan_lanes(g.ny)) ? g.ny : 0;
  g.mr_x = (!g.reg_x && xy_mixed_radix() && xyplan_lanes(g.nx)) ? g.nx : 0;
  g.wg_z = g.nx * g.nyt_z;
  if (g.reg_y) {
    const int lpw = 256 / (g.ny / 8);
    g.wg_y = (int)((P * g.nx + lpw - 1) / lpw);
  } else if (g.mr_y) {
    const int lpw = xyplan_threads(g.ny) / xyplan_lanes(g.ny);
    g.wg_y = (int)((P * g.nx + lpw - 1) / lpw);
  } else {
    g.wg_y = (int)((P * g.nx + g.Ty - 1) / g.Ty);
  }
  if (g.reg_x) {
    g.Tr = xreg_cols(g.reg_x);
    g.wg_x = (int)(P * ((g.ny + g.Tr - 1) / g.Tr));
  } else if (g.mr_x) {
    g.Tr = xyplan_threads(g.nx) / xyplan_lanes(g.nx);
    g.wg_x = (int)(P * ((g.ny + g.Tr - 1) / g.Tr));
  } else {
    g.Tr = 0;
    g.wg_x = (int)(P * g.nyt_x);
  }
  lds_z = (size_t)(g.Tz + 1) * nzc * 8;
  lds_zinv = lds_z + (size_t)g.Tz * nz_acc * 4 * (1 + kahan);
  g.tile_z = g.nblocks == 1 && g.nz % 4 == 0 && g.Tz * g.nz <= 4 * ZT_U * Z_NTH;
  int zr_lines = 0;
  g.zreg = g.nblocks == 1 && zplan_lines(g.n2, zr_lines);
  g.wg_zr = g.zreg ? (int)(((int64_t)g.nx * g.ny + zr_lines - 1) / zr_lines) : 0;
  if (g.tile_z) {
    lds_z = std::max(lds_z + (size_t)(g.n2 + g.n2 / 2 + 1) * 8, (size_t)3 * 8 * (Z_NTH / 64));
    lds_zinv = lds_z;
  }
  lds_y = (size_t)((g.Ty + 1) * g.ny + g.ay.bm) * 8;
  lds_x = (size_t)(g.nx * g.Tx + g.ax.bm) * 8;
  // the generic passes of a length with a direct or radix-7 stage read the
  // axis twiddles from LDS when they fit beside the lines
  g.twl_y = radix_set(g.ay) == 2 && lds_y + (size_t)g.ny * 8 <= kLds;
  g.twl_x = radix_set(g.ax) == 2 && lds_x + (size_t)g.nx * 8 <= kLds;
  if (g.twl_y) lds_y += (size_t)g.ny * 8;
  if (g.twl_x) lds_x += (size_t)g.nx * 8;
}

// model cost of a z transform length: spectral planes nblocks*(n/2+1), each
// costing its x/y passes plus its z passes; the z passes of a length with a
// register plan cost about a quarter of the LDS-staged ones (measured per
// plane at 2048x2048x115: x/y 80 us, register z 38 us, LDS-staged z 135 us),
// so a plane weighs 5 with a register plan and 9 without; -1 if n is not
// usable
static int64_t nfft_cost(int64_t nz, int64_t kz, int64_t n) {
  if (n < 2 || n % 2 || n < kz || n / 2 + 1 > FFT_LDSE || !is_smooth7(n / 2)) return -1;
  const int64_t L = n - kz + 1;
  const int64_t nb = (nz + L - 1) / L;
  int lines = 0;
  const bool reg = nb == 1 && zplan_lines((int)(n / 2), lines);
  return nb * (n / 2 + 1) * (reg ? 5 : 9);
}

// candidate lengths in increasing model cost (ties: shorter first)
static std::vector<int64_t> nfft_candidates(int64_t nz, int64_t kz) {
  std::vector<std::pair<int64_t, int64_t>> c;
  const int64_t lo = std::max<int64_t>(2, kz);
  const int64_t hi = std::max<int64_t>(lo, 2 * (nz + kz));
  for (int64_t n = lo; n <= hi; ++n) {
    const int64_t cost = nfft_cost(nz, kz, n);
    if (cost >= 0) c.push_back({cost, n});
  }
  std::sort(c.begin(), c.end());
  std::vector<int64_t> out;
  for (auto& v : c) out.push_back(v.second);
  return out;
}

static int64_t choose_nfft(int64_t nz, int64_t kz) {
  const std::vector<int64_t> c = nfft_candidates(nz, kz);
  return c.empty() ? -1 : c[0];
}

extern "C" int steam_conv_nfft_candidates(int64_t nz, int64_t kz, int64_t* out,
                                          int32_t max_out, int32_t* count) {
  if (nz < 1 || kz < 1 || !count || (max_out > 0 && !out))
    return set_error(STEAM_EINVAL, "steam_conv_nfft_candidates: bad arguments");
  const std::vector<int64_t> c = nfft_candidates(nz, kz);
  const int n = std::min<int>((int)c.size(), std::max(0, max_out));
  for (int i = 0; i < n; ++i) out[i] = c[i];
  *count = n;
  return STEAM_OK;
}

extern "C" int steam_conv_zplans(int32_t* out, int32_t max_out, int32_t* count) {
  if (!count || (max_out > 0 && !out))
    return set_error(STEAM_EINVAL, "steam_conv_zplans: bad arguments");
  int n = 0;
  for (int n2 = 2; n2 <= FFT_LDSE; n2 += 2) {
    int lines = 0;
    if (!zplan_lines(n2, lines)) continue;
    if (n < max_out) out[n] = n2;
    ++n;
  }
  *count = std::min(n, std::max(0, (int)max_out));
  return STEAM_OK;
}

extern "C" int steam_conv_xyplans(int32_t* out, int32_t max_out, int32_t* count) {
  if (!count || (max_out > 0 && !out))
    return set_error(STEAM_EINVAL, "steam_conv_xyplans: bad arguments");
  int n = 0;
  for (int len = 2; len <= FFT_LDSE; len += 2) {
    if (!xyplan_lanes(len) || reg_log2(len)) continue;
    if (n < max_out) out[n] = len;
    ++n;
  }
  *count = std::min(n, std::max(0, (int)max_out));
  return STEAM_OK;
}

extern "C" int steam_conv_plan_create(int64_t nx, int64_t ny, int64_t nz,
                                      int64_t kx, int64_t ky, int64_t kz,
                                      int64_t n_fft, int precision,
                                      steam_conv_plan** out) {
  if (!out) return set_error(STEAM_EINVAL, "plan pointer is NULL");
  *out = nullptr;
  if (nx < 1 || ny < 1 || nz < 1 || kx < 1 || ky < 1 || kz < 1)
    return set_error(STEAM_EINVAL, "conv plan: all sizes must be positive");
  if (nx > kMaxAxis || ny > kMaxAxis)
    return set_error(STEAM_ENOTIMPL,
                     "conv plan: x/y extents above 4096 cells are not supported "
                     "by one workgroup's LDS line; tile the domain");
  if (precision != STEAM_PREC_PLAIN && precision != STEAM_PREC_KAHAN)
    return set_error(STEAM_EINVAL, "conv plan: unknown precision");
  if (n_fft == 0) n_fft = choose_nfft(nz, kz);
  if (n_fft < 2 || n_fft % 2 || n_fft < kz || n_fft / 2 + 1 > FFT_LDSE)
    return set_error(STEAM_EINVAL,
                     "conv plan: n_fft must be even, >= kz and <= 8190");
  if (nz > 16384) return set_error(STEAM_ENOTIMPL, "conv plan: nz above 16384");

  steam_conv_plan* p = new steam_conv_plan();
  p->nx = nx; p->ny = ny; p->nz = nz; p->kx = kx; p->ky = ky; p->kz = kz;
  p->precision = precision;
  ConvGeom& g = p->g;
  memset(&g, 0, sizeof(g));
  g.nx = (int)nx; g.ny = (int)ny; g.nz = (int)nz;
  g.kx = (int)kx; g.ky = (int)ky; g.kz = (int)kz;
  // kernel placement (fold_kernel_to_field then roll by -(k'-1)/2)
  g.offx = (kx <= nx) ? (int)((kx - 1) / 2)
                      : (int)(kx / 2 + (nx - 1) / 2 - nx / 2);
  g.offy = (ky <= ny) ? (int)((ky - 1) / 2)
                      : (int)(ky / 2 + (ny - 1) / 2 - ny / 2);
  g.cz = (int)((kz - 1) / 2);
  g.nfft = (int)n_fft;
  g.n2 = (int)(n_fft / 2);
  g.nzc = g.n2 + 1;
  g.L = (int)(n_fft - kz + 1);
  g.nblocks = (int)((nz + g.L - 1) / g.L);
  factor_xy_axis((int)nx, g.ax);
  factor_xy_axis((int)ny, g.ay);
  factor_axis(g.n2, g.az);
  g.az.bm = 0;
  const int kah = precision == STEAM_PREC_KAHAN ? 1 : 0;
  set_tiles(g, (int)nz, kah, p->lds_z, p->lds_zinv, p->lds_y, p->lds_x);
  if (p->lds_zinv > 64 * 1024) {
    delete p;
    return set_error(STEAM_ENOTIMPL, "conv plan: z line too long for LDS");
  }
  // kernel build geometry: one block covering kz taps
  p->gk = g;
  p->gk.nblocks = 1;
  p->gk.L = g.nfft;
  {
    // its own tile heights: one accumulated level and no compensation let
    // the build take taller z tiles than the field's, with more LDS
    size_t unused;
    set_tiles(p->gk, 1, 0, p->lds_zk, unused, p->lds_yk, p->lds_xk);
  }

  // twiddles: [x: nx][y: ny][z: n2][r2c: n2+1], then for a Bluestein x and
  // y axis [chirp: n][chirp transform / M: M][M-point twiddles: M]
  const int64_t ntw0 = nx + ny + g.n2 + g.n2 + 1;
  std::vector<float2> tw(ntw0);
  auto fill = [&](int64_t off, int64_t n, int64_t period) {
    for (int64_t m = 0; m < n; ++m) {
      double a = -2.0 * M_PI * (double)m / (double)period;
      tw[off + m] = make_float2((float)cos(a), (float)sin(a));
    }
  };
  fill(0, nx, nx);
  fill(nx, ny, ny);
  fill(nx + ny, g.n2, g.n2);
  fill(nx + ny + g.n2, g.n2 + 1, g.nfft);
  const int64_t bx_off = (int64_t)tw.size();
  if (g.ax.bm) bluestein_tables(g.ax, tw);
  const int64_t by_off = (int64_t)tw.size();
  if (g.ay.bm) bluestein_tables(g.ay, tw);
  const int64_t ntw = (int64_t)tw.size();
  hipError_t e = hipMalloc(&p->tw_dev, ntw * sizeof(float2));
  if (e != hipSuccess) {
    delete p;
    return check_hip(e, "hipMalloc(twiddles)");
  }
  e = hipMemcpy(p->tw_dev, tw.data(), ntw * sizeof(float2), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(p->tw_dev);
    delete p;
    return check_hip(e, "hipMemcpy(twiddles)");
  }
  g.ax.tw = p->tw_dev;
  g.ay.tw = p->tw_dev + nx;
  g.az.tw = p->tw_dev + nx + ny;
  g.wr2c = p->tw_dev + nx + ny + g.n2;
  auto blue_ptrs = [&](AxisPlan& ap, int64_t off) {
    if (!ap.bm) return;
    ap.chirp = p->tw_dev + off;
    ap.bk = p->tw_dev + off + ap.n;
    ap.btw = p->tw_dev + off + ap.n + ap.bm;
  };
  blue_ptrs(g.ax, bx_off);
  blue_ptrs(g.ay, by_off);
  p->gk.ax = g.ax;
  p->gk.ay = g.ay;
  p->gk.az.tw = g.az.tw;
  p->gk.wr2c = g.wr2c;

  p->s_bytes = align256((int64_t)g.nblocks * g.nzc * nx * ny * 8);
  p->pz_off = p->s_bytes;
  const int64_t nparts = std::max(g.wg_z, g.wg_zr);
  p->po_off = p->pz_off + align256(nparts * 3 * 8);
  p->led_off = p->po_off + align256(nparts * 2 * 8);
  p->nf_off = p->led_off + align256(sizeof(DevLedger));
  p->ws_bytes = p->nf_off + align256(nx * ny * nz * 4);
  {
    const int lpw = noise_lines_per_wg((int)nz);
    p->wg_noise = (int)((nx * ny + lpw - 1) / lpw);
  }
  // the register x pass reads the kernel spectrum in its own column tiles,
  // each workgroup's block one contiguous run; the build writes the usual
  // [nzc][nx][ny] order after them and is then permuted
  // a spectrum buffer starts with its format word and the symmetry test's
  // counts (KSPEC_DATA_OFF bytes), then holds the spectrum
  p->ktile = (g.reg_x || g.mr_x) ? g.Tr : 0;
  if (p->ktile) {
    const int64_t nyt = (ny + p->ktile - 1) / p->ktile;
    p->kstd_off = KSPEC_DATA_OFF + align256((int64_t)g.nzc * nyt * p->ktile * nx * 8);
    p->spec_bytes = p->kstd_off + align256((int64_t)g.nzc * nx * ny * 8);
  } else {
    p->kstd_off = KSPEC_DATA_OFF;
    p->spec_bytes = KSPEC_DATA_OFF + align256((int64_t)g.nzc * nx * ny * 8);
  }
  // the real, x-even format needs the register x pass and a kernel that is
  // neither folded nor even-sized on any axis (then mirror symmetry of its
  // taps makes the spectrum, with the 'same' trim phase, real and even)
  p->ksym_shape = p->ktile && (kx & 1) && (ky & 1) && (kz & 1) && kx <= nx && ky <= ny;
  p->wg_post = (int)((nx * ny * nz + 4 * FFT_NTH - 1) / (4 * FFT_NTH));
  *out = p;
  return STEAM_OK;
}

extern "C" int steam_conv_plan_destroy(steam_conv_plan* p) {
  if (!p) return STEAM_OK;
  if (p->tw_dev) (void)hipFree(p->tw_dev);
  delete p;
  return STEAM_OK;
}

extern "C" int steam_conv_plan_info(const steam_conv_plan* p,
                                    steam_conv_info* info) {
  if (!p || !info) return set_error(STEAM_EINVAL, "NULL plan or info");
  info->n_fft = p->g.nfft;
  info->n_blocks = p->g.nblocks;
  info->block_len = p->g.L;
  info->workspace_bytes = p->ws_bytes;
  info->spectrum_bytes = p->spec_bytes;
  info->spectral_elems = (int64_t)p->g.nblocks * p->g.nzc * p->nx * p->ny;
  return STEAM_OK;
}


namespace steam {

// kernel spectrum [nzc][nx][ny] -> the register x pass's column tiles
// [nzc][ceil(ny/T)][nx][T] (the tail of a partial last tile is never read),
// or with the real, x-even format the real parts of bins x <= nx/2 as
// floats [nzc][ceil(ny/T)][nx/2 + 1][T]
__global__ __launch_bounds__(256) void ktile_kernel(const float2* __restrict__ src,
                                                    void* __restrict__ dst_v, const int* fmt,
                                                    int nzc, int nx, int ny, int T) {
  const int64_t plane = (int64_t)nx * ny;
  const int64_t n = plane * nzc;
  const int64_t nyt = (ny + T - 1) / T;
  const bool real = *fmt == KFMT_REAL_EVEN;
  const int xh = nx / 2 + 1;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const int64_t p = i / plane, r = i - p * plane;
    const int x = (int)(r / ny), y = (int)(r - (int64_t)x * ny);
    if (real) {
      if (x < xh)
        reinterpret_cast<float*>(dst_v)[((p * nyt + y / T) * xh + x) * T + y % T] = src[i].x;
    } else {
      reinterpret_cast<float2*>(dst_v)[((p * nyt + y / T) * nx + x) * T + y % T] = src[i];
    }
  }
}

// mirror-symmetry test of the kernel taps (kx, ky, kz): per workgroup, the
// number of taps that differ from their mirror image on some axis
__global__ __launch_bounds__(256) void ksym_count_kernel(const float* __restrict__ k, int kx,
                                                         int ky, int kz, int* counts) {
  __shared__ int red[256];
  const int64_t n = (int64_t)kx * ky * kz;
  int bad = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const int c = (int)(i % kz);
    const int64_t ab = i / kz;
    const int b = (int)(ab % ky), a = (int)(ab / ky);
    const float v = k[i];
    const float mx = k[((int64_t)(kx - 1 - a) * ky + b) * kz + c];
    const float my = k[((int64_t)a * ky + (ky - 1 - b)) * kz + c];
    const float mz = k[((int64_t)a * ky + b) * kz + (kz - 1 - c)];
    bad += (v != mx) || (v != my) || (v != mz);
  }
  red[threadIdx.x] = bad;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[blockIdx.x] = red[0];
}

// the format word: REAL_EVEN when the shape allows it and no tap failed
__global__ __launch_bounds__(256) void ksym_format_kernel(const int* counts, int n,
                                                          int shape_ok, int* fmt) {
  __shared__ int red[256];
  int bad = 0;
  for (int i = threadIdx.x; i < n; i += 256) bad += counts[i];
  red[threadIdx.x] = bad;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *fmt = (shape_ok && red[0] == 0) ? KFMT_REAL_EVEN : KFMT_COMPLEX;
}

}  // namespace steam

extern "C" int steam_conv_set_kernel(const steam_conv_plan* p,
                                     const float* kernel_dev,
                                     void* spectrum_dev, void* workspace_dev,
                                     steam_stream_t stream) {
  (void)workspace_dev;
  if (!p || !kernel_dev || !spectrum_dev)
    return set_error(STEAM_EINVAL, "steam_conv_set_kernel: NULL argument");
  hipStream_t s = (hipStream_t)stream;
  PassDesc d;
  memset(&d, 0, sizeof(d));
  d.g = p->gk;
  d.mode = MODE_KERNEL;
  d.S = (float2*)((char*)spectrum_dev + p->kstd_off);
  d.in = kernel_dev;
  d.kscale = (float)(1.0 / (4.0 * (double)p->g.nfft * (double)p->nx * (double)p->ny));
  SingleSrc src{d};
  const ConvGeom& g = d.g;
  const int rsz = radix_set(g.az);
  STEAM_RS_DISPATCH(rsz, hipLaunchKernelGGL((zfwd_kernel<SingleSrc, MODE_KERNEL, RS_>),
                                            dim3(g.wg_z), dim3(Z_NTH), p->lds_zk, s, src));
  STEAM_LAUNCH_CHECK("zfwd_kernel(kernel)");
  launch_ypass(g, p->lds_yk, s, src, 0);
  STEAM_LAUNCH_CHECK("ypass_kernel(kernel)");
  launch_xpass<SingleSrc, 1>(g, p->lds_xk, s, src);
  STEAM_LAUNCH_CHECK("xpass_kernel(kernel)");
  // the format word, then the spectrum in that format
  int* fmt = (int*)spectrum_dev;
  int* counts = fmt + 64;
  int nsym = 0;
  if (p->ksym_shape) {
    nsym = (int)std::min<int64_t>((p->kx * p->ky * p->kz + 255) / 256, KSYM_WGS);
    hipLaunchKernelGGL(ksym_count_kernel, dim3(nsym), dim3(256), 0, s, kernel_dev, (int)p->kx,
                       (int)p->ky, (int)p->kz, counts);
    STEAM_LAUNCH_CHECK("ksym_count_kernel");
  }
  hipLaunchKernelGGL(ksym_format_kernel, dim3(1), dim3(256), 0, s, (const int*)counts, nsym,
                     (int)p->ksym_shape, fmt);
  STEAM_LAUNCH_CHECK("ksym_format_kernel");
  if (p->ktile) {
    const int64_t ne = (int64_t)g.nzc * p->nx * p->ny;
    hipLaunchKernelGGL(ktile_kernel, dim3((unsigned)std::min<int64_t>((ne + 255) / 256, 1 << 16)),
                       dim3(256), 0, s, (const float2*)d.S,
                       (void*)((char*)spectrum_dev + KSPEC_DATA_OFF), (const int*)fmt,
                       (int)g.nzc, (int)p->nx, (int)p->ny, p->ktile);
    STEAM_LAUNCH_CHECK("ktile_kernel");
  }
  return STEAM_OK;
}

extern "C" int steam_conv_execute(const steam_conv_plan* p, const float* in_dev,
                                  float* out_dev, const void* spectrum_dev,
                                  void* workspace_dev, steam_stream_t stream) {
  if (!p || !in_dev || !out_dev || !spectrum_dev || !workspace_dev)
    return set_error(STEAM_EINVAL, "steam_conv_execute: NULL argument");
  PassDesc d = base_desc(p, workspace_dev);
  d.mode = MODE_PLAIN;
  d.K = (const float2*)((const char*)spectrum_dev + KSPEC_DATA_OFF);
  d.kfmt = (const int*)spectrum_dev;
  d.ktile = p->ktile;
  d.in = in_dev;
  d.out = out_dev;
  return run_operator<MODE_PLAIN>(p, d, (hipStream_t)stream);
}

namespace steam {

NoiseDev make_noise_dev(const steam_noise_params* np) {
  NoiseDev n;
  memset(&n, 0, sizeof(n));
  n.key0 = np->region.key[0];
  n.key1 = np->region.key[1];
  n.key2 = np->region.key[2];
  n.ox = np->region.origin[0];
  n.oy = np->region.origin[1];
  n.oz = np->region.origin[2];
  n.wx = np->region.world_shape[0];
  n.wy = np->region.world_shape[1];
  n.sx = np->region.sparsity[0];
  n.sy = np->region.sparsity[1];
  n.sz = np->region.sparsity[2];
  n.cms = make_cms(np->alpha);
  n.pcs = np->per_class_scale;
  n.shift = np->shift;
  n.n_zero = np->n_zero;
  n.zero_bottom = np->zero_bottom;
  n.zero_top = np->zero_top;
  return n;
}

int check_noise_params(const steam_noise_params* np) {
  if (!np) return set_error(STEAM_EINVAL, "noise params are NULL");
  for (int a = 0; a < 3; ++a)
    if (np->region.origin[a] < -(1ll << 30) || np->region.origin[a] > (1ll << 30) ||
        np->region.world_shape[a] > (1ll << 30))
      return set_error(STEAM_EINVAL, "world origin/shape beyond +-2^30 cells");
  if (np->alpha == 1.0)
    return set_error(STEAM_EINVAL, "alpha = 1 is outside the generator's domain");
  for (int a = 0; a < 3; ++a)
    if (np->region.sparsity[a] < 1)
      return set_error(STEAM_EINVAL, "sparsity factors must be positive");
  if (np->region.world_shape[0] < 1 || np->region.world_shape[1] < 1)
    return set_error(STEAM_EINVAL, "world shape must be positive in x and y");
  if (!(np->alpha > 0.0 && np->alpha <= 2.0))
    return set_error(STEAM_EINVAL, "alpha must lie in (0, 2]");
  return STEAM_OK;
}

DevLedger ledger_to_dev(const steam_flux_ledger* h) {
  DevLedger L;
  memset(&L, 0, sizeof(L));
  L.entering_total = h->entering_total;
  L.entering_count = (double)h->entering_count;
  L.noise_abs_total = h->noise_abs_total;
  L.noise_abs_count = (double)h->noise_abs_count;
  L.realized_total = h->realized_total;
  L.realized_count = (double)h->realized_count;
  L.n_clipped = (double)h->n_clipped;
  return L;
}

// Apply mode: the rescale the finalize kernel derives from a recorded
// ledger (the same float64 operations), set into the descriptor so the
// inverse z pass stores the rescaled flux.
void set_out_scale(PassDesc& d, const DevLedger& L) {
  const double em = L.entering_total / std::fmax(L.entering_count, 1.0);
  const double rm = L.realized_total / std::fmax(L.realized_count, 1.0);
  d.out_scale = 1;
  d.out_flat = rm > 0.0 ? 0 : 1;
  d.out_rs = rm > 0.0 ? (float)(em / rm) : 1.0f;
  d.out_em = (float)em;
}

void ledger_to_host(const DevLedger& L, steam_flux_ledger* h) {
  h->entering_total = L.entering_total;
  h->entering_count = (int64_t)L.entering_count;
  h->noise_abs_total = L.noise_abs_total;
  h->noise_abs_count = (int64_t)L.noise_abs_count;
  h->realized_total = L.realized_total;
  h->realized_count = (int64_t)L.realized_count;
  h->n_clipped = (int64_t)L.n_clipped;
  h->rescale = L.rescale;
  h->flatten = L.flatten;
}

}  // namespace steam

int steam::noise_wgs_cap() {
  static const int cap = [] {
    const char* e = getenv("STEAM_NOISE_WGS");
    return e && *e ? atoi(e) : STEAM_NOISE_WGS;
  }();
  return cap;
}

extern "C" int steam_flux_noise(const steam_conv_plan* p,
                                const steam_noise_params* noise, float* out_dev,
                                steam_stream_t stream) {
  if (!p || !out_dev)
    return set_error(STEAM_EINVAL, "steam_flux_noise: NULL argument");
  int rc = check_noise_params(noise);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  PassDesc d;
  memset(&d, 0, sizeof(d));
  d.g = p->g;
  d.nfield = out_dev;
  d.noise = make_noise_dev(noise);
  {
    // the field generated ahead, usually beside another class's transforms:
    // at most STEAM_NOISE_WGS workgroups (0: one per four lines)
    ProfScope ps(STEAM_KF_NOISE, s);
    const int cap = noise_wgs_cap();
    const int grid = (cap > 0 && p->wg_noise > cap) ? cap : p->wg_noise;
    if (noise_lines_per_pass((int)p->nz) > 1)
      hipLaunchKernelGGL((noise_field_kernel<SingleSrc, true, NOISE_SHORT>), dim3(grid),
                         dim3(256), 0, s, SingleSrc{d});
    else
      hipLaunchKernelGGL((noise_field_kernel<SingleSrc, true, NOISE_LONG>), dim3(grid),
                         dim3(256), 0, s, SingleSrc{d});
  }
  STEAM_LAUNCH_CHECK("noise_field_kernel");
  return STEAM_OK;
}

extern "C" int steam_flux_advance(const steam_conv_plan* p,
                                  const void* spectrum_dev, void* workspace_dev,
                                  float* flux_dev, float sigma_in,
                                  const steam_noise_params* noise,
                                  const float* noise_dev,
                                  const int64_t* window, int mode,
                                  steam_flux_ledger* ledger, float* sk_dev,
                                  int flags, steam_stream_t stream) {
  const int defer_rescale = (flags & STEAM_DEFER_RESCALE) != 0;
  const int chain_pending = (flags & STEAM_CHAIN_PENDING) != 0;
  const int clip_win = (flags & STEAM_CLIP_IN_WINDOW) != 0;
  const bool sk_scale = sk_dev && !(flags & STEAM_SK_RAW);
  const int unit_in = (flags & STEAM_UNIT_FLUX) != 0;
  if (!p || !spectrum_dev || !workspace_dev || !flux_dev)
    return set_error(STEAM_EINVAL, "steam_flux_advance: NULL argument");
  int rc = check_noise_params(noise);
  if (rc) return rc;
  if (mode != STEAM_MEASURE && mode != STEAM_APPLY)
    return set_error(STEAM_EINVAL, "steam_flux_advance: unknown mode");
  if (mode == STEAM_APPLY && !ledger)
    return set_error(STEAM_EINVAL, "steam_flux_advance: apply mode needs a ledger");
  hipStream_t s = (hipStream_t)stream;
  PassDesc d = base_desc(p, workspace_dev);
  d.mode = MODE_FLUX;
  d.K = (const float2*)((const char*)spectrum_dev + KSPEC_DATA_OFF);
  d.kfmt = (const int*)spectrum_dev;
  d.ktile = p->ktile;
  d.out = flux_dev;
  d.sigma_in = sigma_in;
  d.pend = chain_pending ? (const DevLedger*)((const char*)workspace_dev + p->led_off)
                         : nullptr;
  d.noise = make_noise_dev(noise);
  d.sk = sk_dev;
  d.apply = mode == STEAM_APPLY;
  d.clip_win = clip_win;
  if (window) {
    if (window[0] < 0 || window[1] > p->nx || window[0] >= window[1] ||
        window[2] < 0 || window[3] > p->ny || window[2] >= window[3])
      return set_error(STEAM_EINVAL, "steam_flux_advance: window out of range");
    for (int i = 0; i < 4; ++i) d.win[i] = (int)window[i];
  }
  if (d.apply) d.led_in = ledger_to_dev(ledger);
  d.unit_in = unit_in;
  if (unit_in && chain_pending)
    return set_error(STEAM_EINVAL, "steam_flux_advance: a unit entering flux has no pending "
                                   "rescale to chain");
  // a served tile (unit entering flux, apply mode) knows its rescale up
  // front: the inverse z pass stores the rescaled flux and no post pass
  // rescales it.  Other apply-mode advances keep the post pass, so a replay
  // runs the kernels of the run it replays
  if (d.apply && !defer_rescale && unit_in) set_out_scale(d, d.led_in);
  if (noise_dev) {
    d.nfield = const_cast<float*>(noise_dev);
  } else {
    ProfScope ps(STEAM_KF_NOISE, s);
    if (noise_lines_per_pass((int)p->nz) > 1)
      hipLaunchKernelGGL((noise_field_kernel<SingleSrc, false, NOISE_SHORT>),
                         dim3(p->wg_noise), dim3(256), 0, s, SingleSrc{d});
    else
      hipLaunchKernelGGL((noise_field_kernel<SingleSrc, false, NOISE_LONG>),
                         dim3(p->wg_noise), dim3(256), 0, s, SingleSrc{d});
  }
  STEAM_LAUNCH_CHECK("noise_field_kernel");
  rc = run_operator<MODE_FLUX>(p, d, s);
  if (rc) return rc;
  SingleSrc src{d};
  {
    ProfScope ps(STEAM_KF_FINALIZE, s);
    if (d.apply)
      hipLaunchKernelGGL((finalize_apply_kernel<SingleSrc>), dim3(1), dim3(FIN_APPLY_NTH), 0, s,
                         src);
    else
      hipLaunchKernelGGL((finalize_kernel<SingleSrc>), dim3(1), dim3(FIN_NTH), 0, s, src);
  }
  STEAM_LAUNCH_CHECK("finalize_kernel");
  const bool post_rescale = !defer_rescale && !d.out_scale;
  if (post_rescale || sk_scale) {
    PassDesc dp = d;
    if (!sk_scale) dp.sk = nullptr;
    ProfScope ps(STEAM_KF_POST, s);
    hipLaunchKernelGGL((post_kernel<SingleSrc>), dim3(p->wg_post), dim3(FFT_NTH), 0,
                       s, SingleSrc{dp}, post_rescale ? 1 : 0);
    STEAM_LAUNCH_CHECK("post_kernel");
  }
  if (ledger && mode == STEAM_MEASURE) {
    DevLedger L;
    STEAM_HIP(hipMemcpyAsync(&L, d.led, sizeof(L), hipMemcpyDeviceToHost, s));
    STEAM_HIP(hipStreamSynchronize(s));
    ledger_to_host(L, ledger);
  }
  return STEAM_OK;
}

extern "C" int steam_flux_ledger_read(const steam_conv_plan* p,
                                      const void* workspace_dev,
                                      steam_flux_ledger* out,
                                      steam_stream_t stream) {
  if (!p || !workspace_dev || !out)
    return set_error(STEAM_EINVAL, "steam_flux_ledger_read: NULL argument");
  hipStream_t s = (hipStream_t)stream;
  DevLedger L;
  STEAM_HIP(hipMemcpyAsync(&L, (const char*)workspace_dev + p->led_off, sizeof(L),
                           hipMemcpyDeviceToHost, s));
  STEAM_HIP(hipStreamSynchronize(s));
  ledger_to_host(L, out);
  return STEAM_OK;
}
