/*
 * CPU oracle for the keyed-noise family: Philox4x32-10, the float32 uniform
 * conversion, and the Chambers-Mallows-Stuck extremal-Levy transform.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP kernels are
 * compared against; nothing in the steam_amd package links it or calls it.
 *
 * Restated from the reference's Numba/NumPy code:
 *   philox4x32_10      steam/noise.py:67-85   (round function, key bump)
 *   uniform pair       steam/noise.py:88-100  (words 0,1 -> (w>>8)*2^-24)
 *   keyed_uniforms     steam/noise.py:103-130 (C order over ix, iy, iz)
 *   CMS, float32       steam/simulate.py:413-479 (same op order, eps=1e-6f)
 *   CMS, float64       same formula evaluated in double
 *
 * Integer semantics follow the reference exactly: every state word is a
 * uint64 (numba np.uint64), products wrap mod 2^64, and only the key words are
 * masked to 32 bits.  For lattice indices in [0, 2^32) every word stays below
 * 2^32 and this is textbook Philox4x32-10; a negative world index (possible
 * only for z of an elevated nest) reproduces the reference's uint64 cast.
 */
#include <math.h>
#include <stdint.h>

#define M0 0xD2511F53ull
#define M1 0xCD9E8D57ull
#define W0 0x9E3779B9ull
#define W1 0xBB67AE85ull
#define MASK32 0xFFFFFFFFull

static inline void philox4x32_10(uint64_t c[4], uint64_t k0, uint64_t k1) {
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 = (k0 + W0) & MASK32;
      k1 = (k1 + W1) & MASK32;
    }
    uint64_t p0 = M0 * c[0];
    uint64_t p1 = M1 * c[2];
    uint64_t hi0 = p0 >> 32, lo0 = p0 & MASK32;
    uint64_t hi1 = p1 >> 32, lo1 = p1 & MASK32;
    uint64_t n0 = hi1 ^ c[1] ^ k0;
    uint64_t n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
  }
}

void oracle_philox4x32_10(const uint64_t in[4], uint64_t k0, uint64_t k1,
                          uint64_t out[4]) {
  uint64_t c[4] = {in[0], in[1], in[2], in[3]};
  philox4x32_10(c, k0, k1);
  for (int i = 0; i < 4; ++i) out[i] = c[i];
}

static inline void uniform_pair(int64_t ix, int64_t iy, int64_t iz,
                                uint64_t key0, uint64_t key1, uint64_t key2,
                                float* u0, float* u1) {
  uint64_t c[4] = {(uint64_t)ix, (uint64_t)iy, (uint64_t)iz, key2};
  philox4x32_10(c, key0, key1);
  const float scale = 1.0f / 16777216.0f;
  *u0 = (float)(c[0] >> 8) * scale;
  *u1 = (float)(c[1] >> 8) * scale;
}

void oracle_keyed_uniforms(const int64_t* ix, int64_t nx, const int64_t* iy,
                           int64_t ny, const int64_t* iz, int64_t nz,
                           uint64_t key0, uint64_t key1, uint64_t key2,
                           float* lane0, float* lane1) {
#pragma omp parallel for schedule(static)
  for (int64_t a = 0; a < nx; ++a)
    for (int64_t b = 0; b < ny; ++b) {
      int64_t base = (a * ny + b) * nz;
      for (int64_t c = 0; c < nz; ++c)
        uniform_pair(ix[a], iy[b], iz[c], key0, key1, key2, &lane0[base + c],
                     &lane1[base + c]);
    }
}

/* CMS in float32 with the reference's operation order (simulate.py:413-438,
 * constants from 441-470).  libm float functions stand in for numpy's float32
 * ufuncs; both are faithful to a few ulp, so this leg is compared at ulp-level
 * tolerance, never bitwise. */
static inline float cms_f32(float u0, float u1, float alpha_f, float phi0,
                            float sign, float abs_am1, float neg_inv_alpha,
                            float tail_exp) {
  const float eps = 1e-6f;
  const float pi_f = (float)3.141592653589793;
  float phi = (u0 - 0.5f) * pi_f;
  float R = u1 < eps ? eps : u1;
  R = logf(R);
  R *= -1.0f;
  float factor = cosf(phi);
  factor = factor < eps ? eps : factor;
  factor *= abs_am1;
  factor = powf(factor, neg_inv_alpha);
  float shifted = (phi - phi0) * alpha_f;
  float p = phi - shifted;
  p = cosf(p);
  p = p < eps ? eps : p;
  p /= R;
  p = powf(p, tail_exp);
  float s = sinf(shifted);
  s *= sign;
  s *= factor;
  s *= p;
  return s;
}

/* CMS in float64 on the same float32 uniforms: the FP64 oracle. */
static inline double cms_f64(float u0f, float u1f, double alpha) {
  const double eps = (double)1e-6f;
  const double pi = 3.141592653589793;
  double phi0 = -(pi / 2.0) * (1.0 - fabs(1.0 - alpha)) / alpha;
  double sign = alpha > 1.0 ? 1.0 : -1.0;
  double phi = pi * ((double)u0f - 0.5);
  double R = -log((double)u1f < eps ? eps : (double)u1f);
  double c = cos(phi);
  double factor = pow((c < eps ? eps : c) * fabs(alpha - 1.0), -1.0 / alpha);
  double shifted = alpha * (phi - phi0);
  double c2 = cos(phi - shifted);
  double tail = pow((c2 < eps ? eps : c2) / R, (1.0 - alpha) / alpha);
  return sign * sin(shifted) * factor * tail;
}

void oracle_levy_f32(double alpha, const float* u0, const float* u1, float* out,
                     int64_t n) {
  /* the float32 constants are rounded from float64 expressions of alpha, as
   * the reference forms them (steam/simulate.py:471-473, 421-433) */
  double a = alpha;
  float phi0 = (float)(-(3.141592653589793 / 2.0) * (1.0 - fabs(1.0 - a)) / a);
  float sign = a > 1.0 ? 1.0f : -1.0f;
  float abs_am1 = (float)fabs(a - 1.0);
  float neg_inv_alpha = (float)(-1.0 / a);
  float tail_exp = (float)((1.0 - a) / a);
  float alpha_f = (float)a;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
    out[i] = cms_f32(u0[i], u1[i], alpha_f, phi0, sign, abs_am1, neg_inv_alpha,
                     tail_exp);
}

void oracle_levy_f64(double alpha, const float* u0, const float* u1,
                     double* out, int64_t n) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) out[i] = cms_f64(u0[i], u1[i], alpha);
}

/* Fused keyed Levy field over an index box (the centres only), used by the
 * CPU baseline so the timed leg does one pass like the GPU kernel does.
 * mode 0: float32 CMS into out32; mode 1: float64 CMS into out64. */
void oracle_keyed_levy(const int64_t* ix, int64_t nx, const int64_t* iy,
                       int64_t ny, const int64_t* iz, int64_t nz, uint64_t key0,
                       uint64_t key1, uint64_t key2, double alpha, int mode,
                       float* out32, double* out64) {
  double a = alpha;
  float phi0 = (float)(-(3.141592653589793 / 2.0) * (1.0 - fabs(1.0 - a)) / a);
  float sign = a > 1.0 ? 1.0f : -1.0f;
  float abs_am1 = (float)fabs(a - 1.0);
  float neg_inv_alpha = (float)(-1.0 / a);
  float tail_exp = (float)((1.0 - a) / a);
  float alpha_f = (float)a;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nx; ++i)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t k = 0; k < nz; ++k) {
        float u0, u1;
        uniform_pair(ix[i], iy[j], iz[k], key0, key1, key2, &u0, &u1);
        int64_t o = (i * ny + j) * nz + k;
        if (mode == 0)
          out32[o] = cms_f32(u0, u1, alpha_f, phi0, sign, abs_am1,
                             neg_inv_alpha, tail_exp);
        else
          out64[o] = cms_f64(u0, u1, a);
      }
}
