// Shared device helpers for libsteam_hip: Philox4x32-10, the float32 uniform
// conversion, the Chambers-Mallows-Stuck transform, complex arithmetic and
// the error channel of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/steam_hip.h"

namespace steam {

// ---------------------------------------------------------------------------
// error channel (thread-local message, status code returned by every entry)
// ---------------------------------------------------------------------------
int set_error(int code, const std::string& msg);
int check_hip(hipError_t e, const char* what);

#define STEAM_HIP(call)                                        \
  do {                                                         \
    hipError_t _e = (call);                                    \
    if (_e != hipSuccess) return ::steam::check_hip(_e, #call); \
  } while (0)

#define STEAM_LAUNCH_CHECK(what)                              \
  do {                                                        \
    hipError_t _e = hipGetLastError();                        \
    if (_e != hipSuccess) return ::steam::check_hip(_e, what); \
  } while (0)

// ---------------------------------------------------------------------------
// instrumentation (no-ops unless steam_profile_enable(1) was called)
// ---------------------------------------------------------------------------
extern bool g_profiling;
void prof_begin(int family, hipStream_t s);
void prof_end(int family, hipStream_t s);

struct ProfScope {
  int fam;
  hipStream_t s;
  ProfScope(int f, hipStream_t st) : fam(f), s(st) {
    if (g_profiling) prof_begin(fam, s);
  }
  ~ProfScope() {
    if (g_profiling) prof_end(fam, s);
  }
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (steam/noise.py:48-85).  Counter (ix, iy, iz, key2), key
// (key0, key1).  Words are carried the way the reference carries them: as
// uint64 values whose products wrap mod 2^64.  Lattice indices in [0, 2^32)
// keep every word below 2^32, which is the fast path (32-bit mul_hi/mul_lo);
// anything else (a negative world z index) takes the exact 64-bit path.
// ---------------------------------------------------------------------------
constexpr uint32_t PHILOX_M0 = 0xD2511F53u;
constexpr uint32_t PHILOX_M1 = 0xCD9E8D57u;
constexpr uint32_t PHILOX_W0 = 0x9E3779B9u;
constexpr uint32_t PHILOX_W1 = 0xBB67AE85u;

__device__ __forceinline__ void philox_fast(uint32_t& c0, uint32_t& c1,
                                            uint32_t& c2, uint32_t& c3,
                                            uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += PHILOX_W0;
      k1 += PHILOX_W1;
    }
    // one 32x32->64 multiply per product (v_mad_u64_u32) instead of separate
    // high and low multiplies
    const uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
    const uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    // three-input XOR in one instruction (gfx950 v_bitop3_b32, truth table
    // 0x96 = a ^ b ^ c)
    uint32_t n0 = __builtin_amdgcn_bitop3_b32(hi1, c1, k0, 0x96);
    uint32_t n2 = __builtin_amdgcn_bitop3_b32(hi0, c3, k1, 0x96);
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
}

__device__ __forceinline__ void philox_wide(uint64_t& c0, uint64_t& c1,
                                            uint64_t& c2, uint64_t& c3,
                                            uint64_t k0, uint64_t k1) {
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 = (k0 + PHILOX_W0) & 0xFFFFFFFFull;
      k1 = (k1 + PHILOX_W1) & 0xFFFFFFFFull;
    }
    const uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
    const uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
    const uint64_t n0 = (p1 >> 32) ^ c1 ^ k0;
    const uint64_t n2 = (p0 >> 32) ^ c3 ^ k1;
    c0 = n0;
    c1 = p1 & 0xFFFFFFFFull;
    c2 = n2;
    c3 = p0 & 0xFFFFFFFFull;
  }
}

// Two uniforms of world site (ix, iy, iz): words 0 and 1 -> (w >> 8) * 2^-24
// (steam/noise.py:88-100).
__device__ __forceinline__ void uniform_pair(int64_t ix, int64_t iy,
                                             int64_t iz, uint32_t key0,
                                             uint32_t key1, uint32_t key2,
                                             float& u0, float& u1) {
  const float scale = 1.0f / 16777216.0f;
  uint64_t hi = ((uint64_t)ix | (uint64_t)iy | (uint64_t)iz) >> 32;
  if (hi == 0) {
    uint32_t c0 = (uint32_t)ix, c1 = (uint32_t)iy, c2 = (uint32_t)iz,
             c3 = key2;
    philox_fast(c0, c1, c2, c3, key0, key1);
    u0 = (float)(c0 >> 8) * scale;
    u1 = (float)(c1 >> 8) * scale;
  } else {
    uint64_t c0 = (uint64_t)ix, c1 = (uint64_t)iy, c2 = (uint64_t)iz, c3 = key2;
    philox_wide(c0, c1, c2, c3, key0, key1);
    u0 = (float)(c0 >> 8) * scale;
    u1 = (float)(c1 >> 8) * scale;
  }
}

// ---------------------------------------------------------------------------
// Elementary functions for the noise path.
//   log2 / exp2: the transcendental unit (v_log_f32 / v_exp_f32, ~1 ulp).
//   sinpi(t) = sin(pi t) for the CMS factors, whose arguments lie in [0, 2]:
//     reflected exactly onto [0, 1/2] (1 - t and t - 1 are exact in float32
//     there), then one odd polynomial t (a1 + a3 t^2 + ... + a9 t^8) fitted
//     here in relative error (5e-9 for the fit; at most 2.8 ulp evaluated
//     in float32, relative error below 1.7e-7 down to the smallest inputs).
//   expm1: degree-9 Taylor polynomial for |x| < 1/2, exp2 otherwise.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float hw_log2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float hw_exp2(float x) { return __builtin_amdgcn_exp2f(x); }

// sin(pi t), t in [0, 1/2]
__device__ __forceinline__ float sinpi_half(float t) {
  const float t2 = t * t;
  float p = 0.0775611624121666f;
  p = fmaf(p, t2, -0.598242461681366f);
  p = fmaf(p, t2, 2.550069808959961f);
  p = fmaf(p, t2, -5.167709827423096f);
  p = fmaf(p, t2, 3.1415927410125732f);
  return p * t;
}
// sin(pi t), t in [0, 1]: sin(pi t) = sin(pi (1 - t))
__device__ __forceinline__ float sinpi_unit(float t) { return sinpi_half(fminf(t, 1.0f - t)); }
// sin(pi t), t in [0, 2]: sin(pi t) = -sin(pi (t - 1)) from t = 1 on
__device__ __forceinline__ float sinpi_two(float t) {
  const bool hi = t >= 1.0f;
  const float s = sinpi_unit(hi ? t - 1.0f : t);
  return hi ? -s : s;
}

__device__ __forceinline__ float expm1_f32(float x) {
  if (fabsf(x) < 0.5f) {
    float p = 1.0f / 362880.0f;
    p = fmaf(p, x, 1.0f / 40320.0f);
    p = fmaf(p, x, 1.0f / 5040.0f);
    p = fmaf(p, x, 1.0f / 720.0f);
    p = fmaf(p, x, 1.0f / 120.0f);
    p = fmaf(p, x, 1.0f / 24.0f);
    p = fmaf(p, x, 1.0f / 6.0f);
    p = fmaf(p, x, 0.5f);
    p = fmaf(p, x, 1.0f);
    return p * x;
  }
  return hw_exp2(x * 1.4426950408889634f) - 1.0f;
}

// ---------------------------------------------------------------------------
// CMS extremal-Levy transform (steam/simulate.py:413-479), float32, in a
// cancellation-free form.  With phi = pi(u0 - 1/2), d = 1 - u0 (exact) and
// phi0 = -(pi/2)(1 - |1 - alpha|)/alpha the reference's three trigonometric
// factors are, for alpha > 1,
//   cos(phi)                   = sin(pi min(u0, d))
//   sin(alpha (phi - phi0))    = sin(pi alpha d)
//   cos(phi - alpha(phi-phi0)) = sin(pi (alpha - 1) d)
// and for alpha < 1 the last two are sin(pi alpha u0), sin(pi (1-alpha) u0).
// Evaluating these instead of the reference's float32 chain removes the
// cancellation near u0 -> 0, 1 that puts the reference's own float32 draws
// ~2e-3 (rel-L2) from the float64 transform; here they are ~5e-7.  Clamps
// (eps = 1e-6) and the sign convention are the reference's.
// ---------------------------------------------------------------------------
struct CmsConst {
  float alpha;      // float32(alpha)
  float am1;        // float32(|alpha - 1|)
  float neg_inv_a;  // float32(-1/alpha)
  float tail_exp;   // float32((1-alpha)/alpha)
  float c0;         // -log2|alpha-1|/alpha - tail_exp * log2(ln 2)
  float sign;       // +1 for alpha > 1, -1 otherwise
  int above_one;    // alpha > 1
};

inline CmsConst make_cms(double a) {
  CmsConst c;
  c.alpha = (float)a;
  c.am1 = (float)fabs(a - 1.0);
  c.neg_inv_a = (float)(-1.0 / a);
  c.tail_exp = (float)((1.0 - a) / a);
  c.c0 = (float)(-log2(fabs(a - 1.0)) / a - (1.0 - a) / a * log2(log(2.0)));
  c.sign = a > 1.0 ? 1.0f : -1.0f;
  c.above_one = a > 1.0;
  return c;
}

// X = sign sin(alpha(phi-phi0)) (cos(phi)|alpha-1|)^(-1/alpha)
//       (cos(phi - alpha(phi-phi0)) / R)^((1-alpha)/alpha),  R = -ln u1,
// with the two powers folded into one exp2 of a sum of log2 terms.
__device__ __forceinline__ float cms_f32(float u0, float u1, const CmsConst& k) {
  const float eps = 1e-6f;
  const float d = 1.0f - u0;
  const float w = k.above_one ? d : u0;
  const float lu = hw_log2(fmaxf(u1, eps));          // log2 u1 < 0
  const float lr = hw_log2(-lu);                      // log2(R / ln 2)
  // arguments: min(u0, d) in [0, 1/2], |alpha - 1| w in [0, 1], alpha w in [0, 2]
  const float lc = hw_log2(fmaxf(sinpi_half(fminf(u0, d)), eps));
  const float lp = hw_log2(fmaxf(sinpi_unit(k.am1 * w), eps));
  const float ss = sinpi_two(k.alpha * w);
  const float e = fmaf(k.neg_inv_a, lc, fmaf(k.tail_exp, lp - lr, k.c0));
  return k.sign * ss * hw_exp2(e);
}

// Python-style (floor) modulo for a positive modulus.
__device__ __forceinline__ int64_t pmod(int64_t a, int64_t n) {
  int64_t r = a % n;
  return r < 0 ? r + n : r;
}

// ---------------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 csub(float2 a, float2 b) {
  return make_float2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) {
  return make_float2(a.x * s, a.y * s);
}

}  // namespace steam
