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
    uint32_t hi0 = __umulhi(PHILOX_M0, c0);
    uint32_t lo0 = PHILOX_M0 * c0;
    uint32_t hi1 = __umulhi(PHILOX_M1, c2);
    uint32_t lo1 = PHILOX_M1 * c2;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
}

__device__ __noinline__ void philox_wide(uint64_t c[4], uint64_t k0,
                                         uint64_t k1) {
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 = (k0 + PHILOX_W0) & 0xFFFFFFFFull;
      k1 = (k1 + PHILOX_W1) & 0xFFFFFFFFull;
    }
    uint64_t p0 = (uint64_t)PHILOX_M0 * c[0];
    uint64_t p1 = (uint64_t)PHILOX_M1 * c[2];
    uint64_t n0 = (p1 >> 32) ^ c[1] ^ k0;
    uint64_t n2 = (p0 >> 32) ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = p1 & 0xFFFFFFFFull;
    c[2] = n2;
    c[3] = p0 & 0xFFFFFFFFull;
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
    uint64_t c[4] = {(uint64_t)ix, (uint64_t)iy, (uint64_t)iz, key2};
    philox_wide(c, key0, key1);
    u0 = (float)(c[0] >> 8) * scale;
    u1 = (float)(c[1] >> 8) * scale;
  }
}

// ---------------------------------------------------------------------------
// CMS extremal-Levy transform in float32, in the reference's operation order
// (steam/simulate.py:413-438 with the constants of 441-470).
// ---------------------------------------------------------------------------
struct CmsConst {
  float alpha;      // float32(alpha)
  float phi0;       // float32(-(pi/2)(1-|1-alpha|)/alpha)
  float sign;       // +1 for alpha > 1
  float abs_am1;    // float32(|alpha-1|)
  float neg_inv_a;  // float32(-1/alpha)
  float tail_exp;   // float32((1-alpha)/alpha)
};

inline CmsConst make_cms(double a) {
  CmsConst c;
  c.alpha = (float)a;
  c.phi0 = (float)(-(3.141592653589793 / 2.0) * (1.0 - fabs(1.0 - a)) / a);
  c.sign = a > 1.0 ? 1.0f : -1.0f;
  c.abs_am1 = (float)fabs(a - 1.0);
  c.neg_inv_a = (float)(-1.0 / a);
  c.tail_exp = (float)((1.0 - a) / a);
  return c;
}

__device__ __forceinline__ float cms_f32(float u0, float u1, const CmsConst& k) {
  const float eps = 1e-6f;
  const float pi_f = 3.14159265358979323846f;
  float phi = __fmul_rn(__fsub_rn(u0, 0.5f), pi_f);
  float R = fmaxf(u1, eps);
  R = -logf(R);
  float factor = fmaxf(cosf(phi), eps);
  factor = __fmul_rn(factor, k.abs_am1);
  factor = powf(factor, k.neg_inv_a);
  float shifted = __fmul_rn(__fsub_rn(phi, k.phi0), k.alpha);
  float p = __fsub_rn(phi, shifted);
  p = fmaxf(cosf(p), eps);
  p = __fdiv_rn(p, R);
  p = powf(p, k.tail_exp);
  float s = sinf(shifted);
  s = __fmul_rn(s, k.sign);
  s = __fmul_rn(s, factor);
  s = __fmul_rn(s, p);
  return s;
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
