// LDS-resident multi-line Stockham FFT for gfx950 (64-wide wavefronts).
//
// A workgroup of NTH threads holds T lines of length N (N*T <= EM*nth complex
// values; EM values per thread) in LDS and transforms all of them together.  Mixed radix
// 2/3/4/5/7/8 stages (Stockham autosort, natural order in and out); a length
// with prime factors above 7 takes one direct stage of their product.  Each
// stage stages its butterflies' inputs in registers, synchronises, and
// writes the outputs back in place, so one LDS buffer suffices.
//
// Element (line t, index n) lives at buf[n*ldn + t*ldt].  TFAST selects the
// thread mapping: consecutive threads take consecutive t (lines interleaved
// in LDS, [n][t] layout) or consecutive n within a line ([t][n] layout);
// either way consecutive lanes touch consecutive LDS words.
#pragma once

#include "common.h"

namespace steam {

constexpr int FFT_NTH = 256;              // threads per FFT workgroup
constexpr int FFT_EMAX = 16;              // complex values per thread
constexpr int FFT_LDSE = FFT_NTH * FFT_EMAX;  // 4096 complex = 32 KiB
constexpr int FFT_MAXST = 6;   // stages per length (factor 8 first)

struct AxisPlan {
  int n;                   // transform length
  int nst;                 // number of stages
  int radix[FFT_MAXST];    // 2,3,4,5,7,8; above 8: direct stage of that radix;
                           // 0: direct DFT of the whole line
  const float2* tw;        // n entries exp(-2 pi i m / n)
  // Bluestein (chirp-z) form, for an x/y length whose direct stage would be
  // long: bm > 0 is the smooth length of the circular convolution that
  // replaces it (bm >= 2n - 1), with its own stages and tables
  int bm, bnst;
  int bradix[FFT_MAXST];
  const float2* btw;       // bm entries exp(-2 pi i m / bm)
  const float2* chirp;     // n entries exp(-pi i m^2 / n)
  const float2* bk;        // bm entries: the transform of the wrapped conjugate
                           // chirp, divided by bm
};

// constant rotations for the odd radices: cos/sin(2 pi m / R)
__device__ __forceinline__ float2 rot_const(int R, int m) {
  // tables for R = 3, 5, 7 (m < R)
  const float c3[3] = {1.0f, -0.5f, -0.5f};
  const float s3[3] = {0.0f, 0.86602540378443865f, -0.86602540378443865f};
  const float c5[5] = {1.0f, 0.30901699437494742f, -0.80901699437494742f,
                       -0.80901699437494742f, 0.30901699437494742f};
  const float s5[5] = {0.0f, 0.95105651629515357f, 0.58778525229247313f,
                       -0.58778525229247313f, -0.95105651629515357f};
  const float c7[7] = {1.0f,
                       0.62348980185873353f,
                       -0.22252093395631440f,
                       -0.90096886790241913f,
                       -0.90096886790241913f,
                       -0.22252093395631440f,
                       0.62348980185873353f};
  const float s7[7] = {0.0f,
                       0.78183148246802981f,
                       0.97492791218182361f,
                       0.43388373911755812f,
                       -0.43388373911755812f,
                       -0.97492791218182361f,
                       -0.78183148246802981f};
  if (R == 3) return make_float2(c3[m], s3[m]);
  if (R == 5) return make_float2(c5[m], s5[m]);
  return make_float2(c7[m], s7[m]);
}

// In-register DFT of size R; forward uses exp(-2 pi i rq/R).
template <int R>
__device__ __forceinline__ void dft_small(float2* v, bool inv) {
  if constexpr (R == 2) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    float2 t2 = cadd(v[1], v[3]), t3 = csub(v[1], v[3]);
    // forward: y1 = t1 - i t3, y3 = t1 + i t3
    float2 mt3 = inv ? make_float2(-t3.y, t3.x) : make_float2(t3.y, -t3.x);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, mt3);
    v[3] = csub(t1, mt3);
  } else if constexpr (R == 8) {
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    dft_small<4>(e, inv);
    dft_small<4>(o, inv);
    const float c = 0.70710678118654752f;
    float sg = inv ? 1.0f : -1.0f;
    // w^q = exp(sg * i pi q / 4)
    float2 w1 = make_float2(c, sg * c);
    float2 w3 = make_float2(-c, sg * c);
    float2 o1 = cmul(o[1], w1);
    float2 o2 = inv ? make_float2(-o[2].y, o[2].x) : make_float2(o[2].y, -o[2].x);
    float2 o3 = cmul(o[3], w3);
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  } else {
    float2 y[R];
    float sg = inv ? 1.0f : -1.0f;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      float2 acc = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = rot_const(R, (r * q) % R);
        w.y *= sg;
        acc = cadd(acc, cmul(v[r], w));
      }
      y[q] = acc;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = y[q];
  }
}

template <bool TFAST>
__device__ __forceinline__ void line_index(int b, int T, int logT, int NR,
                                           int& t, int& j) {
  if (TFAST) {
    t = b & (T - 1);
    j = b >> logT;
  } else {
    t = b / NR;
    j = b - t * NR;
  }
}

template <int R, bool TFAST, int EM, int NTH>
__device__ __forceinline__ void fft_stage(float2* buf, int N, int T, int logT, int ldn,
                          int ldt, int Ns, const float2* __restrict__ tw,
                          bool inv, int tid) {
  constexpr int MAXB = (EM + R - 1) / R;
  const int NR = N / R;
  const int nb = NR * T;
  constexpr int nth = NTH;
  float2 v[MAXB][R];
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    int b = tid + i * nth;
    if (b < nb) {
      int t, j;
      line_index<TFAST>(b, T, logT, NR, t, j);
      const float2* src = buf + j * ldn + t * ldt;
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = src[r * NR * ldn];
    }
  }
  __syncthreads();
  const int step = N / (Ns * R);
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    int b = tid + i * nth;
    if (b < nb) {
      int t, j;
      line_index<TFAST>(b, T, logT, NR, t, j);
      int k = j % Ns;
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[r * k * step];
          if (inv) w.y = -w.y;
          v[i][r] = cmul(v[i][r], w);
        }
      }
      dft_small<R>(v[i], inv);
      int base = (j - k) * R + k;
      float2* dst = buf + base * ldn + t * ldt;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[r * Ns * ldn] = v[i][r];
    }
  }
  __syncthreads();
}

template <int R, bool TFAST, int EM, int NTH>
__device__ __noinline__ void fft_stage_call(float2* buf, int N, int T, int logT,
                                           int ldn, int ldt, int Ns,
                                           const float2* __restrict__ tw,
                                           bool inv, int tid) {
  fft_stage<R, TFAST, EM, NTH>(buf, N, T, logT, ldn, ldt, Ns, tw, inv, tid);
}

// Direct stage of a runtime radix R (the product of a length's prime
// factors above 7; R = N is the whole line): output q of a line, written as
// q = (j - k) R + k + s Ns with k = j mod Ns, j < N/R and s < R, is
//   sum_r buf[j + r N/R] exp(-2 pi i (r k / (Ns R) + r s / R)),
// the stage twiddle and the R-point DFT in one exponent mod N of the axis
// table.  Cost: R complex products per element.
template <bool TFAST, int EM, int NTH>
__device__ __noinline__ void fft_direct(float2* buf, int N, int R, int Ns, int T, int logT,
                                        int ldn, int ldt, const float2* __restrict__ tw,
                                        bool inv, int tid) {
  float2 o[EM];
  const int ne = N * T;
  const int NR = N / R;
  const int step = N / (Ns * R);
  constexpr int nth = NTH;
#pragma unroll
  for (int i = 0; i < EM; ++i) {
    int e = tid + i * nth;
    if (e < ne) {
      int t, q;
      line_index<TFAST>(e, T, logT, N, t, q);
      const int k = q % Ns;
      const int s = (q / Ns) % R;
      const int j = (q / (Ns * R)) * Ns + k;
      const float2* src = buf + t * ldt + j * ldn;
      const int dm = (k * step + s * NR) % N;
      float2 acc = make_float2(0.f, 0.f);
      int m = 0;
      for (int r = 0; r < R; ++r) {
        float2 w = tw[m];
        if (inv) w.y = -w.y;
        acc = cadd(acc, cmul(src[r * NR * ldn], w));
        m += dm;
        if (m >= N) m -= N;
      }
      o[i] = acc;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < EM; ++i) {
    int e = tid + i * nth;
    if (e < ne) {
      int t, q;
      line_index<TFAST>(e, T, logT, N, t, q);
      buf[q * ldn + t * ldt] = o[i];
    }
  }
  __syncthreads();
}

// Radix sets a kernel is instantiated for: 0 = {2,4,8}, 1 = {2,3,4,5,8},
// 2 = every stage kind, 3 = the Bluestein form (its own kernels, so the
// others do not carry its registers).  The stage loop is unrolled into fixed slots so the
// register allocator sees straight-line code (a rolled loop around the radix
// switch keeps every case's registers live); the rare radix-7 / direct set
// calls out-of-line stages instead.
inline int radix_set(const AxisPlan& ap) {
  if (ap.bm > 0) return 3;  // the Bluestein form (x/y only)
  int rs = 0;
  for (int s = 0; s < ap.nst; ++s) {
    int r = ap.radix[s];
    if (r == 3 || r == 5) rs = rs < 1 ? 1 : rs;
    if (r == 7 || r == 0 || r > 8) rs = 2;
  }
  return rs;
}

// The stages of a length (its radix list).  Caller synchronises before.
template <bool TFAST, int RSET, bool CALLS = false, int EM = FFT_EMAX, int NTH = FFT_NTH>
__device__ __forceinline__ void fft_lines_stages(float2* buf, const AxisPlan& ap, int T,
                                                 int logT, int ldn, int ldt, bool inv,
                                                 int tid) {
  int Ns = 1;
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    if (s >= ap.nst) break;
    const int R = ap.radix[s];
    if (RSET < 2 && !CALLS) {
      switch (R) {
        case 2: fft_stage<2, TFAST, EM, NTH>(buf, ap.n, T, logT, ldn, ldt, Ns, ap.tw, inv, tid); break;
        case 4: fft_stage<4, TFAST, EM, NTH>(buf, ap.n, T, logT, ldn, ldt, Ns, ap.tw, inv, tid); break;
        case 8: fft_stage<8, TFAST, EM, NTH>(buf, ap.n, T, logT, ldn, ldt, Ns, ap.tw, inv, tid); break;
        case 3: if (RSET >= 1) fft_stage<3, TFAST, EM, NTH>(buf, ap.n, T, logT, ldn, ldt, Ns, ap.tw, inv, tid); break;
        case 5: if (RSET >= 1) fft_stage<5, TFAST, EM, NTH>(buf, ap.n, T, logT, ldn, ldt, Ns, ap.tw, inv, tid); break;
        default: break;
      }
    } else {
      switch (R) {
        case 2: fft_stage_call<2, TFAST, EM, NTH>(buf, ap.n, T, logT, ldn, ldt, Ns, ap.tw, inv, tid); break;
        case 3: fft_stage_call<3, TFAST, EM, NTH>(buf, ap.n, T, logT, ldn, ldt, Ns, ap.tw, inv, tid); break;
        case 4: fft_stage_call<4, TFAST, EM, NTH>(buf, ap.n, T, logT, ldn, ldt, Ns, ap.tw, inv, tid); break;
        case 5: fft_stage_call<5, TFAST, EM, NTH>(buf, ap.n, T, logT, ldn, ldt, Ns, ap.tw, inv, tid); break;
        case 7: fft_stage_call<7, TFAST, EM, NTH>(buf, ap.n, T, logT, ldn, ldt, Ns, ap.tw, inv, tid); break;
        case 8: fft_stage_call<8, TFAST, EM, NTH>(buf, ap.n, T, logT, ldn, ldt, Ns, ap.tw, inv, tid); break;
        default: fft_direct<TFAST, EM, NTH>(buf, ap.n, R == 0 ? ap.n : R, Ns, T, logT, ldn, ldt,
                                            ap.tw, inv, tid); break;
      }
    }
    Ns *= (R == 0) ? ap.n : R;
  }
}

// Bluestein's chirp-z transform of T lines of length N = ap.n, one line at a
// time through a scratch line of M = ap.bm values placed after the lines
// (buf + N * ldn; the launch sizes its LDS for it).  With c[m] =
// exp(-pi i m^2 / N), nk = (n^2 + k^2 - (k - n)^2) / 2 gives
//   X[k] = c[k] sum_n (x[n] c[n]) conj(c[k - n]),
// a linear convolution over |k - n| < N, taken circularly at length M >= 2N - 1
// by forward and unnormalised inverse transforms of M (its stages are 2..8
// radices) with the conjugate chirp's transform (/ M) from the table.  The
// inverse transform is conj(DFT(conj x)).  Cost: two length-M transforms per
// line, against N R complex products for a direct stage of radix R.
// What the Bluestein form reads of its axis plan, passed by value (a
// reference into the kernel's argument block made the caller copy the whole
// descriptor to scratch memory).
struct BlueArgs {
  int n, m, nst;
  int radix[FFT_MAXST];
  const float2* btw;
  const float2* chirp;
  const float2* bk;
};

__device__ __forceinline__ BlueArgs blue_args(const AxisPlan& ap) {
  BlueArgs b;
  b.n = ap.n;
  b.m = ap.bm;
  b.nst = ap.bnst;
#pragma unroll
  for (int s = 0; s < FFT_MAXST; ++s) b.radix[s] = ap.bradix[s];
  b.btw = ap.btw;
  b.chirp = ap.chirp;
  b.bk = ap.bk;
  return b;
}

// The stages of the Bluestein length on one scratch line, inlined (called
// out of line with 32 values per thread they spilled to scratch memory).
template <int NTH>
__device__ __forceinline__ void fft_blue_stages(float2* sc, const BlueArgs& ap, bool inv,
                                                int tid) {
  constexpr int EM = 8192 / NTH;  // values per thread: lengths below 2^13
  const int M = ap.m;
  int Ns = 1;
  // a rolled loop: each radix's stage body appears once
#pragma nounroll
  for (int s = 0; s < ap.nst; ++s) {
    const int R = ap.radix[s];
    switch (R) {
      case 2: fft_stage<2, false, EM, NTH>(sc, M, 1, 0, 1, 0, Ns, ap.btw, inv, tid); break;
      case 3: fft_stage<3, false, EM, NTH>(sc, M, 1, 0, 1, 0, Ns, ap.btw, inv, tid); break;
      case 4: fft_stage<4, false, EM, NTH>(sc, M, 1, 0, 1, 0, Ns, ap.btw, inv, tid); break;
      case 5: fft_stage<5, false, EM, NTH>(sc, M, 1, 0, 1, 0, Ns, ap.btw, inv, tid); break;
      case 7: fft_stage<7, false, EM, NTH>(sc, M, 1, 0, 1, 0, Ns, ap.btw, inv, tid); break;
      default: fft_stage<8, false, EM, NTH>(sc, M, 1, 0, 1, 0, Ns, ap.btw, inv, tid); break;
    }
    Ns *= R;
  }
}

template <int NTH>
__device__ __noinline__ void fft_bluestein(float2* buf, BlueArgs ap, int T, int ldn, int ldt,
                                           bool inv, int tid) {
  const int N = ap.n, M = ap.m;
  float2* sc = buf + (int64_t)N * ldn;
  for (int t = 0; t < T; ++t) {
    float2* line = buf + t * ldt;
    for (int m = tid; m < M; m += NTH) {
      float2 v = make_float2(0.f, 0.f);
      if (m < N) {
        float2 x = line[m * ldn];
        if (inv) x.y = -x.y;
        v = cmul(x, ap.chirp[m]);
      }
      sc[m] = v;
    }
    __syncthreads();
    fft_blue_stages<NTH>(sc, ap, false, tid);
    for (int m = tid; m < M; m += NTH) sc[m] = cmul(sc[m], ap.bk[m]);
    __syncthreads();
    fft_blue_stages<NTH>(sc, ap, true, tid);
    for (int k = tid; k < N; k += NTH) {
      float2 y = cmul(sc[k], ap.chirp[k]);
      if (inv) y.y = -y.y;
      line[k * ldn] = y;
    }
    __syncthreads();
  }
}

// Transform T lines of length ap.n held in LDS.  Caller synchronises before.
// Radix set 3 (an axis with a Bluestein form, ap.bm > 0) takes that form.
template <bool TFAST, int RSET, bool CALLS = false, int EM = FFT_EMAX, int NTH = FFT_NTH>
__device__ __forceinline__ void fft_lines(float2* buf, const AxisPlan& ap, int T,
                                          int logT, int ldn, int ldt, bool inv,
                                          int tid) {
  if constexpr (RSET == 3)
    fft_bluestein<NTH>(buf, blue_args(ap), T, ldn, ldt, inv, tid);
  else
    fft_lines_stages<TFAST, RSET, CALLS, EM, NTH>(buf, ap, T, logT, ldn, ldt, inv, tid);
}

}  // namespace steam
