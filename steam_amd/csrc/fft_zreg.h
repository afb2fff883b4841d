// Register-resident mixed-radix Stockham FFT for the z lines of the
// convolution (compile-time length N = n_fft / 2, the half-length complex
// transform of the real z line).
//
// A line is held by P = N / V lanes of one wavefront; lane q holds the V
// values with natural indices q + m P (m = 0..V-1).  A stage of radix R is
// N / R butterflies; lane q does butterflies j = q + b P (b < V / R), whose
// inputs j + r N/R are exactly its own values v[b + (V/R) r], so every stage
// computes in registers.  Between stages the outputs go through a
// wave-private LDS line (Stockham autosort order) and come back in the
// natural pattern; the last stage's outputs are already natural.  Only the
// lanes of one wavefront share a line, so the exchanges need no workgroup
// barrier.  Small DFTs are unrolled at compile time with literal twiddles;
// stage twiddles come from an LDS table of exp(-2 pi i k / 2N).
#pragma once

#include <array>

#include "fft.h"

namespace steam {

// ---------------------------------------------------------------------------
// compile-time cos/sin of 2 pi k / n (quadrant reduction, Taylor series in
// double precision, rounded once to float)
// ---------------------------------------------------------------------------
constexpr double kPiD = 3.14159265358979323846264338327950288;

constexpr double ct_sin(double x) {
  double x2 = x * x, t = x, s = x;
  for (int i = 1; i <= 14; ++i) {
    t *= -x2 / ((2.0 * i) * (2.0 * i + 1.0));
    s += t;
  }
  return s;
}
constexpr double ct_cos(double x) {
  double x2 = x * x, t = 1.0, s = 1.0;
  for (int i = 1; i <= 14; ++i) {
    t *= -x2 / ((2.0 * i - 1.0) * (2.0 * i));
    s += t;
  }
  return s;
}
// cos and sin of 2 pi k / n
constexpr void ct_cis(long k, long n, double& c, double& s) {
  k %= n;
  if (k < 0) k += n;
  const long qd = (4 * k) / n, rem = 4 * k - qd * n;
  const double x = (kPiD / 2.0) * (double)rem / (double)n;
  const double cx = ct_cos(x), sx = ct_sin(x);
  switch (qd) {
    case 0: c = cx; s = sx; break;
    case 1: c = -sx; s = cx; break;
    case 2: c = -cx; s = -sx; break;
    default: c = sx; s = -cx; break;
  }
}
// table of W_R^e = exp(-2 pi i e / R): [2e] = cos, [2e+1] = -sin
template <int R>
constexpr std::array<float, 2 * R> ct_wtab() {
  std::array<float, 2 * R> a{};
  for (int e = 0; e < R; ++e) {
    double c = 0, s = 0;
    ct_cis(e, R, c, s);
    a[2 * e] = (float)c;
    a[2 * e + 1] = (float)(-s);
  }
  return a;
}

// x * W_R^e (conjugate twiddle for the inverse), trivial rotations exact
template <int R, bool INV>
__device__ __forceinline__ float2 rot_by(float2 x, int e) {
  e %= R;
  if (e == 0) return x;
  if (4 * e == R) return INV ? make_float2(-x.y, x.x) : make_float2(x.y, -x.x);
  if (2 * e == R) return make_float2(-x.x, -x.y);
  if (4 * e == 3 * R) return INV ? make_float2(x.y, -x.x) : make_float2(-x.y, x.x);
  constexpr std::array<float, 2 * R> W = ct_wtab<R>();
  const float c = W[2 * e], s = INV ? -W[2 * e + 1] : W[2 * e + 1];
  return make_float2(x.x * c - x.y * s, x.x * s + x.y * c);
}

// ---------------------------------------------------------------------------
// in-register DFTs of compile-time size (forward exp(-2 pi i n k / R))
// ---------------------------------------------------------------------------
template <int R, bool INV>
struct Dft;

template <bool INV>
struct Dft<2, INV> {
  static __device__ __forceinline__ void run(float2 (&v)[2]) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <bool INV>
struct Dft<3, INV> {
  static __device__ __forceinline__ void run(float2 (&v)[3]) {
    const float s3 = 0.86602540378443865f;
    const float2 a = v[0], t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    const float2 m = make_float2(a.x - 0.5f * t1.x, a.y - 0.5f * t1.y);
    // forward: X1 = m - i s3 t2, X2 = m + i s3 t2
    const float2 it = INV ? make_float2(-s3 * t2.y, s3 * t2.x) : make_float2(s3 * t2.y, -s3 * t2.x);
    v[0] = cadd(a, t1);
    v[1] = cadd(m, it);
    v[2] = csub(m, it);
  }
};

template <bool INV>
struct Dft<4, INV> {
  static __device__ __forceinline__ void run(float2 (&v)[4]) {
    const float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const float2 t2 = cadd(v[1], v[3]), t3 = csub(v[1], v[3]);
    const float2 mt3 = INV ? make_float2(-t3.y, t3.x) : make_float2(t3.y, -t3.x);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, mt3);
    v[3] = csub(t1, mt3);
  }
};

template <bool INV>
struct Dft<5, INV> {
  static __device__ __forceinline__ void run(float2 (&v)[5]) {
    const float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;
    const float s1 = 0.95105651629515357f, s2 = 0.58778525229247313f;
    const float2 a = v[0];
    const float2 p1 = cadd(v[1], v[4]), m1 = csub(v[1], v[4]);
    const float2 p2 = cadd(v[2], v[3]), m2 = csub(v[2], v[3]);
    const float2 r1 = make_float2(a.x + c1 * p1.x + c2 * p2.x, a.y + c1 * p1.y + c2 * p2.y);
    const float2 r2 = make_float2(a.x + c2 * p1.x + c1 * p2.x, a.y + c2 * p1.y + c1 * p2.y);
    // sine parts: forward X_k = r_k - i (s_k1 m1 + s_k2 m2)
    const float2 q1 = make_float2(s1 * m1.x + s2 * m2.x, s1 * m1.y + s2 * m2.y);
    const float2 q2 = make_float2(s2 * m1.x - s1 * m2.x, s2 * m1.y - s1 * m2.y);
    const float2 iq1 = INV ? make_float2(-q1.y, q1.x) : make_float2(q1.y, -q1.x);
    const float2 iq2 = INV ? make_float2(-q2.y, q2.x) : make_float2(q2.y, -q2.x);
    v[0] = make_float2(a.x + p1.x + p2.x, a.y + p1.y + p2.y);
    v[1] = cadd(r1, iq1);
    v[4] = csub(r1, iq1);
    v[2] = cadd(r2, iq2);
    v[3] = csub(r2, iq2);
  }
};

template <bool INV>
struct Dft<8, INV> {
  static __device__ __forceinline__ void run(float2 (&v)[8]) {
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, INV>::run(e);
    Dft<4, INV>::run(o);
    const float c = 0.70710678118654752f;
    // o1 * W_8^1, o2 * W_8^2, o3 * W_8^3 (conjugated for the inverse)
    const float2 o1 = INV ? make_float2(c * (o[1].x - o[1].y), c * (o[1].x + o[1].y))
                          : make_float2(c * (o[1].x + o[1].y), c * (o[1].y - o[1].x));
    const float2 o2 = INV ? make_float2(-o[2].y, o[2].x) : make_float2(o[2].y, -o[2].x);
    const float2 o3 = INV ? make_float2(-c * (o[3].x + o[3].y), c * (o[3].x - o[3].y))
                          : make_float2(c * (o[3].y - o[3].x), -c * (o[3].x + o[3].y));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
};

// composite R = A * B (Cooley-Tukey, input n = B n1 + n2, output k1 + A k2)
template <int A, int B, bool INV>
struct DftCT {
  static __device__ __forceinline__ void run(float2 (&v)[A * B]) {
    constexpr int R = A * B;
    float2 y[R];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) {
      float2 a[A];
#pragma unroll
      for (int n1 = 0; n1 < A; ++n1) a[n1] = v[B * n1 + n2];
      Dft<A, INV>::run(a);
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) y[k1 * B + n2] = rot_by<R, INV>(a[k1], n2 * k1);
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
      float2 b[B];
#pragma unroll
      for (int n2 = 0; n2 < B; ++n2) b[n2] = y[k1 * B + n2];
      Dft<B, INV>::run(b);
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = b[k2];
    }
  }
};

template <bool INV> struct Dft<6, INV> : DftCT<2, 3, INV> {};
template <bool INV> struct Dft<10, INV> : DftCT<2, 5, INV> {};
template <bool INV> struct Dft<12, INV> : DftCT<4, 3, INV> {};
template <bool INV> struct Dft<16, INV> : DftCT<4, 4, INV> {};
template <bool INV> struct Dft<20, INV> : DftCT<4, 5, INV> {};
template <bool INV> struct Dft<24, INV> : DftCT<8, 3, INV> {};

// ---------------------------------------------------------------------------
// wave-level LDS ordering: the lanes of one wavefront exchange through LDS;
// LDS serves a wave's accesses in issue order, so a compiler barrier plus a
// drained LDS counter is all the exchange needs
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// padded LDS slot of line index i (one pad slot per 8 values)
__device__ __forceinline__ int zpad(int i) { return i + (i >> 3); }

// line pitch in LDS (complex): padded length rounded to 8 mod 16 so that the
// lines of a wave start in alternating bank halves
constexpr int zline_pitch(int n) {
  const int p = n + n / 8;
  return p + ((8 - p % 16) + 16) % 16;
}

// stage NS (product of earlier radices) of radix R, then the rest
template <int N, int V, bool INV, int NS, int R, int... REST>
struct ZStages {
  static __device__ __forceinline__ void run(float2 (&v)[V], int q, float2* line,
                                             const float2* tw2) {
    constexpr int P = N / V;
    constexpr int NB = V / R;
    static_assert(V % R == 0, "radix must divide the values per lane");
    constexpr int STEP = N / (NS * R);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = q + b * P;
      const int k = (NS == 1) ? 0 : j % NS;
      float2 u[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float2 x = v[b + NB * r];
        if (NS > 1 && r > 0) {
          float2 w = tw2[2 * r * k * STEP];  // W_N^{r k STEP} = W_2N^{2 r k STEP}
          if (INV) w.y = -w.y;
          x = cmul(x, w);
        }
        u[r] = x;
      }
      Dft<R, INV>::run(u);
#pragma unroll
      for (int r = 0; r < R; ++r) v[b + NB * r] = u[r];
    }
    if constexpr (sizeof...(REST) > 0) {
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int j = q + b * P;
        const int k = (NS == 1) ? 0 : j % NS;
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) line[zpad(base + r * NS)] = v[b + NB * r];
      }
      wave_lds_sync();
#pragma unroll
      for (int m = 0; m < V; ++m) v[m] = line[zpad(q + m * P)];
      wave_lds_sync();
      ZStages<N, V, INV, NS * R, REST...>::run(v, q, line, tw2);
    }
  }
};

// The same stages for x/y lines, with the line in LDS at
// lds[zpad(i) * ls + loff] and stage twiddles from a table of W_N^k (k < N).
// WG: the lanes of a line span wavefronts (a column of the x pass), so an
// exchange is bracketed by workgroup barriers; otherwise the line is
// wave-private as in ZStages.  Every thread of the workgroup calls it.
template <int N, int V, bool INV, bool WG, int NS, int R, int... REST>
struct XYStages {
  static __device__ __forceinline__ void run(float2 (&v)[V], int q, float2* lds, int ls,
                                             int loff, const float2* tw) {
    constexpr int P = N / V;
    constexpr int NB = V / R;
    static_assert(V % R == 0, "radix must divide the values per lane");
    constexpr int STEP = N / (NS * R);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = q + b * P;
      const int k = (NS == 1) ? 0 : j % NS;
      float2 u[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float2 x = v[b + NB * r];
        if (NS > 1 && r > 0) {
          float2 w = tw[r * k * STEP];
          if (INV) w.y = -w.y;
          x = cmul(x, w);
        }
        u[r] = x;
      }
      Dft<R, INV>::run(u);
#pragma unroll
      for (int r = 0; r < R; ++r) v[b + NB * r] = u[r];
    }
    if constexpr (sizeof...(REST) > 0) {
      if (WG) __syncthreads();  // the previous exchange's reads are done
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int j = q + b * P;
        const int k = (NS == 1) ? 0 : j % NS;
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) lds[zpad(base + r * NS) * ls + loff] = v[b + NB * r];
      }
      if (WG) __syncthreads(); else wave_lds_sync();
#pragma unroll
      for (int m = 0; m < V; ++m) v[m] = lds[zpad(q + m * P) * ls + loff];
      if (!WG) wave_lds_sync();
      XYStages<N, V, INV, WG, NS * R, REST...>::run(v, q, lds, ls, loff, tw);
    }
  }
};

// a z-line plan: length N, V values per lane, radices in stage order
template <int N_, int V_, int... RS>
struct ZPlan {
  static constexpr int N = N_;
  static constexpr int V = V_;
  static constexpr int P = N_ / V_;
  static constexpr int LINES = 64 / P;          // lines per wavefront
  static constexpr int PITCH = zline_pitch(N_);  // LDS complex per line
  static_assert(N_ % V_ == 0 && 64 % (N_ / V_) == 0, "lanes per line must divide 64");
  template <bool INV>
  static __device__ __forceinline__ void fft(float2 (&v)[V_], int q, float2* line,
                                             const float2* tw2) {
    ZStages<N_, V_, INV, 1, RS...>::run(v, q, line, tw2);
  }
  // the same plan over an x/y line (XYStages)
  template <bool INV, bool WG>
  static __device__ __forceinline__ void fft_xy(float2 (&v)[V_], int q, float2* lds, int ls,
                                                int loff, const float2* tw) {
    XYStages<N_, V_, INV, WG, 1, RS...>::run(v, q, lds, ls, loff, tw);
  }
};

}  // namespace steam
