// Register-resident radix-8 Stockham FFT for power-of-two lines
// (N = 2^LOGN, 64 <= N <= 2048).
//
// A line is held by P = N/8 threads; thread q holds the 8 values with natural
// indices q + m P (m = 0..7), which is both the input pattern of every stage
// and the output pattern of the last one.  So the first stage reads straight
// from HBM and the last stage's results are written straight back, both
// coalesced across q; only the NST - 1 intermediate exchanges go through LDS.
// Stages are radix 8, with a final radix-4 or radix-2 stage when LOGN is not a
// multiple of 3 (each thread then does 2 or 4 butterflies of that stage).
// LDS index i is stored at i + i/8 (one pad slot per 8 values), which makes
// the strided writes of the early stages bank-conflict free.
#pragma once

#include "fft.h"

namespace steam {

template <int LOGN>
struct RegFFT {
  static constexpr int N = 1 << LOGN;
  static constexpr int P = N / 8;
  static constexpr int N8 = LOGN / 3;
  static constexpr int TAIL = 1 << (LOGN % 3);
  static constexpr int NST = N8 + (TAIL > 1 ? 1 : 0);
  static constexpr int LDSN = N + N / 8;  // padded line length
};

__device__ __forceinline__ int pad8(int i) { return i + (i >> 3); }

// v[r] *= w[r] (conjugated for the inverse), r = 1..7
__device__ __forceinline__ void mul_twiddles8(float2 (&v)[8], const float2 (&w)[8], bool inv) {
#pragma unroll
  for (int r = 1; r < 8; ++r) {
    float2 x = w[r];
    if (inv) x.y = -x.y;
    v[r] = cmul(v[r], x);
  }
}

// The seven twiddles W^(r e), r = 1..7, of a stage from its bases W^e, W^2e
// and W^4e: the other four are products, at most two deep, so each stays
// within a few float32 ulps of the correctly rounded table value.
__device__ __forceinline__ void expand_twiddles8(const float2 (&b)[3], float2 (&w)[8]) {
  w[0] = make_float2(1.f, 0.f);
  w[1] = b[0];
  w[2] = b[1];
  w[4] = b[2];
  w[3] = cmul(b[0], b[1]);
  w[5] = cmul(b[0], b[2]);
  w[6] = cmul(b[1], b[2]);
  w[7] = cmul(w[3], b[2]);
}

// Twiddle bases of a thread lane q for the radix-8 stages s >= 1: W^e, W^2e
// and W^4e with e = k * step (k = q mod 8^s, step = N / 8^(s+1)), read from
// the table once; apply() forms a stage's seven twiddles from them.
template <int LOGN>
struct RegTw {
  static constexpr int NS = RegFFT<LOGN>::N8 - 1;
  float2 b[NS > 0 ? NS : 1][3];
  __device__ __forceinline__ void apply(float2 (&v)[8], int s, bool inv) const {
    float2 w[8];
    expand_twiddles8(b[s], w);
    mul_twiddles8(v, w, inv);
  }
};

// The same with every stage's seven twiddles formed once, for a pass that
// transforms twice with one lane's twiddles (forward and inverse): the same
// values, without re-forming the products in the second transform.
template <int LOGN>
struct RegTwFull {
  static constexpr int NS = RegFFT<LOGN>::N8 - 1;
  float2 w[NS > 0 ? NS : 1][8];
  __device__ __forceinline__ explicit RegTwFull(const RegTw<LOGN>& t) {
#pragma unroll
    for (int s = 0; s < NS; ++s) expand_twiddles8(t.b[s], w[s]);
  }
  __device__ __forceinline__ void apply(float2 (&v)[8], int s, bool inv) const {
    mul_twiddles8(v, w[s], inv);
  }
};

template <int LOGN>
__device__ __forceinline__ void reg_tw_load(RegTw<LOGN>& t, int q,
                                            const float2* __restrict__ tw) {
  using F = RegFFT<LOGN>;
#pragma unroll
  for (int s = 1; s < F::N8; ++s) {
    const int Ns = 1 << (3 * s);
    const int e = (q & (Ns - 1)) * (F::N / (Ns * 8));
    t.b[s - 1][0] = tw[e];
    t.b[s - 1][1] = tw[2 * e];
    t.b[s - 1][2] = tw[4 * e];
  }
}

// Transform the line held in v (thread q of the line; twb its twiddle bases,
// reg_tw_load).  LDS element i of the line lives at
// lds[pad8(i) * lstride + loff].  Every thread of the workgroup must call
// this (it synchronises the workgroup between stages).
template <int LOGN>
__device__ __forceinline__ void fft_reg(float2 (&v)[8], int q, float2* lds,
                                        int lstride, int loff,
                                        const float2* __restrict__ tw,
                                        const RegTw<LOGN>& twb, bool inv) {
  using F = RegFFT<LOGN>;
  constexpr int P = F::P;
#pragma unroll
  for (int s = 0; s < F::N8; ++s) {
    const int Ns = 1 << (3 * s);
    const int k = q & (Ns - 1);
    if (s > 0) twb.apply(v, s - 1, inv);
    dft_small<8>(v, inv);
    const bool last = (s == F::NST - 1);
    if (!last) {
      const int base = (q - k) * 8 + k;
      __syncthreads();
#pragma unroll
      for (int r = 0; r < 8; ++r) lds[pad8(base + r * Ns) * lstride + loff] = v[r];
      __syncthreads();
#pragma unroll
      for (int m = 0; m < 8; ++m) v[m] = lds[pad8(q + m * P) * lstride + loff];
    }
  }
  if constexpr (F::TAIL > 1) {
    constexpr int R = F::TAIL;
    constexpr int NB = 8 / R;
    // butterfly b: j = q + b P (j < Ns = N/R), inputs v[b + NB r]
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = q + b * P;
      float2 u[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float2 x = v[b + NB * r];
        if (r > 0) {
          float2 w = tw[r * j];
          if (inv) w.y = -w.y;
          x = cmul(x, w);
        }
        u[r] = x;
      }
      dft_small<R>(u, inv);
#pragma unroll
      for (int r = 0; r < R; ++r) v[b + NB * r] = u[r];
    }
  }
}

// The same transform with each thread holding C lanes of the line: lane
// q0 + c * qstep (c < C) holds v[c].  The LDS exchanges of all C lanes share
// one pair of workgroup barriers per stage.
template <int LOGN, int C, class TW>
__device__ __forceinline__ void fft_reg_lanes(float2 (&v)[C][8], int q0, int qstep,
                                              float2* lds, int lstride, int loff,
                                              const float2* __restrict__ tw,
                                              const TW (&twb)[C], bool inv) {
  using F = RegFFT<LOGN>;
  constexpr int P = F::P;
#pragma unroll
  for (int s = 0; s < F::N8; ++s) {
    const int Ns = 1 << (3 * s);
#pragma unroll
    for (int c = 0; c < C; ++c) {
      if (s > 0) twb[c].apply(v[c], s - 1, inv);
      dft_small<8>(v[c], inv);
    }
    const bool last = (s == F::NST - 1);
    if (!last) {
      __syncthreads();
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int q = q0 + c * qstep;
        const int k = q & (Ns - 1);
        const int base = (q - k) * 8 + k;
#pragma unroll
        for (int r = 0; r < 8; ++r) lds[pad8(base + r * Ns) * lstride + loff] = v[c][r];
      }
      __syncthreads();
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int q = q0 + c * qstep;
#pragma unroll
        for (int m = 0; m < 8; ++m) v[c][m] = lds[pad8(q + m * P) * lstride + loff];
      }
    }
  }
  if constexpr (F::TAIL > 1) {
    constexpr int R = F::TAIL;
    constexpr int NB = 8 / R;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int q = q0 + c * qstep;
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int j = q + b * P;
        float2 u[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          float2 x = v[c][b + NB * r];
          if (r > 0) {
            float2 w = tw[r * j];
            if (inv) w.y = -w.y;
            x = cmul(x, w);
          }
          u[r] = x;
        }
        dft_small<R>(u, inv);
#pragma unroll
        for (int r = 0; r < R; ++r) v[c][b + NB * r] = u[r];
      }
    }
  }
}

}  // namespace steam
