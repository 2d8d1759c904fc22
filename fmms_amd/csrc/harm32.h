// harm32.h — FP32 solid-harmonic helpers of the mixed form (radius-scaled
// operands): the Upsilon recursion of harmonic.rs:141-173 degree by degree,
// and complex multiply-adds on the m >= 0 halves of the expansions
// (X_n^{-m} = (-1)^m conj(X_n^m)).
#pragma once
#include "common.h"

namespace fmmh {

// Upsilon_k^q for q = 0..k from the two previous degrees (U1 = k-1, U2 = k-2).
template <int P>
__device__ __forceinline__ void upsilon_degree32(int k, float x, float y, float z, float r2,
                                                 cplxf* U, const cplxf* U1, const cplxf* U2) {
  if (k == 0) {
    U[0] = cplxf{1.0f, 0.0f};
    return;
  }
  const float a = (float)(2 * k - 1) * z;
#pragma unroll
  for (int q = 0; q <= P; q++) {
    if (q > k) continue;
    if (q == k) {
      const float f = 1.0f / (float)(2 * k);
      const float wr = x * f, wi = y * f;
      U[q].re = wr * U1[k - 1].re - wi * U1[k - 1].im;
      U[q].im = wr * U1[k - 1].im + wi * U1[k - 1].re;
    } else {
      const float ic = 1.0f / (float)(k * k - q * q);
      if (q + 2 <= k) {
        U[q].re = (a * U1[q].re - r2 * U2[q].re) * ic;
        U[q].im = (a * U1[q].im - r2 * U2[q].im) * ic;
      } else {
        U[q].re = a * U1[q].re * ic;
        U[q].im = a * U1[q].im * ic;
      }
    }
  }
}

__device__ __forceinline__ void conj_y32(const cplxf* U, int l, float* ar, float* ai) {
  if (l >= 0) {
    *ar = U[l].re;
    *ai = -U[l].im;
  } else {
    const float sg = ((-l) & 1) ? -1.0f : 1.0f;
    *ar = sg * U[-l].re;
    *ai = sg * U[-l].im;
  }
}

template <int P>
__device__ __forceinline__ void f_at32(const cplxf* F, int s, int q, float* br, float* bi) {
  if (q >= 0) {
    *br = F[hpos(s, q)].re;
    *bi = F[hpos(s, q)].im;
  } else {
    const float sg = ((-q) & 1) ? -1.0f : 1.0f;
    *br = sg * F[hpos(s, -q)].re;
    *bi = -sg * F[hpos(s, -q)].im;
  }
}

__device__ __forceinline__ void cmac2f(cplxf& f, float ar, float ai, float br, float bi) {
  f.re = fmaf(ar, br, f.re);
  f.re = fmaf(-ai, bi, f.re);
  f.im = fmaf(ar, bi, f.im);
  f.im = fmaf(ai, br, f.im);
}

// Upsilon_k^l itself (not conjugated) from the m >= 0 column U.
__device__ __forceinline__ void y32(const cplxf* U, int l, float* ar, float* ai) {
  if (l >= 0) {
    *ar = U[l].re;
    *ai = U[l].im;
  } else {
    const float sg = ((-l) & 1) ? -1.0f : 1.0f;
    *ar = sg * U[-l].re;
    *ai = -sg * U[-l].im;
  }
}

}  // namespace fmmh
