// m2l_kernel.h — far-field M2L on gfx950 (FP64, FMA contraction on).
//
// F_B[n,m] += sum_{k<=p-n} sum_{|l|<=k} conj(M_A[k,l]) Theta_{n+k}^{m+l}(s_B - z_A)
// (crates/fmm translate.rs:47-136). The reference evaluates each accepted
// pair mutually into both cells' locals; here the directed entries are
// grouped by target (sorted (target, source) keys), so every local is
// written by exactly one reduction, in a fixed order, with no atomics.
//
// Mapping: one thread per chunk of <= M2L_CH sources of one target. Per
// source the thread streams Theta degree by degree (three-term recursion,
// two previous columns in registers) against the source multipole held in
// registers, accumulating the m >= 0 half of F (the m < 0 half follows from
// F_n^{-m} = (-1)^m conj(F_n^m)). P is a template parameter so every index
// is static and all operands live in registers. Partials of the chunks of a
// target are then summed in chunk order by k_m2l_reduce.
#pragma once
#include "common.h"
#include "internal.h"

namespace fmmh {

#ifndef FMMH_M2L_CH
#define FMMH_M2L_CH 16
#endif
constexpr int M2L_CH = FMMH_M2L_CH;  // sources per chunk (one thread's serial work)

__device__ __forceinline__ void cmac(cplx& f, double ar, double ai, double br, double bi) {
  f.re = fma(ar, br, f.re);
  f.re = fma(-ai, bi, f.re);
  f.im = fma(ar, bi, f.im);
  f.im = fma(ai, br, f.im);
}

// F (m>=0 half, hpos layout) += M2L of source multipole M (m>=0 half) at r.
// H = 2: every degree n of F; H = 0 / 1: only even / odd n (the two halves
// of a chunk, run by two waves: about equal work, half the accumulators).
template <int P, int H>
__device__ __forceinline__ void m2l_accumulate(double x, double y, double z, double soft2,
                                               const cplx* M, cplx* F) {
  // Softening: the harmonic derivatives of the Plummer kernel are those of
  // 1/r times (r/s)^{2n+1}, s^2 = r^2 + eps^2; in the recursion that is
  // 1/r^2 -> 1/s^2 and the (n-2) term scaled by ks = r^2/s^2 (soft2 = 0:
  // ks = 1, pure 1/r).
  const double rr2 = x * x + y * y + z * z;
  const double r2 = rr2 + soft2;
  const double ir = rsqrt(r2);
  const double ir2 = ir * ir;
  const double ks = soft2 > 0.0 ? rr2 * ir2 : 1.0;
  cplx T[P + 1], T1[P + 1], T2[P + 1];
#pragma unroll
  for (int s = 0; s <= P; s++) {
    // Theta_s^q, q = 0..s (harmonic.rs:91-129 recursion, Dehnen eq. 52/54)
    if (s == 0) {
      T[0] = cplx{ir, 0.0};
    } else {
      const double a = (double)(2 * s - 1) * z;
#pragma unroll
      for (int q = 0; q <= P; q++) {
        if (q > s) continue;
        if (q + 2 <= s) {
          const double bq = (double)((s - 1) * (s - 1) - q * q) * ks;
          T[q].re = (a * T1[q].re - bq * T2[q].re) * ir2;
          T[q].im = (a * T1[q].im - bq * T2[q].im) * ir2;
        } else if (q == s - 1) {
          T[q].re = a * T1[q].re * ir2;
          T[q].im = a * T1[q].im * ir2;
        } else {  // q == s: (2s-1)(x+iy)/r^2 * Theta_{s-1}^{s-1}
          const double f = (double)(2 * s - 1) * ir2;
          const double wr = x * f, wi = y * f;
          T[q].re = wr * T1[s - 1].re - wi * T1[s - 1].im;
          T[q].im = wr * T1[s - 1].im + wi * T1[s - 1].re;
        }
      }
    }
    // all terms with n + k == s
#pragma unroll
    for (int n = 0; n <= P; n++) {
      if (n > s) continue;
      if (H != 2 && (n & 1) != H) continue;
      const int k = s - n;
#pragma unroll
      for (int m = 0; m <= P; m++) {
        if (m > n) continue;
#pragma unroll
        for (int l = -P; l <= P; l++) {
          if (l < -k || l > k) continue;
          // conj(M_k^l): l >= 0 direct; l < 0 via M_k^l = (-1)^l conj(M_k^{-l})
          double ar, ai;
          if (l >= 0) {
            ar = M[hpos(k, l)].re;
            ai = -M[hpos(k, l)].im;
          } else {
            const double sg = ((-l) & 1) ? -1.0 : 1.0;
            ar = sg * M[hpos(k, -l)].re;
            ai = sg * M[hpos(k, -l)].im;
          }
          const int q = m + l;
          double br, bi;
          if (q >= 0) {
            br = T[q].re;
            bi = T[q].im;
          } else {
            const double sg = ((-q) & 1) ? -1.0 : 1.0;
            br = sg * T[-q].re;
            bi = -sg * T[-q].im;
          }
          cmac(F[hpos(n, m)], ar, ai, br, bi);
        }
      }
    }
#pragma unroll
    for (int q = 0; q <= P; q++) {
      T2[q] = T1[q];
      T1[q] = T[q];
    }
  }
}

struct M2lChunk {
  uint32_t target;
  uint32_t begin;
  uint32_t end;
};

template <int P, int H>
__device__ __forceinline__ void m2l_chunk(const M2lChunk ck, uint32_t ci,
                                          const uint64_t* __restrict__ skeys, uint64_t smask,
                                          const double* __restrict__ ccx,
                                          const double* __restrict__ ccy,
                                          const double* __restrict__ ccz,
                                          const cplx* __restrict__ moments, double soft2,
                                          cplx* __restrict__ part) {
  constexpr int D = hdim(P), NF = hnpos(P);
  const double tx = ccx[ck.target], ty = ccy[ck.target], tz = ccz[ck.target];
  cplx F[NF];
#pragma unroll
  for (int i = 0; i < NF; i++) F[i] = cz();
  for (uint32_t j = ck.begin; j < ck.end; j++) {
    const uint32_t src = (uint32_t)(skeys[j] & smask);
    const cplx* ms = moments + (size_t)src * D;
    cplx M[NF];
#pragma unroll
    for (int k = 0; k <= P; k++)
#pragma unroll
      for (int l = 0; l <= k; l++) M[hpos(k, l)] = ms[hidx(k, l)];
    m2l_accumulate<P, H>(tx - ccx[src], ty - ccy[src], tz - ccz[src], soft2, M, F);
  }
  cplx* out = part + (size_t)ci * NF;
#pragma unroll
  for (int n = 0; n <= P; n++) {
    if (H != 2 && (n & 1) != H) continue;
#pragma unroll
    for (int m = 0; m <= n; m++) out[hpos(n, m)] = F[hpos(n, m)];
  }
}

// SPLIT: two blocks per 64 chunks, the even-n and the odd-n half of F (the
// pair of blocks is adjacent after the XCD remap, so on one XCD's L2).
// Orders whose accumulators are split over two waves (register pressure),
// and the waves per SIMD the split kernels are compiled for.
#ifndef FMMH_M2L_SPLIT_FROM
#define FMMH_M2L_SPLIT_FROM 8
#endif
#ifndef FMMH_M2L_SPLIT_WAVES
#define FMMH_M2L_SPLIT_WAVES 1
#endif

template <int P, bool SPLIT>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(SPLIT ? FMMH_M2L_SPLIT_WAVES : 1))) k_m2l(const M2lChunk* __restrict__ chunks, uint32_t nchunks,
                                           const uint64_t* __restrict__ skeys, uint64_t smask,
                                           const double* __restrict__ ccx,
                                           const double* __restrict__ ccy,
                                           const double* __restrict__ ccz,
                                           const cplx* __restrict__ moments, double soft2,
                                           cplx* __restrict__ part) {
  const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
  if constexpr (SPLIT) {
    const uint32_t ci = (lb >> 1) * blockDim.x + threadIdx.x;
    if (ci >= nchunks) return;
    if (lb & 1)
      m2l_chunk<P, 1>(chunks[ci], ci, skeys, smask, ccx, ccy, ccz, moments, soft2, part);
    else
      m2l_chunk<P, 0>(chunks[ci], ci, skeys, smask, ccx, ccy, ccz, moments, soft2, part);
  } else {
    const uint32_t ci = lb * blockDim.x + threadIdx.x;
    if (ci >= nchunks) return;
    m2l_chunk<P, 2>(chunks[ci], ci, skeys, smask, ccx, ccy, ccz, moments, soft2, part);
  }
}

template <int P>
void m2l_t(const M2lChunk* chunks, uint32_t nchunks, const uint64_t* skeys, uint64_t smask,
                  DevCells c, const cplx* moments, double soft2, cplx* part, hipStream_t s) {
  if (!nchunks) return;
  constexpr bool split = P >= FMMH_M2L_SPLIT_FROM;
  const uint32_t nb = (nchunks + 63) / 64 * (split ? 2 : 1);
  hipLaunchKernelGGL((k_m2l<P, split>), dim3(nb), dim3(64), 0, s, chunks, nchunks, skeys, smask,
                     c.cx, c.cy, c.cz, moments, soft2, part);
}


// One translation unit per range of orders instantiates the launches.
#define FMMH_M2L_INSTANTIATE(P)                                                             \
  template void m2l_t<P>(const M2lChunk*, uint32_t, const uint64_t*, uint64_t, DevCells,    \
                         const cplx*, double, cplx*, hipStream_t);

}  // namespace fmmh
