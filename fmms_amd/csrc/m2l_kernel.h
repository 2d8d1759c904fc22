// m2l_kernel.h — far-field M2L on gfx950 (FP64, FMA contraction on).
//
// F_B[n,m] += sum_{k<=p-n} sum_{|l|<=k} conj(M_A[k,l]) Theta_{n+k}^{m+l}(s_B - z_A)
// (crates/fmm translate.rs:47-136). The reference evaluates each accepted
// pair mutually into both cells' locals; here the directed entries are
// grouped by target (sorted (target, source) keys), so every local is
// written by exactly one reduction, in a fixed order, with no atomics.
//
// Mapping: one thread per chunk of at most m2l_chunk sources of one target. Per
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

// The chunk length (sources per thread) is a context parameter,
// fmmh_tuning.m2l_chunk (default 32).

__device__ __forceinline__ void cmac(cplx& f, double ar, double ai, double br, double bi) {
  f.re = fma(ar, br, f.re);
  f.re = fma(-ai, bi, f.re);
  f.im = fma(ar, bi, f.im);
  f.im = fma(ai, br, f.im);
}

__device__ __forceinline__ void cmac(cplxf& f, float ar, float ai, float br, float bi) {
  f.re = fmaf(ar, br, f.re);
  f.re = fmaf(-ai, bi, f.re);
  f.im = fmaf(ar, bi, f.im);
  f.im = fmaf(ai, br, f.im);
}

// f += a * (br + i bi) for a real factor a.
__device__ __forceinline__ void cmac_real(cplx& f, double a, double br, double bi) {
  f.re = fma(a, br, f.re);
  f.im = fma(a, bi, f.im);
}
__device__ __forceinline__ void cmac_real(cplxf& f, float a, float br, float bi) {
  f.re = fmaf(a, br, f.re);
  f.im = fmaf(a, bi, f.im);
}

__device__ __forceinline__ double m2l_rsqrt(double x) { return rsqrt(x); }
__device__ __forceinline__ float m2l_rsqrt(float x) {
  const float y = rsqrtf(x);
  return y * fmaf(-0.5f * x * y, y, 1.5f);  // one Newton step on the hardware estimate
}

// F (m>=0 half, hpos layout) += M2L of source multipole M (m>=0 half) at r.
// H = 2: every degree n of F; H = 0 / 1: only even / odd n (the two halves
// of a chunk, run by two waves: about equal work, half the accumulators).
template <int P, int H, typename R, typename C>
__device__ __forceinline__ void m2l_accumulate(R x, R y, R z, R soft2, const C* M, C* F) {
  // Softening: the harmonic derivatives of the Plummer kernel are those of
  // 1/r times (r/s)^{2n+1}, s^2 = r^2 + eps^2; in the recursion that is
  // 1/r^2 -> 1/s^2 and the (n-2) term scaled by ks = r^2/s^2 (soft2 = 0:
  // ks = 1, pure 1/r).
  const R rr2 = x * x + y * y + z * z;
  const R r2 = rr2 + soft2;
  const R ir = m2l_rsqrt(r2);
  const R ir2 = ir * ir;
  const R ks = soft2 > R(0) ? rr2 * ir2 : R(1);
  C T[P + 1], T1[P + 1], T2[P + 1];
#pragma unroll
  for (int s = 0; s <= P; s++) {
    // Theta_s^q, q = 0..s (harmonic.rs:91-129 recursion, Dehnen eq. 52/54)
    if (s == 0) {
      T[0] = C{ir, R(0)};
    } else {
      const R a = (R)(2 * s - 1) * z;
#pragma unroll
      for (int q = 0; q <= P; q++) {
        if (q > s) continue;
        if (q + 2 <= s) {
          const R bq = (R)((s - 1) * (s - 1) - q * q) * ks;
          T[q].re = (a * T1[q].re - bq * T2[q].re) * ir2;
          T[q].im = q == 0 ? R(0) : (a * T1[q].im - bq * T2[q].im) * ir2;  // Theta_s^0 is real
        } else if (q == s - 1) {
          T[q].re = a * T1[q].re * ir2;
          T[q].im = q == 0 ? R(0) : a * T1[q].im * ir2;
        } else {  // q == s: (2s-1)(x+iy)/r^2 * Theta_{s-1}^{s-1}
          const R f = (R)(2 * s - 1) * ir2;
          const R wr = x * f, wi = y * f;
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
          R ar, ai;
          if (l >= 0) {
            ar = M[hpos(k, l)].re;
            ai = -M[hpos(k, l)].im;
          } else {
            const R sg = ((-l) & 1) ? R(-1) : R(1);
            ar = sg * M[hpos(k, -l)].re;
            ai = sg * M[hpos(k, -l)].im;
          }
          const int q = m + l;
          R br, bi;
          if (q >= 0) {
            br = T[q].re;
            bi = T[q].im;
          } else {
            const R sg = ((-q) & 1) ? R(-1) : R(1);
            br = sg * T[-q].re;
            bi = -sg * T[-q].im;
          }
          // Theta_s^0 is real (its imaginary part is an exact zero), and so is
          // M_k^0 up to rounding residue in cells built by M2M: those terms
          // take the real-factor product (2 FMAs instead of 4).
          if (q == 0) {
            cmac_real(F[hpos(n, m)], br, ar, ai);  // (ar + i ai) * br
          } else if (l == 0) {
            cmac_real(F[hpos(n, m)], ar, br, bi);  // ar * (br + i bi)
          } else {
            cmac(F[hpos(n, m)], ar, ai, br, bi);
          }
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
    m2l_accumulate<P, H, double, cplx>(tx - ccx[src], ty - ccy[src], tz - ccz[src], soft2, M, F);
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


// ------------------------------------------------------- FP32 (mixed form)
// The same chunked M2L on radius-scaled FP32 operands. Per target t a scale
// lam_t (its radius, floored at a quarter of its octant edge) gives
// dimensionless offsets v = r / lam_t, and by homogeneity
// Theta_s(r) = lam_t^-(s+1) Theta_s(v) (softened: eps -> eps / lam_t), so
//   F_n = lam_t^-(n+1) sum_k (M_k lam_t^-k) Theta_{n+k}(v).
// Moments are stored pre-scaled per source cell, Mt_k = M_k lam_s^-k (FP32),
// and M_k lam_t^-k = Mt_k kappa^k with kappa = lam_s / lam_t. Every operand
// is then O(1) up to the MAC's (rho / r)^k decay; the recursion, products and
// the chunk's sums run in FP32. Chunks write their scaled F in FP32; the
// reduce adds a target's chunks in FP64 and applies lam_t^-(n+1).
template <int P, int H>
__device__ __forceinline__ void m2l_chunk32(const M2lChunk ck, uint32_t ci,
                                            const uint64_t* __restrict__ skeys, uint64_t smask,
                                            const double* __restrict__ ccx,
                                            const double* __restrict__ ccy,
                                            const double* __restrict__ ccz,
                                            const double* __restrict__ lam,
                                            const cplxf* __restrict__ m32, double soft2,
                                            cplxf* __restrict__ part) {
  constexpr int NF = hnpos(P);
  const double tx = ccx[ck.target], ty = ccy[ck.target], tz = ccz[ck.target];
  const double ilt = 1.0 / lam[ck.target];
  const float s2s = (float)(soft2 * ilt * ilt);
  cplxf F[NF];
#pragma unroll
  for (int i = 0; i < NF; i++) F[i] = cplxf{0.f, 0.f};
  for (uint32_t j = ck.begin; j < ck.end; j++) {
    const uint32_t src = (uint32_t)(skeys[j] & smask);
    const cplxf* ms = m32 + (size_t)src * NF;
    const float kap = (float)(lam[src] * ilt);
    cplxf M[NF];
    float kk = 1.f;
#pragma unroll
    for (int k = 0; k <= P; k++) {
#pragma unroll
      for (int l = 0; l <= k; l++) {
        const cplxf v = ms[hpos(k, l)];
        M[hpos(k, l)] = cplxf{v.re * kk, v.im * kk};
      }
      kk *= kap;
    }
    m2l_accumulate<P, H, float, cplxf>((float)((tx - ccx[src]) * ilt),
                                       (float)((ty - ccy[src]) * ilt),
                                       (float)((tz - ccz[src]) * ilt), s2s, M, F);
  }
  cplxf* out = part + (size_t)ci * NF;
#pragma unroll
  for (int n = 0; n <= P; n++) {
    if (H != 2 && (n & 1) != H) continue;
#pragma unroll
    for (int m = 0; m <= n; m++) out[hpos(n, m)] = F[hpos(n, m)];
  }
}

template <int P, bool SPLIT>
__global__ void __launch_bounds__(64) k_m2l32(const M2lChunk* __restrict__ chunks,
                                             uint32_t nchunks, const uint64_t* __restrict__ skeys,
                                             uint64_t smask, const double* __restrict__ ccx,
                                             const double* __restrict__ ccy,
                                             const double* __restrict__ ccz,
                                             const double* __restrict__ lam,
                                             const cplxf* __restrict__ m32, double soft2,
                                             cplxf* __restrict__ part) {
  const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
  if constexpr (SPLIT) {
    const uint32_t ci = (lb >> 1) * blockDim.x + threadIdx.x;
    if (ci >= nchunks) return;
    if (lb & 1)
      m2l_chunk32<P, 1>(chunks[ci], ci, skeys, smask, ccx, ccy, ccz, lam, m32, soft2, part);
    else
      m2l_chunk32<P, 0>(chunks[ci], ci, skeys, smask, ccx, ccy, ccz, lam, m32, soft2, part);
  } else {
    const uint32_t ci = lb * blockDim.x + threadIdx.x;
    if (ci >= nchunks) return;
    m2l_chunk32<P, 2>(chunks[ci], ci, skeys, smask, ccx, ccy, ccz, lam, m32, soft2, part);
  }
}

// Orders whose FP32 accumulators are split over two waves.
#ifndef FMMH_M2L32_SPLIT_FROM
#define FMMH_M2L32_SPLIT_FROM 9
#endif

template <int P>
void m2l32_t(const M2lChunk* chunks, uint32_t nchunks, const uint64_t* skeys, uint64_t smask,
             DevCells c, const double* lam, const cplxf* m32, double soft2, cplxf* part,
             hipStream_t s) {
  if (!nchunks) return;
  constexpr bool split = P >= FMMH_M2L32_SPLIT_FROM;
  const uint32_t nb = (nchunks + 63) / 64 * (split ? 2 : 1);
  hipLaunchKernelGGL((k_m2l32<P, split>), dim3(nb), dim3(64), 0, s, chunks, nchunks, skeys, smask,
                     c.cx, c.cy, c.cz, lam, m32, soft2, part);
}

// One translation unit per range of orders instantiates the launches.
#define FMMH_M2L_INSTANTIATE(P)                                                             \
  template void m2l_t<P>(const M2lChunk*, uint32_t, const uint64_t*, uint64_t, DevCells,    \
                         const cplx*, double, cplx*, hipStream_t);                          \
  template void m2l32_t<P>(const M2lChunk*, uint32_t, const uint64_t*, uint64_t, DevCells,  \
                           const double*, const cplxf*, double, cplxf*, hipStream_t);

}  // namespace fmmh
