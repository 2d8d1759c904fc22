// common.h — device helpers shared by the gfx950 kernels of the FMM
// evaluator: complex FP64 arithmetic in the reference's operation order,
// index helpers of the solid-harmonic layout, wave64 utilities.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FMMH_WAVE 64

struct cplx {
  double re, im;
};

// Component-wise complex ops in the order of crates/fmm harmonic.rs:22-74.
// Files that need bit-exact results are compiled with -ffp-contract=off.
// FP32 complex of the mixed form (radius-scaled expansions).
struct cplxf {
  float re, im;
};

__device__ __forceinline__ cplx cz() { return cplx{0.0, 0.0}; }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return cplx{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return cplx{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cplx cscale(cplx a, double s) { return cplx{a.re * s, a.im * s}; }
__device__ __forceinline__ cplx cconj(cplx a) { return cplx{a.re, -a.im}; }

// Flat index of (n, m), -n <= m <= n: n(n+1)+m (harmonic.rs:84); the order-n
// block occupies [n^2, n^2+2n] and dim(p) = (p+1)^2.
__host__ __device__ constexpr int hidx(int n, int m) { return n * (n + 1) + m; }
__host__ __device__ constexpr int hdim(int p) { return (p + 1) * (p + 1); }
// Index among the m >= 0 coefficients only: n(n+1)/2 + m; count (p+1)(p+2)/2.
__host__ __device__ constexpr int hpos(int n, int m) { return n * (n + 1) / 2 + m; }
__host__ __device__ constexpr int hnpos(int p) { return (p + 1) * (p + 2) / 2; }

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int lane = threadIdx.x & (FMMH_WAVE - 1);
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Wave-uniform value (forces scalar registers / scalar loads downstream).
__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Exclusive prefix sum of a per-lane u32 over the wave; *total gets the sum.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t* total) {
  const int lane = threadIdx.x & (FMMH_WAVE - 1);
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < FMMH_WAVE; d <<= 1) {
    uint32_t y = __shfl_up(x, d, FMMH_WAVE);
    if (lane >= d) x += y;
  }
  *total = __shfl(x, FMMH_WAVE - 1, FMMH_WAVE);
  return x - v;
}

// Lane k's value (k wave-uniform) in every lane.
__device__ __forceinline__ double lane_read(double v, uint32_t k) {
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, (int)k);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), (int)k);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Workgroups are dispatched to the 8 XCDs round-robin (block b runs on XCD
// b % 8). Remapping gives each XCD a contiguous range of logical blocks, so
// neighbouring work (which shares source data) shares one L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  const uint32_t x = b & 7u, i = b >> 3, per = nb >> 3, rem = nb & 7u;
  return x * per + min(x, rem) + i;
}

// 1/sqrt(x) for finite x > 0: the hardware estimate refined by one
// second-order Newton step, y (1 + e/2 + 3e^2/8) with e = 1 - x y^2.
__device__ __forceinline__ double rsqrt_pos(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}

struct BBox {
  double lo[3];
  double hi[3];
  double side;
  double inv_side;
};
