// m2l.hip — far-field M2L on gfx950 (FP64, FMA contraction on).
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
#include "common.h"
#include "internal.h"

namespace fmmh {

constexpr int M2L_CH = 16;

__device__ __forceinline__ void cmac(cplx& f, double ar, double ai, double br, double bi) {
  f.re = fma(ar, br, f.re);
  f.re = fma(-ai, bi, f.re);
  f.im = fma(ar, bi, f.im);
  f.im = fma(ai, br, f.im);
}

// F (m>=0 half, hpos layout) += M2L of source multipole M (m>=0 half) at r.
template <int P>
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

template <int P>
__global__ void __launch_bounds__(64) k_m2l(const M2lChunk* __restrict__ chunks, uint32_t nchunks,
                                           const uint64_t* __restrict__ skeys, uint64_t smask,
                                           const double* __restrict__ ccx,
                                           const double* __restrict__ ccy,
                                           const double* __restrict__ ccz,
                                           const cplx* __restrict__ moments, double soft2,
                                           cplx* __restrict__ part) {
  constexpr int D = hdim(P), NF = hnpos(P);
  const uint32_t ci = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (ci >= nchunks) return;
  const M2lChunk ck = chunks[ci];
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
    m2l_accumulate<P>(tx - ccx[src], ty - ccy[src], tz - ccz[src], soft2, M, F);
  }
  cplx* out = part + (size_t)ci * NF;
#pragma unroll
  for (int i = 0; i < NF; i++) out[i] = F[i];
}

// Chunks: per target ceil(S_t / M2L_CH) consecutive source ranges.
__global__ void __launch_bounds__(256) k_m2l_nchunks(const uint32_t* __restrict__ tptr, uint32_t nc,
                                                    uint32_t* __restrict__ nch) {
  uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nc) return;
  uint32_t s = tptr[t + 1] - tptr[t];
  nch[t] = (s + M2L_CH - 1) / M2L_CH;
}

__global__ void __launch_bounds__(256) k_m2l_emit(const uint32_t* __restrict__ tptr, uint32_t nc,
                                                 const uint32_t* __restrict__ choff,
                                                 M2lChunk* __restrict__ chunks) {
  uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nc) return;
  uint32_t b = tptr[t], e = tptr[t + 1];
  uint32_t w = choff[t];
  for (uint32_t j = b; j < e; j += M2L_CH) chunks[w++] = M2lChunk{t, j, min(e, j + M2L_CH)};
}

// locals[t] = sum over the target's chunk partials in chunk order; both
// halves of m written (F_n^{-m} = (-1)^m conj(F_n^m)); empty targets get 0.
template <int P>
__global__ void __launch_bounds__(256) k_m2l_reduce(uint32_t nc, const uint32_t* __restrict__ choff,
                                                   const cplx* __restrict__ part,
                                                   cplx* __restrict__ locals) {
  constexpr int D = hdim(P), NF = hnpos(P);
  size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= (size_t)nc * NF) return;
  uint32_t t = (uint32_t)(g / NF);
  int e = (int)(g % NF);
  int n = 0;
  while ((n + 1) * (n + 2) / 2 <= e) n++;
  int m = e - n * (n + 1) / 2;
  cplx s = cplx{0.0, 0.0};
  for (uint32_t c = choff[t]; c < choff[t + 1]; c++) {
    cplx v = part[(size_t)c * NF + e];
    s.re += v.re;
    s.im += v.im;
  }
  cplx* out = locals + (size_t)t * D;
  out[hidx(n, m)] = s;
  if (m > 0) {
    double sg = (m & 1) ? -1.0 : 1.0;
    out[hidx(n, -m)] = cplx{sg * s.re, -sg * s.im};
  }
}

// Scratch layout handed in by the context: nch/choff (nc+1 words each) and
// chunk/partial arrays sized for the worst case sum_t ceil(S_t/CH).
size_t m2l_max_chunks(uint32_t nnz, uint32_t nc) { return (size_t)nnz / M2L_CH + nc + 1; }
size_t m2l_part_bytes(int p, size_t maxchunks) { return maxchunks * hnpos(p) * sizeof(cplx); }

template <int P>
static void m2l_t(const M2lChunk* chunks, uint32_t nchunks, const uint64_t* skeys, uint64_t smask,
                  DevCells c, const cplx* moments, double soft2, cplx* part, hipStream_t s) {
  if (nchunks)
    hipLaunchKernelGGL(k_m2l<P>, dim3((nchunks + 63) / 64), dim3(64), 0, s, chunks, nchunks, skeys,
                       smask, c.cx, c.cy, c.cz, moments, soft2, part);
}

template <int P>
static void m2l_reduce_t(uint32_t nc, const uint32_t* choff, const cplx* part, cplx* locals,
                         hipStream_t s) {
  size_t th = (size_t)nc * hnpos(P);
  hipLaunchKernelGGL(k_m2l_reduce<P>, dim3((th + 255) / 256), dim3(256), 0, s, nc, choff, part,
                     locals);
}

void m2l_build_chunks(const uint32_t* tptr, uint32_t nc, uint32_t* nch, uint32_t* choff,
                      uint32_t* scan_tmp, void* chunks, hipStream_t s, uint32_t** total_dev) {
  hipLaunchKernelGGL(k_m2l_nchunks, dim3((nc + 255) / 256), dim3(256), 0, s, tptr, nc, nch);
  scan_exclusive_u32(nch, choff, (size_t)nc, scan_tmp, s, total_dev);
  hipLaunchKernelGGL(k_m2l_emit, dim3((nc + 255) / 256), dim3(256), 0, s, tptr, nc, choff,
                     (M2lChunk*)chunks);
}

#define FMMH_P_SWITCH3(p, CALL) \
  switch (p) {                  \
    case 1: CALL(1); break;     \
    case 2: CALL(2); break;     \
    case 3: CALL(3); break;     \
    case 4: CALL(4); break;     \
    case 5: CALL(5); break;     \
    case 6: CALL(6); break;     \
    case 7: CALL(7); break;     \
    case 8: CALL(8); break;     \
    case 9: CALL(9); break;     \
    case 10: CALL(10); break;   \
    case 11: CALL(11); break;   \
    case 12: CALL(12); break;   \
    default: return false;      \
  }

bool launch_m2l_kernel(int p, const void* chunks, uint32_t nchunks, const uint64_t* skeys,
                       int sbits, DevCells c, const cplx* moments, double soft2, cplx* part,
                       hipStream_t s) {
  const uint64_t smask = (1ull << sbits) - 1ull;
  const M2lChunk* ck = (const M2lChunk*)chunks;
#define CALL(P) m2l_t<P>(ck, nchunks, skeys, smask, c, moments, soft2, part, s)
  FMMH_P_SWITCH3(p, CALL)
#undef CALL
  return true;
}

bool launch_m2l_reduce(int p, uint32_t nc, const uint32_t* choff, const cplx* part, cplx* locals,
                       hipStream_t s) {
#define CALL(P) m2l_reduce_t<P>(nc, choff, part, locals, s)
  FMMH_P_SWITCH3(p, CALL)
#undef CALL
  return true;
}

size_t m2l_chunk_bytes() { return sizeof(M2lChunk); }

}  // namespace fmmh
