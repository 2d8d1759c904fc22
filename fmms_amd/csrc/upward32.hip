// upward32.hip — the mixed form's upward pass below the top tree levels: P2M
// and M2M in FP32 on radius-scaled moments Mt_n^m = M_n^m lam^-n (lam: the
// cell's M2L scale), m >= 0 half, written straight into the FP32 moments the
// M2L contraction reads. Offsets are formed in FP64 and scaled before the
// FP32 sums (P2M moments.rs:30-53, M2M moments.rs:56-90, rescaled):
//   P2M  Mt_n^m = sum_bodies m Y_n^m((x - z) / lam),
//   M2M  Mt_n^m(parent) = sum_children sum_k sum_l Y_k^l((z_c - z_p) / lam_p)
//                         kappa^(n-k) Mt_{n-k}^{m-l}(child),  kappa = lam_c / lam_p <= 1.
// The cells within FP64_TOP levels of a system root keep FP64 moments: their
// children's scaled moments are converted back (M = Mt lam^n) and the FP64
// M2M of upward.hip runs there. Compiled without multiply-add contraction
// (the sums use explicit fmaf), so the two P2M mappings below give the same
// bits and a system's moments do not depend on the size of its ensemble.
#include "harm32.h"
#include "internal.h"

namespace fmmh {

// P2M with one thread per leaf (leaves taken in body-count order, so a
// wave's lanes loop over nearly equal numbers of bodies): each body is read
// once; for trees with enough leaves to fill the device.
template <int P>
__global__ void __launch_bounds__(64) k_p2m32_leaf(const uint32_t* __restrict__ leaf_ids,
                                             uint32_t nleaves, const double* __restrict__ cx,
                                             const double* __restrict__ cy,
                                             const double* __restrict__ cz,
                                             const uint32_t* __restrict__ first,
                                             const uint32_t* __restrict__ count,
                                             const double4* __restrict__ xyzm,
                                             const double* __restrict__ lam,
                                             cplxf* __restrict__ m32) {
  constexpr int NF = hnpos(P);
  const uint32_t i = blockIdx.x * 64 + threadIdx.x;
  if (i >= nleaves) return;
  const uint32_t id = leaf_ids[i];
  const double zx = cx[id], zy = cy[id], zz = cz[id], il = 1.0 / lam[id];
  cplxf acc[NF];
#pragma unroll
  for (int e = 0; e < NF; e++) acc[e] = cplxf{0.0f, 0.0f};
  const uint32_t f = first[id], e = f + count[id];
  for (uint32_t k = f; k < e; k++) {
    const double4 q = xyzm[k];
    const float x = (float)((q.x - zx) * il), y = (float)((q.y - zy) * il),
                z = (float)((q.z - zz) * il);
    const float r2 = x * x + y * y + z * z;
    const float m = (float)q.w;
    cplxf U[P + 1], U1[P + 1], U2[P + 1];
#pragma unroll
    for (int deg = 0; deg <= P; deg++) {
      upsilon_degree32<P>(deg, x, y, z, r2, U, U1, U2);
#pragma unroll
      for (int q2 = 0; q2 <= P; q2++) {
        if (q2 > deg) continue;
        acc[hpos(deg, q2)].re = fmaf(m, U[q2].re, acc[hpos(deg, q2)].re);
        acc[hpos(deg, q2)].im = fmaf(m, U[q2].im, acc[hpos(deg, q2)].im);
      }
#pragma unroll
      for (int q2 = 0; q2 <= P; q2++) {
        U2[q2] = U1[q2];
        U1[q2] = U[q2];
      }
    }
  }
  cplxf* o = m32 + (size_t)id * NF;
#pragma unroll
  for (int e2 = 0; e2 < NF; e2++) o[e2] = acc[e2];
}

// P2M with one thread per (leaf, order m), for trees with too few leaves to
// fill the device with one thread per leaf: thread t takes column m = t / L of leaf
// t % L (L = leaves rounded up to whole waves), so a wave's lanes share m and
// run the same recursion, over leaves taken in body-count order (nearly
// equal trip counts). Column m: Y_m^m by the diagonal chain, then
// Y_n^m = ((2n-1) z Y_{n-1}^m - r^2 Y_{n-2}^m) / (n^2 - m^2) (harmonic.rs:141-173).
template <int P>
__global__ void __launch_bounds__(64) k_p2m32_col(const uint32_t* __restrict__ leaf_ids,
                                             uint32_t nleaves, uint32_t lpad,
                                             const double* __restrict__ cx,
                                             const double* __restrict__ cy,
                                             const double* __restrict__ cz,
                                             const uint32_t* __restrict__ first,
                                             const uint32_t* __restrict__ count,
                                             const double4* __restrict__ xyzm,
                                             const double* __restrict__ lam,
                                             cplxf* __restrict__ m32) {
  constexpr int NF = hnpos(P);
  const uint32_t t = blockIdx.x * 64 + threadIdx.x;
  const uint32_t li = t % lpad;
  const int m = (int)(t / lpad);
  if (li >= nleaves || m > P) return;
  const uint32_t id = leaf_ids[li];
  const double zx = cx[id], zy = cy[id], zz = cz[id], il = 1.0 / lam[id];
  cplxf acc[P + 1];
#pragma unroll
  for (int n = 0; n <= P; n++) acc[n] = cplxf{0.0f, 0.0f};
  const uint32_t f = first[id], e = f + count[id];
  for (uint32_t k = f; k < e; k++) {
    const double4 q = xyzm[k];
    const float x = (float)((q.x - zx) * il), y = (float)((q.y - zy) * il),
                z = (float)((q.z - zz) * il);
    const float r2 = x * x + y * y + z * z;
    const float ms = (float)q.w;
    cplxf u = cplxf{1.0f, 0.0f};  // Y_j^j up to j = m
    for (int j = 1; j <= m; j++) {
      const float g = 1.0f / (float)(2 * j);
      const float wr = x * g, wi = y * g;
      u = cplxf{wr * u.re - wi * u.im, wr * u.im + wi * u.re};
    }
    cplxf u1 = u, u2 = cplxf{0.0f, 0.0f};
#pragma unroll
    for (int n = 0; n <= P; n++) {
      if (n < m) continue;
      cplxf un;
      if (n == m) {
        un = u;
      } else {
        const float a = (float)(2 * n - 1) * z;
        const float ic = 1.0f / (float)(n * n - m * m);
        un = cplxf{(a * u1.re - r2 * u2.re) * ic, (a * u1.im - r2 * u2.im) * ic};
        u2 = u1;
        u1 = un;
      }
      acc[n].re = fmaf(ms, un.re, acc[n].re);
      acc[n].im = fmaf(ms, un.im, acc[n].im);
    }
  }
  cplxf* o = m32 + (size_t)id * NF;
#pragma unroll
  for (int n = 0; n <= P; n++)
    if (n >= m) o[hpos(n, m)] = acc[n];
}

// One lane per (internal cell of a level, octant): eight lanes form one
// parent; each computes its child's shifted moments, and the eight are added
// by a fixed xor-shuffle tree (children's levels done before).
template <int P>
__global__ void __launch_bounds__(64) k_m2m32(const uint32_t* __restrict__ cells, uint32_t cnt,
                                             const int32_t* __restrict__ children,
                                             const double* __restrict__ cx,
                                             const double* __restrict__ cy,
                                             const double* __restrict__ cz,
                                             const double* __restrict__ lam,
                                             cplxf* __restrict__ m32) {
  constexpr int NF = hnpos(P);
  const uint32_t g = blockIdx.x * 64 + threadIdx.x;
  const uint32_t pi = g >> 3, o = g & 7u;
  const bool valid = pi < cnt;
  const uint32_t id = valid ? cells[pi] : 0u;
  const int ch = valid ? children[(size_t)id * 8 + o] : -1;
  cplxf acc[NF];
#pragma unroll
  for (int e = 0; e < NF; e++) acc[e] = cplxf{0.0f, 0.0f};
  if (ch >= 0) {
    const double il = 1.0 / lam[id];
    const float x = (float)((cx[ch] - cx[id]) * il), y = (float)((cy[ch] - cy[id]) * il),
                z = (float)((cz[ch] - cz[id]) * il);
    const float r2 = x * x + y * y + z * z;
    const float kap = (float)(lam[ch] * il);
    cplxf M[NF];
    const cplxf* mc = m32 + (size_t)ch * NF;
    float kj = 1.0f;
#pragma unroll
    for (int j = 0; j <= P; j++) {
#pragma unroll
      for (int q = 0; q <= j; q++) {
        const cplxf v = mc[hpos(j, q)];
        M[hpos(j, q)] = cplxf{v.re * kj, v.im * kj};
      }
      kj *= kap;
    }
    cplxf U[P + 1], U1[P + 1], U2[P + 1];
#pragma unroll
    for (int k = 0; k <= P; k++) {
      upsilon_degree32<P>(k, x, y, z, r2, U, U1, U2);
#pragma unroll
      for (int n = 0; n <= P; n++) {
        if (n < k) continue;
#pragma unroll
        for (int m = 0; m <= P; m++) {
          if (m > n) continue;
#pragma unroll
          for (int l = -P; l <= P; l++) {
            if (l < -k || l > k) continue;
            if (m - l > n - k || l - m > n - k) continue;  // |m - l| <= n - k
            float ar, ai, br, bi;
            y32(U, l, &ar, &ai);
            f_at32<P>(M, n - k, m - l, &br, &bi);
            cmac2f(acc[hpos(n, m)], ar, ai, br, bi);
          }
        }
      }
#pragma unroll
      for (int q = 0; q <= P; q++) {
        U2[q] = U1[q];
        U1[q] = U[q];
      }
    }
  }
#pragma unroll
  for (int sft = 1; sft < 8; sft <<= 1)
#pragma unroll
    for (int e = 0; e < NF; e++) {
      acc[e].re += __shfl_xor(acc[e].re, sft, 64);
      acc[e].im += __shfl_xor(acc[e].im, sft, 64);
    }
  if (valid && o == 0) {
    cplxf* out = m32 + (size_t)id * NF;
#pragma unroll
    for (int e = 0; e < NF; e++) out[e] = acc[e];
  }
}

// FP64 moments M = Mt lam^n (both halves of m) of the cells that sel picks:
// sel >= 0, the cells at level sel and the leaves above it (the children of
// the FP64 top levels); sel < 0, every cell (the parity view).
template <int P>
__global__ void __launch_bounds__(256) k_m32_to_f64(uint32_t nc, const uint8_t* __restrict__ level,
                                                   const uint8_t* __restrict__ leaf, int sel,
                                                   const double* __restrict__ lam,
                                                   const cplxf* __restrict__ m32,
                                                   cplx* __restrict__ out) {
  constexpr int D = hdim(P), NF = hnpos(P);
  const uint32_t c = blockIdx.x * 256 + threadIdx.x;
  if (c >= nc) return;
  if (sel >= 0 && !((int)level[c] == sel || (leaf[c] && (int)level[c] < sel))) return;
  const double l = lam[c];
  double sc = 1.0;
  for (int n = 0; n <= P; n++) {
    for (int m = 0; m <= n; m++) {
      const cplxf v = m32[(size_t)c * NF + hpos(n, m)];
      const cplx g = cplx{(double)v.re * sc, (double)v.im * sc};
      out[(size_t)c * D + hidx(n, m)] = g;
      if (m > 0) {
        const double sg = (m & 1) ? -1.0 : 1.0;
        out[(size_t)c * D + hidx(n, -m)] = cplx{sg * g.re, -sg * g.im};
      }
    }
    sc *= l;
  }
}

#define FMMH_P_SWITCH4(p, CALL) \
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
    default: return false;      \
  }

bool launch_p2m32(int p, const uint32_t* leaf_ids, uint32_t nleaves, DevCells c, DevBodies b,
                  const double* lam, void* m32, hipStream_t s) {
  if (nleaves == 0) return true;
  // one thread per leaf once there are enough leaves for 1024 waves
  const uint32_t lpad = (nleaves + 63) / 64 * 64;
  const bool per_leaf = nleaves >= 65536;
#define CALL(P)                                                                                    \
  if (per_leaf)                                                                                    \
    hipLaunchKernelGGL(k_p2m32_leaf<P>, dim3((nleaves + 63) / 64), dim3(64), 0, s, leaf_ids,       \
                       nleaves, c.cx, c.cy, c.cz, c.first, c.count, b.xyzm, lam, (cplxf*)m32);     \
  else                                                                                             \
    hipLaunchKernelGGL(k_p2m32_col<P>, dim3(lpad / 64 * (P + 1)), dim3(64), 0, s, leaf_ids,        \
                       nleaves, lpad, c.cx, c.cy, c.cz, c.first, c.count, b.xyzm, lam, (cplxf*)m32)
  FMMH_P_SWITCH4(p, CALL)
#undef CALL
  return true;
}

bool launch_m2m32(int p, const uint32_t* cells, uint32_t cnt, DevCells c, const double* lam,
                  void* m32, hipStream_t s) {
  if (cnt == 0) return true;
#define CALL(P)                                                                                 \
  hipLaunchKernelGGL(k_m2m32<P>, dim3((cnt + 7) / 8), dim3(64), 0, s, cells, cnt, c.children,    \
                     c.cx, c.cy, c.cz, lam, (cplxf*)m32)
  FMMH_P_SWITCH4(p, CALL)
#undef CALL
  return true;
}

bool launch_m32_to_f64(int p, uint32_t nc, DevCells c, int sel, const double* lam,
                       const void* m32, cplx* out, hipStream_t s) {
  if (nc == 0) return true;
#define CALL(P)                                                                                 \
  hipLaunchKernelGGL(k_m32_to_f64<P>, dim3((nc + 255) / 256), dim3(256), 0, s, nc, c.level, \
                     c.leaf, sel, lam, (const cplxf*)m32, out)
  FMMH_P_SWITCH4(p, CALL)
#undef CALL
  return true;
}

}  // namespace fmmh
