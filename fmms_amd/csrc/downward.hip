// downward.hip — L2L and L2P on gfx950 (FP64, and the mixed form's FP32 on
// radius-scaled locals; FMA contraction on).
//
// L2L (translate.rs:175-206): F'_n^m(s') += sum_k sum_l conj(Y_k^l(s - s')) F_{n+k}^{m+l}(s)
// L2P (translate.rs:271-309): Psi_n^m(x) = sum_k sum_l conj(Y_k^l(s - x)) F_{n+k}^{m+l}(s),
//     pot += Re Psi_0^0, a -= (Re Psi_1^1, Im Psi_1^1, Re Psi_1^0).
// Upsilon is streamed degree by degree (harmonic.rs:141-173 recursion) and
// only the m >= 0 halves of the expansions are held (registers, P static).
#include "common.h"
#include "harm32.h"
#include "internal.h"

namespace fmmh {

__device__ __forceinline__ void cmac2(cplx& f, double ar, double ai, double br, double bi) {
  f.re = fma(ar, br, f.re);
  f.re = fma(-ai, bi, f.re);
  f.im = fma(ar, bi, f.im);
  f.im = fma(ai, br, f.im);
}

// Upsilon_k^q for q = 0..k from the two previous degrees (U1 = k-1, U2 = k-2).
template <int P>
__device__ __forceinline__ void upsilon_degree(int k, double x, double y, double z, double r2,
                                               cplx* U, const cplx* U1, const cplx* U2) {
  if (k == 0) {
    U[0] = cplx{1.0, 0.0};
    return;
  }
  const double a = (double)(2 * k - 1) * z;
#pragma unroll
  for (int q = 0; q <= P; q++) {
    if (q > k) continue;
    if (q == k) {  // (x+iy)/(2k) * Upsilon_{k-1}^{k-1}
      const double f = 1.0 / (double)(2 * k);
      const double wr = x * f, wi = y * f;
      U[q].re = wr * U1[k - 1].re - wi * U1[k - 1].im;
      U[q].im = wr * U1[k - 1].im + wi * U1[k - 1].re;
    } else {
      const double ic = 1.0 / (double)(k * k - q * q);
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

// conj(Y_k^l) from the m>=0 column U (Y_k^{-l} = (-1)^l conj(Y_k^l)).
__device__ __forceinline__ void conj_y(const cplx* U, int l, double* ar, double* ai) {
  if (l >= 0) {
    *ar = U[l].re;
    *ai = -U[l].im;
  } else {
    const double sg = ((-l) & 1) ? -1.0 : 1.0;
    *ar = sg * U[-l].re;
    *ai = sg * U[-l].im;
  }
}

// F_s^q from the m>=0 half (F_s^{-q} = (-1)^q conj(F_s^q)).
template <int P>
__device__ __forceinline__ void f_at(const cplx* F, int s, int q, double* br, double* bi) {
  if (q >= 0) {
    *br = F[hpos(s, q)].re;
    *bi = F[hpos(s, q)].im;
  } else {
    const double sg = ((-q) & 1) ? -1.0 : 1.0;
    *br = sg * F[hpos(s, -q)].re;
    *bi = -sg * F[hpos(s, -q)].im;
  }
}

// L2L, one thread per child cell of a level: locals[child] += shift(parent).
template <int P>
__global__ void __launch_bounds__(64) k_l2l(const uint32_t* __restrict__ cells, uint32_t cnt,
                                           const int32_t* __restrict__ parent,
                                           const double* __restrict__ cx,
                                           const double* __restrict__ cy,
                                           const double* __restrict__ cz,
                                           cplx* __restrict__ locals) {
  constexpr int D = hdim(P), NF = hnpos(P);
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const uint32_t c = cells[i];
  const int32_t pa = parent[c];
  if (pa < 0) return;
  const double x = cx[pa] - cx[c], y = cy[pa] - cy[c], z = cz[pa] - cz[c];
  const double r2 = x * x + y * y + z * z;
  cplx Fp[NF], G[NF];
  const cplx* lp = locals + (size_t)pa * D;
  cplx* lc = locals + (size_t)c * D;
#pragma unroll
  for (int n = 0; n <= P; n++)
#pragma unroll
    for (int m = 0; m <= n; m++) {
      Fp[hpos(n, m)] = lp[hidx(n, m)];
      G[hpos(n, m)] = lc[hidx(n, m)];
    }
  cplx U[P + 1], U1[P + 1], U2[P + 1];
#pragma unroll
  for (int k = 0; k <= P; k++) {
    upsilon_degree<P>(k, x, y, z, r2, U, U1, U2);
#pragma unroll
    for (int n = 0; n <= P; n++) {
      if (n + k > P) continue;
#pragma unroll
      for (int m = 0; m <= P; m++) {
        if (m > n) continue;
#pragma unroll
        for (int l = -P; l <= P; l++) {
          if (l < -k || l > k) continue;
          double ar, ai, br, bi;
          conj_y(U, l, &ar, &ai);
          f_at<P>(Fp, n + k, m + l, &br, &bi);
          cmac2(G[hpos(n, m)], ar, ai, br, bi);
        }
      }
    }
#pragma unroll
    for (int q = 0; q <= P; q++) {
      U2[q] = U1[q];
      U1[q] = U[q];
    }
  }
#pragma unroll
  for (int n = 0; n <= P; n++)
#pragma unroll
    for (int m = 0; m <= n; m++) {
      cplx g = G[hpos(n, m)];
      lc[hidx(n, m)] = g;
      if (m > 0) {
        const double sg = (m & 1) ? -1.0 : 1.0;
        lc[hidx(n, -m)] = cplx{sg * g.re, -sg * g.im};
      }
    }
}

// L2P, one thread per body (slot): far-field acceleration and potential of
// the local expansion (the softened trace terms are added by k_l2p_trace).
template <int P>
__global__ void __launch_bounds__(64) k_l2p(uint32_t n, const uint32_t* __restrict__ leaf_of_slot,
                                           const double* __restrict__ cx,
                                           const double* __restrict__ cy,
                                           const double* __restrict__ cz,
                                           const double* __restrict__ bx,
                                           const double* __restrict__ by,
                                           const double* __restrict__ bz,
                                           const cplx* __restrict__ locals,
                                           double4* __restrict__ far) {
  constexpr int D = hdim(P), NF = hnpos(P);
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t c = leaf_of_slot[k];
  const double x = cx[c] - bx[k], y = cy[c] - by[k], z = cz[c] - bz[k];
  const double r2 = x * x + y * y + z * z;
  const cplx* lf = locals + (size_t)c * D;
  cplx F[NF];
#pragma unroll
  for (int s = 0; s <= P; s++)
#pragma unroll
    for (int q = 0; q <= s; q++) F[hpos(s, q)] = lf[hidx(s, q)];
  double p00 = 0.0, p11r = 0.0, p11i = 0.0, p10r = 0.0;
  cplx U[P + 1], U1[P + 1], U2[P + 1];
#pragma unroll
  for (int kk = 0; kk <= P; kk++) {
    upsilon_degree<P>(kk, x, y, z, r2, U, U1, U2);
#pragma unroll
    for (int l = -P; l <= P; l++) {
      if (l < -kk || l > kk) continue;
      double ar, ai, br, bi;
      conj_y(U, l, &ar, &ai);
      f_at<P>(F, kk, l, &br, &bi);
      p00 = fma(ar, br, p00);
      p00 = fma(-ai, bi, p00);
      if (kk + 1 <= P) {
        f_at<P>(F, kk + 1, l + 1, &br, &bi);
        p11r = fma(ar, br, p11r);
        p11r = fma(-ai, bi, p11r);
        p11i = fma(ar, bi, p11i);
        p11i = fma(ai, br, p11i);
        f_at<P>(F, kk + 1, l, &br, &bi);
        p10r = fma(ar, br, p10r);
        p10r = fma(-ai, bi, p10r);
      }
    }
#pragma unroll
    for (int q = 0; q <= P; q++) {
      U2[q] = U1[q];
      U1[q] = U[q];
    }
  }
  far[k] = make_double4(-p11r, -p11i, -p10r, p00);
}

// ----------------------------------------------------------- FP32 bulk form
// Locals held per cell as Ft_n^m = F_n^m lam^n (lam: the cell's M2L scale,
// its radius floored at a quarter of its octant edge), m >= 0 half, FP32. A
// body at offset d from its leaf centre sees Upsilon_k(d) = lam^k
// Upsilon_k(d / lam), so with u = d / lam every operand is O(1):
//   L2L  Gt_n(child) = kappa^n sum_k sum_l conj(Y_k^l(u)) Ft_{n+k}^{m+l}(parent),
//        u = (s_parent - s_child) / lam_parent, kappa = lam_child / lam_parent <= 1;
//   L2P  Psi_n = lam^-n sum_k sum_l conj(Y_k^l(u)) Ft_{n+k}^{m+l}, u = (s - x) / lam.
// One thread per cell of a level (parents' levels done before): the cell's
// own FP64 locals (M2L and the softened source-side terms) scaled to Ft, plus
// its parent's Ft shifted to the cell. The own part enters divided by
// kappa^n, so the shifted sum needs no per-term factor; the result is
// multiplied by kappa^n once.
template <int P>
__global__ void __launch_bounds__(64) k_l2l32(const uint32_t* __restrict__ cells, uint32_t cnt,
                                             const int32_t* __restrict__ parent,
                                             const double* __restrict__ cx,
                                             const double* __restrict__ cy,
                                             const double* __restrict__ cz,
                                             const double* __restrict__ lam,
                                             const cplx* __restrict__ own,
                                             cplxf* __restrict__ loc32) {
  constexpr int D = hdim(P), NF = hnpos(P);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const uint32_t c = cells[i];
  const int32_t pa = parent[c];
  const double lc = lam[c];
  const cplx* lo = own + (size_t)c * D;
  cplxf* out = loc32 + (size_t)c * NF;
  if (pa < 0) {
    double sc = 1.0;
#pragma unroll
    for (int n = 0; n <= P; n++) {
#pragma unroll
      for (int m = 0; m <= n; m++) {
        const cplx v = lo[hidx(n, m)];
        out[hpos(n, m)] = cplxf{(float)(v.re * sc), (float)(v.im * sc)};
      }
      sc *= lc;
    }
    return;
  }
  const double lp = lam[pa], il = 1.0 / lp;
  const double kap = lc * il;
  cplxf G[NF], Fp[NF];
  {
    const double ik = 1.0 / kap;
    double sc = 1.0;  // lam_c^n kappa^-n = lam_p^n
#pragma unroll
    for (int n = 0; n <= P; n++) {
#pragma unroll
      for (int m = 0; m <= n; m++) {
        const cplx v = lo[hidx(n, m)];
        G[hpos(n, m)] = cplxf{(float)(v.re * sc), (float)(v.im * sc)};
      }
      sc *= lc * ik;
    }
  }
  const cplxf* lpp = loc32 + (size_t)pa * NF;
#pragma unroll
  for (int e = 0; e < NF; e++) Fp[e] = lpp[e];
  const float x = (float)((cx[pa] - cx[c]) * il), y = (float)((cy[pa] - cy[c]) * il),
              z = (float)((cz[pa] - cz[c]) * il);
  const float r2 = x * x + y * y + z * z;
  cplxf U[P + 1], U1[P + 1], U2[P + 1];
#pragma unroll
  for (int k = 0; k <= P; k++) {
    upsilon_degree32<P>(k, x, y, z, r2, U, U1, U2);
#pragma unroll
    for (int n = 0; n <= P; n++) {
      if (n + k > P) continue;
#pragma unroll
      for (int m = 0; m <= P; m++) {
        if (m > n) continue;
#pragma unroll
        for (int l = -P; l <= P; l++) {
          if (l < -k || l > k) continue;
          float ar, ai, br, bi;
          conj_y32(U, l, &ar, &ai);
          f_at32<P>(Fp, n + k, m + l, &br, &bi);
          cmac2f(G[hpos(n, m)], ar, ai, br, bi);
        }
      }
    }
#pragma unroll
    for (int q = 0; q <= P; q++) {
      U2[q] = U1[q];
      U1[q] = U[q];
    }
  }
  float kn = 1.0f;
  const float kf = (float)kap;
#pragma unroll
  for (int n = 0; n <= P; n++) {
#pragma unroll
    for (int m = 0; m <= n; m++) out[hpos(n, m)] = cplxf{G[hpos(n, m)].re * kn, G[hpos(n, m)].im * kn};
    kn *= kf;
  }
}

// L2P of the scaled locals, one thread per body; the offset from the leaf
// centre is formed in FP64 and scaled before the FP32 sums.
template <int P>
__global__ void __launch_bounds__(64) k_l2p32(uint32_t n, const uint32_t* __restrict__ leaf_of_slot,
                                             const double* __restrict__ cx,
                                             const double* __restrict__ cy,
                                             const double* __restrict__ cz,
                                             const double* __restrict__ bx,
                                             const double* __restrict__ by,
                                             const double* __restrict__ bz,
                                             const double* __restrict__ lam,
                                             const cplxf* __restrict__ loc32,
                                             double4* __restrict__ far) {
  constexpr int NF = hnpos(P);
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t c = leaf_of_slot[k];
  const double il = 1.0 / lam[c];
  const float x = (float)((cx[c] - bx[k]) * il), y = (float)((cy[c] - by[k]) * il),
              z = (float)((cz[c] - bz[k]) * il);
  const float r2 = x * x + y * y + z * z;
  const cplxf* lf = loc32 + (size_t)c * NF;
  cplxf F[NF];
#pragma unroll
  for (int e = 0; e < NF; e++) F[e] = lf[e];
  float p00 = 0.0f, p11r = 0.0f, p11i = 0.0f, p10r = 0.0f;
  cplxf U[P + 1], U1[P + 1], U2[P + 1];
#pragma unroll
  for (int kk = 0; kk <= P; kk++) {
    upsilon_degree32<P>(kk, x, y, z, r2, U, U1, U2);
#pragma unroll
    for (int l = -P; l <= P; l++) {
      if (l < -kk || l > kk) continue;
      float ar, ai, br, bi;
      conj_y32(U, l, &ar, &ai);
      f_at32<P>(F, kk, l, &br, &bi);
      p00 = fmaf(ar, br, p00);
      p00 = fmaf(-ai, bi, p00);
      if (kk + 1 <= P) {
        f_at32<P>(F, kk + 1, l + 1, &br, &bi);
        p11r = fmaf(ar, br, p11r);
        p11r = fmaf(-ai, bi, p11r);
        p11i = fmaf(ar, bi, p11i);
        p11i = fmaf(ai, br, p11i);
        f_at32<P>(F, kk + 1, l, &br, &bi);
        p10r = fmaf(ar, br, p10r);
        p10r = fmaf(-ai, bi, p10r);
      }
    }
#pragma unroll
    for (int q = 0; q <= P; q++) {
      U2[q] = U1[q];
      U1[q] = U[q];
    }
  }
  far[k] = make_double4(-(double)p11r * il, -(double)p11i * il, -(double)p10r * il, (double)p00);
}

// FP64 locals F_n^m = Ft_n^m lam^-n, both halves of m (the parity view of
// the mixed form's locals after L2L).
template <int P>
__global__ void __launch_bounds__(256) k_loc32_export(uint32_t nc, const double* __restrict__ lam,
                                                     const cplxf* __restrict__ loc32,
                                                     cplx* __restrict__ out) {
  constexpr int D = hdim(P), NF = hnpos(P);
  const uint32_t c = blockIdx.x * 256 + threadIdx.x;
  if (c >= nc) return;
  const double il = 1.0 / lam[c];
  double sc = 1.0;
  for (int n = 0; n <= P; n++) {
    for (int m = 0; m <= n; m++) {
      const cplxf v = loc32[(size_t)c * NF + hpos(n, m)];
      const cplx g = cplx{(double)v.re * sc, (double)v.im * sc};
      out[(size_t)c * D + hidx(n, m)] = g;
      if (m > 0) {
        const double sg = (m & 1) ? -1.0 : 1.0;
        out[(size_t)c * D + hidx(n, -m)] = cplx{sg * g.re, -sg * g.im};
      }
    }
    sc *= il;
  }
}

template <int P>
static void l2l_t(const uint32_t* cells, uint32_t cnt, DevCells c, cplx* locals, hipStream_t s) {
  hipLaunchKernelGGL(k_l2l<P>, dim3((cnt + 63) / 64), dim3(64), 0, s, cells, cnt, c.parent, c.cx,
                     c.cy, c.cz, locals);
}

template <int P>
static void l2p_t(uint32_t n, const uint32_t* leaf_of_slot, DevCells c, DevBodies b,
                  const cplx* locals, double4* far, hipStream_t s) {
  hipLaunchKernelGGL(k_l2p<P>, dim3((n + 63) / 64), dim3(64), 0, s, n, leaf_of_slot, c.cx, c.cy,
                     c.cz, b.x, b.y, b.z, locals, far);
}

#define FMMH_P_SWITCH2(p, CALL) \
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

bool launch_l2l32(int p, const uint32_t* cells, uint32_t cnt, DevCells c, const double* lam,
                  const cplx* own, void* loc32, hipStream_t s) {
  if (cnt == 0) return true;
#define CALL(P)                                                                                  \
  hipLaunchKernelGGL(k_l2l32<P>, dim3((cnt + 63) / 64), dim3(64), 0, s, cells, cnt, c.parent, \
                     c.cx, c.cy, c.cz, lam, own, (cplxf*)loc32)
  FMMH_P_SWITCH2(p, CALL)
#undef CALL
  return true;
}

bool launch_loc32_export(int p, uint32_t nc, const double* lam, const void* loc32, cplx* out,
                         hipStream_t s) {
  if (nc == 0) return true;
#define CALL(P)                                                                          \
  hipLaunchKernelGGL(k_loc32_export<P>, dim3((nc + 255) / 256), dim3(256), 0, s, nc, lam, \
                     (const cplxf*)loc32, out)
  FMMH_P_SWITCH2(p, CALL)
#undef CALL
  return true;
}

bool launch_l2l(int p, const uint32_t* cells, uint32_t cnt, DevCells c, cplx* locals,
                hipStream_t s) {
  if (cnt == 0) return true;
#define CALL(P) l2l_t<P>(cells, cnt, c, locals, s)
  FMMH_P_SWITCH2(p, CALL)
#undef CALL
  return true;
}

// Softened far field's trace polynomial of every body's leaf and ancestors,
// added to the body's L2P result (the same additions, in the same order, as
// inside k_l2p; a kernel of its own keeps the ancestor walk's dependent loads
// away from k_l2p's register-heavy harmonic sum, so many more waves hide them).
__global__ void __launch_bounds__(256) k_l2p_trace(uint32_t n,
                                                  const uint32_t* __restrict__ leaf_of_slot,
                                                  const double* __restrict__ cx,
                                                  const double* __restrict__ cy,
                                                  const double* __restrict__ cz,
                                                  const double* __restrict__ bx,
                                                  const double* __restrict__ by,
                                                  const double* __restrict__ bz,
                                                  const double* __restrict__ lap,
                                                  const int32_t* __restrict__ parent,
                                                  double4* __restrict__ far) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  double4 o = far[k];
  l2p_trace_walk(lap, parent, cx, cy, cz, leaf_of_slot[k], bx[k], by[k], bz[k], o);
  far[k] = o;
}

bool launch_l2p(int p, uint32_t n, const uint32_t* leaf_of_slot, DevCells c, DevBodies b,
                const cplx* locals, const double* lap, double4* far, hipStream_t s,
                const double* lam, const void* loc32) {
  if (n == 0) return true;
  if (loc32) {
#define CALL(P)                                                                                  \
  hipLaunchKernelGGL(k_l2p32<P>, dim3((n + 63) / 64), dim3(64), 0, s, n, leaf_of_slot, c.cx, c.cy, \
                     c.cz, b.x, b.y, b.z, lam, (const cplxf*)loc32, far)
    FMMH_P_SWITCH2(p, CALL)
#undef CALL
  } else {
#define CALL(P) l2p_t<P>(n, leaf_of_slot, c, b, locals, far, s)
    FMMH_P_SWITCH2(p, CALL)
#undef CALL
  }
  if (lap)
    hipLaunchKernelGGL(k_l2p_trace, dim3((n + 255) / 256), dim3(256), 0, s, n, leaf_of_slot, c.cx,
                       c.cy, c.cz, b.x, b.y, b.z, lap, c.parent, far);
  return true;
}

}  // namespace fmmh
