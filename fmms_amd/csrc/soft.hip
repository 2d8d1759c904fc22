// soft.hip — trace terms of the Plummer-softened far field.
//
// The softened kernel K = (r^2 + eps^2)^{-1/2} is not harmonic, so its Taylor
// expansions about the source and the target centre carry, beside the
// solid-harmonic parts that M2L / L2P evaluate (Theta scaled by (r/s)^{2n+1},
// s^2 = r^2 + eps^2), trace terms |d|^{2k} Y_{n-2k}(d) Lap^k K
// (Lap K = -3 eps^2 s^{-5}). Kept:
//   * source side, the field (q/6) Lap K of every M2L source (q = sum m |x - z|^2
//     of the source cell), Taylor-expanded at the target centre to second
//     order: value into F_0^0, gradient into F_1^m, traceless Hessian part
//     into F_2^m, its trace into the target-side Laplacian;
//   * target side, the trace parts of the Taylor series of the cell's own M2L
//     far field to fourth order,
//       |d|^2 L/6 + |d|^2 (d.G)/10 + |d|^2 (d.H.d)/28 - |d|^4 tr(H)/280,
//     from L = Lap Phi, G = grad L, H = Hess L at the centre (10 doubles per
//     cell). A centre shift would turn part of this polynomial into harmonic
//     terms, so it is not moved through L2L: L2P sums it over the body's
//     leaf and every ancestor, each about its own centre.
// Oracle and device apply the same rule; eps = 0 skips all of it.
#include "common.h"
#include "internal.h"

namespace fmmh {

constexpr int ST = 256;

// q of each leaf from its bodies, one wave per leaf: the lanes form each
// body's |x - z|^2 (64 bodies at a time, coalesced) and stage them in LDS,
// and lane 0 sums them in slot order.
__global__ void __launch_bounds__(ST) k_trace_leaf(const uint32_t* __restrict__ leaf_ids,
                                                  uint32_t nleaves, DevCells c, DevBodies b,
                                                  double* __restrict__ q) {
  __shared__ double chunk[ST / 64][2][64];  // per wave: m and |x - z|^2 of 64 bodies
  const uint32_t wv = threadIdx.x >> 6;
  const uint32_t i = blockIdx.x * (ST / 64) + wv;
  const uint32_t lane = threadIdx.x & 63;
  if (i >= nleaves) return;
  const uint32_t id = leaf_ids[i];
  const double zx = c.cx[id], zy = c.cy[id], zz = c.cz[id];
  double acc = 0.0;
  const uint32_t f = c.first[id], n = c.count[id];
  for (uint32_t base = 0; base < n; base += 64) {
    const uint32_t k = f + base + lane;
    const bool have = base + lane < n;
    double mk = 0.0, d2 = 0.0;
    if (have) {
      const double dx = b.x[k] - zx, dy = b.y[k] - zy, dz = b.z[k] - zz;
      mk = b.m[k];
      d2 = dx * dx + dy * dy + dz * dz;
    }
    chunk[wv][0][lane] = mk;
    chunk[wv][1][lane] = d2;
    __builtin_amdgcn_wave_barrier();
    const uint32_t cnt = min(64u, n - base);
    if (lane == 0)
      for (uint32_t j = 0; j < cnt; j++) acc = fma(chunk[wv][0][j], chunk[wv][1][j], acc);
    __builtin_amdgcn_wave_barrier();
  }
  if (lane == 0) q[id] = acc;
}

// Parallel-axis upward pass: q_P = sum_c q_c + 2 D_c.(z_c - z_P) + M_c |z_c - z_P|^2,
// D_c = the child's dipole about its centre, from its multipole moments
// (M_1^0 = sum m dz, M_1^1 = sum m (dx + i dy) / 2).
__global__ void __launch_bounds__(ST) k_trace_up(const uint32_t* __restrict__ cells, uint32_t cnt,
                                                DevCells c, int D, const cplx* __restrict__ mom,
                                                int NF, const cplxf* __restrict__ m32,
                                                const double* __restrict__ lam, int m32_below,
                                                double* __restrict__ q) {
  const uint32_t i = blockIdx.x * ST + threadIdx.x;
  if (i >= cnt) return;
  const uint32_t id = cells[i];
  if (c.leaf[id]) return;
  const double zx = c.cx[id], zy = c.cy[id], zz = c.cz[id];
  double acc = 0.0;
  for (int o = 0; o < 8; o++) {
    const int ch = c.children[(size_t)id * 8 + o];
    if (ch < 0) continue;
    const double dx = c.cx[ch] - zx, dy = c.cy[ch] - zy, dz = c.cz[ch] - zz;
    double px, py, pz;
    // cells below level m32_below hold only scaled FP32 moments (the mixed
    // form's FP32 upward pass): M_1 = Mt_1 lam (m >= 0 layout: (1,0) at 1,
    // (1,1) at 2); the others hold FP64 moments
    if (m32 && (int)c.level[ch] > m32_below) {
      const cplxf* mc = m32 + (size_t)ch * NF;
      const double l = lam[ch];
      px = 2.0 * (double)mc[2].re * l;
      py = 2.0 * (double)mc[2].im * l;
      pz = (double)mc[1].re * l;
    } else {
      const cplx* mc = mom + (size_t)ch * D;
      px = 2.0 * mc[3].re;
      py = 2.0 * mc[3].im;
      pz = mc[2].re;
    }
    acc += q[ch] + 2.0 * (px * dx + py * dy + pz * dz) + c.mass[ch] * (dx * dx + dy * dy + dz * dz);
  }
  q[id] = acc;
}

// Per target (one wave, lanes striding over its sources): the source-side
// field of every M2L source into the local expansion (F_0^0, F_1^m, F_2^m)
// and the target-side Laplacian field (L, G, H) at the target centre.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

constexpr int NACC = 19;  // F00, F1 (3), F2 (5), L, G (3), H (6)

__global__ void __launch_bounds__(256) k_m2l_trace(uint32_t nc, const uint32_t* __restrict__ tptr,
                                                  const uint64_t* __restrict__ skeys, uint64_t smask,
                                                  const double4* __restrict__ cpm,
                                                  const double* __restrict__ q,
                                                  double soft2, int D, cplx* __restrict__ locals,
                                                  double* __restrict__ lap) {
  // neighbouring targets share sources: each XCD takes a contiguous range
  const uint32_t t = xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  if (t >= nc) return;
  const double4 tc = cpm[t];
  const double tx = tc.x, ty = tc.y, tz = tc.z;
  double acc[NACC];
#pragma unroll
  for (int k = 0; k < NACC; k++) acc[k] = 0.0;
  const uint32_t jend = tptr[t + 1];
  // the lane's sources j, j + 64, j + 128, ... in order; the next one's key
  // is loaded while this one is computed (prefetching its record too costs
  // a wave of occupancy)
  uint32_t jn = tptr[t] + lane;
  uint32_t src_n = jn < jend ? (uint32_t)(skeys[jn] & smask) : 0u;
  for (uint32_t j = jn; j < jend; j += 64) {
    const uint32_t src = src_n;
    const double4 sc = cpm[src];
    const double qs = q[src];
    jn = j + 64;
    if (jn < jend) src_n = (uint32_t)(skeys[jn] & smask);
    const double rx = tx - sc.x, ry = ty - sc.y, rz = tz - sc.z;
    const double r2 = rx * rx + ry * ry + rz * rz;
    const double s2 = r2 + soft2;
    const double is = rsqrt_pos(s2);  // s2 >= softening^2 > 0
    const double is2 = is * is;
    const double is5 = is2 * is2 * is, is7 = is5 * is2, is9 = is7 * is2, is11 = is9 * is2;
    const double lk = -3.0 * soft2 * is5;                              // Lap K
    const double gk = 15.0 * soft2 * is7;                              // grad Lap K / r
    const double hk = 105.0 * soft2 * is9;                             // Hess Lap K, r r^T part
    const double l4 = 15.0 * soft2 * (3.0 * is7 - 7.0 * r2 * is9);     // Lap^2 K
    const double g4 = 15.0 * soft2 * (63.0 * r2 * is11 - 35.0 * is9);  // grad Lap^2 K / r
    const double ms = sc.w;
    const double w = qs * (1.0 / 6.0);
    // source side: w Lap K, its gradient, the traceless part of its Hessian
    const double gw = w * gk, hw = 0.5 * w * hk;
    acc[0] += w * lk;
    acc[1] -= gw * rz;                                  // F_1^0
    acc[2] -= gw * rx;                                  // Re F_1^1
    acc[3] -= gw * ry;                                  // Im F_1^1
    acc[4] += 2.0 * hw * (r2 / 3.0 - rz * rz);          // F_2^0
    acc[5] += -2.0 * hw * rx * rz;                      // Re F_2^1
    acc[6] += -2.0 * hw * ry * rz;                      // Im F_2^1
    acc[7] += -2.0 * hw * (rx * rx - ry * ry);          // Re F_2^2
    acc[8] += -4.0 * hw * rx * ry;                      // Im F_2^2
    // target side: Laplacian field of the source (monopole + source-side field)
    acc[9] += ms * lk + w * l4;
    const double gg = ms * gk + w * g4;
    acc[10] += gg * rx;
    acc[11] += gg * ry;
    acc[12] += gg * rz;
    const double mh = ms * hk, mg = ms * gk;
    acc[13] += mg - mh * rx * rx;
    acc[14] += mg - mh * ry * ry;
    acc[15] += mg - mh * rz * rz;
    acc[16] -= mh * rx * ry;
    acc[17] -= mh * rx * rz;
    acc[18] -= mh * ry * rz;
  }
#pragma unroll
  for (int k = 0; k < NACC; k++) acc[k] = wave_sum(acc[k]);
  if (lane != 0) return;
  cplx* F = locals + (size_t)t * D;
  F[0].re += acc[0];
  if (D > 1) {
    F[2].re += acc[1];  // (1, 0)
    F[3].re += acc[2];  // (1, 1)
    F[3].im += acc[3];
    F[1].re -= acc[2];  // (1, -1) = -conj(F_1^1)
    F[1].im += acc[3];
  }
  if (D > 4) {
    F[6].re += acc[4];  // (2, 0)
    F[7].re += acc[5];  // (2, 1)
    F[7].im += acc[6];
    F[5].re -= acc[5];  // (2, -1) = -conj(F_2^1)
    F[5].im += acc[6];
    F[8].re += acc[7];  // (2, 2)
    F[8].im += acc[8];
    F[4].re += acc[7];  // (2, -2) = conj(F_2^2)
    F[4].im -= acc[8];
  }
  double* o = lap + (size_t)SOFT_LAP_WORDS * t;
#pragma unroll
  for (int k = 0; k < SOFT_LAP_WORDS; k++) o[k] = acc[9 + k];
}

void launch_trace_leaf(const uint32_t* leaf_ids, uint32_t nleaves, DevCells c, DevBodies b,
                       double* q, hipStream_t s) {
  if (nleaves == 0) return;
  hipLaunchKernelGGL(k_trace_leaf, dim3((nleaves + ST / 64 - 1) / (ST / 64)), dim3(ST), 0, s,
                     leaf_ids, nleaves, c, b, q);
}

void launch_trace_up(const uint32_t* cells, uint32_t cnt, DevCells c, int p, const cplx* mom,
                     double* q, hipStream_t s, const void* m32, const double* lam,
                     int m32_below) {
  if (cnt == 0) return;
  hipLaunchKernelGGL(k_trace_up, dim3((cnt + ST - 1) / ST), dim3(ST), 0, s, cells, cnt, c, hdim(p),
                     mom, hnpos(p), (const cplxf*)m32, lam, m32_below, q);
}

void launch_m2l_trace(uint32_t nc, const uint32_t* tptr, const uint64_t* skeys, int sbits,
                      const double4* cpm, const double* q, double soft2, int p, cplx* locals,
                      double* lap, hipStream_t s) {
  const uint64_t smask = (1ull << sbits) - 1ull;
  hipLaunchKernelGGL(k_m2l_trace, dim3((nc + 3) / 4), dim3(256), 0, s, nc, tptr, skeys, smask,
                     cpm, q, soft2, hdim(p), locals, lap);
}

}  // namespace fmmh
