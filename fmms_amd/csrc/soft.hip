// soft.hip — trace terms of the Plummer-softened far field.
//
// The softened kernel K = (r^2 + eps^2)^{-1/2} is not harmonic, so its Taylor
// expansions about the source and the target centre carry, beside the
// solid-harmonic parts that M2L / L2P evaluate (Theta scaled by (r/s)^{2n+1},
// s^2 = r^2 + eps^2), trace terms |d|^{2k} Y_{n-2k}(d) Lap^k K. The leading
// ones are kept (Lap K = -3 eps^2 s^{-5}):
//   * source side, (q/6) Lap K, q = sum m |x - z|^2 of the source cell, added
//     to each M2L target's local expansion to first order (F_0^0 and F_1^m);
//   * target side, the second- and third-order trace parts of the Taylor
//     series about the target centre, (|d|^2 / 6) L + (|d|^2 / 10) d.G, with
//     L = Lap Phi and G = grad Lap Phi of the cell's (and its ancestors')
//     M2L sources' monopole fields at the centre (4 doubles per cell),
//     applied at L2P.
// Oracle and device apply the same rule; eps = 0 skips all of it.
#include "common.h"
#include "internal.h"

namespace fmmh {

constexpr int ST = 256;

// q of each leaf from its bodies (slot order).
__global__ void __launch_bounds__(ST) k_trace_leaf(const uint32_t* __restrict__ leaf_ids,
                                                  uint32_t nleaves, DevCells c, DevBodies b,
                                                  double* __restrict__ q) {
  const uint32_t i = blockIdx.x * ST + threadIdx.x;
  if (i >= nleaves) return;
  const uint32_t id = leaf_ids[i];
  const double zx = c.cx[id], zy = c.cy[id], zz = c.cz[id];
  double acc = 0.0;
  const uint32_t f = c.first[id], e = f + c.count[id];
  for (uint32_t k = f; k < e; k++) {
    const double dx = b.x[k] - zx, dy = b.y[k] - zy, dz = b.z[k] - zz;
    acc = fma(b.m[k], dx * dx + dy * dy + dz * dz, acc);
  }
  q[id] = acc;
}

// Parallel-axis upward pass: q_P = sum_c q_c + 2 D_c.(z_c - z_P) + M_c |z_c - z_P|^2,
// D_c = the child's dipole about its centre, from its multipole moments
// (M_1^0 = sum m dz, M_1^1 = sum m (dx + i dy) / 2).
__global__ void __launch_bounds__(ST) k_trace_up(const uint32_t* __restrict__ cells, uint32_t cnt,
                                                DevCells c, int D, const cplx* __restrict__ mom,
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
    const cplx* mc = mom + (size_t)ch * D;
    const double px = 2.0 * mc[3].re, py = 2.0 * mc[3].im, pz = mc[2].re;
    acc += q[ch] + 2.0 * (px * dx + py * dy + pz * dz) + c.mass[ch] * (dx * dx + dy * dy + dz * dz);
  }
  q[id] = acc;
}

// Per target (one wave, lanes striding over its sources): the source-side
// trace correction of every M2L source into the local expansion (F_0^0,
// F_1^0, F_1^{+-1}) and the target's Lap Phi.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ void __launch_bounds__(256) k_m2l_trace(uint32_t nc, const uint32_t* __restrict__ tptr,
                                                  const uint64_t* __restrict__ skeys, uint64_t smask,
                                                  DevCells c, const double* __restrict__ q,
                                                  double soft2, int D, cplx* __restrict__ locals,
                                                  double* __restrict__ lap) {
  const uint32_t t = blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  if (t >= nc) return;
  const double tx = c.cx[t], ty = c.cy[t], tz = c.cz[t];
  double f00 = 0.0, f10 = 0.0, f11r = 0.0, f11i = 0.0, L = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
  for (uint32_t j = tptr[t] + lane; j < tptr[t + 1]; j += 64) {
    const uint32_t src = (uint32_t)(skeys[j] & smask);
    const double rx = tx - c.cx[src], ry = ty - c.cy[src], rz = tz - c.cz[src];
    const double s2 = rx * rx + ry * ry + rz * rz + soft2;
    const double is = 1.0 / sqrt(s2);
    const double is2 = is * is;
    const double is5 = is2 * is2 * is;
    const double ms = c.mass[src];
    L -= 3.0 * soft2 * ms * is5;             // Lap of the source monopole's field
    const double gl = 15.0 * soft2 * ms * is5 / s2;  // its gradient / r
    gx += gl * rx;
    gy += gl * ry;
    gz += gl * rz;
    const double cc = 0.5 * q[src] * soft2 * is5;  // -(q/6) Lap K at the target centre
    const double g = 5.0 * cc / s2;                 // its gradient / r
    f00 -= cc;
    f10 -= g * rz;
    f11r -= g * rx;
    f11i -= g * ry;
  }
  f00 = wave_sum(f00);
  f10 = wave_sum(f10);
  f11r = wave_sum(f11r);
  f11i = wave_sum(f11i);
  L = wave_sum(L);
  gx = wave_sum(gx);
  gy = wave_sum(gy);
  gz = wave_sum(gz);
  if (lane != 0) return;
  cplx* F = locals + (size_t)t * D;
  F[0].re += f00;
  if (D > 1) {
    F[2].re += f10;           // (1, 0)
    F[3].re += f11r;          // (1, 1)
    F[3].im += f11i;
    F[1].re -= f11r;          // (1, -1) = -conj(F_1^1)
    F[1].im += f11i;
  }
  lap[4 * (size_t)t] = L;
  lap[4 * (size_t)t + 1] = gx;
  lap[4 * (size_t)t + 2] = gy;
  lap[4 * (size_t)t + 3] = gz;
}

// Each cell of a level adds its parent's linear Laplacian field shifted to
// its centre: L += L_p + G_p . (z - z_p), G += G_p (levels top-down).
__global__ void __launch_bounds__(ST) k_lap_down(const uint32_t* __restrict__ cells, uint32_t cnt,
                                                DevCells c, double* __restrict__ lap) {
  const uint32_t i = blockIdx.x * ST + threadIdx.x;
  if (i >= cnt) return;
  const uint32_t id = cells[i];
  const int32_t pa = c.parent[id];
  if (pa < 0) return;
  const double* lp = lap + 4 * (size_t)pa;
  double* lc = lap + 4 * (size_t)id;
  lc[0] += lp[0] + lp[1] * (c.cx[id] - c.cx[pa]) + lp[2] * (c.cy[id] - c.cy[pa]) +
           lp[3] * (c.cz[id] - c.cz[pa]);
  lc[1] += lp[1];
  lc[2] += lp[2];
  lc[3] += lp[3];
}

void launch_trace_leaf(const uint32_t* leaf_ids, uint32_t nleaves, DevCells c, DevBodies b,
                       double* q, hipStream_t s) {
  if (nleaves == 0) return;
  hipLaunchKernelGGL(k_trace_leaf, dim3((nleaves + ST - 1) / ST), dim3(ST), 0, s, leaf_ids,
                     nleaves, c, b, q);
}

void launch_trace_up(const uint32_t* cells, uint32_t cnt, DevCells c, int p, const cplx* mom,
                     double* q, hipStream_t s) {
  if (cnt == 0) return;
  hipLaunchKernelGGL(k_trace_up, dim3((cnt + ST - 1) / ST), dim3(ST), 0, s, cells, cnt, c, hdim(p),
                     mom, q);
}

void launch_m2l_trace(uint32_t nc, const uint32_t* tptr, const uint64_t* skeys, int sbits,
                      DevCells c, const double* q, double soft2, int p, cplx* locals, double* lap,
                      hipStream_t s) {
  const uint64_t smask = (1ull << sbits) - 1ull;
  hipLaunchKernelGGL(k_m2l_trace, dim3((nc + 3) / 4), dim3(256), 0, s, nc, tptr, skeys, smask, c,
                     q, soft2, hdim(p), locals, lap);
}

void launch_lap_down(const uint32_t* cells, uint32_t cnt, DevCells c, double* lap, hipStream_t s) {
  if (cnt == 0) return;
  hipLaunchKernelGGL(k_lap_down, dim3((cnt + ST - 1) / ST), dim3(ST), 0, s, cells, cnt, c, lap);
}

}  // namespace fmmh
