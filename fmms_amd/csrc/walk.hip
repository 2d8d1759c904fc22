// walk.hip — the mutual dual-tree walk on gfx950 (compiled -ffp-contract=off).
//
// crates/fmm walk.rs:94-190 decides, for a cell pair (a, b): self pair of a
// leaf -> P2P (a,a); self pair of an internal cell -> its child pairs
// (i <= j); both MAC accepts -> mutual M2L (a,b); two leaves -> P2P (a,b);
// otherwise split the leaf's partner, or the larger-rho cell (ties split a).
// The pair set this produces does not depend on traversal order, so the GPU
// walks it breadth-first, one thread per frontier pair per round: children
// are written to the next frontier and M2L / P2P pairs to their lists at
// offsets from scans of the per-block counts, in frontier order (so the lists
// come out in one fixed order). The MAC (mac.rs:23-95) is evaluated in the
// reference's operation order, so decisions agree bit for bit.

#include <algorithm>

#include "common.h"
#include "internal.h"

namespace fmmh {

// One 64-byte record per cell with everything a walk decision reads:
// centre, rho and the child ids (all -1 <=> leaf, since an internal cell
// always has a non-empty octant).
struct __align__(64) WalkRec {
  double cx, cy, cz, rho;
  int32_t ch[8];
};

__global__ void __launch_bounds__(256) k_pack_walk(uint32_t nc, DevCells c,
                                                  WalkRec* __restrict__ rec,
                                                  double4* __restrict__ cpm) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nc) return;
  WalkRec r;
  r.cx = c.cx[i];
  r.cy = c.cy[i];
  r.cz = c.cz[i];
  r.rho = c.rho[i];
#pragma unroll
  for (int o = 0; o < 8; o++) r.ch[o] = c.children[(size_t)i * 8 + o];
  rec[i] = r;
  cpm[i] = make_double4(r.cx, r.cy, r.cz, c.mass[i]);
}

size_t walk_rec_bytes() { return sizeof(WalkRec); }

// Initial walk frontier: the self pairs (c, c) of the given cells (the root,
// or every system root of an ensemble).
__global__ void __launch_bounds__(256) k_front_init(const uint32_t* __restrict__ cells,
                                                   uint32_t cnt, uint2* __restrict__ front) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < cnt) front[i] = make_uint2(cells[i], cells[i]);
}

void launch_front_init(const uint32_t* cells, uint32_t cnt, void* front, hipStream_t s) {
  hipLaunchKernelGGL(k_front_init, dim3((cnt + 255) / 256), dim3(256), 0, s, cells, cnt,
                     (uint2*)front);
}

void launch_pack_walk(uint32_t nc, DevCells c, void* rec, double4* cpm, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_walk, dim3((nc + 255) / 256), dim3(256), 0, s, nc, c, (WalkRec*)rec,
                     cpm);
}

// mac.rs:23-61 error estimate of cell a acting on cell b (r > 0 here).
__device__ __forceinline__ double error_estimate(uint32_t a, double rho_b, double mass_a,
                                                 double rho_a, double r, const MacParams& mp) {
  const int p = mp.p;
  double r2 = r * r;
  double rp = 1.0;
  for (int i = 0; i < p / 2; i++) rp *= r2;
  if (p % 2 == 1) rp *= r;
  double rpows[FMMH_MAX_P + 1];
  rpows[0] = 1.0;
  for (int j = 1; j <= p; j++) rpows[j] = rpows[j - 1] * rho_b;
  const double* pw = mp.pows + (size_t)a * (p + 1);
  double s = 0.0;
  for (int k = 0; k <= p; k++) s += mp.binom[k] * pw[k] * rpows[p - k];
  double e = s / (mass_a * rp);
  double denom = rho_a + rho_b;
  if (denom == 0.0) return 0.0;
  return 8.0 * fmax(rho_a, rho_b) / denom * e;
}

// mac.rs:64-95 for both directions of one pair. theta_max extension:
// r2*theta_max^2 <= (rhoA+rhoB)^2 rejects; theta_max = 1 multiplies by 1.
// Softening guard extension: eps^2 (rhoA+rhoB)^3 > tau r^5 rejects (the
// softened far field's leading residual relative to the pair's field).
__device__ __forceinline__ bool accept(uint32_t a, const WalkRec& ra, uint32_t b,
                                       const WalkRec& rb, double r2, const double* __restrict__ mass,
                                       const MacParams& mp) {
  double sum_rho = ra.rho + rb.rho;
  if (r2 * mp.theta2 <= sum_rho * sum_rho) return false;
  if (mp.soft_mac > 0.0 && mp.soft2 > 0.0 &&
      mp.soft2 * (sum_rho * sum_rho * sum_rho) > mp.soft_mac * (r2 * r2 * sqrt(r2)))
    return false;
  if (!mp.use_err) return true;
  double r = sqrt(r2);
  if (r == 0.0) return false;
  const double ma = mass[a];
  double est = error_estimate(a, rb.rho, ma, ra.rho, r, mp);
  if (est == 0.0) return true;
  return est * ma / r2 < mp.eps * mp.fmin[b];
}

enum : int { A_NONE = 0, A_P2P = 1, A_M2L = 2, A_SELF_SPLIT = 3, A_SPLIT_A = 4, A_SPLIT_B = 5 };

__device__ __forceinline__ uint32_t nchild(const WalkRec& r) {
  uint32_t k = 0;
#pragma unroll
  for (int o = 0; o < 8; o++) k += r.ch[o] >= 0 ? 1u : 0u;
  return k;
}

// walk.rs:94-190 decision of one frontier pair; nch = its child pairs.
__device__ __forceinline__ int decide(uint32_t a, uint32_t b, const WalkRec& ra,
                                      const WalkRec& rb, const double* __restrict__ mass,
                                      const MacParams& mp, uint32_t* nch) {
  const uint32_t na = nchild(ra), nb = nchild(rb);
  const bool la = na == 0, lb = nb == 0;
  *nch = 0;
  if (a == b) {
    if (la) return A_P2P;
    *nch = na * (na + 1) / 2;
    return A_SELF_SPLIT;
  }
  double dx = rb.cx - ra.cx, dy = rb.cy - ra.cy, dz = rb.cz - ra.cz;
  double r2ab = dx * dx + dy * dy + dz * dz;
  double ex = ra.cx - rb.cx, ey = ra.cy - rb.cy, ez = ra.cz - rb.cz;
  double r2ba = ex * ex + ey * ey + ez * ez;
  if (accept(a, ra, b, rb, r2ab, mass, mp) && accept(b, rb, a, ra, r2ba, mass, mp)) return A_M2L;
  if (la && lb) return A_P2P;
  bool split_b;
  if (la)
    split_b = true;
  else if (lb)
    split_b = false;
  else
    split_b = !(ra.rho >= rb.rho);
  *nch = split_b ? nb : na;
  return split_b ? A_SPLIT_B : A_SPLIT_A;
}

constexpr int CW_T = 256;

// Exclusive block scan of three counters (CW_T threads); totals in tot[3].
__device__ __forceinline__ void block_scan3(const uint32_t v[3], uint32_t x[3], uint32_t tot[3],
                                            uint32_t (*sh)[3]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t wt[3];
  for (int k = 0; k < 3; k++) x[k] = wave_excl_scan(v[k], &wt[k]);
  if (lane == 0)
    for (int k = 0; k < 3; k++) sh[w][k] = wt[k];
  __syncthreads();
  for (int k = 0; k < 3; k++) {
    uint32_t off = 0, t = 0;
    for (int q = 0; q < CW_T / 64; q++) {
      if (q < w) off += sh[q][k];
      t += sh[q][k];
    }
    x[k] += off;
    tot[k] = t;
  }
  __syncthreads();
}

// One round of the breadth-first walk in three launches, with outputs in
// frontier order: (1) every frontier pair is decided, its decision stored
// and its block's output counts written; (2) one block scans the per-block
// counts into offsets and round totals; (3) every pair writes its outputs
// (child pairs to the next frontier, M2L and P2P pairs appended to the
// lists) at its block's offset plus its rank in the block. The lists thus
// come out in one fixed order, and a system's pairs keep their relative
// order whether it is walked alone or in an ensemble.
__global__ void __launch_bounds__(CW_T) k_walk_decide(const uint2* __restrict__ fin, uint32_t nf,
                                                    uint16_t* __restrict__ code,
                                                    uint32_t* __restrict__ blk,
                                                    const WalkRec* __restrict__ rec,
                                                    const double* __restrict__ mass, MacParams mp) {
  __shared__ uint32_t sh[CW_T / 64][3];
  const uint32_t t = blockIdx.x * CW_T + threadIdx.x;
  uint32_t v[3] = {0u, 0u, 0u};  // child pairs, M2L, P2P
  if (t < nf) {
    const uint2 pr = fin[t];
    const WalkRec ra = rec[pr.x], rb = rec[pr.y];
    uint32_t nch;
    const int act = decide(pr.x, pr.y, ra, rb, mass, mp, &nch);
    code[t] = (uint16_t)(act | (nch << 3));
    v[0] = nch;
    v[1] = act == A_M2L ? 1u : 0u;
    v[2] = act == A_P2P ? 1u : 0u;
  }
  uint32_t x[3], tot[3];
  block_scan3(v, x, tot, sh);
  if (threadIdx.x < 3) blk[(size_t)gridDim.x * threadIdx.x + blockIdx.x] = tot[threadIdx.x];
}

// One block: exclusive scans of the three per-block count columns (nb each,
// in place), 4096 entries of each column per pass (four per thread); res
// gets the round totals added to the running list lengths.
__global__ void __launch_bounds__(1024) k_walk_scan(uint32_t* __restrict__ blk, uint32_t nb,
                                                  WalkResult* __restrict__ res) {
  __shared__ uint32_t sh[16][3];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t carry[3] = {0u, 0u, 0u};
  for (uint32_t b0 = 0; b0 < nb; b0 += 4096) {
    const uint32_t j0 = b0 + 4 * threadIdx.x;
    uint32_t v[3][4], sum[3];
    for (int k = 0; k < 3; k++) {
      const uint32_t* col = blk + (size_t)nb * k;
      sum[k] = 0;
      for (int q = 0; q < 4; q++) {
        v[k][q] = j0 + q < nb ? col[j0 + q] : 0u;
        sum[k] += v[k][q];
      }
    }
    uint32_t x[3], wt[3];
    for (int k = 0; k < 3; k++) x[k] = wave_excl_scan(sum[k], &wt[k]);
    if (lane == 0)
      for (int k = 0; k < 3; k++) sh[w][k] = wt[k];
    __syncthreads();
    for (int k = 0; k < 3; k++) {
      uint32_t off = 0, t = 0;
      for (int q = 0; q < 16; q++) {
        if (q < w) off += sh[q][k];
        t += sh[q][k];
      }
      uint32_t* col = blk + (size_t)nb * k;
      uint32_t run = carry[k] + off + x[k];
      for (int q = 0; q < 4; q++) {
        if (j0 + q < nb) col[j0 + q] = run;
        run += v[k][q];
      }
      carry[k] += t;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    res->next = carry[0];
    res->n_m2l += carry[1];
    res->n_p2p += carry[2];
  }
}

__global__ void __launch_bounds__(CW_T) k_walk_emit(const uint2* __restrict__ fin, uint32_t nf,
                                                  uint2* __restrict__ fout, uint32_t cap_f,
                                                  uint2* __restrict__ m2l, uint32_t cap_m2l,
                                                  uint2* __restrict__ p2p, uint32_t cap_p2p,
                                                  const uint16_t* __restrict__ code,
                                                  const uint32_t* __restrict__ blk,
                                                  uint32_t m2l0, uint32_t p2p0,
                                                  WalkResult* __restrict__ res,
                                                  const WalkRec* __restrict__ rec) {
  __shared__ uint32_t sh[CW_T / 64][3];
  const uint32_t t = blockIdx.x * CW_T + threadIdx.x;
  const bool valid = t < nf;
  const uint32_t cd = valid ? (uint32_t)code[t] : 0u;
  const int act = valid ? (int)(cd & 7u) : A_NONE;
  const uint32_t v[3] = {cd >> 3, act == A_M2L ? 1u : 0u, act == A_P2P ? 1u : 0u};
  uint32_t x[3], tot[3];
  block_scan3(v, x, tot, sh);
  if (!valid || act == A_NONE) return;
  const size_t nb = gridDim.x;
  const uint2 pr = fin[t];
  if (act == A_M2L) {
    const uint32_t wp = m2l0 + blk[nb + blockIdx.x] + x[1];
    if (wp < cap_m2l)
      m2l[wp] = pr;
    else
      atomicOr(&res->status, 2u);
    return;
  }
  if (act == A_P2P) {
    const uint32_t wp = p2p0 + blk[2 * nb + blockIdx.x] + x[2];
    if (wp < cap_p2p)
      p2p[wp] = pr;
    else
      atomicOr(&res->status, 4u);
    return;
  }
  uint32_t wp = blk[blockIdx.x] + x[0];
  if (wp + v[0] > cap_f) {
    atomicOr(&res->status, 1u);
    return;
  }
  if (act == A_SELF_SPLIT) {
    const WalkRec ra = rec[pr.x];
    for (int i = 0; i < 8; i++) {
      if (ra.ch[i] < 0) continue;
      for (int j = i; j < 8; j++)
        if (ra.ch[j] >= 0) fout[wp++] = make_uint2((uint32_t)ra.ch[i], (uint32_t)ra.ch[j]);
    }
  } else if (act == A_SPLIT_B) {
    const WalkRec rb = rec[pr.y];
    for (int j = 0; j < 8; j++)
      if (rb.ch[j] >= 0) fout[wp++] = make_uint2(pr.x, (uint32_t)rb.ch[j]);
  } else {
    const WalkRec ra = rec[pr.x];
    for (int i = 0; i < 8; i++)
      if (ra.ch[i] >= 0) fout[wp++] = make_uint2((uint32_t)ra.ch[i], pr.y);
  }
}

uint32_t walk_blocks(uint32_t nf) { return (nf + CW_T - 1) / CW_T; }

void launch_walk_decide(const uint2* fin, uint32_t nf, uint16_t* code, uint32_t* blk,
                        const void* rec, const double* mass, MacParams mac, WalkResult* res,
                        hipStream_t s) {
  const uint32_t nb = walk_blocks(nf);
  hipLaunchKernelGGL(k_walk_decide, dim3(nb), dim3(CW_T), 0, s, fin, nf, code, blk,
                     (const WalkRec*)rec, mass, mac);
  hipLaunchKernelGGL(k_walk_scan, dim3(1), dim3(1024), 0, s, blk, nb, res);
}

void launch_walk_emit(const uint2* fin, uint32_t nf, uint2* fout, uint32_t cap_f, uint2* m2l,
                      uint32_t cap_m2l, uint2* p2p, uint32_t cap_p2p, const uint16_t* code,
                      const uint32_t* blk, uint32_t m2l0, uint32_t p2p0, WalkResult* res,
                      const void* rec, hipStream_t s) {
  hipLaunchKernelGGL(k_walk_emit, dim3(walk_blocks(nf)), dim3(CW_T), 0, s, fin, nf, fout, cap_f,
                     m2l, cap_m2l, p2p, cap_p2p, code, blk, m2l0, p2p0, res, (const WalkRec*)rec);
}

// Row pointers of the target-sorted directed entries over all cells:
// ptr[t] = first entry of target t, ptr[ntarget] = entries with a target
// below ntarget. Entries past those (sentinel targets >= ntarget, sorted
// last) are not counted.
__global__ void __launch_bounds__(256) k_csr_ptr(const uint64_t* __restrict__ keys, uint32_t nnz,
                                                int sbits, uint32_t ntarget,
                                                uint32_t* __restrict__ ptr) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (nnz == 0) {
    for (uint32_t x = i; x <= ntarget; x += gridDim.x * 256) ptr[x] = 0;
    return;
  }
  if (i >= nnz) return;
  const int64_t t = std::min<int64_t>((int64_t)(keys[i] >> sbits), (int64_t)ntarget);
  const int64_t tp = i == 0 ? -1 : std::min<int64_t>((int64_t)(keys[i - 1] >> sbits), ntarget);
  for (int64_t x = tp + 1; x <= t; x++) ptr[x] = i;
  if (i == nnz - 1)
    for (int64_t x = t + 1; x <= (int64_t)ntarget; x++) ptr[x] = nnz;
}

void launch_csr_ptr(const uint64_t* keys, uint32_t nnz, int sbits, uint32_t ntarget, uint32_t* ptr,
                    hipStream_t s) {
  uint32_t threads = nnz ? nnz : ntarget + 1;
  hipLaunchKernelGGL(k_csr_ptr, dim3((threads + 255) / 256), dim3(256), 0, s, keys, nnz, sbits,
                     ntarget, ptr);
}

}  // namespace fmmh
