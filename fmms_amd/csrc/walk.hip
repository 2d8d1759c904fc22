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
// come out in one fixed order). Each round runs on fixed grids that read the
// frontier size from device memory, so the host checks the walk's progress
// only once per batch of rounds. The MAC (mac.rs:23-95) is evaluated in the
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
// or every system root of an ensemble); its size goes to the walk state.
__global__ void __launch_bounds__(256) k_front_init(const uint32_t* __restrict__ cells,
                                                   uint32_t cnt, uint2* __restrict__ front,
                                                   WalkState* __restrict__ st) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < cnt) front[i] = make_uint2(cells[i], cells[i]);
  if (i == 0) st->fsize[0] = cnt;
}

void launch_front_init(const uint32_t* cells, uint32_t cnt, void* front, WalkState* st,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_front_init, dim3((cnt + 255) / 256), dim3(256), 0, s, cells, cnt,
                     (uint2*)front, st);
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
template <bool ERR>
__device__ __forceinline__ bool accept(uint32_t a, const WalkRec& ra, uint32_t b,
                                       const WalkRec& rb, double r2, const double* __restrict__ mass,
                                       const MacParams& mp) {
  double sum_rho = ra.rho + rb.rho;
  if (r2 * mp.theta2 <= sum_rho * sum_rho) return false;
  if (mp.soft_mac > 0.0 && mp.soft2 > 0.0 &&
      mp.soft2 * (sum_rho * sum_rho * sum_rho) > mp.soft_mac * (r2 * r2 * sqrt(r2)))
    return false;
  if (!ERR) return true;
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

// walk.rs:94-190 decision of one frontier pair; nch = its child pairs. ERR:
// the error-controlled MAC (mp.use_err); the theta-only walk is compiled
// without it (no scratch for the error estimate's power table).
template <bool ERR>
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
  if (accept<ERR>(a, ra, b, rb, r2ab, mass, mp) && accept<ERR>(b, rb, a, ra, r2ba, mass, mp))
    return A_M2L;
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

// Sum of three counters over the block (CW_T threads); every thread gets it.
__device__ __forceinline__ void block_sum3(uint32_t v[3], uint32_t (*sh)[3]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int k = 0; k < 3; k++) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
  }
  if (lane == 0)
    for (int k = 0; k < 3; k++) sh[w][k] = v[k];
  __syncthreads();
  for (int k = 0; k < 3; k++) {
    uint32_t t = 0;
    for (int q = 0; q < CW_T / 64; q++) t += sh[q][k];
    v[k] = t;
  }
  __syncthreads();
}

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

// The round's frontier slice of logical block lb: WALK_G blocks, each a
// contiguous run of ceil(nf / WALK_G) pairs, in frontier order. Logical
// blocks are spread so that each XCD takes a contiguous range of slices
// (neighbouring pairs share cells, whose records then share one L2).
__device__ __forceinline__ uint32_t walk_block() { return xcd_block(blockIdx.x, gridDim.x); }
__device__ __forceinline__ void walk_slice(uint32_t nf, uint32_t lb, uint32_t* b0, uint32_t* b1) {
  const uint32_t chunk = (nf + WALK_G - 1) / WALK_G;
  *b0 = min(lb * chunk, nf);
  *b1 = min(*b0 + chunk, nf);
}

// One round of the breadth-first walk in three launches of fixed grids, with
// outputs in frontier order and no host round trip: (1) every pair of the
// round's frontier (its size read from the walk state) is decided, its
// decision stored and its block's output counts written; (2) one block scans
// the per-block counts and, when the round's outputs fit, advances the walk
// state (next frontier size, list lengths); otherwise it records the sizes
// needed and halts the walk at this round (every later kernel returns at
// once); (3) every pair writes its outputs (child pairs to the next
// frontier, M2L and P2P pairs appended to the lists) at its block's offset
// plus its rank in the block. The lists thus come out in one fixed order,
// and a system's pairs keep their relative order whether it is walked alone
// or in an ensemble.
template <bool ERR>
__global__ void __launch_bounds__(CW_T) k_walk_decide(const uint2* __restrict__ fin, uint32_t r,
                                                    const WalkState* __restrict__ st,
                                                    uint16_t* __restrict__ code,
                                                    uint32_t* __restrict__ blk,
                                                    const WalkRec* __restrict__ rec,
                                                    const double* __restrict__ mass, MacParams mp) {
  __shared__ uint32_t sh[CW_T / 64][3];
  if (st->status != 0u) return;
  const uint32_t lb = walk_block();
  uint32_t b0, b1;
  walk_slice(st->fsize[r & 1u], lb, &b0, &b1);
  uint32_t v[3] = {0u, 0u, 0u};  // child pairs, M2L, P2P
  for (uint32_t t = b0 + threadIdx.x; t < b1; t += CW_T) {
    const uint2 pr = fin[t];
    const WalkRec ra = rec[pr.x], rb = rec[pr.y];
    uint32_t nch;
    const int act = decide<ERR>(pr.x, pr.y, ra, rb, mass, mp, &nch);
    code[t] = (uint16_t)(act | (nch << 3));
    v[0] += nch;
    v[1] += act == A_M2L ? 1u : 0u;
    v[2] += act == A_P2P ? 1u : 0u;
  }
  block_sum3(v, sh);
  if (threadIdx.x < 3) blk[WALK_G * threadIdx.x + lb] = v[threadIdx.x];
}

// One block of 1024 threads: exclusive scans of the three count columns
// (WALK_G entries each, in place), then the walk state update.
__global__ void __launch_bounds__(1024) k_walk_scan(uint32_t* __restrict__ blk, uint32_t r,
                                                  WalkState* __restrict__ st, uint32_t cap_f,
                                                  uint32_t cap_m2l, uint32_t cap_p2p) {
  constexpr int PER = WALK_G / 1024;
  __shared__ uint32_t sh[16][3];
  if (st->status != 0u) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t v[3][PER], sum[3], x[3], wt[3], tot[3];
  for (int k = 0; k < 3; k++) {
    sum[k] = 0;
    for (int q = 0; q < PER; q++) {
      v[k][q] = blk[WALK_G * k + PER * threadIdx.x + q];
      sum[k] += v[k][q];
    }
    x[k] = wave_excl_scan(sum[k], &wt[k]);
  }
  if (lane == 0)
    for (int k = 0; k < 3; k++) sh[w][k] = wt[k];
  __syncthreads();
  for (int k = 0; k < 3; k++) {
    uint32_t off = 0, t = 0;
    for (int q = 0; q < 16; q++) {
      if (q < w) off += sh[q][k];
      t += sh[q][k];
    }
    tot[k] = t;
    uint32_t run = off + x[k];
    for (int q = 0; q < PER; q++) {
      blk[WALK_G * k + PER * threadIdx.x + q] = run;
      run += v[k][q];
    }
  }
  if (threadIdx.x != 0) return;
  const uint64_t need_m2l = (uint64_t)st->n_m2l + tot[1];
  const uint64_t need_p2p = (uint64_t)st->n_p2p + tot[2];
  uint32_t bad = 0u;
  if (tot[0] > cap_f) bad |= 1u;
  if (need_m2l > cap_m2l) bad |= 2u;
  if (need_p2p > cap_p2p) bad |= 4u;
  if (bad) {
    st->need_f = tot[0];
    st->need_m2l = (uint32_t)min(need_m2l, (uint64_t)0xFFFFFFFFu);
    st->need_p2p = (uint32_t)min(need_p2p, (uint64_t)0xFFFFFFFFu);
    st->status = bad;
    return;
  }
  if (st->fsize[r & 1u] > 0u) st->rounds++;
  st->fsize[(r + 1u) & 1u] = tot[0];
  st->m2l_base = st->n_m2l;
  st->p2p_base = st->n_p2p;
  st->n_m2l = (uint32_t)need_m2l;
  st->n_p2p = (uint32_t)need_p2p;
  st->round = r + 1u;
}

constexpr uint32_t WALK_STAGE = CW_T * 8;  // child pairs staged per batch

__global__ void __launch_bounds__(CW_T) k_walk_emit(const uint2* __restrict__ fin, uint32_t r,
                                                  const WalkState* __restrict__ st,
                                                  uint2* __restrict__ fout, uint2* __restrict__ m2l,
                                                  uint2* __restrict__ p2p,
                                                  const uint16_t* __restrict__ code,
                                                  const uint32_t* __restrict__ blk,
                                                  const WalkRec* __restrict__ rec) {
  __shared__ uint32_t sh[CW_T / 64][3];
  __shared__ uint2 stage[WALK_STAGE];
  if (st->status != 0u) return;
  const uint32_t lb = walk_block();
  uint32_t b0, b1;
  walk_slice(st->fsize[r & 1u], lb, &b0, &b1);
  // the scan checked that the whole round fits its buffers
  uint32_t fo = blk[lb];
  uint32_t mo = st->m2l_base + blk[WALK_G + lb];
  uint32_t po = st->p2p_base + blk[2 * WALK_G + lb];
  for (uint32_t t0 = b0; t0 < b1; t0 += CW_T) {
    const uint32_t t = t0 + threadIdx.x;
    const bool valid = t < b1;
    const uint32_t cd = valid ? (uint32_t)code[t] : 0u;
    const int act = valid ? (int)(cd & 7u) : A_NONE;
    const uint32_t v[3] = {cd >> 3, act == A_M2L ? 1u : 0u, act == A_P2P ? 1u : 0u};
    uint32_t x[3], tot[3];
    block_scan3(v, x, tot, sh);
    const uint2 pr = valid ? fin[t] : make_uint2(0u, 0u);
    if (act == A_M2L) m2l[mo + x[1]] = pr;
    if (act == A_P2P) p2p[po + x[2]] = pr;
    // child pairs: staged in LDS and written as one contiguous run when the
    // batch's fit (always, unless self pairs of cells with many children
    // crowd it), else written by each pair directly
    const bool staged = tot[0] <= WALK_STAGE;
    uint2* dst = staged ? stage : fout + fo;
    uint32_t wp = x[0];
    if (act == A_SELF_SPLIT) {
      const WalkRec ra = rec[pr.x];
      for (int i = 0; i < 8; i++) {
        if (ra.ch[i] < 0) continue;
        for (int j = i; j < 8; j++)
          if (ra.ch[j] >= 0) dst[wp++] = make_uint2((uint32_t)ra.ch[i], (uint32_t)ra.ch[j]);
      }
    } else if (act == A_SPLIT_B) {
      const WalkRec rb = rec[pr.y];
      for (int j = 0; j < 8; j++)
        if (rb.ch[j] >= 0) dst[wp++] = make_uint2(pr.x, (uint32_t)rb.ch[j]);
    } else if (act == A_SPLIT_A) {
      const WalkRec ra = rec[pr.x];
      for (int i = 0; i < 8; i++)
        if (ra.ch[i] >= 0) dst[wp++] = make_uint2((uint32_t)ra.ch[i], pr.y);
    }
    if (staged) {
      __syncthreads();
      for (uint32_t i = threadIdx.x; i < tot[0]; i += CW_T) fout[fo + i] = stage[i];
      __syncthreads();
    }
    fo += tot[0];
    mo += tot[1];
    po += tot[2];
  }
}

void launch_walk_round(const void* fin, void* fout, uint32_t r, WalkState* st, uint32_t cap_f,
                       uint2* m2l, uint32_t cap_m2l, uint2* p2p, uint32_t cap_p2p, uint16_t* code,
                       uint32_t* blk, const void* rec, const double* mass, MacParams mac,
                       hipStream_t s) {
  if (mac.use_err)
    hipLaunchKernelGGL(k_walk_decide<true>, dim3(WALK_G), dim3(CW_T), 0, s, (const uint2*)fin, r,
                       st, code, blk, (const WalkRec*)rec, mass, mac);
  else
    hipLaunchKernelGGL(k_walk_decide<false>, dim3(WALK_G), dim3(CW_T), 0, s, (const uint2*)fin, r,
                       st, code, blk, (const WalkRec*)rec, mass, mac);
  hipLaunchKernelGGL(k_walk_scan, dim3(1), dim3(1024), 0, s, blk, r, st, cap_f, cap_m2l, cap_p2p);
  hipLaunchKernelGGL(k_walk_emit, dim3(WALK_G), dim3(CW_T), 0, s, (const uint2*)fin, r, st,
                     (uint2*)fout, m2l, p2p, code, blk, (const WalkRec*)rec);
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
