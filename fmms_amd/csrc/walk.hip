// walk.hip — the mutual dual-tree walk on gfx950 (compiled -ffp-contract=off).
//
// crates/fmm walk.rs:94-190 decides, for a cell pair (a, b): self pair of a
// leaf -> P2P (a,a); self pair of an internal cell -> its child pairs
// (i <= j); both MAC accepts -> mutual M2L (a,b); two leaves -> P2P (a,b);
// otherwise split the leaf's partner, or the larger-rho cell (ties split a).
// The pair set this produces does not depend on traversal order, so the GPU
// walks it breadth-first: one thread per frontier pair per round, children
// appended to the next frontier, M2L / P2P pairs appended to their lists with
// wave-aggregated atomics. The MAC (mac.rs:23-95) is evaluated in the
// reference's operation order, so decisions agree bit for bit.
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

constexpr int WALK_T = 1024;

__global__ void __launch_bounds__(WALK_T) k_walk_round(const uint2* __restrict__ fin, uint32_t nf,
                                                   uint2* __restrict__ fout, uint32_t cap_f,
                                                   uint2* __restrict__ m2l, uint32_t cap_m2l,
                                                   uint2* __restrict__ p2p, uint32_t cap_p2p,
                                                   WalkCounters* __restrict__ ctr,
                                                   const WalkRec* __restrict__ rec,
                                                   const double* __restrict__ mass, MacParams mp) {
  const int lane = threadIdx.x & 63;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = t < nf;
  uint32_t a = 0, b = 0;
  int act = A_NONE;
  uint32_t nch = 0;
  WalkRec ra, rb;
  if (valid) {
    uint2 pr = fin[t];
    a = pr.x;
    b = pr.y;
    ra = rec[a];
    rb = rec[b];
    const uint32_t na = nchild(ra), nb = nchild(rb);
    const bool la = na == 0, lb = nb == 0;
    if (a == b) {
      if (la) {
        act = A_P2P;
      } else {
        act = A_SELF_SPLIT;
        nch = na * (na + 1) / 2;
      }
    } else {
      // walk.rs:138-139: accept(a,b) && accept(b,a); mac.rs:75-78 forms
      // r = cen_b - cen_a for the first call and cen_a - cen_b for the second.
      double dx = rb.cx - ra.cx, dy = rb.cy - ra.cy, dz = rb.cz - ra.cz;
      double r2ab = dx * dx + dy * dy + dz * dz;
      double ex = ra.cx - rb.cx, ey = ra.cy - rb.cy, ez = ra.cz - rb.cz;
      double r2ba = ex * ex + ey * ey + ez * ez;
      if (accept(a, ra, b, rb, r2ab, mass, mp) && accept(b, rb, a, ra, r2ba, mass, mp)) {
        act = A_M2L;
      } else if (la && lb) {
        act = A_P2P;
      } else {
        bool split_b;
        if (la)
          split_b = true;
        else if (lb)
          split_b = false;
        else
          split_b = !(ra.rho >= rb.rho);
        act = split_b ? A_SPLIT_B : A_SPLIT_A;
        nch = split_b ? nb : na;
      }
    }
  }
  // --- appends, aggregated per block: one atomic per counter per block
  __shared__ uint32_t wsum[WALK_T / 64][3];
  __shared__ uint32_t wbase[WALK_T / 64][3];
  const int w = threadIdx.x >> 6;
  const uint64_t lt = lanemask_lt();
  const uint64_t bm = __ballot(act == A_M2L);
  const uint64_t bp = __ballot(act == A_P2P);
  uint32_t wtot;
  const uint32_t pre = wave_excl_scan(nch, &wtot);
  if (lane == 0) {
    wsum[w][0] = wtot;
    wsum[w][1] = (uint32_t)__popcll(bm);
    wsum[w][2] = (uint32_t)__popcll(bp);
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int k = threadIdx.x;
    uint32_t tot = 0;
    for (int q = 0; q < WALK_T / 64; q++) {
      wbase[q][k] = tot;
      tot += wsum[q][k];
    }
    uint32_t* ctrk = k == 0 ? &ctr->next : (k == 1 ? &ctr->m2l : &ctr->p2p);
    const uint32_t b0 = tot ? atomicAdd(ctrk, tot) : 0u;
    for (int q = 0; q < WALK_T / 64; q++) wbase[q][k] += b0;
  }
  __syncthreads();
  if (nch) {
    uint32_t wp = wbase[w][0] + pre;
    if (wp + nch > cap_f) {
      atomicOr(&ctr->overflow, 1u);
    } else if (act == A_SELF_SPLIT) {
      for (int i = 0; i < 8; i++) {
        if (ra.ch[i] < 0) continue;
        for (int j = i; j < 8; j++)
          if (ra.ch[j] >= 0) fout[wp++] = make_uint2((uint32_t)ra.ch[i], (uint32_t)ra.ch[j]);
      }
    } else if (act == A_SPLIT_B) {
      for (int j = 0; j < 8; j++)
        if (rb.ch[j] >= 0) fout[wp++] = make_uint2(a, (uint32_t)rb.ch[j]);
    } else {
      for (int i = 0; i < 8; i++)
        if (ra.ch[i] >= 0) fout[wp++] = make_uint2((uint32_t)ra.ch[i], b);
    }
  }
  if (act == A_M2L) {
    const uint32_t wp = wbase[w][1] + (uint32_t)__popcll(bm & lt);
    if (wp < cap_m2l)
      m2l[wp] = make_uint2(a, b);
    else
      atomicOr(&ctr->overflow, 2u);
  }
  if (act == A_P2P) {
    const uint32_t wp = wbase[w][2] + (uint32_t)__popcll(bp & lt);
    if (wp < cap_p2p)
      p2p[wp] = make_uint2(a, b);
    else
      atomicOr(&ctr->overflow, 4u);
  }
}

void launch_walk_round(const uint2* fin, uint32_t nf, uint2* fout, uint32_t cap_f, uint2* m2l,
                       uint32_t cap_m2l, uint2* p2p, uint32_t cap_p2p, WalkCounters* ctr,
                       const void* rec, const double* mass, MacParams mac, hipStream_t s) {
  hipLaunchKernelGGL(k_walk_round, dim3((nf + WALK_T - 1) / WALK_T), dim3(WALK_T), 0, s, fin, nf,
                     fout, cap_f,
                     m2l, cap_m2l, p2p, cap_p2p, ctr, (const WalkRec*)rec, mass, mac);
}

// Directed M2L entries (target << sbits | source), both directions per pair.
__global__ void __launch_bounds__(256) k_keys_m2l(const uint2* __restrict__ pairs, uint32_t np,
                                                 int sbits, uint64_t* __restrict__ keys) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= np) return;
  uint2 pr = pairs[i];
  keys[2 * (size_t)i] = ((uint64_t)pr.y << sbits) | pr.x;
  keys[2 * (size_t)i + 1] = ((uint64_t)pr.x << sbits) | pr.y;
}

void launch_directed_keys_m2l(const uint2* pairs, uint32_t npairs, int sbits, uint64_t* keys,
                              hipStream_t s) {
  if (npairs == 0) return;
  hipLaunchKernelGGL(k_keys_m2l, dim3((npairs + 255) / 256), dim3(256), 0, s, pairs, npairs, sbits,
                     keys);
}

// Directed P2P entries; a self pair (a,a) yields one entry and a sentinel
// (all ones in the key range) that sorts past every real entry.
__global__ void __launch_bounds__(256) k_keys_p2p(const uint2* __restrict__ pairs, uint32_t np,
                                                 int sbits, uint64_t* __restrict__ keys,
                                                 uint32_t* __restrict__ count) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= np) return;
  uint2 pr = pairs[i];
  keys[2 * (size_t)i] = ((uint64_t)pr.y << sbits) | pr.x;
  if (pr.x == pr.y) {
    keys[2 * (size_t)i + 1] = (1ull << (2 * sbits)) - 1ull;
    atomicAdd(count, 1u);
  } else {
    keys[2 * (size_t)i + 1] = ((uint64_t)pr.x << sbits) | pr.y;
  }
}

void launch_directed_keys_p2p(const uint2* pairs, uint32_t npairs, int sbits, uint64_t* keys,
                              uint32_t* count, hipStream_t s) {
  if (npairs == 0) return;
  hipLaunchKernelGGL(k_keys_p2p, dim3((npairs + 255) / 256), dim3(256), 0, s, pairs, npairs, sbits,
                     keys, count);
}

// Row pointers of the target-sorted directed entries over all cells.
__global__ void __launch_bounds__(256) k_csr_ptr(const uint64_t* __restrict__ keys, uint32_t nnz,
                                                int sbits, uint32_t ntarget,
                                                uint32_t* __restrict__ ptr) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (nnz == 0) {
    for (uint32_t x = i; x <= ntarget; x += gridDim.x * 256) ptr[x] = 0;
    return;
  }
  if (i >= nnz) return;
  uint32_t t = (uint32_t)(keys[i] >> sbits);
  int64_t tp = i == 0 ? -1 : (int64_t)(keys[i - 1] >> sbits);
  for (int64_t x = tp + 1; x <= (int64_t)t; x++) ptr[x] = i;
  if (i == nnz - 1)
    for (uint32_t x = t + 1; x <= ntarget; x++) ptr[x] = nnz;
}

void launch_csr_ptr(const uint64_t* keys, uint32_t nnz, int sbits, uint32_t ntarget, uint32_t* ptr,
                    hipStream_t s) {
  uint32_t threads = nnz ? nnz : ntarget + 1;
  hipLaunchKernelGGL(k_csr_ptr, dim3((threads + 255) / 256), dim3(256), 0, s, keys, nnz, sbits,
                     ntarget, ptr);
}

}  // namespace fmmh
