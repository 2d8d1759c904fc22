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

struct WalkCells {
  const uint8_t* __restrict__ leaf;
  const int32_t* __restrict__ children;
  const double* __restrict__ cx;
  const double* __restrict__ cy;
  const double* __restrict__ cz;
  const double* __restrict__ rho;
  const double* __restrict__ mass;
};

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

// mac.rs:64-95 (+ theta_max: r2*theta_max^2 <= (rhoA+rhoB)^2 rejects; with
// theta_max = 1 the product is r2 itself).
__device__ __forceinline__ bool accept(uint32_t a, uint32_t b, const WalkCells& c,
                                       const MacParams& mp) {
  double dx = c.cx[b] - c.cx[a], dy = c.cy[b] - c.cy[a], dz = c.cz[b] - c.cz[a];
  double r2 = dx * dx + dy * dy + dz * dz;
  const double ra = c.rho[a], rb = c.rho[b];
  double sum_rho = ra + rb;
  if (r2 * mp.theta2 <= sum_rho * sum_rho) return false;
  if (!mp.use_err) return true;
  double r = sqrt(r2);
  if (r == 0.0) return false;
  const double ma = c.mass[a];
  double est = error_estimate(a, rb, ma, ra, r, mp);
  if (est == 0.0) return true;
  return est * ma / r2 < mp.eps * mp.fmin[b];
}

enum : int { A_NONE = 0, A_P2P = 1, A_M2L = 2, A_SELF_SPLIT = 3, A_SPLIT_A = 4, A_SPLIT_B = 5 };

__global__ void __launch_bounds__(256) k_walk_round(const uint2* __restrict__ fin, uint32_t nf,
                                                   uint2* __restrict__ fout, uint32_t cap_f,
                                                   uint2* __restrict__ m2l, uint32_t cap_m2l,
                                                   uint2* __restrict__ p2p, uint32_t cap_p2p,
                                                   WalkCounters* __restrict__ ctr, WalkCells c,
                                                   MacParams mp) {
  const int lane = threadIdx.x & 63;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = t < nf;
  uint32_t a = 0, b = 0;
  int act = A_NONE;
  uint32_t nch = 0;
  if (valid) {
    uint2 pr = fin[t];
    a = pr.x;
    b = pr.y;
    const bool la = c.leaf[a], lb = c.leaf[b];
    if (a == b) {
      if (la) {
        act = A_P2P;
      } else {
        act = A_SELF_SPLIT;
        uint32_t k = 0;
        for (int o = 0; o < 8; o++) k += c.children[(size_t)a * 8 + o] >= 0 ? 1u : 0u;
        nch = k * (k + 1) / 2;
      }
    } else if (accept(a, b, c, mp) && accept(b, a, c, mp)) {
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
        split_b = !(c.rho[a] >= c.rho[b]);
      act = split_b ? A_SPLIT_B : A_SPLIT_A;
      uint32_t s = split_b ? b : a;
      for (int o = 0; o < 8; o++) nch += c.children[(size_t)s * 8 + o] >= 0 ? 1u : 0u;
    }
  }
  // --- child pairs: wave-aggregated reservation in the next frontier
  uint32_t wtot;
  uint32_t pre = wave_excl_scan(nch, &wtot);
  uint32_t base = 0;
  if (lane == 0 && wtot) base = atomicAdd(&ctr->next, wtot);
  base = __shfl(base, 0);
  if (nch) {
    uint32_t w = base + pre;
    if (w + nch > cap_f) {
      atomicOr(&ctr->overflow, 1u);
    } else if (act == A_SELF_SPLIT) {
      const int32_t* ch = c.children + (size_t)a * 8;
      for (int i = 0; i < 8; i++) {
        if (ch[i] < 0) continue;
        for (int j = i; j < 8; j++)
          if (ch[j] >= 0) fout[w++] = make_uint2((uint32_t)ch[i], (uint32_t)ch[j]);
      }
    } else if (act == A_SPLIT_B) {
      const int32_t* ch = c.children + (size_t)b * 8;
      for (int j = 0; j < 8; j++)
        if (ch[j] >= 0) fout[w++] = make_uint2(a, (uint32_t)ch[j]);
    } else {
      const int32_t* ch = c.children + (size_t)a * 8;
      for (int i = 0; i < 8; i++)
        if (ch[i] >= 0) fout[w++] = make_uint2((uint32_t)ch[i], b);
    }
  }
  // --- M2L and P2P appends
  const uint64_t lt = lanemask_lt();
  uint64_t bm = __ballot(act == A_M2L);
  if (bm) {
    uint32_t mb = 0;
    if (lane == 0) mb = atomicAdd(&ctr->m2l, (uint32_t)__popcll(bm));
    mb = __shfl(mb, 0);
    if (act == A_M2L) {
      uint32_t w = mb + (uint32_t)__popcll(bm & lt);
      if (w < cap_m2l)
        m2l[w] = make_uint2(a, b);
      else
        atomicOr(&ctr->overflow, 2u);
    }
  }
  uint64_t bp = __ballot(act == A_P2P);
  if (bp) {
    uint32_t pb = 0;
    if (lane == 0) pb = atomicAdd(&ctr->p2p, (uint32_t)__popcll(bp));
    pb = __shfl(pb, 0);
    if (act == A_P2P) {
      uint32_t w = pb + (uint32_t)__popcll(bp & lt);
      if (w < cap_p2p)
        p2p[w] = make_uint2(a, b);
      else
        atomicOr(&ctr->overflow, 4u);
    }
  }
}

void launch_walk_round(const uint2* fin, uint32_t nf, uint2* fout, uint32_t cap_f, uint2* m2l,
                       uint32_t cap_m2l, uint2* p2p, uint32_t cap_p2p, WalkCounters* ctr,
                       DevCells c, MacParams mac, hipStream_t s) {
  WalkCells wc{c.leaf, c.children, c.cx, c.cy, c.cz, c.rho, c.mass};
  hipLaunchKernelGGL(k_walk_round, dim3((nf + 255) / 256), dim3(256), 0, s, fin, nf, fout, cap_f,
                     m2l, cap_m2l, p2p, cap_p2p, ctr, wc, mac);
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
