// tree.hip — Morton-key octree build on gfx950 (compiled -ffp-contract=off).
//
// Follows crates/fmm tree.rs:94-227 and reproduces its output bit for bit:
//   bbox (min/max, order-free) -> 63-bit Morton keys with the reference's
//   quantiser (every axis measured from the x minimum, tree.rs:37-46) ->
//   stable LSD radix sort (prims.hip) -> level-by-level octant splits ->
//   DFS-preorder ids -> geometry in the reference's arithmetic order.
//
// DFS preorder without recursion: preorder sorts cells by (first slot,
// level), because the only cells sharing a first slot form a parent->first-
// child chain at consecutive levels. So id = (cells whose range starts
// before `first`) + (level - top level of the chain at `first`); the first
// term is an exclusive scan over slots of the per-slot chain lengths.
#include "common.h"
#include "internal.h"

namespace fmmh {

constexpr int BITS = 21;  // tree.rs:16
constexpr int TB = 256;

__global__ void __launch_bounds__(TB) k_bbox_partial(const double* __restrict__ pos, size_t n,
                                                    double* __restrict__ partial) {
  __shared__ double sh[6][TB];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = (size_t)blockIdx.x * TB + threadIdx.x; i < n; i += (size_t)gridDim.x * TB) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
      double v = pos[3 * i + d];
      lo[d] = fmin(lo[d], v);
      hi[d] = fmax(hi[d], v);
    }
  }
#pragma unroll
  for (int d = 0; d < 3; d++) {
    sh[d][threadIdx.x] = lo[d];
    sh[3 + d][threadIdx.x] = hi[d];
  }
  __syncthreads();
  for (int st = TB / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
#pragma unroll
      for (int d = 0; d < 3; d++) {
        sh[d][threadIdx.x] = fmin(sh[d][threadIdx.x], sh[d][threadIdx.x + st]);
        sh[3 + d][threadIdx.x] = fmax(sh[3 + d][threadIdx.x], sh[3 + d][threadIdx.x + st]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) partial[(size_t)blockIdx.x * 6 + threadIdx.x] = sh[threadIdx.x][0];
}

// tree.rs:99-111: side = max extent (at least f64::MIN_POSITIVE), inv = 1/side.
__global__ void __launch_bounds__(TB) k_bbox_final(const double* __restrict__ partial, int nb,
                                                  BBox* __restrict__ out) {
  __shared__ double sh[6][TB];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int b = threadIdx.x; b < nb; b += TB)
    for (int d = 0; d < 3; d++) {
      lo[d] = fmin(lo[d], partial[6 * b + d]);
      hi[d] = fmax(hi[d], partial[6 * b + 3 + d]);
    }
  for (int d = 0; d < 3; d++) {
    sh[d][threadIdx.x] = lo[d];
    sh[3 + d][threadIdx.x] = hi[d];
  }
  __syncthreads();
  for (int st = TB / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int d = 0; d < 3; d++) {
        sh[d][threadIdx.x] = fmin(sh[d][threadIdx.x], sh[d][threadIdx.x + st]);
        sh[3 + d][threadIdx.x] = fmax(sh[3 + d][threadIdx.x], sh[3 + d][threadIdx.x + st]);
      }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  for (int d = 0; d < 3; d++) {
    lo[d] = sh[d][0];
    hi[d] = sh[3 + d][0];
  }
  double side = fmax(fmax(fmax(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]), 2.2250738585072014e-308);
  for (int d = 0; d < 3; d++) {
    out->lo[d] = lo[d];
    out->hi[d] = hi[d];
  }
  out->side = side;
  out->inv_side = 1.0 / side;
}

void launch_bbox(const double* pos, size_t n, double* partial, BBox* out, hipStream_t s) {
  int nb = (int)((n + TB - 1) / TB);
  if (nb > 1024) nb = 1024;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_bbox_partial, dim3(nb), dim3(TB), 0, s, pos, n, partial);
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(TB), 0, s, partial, nb, out);
}

// Ensembles: one block per system reduces its body range [off[s], off[s+1]).
__global__ void __launch_bounds__(TB) k_bbox_sys(const double* __restrict__ pos,
                                                const uint64_t* __restrict__ off,
                                                BBox* __restrict__ out) {
  __shared__ double sh[6][TB];
  const uint32_t sys = blockIdx.x;
  const size_t b = off[sys], e = off[sys + 1];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = b + threadIdx.x; i < e; i += TB)
    for (int d = 0; d < 3; d++) {
      double v = pos[3 * i + d];
      lo[d] = fmin(lo[d], v);
      hi[d] = fmax(hi[d], v);
    }
  for (int d = 0; d < 3; d++) {
    sh[d][threadIdx.x] = lo[d];
    sh[3 + d][threadIdx.x] = hi[d];
  }
  __syncthreads();
  for (int st = TB / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int d = 0; d < 3; d++) {
        sh[d][threadIdx.x] = fmin(sh[d][threadIdx.x], sh[d][threadIdx.x + st]);
        sh[3 + d][threadIdx.x] = fmax(sh[3 + d][threadIdx.x], sh[3 + d][threadIdx.x + st]);
      }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  for (int d = 0; d < 3; d++) {
    lo[d] = sh[d][0];
    hi[d] = sh[3 + d][0];
  }
  double side = fmax(fmax(fmax(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]), 2.2250738585072014e-308);
  BBox o;
  for (int d = 0; d < 3; d++) {
    o.lo[d] = lo[d];
    o.hi[d] = hi[d];
  }
  o.side = side;
  o.inv_side = 1.0 / side;
  out[sys] = o;
}

void launch_bbox_sys(const double* pos, const uint64_t* off, uint32_t nsys, BBox* out,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_bbox_sys, dim3(nsys), dim3(TB), 0, s, pos, off, out);
}

// Spread the low 21 bits of x to every third bit (tree.rs:28-35).
__device__ __forceinline__ uint64_t dilate21(uint64_t x) {
  x &= (1ull << BITS) - 1;
  x = (x | (x << 32)) & 0x001f00000000ffffull;
  x = (x | (x << 16)) & 0x001f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  return (x | (x << 2)) & 0x1249249249249249ull;
}

// tree.rs:37-46: clamp((c - lo[0]) * inv_side, 0, 1) * (2^21-1), truncated;
// NaN keeps through the clamp and converts to 0.
__device__ __forceinline__ uint64_t quantise(double c, double lo0, double inv_side) {
  double v = (c - lo0) * inv_side;
  if (v < 0.0) v = 0.0;
  if (v > 1.0) v = 1.0;
  double w = v * 2097151.0;
  if (!(w == w)) return 0;
  return (uint64_t)w;
}

__global__ void __launch_bounds__(TB) k_keys(const double* __restrict__ pos,
                                            const double* __restrict__ mass, size_t n,
                                            const BBox* __restrict__ bb,
                                            uint64_t* __restrict__ keys,
                                            uint32_t* __restrict__ vals,
                                            uint32_t* __restrict__ flags,
                                            double4* __restrict__ pk) {
  size_t i = (size_t)blockIdx.x * TB + threadIdx.x;
  if (i >= n) return;
  const double lo0 = bb->lo[0], inv = bb->inv_side;
  double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2], m = mass[i];
  uint64_t key = dilate21(quantise(z, lo0, inv));
  key |= dilate21(quantise(y, lo0, inv)) << 1;
  key |= dilate21(quantise(x, lo0, inv)) << 2;
  keys[i] = key;
  vals[i] = (uint32_t)i;
  pk[i] = make_double4(x, y, z, m);
  if (!(isfinite(x) && isfinite(y) && isfinite(z) && isfinite(m))) atomicOr(flags, 1u);
}

__device__ __forceinline__ uint32_t system_of(const uint64_t* __restrict__ off, uint32_t nsys,
                                              uint64_t i) {
  uint32_t lo = 0, hi = nsys;  // off[lo] <= i < off[lo + 1]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= i)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Ensemble keys, first half: the body's Morton key within its own system's
// box (its standalone key). The bodies are sorted by these keys first, so
// bodies whose keys agree above the low 3V bits keep their standalone order.
__global__ void __launch_bounds__(TB) k_keys_sys(const double* __restrict__ pos,
                                                const double* __restrict__ mass, size_t n,
                                                const uint64_t* __restrict__ off, uint32_t nsys,
                                                int vlev, const BBox* __restrict__ bbs,
                                                uint64_t* __restrict__ keys,
                                                uint32_t* __restrict__ vals,
                                                uint32_t* __restrict__ flags,
                                                double4* __restrict__ pk) {
  size_t i = (size_t)blockIdx.x * TB + threadIdx.x;
  if (i >= n) return;
  const BBox* bb = bbs + system_of(off, nsys, i);
  const double lo0 = bb->lo[0], inv = bb->inv_side;
  double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2], m = mass[i];
  uint64_t key = dilate21(quantise(z, lo0, inv));
  key |= dilate21(quantise(y, lo0, inv)) << 1;
  key |= dilate21(quantise(x, lo0, inv)) << 2;
  keys[i] = key;
  vals[i] = (uint32_t)i;
  pk[i] = make_double4(x, y, z, m);
  if (!(isfinite(x) && isfinite(y) && isfinite(z) && isfinite(m))) atomicOr(flags, 1u);
}

// Ensemble keys, second half (on the bodies sorted by standalone key): the
// standalone key with its low 3V bits dropped, under a 3V-bit system prefix
// (V virtual levels, nsys <= 8^V). A stable sort on the prefix bits then
// orders the bodies by (system, standalone key), and below level V the
// forest splits exactly like each system's standalone tree.
__global__ void __launch_bounds__(TB) k_keys_prefix(uint64_t* __restrict__ keys,
                                                   const uint32_t* __restrict__ ids, size_t n,
                                                   const uint64_t* __restrict__ off, uint32_t nsys,
                                                   int vlev) {
  size_t k = (size_t)blockIdx.x * TB + threadIdx.x;
  if (k >= n) return;
  const uint32_t sys = system_of(off, nsys, ids[k]);
  keys[k] = ((uint64_t)sys << (63 - 3 * vlev)) | (keys[k] >> (3 * vlev));
}

void launch_keys_prefix(uint64_t* keys, const uint32_t* ids, size_t n, const uint64_t* off,
                        uint32_t nsys, int vlev, hipStream_t s) {
  hipLaunchKernelGGL(k_keys_prefix, dim3((n + TB - 1) / TB), dim3(TB), 0, s, keys, ids, n, off,
                     nsys, vlev);
}

void launch_keys_sys(const double* pos, const double* mass, size_t n, const uint64_t* off,
                     uint32_t nsys, int vlev, const BBox* bbs, uint64_t* keys, uint32_t* vals,
                     uint32_t* flags, double4* pk, hipStream_t s) {
  hipLaunchKernelGGL(k_keys_sys, dim3((n + TB - 1) / TB), dim3(TB), 0, s, pos, mass, n, off, nsys,
                     vlev, bbs, keys, vals, flags, pk);
}

void launch_keys(const double* pos, const double* mass, size_t n, const BBox* bb, uint64_t* keys,
                 uint32_t* vals, uint32_t* flags, double4* pk, hipStream_t s) {
  hipLaunchKernelGGL(k_keys, dim3((n + TB - 1) / TB), dim3(TB), 0, s, pos, mass, n, bb, keys, vals,
                     flags, pk);
}

// Bodies::from_parts (bodies.rs:27-47): id order AoS -> slot order SoA.
__global__ void __launch_bounds__(TB) k_gather(const double* __restrict__ pos,
                                              const double* __restrict__ mass,
                                              const uint32_t* __restrict__ order, size_t n,
                                              DevBodies b) {
  size_t k = (size_t)blockIdx.x * TB + threadIdx.x;
  if (k >= n) return;
  size_t id = order[k];
  const double x = pos[3 * id], y = pos[3 * id + 1], z = pos[3 * id + 2], m = mass[id];
  b.x[k] = x;
  b.y[k] = y;
  b.z[k] = z;
  b.m[k] = m;
  b.xyzm[k] = make_double4(x, y, z, m);
}

void launch_gather(const double* pos, const double* mass, const uint32_t* order, size_t n,
                   DevBodies b, hipStream_t s) {
  hipLaunchKernelGGL(k_gather, dim3((n + TB - 1) / TB), dim3(TB), 0, s, pos, mass, order, n, b);
}

// The same gather from the keys kernel's packed copy (x, y, z, m in id
// order, 32 bytes per body): one aligned 32-byte read per body instead of
// the position triple and the mass from two arrays.
__global__ void __launch_bounds__(TB) k_gather_packed(const double4* __restrict__ pk,
                                                     const uint32_t* __restrict__ order, size_t n,
                                                     DevBodies b) {
  size_t k = (size_t)blockIdx.x * TB + threadIdx.x;
  if (k >= n) return;
  const double4 v = pk[order[k]];
  b.x[k] = v.x;
  b.y[k] = v.y;
  b.z[k] = v.z;
  b.m[k] = v.w;
  b.xyzm[k] = v;
}

void launch_gather_packed(const double4* pk, const uint32_t* order, size_t n, DevBodies b,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_gather_packed, dim3((n + TB - 1) / TB), dim3(TB), 0, s, pk, order, n, b);
}

__global__ void k_tree_init(DevBfs bfs, uint32_t n, uint32_t* chain, uint8_t* top) {
  if (threadIdx.x != 0) return;
  bfs.first[0] = 0;
  bfs.count[0] = n;
  bfs.parent[0] = -1;
  bfs.oct[0] = 0;
  bfs.level[0] = 0;
  chain[0] = 1;
  top[0] = 0;
}

void launch_tree_init(DevBfs bfs, size_t n, uint32_t* chain, uint8_t* top, hipStream_t s) {
  hipLaunchKernelGGL(k_tree_init, dim3(1), dim3(64), 0, s, bfs, (uint32_t)n, chain, top);
}

// tree.rs:414 leaf rule; the virtual levels of an ensemble (level < vlev,
// cells holding several systems) always split.
__device__ __forceinline__ bool is_internal(uint32_t count, int level, uint32_t n_leaf, int vlev) {
  return level < vlev || !(count <= n_leaf || level >= BITS);
}

// Octant boundaries of an internal cell (tree.rs:420-443): bnd[o] is the
// first slot whose octant digit at this level is >= o, so the children are
// bnd[o]..bnd[o+1]. One wave finds them by a 65-ary search per boundary
// (64 probes a round, all seven boundaries at once): a cell of N bodies
// takes about log65(N) rounds of loads instead of 7 log2(N) dependent ones.
__device__ __forceinline__ void octant_bounds_wave(const uint64_t* __restrict__ keys, uint32_t f,
                                                   uint32_t c, int level, int lane,
                                                   uint32_t bnd[9]) {
  const int shift = 3 * (BITS - (level + 1));
  uint32_t lo[8], hi[8];
#pragma unroll
  for (int o = 1; o < 8; o++) {
    lo[o] = f;
    hi[o] = f + c;
  }
  // digits are non-decreasing over the cell's slots, so every ballot below
  // is a prefix of the wave; a round shrinks a bracket of length L > 64 to
  // at most L/65 + 1 and settles one of length <= 64 (7 rounds cover 2^32)
  for (int it = 0; it < 8; it++) {
    bool more = false;
#pragma unroll
    for (int o = 1; o < 8; o++) {
      const uint32_t len = hi[o] - lo[o];
      if (len == 0) continue;
      more = true;
      if (len <= 64) {
        const bool v = (uint32_t)lane < len;
        const uint32_t d = v ? (uint32_t)((keys[lo[o] + (uint32_t)lane] >> shift) & 7u) : 8u;
        lo[o] += (uint32_t)__popcll(__ballot(d < (uint32_t)o));
        hi[o] = lo[o];
      } else {
        const uint32_t q = lo[o] + (uint32_t)(((uint64_t)(lane + 1) * len) / 65u);
        const uint32_t d = (uint32_t)((keys[q] >> shift) & 7u);
        const uint32_t t = (uint32_t)__popcll(__ballot(d < (uint32_t)o));
        const uint32_t nlo = t == 0 ? lo[o] : lo[o] + (uint32_t)(((uint64_t)t * len) / 65u) + 1u;
        const uint32_t nhi = t == 64 ? hi[o] : lo[o] + (uint32_t)(((uint64_t)(t + 1) * len) / 65u);
        lo[o] = nlo;
        hi[o] = nhi;
      }
    }
    if (!more) break;
  }
  bnd[0] = f;
#pragma unroll
  for (int o = 1; o < 8; o++) bnd[o] = lo[o];
  bnd[8] = f + c;
}

constexpr int SPLIT_WAVES = 4;

// One wave per cell of a level: child count and octant boundaries, which
// are kept (9 words a cell) for k_split_emit.
__global__ void __launch_bounds__(64 * SPLIT_WAVES) k_split_count(DevBfs bfs, uint32_t s0,
                                                                uint32_t cnt,
                                                                const uint64_t* __restrict__ keys,
                                                                uint32_t n_leaf, int vlev,
                                                                uint32_t* __restrict__ nchild,
                                                                uint32_t* __restrict__ bounds) {
  const uint32_t i = blockIdx.x * SPLIT_WAVES + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= cnt) return;
  const uint32_t b = s0 + i;
  const uint32_t f = bfs.first[b], c = bfs.count[b];
  const int lvl = bfs.level[b];
  uint32_t k = 0;
  if (is_internal(c, lvl, n_leaf, vlev)) {
    uint32_t bnd[9];
    octant_bounds_wave(keys, f, c, lvl, lane, bnd);
#pragma unroll
    for (int o = 0; o < 8; o++) k += bnd[o + 1] > bnd[o] ? 1u : 0u;
#pragma unroll
    for (int o = 0; o < 9; o++)
      if (lane == o) bounds[(size_t)i * 9 + o] = bnd[o];
  }
  if (lane == 0) nchild[i] = k;
}

void launch_split_count(DevBfs bfs, uint32_t s0, uint32_t cnt, const uint64_t* keys,
                        uint32_t n_leaf, int vlev, uint32_t* nchild, uint32_t* bounds,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_split_count, dim3((cnt + SPLIT_WAVES - 1) / SPLIT_WAVES),
                     dim3(64 * SPLIT_WAVES), 0, s, bfs, s0, cnt, keys, n_leaf, vlev, nchild,
                     bounds);
}

__global__ void __launch_bounds__(TB) k_split_emit(DevBfs bfs, uint32_t s0, uint32_t cnt,
                                                  uint32_t e0, uint32_t n_leaf, int vlev,
                                                  const uint32_t* __restrict__ bounds,
                                                  const uint32_t* __restrict__ child_off,
                                                  uint32_t* __restrict__ chain,
                                                  uint8_t* __restrict__ top) {
  uint32_t i = blockIdx.x * TB + threadIdx.x;
  if (i >= cnt) return;
  uint32_t b = s0 + i;
  uint32_t c = bfs.count[b];
  int lvl = bfs.level[b];
  if (!is_internal(c, lvl, n_leaf, vlev)) return;
  const uint32_t* bnd = bounds + (size_t)i * 9;
  uint32_t w = e0 + child_off[i];
  for (int o = 0; o < 8; o++) {
    uint32_t lo = bnd[o], hi = bnd[o + 1];
    if (hi <= lo) continue;
    bfs.first[w] = lo;
    bfs.count[w] = hi - lo;
    bfs.parent[w] = (int32_t)b;
    bfs.oct[w] = (uint8_t)o;
    bfs.level[w] = (uint8_t)(lvl + 1);
    // Ranges of one level are disjoint, so each slot is touched at most once
    // per launch; a slot seen for the first time starts a chain here.
    if (chain[lo] == 0) top[lo] = (uint8_t)(lvl + 1);
    chain[lo] += 1;
    w++;
  }
}

void launch_split_emit(DevBfs bfs, uint32_t s0, uint32_t cnt, uint32_t e0, uint32_t n_leaf,
                       int vlev, const uint32_t* bounds, const uint32_t* child_off,
                       uint32_t* chain, uint8_t* top, hipStream_t s) {
  hipLaunchKernelGGL(k_split_emit, dim3((cnt + TB - 1) / TB), dim3(TB), 0, s, bfs, s0, cnt, e0,
                     n_leaf, vlev, bounds, child_off, chain, top);
}

__global__ void __launch_bounds__(TB) k_dfs_ids(DevBfs bfs, uint32_t nc,
                                               const uint32_t* __restrict__ cstart,
                                               const uint8_t* __restrict__ top) {
  uint32_t b = blockIdx.x * TB + threadIdx.x;
  if (b >= nc) return;
  uint32_t f = bfs.first[b];
  bfs.dfs[b] = cstart[f] + (uint32_t)bfs.level[b] - (uint32_t)top[f];
}

void launch_dfs_ids(DevBfs bfs, uint32_t nc, const uint32_t* cstart, const uint8_t* top,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_dfs_ids, dim3((nc + TB - 1) / TB), dim3(TB), 0, s, bfs, nc, cstart, top);
}

__global__ void __launch_bounds__(TB) k_scatter_cells(DevBfs bfs, uint32_t nc, uint32_t n_leaf,
                                                     int vlev, DevCells c,
                                                     uint32_t* __restrict__ lvl_cells) {
  uint32_t b = blockIdx.x * TB + threadIdx.x;
  if (b >= nc) return;
  uint32_t d = bfs.dfs[b];
  uint32_t cnt = bfs.count[b];
  int lvl = bfs.level[b];
  c.first[d] = bfs.first[b];
  c.count[d] = cnt;
  c.level[d] = (uint8_t)lvl;
  c.leaf[d] = is_internal(cnt, lvl, n_leaf, vlev) ? 0 : 1;
  int32_t pb = bfs.parent[b];
  if (pb >= 0) {
    uint32_t pd = bfs.dfs[pb];
    c.parent[d] = (int32_t)pd;
    c.children[(size_t)pd * 8 + bfs.oct[b]] = (int32_t)d;
  } else {
    c.parent[d] = -1;
  }
  lvl_cells[b] = d;
}

void launch_scatter_cells(DevBfs bfs, uint32_t nc, uint32_t n_leaf, int vlev, DevCells c,
                          uint32_t* lvl_cells, hipStream_t s) {
  hipLaunchKernelGGL(k_scatter_cells, dim3((nc + TB - 1) / TB), dim3(TB), 0, s, bfs, nc, n_leaf,
                     vlev, c, lvl_cells);
}

// ---------------------------------------------------------------- geometry
__device__ __forceinline__ double dist3(double ax, double ay, double az, double bx, double by,
                                        double bz) {  // tree.rs:392-402
  double dx = ax - bx, dy = ay - by, dz = az - bz;
  return sqrt(dx * dx + dy * dy + dz * dz);
}

// Ritter grow/enclose sphere with the bbox-centred fallback, tree.rs:290-389.
// `get(i, c, r)` yields item i's centre and radius.
template <class Get>
__device__ void ritter(int nitems, Get get, double out[3]) {
  const double MINP = 2.2250738585072014e-308;
  double c0[3], r0;
  int a = 0, b = 0;
  double best = -1.0;
  double ca[3], ra;
  get(0, ca, &ra);
  for (int i = 0; i < nitems; i++) {
    double ci[3], ri;
    get(i, ci, &ri);
    double dx = ci[0] - ca[0], dy = ci[1] - ca[1], dz = ci[2] - ca[2];
    double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 > best) {
      best = d2;
      b = i;
    }
  }
  a = b;
  best = -1.0;
  get(a, ca, &ra);
  for (int i = 0; i < nitems; i++) {
    double ci[3], ri;
    get(i, ci, &ri);
    double dx = ci[0] - ca[0], dy = ci[1] - ca[1], dz = ci[2] - ca[2];
    double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 > best) {
      best = d2;
      b = i;
    }
  }
  double pa[3], pb[3], rad_a, rad_b;
  get(a, pa, &rad_a);
  get(b, pb, &rad_b);
  double dir[3];
  {
    double dx = pb[0] - pa[0], dy = pb[1] - pa[1], dz = pb[2] - pa[2];
    double d = fmax(sqrt(dx * dx + dy * dy + dz * dz), MINP);
    dir[0] = dx / d;
    dir[1] = dy / d;
    dir[2] = dz / d;
  }
  double fa[3] = {pa[0] + rad_a * dir[0], pa[1] + rad_a * dir[1], pa[2] + rad_a * dir[2]};
  double fb[3] = {pb[0] - rad_b * dir[0], pb[1] - rad_b * dir[1], pb[2] - rad_b * dir[2]};
  c0[0] = (fa[0] + fb[0]) * 0.5;
  c0[1] = (fa[1] + fb[1]) * 0.5;
  c0[2] = (fa[2] + fb[2]) * 0.5;
  r0 = dist3(fa[0], fa[1], fa[2], fb[0], fb[1], fb[2]) * 0.5;
  for (int pass = 0; pass < 2; pass++) {
    for (int i = 0; i < nitems; i++) {
      double q[3], rad;
      get(i, q, &rad);
      double d = dist3(q[0], q[1], q[2], c0[0], c0[1], c0[2]);
      if (d + rad > r0) {
        double dd = fmax(d, MINP);
        double dr[3] = {(q[0] - c0[0]) / dd, (q[1] - c0[1]) / dd, (q[2] - c0[2]) / dd};
        double pnt[3] = {q[0] + rad * dr[0], q[1] + rad * dr[1], q[2] + rad * dr[2]};
        double d_far = d + rad;
        double t = (d_far - r0) / (2.0 * d_far);
        for (int k = 0; k < 3; k++) c0[k] += t * (pnt[k] - c0[k]);
        r0 = (d_far + r0) * 0.5;
      }
    }
  }
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = 0; i < nitems; i++) {
    double q[3], rad;
    get(i, q, &rad);
    for (int d = 0; d < 3; d++) {
      lo[d] = fmin(lo[d], q[d] - rad);
      hi[d] = fmax(hi[d], q[d] + rad);
    }
  }
  double cb[3] = {(lo[0] + hi[0]) * 0.5, (lo[1] + hi[1]) * 0.5, (lo[2] + hi[2]) * 0.5};
  double rb = 0.0;
  for (int i = 0; i < nitems; i++) {
    double q[3], rad;
    get(i, q, &rad);
    rb = fmax(rb, dist3(q[0], q[1], q[2], cb[0], cb[1], cb[2]) + rad);
  }
  if (rb < r0) {
    out[0] = cb[0];
    out[1] = cb[1];
    out[2] = cb[2];
  } else {
    out[0] = c0[0];
    out[1] = c0[1];
    out[2] = c0[2];
  }
}

// tree.rs:152-227, one thread per cell of a level (children done first).
// which = 0: every cell of the list; 1: only its leaves; 2: only its internal cells.
__global__ void __launch_bounds__(TB) k_geometry(const uint32_t* __restrict__ cells, uint32_t cnt,
                                                DevCells c, DevBodies b, int mode, int which) {
  uint32_t i = blockIdx.x * TB + threadIdx.x;
  if (i >= cnt) return;
  uint32_t id = cells[i];
  if ((which == 1 && !c.leaf[id]) || (which == 2 && c.leaf[id])) return;
  double ms = 0.0;
  double cen[3];
  double rho = 0.0;
  if (c.leaf[id]) {
    uint32_t f = c.first[id], e = f + c.count[id];
    for (uint32_t k = f; k < e; k++) ms += b.m[k];
    if (mode == 0) {  // com_of_slots, tree.rs:257-275
      double com[3] = {0.0, 0.0, 0.0};
      for (uint32_t k = f; k < e; k++) {
        double m = b.m[k];
        com[0] += m * b.x[k];
        com[1] += m * b.y[k];
        com[2] += m * b.z[k];
      }
      cen[0] = com[0] / ms;
      cen[1] = com[1] / ms;
      cen[2] = com[2] / ms;
    } else {
      ritter(
          (int)(e - f),
          [&](int q, double* qc, double* qr) {
            qc[0] = b.x[f + q];
            qc[1] = b.y[f + q];
            qc[2] = b.z[f + q];
            *qr = 0.0;
          },
          cen);
    }
    for (uint32_t k = f; k < e; k++) rho = fmax(rho, dist3(b.x[k], b.y[k], b.z[k], cen[0], cen[1], cen[2]));
  } else {
    const int32_t* ch = c.children + (size_t)id * 8;
    for (int o = 0; o < 8; o++)
      if (ch[o] >= 0) ms += c.mass[ch[o]];
    if (mode == 0) {
      double com[3] = {0.0, 0.0, 0.0};
      for (int o = 0; o < 8; o++) {
        int q = ch[o];
        if (q < 0) continue;
        double m = c.mass[q];
        com[0] += m * c.cx[q];
        com[1] += m * c.cy[q];
        com[2] += m * c.cz[q];
      }
      cen[0] = com[0] / ms;
      cen[1] = com[1] / ms;
      cen[2] = com[2] / ms;
    } else {
      int list[8];
      int nl = 0;
      for (int o = 0; o < 8; o++)
        if (ch[o] >= 0) list[nl++] = ch[o];
      ritter(
          nl,
          [&](int q, double* qc, double* qr) {
            int cc = list[q];
            qc[0] = c.cx[cc];
            qc[1] = c.cy[cc];
            qc[2] = c.cz[cc];
            *qr = c.rho[cc];
          },
          cen);
    }
    for (int o = 0; o < 8; o++) {
      int q = ch[o];
      if (q < 0) continue;
      rho = fmax(rho, c.rho[q] + dist3(c.cx[q], c.cy[q], c.cz[q], cen[0], cen[1], cen[2]));
    }
  }
  c.mass[id] = ms;
  c.cx[id] = cen[0];
  c.cy[id] = cen[1];
  c.cz[id] = cen[2];
  c.rho[id] = rho;
}

// Leaf geometry with one wave per leaf: the same values as k_geometry's
// leaf branch (tree.rs:152-227, 257-275, 290-389). Order-dependent sums
// (mass, centre of mass) run in slot order over bodies read from the lanes
// that hold them; the farthest-point searches, the box and the radius and
// rho maxima spread over the lanes (a maximum, ties to the lowest slot, does
// not depend on the order); the grow passes test a chunk of bodies against
// the current sphere at once and apply the first failing body's update, then
// test the bodies after it again, which is exactly the sequential pass.
constexpr int GEO_WAVES = 4;

__device__ __forceinline__ double wave_max_pos(double v) {  // v >= +0 in every lane
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Farthest body from slot a (first slot on ties), over the leaf's bodies.
__device__ __forceinline__ uint32_t wave_farthest(DevBodies b, uint32_t f, uint32_t n, uint32_t a,
                                                  int lane) {
  const double ax = b.x[f + a], ay = b.y[f + a], az = b.z[f + a];
  double best = -1.0;
  uint32_t bi = 0;
  for (uint32_t i = (uint32_t)lane; i < n; i += 64) {
    const double dx = b.x[f + i] - ax, dy = b.y[f + i] - ay, dz = b.z[f + i] - az;
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 > best) {
      best = d2;
      bi = i;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double ob = __shfl_xor(best, o, 64);
    const uint32_t oi = (uint32_t)__shfl_xor((int)bi, o, 64);
    if (ob > best || (ob == best && oi < bi)) {
      best = ob;
      bi = oi;
    }
  }
  return bi;
}

__global__ void __launch_bounds__(64 * GEO_WAVES) k_geometry_leaf(
    const uint32_t* __restrict__ leaf_ids, uint32_t nleaves, DevCells c, DevBodies b, int mode) {
  // per wave: one chunk of the leaf's bodies (mass, x, y, z) for the serial sums
  __shared__ double chunk[GEO_WAVES][4][64];
  const uint32_t wv = threadIdx.x >> 6;
  const uint32_t li = blockIdx.x * GEO_WAVES + wv;
  const int lane = threadIdx.x & 63;
  if (li >= nleaves) return;
  const uint32_t id = leaf_ids[li];
  const uint32_t f = c.first[id], n = c.count[id];
  // Slot-ordered sums over chunks of 64 bodies: the lanes stage the chunk in
  // LDS, then lane 0 sums the masses (m * 1) and lanes 1-3 the mass-weighted
  // coordinates, each in slot order (the wave's LDS operations run in program
  // order; the wave barrier keeps the compiler from reordering them).
  double sum = 0.0;  // lane 0: mass; lanes 1-3: sum of m x, m y, m z
  for (uint32_t base = 0; base < n; base += 64) {
    const uint32_t i = base + (uint32_t)lane;
    const bool have = i < n;
    const uint32_t cnt = min(64u, n - base);
    chunk[wv][0][lane] = have ? b.m[f + i] : 0.0;
    if (mode == 0) {  // com_of_slots, tree.rs:257-275
      chunk[wv][1][lane] = have ? b.x[f + i] : 0.0;
      chunk[wv][2][lane] = have ? b.y[f + i] : 0.0;
      chunk[wv][3][lane] = have ? b.z[f + i] : 0.0;
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < (mode == 0 ? 4 : 1)) {
      for (uint32_t k = 0; k < cnt; k++) {
        const double m = chunk[wv][0][k];
        sum += m * (lane == 0 ? 1.0 : chunk[wv][lane][k]);
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  const double ms = lane_read(sum, 0);
  double cen[3];
  if (mode == 0) {
    cen[0] = lane_read(sum, 1) / ms;
    cen[1] = lane_read(sum, 2) / ms;
    cen[2] = lane_read(sum, 3) / ms;
  } else {  // Ritter sphere, tree.rs:290-389 (bodies have radius 0)
    const double MINP = 2.2250738585072014e-308;
    const double rad = 0.0;
    const uint32_t a = wave_farthest(b, f, n, 0, lane);
    const uint32_t bb = wave_farthest(b, f, n, a, lane);
    const double pa[3] = {b.x[f + a], b.y[f + a], b.z[f + a]};
    const double pb[3] = {b.x[f + bb], b.y[f + bb], b.z[f + bb]};
    double dir[3];
    {
      const double dx = pb[0] - pa[0], dy = pb[1] - pa[1], dz = pb[2] - pa[2];
      const double d = fmax(sqrt(dx * dx + dy * dy + dz * dz), MINP);
      dir[0] = dx / d;
      dir[1] = dy / d;
      dir[2] = dz / d;
    }
    const double fa[3] = {pa[0] + rad * dir[0], pa[1] + rad * dir[1], pa[2] + rad * dir[2]};
    const double fb[3] = {pb[0] - rad * dir[0], pb[1] - rad * dir[1], pb[2] - rad * dir[2]};
    double c0[3] = {(fa[0] + fb[0]) * 0.5, (fa[1] + fb[1]) * 0.5, (fa[2] + fb[2]) * 0.5};
    double r0 = dist3(fa[0], fa[1], fa[2], fb[0], fb[1], fb[2]) * 0.5;
    for (int pass = 0; pass < 2; pass++) {
      for (uint32_t base = 0; base < n; base += 64) {
        const uint32_t i = base + (uint32_t)lane;
        const bool have = i < n;
        const double qx = have ? b.x[f + i] : 0.0, qy = have ? b.y[f + i] : 0.0,
                     qz = have ? b.z[f + i] : 0.0;
        int start = 0;
        // each round either ends the chunk or moves start past one lane
        while (start < 64) {
          double d = 0.0;
          bool out = false;
          if (have && lane >= start) {
            d = dist3(qx, qy, qz, c0[0], c0[1], c0[2]);
            out = d + rad > r0;
          }
          const uint64_t m = __ballot(out);
          if (m == 0) break;
          const int j = __ffsll((long long)m) - 1;
          const double q[3] = {__shfl(qx, j, 64), __shfl(qy, j, 64), __shfl(qz, j, 64)};
          const double dj = __shfl(d, j, 64);
          const double dd = fmax(dj, MINP);
          const double dr[3] = {(q[0] - c0[0]) / dd, (q[1] - c0[1]) / dd, (q[2] - c0[2]) / dd};
          const double pnt[3] = {q[0] + rad * dr[0], q[1] + rad * dr[1], q[2] + rad * dr[2]};
          const double d_far = dj + rad;
          const double t = (d_far - r0) / (2.0 * d_far);
          for (int k = 0; k < 3; k++) c0[k] += t * (pnt[k] - c0[k]);
          r0 = (d_far + r0) * 0.5;
          start = j + 1;
        }
      }
    }
    // the box: a min/max does not depend on the order (up to the sign of a
    // zero extreme, which changes no centre unless both extremes are zero)
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = (uint32_t)lane; i < n; i += 64) {
      const double q[3] = {b.x[f + i], b.y[f + i], b.z[f + i]};
      for (int d = 0; d < 3; d++) {
        lo[d] = fmin(lo[d], q[d] - rad);
        hi[d] = fmax(hi[d], q[d] + rad);
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
      for (int d = 0; d < 3; d++) {
        lo[d] = fmin(lo[d], __shfl_xor(lo[d], o, 64));
        hi[d] = fmax(hi[d], __shfl_xor(hi[d], o, 64));
      }
    const double cb[3] = {(lo[0] + hi[0]) * 0.5, (lo[1] + hi[1]) * 0.5, (lo[2] + hi[2]) * 0.5};
    double rb = 0.0;
    for (uint32_t i = (uint32_t)lane; i < n; i += 64)
      rb = fmax(rb, dist3(b.x[f + i], b.y[f + i], b.z[f + i], cb[0], cb[1], cb[2]) + rad);
    rb = wave_max_pos(rb);
    for (int k = 0; k < 3; k++) cen[k] = rb < r0 ? cb[k] : c0[k];
  }
  double rho = 0.0;
  for (uint32_t i = (uint32_t)lane; i < n; i += 64)
    rho = fmax(rho, dist3(b.x[f + i], b.y[f + i], b.z[f + i], cen[0], cen[1], cen[2]));
  rho = wave_max_pos(rho);
  if (lane == 0) {
    c.mass[id] = ms;
    c.cx[id] = cen[0];
    c.cy[id] = cen[1];
    c.cz[id] = cen[2];
    c.rho[id] = rho;
  }
}

void launch_geometry_leaves(const uint32_t* leaf_ids, uint32_t nleaves, DevCells c, DevBodies b,
                            int center_mode, hipStream_t s) {
  if (nleaves == 0) return;
  hipLaunchKernelGGL(k_geometry_leaf, dim3((nleaves + GEO_WAVES - 1) / GEO_WAVES),
                     dim3(64 * GEO_WAVES), 0, s, leaf_ids, nleaves, c, b, center_mode);
}

void launch_geometry(const uint32_t* lvl_cells, uint32_t cnt, DevCells c, DevBodies b,
                     int center_mode, hipStream_t s, int which) {
  if (cnt == 0) return;
  hipLaunchKernelGGL(k_geometry, dim3((cnt + TB - 1) / TB), dim3(TB), 0, s, lvl_cells, cnt, c, b,
                     center_mode, which);
}

// Ropes of the mirrored DFS (children visited 7..0): the cell visited after
// a child's subtree is its next lower existing sibling, else the parent's rope.
__global__ void __launch_bounds__(TB) k_ropes(const uint32_t* __restrict__ cells, uint32_t cnt,
                                             DevCells c) {
  uint32_t i = blockIdx.x * TB + threadIdx.x;
  if (i >= cnt) return;
  uint32_t id = cells[i];
  if (id == 0) c.rope[0] = -1;
  if (c.leaf[id]) {
    c.last_child[id] = -1;
    return;
  }
  int32_t prev = (id == 0) ? -1 : c.rope[id];
  const int32_t* ch = c.children + (size_t)id * 8;
  for (int o = 0; o < 8; o++) {
    int q = ch[o];
    if (q < 0) continue;
    c.rope[q] = prev;
    prev = q;
  }
  c.last_child[id] = prev;
}

void launch_ropes(const uint32_t* lvl_cells, uint32_t cnt, DevCells c, hipStream_t s) {
  if (cnt == 0) return;
  hipLaunchKernelGGL(k_ropes, dim3((cnt + TB - 1) / TB), dim3(TB), 0, s, lvl_cells, cnt, c);
}

__global__ void __launch_bounds__(TB) k_leaf_flags(DevCells c, uint32_t nc, uint32_t* flags) {
  uint32_t i = blockIdx.x * TB + threadIdx.x;
  if (i < nc) flags[i] = c.leaf[i];
}

void launch_leaf_flags(DevCells c, uint32_t nc, uint32_t* flags, hipStream_t s) {
  hipLaunchKernelGGL(k_leaf_flags, dim3((nc + TB - 1) / TB), dim3(TB), 0, s, c, nc, flags);
}

__global__ void __launch_bounds__(TB) k_leaf_compact(DevCells c, uint32_t nc,
                                                    const uint32_t* __restrict__ off,
                                                    uint32_t* __restrict__ leaf_ids,
                                                    uint32_t* __restrict__ leaf_of_slot) {
  uint32_t i = blockIdx.x * TB + threadIdx.x;
  if (i >= nc || !c.leaf[i]) return;
  leaf_ids[off[i]] = i;
  uint32_t f = c.first[i], e = f + c.count[i];
  for (uint32_t k = f; k < e; k++) leaf_of_slot[k] = i;
}

void launch_leaf_compact(DevCells c, uint32_t nc, const uint32_t* off, uint32_t* leaf_ids,
                         uint32_t* leaf_of_slot, hipStream_t s) {
  hipLaunchKernelGGL(k_leaf_compact, dim3((nc + TB - 1) / TB), dim3(TB), 0, s, c, nc, off, leaf_ids,
                     leaf_of_slot);
}

// Internal cells of every level in BFS order (levels stay contiguous):
// flag, exclusive scan, compact; then each level's start in the compacted
// list is picked at the level's BFS offset.
__global__ void __launch_bounds__(TB) k_internal_flags(const uint32_t* __restrict__ lvl_cells,
                                                      uint32_t nc, DevCells c,
                                                      uint32_t* __restrict__ flags) {
  uint32_t i = blockIdx.x * TB + threadIdx.x;
  if (i < nc) flags[i] = c.leaf[lvl_cells[i]] ? 0u : 1u;
}

__global__ void __launch_bounds__(TB) k_internal_compact(const uint32_t* __restrict__ lvl_cells,
                                                        uint32_t nc, DevCells c,
                                                        const uint32_t* __restrict__ pos,
                                                        uint32_t* __restrict__ out) {
  uint32_t i = blockIdx.x * TB + threadIdx.x;
  if (i >= nc) return;
  const uint32_t id = lvl_cells[i];
  if (!c.leaf[id]) out[pos[i]] = id;
}

__global__ void k_level_pick(const uint32_t* __restrict__ pos, const uint32_t* __restrict__ total,
                             uint32_t nc, const uint32_t* __restrict__ at, uint32_t nl,
                             uint32_t* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nl) out[i] = at[i] < nc ? pos[at[i]] : *total;
}

// Leaves ordered by body count (for per-leaf kernels whose waves should see
// leaves of similar size): key = min(count, 65535) << 32 | leaf id, sorted
// on bits 32..47, then the ids read back.
__global__ void __launch_bounds__(TB) k_leaf_count_keys(const uint32_t* __restrict__ leaf_ids,
                                                       uint32_t nleaves, DevCells c,
                                                       uint64_t* __restrict__ keys) {
  uint32_t i = blockIdx.x * TB + threadIdx.x;
  if (i >= nleaves) return;
  const uint32_t id = leaf_ids[i];
  keys[i] = ((uint64_t)min(c.count[id], 65535u) << 32) | id;
}

__global__ void __launch_bounds__(TB) k_low32(const uint64_t* __restrict__ keys, uint32_t n,
                                             uint32_t* __restrict__ out) {
  uint32_t i = blockIdx.x * TB + threadIdx.x;
  if (i < n) out[i] = (uint32_t)keys[i];
}

void launch_leaves_by_count(const uint32_t* leaf_ids, uint32_t nleaves, DevCells c,
                            uint64_t* keys, uint64_t* keys_alt, uint32_t* tmp, uint32_t* out,
                            hipStream_t s) {
  if (nleaves == 0) return;
  const dim3 g((nleaves + TB - 1) / TB);
  hipLaunchKernelGGL(k_leaf_count_keys, g, dim3(TB), 0, s, leaf_ids, nleaves, c, keys);
  const int fl = radix_sort_u64(keys, nullptr, keys_alt, nullptr, nleaves, 32, 48, tmp, s);
  hipLaunchKernelGGL(k_low32, g, dim3(TB), 0, s, fl ? keys_alt : keys, nleaves, out);
}

void launch_internal_flags(const uint32_t* lvl_cells, uint32_t nc, DevCells c, uint32_t* flags,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_internal_flags, dim3((nc + TB - 1) / TB), dim3(TB), 0, s, lvl_cells, nc, c,
                     flags);
}

void launch_internal_compact(const uint32_t* lvl_cells, uint32_t nc, DevCells c,
                             const uint32_t* pos, uint32_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_internal_compact, dim3((nc + TB - 1) / TB), dim3(TB), 0, s, lvl_cells, nc,
                     c, pos, out);
}

void launch_level_pick(const uint32_t* pos, const uint32_t* total, uint32_t nc, const uint32_t* at,
                       uint32_t nl, uint32_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_level_pick, dim3((nl + 63) / 64), dim3(64), 0, s, pos, total, nc, at, nl,
                     out);
}

}  // namespace fmmh
