// near.hip — P2P near field on gfx950 (FP64 and mixed forms), fused with the
// final near + far sum and the slot -> particle-id scatter.
//
// Reference: crates/fmm p2p.rs:13-88 (mutual a_i += m_j d/r^3, a_j -= m_i
// d/r^3, pot += m/r) and translate.rs:339-342 (accel[id] += near + far).
// Here each target leaf owns its sums: one wave per work item (a leaf, or a
// slice of a heavy leaf's neighbour list) streams the item's source bodies
// (its own leaf included, self excluded by slot) through LDS in tiles of 64.
// The leaf's targets run as power-of-two chunks of n bodies with 64/n lanes
// per body, each lane on a strided share of every tile, combined by
// xor-shuffles at the end.
// Softening (an extension; 0 = the reference): r^2 += eps^2.
#include "common.h"
#include "internal.h"

namespace fmmh {

constexpr uint32_t P2P_RMAX = 256;  // source ranges staged in LDS per pass

// P2P work items (uint4): x = target leaf | P2P_OWN when the item holds the
// leaf's own bodies as sources, [y, z) = its slice of the leaf's neighbour
// list, w = P2P_DIRECT (the item is the whole leaf: results go straight to
// accel / pot) or the offset of its partial sums (one double4 per target).
constexpr uint32_t P2P_OWN = 0x80000000u;
constexpr uint32_t P2P_LEAF_MASK = 0x7fffffffu;
constexpr uint32_t P2P_DIRECT = 0xffffffffu;

struct TileLoad {
  double4 v;
  uint32_t fill;
};

// Next tile of the packed source stream: advances the (range, offset)
// cursor over up to 64 bodies; the lane's body is loaded (not yet waited on).
__device__ __forceinline__ TileLoad gather_tile(const uint2* rng, uint32_t nr, uint32_t& r,
                                                uint32_t& off, uint32_t lane,
                                                const double4* __restrict__ body, double4 pad) {
  uint32_t fill = 0;
  int slot = -1;
  while (fill < 64 && r < nr) {
    const uint2 g = rng[r];
    const uint32_t take = min(g.y - off, 64u - fill);
    if (lane >= fill && lane < fill + take) slot = (int)(g.x + off + (lane - fill));
    fill += take;
    off += take;
    if (off == g.y) {
      r++;
      off = 0;
    }
  }
  TileLoad t;
  t.v = body[slot >= 0 ? slot : 0];  // in-bounds load, then padded by value
  if (slot < 0) t.v = pad;
  t.fill = fill;
  return t;
}

// The two forms of the pairwise terms. FP64: absolute positions, the
// Newton-refined rsqrt. Mixed: float4 positions relative to the target
// leaf's centre (the subtraction done in FP64), the hardware FP32 rsqrt, and
// each tile's FP32 partial sums added into FP64 accumulators.
template <bool F32> struct P2PForm;
template <> struct P2PForm<false> {
  using T = double;
  using T4 = double4;
  static __device__ __forceinline__ double rsq(double x) { return rsqrt_pos(x); }
  static __device__ __forceinline__ double fmat(double a, double b, double c) {
    return __builtin_fma(a, b, c);
  }
  static __device__ __forceinline__ double4 stage(double4 v, double, double, double) { return v; }
};
template <> struct P2PForm<true> {
  using T = float;
  using T4 = float4;
  static __device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
  static __device__ __forceinline__ float fmat(float a, float b, float c) {
    return __builtin_fmaf(a, b, c);
  }
  static __device__ __forceinline__ float4 stage(double4 v, double cx, double cy, double cz) {
    return make_float4((float)(v.x - cx), (float)(v.y - cy), (float)(v.z - cz), (float)v.w);
  }
};

// Bodies of a 64-entry LDS tile against the lane's target: the lane takes
// entries sub, sub + K, ... (ts = tile + sub), 64 / K of them. K is a
// template argument, so every lane runs the same trip count (no per-lane
// exit test) and the LDS reads take immediate offsets. Padding entries have
// mass 0 far away and contribute exactly 0. SELF: the entry at offset
// self_i (own slot - tile start - sub) is the target itself and is skipped.
template <bool F32, uint32_t K, bool SELF>
__device__ __forceinline__ void p2p_tile(const typename P2PForm<F32>::T4* __restrict__ ts,
                                         typename P2PForm<F32>::T xi, typename P2PForm<F32>::T yi,
                                         typename P2PForm<F32>::T zi,
                                         typename P2PForm<F32>::T soft2, int self_i,
                                         typename P2PForm<F32>::T& ax,
                                         typename P2PForm<F32>::T& ay,
                                         typename P2PForm<F32>::T& az,
                                         typename P2PForm<F32>::T& ph) {
  using F = P2PForm<F32>;
  using T = typename F::T;
  // FP32: 8 reads in flight per lane; FP64 (twice the registers per entry,
  // longer rsqrt): 4
#pragma unroll(F32 ? 8 : 4)
  for (uint32_t i = 0; i < 64 / K; i++) {
    const auto sj = ts[i * K];
    const T dx = sj.x - xi, dy = sj.y - yi, dz = sj.z - zi;
    T r2 = F::fmat(dx, dx, F::fmat(dy, dy, F::fmat(dz, dz, soft2)));
    T mj = sj.w;
    if (SELF) {
      const bool self = (int)(i * K) == self_i;
      r2 = self ? T(1) : r2;
      mj = self ? T(0) : mj;
    }
    // r2 is finite and > 0 here (softening, or distinct bodies, or padding)
    const T inv = F::rsq(r2);
    const T minv = mj * inv;
    const T s3 = minv * inv * inv;
    ax = F::fmat(s3, dx, ax);
    ay = F::fmat(s3, dy, ay);
    az = F::fmat(s3, dz, az);
    ph += minv;
  }
}

// One chunk's pass over a staged tile: n targets (a power of two), 64 / n
// lanes per target, n iterations per lane.
template <bool F32, bool SELF>
__device__ __forceinline__ void p2p_chunk(const typename P2PForm<F32>::T4* __restrict__ tile,
                                          uint32_t n, uint32_t sub, typename P2PForm<F32>::T xi,
                                          typename P2PForm<F32>::T yi,
                                          typename P2PForm<F32>::T zi,
                                          typename P2PForm<F32>::T soft2, int self_i,
                                          typename P2PForm<F32>::T& ax,
                                          typename P2PForm<F32>::T& ay,
                                          typename P2PForm<F32>::T& az,
                                          typename P2PForm<F32>::T& ph) {
  const auto* ts = tile + sub;
  switch (n) {
    case 64: p2p_tile<F32, 1, SELF>(ts, xi, yi, zi, soft2, self_i, ax, ay, az, ph); break;
    case 32: p2p_tile<F32, 2, SELF>(ts, xi, yi, zi, soft2, self_i, ax, ay, az, ph); break;
    case 16: p2p_tile<F32, 4, SELF>(ts, xi, yi, zi, soft2, self_i, ax, ay, az, ph); break;
    case 8: p2p_tile<F32, 8, SELF>(ts, xi, yi, zi, soft2, self_i, ax, ay, az, ph); break;
    case 4: p2p_tile<F32, 16, SELF>(ts, xi, yi, zi, soft2, self_i, ax, ay, az, ph); break;
    case 2: p2p_tile<F32, 32, SELF>(ts, xi, yi, zi, soft2, self_i, ax, ay, az, ph); break;
    default: p2p_tile<F32, 64, SELF>(ts, xi, yi, zi, soft2, self_i, ax, ay, az, ph); break;
  }
}

// A group of r <= 64 targets of one leaf is cut into at most P2P_NQ chunks of
// power-of-two sizes n_q (64 / n_q lanes per target), all run on every
// staged source tile: a tile costs sum(n_q) iterations per lane, so a leaf
// of 37 bodies runs as 32 + 4 + 1 instead of one 64-target chunk with 27
// idle targets. The plan weighs the padding against a fixed cost per chunk
// and tile (its sums and loop set-up), in VALU issues.
constexpr int P2P_NQ = 3;
#ifndef FMMH_P2P_WAVES_F32
#define FMMH_P2P_WAVES_F32 5  // waves per SIMD the mixed-form P2P kernel is built for
#endif

__device__ __forceinline__ uint32_t pow2ceil(uint32_t x) {
  return x <= 1u ? 1u : 1u << (32 - __clz(x - 1u));
}
__device__ __forceinline__ uint32_t pow2floor(uint32_t x) { return 1u << (31 - __clz(x)); }

__device__ __forceinline__ void chunk_plan(uint32_t r, uint32_t (&n)[P2P_NQ]) {
  constexpr uint32_t IT = 14, CH = 12;  // per iteration, per chunk
  n[0] = pow2ceil(r);
  n[1] = n[2] = 0;
  uint32_t best = IT * n[0] + CH;
  const uint32_t a = pow2floor(r), ra = r - a;
  if (ra == 0) return;
  const uint32_t b2 = pow2ceil(ra);
  if (IT * (a + b2) + 2 * CH < best) {
    best = IT * (a + b2) + 2 * CH;
    n[0] = a;
    n[1] = b2;
  }
  const uint32_t b = pow2floor(ra), rb = ra - b;
  if (rb == 0) return;
  const uint32_t c = pow2ceil(rb);
  if (IT * (a + b + c) + 3 * CH < best) {
    n[0] = a;
    n[1] = b;
    n[2] = c;
  }
}

// Compensated FP64 accumulation (Knuth's TwoSum: s + v = t + e exactly, e
// kept in c): the mixed form's final per-body sums.
__device__ __forceinline__ void acc2(double& s, double& c, double v) {
  const double t = s + v;
  const double bp = t - s;
  c += (s - (t - bp)) + (v - bp);
  s = t;
}

// Slot ranges of one segment of an item's source stream into LDS: rng[0] =
// the leaf's own bodies when the stream starts with them (else empty), then
// each directed entry's source leaf (first slot, count), the own leaf's entry
// emptied. The barriers keep the previous pass's readers and this pass's
// readers apart.
__device__ __forceinline__ void stage_ranges(uint2* rng, const uint64_t* __restrict__ skeys,
                                             uint64_t smask, const uint32_t* __restrict__ cfirst,
                                             const uint32_t* __restrict__ ccount, uint32_t ft,
                                             uint32_t nt_all, bool pre, uint32_t seg, uint32_t nr,
                                             uint32_t lane) {
  __syncthreads();
  if (lane == 0) rng[0] = make_uint2(ft, pre ? nt_all : 0u);
  for (uint32_t q = lane; q < nr; q += 64) {
    const uint32_t src = (uint32_t)(skeys[seg + q] & smask);
    uint2 g = make_uint2(cfirst[src], ccount[src]);
    if (g.x == ft) g.y = 0;
    rng[1 + q] = g;
  }
  __syncthreads();
}

// One wave per work item. The targets of the item's leaf are taken in groups
// of 64, each group cut into chunks (chunk_plan); the item's source stream
// (the leaf's own bodies first when the item holds them, each target
// skipping its own slot there; then the neighbour leaves' ranges with the own
// leaf's entry emptied) is packed back to back into 64-body LDS tiles, and
// every chunk runs on every tile. The next tile's global loads are issued
// before the current one is computed. The mixed form (F32) sums each tile in
// FP32 and compensates the final sums (sub-lane partials, then the far field).
#define FMMH_P2P_PARAMS                                                                     \
  const uint4 *__restrict__ items, const uint64_t *__restrict__ skeys, uint64_t smask,       \
      const uint32_t *__restrict__ cfirst, const uint32_t *__restrict__ ccount,               \
      const double *__restrict__ ccx, const double *__restrict__ ccy,                         \
      const double *__restrict__ ccz, const double4 *__restrict__ body, double soft2,         \
      const double4 *__restrict__ far, const uint32_t *__restrict__ order,                    \
      double *__restrict__ accel, double *__restrict__ pot, double4 *__restrict__ partial
#define FMMH_P2P_ARGS \
  items, skeys, smask, cfirst, ccount, ccx, ccy, ccz, body, soft2, far, order, accel, pot, partial

template <bool F32>
__device__ __forceinline__ void p2p_item(FMMH_P2P_PARAMS) {
  using F = P2PForm<F32>;
  using T = typename F::T;
  __shared__ typename F::T4 tile[64];
  __shared__ uint2 rng[P2P_RMAX + 1];
  const uint32_t lane = threadIdx.x;
  const uint4 it = items[xcd_block(blockIdx.x, gridDim.x)];
  const uint32_t leaf = it.x & P2P_LEAF_MASK;
  const bool own = (it.x & P2P_OWN) != 0u;
  const uint32_t ft = cfirst[leaf], nt_all = ccount[leaf];
  const uint32_t jb = it.y, je = it.z;
  // the mixed form works relative to the leaf's centre; FP64 in place
  const double cx = F32 ? ccx[leaf] : 0.0, cy = F32 ? ccy[leaf] : 0.0,
               cz = F32 ? ccz[leaf] : 0.0;
  const double4 pad = make_double4(cx + 1e18, cy + 1e18, cz + 1e18, 0.0);
  const T s2 = (T)soft2;
  const uint32_t nseg = (je - jb + P2P_RMAX - 1) / P2P_RMAX;
  for (uint32_t t0 = 0; t0 < nt_all; t0 += 64) {
    const uint32_t r = min(64u, nt_all - t0);
    uint32_t n[P2P_NQ];
    chunk_plan(r, n);
    // lane state per chunk: target index within the leaf (t0 + base + ti)
    // and the lane's strided share sub of each tile (64 / n lanes per target)
    uint32_t tix[P2P_NQ], sub[P2P_NQ];
    bool tval[P2P_NQ];
    T mx[P2P_NQ], my[P2P_NQ], mz[P2P_NQ];
    double ax[P2P_NQ], ay[P2P_NQ], az[P2P_NQ], ph[P2P_NQ];
    uint32_t base = 0;
#pragma unroll
    for (int q = 0; q < P2P_NQ; q++) {
      const uint32_t nq = n[q] ? n[q] : 1u;
      const uint32_t ti = lane & (nq - 1);
      sub[q] = lane / nq;
      tval[q] = n[q] != 0 && base + ti < r;
      tix[q] = t0 + base + ti;
      const auto me = F::stage(body[ft + (tval[q] ? tix[q] : 0)], cx, cy, cz);
      mx[q] = me.x;
      my[q] = me.y;
      mz[q] = me.z;
      ax[q] = ay[q] = az[q] = ph[q] = 0.0;
      base += n[q];
    }
    const uint32_t npass = max(nseg, own ? 1u : 0u);
    for (uint32_t sg = 0; sg < npass; sg++) {
      const uint32_t seg = jb + sg * P2P_RMAX;
      const uint32_t nr = seg < je ? min(P2P_RMAX, je - seg) : 0u;
      const bool pre = own && sg == 0;  // stream starts with the own leaf
      // an item of one segment stages its ranges once, for every group
      if (nseg > 1 || t0 == 0) stage_ranges(rng, skeys, smask, cfirst, ccount, ft, nt_all, pre,
                                            seg, nr, lane);
      uint32_t rr = 0, off = 0, pos = 0;
      TileLoad cur = gather_tile(rng, nr + 1, rr, off, lane, body, pad);
      while (cur.fill > 0) {
        tile[lane] = F::stage(cur.v, cx, cy, cz);
        const uint32_t fill = cur.fill;
        __syncthreads();
        cur = gather_tile(rng, nr + 1, rr, off, lane, body, pad);
        const bool self_tile = pre && pos < nt_all;
#pragma unroll
        for (int q = 0; q < P2P_NQ; q++) {
          if (n[q] == 0) break;
          const int self_i = (int)tix[q] - (int)pos - (int)sub[q];
          if constexpr (F32) {
            float fx = 0.f, fy = 0.f, fz = 0.f, fp = 0.f;
            if (self_tile)
              p2p_chunk<true, true>(tile, n[q], sub[q], mx[q], my[q], mz[q], s2, self_i, fx, fy,
                                    fz, fp);
            else
              p2p_chunk<true, false>(tile, n[q], sub[q], mx[q], my[q], mz[q], s2, 0, fx, fy, fz,
                                     fp);
            ax[q] += fx;
            ay[q] += fy;
            az[q] += fz;
            ph[q] += fp;
          } else {
            if (self_tile)
              p2p_chunk<false, true>(tile, n[q], sub[q], mx[q], my[q], mz[q], s2, self_i, ax[q],
                                     ay[q], az[q], ph[q]);
            else
              p2p_chunk<false, false>(tile, n[q], sub[q], mx[q], my[q], mz[q], s2, 0, ax[q],
                                      ay[q], az[q], ph[q]);
          }
        }
        pos += fill;
        __syncthreads();
      }
    }
#pragma unroll
    for (int q = 0; q < P2P_NQ; q++) {
      if (n[q] == 0) break;
      double kx = 0.0, ky = 0.0, kz = 0.0, kp = 0.0;
      for (uint32_t o = n[q]; o < 64; o <<= 1) {
        if constexpr (F32) {
          acc2(ax[q], kx, __shfl_xor(ax[q], (int)o, 64));
          acc2(ay[q], ky, __shfl_xor(ay[q], (int)o, 64));
          acc2(az[q], kz, __shfl_xor(az[q], (int)o, 64));
          acc2(ph[q], kp, __shfl_xor(ph[q], (int)o, 64));
          kx += __shfl_xor(kx, (int)o, 64);
          ky += __shfl_xor(ky, (int)o, 64);
          kz += __shfl_xor(kz, (int)o, 64);
          kp += __shfl_xor(kp, (int)o, 64);
        } else {
          ax[q] += __shfl_xor(ax[q], (int)o, 64);
          ay[q] += __shfl_xor(ay[q], (int)o, 64);
          az[q] += __shfl_xor(az[q], (int)o, 64);
          ph[q] += __shfl_xor(ph[q], (int)o, 64);
        }
      }
      if (sub[q] == 0 && tval[q]) {
        const uint32_t si = ft + tix[q];
        if (it.w == P2P_DIRECT) {  // the final near + far sum
          const double4 fv = far[si];
          const uint32_t id = order[si];
          if constexpr (F32) {
            acc2(ax[q], kx, fv.x);
            acc2(ay[q], ky, fv.y);
            acc2(az[q], kz, fv.z);
            acc2(ph[q], kp, fv.w);
          } else {
            ax[q] += fv.x;
            ay[q] += fv.y;
            az[q] += fv.z;
            ph[q] += fv.w;
          }
          accel[3 * (size_t)id] = ax[q] + kx;
          accel[3 * (size_t)id + 1] = ay[q] + ky;
          accel[3 * (size_t)id + 2] = az[q] + kz;
          if (pot) pot[id] = ph[q] + kp;
        } else {
          partial[(size_t)it.w + tix[q]] =
              make_double4(ax[q] + kx, ay[q] + ky, az[q] + kz, ph[q] + kp);
        }
      }
    }
  }
}

// The two forms as kernels of their own, each with the occupancy that
// measured fastest: the mixed form at five waves per SIMD (at most 102 VGPRs:
// the unrolled tile loop keeps fewer reads in flight, and the extra wave
// hides more of the tile staging), FP64 at the compiler's choice.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(FMMH_P2P_WAVES_F32)))
k_p2p_mixed(FMMH_P2P_PARAMS) {
  p2p_item<true>(FMMH_P2P_ARGS);
}

__global__ void __launch_bounds__(64) k_p2p_fp64(FMMH_P2P_PARAMS) {
  p2p_item<false>(FMMH_P2P_ARGS);
}

// ------------------------------------------------------- P2P work plan
// A leaf's P2P work is its target count times its source bodies (own leaf
// plus neighbour leaves). One block per leaf leaves the whole device waiting
// on the few leaves whose neighbourhoods are huge (a large sparse leaf beside
// a dense region), so a leaf whose work exceeds the context's item bound W
// (fmmh_tuning.p2p_item_work, default 2^18) is cut into m = ceil(work / W)
// work items (at most one per source unit).
// Source units are the own leaf (unit 0) and the neighbour ranges in list
// order; unit u with P_u sources before it goes to item
// min(m - 1, P_u / ceil(S / m)) (items may come out empty). Split items write
// partial sums that k_p2p_reduce adds in item order. The cut depends on the
// leaf's own counts only, so a system evaluates the same alone and in an
// ensemble. One wave per leaf in both plan kernels.
constexpr int PLAN_T = 256;

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

struct LeafCut {
  uint32_t leaf, nt, jb, je;
  uint64_t S, m;
};

// Source total and item count of leaf i (all lanes get the result).
__device__ __forceinline__ LeafCut leaf_cut(uint32_t i, const uint32_t* __restrict__ leaf_ids,
                                            const uint32_t* __restrict__ pptr,
                                            const uint64_t* __restrict__ skeys, uint64_t smask,
                                            const uint32_t* __restrict__ ccount, uint64_t W,
                                            uint32_t lane) {
  LeafCut L;
  L.leaf = leaf_ids[i];
  L.nt = ccount[L.leaf];
  L.jb = pptr[L.leaf];
  L.je = pptr[L.leaf + 1];
  uint64_t s = 0;
  for (uint32_t j = L.jb + lane; j < L.je; j += 64) {
    const uint32_t src = (uint32_t)(skeys[j] & smask);
    s += src == L.leaf ? 0u : ccount[src];
  }
  L.S = (uint64_t)L.nt + wave_sum_u64(s);
  const uint64_t work = (uint64_t)L.nt * L.S;
  uint64_t m = work > W ? (work + W - 1) / W : 1;
  const uint64_t units = (uint64_t)(L.je - L.jb) + 1;
  L.m = m > units ? units : m;
  return L;
}

__global__ void __launch_bounds__(PLAN_T) k_p2p_plan_count(const uint32_t* __restrict__ leaf_ids,
                                                          uint32_t nleaves,
                                                          const uint32_t* __restrict__ pptr,
                                                          const uint64_t* __restrict__ skeys,
                                                          uint64_t smask,
                                                          const uint32_t* __restrict__ ccount,
                                                          uint64_t W, uint32_t* __restrict__ nitem,
                                                          uint32_t* __restrict__ npart,
                                                          uint32_t* __restrict__ nsplit) {
  const uint32_t i = blockIdx.x * (PLAN_T / 64) + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  if (i > nleaves) return;
  if (i == nleaves) {  // the scans' closing element
    if (lane == 0) nitem[i] = npart[i] = nsplit[i] = 0u;
    return;
  }
  const LeafCut L = leaf_cut(i, leaf_ids, pptr, skeys, smask, ccount, W, lane);
  if (lane == 0) {
    const uint32_t m = (uint32_t)L.m;
    nitem[i] = m;
    npart[i] = m > 1 ? m * L.nt : 0u;
    nsplit[i] = m > 1 ? 1u : 0u;
  }
}

__global__ void __launch_bounds__(PLAN_T) k_p2p_plan_emit(const uint32_t* __restrict__ leaf_ids,
                                                         uint32_t nleaves,
                                                         const uint32_t* __restrict__ pptr,
                                                         const uint64_t* __restrict__ skeys,
                                                         uint64_t smask,
                                                         const uint32_t* __restrict__ ccount,
                                                         uint64_t W,
                                                         const uint32_t* __restrict__ ioff,
                                                         const uint32_t* __restrict__ poff,
                                                         const uint32_t* __restrict__ roff,
                                                         uint4* __restrict__ items,
                                                         uint4* __restrict__ reds) {
  const uint32_t i = blockIdx.x * (PLAN_T / 64) + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  if (i >= nleaves) return;
  const uint32_t base = ioff[i];
  if (ioff[i + 1] - base == 1) {  // one item (the count pass found m = 1): the whole list
    if (lane == 0) {
      const uint32_t leaf = leaf_ids[i];
      items[base] = make_uint4(leaf | P2P_OWN, pptr[leaf], pptr[leaf + 1], P2P_DIRECT);
    }
    return;
  }
  const LeafCut L = leaf_cut(i, leaf_ids, pptr, skeys, smask, ccount, W, lane);
  const uint32_t m = (uint32_t)L.m;
  const uint32_t pb = poff[i];
  uint32_t* it = (uint32_t*)(items + base);  // [m][4]: leaf|own, begin, end, slot
  for (uint32_t k = lane; k < m; k += 64) {
    it[4 * k] = L.leaf | (k == 0 ? P2P_OWN : 0u);
    it[4 * k + 3] = m > 1 ? pb + k * L.nt : P2P_DIRECT;
  }
  if (lane == 0) {
    it[1] = L.jb;
    if (m > 1) reds[roff[i]] = make_uint4(L.leaf, pb, m, 0u);
  }
  // item boundaries: where the unit's item index steps up
  const uint64_t T = (L.S + L.m - 1) / L.m;
  uint64_t running = L.nt;  // sources before the first neighbour unit
  uint32_t gprev = 0;       // item of unit 0 (the own leaf)
  for (uint32_t c0 = L.jb; c0 < L.je; c0 += 64) {
    const uint32_t j = c0 + lane;
    const bool valid = j < L.je;
    uint64_t cnt = 0;
    if (valid) {
      const uint32_t src = (uint32_t)(skeys[j] & smask);
      cnt = src == L.leaf ? 0u : ccount[src];
    }
    uint64_t incl = cnt;  // inclusive wave scan
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint64_t v = __shfl_up(incl, o, 64);
      if ((int)lane >= o) incl += v;
    }
    const uint64_t P = running + incl - cnt;
    const uint64_t gq = P / T;
    const uint32_t g = valid ? (uint32_t)(gq < L.m - 1 ? gq : L.m - 1) : 0u;
    const uint32_t gl = (uint32_t)__shfl_up(g, 1, 64);
    const uint32_t before = lane == 0 ? gprev : gl;
    if (valid && g > before)
      for (uint32_t k = before + 1; k <= g; k++) {
        it[4 * k + 1] = j;
        it[4 * (k - 1) + 2] = j;
      }
    const uint32_t last = min(63u, L.je - 1 - c0);  // last valid lane of this chunk
    gprev = (uint32_t)__shfl(g, (int)last, 64);
    running += __shfl(incl, (int)last, 64);
  }
  // items after the last unit's item are empty; the last item ends the list
  for (uint32_t k = gprev + lane; k < m; k += 64) {
    if (k > gprev) it[4 * k + 1] = L.je;
    it[4 * k + 2] = L.je;
  }
}

// Split leaves: far field + the items' partial sums in item order, scattered
// to particle-id order.
// COMP (the mixed form): the item sums and the far field are added with
// compensated (TwoSum) accumulation.
template <bool COMP>
__global__ void __launch_bounds__(64) k_p2p_reduce(const uint4* __restrict__ reds,
                                                  const uint32_t* __restrict__ cfirst,
                                                  const uint32_t* __restrict__ ccount,
                                                  const double4* __restrict__ partial,
                                                  const double4* __restrict__ far,
                                                  const uint32_t* __restrict__ order,
                                                  double* __restrict__ accel,
                                                  double* __restrict__ pot) {
  const uint4 r = reds[blockIdx.x];
  const uint32_t ft = cfirst[r.x], nt = ccount[r.x];
  for (uint32_t t = threadIdx.x; t < nt; t += 64) {
    double ax = 0.0, ay = 0.0, az = 0.0, ph = 0.0;
    double kx = 0.0, ky = 0.0, kz = 0.0, kp = 0.0;
    for (uint32_t k = 0; k <= r.z; k++) {  // the items in order, then the far field
      const double4 v = k < r.z ? partial[(size_t)r.y + (size_t)k * nt + t] : far[ft + t];
      if (COMP) {
        acc2(ax, kx, v.x);
        acc2(ay, ky, v.y);
        acc2(az, kz, v.z);
        acc2(ph, kp, v.w);
      } else {
        ax += v.x;
        ay += v.y;
        az += v.z;
        ph += v.w;
      }
    }
    const uint32_t id = order[ft + t];
    accel[3 * (size_t)id] = ax + kx;
    accel[3 * (size_t)id + 1] = ay + ky;
    accel[3 * (size_t)id + 2] = az + kz;
    if (pot) pot[id] = ph + kp;
  }
}

// Work plan of the current P2P lists. nitem / npart / nsplit are nleaves + 1
// words each (scanned in place); the totals come back through the host.
void p2p_plan_count(const uint32_t* leaf_ids, uint32_t nleaves, const uint32_t* pptr,
                    const uint64_t* skeys, int sbits, const uint32_t* ccount, uint64_t item_work,
                    uint32_t* nitem, uint32_t* npart, uint32_t* nsplit, hipStream_t s) {
  if (nleaves == 0) return;
  const uint64_t smask = (1ull << sbits) - 1ull;
  hipLaunchKernelGGL(k_p2p_plan_count, dim3((nleaves + 1 + PLAN_T / 64 - 1) / (PLAN_T / 64)),
                     dim3(PLAN_T), 0, s, leaf_ids,
                     nleaves, pptr, skeys, smask, ccount, item_work, nitem, npart, nsplit);
}

void p2p_plan_emit(const uint32_t* leaf_ids, uint32_t nleaves, const uint32_t* pptr,
                   const uint64_t* skeys, int sbits, const uint32_t* ccount, uint64_t item_work,
                   const uint32_t* ioff, const uint32_t* poff, const uint32_t* roff, uint4* items,
                   uint4* reds, hipStream_t s) {
  if (nleaves == 0) return;
  const uint64_t smask = (1ull << sbits) - 1ull;
  hipLaunchKernelGGL(k_p2p_plan_emit, dim3((nleaves + PLAN_T / 64 - 1) / (PLAN_T / 64)),
                     dim3(PLAN_T), 0, s, leaf_ids,
                     nleaves, pptr, skeys, smask, ccount, item_work, ioff, poff, roff, items, reds);
}

void launch_p2p_f32(const P2PPlan& pl, DevCells c, const double4* body, const uint64_t* skeys,
                    int sbits, double soft2, const double4* far, const uint32_t* order,
                    double* accel, double* pot, hipStream_t s) {
  if (pl.nitems == 0) return;
  const uint64_t smask = (1ull << sbits) - 1ull;
  hipLaunchKernelGGL(k_p2p_mixed, dim3(pl.nitems), dim3(64), 0, s, pl.items, skeys, smask, c.first,
                     c.count, c.cx, c.cy, c.cz, body, soft2, far, order, accel, pot, pl.partial);
}

void launch_p2p_reduce(const P2PPlan& pl, DevCells c, const double4* far, const uint32_t* order,
                       double* accel, double* pot, bool compensated, hipStream_t s) {
  if (pl.nitems == 0 || pl.nsplit == 0) return;
  if (compensated)
    hipLaunchKernelGGL(k_p2p_reduce<true>, dim3(pl.nsplit), dim3(64), 0, s, pl.reds, c.first,
                       c.count, pl.partial, far, order, accel, pot);
  else
    hipLaunchKernelGGL(k_p2p_reduce<false>, dim3(pl.nsplit), dim3(64), 0, s, pl.reds, c.first,
                       c.count, pl.partial, far, order, accel, pot);
}

void launch_p2p(const P2PPlan& pl, DevCells c, const double4* body, const uint64_t* skeys,
                int sbits, double soft2, const double4* far, const uint32_t* order, double* accel,
                double* pot, hipStream_t s) {
  if (pl.nitems == 0) return;
  const uint64_t smask = (1ull << sbits) - 1ull;
  hipLaunchKernelGGL(k_p2p_fp64, dim3(pl.nitems), dim3(64), 0, s, pl.items, skeys, smask, c.first,
                     c.count, c.cx, c.cy, c.cz, body, soft2, far, order, accel, pot, pl.partial);
}

__global__ void __launch_bounds__(256) k_p2p_count(const uint2* __restrict__ pairs, uint32_t np,
                                                  const uint32_t* __restrict__ count,
                                                  unsigned long long* __restrict__ part) {
  __shared__ unsigned long long sh[4];
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  unsigned long long v = 0;
  if (i < np) {
    uint2 pr = pairs[i];
    unsigned long long ca = count[pr.x], cb = count[pr.y];
    v = pr.x == pr.y ? ca * (ca - 1) : 2 * ca * cb;
  }
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ void __launch_bounds__(256) k_sum_u64(const unsigned long long* __restrict__ part,
                                                uint32_t np, unsigned long long* __restrict__ out) {
  __shared__ unsigned long long sh[4];
  unsigned long long v = 0;
  for (uint32_t i = threadIdx.x; i < np; i += 256) v += part[i];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) *out = sh[0] + sh[1] + sh[2] + sh[3];
}

size_t p2p_count_scratch(uint32_t np) { return ((size_t)np + 255) / 256 + 1; }

// out[0] = total; out[1..] scratch of p2p_count_scratch(np) words.
void launch_p2p_count(const uint2* pairs, uint32_t np, const uint32_t* count,
                      unsigned long long* out, hipStream_t s) {
  uint32_t nb = (np + 255) / 256;
  if (nb)
    hipLaunchKernelGGL(k_p2p_count, dim3(nb), dim3(256), 0, s, pairs, np, count, out + 1);
  hipLaunchKernelGGL(k_sum_u64, dim3(1), dim3(256), 0, s, out + 1, nb, out);
}

}  // namespace fmmh
