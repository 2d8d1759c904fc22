// prims.hip — HBM-bound primitives for gfx950: device-wide exclusive scan
// (u32) and a stable LSD radix sort of (u64 key, u32 value) pairs.
//
// Radix sort: digits of up to 8 bits, as even as the passes allow (19 target
// bits as 7 + 7 + 5), tiles of 2048 elements per 256-thread block (4 waves x
// 8 rounds x 64 lanes). Per pass: a per-tile digit histogram, an exclusive
// scan of the digit-major histogram, then a stable scatter in which each wave
// ranks its elements with a w-ballot match (lanes with the same digit) and
// per-wave LDS counters. Elements keep their relative order within each
// digit, so the sort is stable: equal keys end in ascending input order
// (ascending particle id for the Morton keys). Wider digits (10 bits, two
// passes for 19 bits) measured slower: a 4096-key tile then writes runs of
// about 4 keys per digit, and the scattered 32-byte writes cost more than
// the pass they save.
#include <algorithm>

#include "common.h"
#include "internal.h"

namespace fmmh {

constexpr int SCAN_T = 256;
constexpr int SCAN_I = 16;
constexpr int SCAN_TILE = SCAN_T * SCAN_I;

// Block-wide exclusive scan of one value per thread (256 threads).
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* total, uint32_t* sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t wt;
  uint32_t x = wave_excl_scan(v, &wt);
  if (lane == 0) sh[w] = wt;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < SCAN_T / 64; i++) {
    uint32_t s = sh[i];
    if (i < w) off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return x + off;
}

__global__ void __launch_bounds__(SCAN_T) k_scan_tiles(const uint32_t* __restrict__ in,
                                                      uint32_t* __restrict__ out, size_t n,
                                                      uint32_t* __restrict__ tile_sums) {
  __shared__ uint32_t tile[SCAN_TILE];
  __shared__ uint32_t sh[SCAN_T / 64];
  const size_t base = (size_t)blockIdx.x * SCAN_TILE;
#pragma unroll
  for (int j = 0; j < SCAN_I; j++) {
    size_t e = base + (size_t)j * SCAN_T + threadIdx.x;
    tile[j * SCAN_T + threadIdx.x] = e < n ? in[e] : 0u;
  }
  __syncthreads();
  uint32_t v[SCAN_I];
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < SCAN_I; j++) {
    v[j] = s;
    s += tile[threadIdx.x * SCAN_I + j];
  }
  uint32_t total;
  uint32_t off = block_excl_scan(s, &total, sh);
#pragma unroll
  for (int j = 0; j < SCAN_I; j++) tile[threadIdx.x * SCAN_I + j] = v[j] + off;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < SCAN_I; j++) {
    size_t e = base + (size_t)j * SCAN_T + threadIdx.x;
    if (e < n) out[e] = tile[j * SCAN_T + threadIdx.x];
  }
  // one tile: its total is the grand total (tile_sums[ntiles = 1])
  if (threadIdx.x == 0) tile_sums[gridDim.x == 1 ? 1 : blockIdx.x] = total;
}

// One block scans the tile sums in place; tile_sums[ntiles] = grand total.
__global__ void __launch_bounds__(SCAN_T) k_scan_sums(uint32_t* __restrict__ sums, size_t ntiles) {
  __shared__ uint32_t sh[SCAN_T / 64];
  uint32_t carry = 0;
  for (size_t b = 0; b < ntiles; b += SCAN_T) {
    size_t e = b + threadIdx.x;
    uint32_t v = e < ntiles ? sums[e] : 0u;
    uint32_t total;
    uint32_t x = block_excl_scan(v, &total, sh);
    if (e < ntiles) sums[e] = x + carry;
    carry += total;
  }
  if (threadIdx.x == 0) sums[ntiles] = carry;
}

__global__ void __launch_bounds__(SCAN_T) k_scan_add(uint32_t* __restrict__ out, size_t n,
                                                    const uint32_t* __restrict__ sums) {
  const uint32_t add = sums[blockIdx.x];
  const size_t base = (size_t)blockIdx.x * SCAN_TILE;
  if (add == 0) return;
#pragma unroll
  for (int j = 0; j < SCAN_I; j++) {
    size_t e = base + (size_t)j * SCAN_T + threadIdx.x;
    if (e < n) out[e] += add;
  }
}

// Up to SCAN_T tiles: each block adds the totals of the tiles before it
// (one per thread, then a block sum), so no separate pass scans the tile
// totals; the last block stores the grand total in sums[ntiles].
__global__ void __launch_bounds__(SCAN_T) k_scan_add_small(uint32_t* __restrict__ out, size_t n,
                                                          uint32_t* __restrict__ sums,
                                                          uint32_t ntiles) {
  __shared__ uint32_t sh[SCAN_T / 64];
  const uint32_t v = threadIdx.x < blockIdx.x ? sums[threadIdx.x] : 0u;
  uint32_t add;
  block_excl_scan(v, &add, sh);
  if (blockIdx.x == ntiles - 1 && threadIdx.x == 0) sums[ntiles] = add + sums[ntiles - 1];
  if (add == 0) return;
  const size_t base = (size_t)blockIdx.x * SCAN_TILE;
#pragma unroll
  for (int j = 0; j < SCAN_I; j++) {
    size_t e = base + (size_t)j * SCAN_T + threadIdx.x;
    if (e < n) out[e] += add;
  }
}

size_t scan_tmp_words(size_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE + 1; }

// Exclusive scan; in may equal out. tmp needs scan_tmp_words(n) words; the
// grand total lands in tmp[ntiles] (device) and *total_dev points at it.
void scan_exclusive_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* tmp,
                        hipStream_t s, uint32_t** total_dev) {
  size_t ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (ntiles == 0) {
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_T), 0, s, tmp, (size_t)0);
  } else {
    hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(SCAN_T), 0, s, in, out, n, tmp);
    if (ntiles > (size_t)SCAN_T) {
      hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_T), 0, s, tmp, ntiles);
      hipLaunchKernelGGL(k_scan_add, dim3(ntiles), dim3(SCAN_T), 0, s, out, n, tmp);
    } else if (ntiles > 1) {
      hipLaunchKernelGGL(k_scan_add_small, dim3(ntiles), dim3(SCAN_T), 0, s, out, n, tmp,
                         (uint32_t)ntiles);
    }
  }
  if (total_dev) *total_dev = tmp + ntiles;
}

// ------------------------------------------------------------ radix sort
constexpr int RS_T = 256;
constexpr int RS_I = 8;  // keys per thread of one sort block
constexpr int RS_TILE = RS_T * RS_I;
constexpr int RS_WAVE_SPAN = 64 * RS_I;  // consecutive elements per wave
constexpr int RS_MAXW = 8;               // digit bits per pass, at most
constexpr int RS_NB = 1 << RS_MAXW;

// Key sources of a pass: a key array, or (first pass of a list sort) the
// directed entries of undirected pairs: entry 2i = (y << sbits | x), entry
// 2i+1 = (x << sbits | y) of pair i = (x, y); with `sentinel`, the second
// entry of a self pair is all ones in [0, 2 sbits) and sorts past every
// real entry.
struct KeysIn {
  const uint64_t* k;
  __device__ __forceinline__ uint64_t operator()(size_t e) const { return k[e]; }
};
struct DirectedIn {
  const uint2* pairs;
  int sbits;
  bool sentinel;
  __device__ __forceinline__ uint64_t operator()(size_t e) const {
    const uint2 pr = pairs[e >> 1];
    if ((e & 1) == 0) return ((uint64_t)pr.y << sbits) | pr.x;
    if (sentinel && pr.x == pr.y) return (1ull << (2 * sbits)) - 1ull;
    return ((uint64_t)pr.x << sbits) | pr.y;
  }
};

// Lanes of the wave holding the same w-bit digit d (w ballots).
__device__ __forceinline__ uint64_t digit_peers(uint32_t d, bool valid, int w) {
  uint64_t peers = __ballot(valid);
  for (int b = 0; b < w; b++) {
    const uint64_t bb = __ballot((d >> b) & 1u);
    peers &= ((d >> b) & 1u) ? bb : ~bb;
  }
  return peers;
}

// Per-tile digit histogram; each wave adds one count per distinct digit of
// its 64 keys (the lowest lane of each peer group), so keys that are already
// grouped do not serialise on one LDS counter.
template <class KS>
__global__ void __launch_bounds__(RS_T) k_rs_hist(KS ks, size_t n, int shift, int w,
                                                 uint32_t* __restrict__ hist, uint32_t nblk) {
  __shared__ uint32_t h[RS_T / 64][RS_NB];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // neighbouring tiles share the cache lines of each digit row of the
  // histogram: keep them on one XCD's L2
  const uint32_t tile = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t nb = 1u << w;
  for (int i = threadIdx.x; i < (RS_T / 64) * RS_NB; i += RS_T) (&h[0][0])[i] = 0;
  __syncthreads();
  const size_t wbase = (size_t)tile * RS_TILE + (size_t)wv * RS_WAVE_SPAN;
  const uint64_t lt = lanemask_lt();
#pragma unroll 4
  for (int j = 0; j < RS_I; j++) {
    const size_t e = wbase + (size_t)j * 64 + lane;
    const bool valid = e < n;
    const uint32_t d = valid ? (uint32_t)(ks(e) >> shift) & (nb - 1u) : 0u;
    const uint64_t peers = digit_peers(d, valid, w);
    if (valid && (peers & lt) == 0) atomicAdd(&h[wv][d], (uint32_t)__popcll(peers));
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < nb; d += RS_T)
    hist[(size_t)d * nblk + tile] = h[0][d] + h[1][d] + h[2][d] + h[3][d];
}

// Scatter of one pass. Each wave ranks its keys per digit (ballot peer masks,
// in element order); the block's keys are then reordered by digit in LDS
// (stable: wave order, then element order within a wave) and written from
// LDS in that order, so the global writes of one digit are contiguous runs
// instead of one line per lane.
template <class KS, bool VALS>
__global__ void __launch_bounds__(RS_T) k_rs_scatter(KS ks, const uint32_t* __restrict__ vin,
                                                    uint64_t* __restrict__ kout,
                                                    uint32_t* __restrict__ vout, size_t n,
                                                    int shift, int w,
                                                    const uint32_t* __restrict__ offs,
                                                    uint32_t nblk) {
  __shared__ uint32_t wcnt[RS_T / 64][RS_NB];  // per wave and digit: count, then offset
  __shared__ uint32_t dstart[RS_NB];            // digit start in the block, then global base
  __shared__ uint64_t sk[RS_TILE];
  __shared__ uint32_t sv[VALS ? RS_TILE : 1];
  __shared__ uint32_t shs[RS_T / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t tile = xcd_block(blockIdx.x, gridDim.x);  // as in k_rs_hist
  const uint32_t nb = 1u << w;
  for (int i = threadIdx.x; i < (RS_T / 64) * RS_NB; i += RS_T) (&wcnt[0][0])[i] = 0;
  __syncthreads();
  const size_t bbase = (size_t)tile * RS_TILE;
  const size_t wbase = bbase + (size_t)wv * RS_WAVE_SPAN;
  const uint64_t lt = lanemask_lt();
  uint64_t kk[RS_I];
  uint32_t vv[RS_I];
  uint32_t lp[RS_I];
  uint32_t dg[RS_I];
#pragma unroll
  for (int j = 0; j < RS_I; j++) {
    const size_t e = wbase + (size_t)j * 64 + lane;
    const bool valid = e < n;
    const uint64_t key = valid ? ks(e) : 0ull;
    const uint32_t val = (VALS && valid) ? vin[e] : 0u;
    const uint32_t d = (uint32_t)(key >> shift) & (nb - 1u);
    const uint64_t peers = digit_peers(d, valid, w);
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    const uint32_t cnt = (uint32_t)__popcll(peers);
    const uint32_t before = valid ? wcnt[wv][d] : 0u;
    // Every lane of the wave has read its counter before the leaders (lowest
    // lane of each digit group) store the advanced value: LDS operations of
    // one wave execute in program order.
    if (valid && rank == 0) wcnt[wv][d] = before + cnt;
    kk[j] = key;
    vv[j] = val;
    lp[j] = before + rank;
    dg[j] = valid ? d : 0xFFFFFFFFu;
  }
  __syncthreads();
  // per digit: the waves' offsets inside the digit's run (in place) and the
  // digit's start in the block (exclusive scan over the digits, 256 at a time)
  uint32_t carry = 0;
  for (int q = 0; q < RS_NB / RS_T; q++) {
    const uint32_t d = q * RS_T + threadIdx.x;
    uint32_t tot = 0;
    if (d < nb) {
#pragma unroll
      for (int ww = 0; ww < RS_T / 64; ww++) {
        const uint32_t c = wcnt[ww][d];
        wcnt[ww][d] = tot;
        tot += c;
      }
    }
    uint32_t all;
    const uint32_t x = block_excl_scan(tot, &all, shs);
    if (d < nb) dstart[d] = carry + x;
    carry += all;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < RS_I; j++) {
    if (dg[j] == 0xFFFFFFFFu) continue;
    const uint32_t lpos = dstart[dg[j]] + wcnt[wv][dg[j]] + lp[j];
    sk[lpos] = kk[j];
    if (VALS) sv[lpos] = vv[j];
  }
  __syncthreads();
  // global position of block slot i with digit d: offs[d][tile] + i - dstart[d]
  for (uint32_t d = threadIdx.x; d < nb; d += RS_T)
    dstart[d] = offs[(size_t)d * nblk + tile] - dstart[d];
  __syncthreads();
  const uint32_t cnt_blk = (uint32_t)min((size_t)RS_TILE, n - bbase);
  for (uint32_t i = threadIdx.x; i < cnt_blk; i += RS_T) {
    const uint64_t key = sk[i];
    const uint32_t d = (uint32_t)(key >> shift) & (nb - 1u);
    const size_t pos = (size_t)(uint32_t)(dstart[d] + i);
    kout[pos] = key;
    if (VALS) vout[pos] = sv[i];
  }
}

size_t radix_tmp_words(size_t n) {
  size_t nblk = (n + RS_TILE - 1) / RS_TILE;
  size_t h = (size_t)RS_NB * nblk;
  return h + scan_tmp_words(h);
}

// One LSD pass of w bits at `shift`: histogram, scan, scatter.
template <class KS>
static void radix_pass(KS ks, const uint32_t* vin, uint64_t* kout, uint32_t* vout, size_t n,
                       int shift, int w, uint32_t* tmp, hipStream_t s) {
  const uint32_t nblk = (uint32_t)((n + RS_TILE - 1) / RS_TILE);
  uint32_t* hist = tmp;
  const size_t nh = ((size_t)1 << w) * nblk;
  uint32_t* stmp = tmp + (size_t)RS_NB * nblk;
  hipLaunchKernelGGL(k_rs_hist<KS>, dim3(nblk), dim3(RS_T), 0, s, ks, n, shift, w, hist, nblk);
  scan_exclusive_u32(hist, hist, nh, stmp, s, nullptr);
  if (vin)
    hipLaunchKernelGGL((k_rs_scatter<KS, true>), dim3(nblk), dim3(RS_T), 0, s, ks, vin, kout, vout,
                       n, shift, w, hist, nblk);
  else
    hipLaunchKernelGGL((k_rs_scatter<KS, false>), dim3(nblk), dim3(RS_T), 0, s, ks, nullptr, kout,
                       nullptr, n, shift, w, hist, nblk);
}

// Digit width of the passes over `bits` bits: as few passes as RS_MAXW
// allows, their widths as even as possible.
static int radix_width(int bits) {
  const int passes = (bits + RS_MAXW - 1) / RS_MAXW;
  return (bits + passes - 1) / passes;
}

// Sorts (keys, vals) by keys[bit_lo, bit_hi) stably. Ping-pongs between the
// given buffer pairs; returns 1 when the result sits in (k_alt, v_alt).
int radix_sort_u64(uint64_t* keys, uint32_t* vals, uint64_t* k_alt, uint32_t* v_alt, size_t n,
                   int bit_lo, int bit_hi, uint32_t* tmp, hipStream_t s) {
  if (n == 0 || bit_hi <= bit_lo) return 0;
  const int w = radix_width(bit_hi - bit_lo);
  int flip = 0;
  for (int shift = bit_lo; shift < bit_hi; shift += w) {
    const int wp = std::min(w, bit_hi - shift);
    const uint64_t* kin = flip ? k_alt : keys;
    const uint32_t* vin = vals ? (flip ? v_alt : vals) : nullptr;
    uint64_t* kout = flip ? keys : k_alt;
    uint32_t* vout = flip ? vals : v_alt;
    radix_pass(KeysIn{kin}, vin, kout, vout, n, shift, wp, tmp, s);
    flip ^= 1;
  }
  return flip;
}

// The directed entries of np pairs (DirectedIn), sorted stably by target
// (bits [sbits, 2 sbits)); the first pass reads the pairs themselves. keys
// and k_alt hold 2 np entries each; returns 1 when the result sits in k_alt.
int radix_sort_directed(const uint2* pairs, uint32_t np, int sbits, bool sentinel, uint64_t* keys,
                        uint64_t* k_alt, uint32_t* tmp, hipStream_t s) {
  if (np == 0) return 0;
  const size_t n = (size_t)2 * np;
  const int w = radix_width(sbits);
  int flip = 1;  // the first pass writes keys; later passes ping-pong
  for (int shift = sbits; shift < 2 * sbits; shift += w) {
    const int wp = std::min(w, 2 * sbits - shift);
    uint64_t* kout = flip ? keys : k_alt;
    if (shift == sbits)
      radix_pass(DirectedIn{pairs, sbits, sentinel}, nullptr, kout, nullptr, n, shift, wp, tmp, s);
    else
      radix_pass(KeysIn{flip ? k_alt : keys}, nullptr, kout, nullptr, n, shift, wp, tmp, s);
    flip ^= 1;
  }
  return flip;
}

}  // namespace fmmh
