// prims.hip — HBM-bound primitives for gfx950: device-wide exclusive scan
// (u32) and a stable LSD radix sort of (u64 key, u32 value) pairs.
//
// Radix sort: 8-bit digits, tiles of 2048 elements per 256-thread block
// (4 waves x 8 rounds x 64 lanes). Per pass: a per-tile digit histogram,
// an exclusive scan of the digit-major histogram, then a stable scatter in
// which each wave ranks its elements with an 8-ballot match (lanes with the
// same digit) and per-wave LDS counters. Elements keep their relative order
// within each digit, so the sort is stable: equal keys end in ascending
// input order (ascending particle id for the Morton keys).
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
#ifndef FMMH_RS_I
#define FMMH_RS_I 8
#endif
constexpr int RS_I = FMMH_RS_I;  // keys per thread of one sort block
constexpr int RS_TILE = RS_T * RS_I;
constexpr int RS_WAVE_SPAN = 64 * RS_I;  // consecutive elements per wave

__global__ void __launch_bounds__(RS_T) k_rs_hist(const uint64_t* __restrict__ keys, size_t n,
                                                 int shift, uint32_t* __restrict__ hist,
                                                 uint32_t nblk) {
  __shared__ uint32_t h[4][256];
  const int w = threadIdx.x >> 6;
  // neighbouring tiles share the cache lines of each digit row of the
  // histogram: keep them on one XCD's L2
  const uint32_t tile = xcd_block(blockIdx.x, gridDim.x);
  for (int i = threadIdx.x; i < 4 * 256; i += RS_T) (&h[0][0])[i] = 0;
  __syncthreads();
  const size_t base = (size_t)tile * RS_TILE;
#pragma unroll 4
  for (int j = 0; j < RS_I; j++) {
    size_t e = base + (size_t)j * RS_T + threadIdx.x;
    if (e < n) atomicAdd(&h[w][(keys[e] >> shift) & 255u], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 256; i += RS_T)
    hist[(size_t)i * nblk + tile] = h[0][i] + h[1][i] + h[2][i] + h[3][i];
}

// Scatter of one 8-bit pass. Each wave ranks its keys per digit (ballot
// peer masks, in element order); the block's keys are then reordered by digit
// in LDS (stable: wave order, then element order within a wave) and written
// from LDS in that order, so the global writes of one digit are contiguous
// runs instead of one line per lane.
template <bool VALS>
__global__ void __launch_bounds__(RS_T) k_rs_scatter(const uint64_t* __restrict__ kin,
                                                    const uint32_t* __restrict__ vin,
                                                    uint64_t* __restrict__ kout,
                                                    uint32_t* __restrict__ vout, size_t n,
                                                    int shift, const uint32_t* __restrict__ offs,
                                                    uint32_t nblk) {
  __shared__ uint32_t wcnt[4][256];
  __shared__ uint32_t wpre[4][256];
  __shared__ uint32_t dstart[256];
  __shared__ uint64_t sk[RS_TILE];
  __shared__ uint32_t sv[VALS ? RS_TILE : 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t tile = xcd_block(blockIdx.x, gridDim.x);  // as in k_rs_hist
  for (int i = threadIdx.x; i < 4 * 256; i += RS_T) (&wcnt[0][0])[i] = 0;
  __syncthreads();
  const size_t bbase = (size_t)tile * RS_TILE;
  const size_t wbase = bbase + (size_t)w * RS_WAVE_SPAN;
  const uint64_t lt = lanemask_lt();
  uint64_t kk[RS_I];
  uint32_t vv[RS_I];
  uint32_t lp[RS_I];
  uint32_t dg[RS_I];
#pragma unroll
  for (int j = 0; j < RS_I; j++) {
    size_t e = wbase + (size_t)j * 64 + lane;
    bool valid = e < n;
    uint64_t key = valid ? kin[e] : 0ull;
    uint32_t val = (VALS && valid) ? vin[e] : 0u;
    uint32_t d = (uint32_t)(key >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      uint64_t bb = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bb : ~bb;
    }
    uint32_t rank = (uint32_t)__popcll(peers & lt);
    uint32_t cnt = (uint32_t)__popcll(peers);
    uint32_t before = valid ? wcnt[w][d] : 0u;
    // Every lane of the wave has read its counter before the leaders (lowest
    // lane of each digit group) store the advanced value: LDS operations of
    // one wave execute in program order.
    if (valid && rank == 0) wcnt[w][d] = before + cnt;
    kk[j] = key;
    vv[j] = val;
    lp[j] = before + rank;
    dg[j] = valid ? d : 0xFFFFFFFFu;
  }
  __syncthreads();
  // per digit: offsets of the waves inside the digit's block run, and the
  // block's digit total; then the digits' starts in the block (exclusive scan)
  uint32_t tot = 0;
  {
    const int d = threadIdx.x;  // RS_T == 256 digits
#pragma unroll
    for (int ww = 0; ww < 4; ww++) {
      uint32_t c = wcnt[ww][d];
      wpre[ww][d] = tot;
      tot += c;
    }
    dstart[d] = tot;
  }
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {  // inclusive scan of the digit totals
    const uint32_t add = threadIdx.x >= o ? dstart[threadIdx.x - o] : 0u;
    __syncthreads();
    dstart[threadIdx.x] += add;
    __syncthreads();
  }
  dstart[threadIdx.x] -= tot;  // exclusive
  __syncthreads();
#pragma unroll
  for (int j = 0; j < RS_I; j++) {
    if (dg[j] == 0xFFFFFFFFu) continue;
    const uint32_t lpos = dstart[dg[j]] + wpre[w][dg[j]] + lp[j];
    sk[lpos] = kk[j];
    if (VALS) sv[lpos] = vv[j];
  }
  __syncthreads();
  const uint32_t cnt_blk = (uint32_t)min((size_t)RS_TILE, n - bbase);
  for (uint32_t i = threadIdx.x; i < cnt_blk; i += RS_T) {
    const uint64_t key = sk[i];
    const uint32_t d = (uint32_t)(key >> shift) & 255u;
    const size_t pos = (size_t)offs[(size_t)d * nblk + tile] + (i - dstart[d]);
    kout[pos] = key;
    if (VALS) vout[pos] = sv[i];
  }
}

size_t radix_tmp_words(size_t n) {
  size_t nblk = (n + RS_TILE - 1) / RS_TILE;
  size_t h = 256 * nblk;
  return h + scan_tmp_words(h);
}

// Sorts (keys, vals) by keys[bit_lo, bit_hi) stably. Ping-pongs between the
// given buffer pairs; returns 1 when the result sits in (k_alt, v_alt).
int radix_sort_u64(uint64_t* keys, uint32_t* vals, uint64_t* k_alt, uint32_t* v_alt, size_t n,
                   int bit_lo, int bit_hi, uint32_t* tmp, hipStream_t s) {
  if (n == 0) return 0;
  uint32_t nblk = (uint32_t)((n + RS_TILE - 1) / RS_TILE);
  uint32_t* hist = tmp;
  uint32_t* stmp = tmp + (size_t)256 * nblk;
  int flip = 0;
  for (int shift = bit_lo; shift < bit_hi; shift += 8) {
    const uint64_t* kin = flip ? k_alt : keys;
    const uint32_t* vin = flip ? v_alt : vals;
    uint64_t* kout = flip ? keys : k_alt;
    uint32_t* vout = flip ? vals : v_alt;
    hipLaunchKernelGGL(k_rs_hist, dim3(nblk), dim3(RS_T), 0, s, kin, n, shift, hist, nblk);
    scan_exclusive_u32(hist, hist, (size_t)256 * nblk, stmp, s, nullptr);
    if (vals)
      hipLaunchKernelGGL(k_rs_scatter<true>, dim3(nblk), dim3(RS_T), 0, s, kin, vin, kout, vout, n,
                         shift, hist, nblk);
    else
      hipLaunchKernelGGL(k_rs_scatter<false>, dim3(nblk), dim3(RS_T), 0, s, kin, nullptr, kout,
                         nullptr, n, shift, hist, nblk);
    flip ^= 1;
  }
  return flip;
}

}  // namespace fmmh
