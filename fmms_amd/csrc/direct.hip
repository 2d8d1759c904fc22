// direct.hip — FP64 direct sum at sampled targets (the checker of the force
// error: direct.rs:6-31 semantics with optional Plummer softening, the
// sampled error of crates/bench/src/main.rs:34-78), and its allocation-free
// entry point.
//
// Each target's sources are cut into nsplit contiguous ranges, one 256-thread
// block per (target, range): a few sampled targets still fill the device.
// Every block writes its five partial sums (ax, ay, az, pot, f) and a second
// kernel adds them in range order, so the result does not depend on timing.
#include <cmath>
#include <mutex>

#include "internal.h"

namespace fmmh {

constexpr int DT = 256;
// Scratch of the allocation-free entry point: partial sums, 5 doubles per
// (target, range). 8 MiB per device.
constexpr size_t DIRECT_SCRATCH_WORDS = (size_t)1 << 20;
static_assert(5 * (size_t)FMMH_DIRECT_NOALLOC_MAX_TARGETS <= DIRECT_SCRATCH_WORDS,
              "one range per target must fit the scratch");

__global__ void __launch_bounds__(DT) k_direct_part(const double* __restrict__ pos,
                                                   const double* __restrict__ mass, size_t n,
                                                   double soft2, const uint32_t* __restrict__ ids,
                                                   uint32_t nsplit, size_t chunk,
                                                   double* __restrict__ part) {
  __shared__ double sh[5][DT];
  const uint32_t t = blockIdx.y, k = blockIdx.x;
  const uint32_t id = ids[t];
  double ax = 0, ay = 0, az = 0, ph = 0, f = 0;
  if (id < n) {  // ids are device data: an id out of range gives NaN sums
    const double xi = pos[3 * (size_t)id], yi = pos[3 * (size_t)id + 1], zi = pos[3 * (size_t)id + 2];
    const size_t j0 = (size_t)k * chunk, j1 = j0 + chunk < n ? j0 + chunk : n;
    for (size_t j = j0 + threadIdx.x; j < j1; j += DT) {
      if (j == id) continue;
      const double dx = pos[3 * j] - xi, dy = pos[3 * j + 1] - yi, dz = pos[3 * j + 2] - zi;
      const double r2 = dx * dx + dy * dy + dz * dz + soft2;
      const double inv = 1.0 / sqrt(r2);
      const double inv2 = inv * inv;
      const double mj = mass[j];
      const double s3 = mj * inv2 * inv;
      ax += s3 * dx;
      ay += s3 * dy;
      az += s3 * dz;
      ph += mj * inv;
      f += mj * inv2;
    }
  } else {
    ax = ay = az = ph = f = __builtin_nan("");
  }
  sh[0][threadIdx.x] = ax;
  sh[1][threadIdx.x] = ay;
  sh[2][threadIdx.x] = az;
  sh[3][threadIdx.x] = ph;
  sh[4][threadIdx.x] = f;
  __syncthreads();
  for (int st = DT / 2; st > 0; st >>= 1) {  // fixed-order tree reduction
    if (threadIdx.x < st)
      for (int q = 0; q < 5; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < 5) part[((size_t)t * nsplit + k) * 5 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void __launch_bounds__(DT) k_direct_reduce(const double* __restrict__ part, uint32_t nt,
                                                     uint32_t nsplit, double* __restrict__ acc,
                                                     double* __restrict__ pot,
                                                     double* __restrict__ f) {
  const uint32_t t = blockIdx.x * DT + threadIdx.x;
  if (t >= nt) return;
  double s[5] = {0, 0, 0, 0, 0};
  const double* pt = part + (size_t)t * nsplit * 5;
  for (uint32_t k = 0; k < nsplit; k++)
    for (int q = 0; q < 5; q++) s[q] += pt[5 * k + q];
  acc[3 * (size_t)t] = s[0];
  acc[3 * (size_t)t + 1] = s[1];
  acc[3 * (size_t)t + 2] = s[2];
  if (pot) pot[t] = s[3];
  if (f) f[t] = s[4];
}

uint32_t direct_split(size_t n, size_t nt) {
  // about 4096 blocks in all (16 per CU), each block at least 8192 sources
  size_t k = (4096 + nt - 1) / nt;
  k = std::min<size_t>(k, std::max<size_t>(1, n / 8192));
  k = std::min<size_t>(k, 256);
  k = std::min<size_t>(k, std::max<size_t>(1, DIRECT_SCRATCH_WORDS / (5 * nt)));
  return (uint32_t)std::max<size_t>(k, 1);
}

void launch_direct(const double* pos, const double* mass, size_t n, double soft2,
                   const uint32_t* ids, uint32_t nt, uint32_t nsplit, double* part, double* acc,
                   double* pot, double* f, hipStream_t s) {
  if (nt == 0) return;
  const size_t chunk = (n + nsplit - 1) / nsplit;
  hipLaunchKernelGGL(k_direct_part, dim3(nsplit, nt), dim3(DT), 0, s, pos, mass, n, soft2, ids,
                     nsplit, chunk, part);
  hipLaunchKernelGGL(k_direct_reduce, dim3((nt + DT - 1) / DT), dim3(DT), 0, s, part, nt, nsplit,
                     acc, pot, f);
}

}  // namespace fmmh

using namespace fmmh;

// Partial sums of fmmh_direct_sampled_noalloc. In host code this name is a
// stand-in, not a device address: kernels get the address that
// hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) double scratch[DIRECT_SCRATCH_WORDS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// The current device's copy of `scratch`, looked up once per device.
int scratch_of_current_device(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return FMMH_E_HIP;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return FMMH_E_HIP;
    g_scratch[dev] = (double*)p;
  }
  *out = g_scratch[dev];
  return FMMH_OK;
}
}  // namespace

extern "C" int fmmh_direct_sampled_noalloc(size_t n, const double* d_pos, const double* d_mass,
                                           double softening, size_t nt, const uint32_t* d_idx,
                                           double* d_acc, double* d_pot, double* d_f,
                                           void* stream) {
  if (n == 0 || nt == 0) return FMMH_E_EMPTY;
  if (!d_pos || !d_mass || !d_idx || !d_acc) return FMMH_E_PARAM;
  if (nt > FMMH_DIRECT_NOALLOC_MAX_TARGETS || n >= 0xFFFFFFFFull) return FMMH_E_PARAM;
  if (!(softening >= 0.0) || !std::isfinite(softening)) return FMMH_E_PARAM;
  double* part = nullptr;
  if (int rc = scratch_of_current_device(&part)) return rc;
  const uint32_t nsplit = direct_split(n, nt);
  launch_direct(d_pos, d_mass, n, softening * softening, d_idx, (uint32_t)nt, nsplit, part, d_acc,
                d_pot, d_f, (hipStream_t)stream);
  return hipGetLastError() == hipSuccess ? FMMH_OK : FMMH_E_HIP;
}
