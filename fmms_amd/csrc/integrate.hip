// integrate.hip — device-resident kick-drift-kick leapfrog (sim.rs:539-594)
// on id-order state: pos/vel/acc [n][3] FP64. The step size lives in device
// memory so a whole step (dt, kick+drift, force evaluation, kick) can be
// replayed from a captured graph without host involvement. Compiled with
// -ffp-contract=off: v += a*dt*0.5 and x += v*dt round exactly as written.
#include "common.h"
#include "internal.h"

namespace fmmh {

constexpr int IT = 256;
constexpr int AMAX_BLOCKS = 1024;

// Per-block maxima of max_d |a_d| (sim.rs:553-556).
__global__ void __launch_bounds__(IT) k_amax_partial(const double* __restrict__ acc, size_t n,
                                                    double* __restrict__ part) {
  __shared__ double sm[IT];
  double m = 0.0;
  for (size_t i = (size_t)blockIdx.x * IT + threadIdx.x; i < n; i += (size_t)gridDim.x * IT) {
    const double* a = acc + 3 * i;
    m = fmax(m, fmax(fabs(a[0]), fmax(fabs(a[1]), fabs(a[2]))));
  }
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = IT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// dt = min(dt_req, eta * sqrt(eps_len / max(1e-9, amax))) (sim.rs:19-20,553-557);
// eta <= 0 keeps dt_req. st[0] = dt of this step, st[1] += dt (time).
__global__ void __launch_bounds__(IT) k_dt(const double* __restrict__ part, int nparts,
                                          double dt_req, double eta, double eps_len,
                                          double* __restrict__ st) {
  __shared__ double sm[IT];
  double m = 1e-9;
  for (int i = threadIdx.x; i < nparts; i += IT) m = fmax(m, part[i]);
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = IT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double dt = dt_req;
    if (eta > 0.0) dt = fmin(dt_req, eta * sqrt(eps_len / sm[0]));
    st[0] = dt;
    st[1] = st[1] + dt;
  }
}

__global__ void k_dt_fixed(double dt, double* __restrict__ st) {
  if (threadIdx.x == 0) {
    st[0] = dt;
    st[1] = st[1] + dt;
  }
}

// v += a dt/2; x += v dt (sim.rs:558-565)
__global__ void __launch_bounds__(IT) k_kick_drift(size_t n, double* __restrict__ pos,
                                                  double* __restrict__ vel,
                                                  const double* __restrict__ acc,
                                                  const double* __restrict__ st) {
  const size_t i = (size_t)blockIdx.x * IT + threadIdx.x;
  if (i >= 3 * n) return;
  const double dt = st[0];
  const double v = vel[i] + acc[i] * dt * 0.5;
  vel[i] = v;
  pos[i] = pos[i] + v * dt;
}

// v += a dt/2 (sim.rs:568-572)
__global__ void __launch_bounds__(IT) k_kick(size_t n, double* __restrict__ vel,
                                            const double* __restrict__ acc,
                                            const double* __restrict__ st) {
  const size_t i = (size_t)blockIdx.x * IT + threadIdx.x;
  if (i >= 3 * n) return;
  vel[i] = vel[i] + acc[i] * st[0] * 0.5;
}

// sim.rs:22-33: non-finite accelerations zeroed, |a| above cap scaled to cap.
__global__ void __launch_bounds__(IT) k_cap_accel(size_t n, double* __restrict__ acc, double cap) {
  const size_t i = (size_t)blockIdx.x * IT + threadIdx.x;
  if (i >= n) return;
  double* a = acc + 3 * i;
  const double s2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
  if (!isfinite(s2)) {
    a[0] = a[1] = a[2] = 0.0;
  } else if (s2 > cap * cap) {
    const double s = cap / sqrt(s2);
    a[0] = a[0] * s;
    a[1] = a[1] * s;
    a[2] = a[2] * s;
  }
}

// Drift monitor of frozen steps: flag = 1 when a body (slot order, positions
// just gathered) lies farther from its leaf's frozen centre than
// rho + tol * max(rho, side / 2), side = the leaf's octant edge at build.
__global__ void __launch_bounds__(IT) k_drift_check(uint32_t n, const uint32_t* __restrict__ leaf_of_slot,
                                                   DevCells c, DevBodies b,
                                                   const BBox* __restrict__ bb,
                                                   const uint64_t* __restrict__ keys_sorted,
                                                   int vlev, double tol,
                                                   uint32_t* __restrict__ flag) {
  const uint32_t i = blockIdx.x * IT + threadIdx.x;
  if (i >= n) return;
  const uint32_t L = leaf_of_slot[i];
  const double dx = b.x[i] - c.cx[L], dy = b.y[i] - c.cy[L], dz = b.z[i] - c.cz[L];
  const double rho = c.rho[L];
  // the leaf's octant edge in its own system's box (ensembles: V virtual levels)
  const uint32_t sys = vlev ? (uint32_t)(keys_sorted[i] >> (63 - 3 * vlev)) : 0u;
  const double side = ldexp(bb[sys].side, -((int)c.level[L] - vlev));
  const double lim = rho + tol * fmax(rho, 0.5 * side);
  if (dx * dx + dy * dy + dz * dz > lim * lim) flag[0] = 1u;
}

__global__ void k_fill_u64(uint64_t* __restrict__ p, size_t words, uint64_t v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

size_t dt_scratch_words() { return AMAX_BLOCKS; }

void launch_step_dt(const double* acc, size_t n, double dt_req, double eta, double eps_len,
                    double* part, double* st, hipStream_t s) {
  if (eta > 0.0) {
    const int nb = (int)std::min<size_t>(AMAX_BLOCKS, (n + IT - 1) / IT);
    hipLaunchKernelGGL(k_amax_partial, dim3(nb), dim3(IT), 0, s, acc, n, part);
    hipLaunchKernelGGL(k_dt, dim3(1), dim3(IT), 0, s, part, nb, dt_req, eta, eps_len, st);
  } else {
    hipLaunchKernelGGL(k_dt_fixed, dim3(1), dim3(64), 0, s, dt_req, st);
  }
}

void launch_kick_drift(size_t n, double* pos, double* vel, const double* acc, const double* st,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_kick_drift, dim3((3 * n + IT - 1) / IT), dim3(IT), 0, s, n, pos, vel, acc,
                     st);
}

void launch_kick(size_t n, double* vel, const double* acc, const double* st, hipStream_t s) {
  hipLaunchKernelGGL(k_kick, dim3((3 * n + IT - 1) / IT), dim3(IT), 0, s, n, vel, acc, st);
}

void launch_cap_accel(size_t n, double* acc, double cap, hipStream_t s) {
  hipLaunchKernelGGL(k_cap_accel, dim3((n + IT - 1) / IT), dim3(IT), 0, s, n, acc, cap);
}

void launch_drift_check(uint32_t n, const uint32_t* leaf_of_slot, DevCells c, DevBodies b,
                        const BBox* bb, const uint64_t* keys_sorted, int vlev, double tol,
                        uint32_t* flag, hipStream_t s) {
  hipLaunchKernelGGL(k_drift_check, dim3((n + IT - 1) / IT), dim3(IT), 0, s, n, leaf_of_slot, c, b,
                     bb, keys_sorted, vlev, tol, flag);
}

void launch_fill_u64(void* p, size_t words, uint64_t v, hipStream_t s) {
  if (words == 0) return;
  const size_t nb = std::min<size_t>(4096, (words + IT - 1) / IT);
  hipLaunchKernelGGL(k_fill_u64, dim3(nb), dim3(IT), 0, s, (uint64_t*)p, words, v);
}

}  // namespace fmmh
