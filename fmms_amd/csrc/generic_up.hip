// generic_up.hip — P2M and M2M for any expansion order (runtime p), used
// when p exceeds the register-resident templated kernels (compiled
// -ffp-contract=off). Results are bit-identical to the reference
// (moments.rs:30-90): the same recursions and the same summation order,
// accumulating straight into the (zeroed) moments array, each entry owned by
// one thread.
#include "common.h"
#include "internal.h"

namespace fmmh {

// Column m (both signs) of Upsilon(r) into tab (harmonic.rs:141-173).
__device__ void upsilon_column_gen(int p, double x, double y, double z, int m, cplx* tab) {
  const double r2 = x * x + y * y + z * z;
  cplx u = cplx{1.0, 0.0};
  for (int j = 1; j <= m; j++) u = cmul(cscale(cplx{x, y}, 1.0 / (double)(2 * j)), u);
  const double sgn = (m % 2 == 0) ? 1.0 : -1.0;
  cplx u1 = u, u2 = cz();
  for (int n = m; n <= p; n++) {
    cplx un;
    if (n == m) {
      un = u;
    } else {
      const double a = (double)(2 * n - 1) * z;
      const double c = (double)(n * n - m * m);
      const cplx p2 = (n >= m + 2) ? u2 : cz();
      un = cscale(csub(cscale(u1, a), cscale(p2, r2)), 1.0 / c);
      u2 = u1;
      u1 = un;
    }
    tab[hidx(n, m)] = un;
    if (m > 0) tab[hidx(n, -m)] = cscale(cconj(un), sgn);
  }
}

// P2M, one thread per (leaf, m): the column of each body's Upsilon is
// recomputed in order and added to the entries (n, +-m) the thread owns.
__global__ void __launch_bounds__(256) k_p2m_gen(int p, const uint32_t* __restrict__ leaf_ids,
                                                uint32_t nleaves, DevCells c, DevBodies b,
                                                cplx* __restrict__ moments) {
  const size_t D = (size_t)hdim(p);
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t li = t / (p + 1);
  const int m = (int)(t % (p + 1));
  if (li >= nleaves) return;
  const uint32_t id = leaf_ids[li];
  const double zx = c.cx[id], zy = c.cy[id], zz = c.cz[id];
  cplx* out = moments + id * D;
  const double sgn = (m % 2 == 0) ? 1.0 : -1.0;
  const uint32_t f = c.first[id], e = f + c.count[id];
  for (uint32_t k = f; k < e; k++) {
    const double x = b.x[k] - zx, y = b.y[k] - zy, z = b.z[k] - zz;
    const double r2 = x * x + y * y + z * z;
    const double ms = b.m[k];
    cplx u = cplx{1.0, 0.0};
    for (int j = 1; j <= m; j++) u = cmul(cscale(cplx{x, y}, 1.0 / (double)(2 * j)), u);
    cplx u1 = u, u2 = cz();
    for (int n = m; n <= p; n++) {
      cplx un;
      if (n == m) {
        un = u;
      } else {
        const double a = (double)(2 * n - 1) * z;
        const double cc = (double)(n * n - m * m);
        const cplx p2 = (n >= m + 2) ? u2 : cz();
        un = cscale(csub(cscale(u1, a), cscale(p2, r2)), 1.0 / cc);
        u2 = u1;
        u1 = un;
      }
      out[hidx(n, m)] = cadd(out[hidx(n, m)], cscale(un, ms));
      if (m > 0) out[hidx(n, -m)] = cadd(out[hidx(n, -m)], cscale(cscale(cconj(un), sgn), ms));
    }
  }
}

// M2M, one block per cell of a level: children in octant order, the child's
// Upsilon table staged in LDS, each output entry summed in the reference's
// (k, l) order and added to the zeroed moments.
__global__ void __launch_bounds__(256) k_m2m_gen(int p, const uint32_t* __restrict__ cells,
                                                DevCells c, cplx* __restrict__ moments) {
  extern __shared__ cplx u[];
  const int D = hdim(p);
  const uint32_t id = cells[blockIdx.x];
  if (c.leaf[id]) return;
  const double zx = c.cx[id], zy = c.cy[id], zz = c.cz[id];
  const int32_t* ch = c.children + (size_t)id * 8;
  cplx* out = moments + (size_t)id * D;
  for (int o = 0; o < 8; o++) {
    const int q = ch[o];
    if (q < 0) continue;
    __syncthreads();
    for (int m = threadIdx.x; m <= p; m += blockDim.x)
      upsilon_column_gen(p, c.cx[q] - zx, c.cy[q] - zy, c.cz[q] - zz, m, u);
    __syncthreads();
    const cplx* mc = moments + (size_t)q * D;
    for (int e = threadIdx.x; e < D; e += blockDim.x) {
      int n = 0;
      while ((n + 1) * (n + 1) <= e) n++;
      const int m = e - n * (n + 1);
      cplx s = cz();
      for (int k = 0; k <= n; k++) {
        const int nk = n - k;
        const int lo = max(-k, m - nk), hi = min(k, m + nk);
        for (int l = lo; l <= hi; l++) s = cadd(s, cmul(u[hidx(k, l)], mc[hidx(nk, m - l)]));
      }
      out[e] = cadd(out[e], s);
    }
  }
}

void launch_p2m_gen(int p, const uint32_t* leaf_ids, uint32_t nleaves, DevCells c, DevBodies b,
                    cplx* moments, hipStream_t s) {
  if (nleaves == 0) return;
  const size_t threads = (size_t)nleaves * (p + 1);
  hipLaunchKernelGGL(k_p2m_gen, dim3((threads + 255) / 256), dim3(256), 0, s, p, leaf_ids, nleaves,
                     c, b, moments);
}

void launch_m2m_gen(int p, const uint32_t* cells, uint32_t cnt, DevCells c, cplx* moments,
                    hipStream_t s) {
  if (cnt == 0) return;
  const size_t lds = (size_t)hdim(p) * sizeof(cplx);
  hipLaunchKernelGGL(k_m2m_gen, dim3(cnt), dim3(256), lds, s, p, cells, c, moments);
}

}  // namespace fmmh
