// upward.hip — upward pass on gfx950 (compiled -ffp-contract=off):
// P2M and M2M on Dehnen solid harmonics, multipole power, the p=0 f pre-pass
// and per-cell fmin. These feed the error-controlled MAC, whose decisions
// must match the reference's exactly, so every sum runs in the reference's
// order (crates/fmm moments.rs:30-168, mac.rs:99-186) and no product is fused.
#include "common.h"
#include "internal.h"

namespace fmmh {

// Column m of Upsilon(r) (harmonic.rs:141-173) for one body, with the
// diagonal seed chained exactly as the table recursion does, accumulated
// as mass * Upsilon into acc[n] (m) and accn[n] (-m). P compile-time keeps
// the column in registers.
template <int P>
__device__ __forceinline__ void upsilon_column_acc(double x, double y, double z, double r2, int m,
                                                   double ms, cplx* acc, cplx* accn) {
  cplx u = cplx{1.0, 0.0};
#pragma unroll
  for (int j = 1; j <= P; j++)
    if (j <= m) u = cmul(cscale(cplx{x, y}, 1.0 / (double)(2 * j)), u);
  const double sgn = (m % 2 == 0) ? 1.0 : -1.0;
  cplx u1 = u, u2 = cz();
#pragma unroll
  for (int n = 0; n <= P; n++) {
    if (n < m) continue;
    cplx un;
    if (n == m) {
      un = u;
    } else {
      double a = (double)(2 * n - 1) * z;
      double c = (double)(n * n - m * m);
      // n == m+1 reads the ZERO seed: 0*r2 = +0 and x - (+0) = x.
      cplx p2 = (n >= m + 2) ? u2 : cz();
      un = cscale(csub(cscale(u1, a), cscale(p2, r2)), 1.0 / c);
      u2 = u1;
      u1 = un;
    }
    acc[n] = cadd(acc[n], cscale(un, ms));
    if (m > 0) {
      cplx neg = cscale(cconj(un), sgn);
      accn[n] = cadd(accn[n], cscale(neg, ms));
    }
  }
}

// P2M (moments.rs:30-53): one thread per (leaf, m >= 0); bodies in slot order.
template <int P>
__global__ void __launch_bounds__(256) k_p2m(const uint32_t* __restrict__ leaf_ids,
                                            uint32_t nleaves, DevCells c, DevBodies b,
                                            cplx* __restrict__ moments) {
  constexpr int D = hdim(P);
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t li = t / (P + 1);
  int m = (int)(t % (P + 1));
  if (li >= nleaves) return;
  uint32_t id = leaf_ids[li];
  const double zx = c.cx[id], zy = c.cy[id], zz = c.cz[id];
  cplx acc[P + 1], accn[P + 1];
#pragma unroll
  for (int n = 0; n <= P; n++) acc[n] = accn[n] = cz();
  const uint32_t f = c.first[id], e = f + c.count[id];
  for (uint32_t k = f; k < e; k++) {
    double x = b.x[k] - zx, y = b.y[k] - zy, z = b.z[k] - zz;
    double r2 = x * x + y * y + z * z;
    upsilon_column_acc<P>(x, y, z, r2, m, b.m[k], acc, accn);
  }
  cplx* out = moments + (size_t)id * D;
#pragma unroll
  for (int n = 0; n <= P; n++) {
    if (n < m) continue;
    out[hidx(n, m)] = acc[n];
    if (m > 0) out[hidx(n, -m)] = accn[n];
  }
}

// Full Upsilon column m (both signs) written into a table (for M2M's
// per-child tables in LDS).
template <int P>
__device__ __forceinline__ void upsilon_column_store(double x, double y, double z, int m,
                                                     cplx* tab) {
  double r2 = x * x + y * y + z * z;
  cplx u = cplx{1.0, 0.0};
#pragma unroll
  for (int j = 1; j <= P; j++)
    if (j <= m) u = cmul(cscale(cplx{x, y}, 1.0 / (double)(2 * j)), u);
  const double sgn = (m % 2 == 0) ? 1.0 : -1.0;
  cplx u1 = u, u2 = cz();
#pragma unroll
  for (int n = 0; n <= P; n++) {
    if (n < m) continue;
    cplx un;
    if (n == m) {
      un = u;
    } else {
      double a = (double)(2 * n - 1) * z;
      double c = (double)(n * n - m * m);
      cplx p2 = (n >= m + 2) ? u2 : cz();
      un = cscale(csub(cscale(u1, a), cscale(p2, r2)), 1.0 / c);
      u2 = u1;
      u1 = un;
    }
    tab[hidx(n, m)] = un;
    if (m > 0) tab[hidx(n, -m)] = cscale(cconj(un), sgn);
  }
}

constexpr int M2M_T = 128;

// M2M (moments.rs:56-90): one block per cell of a level, one thread per
// output coefficient; children in octant order, the triple sum in the
// reference's (k, l) order. Leaves of the level exit at once.
template <int P>
__global__ void __launch_bounds__(M2M_T) k_m2m(const uint32_t* __restrict__ cells, uint32_t cnt,
                                              DevCells c, cplx* __restrict__ moments) {
  constexpr int D = hdim(P);
  __shared__ cplx u[8][D];
  uint32_t id = cells[blockIdx.x];
  if (c.leaf[id]) return;
  const double zx = c.cx[id], zy = c.cy[id], zz = c.cz[id];
  const int32_t* ch = c.children + (size_t)id * 8;
  for (int t = threadIdx.x; t < 8 * (P + 1); t += M2M_T) {
    int o = t / (P + 1), m = t % (P + 1);
    int q = ch[o];
    if (q < 0) continue;
    upsilon_column_store<P>(c.cx[q] - zx, c.cy[q] - zy, c.cz[q] - zz, m, u[o]);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < D; e += M2M_T) {
    int n = 0;
    while ((n + 1) * (n + 1) <= e) n++;
    int m = e - n * (n + 1);
    cplx acc = cz();
    for (int o = 0; o < 8; o++) {
      int q = ch[o];
      if (q < 0) continue;
      const cplx* mc = moments + (size_t)q * D;
      cplx s = cz();
      for (int k = 0; k <= n; k++) {
        int nk = n - k;
        int lo = max(-k, m - nk), hi = min(k, m + nk);
        for (int l = lo; l <= hi; l++) s = cadd(s, cmul(u[o][hidx(k, l)], mc[hidx(nk, m - l)]));
      }
      acc = cadd(acc, s);
    }
    moments[(size_t)id * D + e] = acc;
  }
}

template <int P>
static void p2m_t(const uint32_t* leaf_ids, uint32_t nleaves, DevCells c, DevBodies b,
                  cplx* moments, hipStream_t s) {
  size_t threads = (size_t)nleaves * (P + 1);
  hipLaunchKernelGGL(k_p2m<P>, dim3((threads + 255) / 256), dim3(256), 0, s, leaf_ids, nleaves, c,
                     b, moments);
}

template <int P>
static void m2m_t(const uint32_t* cells, uint32_t cnt, DevCells c, cplx* moments, hipStream_t s) {
  hipLaunchKernelGGL(k_m2m<P>, dim3(cnt), dim3(M2M_T), 0, s, cells, cnt, c, moments);
}

#define FMMH_P_SWITCH(p, CALL) \
  switch (p) {                 \
    case 1: CALL(1); break;    \
    case 2: CALL(2); break;    \
    case 3: CALL(3); break;    \
    case 4: CALL(4); break;    \
    case 5: CALL(5); break;    \
    case 6: CALL(6); break;    \
    case 7: CALL(7); break;    \
    case 8: CALL(8); break;    \
    case 9: CALL(9); break;    \
    case 10: CALL(10); break;  \
    default: return false;     \
  }

bool launch_p2m(int p, const uint32_t* leaf_ids, uint32_t nleaves, DevCells c, DevBodies b,
                cplx* moments, hipStream_t s) {
  if (nleaves == 0) return true;
#define CALL(P) p2m_t<P>(leaf_ids, nleaves, c, b, moments, s)
  FMMH_P_SWITCH(p, CALL)
#undef CALL
  return true;
}

bool launch_m2m(int p, const uint32_t* cells, uint32_t cnt, DevCells c, cplx* moments,
                hipStream_t s) {
  if (cnt == 0) return true;
#define CALL(P) m2m_t<P>(cells, cnt, c, moments, s)
  FMMH_P_SWITCH(p, CALL)
#undef CALL
  return true;
}

// Multipole power (moments.rs:150-168): one thread per (cell, n).
__global__ void __launch_bounds__(256) k_pows(int p, uint32_t nc, const cplx* __restrict__ moments,
                                             const double* __restrict__ fact,
                                             double* __restrict__ pows) {
  size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (size_t)nc * (p + 1)) return;
  size_t cell = t / (p + 1);
  int n = (int)(t % (p + 1));
  const cplx* mc = moments + cell * (size_t)hdim(p);
  double s = 0.0;
  for (int m = -n; m <= n; m++) {
    int am = m < 0 ? -m : m;
    cplx v = mc[hidx(n, m)];
    double w = fact[n - am] * fact[n + am];
    s += w * (v.re * v.re + v.im * v.im);
  }
  pows[t] = sqrt(s);
}

void launch_pows(int p, uint32_t nc, const cplx* moments, const double* fact, double* pows,
                 hipStream_t s) {
  size_t threads = (size_t)nc * (p + 1);
  hipLaunchKernelGGL(k_pows, dim3((threads + 255) / 256), dim3(256), 0, s, p, nc, moments, fact,
                     pows);
}

// f pre-pass (mac.rs:99-147) as a wave-coherent stackless walk. The 64
// lanes (64 consecutive slots) share one traversal pointer over the
// mirrored DFS (children 7..0, exactly the reference stack's pop order); a
// lane that accepts a cell's monopole sleeps until the wave reaches that
// cell's rope, so each lane accumulates precisely its own reference
// sequence of terms, in order. Cell and body loads are wave-uniform.
// Ensembles (vlev > 0): a lane walks only its own system's subtree. A
// virtual cell (level < vlev) is opened by the lanes whose system it holds
// and skipped by the others; a real cell of another system is skipped; so
// every lane meets its system's cells in the standalone order.
struct FpCells {
  const uint8_t* __restrict__ leaf;
  const uint32_t* __restrict__ first;
  const uint32_t* __restrict__ count;
  const double* __restrict__ cx;
  const double* __restrict__ cy;
  const double* __restrict__ cz;
  const double* __restrict__ rho;
  const double* __restrict__ mass;
  const int32_t* __restrict__ rope;
  const int32_t* __restrict__ last_child;
  const uint8_t* __restrict__ level;
};

__global__ void __launch_bounds__(64) k_fprepass(uint32_t n, uint32_t max_steps, FpCells c,
                                                const double* __restrict__ bx,
                                                const double* __restrict__ by,
                                                const double* __restrict__ bz,
                                                const double* __restrict__ bm,
                                                const uint64_t* __restrict__ keys_sorted,
                                                int vlev, double* __restrict__ f) {
  const int ACTIVE = -2, NEVER = -3;
  const uint32_t i = blockIdx.x * 64 + threadIdx.x;
  const bool valid = i < n;
  const int sh = 63 - 3 * vlev;
  const uint32_t mysys = (valid && vlev) ? (uint32_t)(keys_sorted[i] >> sh) : 0u;
  const double px = valid ? bx[i] : 0.0, py = valid ? by[i] : 0.0, pz = valid ? bz[i] : 0.0;
  double fi = 0.0;
  int resume = valid ? ACTIVE : NEVER;
  int cur = 0;
  // The mirrored DFS visits each cell at most once, so a wave needs at most
  // ncells steps; the cap only guards against a corrupt tree.
  for (uint32_t step = 0; cur >= 0 && step < max_steps; step++) {
    cur = wave_uniform(cur);
    if (resume == cur) resume = ACTIVE;
    bool act = resume == ACTIVE;
    bool virt = false;
    if (vlev) {
      const uint32_t f0 = c.first[cur];
      const uint32_t slo = (uint32_t)(keys_sorted[f0] >> sh);
      virt = (int)c.level[cur] < vlev;
      if (act) {
        const uint32_t shi =
            virt ? (uint32_t)(keys_sorted[f0 + c.count[cur] - 1] >> sh) : slo;
        if (mysys < slo || mysys > shi) {  // not this lane's system
          resume = c.rope[cur];
          act = false;
        }
      }
    }
    if (c.leaf[cur]) {
      const uint32_t f0 = c.first[cur], e = f0 + c.count[cur];
      for (uint32_t k = f0; k < e; k++) {
        const double xk = bx[k], yk = by[k], zk = bz[k], mk = bm[k];
        if (act && k != i) {
          double dx = px - xk, dy = py - yk, dz = pz - zk;
          double r2 = dx * dx + dy * dy + dz * dz;
          if (r2 > 0.0) fi += mk / r2;
        }
      }
      cur = c.rope[cur];
    } else {
      const double cxv = c.cx[cur], cyv = c.cy[cur], czv = c.cz[cur];
      const double rh = c.rho[cur], ms = c.mass[cur];
      const int rope = c.rope[cur];
      bool open = false;
      if (act && virt) {
        open = true;  // a virtual cell on the way to the lane's system root
      } else if (act) {
        double dx = px - cxv, dy = py - cyv, dz = pz - czv;
        double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > rh * rh) {
          fi += ms / d2;
          resume = rope;
        } else {
          open = true;
        }
      }
      cur = __ballot(open) ? c.last_child[cur] : rope;
    }
  }
  if (valid) f[i] = fi;
}

void launch_fprepass(uint32_t n, uint32_t nc, DevCells c, DevBodies b,
                     const uint64_t* keys_sorted, int vlev, double* f, hipStream_t s) {
  FpCells fc{c.leaf, c.first, c.count, c.cx, c.cy, c.cz, c.rho, c.mass, c.rope, c.last_child,
             c.level};
  hipLaunchKernelGGL(k_fprepass, dim3((n + 63) / 64), dim3(64), 0, s, n, nc + 1, fc, b.x, b.y, b.z,
                     b.m, keys_sorted, vlev, f);
}

// fmin (mac.rs:151-173), one level at a time, deepest first.
__global__ void __launch_bounds__(256) k_fmin(const uint32_t* __restrict__ cells, uint32_t cnt,
                                             DevCells c, const double* __restrict__ f,
                                             double* __restrict__ fmin_out) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= cnt) return;
  uint32_t id = cells[i];
  double m = INFINITY;
  if (c.leaf[id]) {
    uint32_t f0 = c.first[id], e = f0 + c.count[id];
    for (uint32_t k = f0; k < e; k++) m = fmin(m, f[k]);
  } else {
    const int32_t* ch = c.children + (size_t)id * 8;
    for (int o = 0; o < 8; o++)
      if (ch[o] >= 0) m = fmin(m, fmin_out[ch[o]]);
  }
  fmin_out[id] = m;
}

void launch_fmin(const uint32_t* cells, uint32_t cnt, DevCells c, const double* f, double* fmin,
                 hipStream_t s) {
  if (cnt == 0) return;
  hipLaunchKernelGGL(k_fmin, dim3((cnt + 255) / 256), dim3(256), 0, s, cells, cnt, c, f, fmin);
}

}  // namespace fmmh
