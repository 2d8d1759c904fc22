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

// P2M (moments.rs:30-53) of column M of one leaf; bodies in slot order.
template <int P, int M>
__device__ __forceinline__ void p2m_column(uint32_t id, DevCells c, DevBodies b,
                                           cplx* __restrict__ moments) {
  constexpr int D = hdim(P);
  const double zx = c.cx[id], zy = c.cy[id], zz = c.cz[id];
  cplx acc[P + 1], accn[P + 1];
#pragma unroll
  for (int n = 0; n <= P; n++) acc[n] = accn[n] = cz();
  const uint32_t f = c.first[id], e = f + c.count[id];
  for (uint32_t k = f; k < e; k++) {
    const double4 q = b.xyzm[k];  // one 32-byte load (a lane's leaf differs from its neighbours')
    double x = q.x - zx, y = q.y - zy, z = q.z - zz;
    double r2 = x * x + y * y + z * z;
    upsilon_column_acc<P>(x, y, z, r2, M, q.w, acc, accn);
  }
  cplx* out = moments + (size_t)id * D;
#pragma unroll
  for (int n = M; n <= P; n++) {
    out[hidx(n, M)] = acc[n];
    if (M > 0) out[hidx(n, -M)] = accn[n];
  }
}

template <int P, int M = 0>
__device__ __forceinline__ void p2m_dispatch(int m, uint32_t id, DevCells c, DevBodies b,
                                             cplx* __restrict__ moments) {
  if constexpr (M <= P) {
    if (m == M)
      p2m_column<P, M>(id, c, b, moments);
    else
      p2m_dispatch<P, M + 1>(m, id, c, b, moments);
  }
}

// One block per 64 leaves, wave m computing column m of each: a wave runs
// one column's exact work, and the leaves come in body-count order, so its
// lanes loop over nearly equal numbers of bodies.
template <int P>
__global__ void __launch_bounds__(64 * (P + 1)) k_p2m(const uint32_t* __restrict__ leaf_ids,
                                                     uint32_t nleaves, DevCells c, DevBodies b,
                                                     cplx* __restrict__ moments) {
  const int m = (int)(threadIdx.x >> 6);
  const uint32_t li = blockIdx.x * 64 + (threadIdx.x & 63);
  if (li >= nleaves) return;
  p2m_dispatch<P>(m, leaf_ids[li], c, b, moments);
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

// M2M work schedule. Output coefficient e = (n, m) of a parent is the sum,
// child by child, of its terms (k, l): Upsilon_k^l(child offset) times the
// child's M_{n-k}^{m-l}, in the reference's order (k ascending, then l).
// The schedule lists every output's terms as LDS byte offsets
// (Upsilon index * 16 | moment index * 16 << 16) and orders the outputs
// ("slots") by term count, longest first, so that the eight outputs a wave
// works on at a time have nearly equal trip counts.
constexpr int m2m_terms(int n, int m) {
  int t = 0;
  for (int k = 0; k <= n; k++) {
    const int nk = n - k;
    const int lo = -k > m - nk ? -k : m - nk, hi = k < m + nk ? k : m + nk;
    if (hi >= lo) t += hi - lo + 1;
  }
  return t;
}
constexpr int m2m_total_terms(int p) {
  int t = 0;
  for (int n = 0; n <= p; n++)
    for (int m = -n; m <= n; m++) t += m2m_terms(n, m);
  return t;
}

template <int P>
struct M2mSched {
  static constexpr int D = hdim(P);
  static constexpr int T = m2m_total_terms(P);
  uint32_t code[T];
  uint16_t e_of[D], base[D], nt[D];
};

template <int P>
constexpr M2mSched<P> make_m2m_sched() {
  M2mSched<P> s{};
  constexpr int D = M2mSched<P>::D;
  int cnt[D] = {};
  int ord[D] = {};
  for (int e = 0; e < D; e++) {
    int n = 0;
    while ((n + 1) * (n + 1) <= e) n++;
    cnt[e] = m2m_terms(n, e - n * (n + 1));
    ord[e] = e;
  }
  for (int i = 1; i < D; i++) {  // stable insertion sort, longest first
    const int v = ord[i];
    int j = i;
    while (j > 0 && cnt[ord[j - 1]] < cnt[v]) {
      ord[j] = ord[j - 1];
      j--;
    }
    ord[j] = v;
  }
  int at = 0;
  for (int slot = 0; slot < D; slot++) {
    const int e = ord[slot];
    int n = 0;
    while ((n + 1) * (n + 1) <= e) n++;
    const int m = e - n * (n + 1);
    s.e_of[slot] = (uint16_t)e;
    s.base[slot] = (uint16_t)at;
    s.nt[slot] = (uint16_t)cnt[e];
    for (int k = 0; k <= n; k++) {
      const int nk = n - k;
      const int lo = -k > m - nk ? -k : m - nk, hi = k < m + nk ? k : m + nk;
      for (int l = lo; l <= hi; l++)
        s.code[at++] = (uint32_t)(hidx(k, l) * 16) | ((uint32_t)(hidx(nk, m - l) * 16) << 16);
    }
  }
  return s;
}

template <int P>
__device__ constexpr M2mSched<P> kM2mSched = make_m2m_sched<P>();

constexpr int M2M_T = 256;

// M2M (moments.rs:56-90): one block per internal cell of a level. The
// children's Upsilon tables and moments are staged in LDS; thread (slot
// group g, child o) computes child o's partial sum of the outputs in slots
// g, g + 32, g + 64, ... (each from zero, terms in the reference's order),
// then each output adds its children's partials in octant order. Every
// output is therefore the reference's sum, bit for bit.
template <int P>
__global__ void __launch_bounds__(M2M_T) k_m2m(const uint32_t* __restrict__ cells, uint32_t cnt,
                                              DevCells c, cplx* __restrict__ moments) {
  using S = M2mSched<P>;
  constexpr int D = S::D;
  __shared__ cplx u[8][D];
  __shared__ cplx mo[8][D];
  __shared__ cplx part[8][D];
  __shared__ uint32_t code[S::T];
  const uint32_t id = cells[blockIdx.x];
  if (c.leaf[id]) return;  // block-uniform
  const double zx = c.cx[id], zy = c.cy[id], zz = c.cz[id];
  const int32_t* ch = c.children + (size_t)id * 8;
  for (int t = threadIdx.x; t < 8 * (P + 1); t += M2M_T) {
    const int o = t / (P + 1), m = t % (P + 1);
    const int q = ch[o];
    if (q < 0) continue;
    upsilon_column_store<P>(c.cx[q] - zx, c.cy[q] - zy, c.cz[q] - zz, m, u[o]);
  }
  for (int t = threadIdx.x; t < 8 * D; t += M2M_T) {
    const int o = t / D, e = t - o * D;
    const int q = ch[o];
    if (q >= 0) mo[o][e] = moments[(size_t)q * D + e];
  }
  for (int t = threadIdx.x; t < S::T; t += M2M_T) code[t] = kM2mSched<P>.code[t];
  __syncthreads();
  const int o = threadIdx.x & 7;
  const bool has = ch[o] >= 0;
  const char* ub = (const char*)&u[o][0];
  const char* mb = (const char*)&mo[o][0];
  for (int slot = threadIdx.x >> 3; slot < D; slot += M2M_T / 8) {
    const int e = kM2mSched<P>.e_of[slot], b = kM2mSched<P>.base[slot];
    const int nt = has ? (int)kM2mSched<P>.nt[slot] : 0;
    cplx acc = cz();
    for (int t = 0; t < nt; t++) {
      const uint32_t cd = code[b + t];
      const cplx uv = *(const cplx*)(ub + (cd & 0xFFFFu));
      const cplx mv = *(const cplx*)(mb + (cd >> 16));
      acc = cadd(acc, cmul(uv, mv));
    }
    part[o][e] = acc;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < D; e += M2M_T) {
    cplx acc = cz();
    for (int q = 0; q < 8; q++)
      if (ch[q] >= 0) acc = cadd(acc, part[q][e]);
    moments[(size_t)id * D + e] = acc;
  }
}

template <int P>
static void p2m_t(const uint32_t* leaf_ids, uint32_t nleaves, DevCells c, DevBodies b,
                  cplx* moments, hipStream_t s) {
  hipLaunchKernelGGL(k_p2m<P>, dim3((nleaves + 63) / 64), dim3(64 * (P + 1)), 0, s, leaf_ids,
                     nleaves, c, b, moments);
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
