// fmm_oracle.cpp — CPU oracle: C++ restatement of the reference solver
// `crates/fmm` (Vedant9500/FMMS, Rust). TEST INFRASTRUCTURE ONLY: it is the
// checker for the GPU product and the CPU baseline, never part of the product.
//
// Build with -ffp-contract=off -fno-fast-math: every floating-point
// expression below is evaluated in the reference's order, unfused, so the
// tree geometry, moments, multipole powers, f pre-pass and MAC decisions are
// bit-identical to the reference's (Rust never contracts a*b+c and calls the
// same IEEE sqrt / division; libm pow/sin/cos only in the samplers).
//
// Each function cites the reference file:line it restates.
#include "fmm_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

// ORC_TIMING: the CPU-baseline build (oracle/Makefile `timing`), compiled
// for the host's own ISA with multiply-adds fused. It times the reference's
// algorithm the way the reference runs it: 8-wide FP64 P2P with an rsqrt
// estimate refined by two Newton steps (p2p.rs:117-326 takes that route on
// AVX-512 hosts) and threads that take work dynamically (walk.rs:342-488
// steals work). Its results differ from the parity build in the last bits;
// it is never the checker.
#if defined(ORC_TIMING) && defined(__AVX512F__)
#include <immintrin.h>
#define ORC_P2P_AVX512 1
#endif

namespace {

struct C {
  double re, im;
};
inline C cz() { return C{0.0, 0.0}; }
// harmonic.rs:22-74 — component-wise, unfused.
inline C cadd(C a, C b) { return C{a.re + b.re, a.im + b.im}; }
inline C csub(C a, C b) { return C{a.re - b.re, a.im - b.im}; }
inline C cmul(C a, C b) { return C{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline C cscale(C a, double s) { return C{a.re * s, a.im * s}; }
inline C cconj(C a) { return C{a.re, -a.im}; }

inline size_t dimp(int p) { return (size_t)(p + 1) * (size_t)(p + 1); }  // harmonic.rs:78
inline size_t idx(int n, int m) { return (size_t)((long)n * (n + 1) + m); }  // harmonic.rs:84

// ---------------------------------------------------------------- models.rs
struct Rng {  // models.rs:13-31 SplitMix64
  uint64_t s;
  uint64_t next_u64() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double unit() { return (double)(next_u64() >> 11) / (double)(1ull << 53); }
};

const double PI = 3.14159265358979323846;  // std::f64::consts::PI

// ------------------------------------------------------------------ tree.rs
const int BITS = 21;  // tree.rs:16

uint64_t dilate(uint64_t x) {  // tree.rs:28-35 (standard 21-bit 3D magic-bit spread)
  x &= (1ull << BITS) - 1;
  x = (x | (x << 32)) & 0x001f00000000ffffull;
  x = (x | (x << 16)) & 0x001f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  return (x | (x << 2)) & 0x1249249249249249ull;
}

// tree.rs:37-46. Quirk kept: every axis subtracts lo[0] (the x minimum).
// Rust clamp keeps NaN; `NaN as u64` saturates to 0.
uint64_t quantise(double c, double lo0, double inv_side) {
  double v = (c - lo0) * inv_side;
  if (v < 0.0) v = 0.0;
  if (v > 1.0) v = 1.0;
  double w = v * (double)((1ull << BITS) - 1);
  if (!(w == w)) return 0;
  return (uint64_t)w;
}
uint64_t morton_key(double x, double y, double z, const double lo[3], double inv_side) {
  uint64_t key = dilate(quantise(z, lo[0], inv_side));
  key |= dilate(quantise(y, lo[0], inv_side)) << 1;
  key |= dilate(quantise(x, lo[0], inv_side)) << 2;
  return key;
}

struct Cell {  // tree.rs:49-58
  size_t first, count;
  int32_t children[8];
  bool leaf;
  uint32_t level;
  double cen[3];
  double mass;
  double rho;
};

inline double dist2(const double* a, const double* b) {  // tree.rs:392-398
  double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return dx * dx + dy * dy + dz * dz;
}
inline double dist(const double* a, const double* b) { return std::sqrt(dist2(a, b)); }

struct Item {
  double c[3];
  double r;
};

// tree.rs:290-389 Ritter grow/enclose sphere with the bbox-centred fallback.
void ritter(const std::vector<Item>& items, double out_c[3], double* out_r) {
  size_t a = 0, b = 0;
  double best = -1.0;
  for (size_t i = 0; i < items.size(); i++) {
    double d2 = dist2(items[i].c, items[a].c);
    if (d2 > best) {
      best = d2;
      b = i;
    }
  }
  a = b;
  best = -1.0;
  for (size_t i = 0; i < items.size(); i++) {
    double d2 = dist2(items[i].c, items[a].c);
    if (d2 > best) {
      best = d2;
      b = i;
    }
  }
  const double* pa = items[a].c;
  const double* pb = items[b].c;
  double dir[3];
  {
    double dx = pb[0] - pa[0], dy = pb[1] - pa[1], dz = pb[2] - pa[2];
    double d = std::sqrt(dx * dx + dy * dy + dz * dz);
    d = std::max(d, std::numeric_limits<double>::min());
    dir[0] = dx / d;
    dir[1] = dy / d;
    dir[2] = dz / d;
  }
  double far_a[3] = {pa[0] + items[a].r * dir[0], pa[1] + items[a].r * dir[1],
                     pa[2] + items[a].r * dir[2]};
  double far_b[3] = {pb[0] - items[b].r * dir[0], pb[1] - items[b].r * dir[1],
                     pb[2] - items[b].r * dir[2]};
  double c[3] = {(far_a[0] + far_b[0]) * 0.5, (far_a[1] + far_b[1]) * 0.5,
                 (far_a[2] + far_b[2]) * 0.5};
  double r = dist(far_a, far_b) * 0.5;
  for (int pass = 0; pass < 2; pass++) {
    for (const Item& it : items) {
      double d = dist(it.c, c);
      if (d + it.r > r) {
        double dd = std::max(d, std::numeric_limits<double>::min());
        double dr[3] = {(it.c[0] - c[0]) / dd, (it.c[1] - c[1]) / dd, (it.c[2] - c[2]) / dd};
        double pnt[3] = {it.c[0] + it.r * dr[0], it.c[1] + it.r * dr[1], it.c[2] + it.r * dr[2]};
        double d_far = d + it.r;
        double t = (d_far - r) / (2.0 * d_far);
        for (int k = 0; k < 3; k++) c[k] += t * (pnt[k] - c[k]);
        r = (d_far + r) * 0.5;
      }
    }
  }
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (const Item& it : items)
    for (int d = 0; d < 3; d++) {
      lo[d] = std::fmin(lo[d], it.c[d] - it.r);
      hi[d] = std::fmax(hi[d], it.c[d] + it.r);
    }
  double cb[3] = {(lo[0] + hi[0]) * 0.5, (lo[1] + hi[1]) * 0.5, (lo[2] + hi[2]) * 0.5};
  double rb = 0.0;
  for (const Item& it : items) rb = std::fmax(rb, dist(it.c, cb) + it.r);
  if (rb < r) {
    std::memcpy(out_c, cb, sizeof cb);
    *out_r = rb;
  } else {
    std::memcpy(out_c, c, sizeof c);
    *out_r = r;
  }
}

// tree.rs:404-445, DFS preorder, children in octant order.
size_t build_rec(const std::vector<uint64_t>& keys, std::vector<Cell>& cells, uint32_t level,
                 size_t start, size_t end, size_t n_leaf, uint32_t max_level) {
  size_t count = end - start;
  bool leaf = count <= n_leaf || level >= max_level;
  Cell c;
  c.first = start;
  c.count = count;
  for (int i = 0; i < 8; i++) c.children[i] = -1;
  c.leaf = leaf;
  c.level = level;
  c.cen[0] = c.cen[1] = c.cen[2] = 0.0;
  c.mass = 0.0;
  c.rho = 0.0;
  cells.push_back(c);
  size_t id = cells.size() - 1;
  if (leaf) return id;
  uint32_t shift = 3 * (max_level - (level + 1));
  size_t lo = start;
  for (uint64_t oct = 0; oct < 8; oct++) {
    size_t hi = lo;
    while (hi < end && ((keys[hi] >> shift) & 7ull) == oct) hi++;
    if (hi > lo) {
      size_t cid = build_rec(keys, cells, level + 1, lo, hi, n_leaf, max_level);
      cells[id].children[oct] = (int32_t)cid;
    }
    lo = hi;
  }
  return id;
}

// --------------------------------------------------------------- harmonic.rs
// harmonic.rs:91-129. For n = m+1 the (n-2, m) read lands on slot m^2 (the
// not-yet-written Theta_m^{-m}) and is multiplied by zero, as in the reference.
void theta_into(const double r[3], int p, C* t, double eps2 = 0.0) {
  double x = r[0], y = r[1], z = r[2];
  double r2 = x * x + y * y + z * z;
  double ir = 1.0 / std::sqrt(r2);
  t[0] = C{ir, 0.0};
  for (int n = 1; n <= p; n++) {
    double a = (double)(2 * n - 1) * z;
    double b = (double)((n - 1) * (n - 1));
    C p1 = t[idx(n - 1, 0)];
    C p2 = n >= 2 ? t[idx(n - 2, 0)] : cz();
    t[idx(n, 0)] = cscale(csub(cscale(p1, a), cscale(p2, b)), 1.0 / r2);
  }
  for (int m = 1; m <= p; m++) {
    t[idx(m, m)] = cmul(cscale(C{x, y}, (double)(2 * m - 1) / r2), t[idx(m - 1, m - 1)]);
    for (int n = m + 1; n <= p; n++) {
      double a = (double)(2 * n - 1) * z;
      double b = (double)((n - 1) * (n - 1) - m * m);
      t[idx(n, m)] =
          cscale(csub(cscale(t[idx(n - 1, m)], a), cscale(t[(size_t)((n - 2) * (n - 1) + m)], b)),
                 1.0 / r2);
    }
  }
  for (int m = 1; m <= p; m++) {
    double sgn = (m % 2 == 0) ? 1.0 : -1.0;
    for (int n = m; n <= p; n++) t[idx(n, -m)] = cscale(cconj(t[idx(n, m)]), sgn);
  }
  // Softening extension: the harmonic (traceless) derivatives of the Plummer
  // kernel (r^2 + eps^2)^{-1/2} are those of 1/r times (r/s)^{2n+1},
  // s^2 = r^2 + eps^2 (Hobson's theorem for a function of r^2: Y_n(grad) F(r^2)
  // = 2^n Y_n(r) F^(n)(r^2)). eps2 == 0 leaves the reference arithmetic intact.
  if (eps2 > 0.0) {
    double q = r2 / (r2 + eps2), w = std::sqrt(q);
    for (int n = 0; n <= p; n++) {
      for (int m = -n; m <= n; m++) t[idx(n, m)] = cscale(t[idx(n, m)], w);
      w *= q;
    }
  }
}

// harmonic.rs:141-173
void upsilon_into(const double r[3], int p, C* u) {
  double x = r[0], y = r[1], z = r[2];
  double r2 = x * x + y * y + z * z;
  u[0] = C{1.0, 0.0};
  for (int m = 0; m <= p; m++) {
    if (m > 0) u[idx(m, m)] = cmul(cscale(C{x, y}, 1.0 / (double)(2 * m)), u[idx(m - 1, m - 1)]);
    for (int n = m + 1; n <= p; n++) {
      double a = (double)(2 * n - 1) * z;
      double b = r2;
      double c = (double)(n * n - m * m);
      C p2 = (n >= m + 2) ? u[idx(n - 2, m)] : cz();
      u[idx(n, m)] = cscale(csub(cscale(u[idx(n - 1, m)], a), cscale(p2, b)), 1.0 / c);
    }
  }
  for (int m = 1; m <= p; m++) {
    double sgn = (m % 2 == 0) ? 1.0 : -1.0;
    for (int n = m; n <= p; n++) u[idx(n, -m)] = cscale(cconj(u[idx(n, m)]), sgn);
  }
}

// ---------------------------------------------------------------- moments.rs
std::vector<double> factorials(int p) {  // moments.rs:11-17
  std::vector<double> f(2 * p + 1, 1.0);
  for (size_t i = 1; i < f.size(); i++) f[i] = f[i - 1] * (double)i;
  return f;
}
std::vector<double> binomials(int p) {  // moments.rs:20-27
  std::vector<double> fact = factorials(p);
  std::vector<double> b(p + 1, 1.0);
  for (int k = 0; k <= p; k++) b[k] = fact[p] / (fact[k] * fact[p - k]);
  return b;
}

struct Bodies {  // bodies.rs:13-22, slot order
  std::vector<double> x, y, z, m, ax, ay, az, pot;
  void resize(size_t n) {
    for (auto* v : {&x, &y, &z, &m, &ax, &ay, &az, &pot}) v->assign(n, 0.0);
  }
};

// ----------------------------------------------------------------- p2p.rs
// Scalar mutual kernels (p2p.rs:13-88, the "exact reference" path). Softening
// (extension): r2 += soft2 when soft2 > 0, so soft2 == 0 is bit-identical.
#ifdef ORC_P2P_AVX512
// Timing build: body i against bodies [j0, j1) in 8-wide blocks, both sides
// updated (the mutual kernel); the tail block is masked.
inline void p2p_row_avx512(Bodies& b, size_t i, size_t j0, size_t j1, double soft2) {
  const __m512d xi = _mm512_set1_pd(b.x[i]), yi = _mm512_set1_pd(b.y[i]),
                zi = _mm512_set1_pd(b.z[i]), mi = _mm512_set1_pd(b.m[i]);
  const __m512d e2 = _mm512_set1_pd(soft2), half = _mm512_set1_pd(0.5),
                onehalf = _mm512_set1_pd(1.5);
  __m512d ax = _mm512_setzero_pd(), ay = ax, az = ax, ph = ax;
  for (size_t j = j0; j < j1; j += 8) {
    const size_t left = j1 - j;
    const __mmask8 k = left >= 8 ? (__mmask8)0xFF : (__mmask8)((1u << left) - 1u);
    const __m512d xj = _mm512_maskz_loadu_pd(k, &b.x[j]);
    const __m512d yj = _mm512_maskz_loadu_pd(k, &b.y[j]);
    const __m512d zj = _mm512_maskz_loadu_pd(k, &b.z[j]);
    const __m512d mj = _mm512_maskz_loadu_pd(k, &b.m[j]);
    const __m512d dx = _mm512_sub_pd(xj, xi), dy = _mm512_sub_pd(yj, yi),
                  dz = _mm512_sub_pd(zj, zi);
    __m512d r2 = _mm512_fmadd_pd(dx, dx, _mm512_fmadd_pd(dy, dy, _mm512_fmadd_pd(dz, dz, e2)));
    r2 = _mm512_mask_blend_pd(k, _mm512_set1_pd(1.0), r2);  // masked lanes: r = 1, m = 0
    __m512d inv = _mm512_rsqrt14_pd(r2);
    for (int it = 0; it < 2; it++) {  // Newton: y (3/2 - r2 y^2 / 2)
      const __m512d hy2 = _mm512_mul_pd(_mm512_mul_pd(half, r2), _mm512_mul_pd(inv, inv));
      inv = _mm512_mul_pd(inv, _mm512_sub_pd(onehalf, hy2));
    }
    const __m512d inv3 = _mm512_mul_pd(_mm512_mul_pd(inv, inv), inv);
    const __m512d fj = _mm512_mul_pd(mj, inv3);
    ax = _mm512_fmadd_pd(fj, dx, ax);
    ay = _mm512_fmadd_pd(fj, dy, ay);
    az = _mm512_fmadd_pd(fj, dz, az);
    ph = _mm512_fmadd_pd(mj, inv, ph);
    const __m512d fi = _mm512_mul_pd(mi, inv3);
    _mm512_mask_storeu_pd(&b.ax[j], k, _mm512_fnmadd_pd(fi, dx, _mm512_maskz_loadu_pd(k, &b.ax[j])));
    _mm512_mask_storeu_pd(&b.ay[j], k, _mm512_fnmadd_pd(fi, dy, _mm512_maskz_loadu_pd(k, &b.ay[j])));
    _mm512_mask_storeu_pd(&b.az[j], k, _mm512_fnmadd_pd(fi, dz, _mm512_maskz_loadu_pd(k, &b.az[j])));
    _mm512_mask_storeu_pd(&b.pot[j], k,
                          _mm512_fmadd_pd(mi, inv, _mm512_maskz_loadu_pd(k, &b.pot[j])));
  }
  b.ax[i] += _mm512_reduce_add_pd(ax);
  b.ay[i] += _mm512_reduce_add_pd(ay);
  b.az[i] += _mm512_reduce_add_pd(az);
  b.pot[i] += _mm512_reduce_add_pd(ph);
}

void p2p_self(Bodies& b, size_t first, size_t count, double soft2) {
  for (size_t i = 0; i + 1 < count; i++) p2p_row_avx512(b, first + i, first + i + 1, first + count, soft2);
}

void p2p_pair(Bodies& b, size_t af, size_t ac, size_t bf, size_t bc, double soft2) {
  for (size_t i = 0; i < ac; i++) p2p_row_avx512(b, af + i, bf, bf + bc, soft2);
}
#else
void p2p_self(Bodies& b, size_t first, size_t count, double soft2) {
  for (size_t i = 0; i < count; i++) {
    double xi = b.x[first + i], yi = b.y[first + i], zi = b.z[first + i], mi = b.m[first + i];
    double axi = 0.0, ayi = 0.0, azi = 0.0, poti = 0.0;
    for (size_t j = i + 1; j < count; j++) {
      double dx = b.x[first + j] - xi;
      double dy = b.y[first + j] - yi;
      double dz = b.z[first + j] - zi;
      double r2 = dx * dx + dy * dy + dz * dz;
      if (soft2 > 0.0) r2 = r2 + soft2;
      double inv = 1.0 / std::sqrt(r2);
      double inv3 = inv * inv * inv;
      double mj = b.m[first + j];
      axi += mj * inv3 * dx;
      ayi += mj * inv3 * dy;
      azi += mj * inv3 * dz;
      poti += mj * inv;
      double fm = mi * inv3;
      b.ax[first + j] -= fm * dx;
      b.ay[first + j] -= fm * dy;
      b.az[first + j] -= fm * dz;
      b.pot[first + j] += mi * inv;
    }
    b.ax[first + i] += axi;
    b.ay[first + i] += ayi;
    b.az[first + i] += azi;
    b.pot[first + i] += poti;
  }
}

void p2p_pair(Bodies& b, size_t af, size_t ac, size_t bf, size_t bc, double soft2) {
  for (size_t i = 0; i < ac; i++) {
    double xi = b.x[af + i], yi = b.y[af + i], zi = b.z[af + i], mi = b.m[af + i];
    double axi = 0.0, ayi = 0.0, azi = 0.0, poti = 0.0;
    for (size_t j = 0; j < bc; j++) {
      double dx = b.x[bf + j] - xi;
      double dy = b.y[bf + j] - yi;
      double dz = b.z[bf + j] - zi;
      double r2 = dx * dx + dy * dy + dz * dz;
      if (soft2 > 0.0) r2 = r2 + soft2;
      double inv = 1.0 / std::sqrt(r2);
      double inv3 = inv * inv * inv;
      double mj = b.m[bf + j];
      axi += mj * inv3 * dx;
      ayi += mj * inv3 * dy;
      azi += mj * inv3 * dz;
      poti += mj * inv;
      double fm = mi * inv3;
      b.ax[bf + j] -= fm * dx;
      b.ay[bf + j] -= fm * dy;
      b.az[bf + j] -= fm * dz;
      b.pot[bf + j] += mi * inv;
    }
    b.ax[af + i] += axi;
    b.ay[af + i] += ayi;
    b.az[af + i] += azi;
    b.pot[af + i] += poti;
  }
}
#endif  // ORC_P2P_AVX512

// -------------------------------------------------------------- translate.rs
// translate.rs:47-136 mutual naive M2L on explicit centres (z = s = cen).
struct M2lScratch {
  std::vector<C> th, ths, cma, cmb;
  explicit M2lScratch(int p) {
    size_t d = dimp(p);
    th.assign(d, cz());
    ths.assign(d, cz());
    cma.assign(d, cz());
    cmb.assign(d, cz());
  }
};

void m2l_mutual(int p, const double* cena, const double* cenb, const C* ma, const C* mb, C* la,
                C* lb, M2lScratch& s, double eps2 = 0.0) {
  size_t d = dimp(p);
  double r[3] = {cenb[0] - cena[0], cenb[1] - cena[1], cenb[2] - cena[2]};
  theta_into(r, p, s.th.data(), eps2);
  for (size_t i = 0; i < d; i++) {
    s.cma[i] = cconj(ma[i]);
    s.cmb[i] = cconj(mb[i]);
  }
  {
    size_t o = 0;
    for (int n = 0; n <= p; n++) {
      double sgn = (n % 2 == 0) ? 1.0 : -1.0;
      for (int q = -n; q <= n; q++) {
        s.ths[o] = cscale(s.th[o], sgn);
        o++;
      }
    }
  }
  const C* th = s.th.data();
  const C* ths = s.ths.data();
  const C* cma = s.cma.data();
  const C* cmb = s.cmb.data();
  for (int n = 0; n <= p; n++) {
    long n2n = (long)n * (n + 1);
    for (int m = -n; m <= n; m++) {
      size_t nidx = (size_t)(n2n + m);
      double fa_r = 0.0, fa_i = 0.0, fb_r = 0.0, fb_i = 0.0;
      for (int k = 0; k <= p - n; k++) {
        long sdeg = n + k;
        long sk = sdeg * (sdeg + 1);
        long kb = (long)k * (k + 1);
        size_t thx = (size_t)(sk + m - k);
        size_t i = (size_t)(kb - k);
        for (int l = -k; l <= k; l++) {
          C t = th[thx], ts = ths[thx], a = cma[i], b = cmb[i];
          fb_r += a.re * t.re - a.im * t.im;
          fb_i += a.re * t.im + a.im * t.re;
          fa_r += b.re * ts.re - b.im * ts.im;
          fa_i += b.re * ts.im + b.im * ts.re;
          thx++;
          i++;
        }
      }
      lb[nidx] = cadd(lb[nidx], C{fb_r, fb_i});
      la[nidx] = cadd(la[nidx], C{fa_r, fa_i});
    }
  }
}

// translate.rs:175-206 (r = s - s', F' += conj(Y) . F)
void l2l_into(int p, const double* cp, const double* cc, const C* fpar, C* fch, C* u) {
  double r[3] = {cp[0] - cc[0], cp[1] - cc[1], cp[2] - cc[2]};
  upsilon_into(r, p, u);
  for (int n = 0; n <= p; n++)
    for (int m = -n; m <= n; m++) {
      size_t nidx = idx(n, m);
      C s = cz();
      for (int k = 0; k <= p - n; k++)
        for (int l = -k; l <= k; l++) s = cadd(s, cmul(cconj(u[idx(k, l)]), fpar[idx(n + k, m + l)]));
      fch[nidx] = cadd(fch[nidx], s);
    }
}

// translate.rs:271-309
void l2p_into(int p, const double* cen, const double* pos, const C* f, double* pot, double* acc,
              C* u) {
  double r[3] = {cen[0] - pos[0], cen[1] - pos[1], cen[2] - pos[2]};
  upsilon_into(r, p, u);
  C psi00 = cz();
  for (int k = 0; k <= p; k++)
    for (int l = -k; l <= k; l++) psi00 = cadd(psi00, cmul(cconj(u[idx(k, l)]), f[idx(k, l)]));
  C psi11 = cz(), psi10 = cz();
  if (p >= 1) {
    for (int k = 0; k <= p - 1; k++)
      for (int l = -k; l <= k; l++) {
        C uc = cconj(u[idx(k, l)]);
        psi11 = cadd(psi11, cmul(uc, f[idx(1 + k, 1 + l)]));
        psi10 = cadd(psi10, cmul(uc, f[idx(1 + k, l)]));
      }
  }
  *pot += psi00.re;
  acc[0] -= psi11.re;
  acc[1] -= psi11.im;
  acc[2] -= psi10.re;
}

// -------------------------------------------------------------- rotation.rs
const int ROT_P = 20;  // rotation.rs:45
size_t blocksum(int p) { return (size_t)(p + 1) * (4 * p * p + 8 * p + 3) / 3; }  // :49-51
size_t block_off(int j) { return j == 0 ? 0 : (size_t)j * (4 * j * j - 1) / 3; }  // :55-61

// rotation.rs:70-103 (4-way accumulator split, then the tail).
inline void dot_row(const double* w, const C* v, int nb, double* sr_out, double* si_out) {
  double a0r = 0, a0i = 0, a1r = 0, a1i = 0, a2r = 0, a2i = 0, a3r = 0, a3i = 0;
  int c = 0;
  while (c + 4 <= nb) {
    a0r += w[c] * v[c].re;
    a0i += w[c] * v[c].im;
    a1r += w[c + 1] * v[c + 1].re;
    a1i += w[c + 1] * v[c + 1].im;
    a2r += w[c + 2] * v[c + 2].re;
    a2i += w[c + 2] * v[c + 2].im;
    a3r += w[c + 3] * v[c + 3].re;
    a3i += w[c + 3] * v[c + 3].im;
    c += 4;
  }
  double sr = (a0r + a1r) + (a2r + a3r);
  double si = (a0i + a1i) + (a2i + a3i);
  while (c < nb) {
    sr += w[c] * v[c].re;
    si += w[c] * v[c].im;
    c++;
  }
  *sr_out = sr;
  *si_out = si;
}

struct RotCtx {  // rotation.rs:106-253
  int p;
  std::vector<double> fact, sw, swt, facr;
  std::vector<C> in_, out_, fin, fout, b, f, tmp, tmp2, pha, phb;
  explicit RotCtx(int p_) : p(p_) {
    size_t d = dimp(p);
    fact.assign(p + 1, 1.0);
    for (int i = 1; i <= p; i++) fact[i] = fact[i - 1] * (double)i;
    std::vector<double> f2(2 * p + 1, 1.0);
    for (size_t i = 1; i < f2.size(); i++) f2[i] = f2[i - 1] * (double)i;
    size_t bs = blocksum(p);
    sw.assign(bs, 0.0);
    swt.assign(bs, 0.0);
    sw[0] = 1.0;
    swt[0] = 1.0;
    for (int n = 1; n <= p; n++) {
      int nb = 2 * n + 1;
      size_t off = block_off(n);
      int nbp = 2 * n - 1;
      size_t offp = block_off(n - 1);
      for (int mr = 0; mr < nb; mr++) {
        int m = mr - n;
        size_t row = off + (size_t)mr * nb;
        size_t rowp = offp + (size_t)(mr > 0 ? mr - 1 : 0) * nbp;
        for (int lc = 0; lc < nb; lc++) {
          int l = lc - n;
          double v;
          auto at = [&](size_t base, int col) -> double {
            return std::abs(col) <= n - 1 ? sw[base + (size_t)(col + (n - 1))] : 0.0;
          };
          if (std::abs(m) <= n - 1) {  // eq 63a
            v = (at(rowp, l - 1) - at(rowp, l + 1)) * 0.5;
          } else if (m == n) {  // eq 63b (+)
            size_t src = offp + (size_t)(2 * n - 2) * nbp;
            v = (at(src, l - 1) + at(src, l + 1) + 2.0 * at(src, l)) * 0.5;
          } else {  // eq 63b (-)
            size_t src = offp;
            v = (at(src, l - 1) + at(src, l + 1) - 2.0 * at(src, l)) * 0.5;
          }
          sw[row + lc] = v;
          swt[off + (size_t)lc * nb + mr] = v;
        }
      }
    }
    in_.assign(d, cz());
    out_.assign(d, cz());
    fin.assign(d, cz());
    fout.assign(d, cz());
    for (int n = 0; n <= p; n++)
      for (int m = -n; m <= n; m++) {
        size_t i = idx(n, m);
        int am = std::abs(m);
        double fm = f2[n - am] * f2[n + am];
        int r4 = ((m % 4) + 4) % 4;
        C ei = r4 == 0 ? C{1.0, 0.0} : r4 == 1 ? C{0.0, -1.0} : r4 == 2 ? C{-1.0, 0.0} : C{0.0, 1.0};
        in_[i] = cscale(ei, fm);
        out_[i] = cscale(cconj(ei), 1.0 / fm);
        fin[i] = cscale(ei, 1.0 / fm);
        fout[i] = cscale(cconj(ei), fm);
      }
    b.assign(d, cz());
    f.assign(d, cz());
    tmp.assign(2 * p + 1, cz());
    tmp2.assign(2 * p + 1, cz());
    pha.assign(2 * p + 1, cz());
    phb.assign(2 * p + 1, cz());
    facr.assign(p + 1, 0.0);
  }

  // rotation.rs:262-301
  void m2l_pair(const double* cena, const double* cenb, const C* ma, const C* mb, C* la, C* lb,
                double eps2 = 0.0) {
    double r[3] = {cenb[0] - cena[0], cenb[1] - cena[1], cenb[2] - cena[2]};
    double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    double rinv = 1.0 / std::sqrt(r2);
    double alpha = std::atan2(r[1], r[0]);
    double s_beta = std::sqrt(r[0] * r[0] + r[1] * r[1]);
    double beta = std::atan2(s_beta, r[2]);
    if (eps2 > 0.0) {
      // softening extension: axial harmonic derivatives of the Plummer kernel,
      // j! R^j / s^{2j+1} (j!/R^{j+1} times (R/s)^{2j+1}, see theta_into)
      double R = std::sqrt(r2), s2 = r2 + eps2;
      facr[0] = 1.0 / std::sqrt(s2);
      for (int j = 1; j <= p; j++) facr[j] = facr[j - 1] * ((double)j * R / s2);
    } else {
      double irp = rinv;
      for (int j = 0; j <= p; j++) {
        facr[j] = fact[j] * irp;
        irp *= rinv;
      }
    }
    for (int m = -p; m <= p; m++) {
      double a = (double)m * alpha;
      pha[m + p] = C{std::cos(a), std::sin(a)};
      double bb = (double)m * beta;
      phb[m + p] = C{std::cos(bb), std::sin(bb)};
    }
    apply(ma, lb);
    for (size_t i = 0; i < pha.size(); i++) {
      int m = (int)i - p;
      double sgn = (((m % 2) + 2) % 2 == 0) ? 1.0 : -1.0;
      pha[i] = cscale(pha[i], sgn);
      phb[i] = cscale(cconj(phb[i]), sgn);
    }
    apply(mb, la);
  }

  // rotation.rs:319-435
  void apply(const C* mom, C* loc) {
    for (int k = 0; k <= p; k++) {
      int nb = 2 * k + 1;
      size_t k2 = (size_t)k * k;
      size_t off = block_off(k);
      int po = p - k;
      for (int c = 0; c < nb; c++) {
        C m = mom[k2 + c];
        C ph = pha[po + c];
        C iv = in_[k2 + c];
        double vr = m.re * ph.re + m.im * ph.im;
        double vi = m.re * ph.im - m.im * ph.re;
        tmp[c] = C{vr * iv.re - vi * iv.im, vr * iv.im + vi * iv.re};
      }
      for (int rr = 0; rr < nb; rr++) {
        double sr, si;
        dot_row(sw.data() + off + (size_t)rr * nb, tmp.data(), nb, &sr, &si);
        tmp2[rr] = C{sr, si};
      }
      for (int rr = 0; rr < nb; rr++) {
        C pbb = phb[po + rr];
        C s = tmp2[rr];
        tmp2[rr] = C{s.re * pbb.re - s.im * pbb.im, s.re * pbb.im + s.im * pbb.re};
      }
      for (int rr = 0; rr < nb; rr++) {
        double sr, si;
        dot_row(sw.data() + off + (size_t)rr * nb, tmp2.data(), nb, &sr, &si);
        C ov = out_[k2 + rr];
        b[k2 + rr] = C{sr * ov.re - si * ov.im, sr * ov.im + si * ov.re};
      }
    }
    for (int n = 0; n <= p; n++) {
      int nb = 2 * n + 1;
      size_t n2 = (size_t)n * n;
      for (int rr = 0; rr < nb; rr++) {
        int m = rr - n;
        int am = std::abs(m);
        double fr = 0.0, fi = 0.0;
        for (int k = am; k <= p - n; k++) {
          C bv = b[(size_t)k * k + (size_t)(k - m)];
          double fv = facr[n + k];
          fr += bv.re * fv;
          fi += bv.im * fv;
        }
        f[n2 + rr] = C{fr, fi};
      }
    }
    for (int n = 0; n <= p; n++) {
      int nb = 2 * n + 1;
      size_t off = block_off(n);
      size_t n2 = (size_t)n * n;
      int po = p - n;
      for (int c = 0; c < nb; c++) {
        C fv = f[n2 + c];
        C iv = fin[n2 + c];
        tmp[c] = C{fv.re * iv.re - fv.im * iv.im, fv.re * iv.im + fv.im * iv.re};
      }
      for (int rr = 0; rr < nb; rr++) {
        double sr, si;
        dot_row(swt.data() + off + (size_t)rr * nb, tmp.data(), nb, &sr, &si);
        tmp2[rr] = C{sr, si};
      }
      for (int rr = 0; rr < nb; rr++) {
        C pb = phb[po + rr];
        C s = tmp2[rr];
        tmp2[rr] = C{s.re * pb.re + s.im * pb.im, s.im * pb.re - s.re * pb.im};
      }
      for (int rr = 0; rr < nb; rr++) {
        double sr, si;
        dot_row(swt.data() + off + (size_t)rr * nb, tmp2.data(), nb, &sr, &si);
        C ov = fout[n2 + rr];
        C pv = pha[po + rr];
        C acc = C{sr * ov.re - si * ov.im, sr * ov.im + si * ov.re};
        C cur = loc[n2 + rr];
        C add = C{acc.re * pv.re - acc.im * pv.im, acc.re * pv.im + acc.im * pv.re};
        loc[n2 + rr] = C{cur.re + add.re, cur.im + add.im};
      }
    }
  }
};

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

template <class F>
void parallel_for(int nthreads, size_t n, F fn) {  // contiguous chunks [lo, hi)
  if (nthreads <= 1 || n < 2) {
    fn(0, (size_t)0, n);
    return;
  }
  size_t chunk = (n + nthreads - 1) / nthreads;
  std::vector<std::thread> ts;
  for (int t = 0; t < nthreads; t++) {
    size_t lo = (size_t)t * chunk, hi = std::min(n, lo + chunk);
    if (lo >= hi) break;
    ts.emplace_back([=, &fn] { fn(t, lo, hi); });
  }
  for (auto& th : ts) th.join();
}

// Loops whose iterations differ in cost. The timing build hands out small
// chunks from a shared counter (a thread may get several chunks, so fn must
// not assume one call per thread); the parity build keeps the contiguous
// split of parallel_for, so its results do not depend on timing.
template <class F>
void parallel_for_dyn(int nthreads, size_t n, F fn) {
#ifdef ORC_TIMING
  if (nthreads <= 1 || n < 2) {
    fn(0, (size_t)0, n);
    return;
  }
  const size_t chunk = std::max<size_t>(1, n / ((size_t)nthreads * 32));
  std::atomic<size_t> next{0};
  std::vector<std::thread> ts;
  for (int t = 0; t < nthreads; t++)
    ts.emplace_back([&, t] {
      for (;;) {
        const size_t lo = next.fetch_add(chunk);
        if (lo >= n) break;
        fn(t, lo, std::min(n, lo + chunk));
      }
    });
  for (auto& th : ts) th.join();
#else
  parallel_for(nthreads, n, fn);
#endif
}

}  // namespace

// ---------------------------------------------------------------- context
struct orc_ctx {
  // tree (Octree, tree.rs:60-70)
  std::vector<Cell> cells;
  std::vector<uint32_t> order;
  std::vector<uint64_t> keys_sorted;
  size_t n_leaf = 16;
  int mode = 0;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, inv_side = 0;
  uint32_t max_level = 0;
  std::vector<size_t> lvl_off, lvl_idx;  // levels_csr, tree.rs:236-254
  // last run
  int p = 0;
  std::vector<C> moments, locals_walk, locals;
  std::vector<double> pows, f, fmin, lap;
  std::vector<int32_t> parent;  // softened far field: L2P walks the ancestors
  double soft_mac = 0.0;        // softening guard of the MAC (orc_set_soft_mac)
  std::vector<uint32_t> m2l, p2p;
  Bodies bodies;
  double times[5] = {0, 0, 0, 0, 0};

  // tree.rs:152-227
  void geometry(const double* pos, const double* mass) {
    for (size_t ii = cells.size(); ii-- > 0;) {
      Cell& cell = cells[ii];
      bool is_leaf = true;
      for (int k = 0; k < 8; k++)
        if (cell.children[k] >= 0) is_leaf = false;
      double ms = 0.0;
      if (is_leaf) {
        for (size_t k = cell.first; k < cell.first + cell.count; k++) ms += mass[order[k]];
        double cen[3];
        if (mode == 0) {  // com_of_slots tree.rs:257-275
          double com[3] = {0.0, 0.0, 0.0};
          for (size_t k = cell.first; k < cell.first + cell.count; k++) {
            uint32_t id = order[k];
            for (int d = 0; d < 3; d++) com[d] += mass[id] * pos[3 * (size_t)id + d];
          }
          for (int d = 0; d < 3; d++) cen[d] = com[d] / ms;
        } else {  // sphere_of_slots tree.rs:278-286
          std::vector<Item> items;
          for (size_t k = cell.first; k < cell.first + cell.count; k++) {
            uint32_t id = order[k];
            items.push_back(Item{{pos[3 * (size_t)id], pos[3 * (size_t)id + 1], pos[3 * (size_t)id + 2]}, 0.0});
          }
          double r;
          ritter(items, cen, &r);
        }
        double rho = 0.0;
        for (size_t k = cell.first; k < cell.first + cell.count; k++)
          rho = std::fmax(rho, dist(pos + 3 * (size_t)order[k], cen));
        cell.mass = ms;
        std::memcpy(cell.cen, cen, sizeof cen);
        cell.rho = rho;
      } else {
        for (int k = 0; k < 8; k++)
          if (cell.children[k] >= 0) ms += cells[cell.children[k]].mass;
        double cen[3];
        if (mode == 0) {
          double com[3] = {0.0, 0.0, 0.0};
          for (int k = 0; k < 8; k++) {
            if (cell.children[k] < 0) continue;
            const Cell& ch = cells[cell.children[k]];
            for (int d = 0; d < 3; d++) com[d] += ch.mass * ch.cen[d];
          }
          for (int d = 0; d < 3; d++) cen[d] = com[d] / ms;
        } else {
          std::vector<Item> items;
          for (int k = 0; k < 8; k++)
            if (cell.children[k] >= 0) {
              const Cell& ch = cells[cell.children[k]];
              items.push_back(Item{{ch.cen[0], ch.cen[1], ch.cen[2]}, ch.rho});
            }
          double r;
          ritter(items, cen, &r);
        }
        double rho = 0.0;
        for (int k = 0; k < 8; k++) {
          if (cell.children[k] < 0) continue;
          const Cell& ch = cells[cell.children[k]];
          rho = std::fmax(rho, ch.rho + dist(ch.cen, cen));
        }
        cell.mass = ms;
        std::memcpy(cell.cen, cen, sizeof cen);
        cell.rho = rho;
      }
    }
  }

  void levels_csr() {
    uint32_t maxl = 0;
    for (auto& c : cells) maxl = std::max(maxl, c.level);
    max_level = maxl;
    lvl_off.assign(maxl + 2, 0);
    for (auto& c : cells) lvl_off[c.level + 1]++;
    for (size_t i = 1; i < lvl_off.size(); i++) lvl_off[i] += lvl_off[i - 1];
    lvl_idx.assign(cells.size(), 0);
    std::vector<size_t> pos(lvl_off.begin(), lvl_off.begin() + maxl + 1);
    for (size_t i = 0; i < cells.size(); i++) lvl_idx[pos[cells[i].level]++] = i;
  }
};

namespace {

// moments.rs:30-53
void p2m_cell(const orc_ctx& t, size_t c, int p, C* moments, C* u) {
  size_t d = dimp(p);
  const Cell& cell = t.cells[c];
  size_t off = c * d;
  const double* z = cell.cen;
  for (size_t k = cell.first; k < cell.first + cell.count; k++) {
    double r[3] = {t.bodies.x[k] - z[0], t.bodies.y[k] - z[1], t.bodies.z[k] - z[2]};
    upsilon_into(r, p, u);
    double m = t.bodies.m[k];
    for (size_t i = 0; i < d; i++) moments[off + i] = cadd(moments[off + i], cscale(u[i], m));
  }
}

// moments.rs:56-90
void m2m_cell(const orc_ctx& t, size_t c, int p, C* moments, C* u) {
  size_t d = dimp(p);
  const Cell& cell = t.cells[c];
  size_t off = c * d;
  const double* z = cell.cen;
  for (int o = 0; o < 8; o++) {
    int ch = cell.children[o];
    if (ch < 0) continue;
    const Cell& child = t.cells[ch];
    double dd[3] = {child.cen[0] - z[0], child.cen[1] - z[1], child.cen[2] - z[2]};
    upsilon_into(dd, p, u);
    size_t coff = (size_t)ch * d;
    for (int n = 0; n <= p; n++)
      for (int m = -n; m <= n; m++) {
        C s = cz();
        for (int k = 0; k <= n; k++) {
          int nk = n - k;
          int lo = std::max(-k, m - nk);
          int hi = std::min(k, m + nk);
          for (int l = lo; l <= hi; l++) s = cadd(s, cmul(u[idx(k, l)], moments[coff + idx(nk, m - l)]));
        }
        moments[off + idx(n, m)] = cadd(moments[off + idx(n, m)], s);
      }
  }
}

// moments.rs:96-146: level sweep, deepest level first (bit-identical to serial).
void compute_moments(const orc_ctx& t, int p, C* moments, int nthreads) {
  for (size_t l = t.lvl_off.size() - 1; l-- > 0;) {
    size_t a = t.lvl_off[l], b = t.lvl_off[l + 1];
    parallel_for(nthreads, b - a, [&](int, size_t lo, size_t hi) {
      std::vector<C> u(dimp(p));
      for (size_t i = lo; i < hi; i++) {
        size_t c = t.lvl_idx[a + i];
        if (t.cells[c].leaf)
          p2m_cell(t, c, p, moments, u.data());
        else
          m2m_cell(t, c, p, moments, u.data());
      }
    });
  }
}

// moments.rs:150-168
std::vector<double> multipole_power(const C* moments, size_t ncells, int p) {
  size_t d = dimp(p);
  std::vector<double> fact = factorials(p);
  std::vector<double> pp(ncells * (p + 1), 0.0);
  for (size_t c = 0; c < ncells; c++)
    for (int n = 0; n <= p; n++) {
      double s = 0.0;
      for (int m = -n; m <= n; m++) {
        int am = std::abs(m);
        C mc = moments[c * d + idx(n, m)];
        double w = fact[n - am] * fact[n + am];
        s += w * (mc.re * mc.re + mc.im * mc.im);
      }
      pp[c * (p + 1) + n] = std::sqrt(s);
    }
  return pp;
}

// mac.rs:99-147 — stack DFS, children pushed 0..7 (popped 7..0).
double f_walk(const orc_ctx& t, size_t i, std::vector<size_t>& stack) {
  const Bodies& b = t.bodies;
  double pb[3] = {b.x[i], b.y[i], b.z[i]};
  stack.clear();
  stack.push_back(0);
  double fi = 0.0;
  while (!stack.empty()) {
    size_t c = stack.back();
    stack.pop_back();
    const Cell& cell = t.cells[c];
    if (cell.leaf) {
      size_t end = cell.first + cell.count;
      size_t below = std::min(i, end);
      size_t above = std::max(cell.first, i + 1);
      for (size_t k = cell.first; k < below; k++) {
        double dx = pb[0] - b.x[k], dy = pb[1] - b.y[k], dz = pb[2] - b.z[k];
        double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 > 0.0) fi += b.m[k] / r2;
      }
      for (size_t k = above; k < end; k++) {
        double dx = pb[0] - b.x[k], dy = pb[1] - b.y[k], dz = pb[2] - b.z[k];
        double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 > 0.0) fi += b.m[k] / r2;
      }
    } else {
      double dx = pb[0] - cell.cen[0], dy = pb[1] - cell.cen[1], dz = pb[2] - cell.cen[2];
      double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 > cell.rho * cell.rho) {
        fi += cell.mass / d2;
      } else {
        for (int k = 0; k < 8; k++)
          if (cell.children[k] >= 0) stack.push_back((size_t)cell.children[k]);
      }
    }
  }
  return fi;
}

// mac.rs:151-173 (min over a subtree; order-free).
std::vector<double> fmin_of(const std::vector<double>& f, const orc_ctx& t) {
  size_t nc = t.cells.size();
  std::vector<double> fm(nc, INFINITY);
  for (size_t c = nc; c-- > 0;) {
    const Cell& cell = t.cells[c];
    double m = INFINITY;
    if (cell.leaf) {
      for (size_t k = cell.first; k < cell.first + cell.count; k++) m = std::fmin(m, f[k]);
    } else {
      for (int k = 0; k < 8; k++)
        if (cell.children[k] >= 0) m = std::fmin(m, fm[cell.children[k]]);
    }
    fm[c] = m;
  }
  return fm;
}

struct Mac {
  const orc_ctx* t;
  int p;
  double eps;
  double theta2;
  bool use_err;
  const double* pows;
  const double* binom;
  const double* fmin;
  double soft2;     // softening^2
  double soft_mac;  // softening guard tau (0 = off)

  // mac.rs:23-61
  bool error_estimate(size_t a, const Cell& cb, double ca_mass, double ca_rho, double r,
                      double* out) const {
    if (r == 0.0) return false;
    double r2 = r * r;
    double rp = 1.0;
    for (int i = 0; i < p / 2; i++) rp *= r2;
    if (p % 2 == 1) rp *= r;
    double rpows[64];
    rpows[0] = 1.0;
    for (int j = 1; j <= p; j++) rpows[j] = rpows[j - 1] * cb.rho;
    size_t base = a * (p + 1);
    double s = 0.0;
    for (int k = 0; k <= p; k++) s += binom[k] * pows[base + k] * rpows[p - k];
    double e = s / (ca_mass * rp);
    double denom = ca_rho + cb.rho;
    if (denom == 0.0) {
      *out = 0.0;
      return true;
    }
    *out = 8.0 * std::fmax(ca_rho, cb.rho) / denom * e;
    return true;
  }

  // mac.rs:64-95 (+ theta_max extension: r2*theta2 <= sum^2 rejects)
  bool accept(size_t a, size_t b) const {
    const Cell& ca = t->cells[a];
    const Cell& cb = t->cells[b];
    double dx = cb.cen[0] - ca.cen[0], dy = cb.cen[1] - ca.cen[1], dz = cb.cen[2] - ca.cen[2];
    double r2 = dx * dx + dy * dy + dz * dz;
    double sum_rho = ca.rho + cb.rho;
    if (r2 * theta2 <= sum_rho * sum_rho) return false;
    // softening guard (extension): the softened far field's leading residual,
    // eps^2 (rhoA + rhoB)^3 / r^5 relative to the pair's field, above tau rejects
    if (soft_mac > 0.0 && soft2 > 0.0 &&
        soft2 * (sum_rho * sum_rho * sum_rho) > soft_mac * (r2 * r2 * std::sqrt(r2)))
      return false;
    if (!use_err) return true;
    double r = std::sqrt(r2);
    double est;
    if (!error_estimate(a, cb, ca.mass, ca.rho, r, &est)) return false;
    if (est == 0.0) return true;
    return est * ca.mass / r2 < eps * fmin[b];
  }
};

enum Action { ACT_SELF_LEAF, ACT_SELF_SPLIT, ACT_M2L, ACT_P2P, ACT_SPLIT_B, ACT_SPLIT_A };

// walk.rs:94-190 decision (and child pushes, in reference order).
template <class Push>
Action decide(const orc_ctx& t, const Mac& mac, size_t a, size_t b, Push push) {
  const Cell& ca = t.cells[a];
  const Cell& cb = t.cells[b];
  if (a == b) {
    if (ca.leaf) return ACT_SELF_LEAF;
    for (int i = 0; i < 8; i++) {
      if (ca.children[i] < 0) continue;
      for (int j = i; j < 8; j++)
        if (ca.children[j] >= 0) push((size_t)ca.children[i], (size_t)ca.children[j]);
    }
    return ACT_SELF_SPLIT;
  }
  if (mac.accept(a, b) && mac.accept(b, a)) return ACT_M2L;
  if (ca.leaf && cb.leaf) return ACT_P2P;
  if (ca.leaf) {
    for (int j = 0; j < 8; j++)
      if (cb.children[j] >= 0) push(a, (size_t)cb.children[j]);
    return ACT_SPLIT_B;
  }
  if (cb.leaf) {
    for (int i = 0; i < 8; i++)
      if (ca.children[i] >= 0) push((size_t)ca.children[i], b);
    return ACT_SPLIT_A;
  }
  if (ca.rho >= cb.rho) {
    for (int i = 0; i < 8; i++)
      if (ca.children[i] >= 0) push((size_t)ca.children[i], b);
    return ACT_SPLIT_A;
  }
  for (int j = 0; j < 8; j++)
    if (cb.children[j] >= 0) push(a, (size_t)cb.children[j]);
  return ACT_SPLIT_B;
}

}  // namespace

// ================================================================== C API
extern "C" {

void orc_plummer(size_t n, const double center[3], uint64_t seed, double* pos, double* mass) {
  Rng rng{seed};  // models.rs:35-58
  double m = 1.0 / (double)n;
  for (size_t i = 0; i < n; i++) {
    double u = rng.unit();
    double w = std::pow(u, 2.0 / 3.0);
    double r = std::sqrt(w / (1.0 - w));
    double z = 2.0 * rng.unit() - 1.0;
    double phi = 2.0 * PI * rng.unit();
    double s = std::sqrt(1.0 - z * z);
    pos[3 * i + 0] = center[0] + r * s * std::cos(phi);
    pos[3 * i + 1] = center[1] + r * s * std::sin(phi);
    pos[3 * i + 2] = center[2] + r * z;
    mass[i] = m;
  }
}

void orc_hernquist(size_t n, const double center[3], uint64_t seed, double* pos, double* mass) {
  Rng rng{seed};  // models.rs:61-82
  double m = 1.0 / (double)n;
  for (size_t i = 0; i < n; i++) {
    double s = std::sqrt(rng.unit());
    double r = s / (1.0 - s);
    double z = 2.0 * rng.unit() - 1.0;
    double phi = 2.0 * PI * rng.unit();
    double sp = std::sqrt(1.0 - z * z);
    pos[3 * i + 0] = center[0] + r * sp * std::cos(phi);
    pos[3 * i + 1] = center[1] + r * sp * std::sin(phi);
    pos[3 * i + 2] = center[2] + r * z;
    mass[i] = m;
  }
}

double orc_rng_unit_first(uint64_t seed, size_t k, double* out) {
  Rng rng{seed};
  double last = 0.0;
  for (size_t i = 0; i < k; i++) last = out[i] = rng.unit();
  return last;
}

uint64_t orc_dilate(uint64_t x) { return dilate(x); }
uint64_t orc_morton_key(double x, double y, double z, const double lo[3], double inv_side) {
  return morton_key(x, y, z, lo, inv_side);
}

void orc_theta_table(const double r[3], int p, double* out) {
  std::vector<C> t(dimp(p), cz());
  theta_into(r, p, t.data());
  std::memcpy(out, t.data(), t.size() * sizeof(C));
}
void orc_upsilon_table(const double r[3], int p, double* out) {
  std::vector<C> u(dimp(p), cz());
  upsilon_into(r, p, u.data());
  std::memcpy(out, u.data(), u.size() * sizeof(C));
}
void orc_factorials(int p, double* out) {
  auto f = factorials(p);
  std::memcpy(out, f.data(), f.size() * sizeof(double));
}
void orc_binomials(int p, double* out) {
  auto b = binomials(p);
  std::memcpy(out, b.data(), b.size() * sizeof(double));
}

static void trace_pair(int p, const double* zs, const double* zt, double q, double mass,
                       double soft2, C* f, double* lap);
static void l2p_trace(const double* cen, const double* x, const double* L10, double* pot,
                      double* a);

void orc_m2l_mutual_soft(int p, const double cen_a[3], const double cen_b[3], const double* ma,
                         const double* mb, double eps2, double* la, double* lb) {
  M2lScratch s(p);
  m2l_mutual(p, cen_a, cen_b, (const C*)ma, (const C*)mb, (C*)la, (C*)lb, s, eps2);
}

void orc_soft_trace_pair(int p, const double cen_src[3], const double cen_dst[3], double q,
                         double mass, double eps2, double* f_dst, double* lap10_dst) {
  trace_pair(p, cen_src, cen_dst, q, mass, eps2, (C*)f_dst, lap10_dst);
}

void orc_l2p_trace(const double cen[3], const double pos[3], const double* lap10, double* pot,
                   double acc[3]) {
  l2p_trace(cen, pos, lap10, pot, acc);
}

void orc_m2l_mutual(int p, const double cen_a[3], const double cen_b[3], const double* ma,
                    const double* mb, double* la, double* lb) {
  M2lScratch s(p);
  m2l_mutual(p, cen_a, cen_b, (const C*)ma, (const C*)mb, (C*)la, (C*)lb, s);
}
void orc_m2l_rotation(int p, const double cen_a[3], const double cen_b[3], const double* ma,
                      const double* mb, double* la, double* lb) {
  RotCtx r(p);
  r.m2l_pair(cen_a, cen_b, (const C*)ma, (const C*)mb, (C*)la, (C*)lb);
}
size_t orc_rot_blocksum(int p) { return blocksum(p); }
void orc_rot_swap(int p, double* out) {
  RotCtx r(p);
  std::memcpy(out, r.sw.data(), r.sw.size() * sizeof(double));
}
void orc_l2l(int p, const double parent_cen[3], const double child_cen[3], const double* fpar,
             double* fchild) {
  std::vector<C> u(dimp(p));
  l2l_into(p, parent_cen, child_cen, (const C*)fpar, (C*)fchild, u.data());
}
void orc_l2p(int p, const double cen[3], const double pos[3], const double* f, double* pot,
             double acc[3]) {
  std::vector<C> u(dimp(p));
  l2p_into(p, cen, pos, (const C*)f, pot, acc, u.data());
}

// direct.rs:6-31 (+ softening extension)
void orc_direct(size_t n, const double* pos, const double* mass, double softening, double* a,
                double* pot) {
  double s2 = softening * softening;
  std::memset(a, 0, 3 * n * sizeof(double));
  std::memset(pot, 0, n * sizeof(double));
  for (size_t i = 0; i < n; i++)
    for (size_t j = i + 1; j < n; j++) {
      const double* pa = pos + 3 * i;
      const double* pb = pos + 3 * j;
      double dx = pb[0] - pa[0], dy = pb[1] - pa[1], dz = pb[2] - pa[2];
      double r2 = dx * dx + dy * dy + dz * dz;
      if (s2 > 0.0) r2 = r2 + s2;
      double inv = 1.0 / std::sqrt(r2);
      double inv3 = inv * inv * inv;
      a[3 * i + 0] += mass[j] * inv3 * dx;
      a[3 * i + 1] += mass[j] * inv3 * dy;
      a[3 * i + 2] += mass[j] * inv3 * dz;
      a[3 * j + 0] -= mass[i] * inv3 * dx;
      a[3 * j + 1] -= mass[i] * inv3 * dy;
      a[3 * j + 2] -= mass[i] * inv3 * dz;
      pot[i] += mass[j] * inv;
      pot[j] += mass[i] * inv;
    }
}

// crates/bench/src/main.rs:34-78 inner loop, per target (+ softening, pot).
void orc_direct_targets(size_t n, const double* pos, const double* mass, double softening,
                        size_t nt, const uint32_t* ids, double* acc, double* pot, double* f,
                        int nthreads) {
  double s2 = softening * softening;
  parallel_for(nthreads, nt, [&](int, size_t lo, size_t hi) {
    for (size_t t = lo; t < hi; t++) {
      size_t i = ids[t];
      double ax = 0.0, ay = 0.0, az = 0.0, ff = 0.0, ph = 0.0;
      const double* pi = pos + 3 * i;
      for (size_t j = 0; j < n; j++) {
        if (j == i) continue;
        const double* pj = pos + 3 * j;
        double dx = pj[0] - pi[0], dy = pj[1] - pi[1], dz = pj[2] - pi[2];
        double r2 = dx * dx + dy * dy + dz * dz;
        if (s2 > 0.0) r2 = r2 + s2;
        double inv = 1.0 / std::sqrt(r2);
        double inv2 = inv * inv;
        double inv3 = inv2 * inv;
        ax += mass[j] * inv3 * dx;
        ay += mass[j] * inv3 * dy;
        az += mass[j] * inv3 * dz;
        ff += mass[j] * inv2;
        ph += mass[j] * inv;
      }
      acc[3 * t] = ax;
      acc[3 * t + 1] = ay;
      acc[3 * t + 2] = az;
      if (pot) pot[t] = ph;
      if (f) f[t] = ff;
    }
  });
}

void orc_sample_ids(size_t n, size_t samp, uint64_t seed, uint32_t* out) {
  uint64_t rng = seed;  // crates/bench/src/main.rs:41-47
  for (size_t s = 0; s < samp; s++) {
    rng = rng * 6364136223846793005ull + 1442695040888963407ull;
    out[s] = (uint32_t)((rng >> 33) % (uint64_t)n);
  }
}

orc_ctx* orc_new(void) { return new orc_ctx(); }
void orc_free(orc_ctx* c) { delete c; }

// tree.rs:94-148 (sort: stable by key, so ties keep ascending particle id)
int orc_build(orc_ctx* t, size_t n, const double* pos, const double* mass, int n_leaf,
              int center_mode) {
  if (n == 0) return 1;
  t->n_leaf = (size_t)n_leaf;
  t->mode = center_mode;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      lo[d] = std::fmin(lo[d], pos[3 * i + d]);
      hi[d] = std::fmax(hi[d], pos[3 * i + d]);
    }
  double side = std::fmax(std::fmax(std::fmax(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]),
                          std::numeric_limits<double>::min());
  double inv_side = 1.0 / side;
  std::memcpy(t->lo, lo, sizeof lo);
  std::memcpy(t->hi, hi, sizeof hi);
  t->inv_side = inv_side;
  std::vector<uint64_t> keys(n);
  for (size_t i = 0; i < n; i++)
    keys[i] = morton_key(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], lo, inv_side);
  std::vector<uint32_t> ix(n);
  for (size_t i = 0; i < n; i++) ix[i] = (uint32_t)i;
  std::stable_sort(ix.begin(), ix.end(), [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
  t->order = ix;
  t->keys_sorted.resize(n);
  for (size_t k = 0; k < n; k++) t->keys_sorted[k] = keys[ix[k]];
  t->cells.clear();
  build_rec(t->keys_sorted, t->cells, 0, 0, n, t->n_leaf, BITS);
  t->geometry(pos, mass);
  t->levels_csr();
  return 0;
}

size_t orc_ncells(const orc_ctx* t) { return t->cells.size(); }
void orc_set_soft_mac(orc_ctx* t, double tau) { t->soft_mac = tau; }
uint32_t orc_max_level(const orc_ctx* t) { return t->max_level; }

void orc_get_tree(const orc_ctx* t, uint32_t* order, uint64_t* keys_sorted, uint32_t* first,
                  uint32_t* count, int32_t* children, uint8_t* leaf, uint32_t* level, double* cen,
                  double* mass, double* rho) {
  size_t n = t->order.size();
  if (order) std::memcpy(order, t->order.data(), n * sizeof(uint32_t));
  if (keys_sorted) std::memcpy(keys_sorted, t->keys_sorted.data(), n * sizeof(uint64_t));
  for (size_t c = 0; c < t->cells.size(); c++) {
    const Cell& cl = t->cells[c];
    if (first) first[c] = (uint32_t)cl.first;
    if (count) count[c] = (uint32_t)cl.count;
    if (children)
      for (int k = 0; k < 8; k++) children[8 * c + k] = cl.children[k];
    if (leaf) leaf[c] = cl.leaf ? 1 : 0;
    if (level) level[c] = cl.level;
    if (cen)
      for (int d = 0; d < 3; d++) cen[3 * c + d] = cl.cen[d];
    if (mass) mass[c] = cl.mass;
    if (rho) rho[c] = cl.rho;
  }
}

void orc_get_bbox(const orc_ctx* t, double lo[3], double hi[3], double* inv_side) {
  std::memcpy(lo, t->lo, 3 * sizeof(double));
  std::memcpy(hi, t->hi, 3 * sizeof(double));
  *inv_side = t->inv_side;
}

// Softening extension, far field: the Taylor expansions of the Plummer
// kernel K = (r^2 + eps^2)^{-1/2} about the source and the target centre
// have, beside the harmonic (solid-harmonic) parts that M2L and L2P carry,
// trace terms |d|^{2k} Y_{n-2k}(d) Lap^k K (Lap K = -3 eps^2 s^{-5}). Kept:
//  * source side, the field (q/6) Lap K of every M2L source (q = sum m |x - z|^2,
//    summed from the bodies), Taylor-expanded at the target centre to second
//    order: value into F_0^0, gradient into F_1^m, the traceless part of
//    the Hessian into F_2^m, its trace into the target-side Laplacian below
//    (l2p_into gives pot = F00 - F10 dz - Re F11 dx - Im F11 dy + ...);
//  * target side, the trace parts of the Taylor series of the cell's own
//    M2L far field to fourth order,
//      |d|^2 L / 6 + |d|^2 (d.G) / 10 + |d|^2 (d.H.d) / 28 - |d|^4 tr(H) / 280,
//    from L = Lap Phi, G = grad L and H = Hess L at the centre (the monopole
//    fields of the sources plus the source-side fields above; 10 doubles per
//    cell). Shifting a centre turns part of this polynomial into harmonic
//    terms, so it is not moved through L2L: each body sums it over its leaf
//    and every ancestor, each about its own centre (the exact re-expansion).
// Derivatives used, with s^2 = r^2 + eps^2:
//   grad Lap K = 15 eps^2 s^-7 r;  Hess Lap K = 15 eps^2 (s^-7 I - 7 s^-9 r r^T);
//   Lap^2 K = 15 eps^2 (3 s^-7 - 7 r^2 s^-9);
//   grad Lap^2 K = 15 eps^2 (63 r^2 s^-11 - 35 s^-9) r.
constexpr int SOFT_LAP_WORDS = 10;  // L, G (3), H (xx, yy, zz, xy, xz, yz)

// One directed M2L pair's trace terms: source-side field into the target's
// local expansion f, the target-side Laplacian field into lap[0..9].
static void trace_pair(int p, const double* zs, const double* zt, double q, double mass,
                       double soft2, C* f, double* lap) {
  double r[3] = {zt[0] - zs[0], zt[1] - zs[1], zt[2] - zs[2]};
  double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  double s2 = r2 + soft2;
  double is = 1.0 / std::sqrt(s2);
  double is2 = is * is;
  double is5 = is2 * is2 * is, is7 = is5 * is2, is9 = is7 * is2, is11 = is9 * is2;
  double lk = -3.0 * soft2 * is5;                          // Lap K
  double gk = 15.0 * soft2 * is7;                          // grad Lap K / r (and H's I part)
  double hk = 105.0 * soft2 * is9;                         // H's r r^T part
  double l4 = 15.0 * soft2 * (3.0 * is7 - 7.0 * r2 * is9);  // Lap^2 K
  double g4 = 15.0 * soft2 * (63.0 * r2 * is11 - 35.0 * is9);  // grad Lap^2 K / r
  double w = q / 6.0;
  // target side: Laplacian field of the source (monopole + source-side field)
  lap[0] += mass * lk + w * l4;
  double gg = mass * gk + w * g4;
  lap[1] += gg * r[0];
  lap[2] += gg * r[1];
  lap[3] += gg * r[2];
  double mh = mass * hk;
  lap[4] += mass * gk - mh * r[0] * r[0];
  lap[5] += mass * gk - mh * r[1] * r[1];
  lap[6] += mass * gk - mh * r[2] * r[2];
  lap[7] -= mh * r[0] * r[1];
  lap[8] -= mh * r[0] * r[2];
  lap[9] -= mh * r[1] * r[2];
  // source side: w Lap K, its gradient, and the traceless part of its Hessian
  // (w/2) d^T (Hess Lap K - tr/3) d = -(w/2) hk (d.r)^2 + (w/2) hk r^2 |d|^2 / 3
  double v = w * lk, gw = w * gk;
  f[idx(0, 0)].re += v;
  if (p >= 1) {
    f[idx(1, 0)].re -= gw * r[2];
    f[idx(1, 1)].re -= gw * r[0];
    f[idx(1, 1)].im -= gw * r[1];
    f[idx(1, -1)].re += gw * r[0];
    f[idx(1, -1)].im -= gw * r[1];
  }
  if (p >= 2) {
    double hw = 0.5 * w * hk;
    double e = 2.0 * hw * (r2 / 3.0 - r[2] * r[2]);   // 2 S_zz
    double ar = -2.0 * hw * r[0] * r[2];             // 2 S_xz
    double ai = -2.0 * hw * r[1] * r[2];             // 2 S_yz
    double br = -2.0 * hw * (r[0] * r[0] - r[1] * r[1]);  // 2 (S_xx - S_yy)
    double bi = -4.0 * hw * r[0] * r[1];             // 4 S_xy
    f[idx(2, 0)].re += e;
    f[idx(2, 1)].re += ar;
    f[idx(2, 1)].im += ai;
    f[idx(2, -1)].re -= ar;
    f[idx(2, -1)].im += ai;
    f[idx(2, 2)].re += br;
    f[idx(2, 2)].im += bi;
    f[idx(2, -2)].re += br;
    f[idx(2, -2)].im -= bi;
  }
}

// Target-side trace polynomial of one cell at L2P (d = x - cen):
// pot += |d|^2 L/6 + |d|^2 (d.G)/10 + |d|^2 (d.H.d)/28 - |d|^4 tr(H)/280,
// a += its gradient.
static void l2p_trace(const double* cen, const double* x, const double* L10, double* pot,
                      double* a) {
  double d[3] = {x[0] - cen[0], x[1] - cen[1], x[2] - cen[2]};
  double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double L = L10[0];
  double dg = d[0] * L10[1] + d[1] * L10[2] + d[2] * L10[3];
  double hd[3] = {L10[4] * d[0] + L10[7] * d[1] + L10[8] * d[2],
                  L10[7] * d[0] + L10[5] * d[1] + L10[9] * d[2],
                  L10[8] * d[0] + L10[9] * d[1] + L10[6] * d[2]};
  double dhd = d[0] * hd[0] + d[1] * hd[1] + d[2] * hd[2];
  double tr = L10[4] + L10[5] + L10[6];
  *pot += d2 * L / 6.0 + d2 * dg / 10.0 + d2 * dhd / 28.0 - d2 * d2 * tr / 280.0;
  for (int k = 0; k < 3; k++)
    a[k] += d[k] * L / 3.0 + (2.0 * d[k] * dg + d2 * L10[1 + k]) / 10.0 +
            (d[k] * dhd + d2 * hd[k]) / 14.0 - d2 * d[k] * tr / 70.0;
}

static void softened_trace(orc_ctx& t, int p, double soft2) {
  const size_t nc = t.cells.size(), d = dimp(p);
  std::vector<double> q(nc, 0.0);
  for (size_t c = 0; c < nc; c++) {
    const Cell& cell = t.cells[c];
    double acc = 0.0;
    for (size_t k = cell.first; k < cell.first + cell.count; k++) {
      double dx = t.bodies.x[k] - cell.cen[0], dy = t.bodies.y[k] - cell.cen[1],
             dz = t.bodies.z[k] - cell.cen[2];
      acc += t.bodies.m[k] * (dx * dx + dy * dy + dz * dz);
    }
    q[c] = acc;
  }
  t.lap.assign(SOFT_LAP_WORDS * nc, 0.0);  // per cell: L, G, H of its own M2L sources
  auto add = [&](size_t src, size_t dst) {
    trace_pair(p, t.cells[src].cen, t.cells[dst].cen, q[src], t.cells[src].mass, soft2,
               t.locals.data() + dst * d, &t.lap[SOFT_LAP_WORDS * dst]);
  };
  for (size_t k = 0; k + 1 < t.m2l.size(); k += 2) {
    add(t.m2l[k], t.m2l[k + 1]);
    add(t.m2l[k + 1], t.m2l[k]);
  }
  t.parent.assign(nc, -1);
  for (size_t c = 0; c < nc; c++)
    for (int k = 0; k < 8; k++)
      if (t.cells[c].children[k] >= 0) t.parent[t.cells[c].children[k]] = (int32_t)c;
}

// fmm.rs:69-142
int orc_run(orc_ctx* t, const double* pos, const double* mass, int p, double eps_mac,
            double theta_max, double softening, int nthreads, double* accel, double* pot) {
  if (t->cells.empty() || p < 0 || p > 60) return 1;
  size_t n = t->order.size();
  size_t nc = t->cells.size();
  size_t d = dimp(p);
  double soft2 = softening * softening;
  t->p = p;
  // Bodies::from_parts (bodies.rs:27-47)
  t->bodies.resize(n);
  for (size_t k = 0; k < n; k++) {
    uint32_t id = t->order[k];
    t->bodies.x[k] = pos[3 * (size_t)id];
    t->bodies.y[k] = pos[3 * (size_t)id + 1];
    t->bodies.z[k] = pos[3 * (size_t)id + 2];
    t->bodies.m[k] = mass[id];
  }
  double t0 = now_ms();
  t->moments.assign(nc * d, cz());
  compute_moments(*t, p, t->moments.data(), nthreads);
  double t1 = now_ms();
  bool use_err = std::isfinite(eps_mac);
  std::vector<double> binom = binomials(p);
  if (use_err) {
    t->pows = multipole_power(t->moments.data(), nc, p);
    t->f.assign(n, 0.0);
    parallel_for_dyn(nthreads, n, [&](int, size_t lo, size_t hi) {
      std::vector<size_t> stack;
      stack.reserve(256);
      for (size_t i = lo; i < hi; i++) t->f[i] = f_walk(*t, i, stack);
    });
    t->fmin = fmin_of(t->f, *t);
  } else {
    t->pows.clear();
    t->f.clear();
    t->fmin.assign(nc, INFINITY);
  }
  double t2 = now_ms();
  Mac mac{t,           p,           eps_mac,         theta_max * theta_max, use_err,
          t->pows.data(), binom.data(), t->fmin.data(), soft2,                 t->soft_mac};
  t->locals.assign(nc * d, cz());
  t->m2l.clear();
  t->p2p.clear();
  bool use_rot = p >= ROT_P;
  const C* mom = t->moments.data();
  if (nthreads <= 1) {
    // dual_walk (walk.rs:238-255): serial stack, inline mutual M2L / P2P.
    M2lScratch scr(p);
    RotCtx* rot = use_rot ? new RotCtx(p) : nullptr;
    std::vector<std::pair<size_t, size_t>> stack;
    stack.reserve(1024);
    stack.emplace_back(0, 0);
    C* loc = t->locals.data();
    while (!stack.empty()) {
      auto [a, b] = stack.back();
      stack.pop_back();
      Action act = decide(*t, mac, a, b, [&](size_t x, size_t y) { stack.emplace_back(x, y); });
      const Cell& ca = t->cells[a];
      const Cell& cb = t->cells[b];
      if (act == ACT_SELF_LEAF) {
        t->p2p.push_back((uint32_t)a);
        t->p2p.push_back((uint32_t)a);
        p2p_self(t->bodies, ca.first, ca.count, soft2);
      } else if (act == ACT_M2L) {
        t->m2l.push_back((uint32_t)a);
        t->m2l.push_back((uint32_t)b);
        if (rot)
          rot->m2l_pair(ca.cen, cb.cen, mom + a * d, mom + b * d, loc + a * d, loc + b * d, soft2);
        else
          m2l_mutual(p, ca.cen, cb.cen, mom + a * d, mom + b * d, loc + a * d, loc + b * d, scr, soft2);
      } else if (act == ACT_P2P) {
        t->p2p.push_back((uint32_t)a);
        t->p2p.push_back((uint32_t)b);
        p2p_pair(t->bodies, ca.first, ca.count, cb.first, cb.count, soft2);
      }
    }
    delete rot;
  } else {
    // dual_walk_par analogue (walk.rs:342-622): expand the top of the walk
    // serially, split the pending pairs over threads that each finish their
    // sub-walks (recording M2L, P2P inline into private bodies), then a
    // threaded mutual M2L replay into private locals and a reduce.
    std::vector<std::pair<size_t, size_t>> seed;
    seed.emplace_back(0, 0);
    std::vector<uint32_t> top_m2l, top_p2p;
    Bodies& B = t->bodies;
    while (!seed.empty() && seed.size() < (size_t)nthreads * 64) {
      std::vector<std::pair<size_t, size_t>> next;
      for (auto [a, b] : seed) {
        Action act = decide(*t, mac, a, b, [&](size_t x, size_t y) { next.emplace_back(x, y); });
        const Cell& ca = t->cells[a];
        const Cell& cb = t->cells[b];
        if (act == ACT_SELF_LEAF) {
          top_p2p.push_back((uint32_t)a);
          top_p2p.push_back((uint32_t)a);
          p2p_self(B, ca.first, ca.count, soft2);
        } else if (act == ACT_M2L) {
          top_m2l.push_back((uint32_t)a);
          top_m2l.push_back((uint32_t)b);
        } else if (act == ACT_P2P) {
          top_p2p.push_back((uint32_t)a);
          top_p2p.push_back((uint32_t)b);
          p2p_pair(B, ca.first, ca.count, cb.first, cb.count, soft2);
        }
      }
      seed.swap(next);
    }
    std::vector<std::vector<uint32_t>> tm2l(nthreads), tp2p(nthreads);
    std::vector<Bodies> tb(nthreads);
    parallel_for_dyn(nthreads, seed.size(), [&](int th, size_t lo, size_t hi) {
      Bodies& my = tb[th];
      if (my.x.empty()) {  // first chunk of this thread: private accumulators
        my = B;
        std::fill(my.ax.begin(), my.ax.end(), 0.0);
        std::fill(my.ay.begin(), my.ay.end(), 0.0);
        std::fill(my.az.begin(), my.az.end(), 0.0);
        std::fill(my.pot.begin(), my.pot.end(), 0.0);
      }
      std::vector<std::pair<size_t, size_t>> st;
      for (size_t s = lo; s < hi; s++) {
        st.push_back(seed[s]);
        while (!st.empty()) {
          auto [a, b] = st.back();
          st.pop_back();
          Action act = decide(*t, mac, a, b, [&](size_t x, size_t y) { st.emplace_back(x, y); });
          const Cell& ca = t->cells[a];
          const Cell& cb = t->cells[b];
          if (act == ACT_SELF_LEAF) {
            tp2p[th].push_back((uint32_t)a);
            tp2p[th].push_back((uint32_t)a);
            p2p_self(my, ca.first, ca.count, soft2);
          } else if (act == ACT_M2L) {
            tm2l[th].push_back((uint32_t)a);
            tm2l[th].push_back((uint32_t)b);
          } else if (act == ACT_P2P) {
            tp2p[th].push_back((uint32_t)a);
            tp2p[th].push_back((uint32_t)b);
            p2p_pair(my, ca.first, ca.count, cb.first, cb.count, soft2);
          }
        }
      }
    });
    t->m2l = top_m2l;
    t->p2p = top_p2p;
    for (int th = 0; th < nthreads; th++) {
      t->m2l.insert(t->m2l.end(), tm2l[th].begin(), tm2l[th].end());
      t->p2p.insert(t->p2p.end(), tp2p[th].begin(), tp2p[th].end());
    }
    // near-field reduce (walk.rs:590-616), thread order
    parallel_for(nthreads, n, [&](int, size_t lo, size_t hi) {
      for (size_t k = lo; k < hi; k++)
        for (int th = 0; th < nthreads; th++) {
          if (tb[th].x.empty()) continue;
          B.ax[k] += tb[th].ax[k];
          B.ay[k] += tb[th].ay[k];
          B.az[k] += tb[th].az[k];
          B.pot[k] += tb[th].pot[k];
        }
    });
    tb.clear();
    // replay: partition by min(a,b) subtree buckets, contiguous ranges per thread
    size_t npairs = t->m2l.size() / 2;
    std::vector<uint32_t> perm(npairs);
    for (size_t i = 0; i < npairs; i++) perm[i] = (uint32_t)i;
    std::sort(perm.begin(), perm.end(), [&](uint32_t x, uint32_t y) {
      uint32_t kx = std::min(t->m2l[2 * x], t->m2l[2 * x + 1]);
      uint32_t ky = std::min(t->m2l[2 * y], t->m2l[2 * y + 1]);
      return kx < ky || (kx == ky && x < y);
    });
    std::vector<std::vector<C>> tls(nthreads);
    parallel_for(nthreads, npairs, [&](int th, size_t lo, size_t hi) {
      tls[th].assign(nc * d, cz());
      M2lScratch scr(p);
      RotCtx* rot = use_rot ? new RotCtx(p) : nullptr;
      C* loc = tls[th].data();
      for (size_t q = lo; q < hi; q++) {
        size_t a = t->m2l[2 * (size_t)perm[q]], b = t->m2l[2 * (size_t)perm[q] + 1];
        const Cell& ca = t->cells[a];
        const Cell& cb = t->cells[b];
        if (rot)
          rot->m2l_pair(ca.cen, cb.cen, mom + a * d, mom + b * d, loc + a * d, loc + b * d, soft2);
        else
          m2l_mutual(p, ca.cen, cb.cen, mom + a * d, mom + b * d, loc + a * d, loc + b * d, scr, soft2);
      }
      delete rot;
    });
    parallel_for(nthreads, nc * d, [&](int, size_t lo, size_t hi) {
      for (size_t i = lo; i < hi; i++) {
        C acc = t->locals[i];
        for (int th = 0; th < nthreads; th++)
          if (!tls[th].empty()) acc = cadd(acc, tls[th][i]);
        t->locals[i] = acc;
      }
    });
  }
  t->lap.clear();
  if (soft2 > 0.0 && !t->m2l.empty()) softened_trace(*t, p, soft2);
  double t3 = now_ms();
  t->locals_walk = t->locals;
  // l2l_down (translate.rs:209-252): ascending level
  for (size_t l = 0; l + 1 < t->lvl_off.size(); l++) {
    size_t a = t->lvl_off[l], b = t->lvl_off[l + 1];
    parallel_for(nthreads, b - a, [&](int, size_t lo, size_t hi) {
      std::vector<C> u(d);
      for (size_t i = lo; i < hi; i++) {
        size_t c = t->lvl_idx[a + i];
        const Cell& cell = t->cells[c];
        if (cell.leaf) continue;
        for (int k = 0; k < 8; k++) {
          int ch = cell.children[k];
          if (ch < 0) continue;
          l2l_into(p, cell.cen, t->cells[ch].cen, t->locals.data() + c * d,
                   t->locals.data() + (size_t)ch * d, u.data());
        }
      }
    });
  }
  double t4 = now_ms();
  // l2p_scatter (translate.rs:312-407)
  std::memset(accel, 0, 3 * n * sizeof(double));
  std::memset(pot, 0, n * sizeof(double));
  parallel_for_dyn(nthreads, nc, [&](int, size_t lo, size_t hi) {
    std::vector<C> u(d);
    for (size_t c = lo; c < hi; c++) {
      const Cell& cell = t->cells[c];
      if (!cell.leaf) continue;
      for (size_t k = cell.first; k < cell.first + cell.count; k++) {
        size_t id = t->order[k];
        double p0 = 0.0;
        double a[3] = {0.0, 0.0, 0.0};
        double x[3] = {t->bodies.x[k], t->bodies.y[k], t->bodies.z[k]};
        l2p_into(p, cell.cen, x, t->locals.data() + c * d, &p0, a, u.data());
        if (!t->lap.empty())  // the leaf's and every ancestor's trace polynomial
          for (int32_t A = (int32_t)c; A >= 0; A = t->parent[A])
            l2p_trace(t->cells[A].cen, x, &t->lap[SOFT_LAP_WORDS * (size_t)A], &p0, a);
        pot[id] += t->bodies.pot[k] + p0;
        accel[3 * id + 0] += t->bodies.ax[k] + a[0];
        accel[3 * id + 1] += t->bodies.ay[k] + a[1];
        accel[3 * id + 2] += t->bodies.az[k] + a[2];
      }
    }
  });
  double t5 = now_ms();
  t->times[0] = t1 - t0;
  t->times[1] = t2 - t1;
  t->times[2] = t3 - t2;
  t->times[3] = t4 - t3;
  t->times[4] = t5 - t4;
  return 0;
}

int orc_dim(const orc_ctx* t) { return (int)dimp(t->p); }
void orc_get_moments(const orc_ctx* t, double* out) {
  std::memcpy(out, t->moments.data(), t->moments.size() * sizeof(C));
}
void orc_get_pows(const orc_ctx* t, double* out) {
  std::memcpy(out, t->pows.data(), t->pows.size() * sizeof(double));
}
void orc_get_f(const orc_ctx* t, double* out) {
  std::memcpy(out, t->f.data(), t->f.size() * sizeof(double));
}
void orc_get_fmin(const orc_ctx* t, double* out) {
  std::memcpy(out, t->fmin.data(), t->fmin.size() * sizeof(double));
}
size_t orc_n_m2l(const orc_ctx* t) { return t->m2l.size() / 2; }
size_t orc_n_p2p(const orc_ctx* t) { return t->p2p.size() / 2; }
void orc_get_m2l(const orc_ctx* t, uint32_t* ab) {
  std::memcpy(ab, t->m2l.data(), t->m2l.size() * sizeof(uint32_t));
}
void orc_get_p2p(const orc_ctx* t, uint32_t* ab) {
  std::memcpy(ab, t->p2p.data(), t->p2p.size() * sizeof(uint32_t));
}
void orc_get_locals_walk(const orc_ctx* t, double* out) {
  std::memcpy(out, t->locals_walk.data(), t->locals_walk.size() * sizeof(C));
}
void orc_get_locals(const orc_ctx* t, double* out) {
  std::memcpy(out, t->locals.data(), t->locals.size() * sizeof(C));
}
void orc_get_near(const orc_ctx* t, double* out) {
  size_t n = t->bodies.x.size();
  for (size_t k = 0; k < n; k++) {
    out[4 * k] = t->bodies.ax[k];
    out[4 * k + 1] = t->bodies.ay[k];
    out[4 * k + 2] = t->bodies.az[k];
    out[4 * k + 3] = t->bodies.pot[k];
  }
}
void orc_get_times(const orc_ctx* t, double* out5) { std::memcpy(out5, t->times, sizeof t->times); }

}  // extern "C"
