// generic_far.hip — far-field kernels for any expansion order (runtime p),
// used when p exceeds the register-resident templated kernels.
//
// M2L for 13 <= p < 20: naive convolution (translate.rs:47-136), one block
// per target with the pair's Theta table and the source multipole in LDS.
// M2L for p >= 20: the rotation route of rotation.rs:106-435 (Dehnen 2014
// A.6.1): z-phase, x<->z swap matrix B_k, z-phase, B_k, axial M2L along z,
// then the inverse rotation with B^T — O(p^3) per pair. As in the reference
// the switch happens at p = 20. Each block owns one target's locals, so the
// directed pairs of a target accumulate without atomics.
#include <vector>

#include "common.h"
#include "internal.h"

namespace fmmh {

__device__ __forceinline__ int deg_of(int e) {
  int n = (int)sqrt((double)e);
  while (n * n > e) n--;
  while ((n + 1) * (n + 1) <= e) n++;
  return n;
}

__device__ __forceinline__ cplx cmul_(cplx a, cplx b) {
  return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// Column m (both signs) of Theta(r) of the (optionally Plummer-softened)
// kernel: the harmonic.rs:91-129 recursion; softened, the harmonic
// derivatives of (r^2 + eps^2)^{-1/2} = those of 1/r times (r/s)^{2n+1}, i.e.
// 1/r^2 -> 1/s^2 and the (n-2) term scaled by r^2/s^2.
__device__ void theta_column_gen(int p, double x, double y, double z, double soft2, int m,
                                 cplx* tab) {
  const double rr2 = x * x + y * y + z * z;
  const double r2 = rr2 + soft2;
  const double ir2 = 1.0 / r2;
  const double ks = soft2 > 0.0 ? rr2 * ir2 : 1.0;
  cplx t = cplx{1.0 / sqrt(r2), 0.0};
  for (int j = 1; j <= m; j++) {
    const double f = (double)(2 * j - 1) * ir2;
    t = cmul_(cplx{x * f, y * f}, t);
  }
  const double sgn = (m & 1) ? -1.0 : 1.0;
  cplx t1 = t, t2 = cplx{0.0, 0.0};
  for (int n = m; n <= p; n++) {
    cplx tn;
    if (n == m) {
      tn = t;
    } else {
      const double a = (double)(2 * n - 1) * z;
      const double b = (double)((n - 1) * (n - 1) - m * m) * ks;
      tn = cplx{(a * t1.re - b * t2.re) * ir2, (a * t1.im - b * t2.im) * ir2};
      t2 = t1;
      t1 = tn;
    }
    tab[hidx(n, m)] = tn;
    if (m > 0) tab[hidx(n, -m)] = cplx{sgn * tn.re, -sgn * tn.im};
  }
}

// ------------------------------------------------------- naive, 13 <= p < 20
constexpr int GEN_T = 256;
constexpr int GEN_EMAX = 2;  // entries per thread: (p+1)^2 <= 400 <= 2 * 256

__global__ void __launch_bounds__(GEN_T) k_m2l_naive_gen(int p, uint32_t nc,
                                                        const uint32_t* __restrict__ tptr,
                                                        const uint64_t* __restrict__ skeys,
                                                        uint64_t smask, DevCells c,
                                                        const cplx* __restrict__ moments,
                                                        double soft2, cplx* __restrict__ locals) {
  extern __shared__ cplx sh[];
  const int D = hdim(p);
  cplx* th = sh;
  cplx* ms = sh + D;
  const uint32_t t = xcd_block(blockIdx.x, gridDim.x);
  const double tx = c.cx[t], ty = c.cy[t], tz = c.cz[t];
  cplx acc[GEN_EMAX];
  for (int q = 0; q < GEN_EMAX; q++) acc[q] = cplx{0.0, 0.0};
  for (uint32_t j = tptr[t]; j < tptr[t + 1]; j++) {
    const uint32_t src = (uint32_t)(skeys[j] & smask);
    __syncthreads();
    for (int e = threadIdx.x; e < D; e += GEN_T) ms[e] = moments[(size_t)src * D + e];
    for (int m = threadIdx.x; m <= p; m += GEN_T)
      theta_column_gen(p, tx - c.cx[src], ty - c.cy[src], tz - c.cz[src], soft2, m, th);
    __syncthreads();
    for (int q = 0; q < GEN_EMAX; q++) {
      const int e = threadIdx.x + q * GEN_T;
      if (e >= D) break;
      const int n = deg_of(e), m = e - n * (n + 1);
      double fr = 0.0, fi = 0.0;
      for (int k = 0; k <= p - n; k++)
        for (int l = -k; l <= k; l++) {
          const cplx a = ms[hidx(k, l)];       // conj(a) * b
          const cplx b = th[hidx(n + k, m + l)];
          fr = fma(a.re, b.re, fma(a.im, b.im, fr));
          fi = fma(a.re, b.im, fma(-a.im, b.re, fi));
        }
      acc[q].re += fr;
      acc[q].im += fi;
    }
  }
  for (int q = 0; q < GEN_EMAX; q++) {
    const int e = threadIdx.x + q * GEN_T;
    if (e < D) locals[(size_t)t * D + e] = acc[q];
  }
}

// --------------------------------------------------------- rotation, p >= 20
struct RotTables {
  const double* sw;   // B_n, block layout, sum (2n+1)^2
  const double* swt;  // B_n^T
  const cplx* din;    // e^{-im pi/2} (n-|m|)!(n+|m|)!
  const cplx* dout;   // e^{+im pi/2} / ((n-|m|)!(n+|m|)!)
  const cplx* fin;    // e^{-im pi/2} / ((n-|m|)!(n+|m|)!)
  const cplx* fout;   // e^{+im pi/2} (n-|m|)!(n+|m|)!
  const double* fact; // j!
};

__device__ __forceinline__ int blk_off(int n) { return n == 0 ? 0 : n * (4 * n * n - 1) / 3; }

__global__ void __launch_bounds__(GEN_T) k_m2l_rot(int p, uint32_t nc,
                                                  const uint32_t* __restrict__ tptr,
                                                  const uint64_t* __restrict__ skeys,
                                                  uint64_t smask, DevCells c,
                                                  const cplx* __restrict__ moments, double soft2,
                                                  RotTables rt, cplx* __restrict__ locals) {
  extern __shared__ cplx sh[];
  const int D = hdim(p);
  cplx* X = sh;
  cplx* Y = sh + D;
  cplx* pha = Y + D;
  cplx* phb = pha + (2 * p + 1);
  double* facr = (double*)(phb + (2 * p + 1));
  const uint32_t t = xcd_block(blockIdx.x, gridDim.x);
  cplx* F = locals + (size_t)t * D;
  for (int e = threadIdx.x; e < D; e += GEN_T) F[e] = cplx{0.0, 0.0};
  const double tx = c.cx[t], ty = c.cy[t], tz = c.cz[t];
  for (uint32_t j = tptr[t]; j < tptr[t + 1]; j++) {
    const uint32_t src = (uint32_t)(skeys[j] & smask);
    const cplx* M = moments + (size_t)src * D;
    // pair geometry (rotation.rs:267-290): r = s_target - z_source
    const double rx = tx - c.cx[src], ry = ty - c.cy[src], rz = tz - c.cz[src];
    const double r2 = rx * rx + ry * ry + rz * rz;
    const double alpha = atan2(ry, rx);
    const double beta = atan2(sqrt(rx * rx + ry * ry), rz);
    __syncthreads();
    for (int i = threadIdx.x; i < 2 * p + 1; i += GEN_T) {
      const int m = i - p;
      double sa, ca, sb, cb;
      sincos((double)m * alpha, &sa, &ca);
      sincos((double)m * beta, &sb, &cb);
      pha[i] = cplx{ca, sa};
      phb[i] = cplx{cb, sb};
    }
    if (threadIdx.x == 0) {
      if (soft2 > 0.0) {  // axial harmonic derivatives of the softened kernel, q! R^q / s^{2q+1}
        const double R = sqrt(r2), s2 = r2 + soft2;
        facr[0] = 1.0 / sqrt(s2);
        for (int q = 1; q <= p; q++) facr[q] = facr[q - 1] * ((double)q * R / s2);
      } else {  // (n+k)!/R^(n+k+1)
        const double rinv = 1.0 / sqrt(r2);
        double irp = rinv;
        for (int q = 0; q <= p; q++) {
          facr[q] = rt.fact[q] * irp;
          irp *= rinv;
        }
      }
    }
    __syncthreads();
    // 1) u = din . pha . conj(M)
    for (int e = threadIdx.x; e < D; e += GEN_T) {
      const int k = deg_of(e), cc = e - k * k;
      const cplx m = M[e], ph = pha[p - k + cc], iv = rt.din[e];
      const double vr = m.re * ph.re + m.im * ph.im;
      const double vi = m.re * ph.im - m.im * ph.re;
      X[e] = cplx{vr * iv.re - vi * iv.im, vr * iv.im + vi * iv.re};
    }
    __syncthreads();
    // 2) Y = phb . (B_k u)
    for (int e = threadIdx.x; e < D; e += GEN_T) {
      const int k = deg_of(e), r = e - k * k, nb = 2 * k + 1;
      const double* w = rt.sw + blk_off(k) + r * nb;
      const cplx* v = X + k * k;
      double sr = 0.0, si = 0.0;
      for (int q = 0; q < nb; q++) {
        sr = fma(w[q], v[q].re, sr);
        si = fma(w[q], v[q].im, si);
      }
      const cplx pb = phb[p - k + r];
      Y[e] = cplx{sr * pb.re - si * pb.im, sr * pb.im + si * pb.re};
    }
    __syncthreads();
    // 3) b = dout . (B_k Y)
    for (int e = threadIdx.x; e < D; e += GEN_T) {
      const int k = deg_of(e), r = e - k * k, nb = 2 * k + 1;
      const double* w = rt.sw + blk_off(k) + r * nb;
      const cplx* v = Y + k * k;
      double sr = 0.0, si = 0.0;
      for (int q = 0; q < nb; q++) {
        sr = fma(w[q], v[q].re, sr);
        si = fma(w[q], v[q].im, si);
      }
      const cplx ov = rt.dout[e];
      X[e] = cplx{sr * ov.re - si * ov.im, sr * ov.im + si * ov.re};
    }
    __syncthreads();
    // 4) axial M2L along z, then fin: Y[n,m] = fin . sum_{k>=|m|} b[k,-m] facr[n+k]
    for (int e = threadIdx.x; e < D; e += GEN_T) {
      const int n = deg_of(e), r = e - n * n, m = r - n, am = m < 0 ? -m : m;
      double fr = 0.0, fi = 0.0;
      for (int k = am; k <= p - n; k++) {
        const cplx bv = X[k * k + (k - m)];
        const double fv = facr[n + k];
        fr = fma(bv.re, fv, fr);
        fi = fma(bv.im, fv, fi);
      }
      const cplx iv = rt.fin[e];
      Y[e] = cplx{fr * iv.re - fi * iv.im, fr * iv.im + fi * iv.re};
    }
    __syncthreads();
    // 5) X = conj(phb) . (B_n^T Y)
    for (int e = threadIdx.x; e < D; e += GEN_T) {
      const int n = deg_of(e), r = e - n * n, nb = 2 * n + 1;
      const double* w = rt.swt + blk_off(n) + r * nb;
      const cplx* v = Y + n * n;
      double sr = 0.0, si = 0.0;
      for (int q = 0; q < nb; q++) {
        sr = fma(w[q], v[q].re, sr);
        si = fma(w[q], v[q].im, si);
      }
      const cplx pb = phb[p - n + r];
      X[e] = cplx{sr * pb.re + si * pb.im, si * pb.re - sr * pb.im};
    }
    __syncthreads();
    // 6) F += fout . pha . (B_n^T X)
    for (int e = threadIdx.x; e < D; e += GEN_T) {
      const int n = deg_of(e), r = e - n * n, nb = 2 * n + 1;
      const double* w = rt.swt + blk_off(n) + r * nb;
      const cplx* v = X + n * n;
      double sr = 0.0, si = 0.0;
      for (int q = 0; q < nb; q++) {
        sr = fma(w[q], v[q].re, sr);
        si = fma(w[q], v[q].im, si);
      }
      const cplx ov = rt.fout[e], pv = pha[p - n + r];
      const cplx a = cplx{sr * ov.re - si * ov.im, sr * ov.im + si * ov.re};
      F[e].re += a.re * pv.re - a.im * pv.im;
      F[e].im += a.re * pv.im + a.im * pv.re;
    }
  }
}

// ------------------------------------------------------------- L2L, L2P
__device__ void upsilon_column_far(int p, double x, double y, double z, int m, cplx* tab) {
  const double r2 = x * x + y * y + z * z;
  cplx u = cplx{1.0, 0.0};
  for (int j = 1; j <= m; j++) {
    const double f = 1.0 / (double)(2 * j);
    u = cmul_(cplx{x * f, y * f}, u);
  }
  const double sgn = (m & 1) ? -1.0 : 1.0;
  cplx u1 = u, u2 = cplx{0.0, 0.0};
  for (int n = m; n <= p; n++) {
    cplx un;
    if (n == m) {
      un = u;
    } else {
      const double a = (double)(2 * n - 1) * z;
      const double ic = 1.0 / (double)(n * n - m * m);
      const double bb = (n >= m + 2) ? r2 : 0.0;
      un = cplx{(a * u1.re - bb * u2.re) * ic, (a * u1.im - bb * u2.im) * ic};
      u2 = u1;
      u1 = un;
    }
    tab[hidx(n, m)] = un;
    if (m > 0) tab[hidx(n, -m)] = cplx{sgn * un.re, -sgn * un.im};
  }
}

__global__ void __launch_bounds__(GEN_T) k_l2l_gen(int p, const uint32_t* __restrict__ cells,
                                                  DevCells c, cplx* __restrict__ locals) {
  extern __shared__ cplx u[];
  const int D = hdim(p);
  const uint32_t ch = cells[blockIdx.x];
  const int32_t pa = c.parent[ch];
  if (pa < 0) return;
  for (int m = threadIdx.x; m <= p; m += GEN_T)
    upsilon_column_far(p, c.cx[pa] - c.cx[ch], c.cy[pa] - c.cy[ch], c.cz[pa] - c.cz[ch], m, u);
  __syncthreads();
  const cplx* fp = locals + (size_t)pa * D;
  cplx* fc = locals + (size_t)ch * D;
  for (int e = threadIdx.x; e < D; e += GEN_T) {
    const int n = deg_of(e), m = e - n * (n + 1);
    double sr = 0.0, si = 0.0;
    for (int k = 0; k <= p - n; k++)
      for (int l = -k; l <= k; l++) {
        const cplx a = u[hidx(k, l)], b = fp[hidx(n + k, m + l)];
        sr = fma(a.re, b.re, fma(a.im, b.im, sr));
        si = fma(a.re, b.im, fma(-a.im, b.re, si));
      }
    fc[e].re += sr;
    fc[e].im += si;
  }
}

constexpr int GEN_PMAX = FMMH_MAX_P;

__global__ void __launch_bounds__(64) k_l2p_gen(int p, const uint32_t* __restrict__ leaf_ids,
                                               DevCells c, DevBodies b,
                                               const cplx* __restrict__ locals,
                                               const double* __restrict__ lap,
                                               double4* __restrict__ far) {
  extern __shared__ cplx F[];
  const int D = hdim(p);
  const uint32_t leaf = leaf_ids[blockIdx.x];
  for (int e = threadIdx.x; e < D; e += 64) F[e] = locals[(size_t)leaf * D + e];
  __syncthreads();
  const double cxl = c.cx[leaf], cyl = c.cy[leaf], czl = c.cz[leaf];
  const uint32_t f0 = c.first[leaf], e0 = f0 + c.count[leaf];
  for (uint32_t k = f0 + threadIdx.x; k < e0; k += 64) {
    const double x = cxl - b.x[k], y = cyl - b.y[k], z = czl - b.z[k];
    const double r2 = x * x + y * y + z * z;
    cplx U[GEN_PMAX + 1], U1[GEN_PMAX + 1], U2[GEN_PMAX + 1];
    double p00 = 0.0, p11r = 0.0, p11i = 0.0, p10r = 0.0;
    for (int kk = 0; kk <= p; kk++) {
      if (kk == 0) {
        U[0] = cplx{1.0, 0.0};
      } else {
        const double a = (double)(2 * kk - 1) * z;
        for (int q = 0; q < kk; q++) {
          const double ic = 1.0 / (double)(kk * kk - q * q);
          const double bb = (q + 2 <= kk) ? r2 : 0.0;
          const cplx u2 = (q + 2 <= kk) ? U2[q] : cplx{0.0, 0.0};
          U[q] = cplx{(a * U1[q].re - bb * u2.re) * ic, (a * U1[q].im - bb * u2.im) * ic};
        }
        const double f = 1.0 / (double)(2 * kk);
        U[kk] = cmul_(cplx{x * f, y * f}, U1[kk - 1]);
      }
      for (int l = -kk; l <= kk; l++) {
        double ar, ai;  // conj(Y_k^l)
        if (l >= 0) {
          ar = U[l].re;
          ai = -U[l].im;
        } else {
          const double sg = ((-l) & 1) ? -1.0 : 1.0;
          ar = sg * U[-l].re;
          ai = sg * U[-l].im;
        }
        const cplx f00 = F[hidx(kk, l)];
        p00 += ar * f00.re - ai * f00.im;
        if (kk + 1 <= p) {
          const cplx f11 = F[hidx(kk + 1, l + 1)], f10 = F[hidx(kk + 1, l)];
          p11r += ar * f11.re - ai * f11.im;
          p11i += ar * f11.im + ai * f11.re;
          p10r += ar * f10.re - ai * f10.im;
        }
      }
      for (int q = 0; q <= kk; q++) {
        U2[q] = U1[q];
        U1[q] = U[q];
      }
    }
    double4 o = make_double4(-p11r, -p11i, -p10r, p00);
    if (lap) l2p_trace_walk(lap, c.parent, c.cx, c.cy, c.cz, leaf, b.x[k], b.y[k], b.z[k], o);
    far[k] = o;
  }
}

// ------------------------------------------------------------- host side
// Swap matrices B_n (Dehnen 2014 eq. 63, rotation.rs:137-211) and the folded
// diagonals (rotation.rs:212-235), built once per context for p >= 20.
void rot_tables_host(int p, std::vector<double>& sw, std::vector<double>& swt,
                     std::vector<cplx>& din, std::vector<cplx>& dout, std::vector<cplx>& fin,
                     std::vector<cplx>& fout, std::vector<double>& fact) {
  auto boff = [](int n) { return n == 0 ? 0 : n * (4 * n * n - 1) / 3; };
  const int total = (p + 1) * (4 * p * p + 8 * p + 3) / 3;
  sw.assign(total, 0.0);
  swt.assign(total, 0.0);
  sw[0] = swt[0] = 1.0;
  for (int n = 1; n <= p; n++) {
    const int nb = 2 * n + 1, pnb = 2 * n - 1, off = boff(n), poff = boff(n - 1);
    // entry (M, L) of B_{n-1}, zero outside |M|,|L| <= n-1
    auto prev = [&](int M, int L) -> double {
      if (M < -(n - 1) || M > n - 1 || L < -(n - 1) || L > n - 1) return 0.0;
      return sw[poff + (M + n - 1) * pnb + (L + n - 1)];
    };
    for (int M = -n; M <= n; M++)
      for (int L = -n; L <= n; L++) {
        double v;
        if (M > -n && M < n)
          v = (prev(M, L - 1) - prev(M, L + 1)) * 0.5;
        else if (M == n)
          v = (prev(n - 1, L - 1) + prev(n - 1, L + 1) + 2.0 * prev(n - 1, L)) * 0.5;
        else
          v = (prev(1 - n, L - 1) + prev(1 - n, L + 1) - 2.0 * prev(1 - n, L)) * 0.5;
        sw[off + (M + n) * nb + (L + n)] = v;
        swt[off + (L + n) * nb + (M + n)] = v;
      }
  }
  std::vector<double> f2(2 * p + 1, 1.0);
  for (size_t i = 1; i < f2.size(); i++) f2[i] = f2[i - 1] * (double)i;
  fact.assign(f2.begin(), f2.begin() + p + 1);
  const int D = (p + 1) * (p + 1);
  din.resize(D);
  dout.resize(D);
  fin.resize(D);
  fout.resize(D);
  for (int n = 0; n <= p; n++)
    for (int m = -n; m <= n; m++) {
      const int i = n * (n + 1) + m, am = m < 0 ? -m : m;
      const double fm = f2[n - am] * f2[n + am];
      const int r4 = ((m % 4) + 4) % 4;  // e^{-i m pi/2} = (-i)^m
      const cplx ei = r4 == 0 ? cplx{1, 0} : r4 == 1 ? cplx{0, -1} : r4 == 2 ? cplx{-1, 0} : cplx{0, 1};
      const cplx ec = cplx{ei.re, -ei.im};
      din[i] = cplx{ei.re * fm, ei.im * fm};
      dout[i] = cplx{ec.re / fm, ec.im / fm};
      fin[i] = cplx{ei.re / fm, ei.im / fm};
      fout[i] = cplx{ec.re * fm, ec.im * fm};
    }
}

void launch_m2l_gen(int p, uint32_t nc, const uint32_t* tptr, const uint64_t* skeys, int sbits,
                    DevCells c, const cplx* moments, double soft2, const void* rot_tables,
                    cplx* locals, hipStream_t s) {
  const uint64_t smask = (1ull << sbits) - 1ull;
  const size_t D = (size_t)hdim(p);
  if (p < 20) {
    hipLaunchKernelGGL(k_m2l_naive_gen, dim3(nc), dim3(GEN_T), 2 * D * sizeof(cplx), s, p, nc, tptr,
                       skeys, smask, c, moments, soft2, locals);
  } else {
    // p <= 40 keeps this within 64 KiB of LDS per workgroup
    const size_t lds = (2 * D + 2 * (2 * p + 1)) * sizeof(cplx) + (p + 1) * sizeof(double);
    hipLaunchKernelGGL(k_m2l_rot, dim3(nc), dim3(GEN_T), lds, s, p, nc, tptr, skeys, smask, c,
                       moments, soft2, *(const RotTables*)rot_tables, locals);
  }
}

size_t rot_tables_struct_bytes() { return sizeof(RotTables); }

void make_rot_tables(const double* sw, const double* swt, const cplx* din, const cplx* dout,
                     const cplx* fin, const cplx* fout, const double* fact, void* out) {
  RotTables* r = (RotTables*)out;
  r->sw = sw;
  r->swt = swt;
  r->din = din;
  r->dout = dout;
  r->fin = fin;
  r->fout = fout;
  r->fact = fact;
}

void launch_l2l_gen(int p, const uint32_t* cells, uint32_t cnt, DevCells c, cplx* locals,
                    hipStream_t s) {
  if (cnt == 0) return;
  hipLaunchKernelGGL(k_l2l_gen, dim3(cnt), dim3(GEN_T), (size_t)hdim(p) * sizeof(cplx), s, p,
                     cells, c, locals);
}

void launch_l2p_gen(int p, const uint32_t* leaf_ids, uint32_t nleaves, DevCells c, DevBodies b,
                    const cplx* locals, const double* lap, double4* far, hipStream_t s) {
  if (nleaves == 0) return;
  hipLaunchKernelGGL(k_l2p_gen, dim3(nleaves), dim3(64), (size_t)hdim(p) * sizeof(cplx), s, p,
                     leaf_ids, c, b, locals, lap, far);
}

}  // namespace fmmh
