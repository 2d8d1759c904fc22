// m2l.hip — M2L work decomposition (chunks of at most m2l_chunk sources per target),
// the per-target reduction of chunk partials and the order dispatch. The
// contraction kernels live in m2l_kernel.h, instantiated per range of orders
// by m2l_p*.hip (separate translation units so they build in parallel).
#include "m2l_kernel.h"

namespace fmmh {

template <int P>
void m2l_t(const M2lChunk* chunks, uint32_t nchunks, const uint64_t* skeys, uint64_t smask,
           DevCells c, const cplx* moments, double soft2, cplx* part, hipStream_t s);
template <int P>
void m2l32_t(const M2lChunk* chunks, uint32_t nchunks, const uint64_t* skeys, uint64_t smask,
             DevCells c, const double* lam, const cplxf* m32, double soft2, cplxf* part,
             hipStream_t s);

// Scale of each cell for the FP32 M2L: its radius, floored at a quarter of
// its octant edge (the system's box side halved per level below the system
// root) so cells of coincident bodies still get a length; 1 if both vanish.
__global__ void __launch_bounds__(256) k_m2l_scale(uint32_t nc, DevCells c,
                                                  const BBox* __restrict__ bb,
                                                  const uint64_t* __restrict__ keys_sorted,
                                                  int vlev, double* __restrict__ lam) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nc) return;
  const int lev = (int)c.level[i];
  double l = 1.0;
  if (lev >= vlev) {
    const uint32_t sys = vlev ? (uint32_t)(keys_sorted[c.first[i]] >> (63 - 3 * vlev)) : 0u;
    l = fmax(c.rho[i], 0.25 * ldexp(bb[sys].side, -(lev - vlev)));
    if (!(l > 0.0) || !isfinite(l)) l = 1.0;
  }
  lam[i] = l;
}

// Moments scaled by lam^-k, m >= 0 half, FP32, of the cells above level
// `below` (the others come from the FP32 upward pass).
template <int P>
__global__ void __launch_bounds__(256) k_m2l_prep32(uint32_t nc, const cplx* __restrict__ moments,
                                                   const double* __restrict__ lam,
                                                   const uint8_t* __restrict__ level, int below,
                                                   cplxf* __restrict__ m32) {
  constexpr int D = hdim(P), NF = hnpos(P);
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nc || (int)level[i] >= below) return;
  const cplx* ms = moments + (size_t)i * D;
  cplxf* o = m32 + (size_t)i * NF;
  const double il = 1.0 / lam[i];
  double sc = 1.0;
#pragma unroll
  for (int k = 0; k <= P; k++) {
#pragma unroll
    for (int l = 0; l <= k; l++) {
      const cplx v = ms[hidx(k, l)];
      o[hpos(k, l)] = cplxf{(float)(v.re * sc), (float)(v.im * sc)};
    }
    sc *= il;
  }
}

__global__ void __launch_bounds__(256) k_m2l_nchunks(const uint32_t* __restrict__ tptr, uint32_t nc,
                                                    uint32_t ch, uint32_t* __restrict__ nch) {
  uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nc) return;
  uint32_t s = tptr[t + 1] - tptr[t];
  nch[t] = (s + ch - 1) / ch;
}

__global__ void __launch_bounds__(256) k_m2l_emit(const uint32_t* __restrict__ tptr, uint32_t nc,
                                                 uint32_t ch, const uint32_t* __restrict__ choff,
                                                 M2lChunk* __restrict__ chunks) {
  uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nc) return;
  uint32_t b = tptr[t], e = tptr[t + 1];
  uint32_t w = choff[t];
  for (uint32_t j = b; j < e; j += ch) chunks[w++] = M2lChunk{t, j, min(e, j + ch)};
}

constexpr int RED_AHEAD = 8;  // chunk partials loaded ahead of the in-order sum

// locals[t] = sum over the target's chunk partials in chunk order; both
// halves of m written (F_n^{-m} = (-1)^m conj(F_n^m)); empty targets get 0.
template <int P>
__global__ void __launch_bounds__(256) k_m2l_reduce(uint32_t nc, const uint32_t* __restrict__ choff,
                                                   const cplx* __restrict__ part,
                                                   cplx* __restrict__ locals) {
  constexpr int D = hdim(P), NF = hnpos(P);
  size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= (size_t)nc * NF) return;
  uint32_t t = (uint32_t)(g / NF);
  int e = (int)(g % NF);
  int n = 0;
  while ((n + 1) * (n + 2) / 2 <= e) n++;
  int m = e - n * (n + 1) / 2;
  cplx s = cplx{0.0, 0.0};
  const uint32_t c0 = choff[t], c1 = choff[t + 1];
  uint32_t c = c0;
  for (; c + RED_AHEAD <= c1; c += RED_AHEAD) {  // loads issued together, added in order
    cplx v[RED_AHEAD];
#pragma unroll
    for (int j = 0; j < RED_AHEAD; j++) v[j] = part[(size_t)(c + j) * NF + e];
#pragma unroll
    for (int j = 0; j < RED_AHEAD; j++) {
      s.re += v[j].re;
      s.im += v[j].im;
    }
  }
  for (; c < c1; c++) {
    cplx v = part[(size_t)c * NF + e];
    s.re += v.re;
    s.im += v.im;
  }
  cplx* out = locals + (size_t)t * D;
  out[hidx(n, m)] = s;
  if (m > 0) {
    double sg = (m & 1) ? -1.0 : 1.0;
    out[hidx(n, -m)] = cplx{sg * s.re, -sg * s.im};
  }
}

// Mixed form: FP32 chunk partials of the scaled F, summed in FP64 in chunk
// order, then F_n = lam_t^-(n+1) * sum.
template <int P>
__global__ void __launch_bounds__(256) k_m2l_reduce32(uint32_t nc, const uint32_t* __restrict__ choff,
                                                     const cplxf* __restrict__ part,
                                                     const double* __restrict__ lam,
                                                     cplx* __restrict__ locals) {
  constexpr int D = hdim(P), NF = hnpos(P);
  size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= (size_t)nc * NF) return;
  uint32_t t = (uint32_t)(g / NF);
  int e = (int)(g % NF);
  int n = 0;
  while ((n + 1) * (n + 2) / 2 <= e) n++;
  int m = e - n * (n + 1) / 2;
  double sr = 0.0, si = 0.0;
  const uint32_t c0 = choff[t], c1 = choff[t + 1];
  uint32_t c = c0;
  for (; c + RED_AHEAD <= c1; c += RED_AHEAD) {  // loads issued together, added in order
    cplxf v[RED_AHEAD];
#pragma unroll
    for (int j = 0; j < RED_AHEAD; j++) v[j] = part[(size_t)(c + j) * NF + e];
#pragma unroll
    for (int j = 0; j < RED_AHEAD; j++) {
      sr += (double)v[j].re;
      si += (double)v[j].im;
    }
  }
  for (; c < c1; c++) {
    const cplxf v = part[(size_t)c * NF + e];
    sr += (double)v.re;
    si += (double)v.im;
  }
  const double il = 1.0 / lam[t];
  double sc = il;
  for (int k = 0; k < n; k++) sc *= il;
  const cplx s = cplx{sr * sc, si * sc};
  cplx* out = locals + (size_t)t * D;
  out[hidx(n, m)] = s;
  if (m > 0) {
    double sg = (m & 1) ? -1.0 : 1.0;
    out[hidx(n, -m)] = cplx{sg * s.re, -sg * s.im};
  }
}

// Scratch layout handed in by the context: nch/choff (nc+1 words each) and
// chunk/partial arrays sized for the worst case sum_t ceil(S_t/CH).
size_t m2l_max_chunks(uint32_t nnz, uint32_t nc, uint32_t ch) { return (size_t)nnz / ch + nc + 1; }
size_t m2l_part_bytes(int p, size_t maxchunks) { return maxchunks * hnpos(p) * sizeof(cplx); }

template <int P>
static void m2l_reduce_t(uint32_t nc, const uint32_t* choff, const cplx* part, cplx* locals,
                         hipStream_t s) {
  size_t th = (size_t)nc * hnpos(P);
  hipLaunchKernelGGL(k_m2l_reduce<P>, dim3((th + 255) / 256), dim3(256), 0, s, nc, choff, part,
                     locals);
}

void m2l_build_chunks(const uint32_t* tptr, uint32_t nc, uint32_t ch, uint32_t* nch,
                      uint32_t* choff, uint32_t* scan_tmp, void* chunks, hipStream_t s,
                      uint32_t** total_dev) {
  hipLaunchKernelGGL(k_m2l_nchunks, dim3((nc + 255) / 256), dim3(256), 0, s, tptr, nc, ch, nch);
  scan_exclusive_u32(nch, choff, (size_t)nc, scan_tmp, s, total_dev);
  hipLaunchKernelGGL(k_m2l_emit, dim3((nc + 255) / 256), dim3(256), 0, s, tptr, nc, ch, choff,
                     (M2lChunk*)chunks);
}

#define FMMH_P_SWITCH3(p, CALL) \
  switch (p) {                  \
    case 1: CALL(1); break;     \
    case 2: CALL(2); break;     \
    case 3: CALL(3); break;     \
    case 4: CALL(4); break;     \
    case 5: CALL(5); break;     \
    case 6: CALL(6); break;     \
    case 7: CALL(7); break;     \
    case 8: CALL(8); break;     \
    case 9: CALL(9); break;     \
    case 10: CALL(10); break;   \
    default: return false;      \
  }

bool launch_m2l_kernel(int p, const void* chunks, uint32_t nchunks, const uint64_t* skeys,
                       int sbits, DevCells c, const cplx* moments, double soft2, cplx* part,
                       hipStream_t s) {
  const uint64_t smask = (1ull << sbits) - 1ull;
  const M2lChunk* ck = (const M2lChunk*)chunks;
#define CALL(P) m2l_t<P>(ck, nchunks, skeys, smask, c, moments, soft2, part, s)
  FMMH_P_SWITCH3(p, CALL)
#undef CALL
  return true;
}

void launch_m2l_scale(uint32_t nc, DevCells c, const BBox* bb, const uint64_t* keys_sorted,
                      int vlev, double* lam, hipStream_t s) {
  if (nc == 0) return;
  hipLaunchKernelGGL(k_m2l_scale, dim3((nc + 255) / 256), dim3(256), 0, s, nc, c, bb, keys_sorted,
                     vlev, lam);
}

bool launch_m2l32(int p, uint32_t nc, const void* chunks, uint32_t nchunks, const uint64_t* skeys,
                  int sbits, DevCells c, const BBox* bb, const uint64_t* keys_sorted, int vlev,
                  const cplx* moments, double soft2, double* lam, void* m32, void* part,
                  hipStream_t s, int prep_below) {
  const uint64_t smask = (1ull << sbits) - 1ull;
  const M2lChunk* ck = (const M2lChunk*)chunks;
  if (nc == 0) return true;
  launch_m2l_scale(nc, c, bb, keys_sorted, vlev, lam, s);
#define CALL(P)                                                                            \
  hipLaunchKernelGGL(k_m2l_prep32<P>, dim3((nc + 255) / 256), dim3(256), 0, s, nc, moments, \
                     (const double*)lam, c.level, prep_below, (cplxf*)m32);                \
  m2l32_t<P>(ck, nchunks, skeys, smask, c, lam, (const cplxf*)m32, soft2, (cplxf*)part, s)
  FMMH_P_SWITCH3(p, CALL)
#undef CALL
  return true;
}

size_t m2l32_moment_bytes(int p) { return (size_t)hnpos(p) * sizeof(cplxf); }

bool launch_m2l_reduce32(int p, uint32_t nc, const uint32_t* choff, const void* part,
                         const double* lam, cplx* locals, hipStream_t s) {
#define CALL(P)                                                                                  \
  hipLaunchKernelGGL(k_m2l_reduce32<P>, dim3(((size_t)nc * hnpos(P) + 255) / 256), dim3(256), 0, \
                     s, nc, choff, (const cplxf*)part, lam, locals)
  FMMH_P_SWITCH3(p, CALL)
#undef CALL
  return true;
}

bool launch_m2l_reduce(int p, uint32_t nc, const uint32_t* choff, const cplx* part, cplx* locals,
                       hipStream_t s) {
#define CALL(P) m2l_reduce_t<P>(nc, choff, part, locals, s)
  FMMH_P_SWITCH3(p, CALL)
#undef CALL
  return true;
}

size_t m2l_chunk_bytes() { return sizeof(M2lChunk); }

}  // namespace fmmh
