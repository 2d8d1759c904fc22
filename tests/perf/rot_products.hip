// rot_products.hip — measures the small dense products of the rotation-based
// M2L (rotation.rs:125-245: the swap matrices B_n, applied four times per
// interacting pair) at expansion order 8 in FP32, on the vector ALUs and on
// the matrix cores (v_mfma_f32_16x16x4_f32), over the same pairs and the
// same matrices. The program prints one JSON line: the time of each kernel,
// the time per pair and application, and the largest difference between
// the two kernels' results.
//
// Work per pair and application: for every degree n <= 8 a real
// (n+1) x (n+1) matrix times the real and the imaginary parts of the
// degree's m >= 0 coefficients, sum_n (n+1)^2 = 285 multiply-adds each.
//
// VALU: one lane per pair, the coefficients in registers, the matrix
// entries uniform (scalar operands).
//
// MFMA: one 16 x 16 x 4 product chain per packed block. The degrees are
// packed block-diagonally into three 16 x 16 matrices, {8, 5}, {7, 6} and
// {4, 3, 2, 1, 0}; the columns are 8 pairs x (real, imaginary). Lane l
// holds the B operand rows {l/16 + 4 s} of column l % 16 (s = 0..3, one per
// K step), and receives the D rows {4 (l/16) + i}. The A rows are stored
// permuted (row 4 g + i holds logical row g + 4 i), so each product's
// output is already the next product's input layout: no lane exchange.
// 285 of the 3 x 256 = 768 entries per application are useful (37 %).
//
// Usage: rot_products [log2 pairs (default 24)] [applications (default 32)]
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#define CK(x)                                                                  \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));             \
      std::exit(1);                                                            \
    }                                                                          \
  } while (0)

constexpr int P = 8;
constexpr int NC = (P + 1) * (P + 2) / 2;  // coefficients (n, m >= 0)
constexpr int MW = 285;                    // sum_n (n+1)^2 matrix entries

__host__ __device__ constexpr int cidx(int n, int m) { return n * (n + 1) / 2 + m; }
__host__ __device__ constexpr int moff(int n) { return n * (n + 1) * (2 * n + 1) / 6; }

// Two matrices per degree, row-major at moff(n): an orthonormal DCT-II
// matrix and its transpose (the products alternate them).
__constant__ float cB[2][MW];

// Coefficient (n, m), part (0 re, 1 im) of pair j: a fixed value in [-1, 1].
__host__ __device__ inline float init_value(uint32_t j, int c, int part) {
  uint32_t h = j * 2654435761u ^ (uint32_t)(c * 2 + part) * 2246822519u;
  h ^= h >> 15;
  h *= 2654435769u;
  h ^= h >> 13;
  return (float)(h & 0xFFFFFF) * (2.0f / 16777216.0f) - 1.0f;
}

__global__ void __launch_bounds__(256) k_valu(uint32_t npairs, int apps, float* __restrict__ out) {
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= npairs) return;
  float re[NC], im[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) {
    re[c] = init_value(j, c, 0);
    im[c] = init_value(j, c, 1);
  }
  for (int a = 0; a < apps; a++) {
    const float* B = cB[a & 1];
#pragma unroll
    for (int n = 0; n <= P; n++) {
      float yr[P + 1], yi[P + 1];
#pragma unroll
      for (int m = 0; m <= n; m++) {
        float sr = 0.f, si = 0.f;
#pragma unroll
        for (int k = 0; k <= n; k++) {
          const float b = B[moff(n) + m * (n + 1) + k];
          sr = fmaf(b, re[cidx(n, k)], sr);
          si = fmaf(b, im[cidx(n, k)], si);
        }
        yr[m] = sr;
        yi[m] = si;
      }
#pragma unroll
      for (int m = 0; m <= n; m++) {
        re[cidx(n, m)] = yr[m];
        im[cidx(n, m)] = yi[m];
      }
    }
  }
  float s = 0.f;
#pragma unroll
  for (int c = 0; c < NC; c++) s += re[c] * (float)(c + 1) + im[c] * (float)(2 * c + 1);
  out[j] = s;
}

// Packed block b, logical row r -> coefficient (n, m), or -1 for padding.
struct Pack {
  int coef[3][16];
  int deg[3][16];
};
static Pack make_pack() {
  Pack pk;
  const int sets[3][5] = {{8, 5, -1, -1, -1}, {7, 6, -1, -1, -1}, {4, 3, 2, 1, 0}};
  for (int b = 0; b < 3; b++) {
    int r = 0;
    for (int q = 0; q < 5 && sets[b][q] >= 0; q++) {
      const int n = sets[b][q];
      for (int m = 0; m <= n; m++, r++) {
        pk.coef[b][r] = cidx(n, m);
        pk.deg[b][r] = n;
      }
    }
    for (; r < 16; r++) pk.coef[b][r] = pk.deg[b][r] = -1;
  }
  return pk;
}

__constant__ int cCoef[3][16];   // logical row -> coefficient
__constant__ float cA[2][3][16][16];  // packed matrices, rows permuted (row 4g+i = logical g+4i)

constexpr int TILES = 4;  // independent column tiles per wave
typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_mfma(uint32_t npairs, int apps, float* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, col = lane & 15;
  const uint32_t wave = (blockIdx.x * 256 + threadIdx.x) >> 6;
  const uint32_t pair0 = wave * (8 * TILES);
  if (pair0 >= npairs) return;
  // A operands: step s of block b, matrix t: A[row = col][k = 4 s + g]
  float A[2][3][4];
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
      for (int s = 0; s < 4; s++) A[t][b][s] = cA[t][b][col][4 * s + g];
  // B operands: tile u, block b, step s: logical row g + 4 s of column col
  float X[TILES][3][4];
#pragma unroll
  for (int u = 0; u < TILES; u++)
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
      for (int s = 0; s < 4; s++) {
        const int c = cCoef[b][g + 4 * s];
        const uint32_t j = pair0 + 8 * u + (col >> 1);
        X[u][b][s] = c >= 0 ? init_value(j, c, col & 1) : 0.f;
      }
  for (int a = 0; a < apps; a++) {
    const int t = a & 1;
#pragma unroll
    for (int u = 0; u < TILES; u++)
#pragma unroll
      for (int b = 0; b < 3; b++) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 4; s++)
          d = __builtin_amdgcn_mfma_f32_16x16x4f32(A[t][b][s], X[u][b][s], d, 0, 0, 0);
#pragma unroll
        for (int s = 0; s < 4; s++) X[u][b][s] = d[s];
      }
  }
  // the same weighted sum as k_valu, per pair: the lane's rows, then the
  // four lane groups of the column, then real and imaginary columns
#pragma unroll
  for (int u = 0; u < TILES; u++) {
    float s = 0.f;
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int c = cCoef[b][g + 4 * q];
        if (c >= 0) s += X[u][b][q] * (float)((col & 1) ? 2 * c + 1 : c + 1);
      }
    s += __shfl_xor(s, 16, 64);
    s += __shfl_xor(s, 32, 64);
    s += __shfl_xor(s, 1, 64);
    const uint32_t j = pair0 + 8 * u + (col >> 1);
    if (lane < 16 && (col & 1) == 0 && j < npairs) out[j] = s;
  }
}

int main(int argc, char** argv) {
  const int lg = argc > 1 ? std::atoi(argv[1]) : 24;
  const int apps = argc > 2 ? std::atoi(argv[2]) : 32;
  if (lg < 8 || lg > 26 || apps < 1 || apps > 4096) {
    std::fprintf(stderr, "usage: rot_products [log2 pairs 8..26] [applications 1..4096]\n");
    return 2;
  }
  const uint32_t npairs = 1u << lg;  // a multiple of 8 TILES pairs per wave
  // matrices: orthonormal DCT-II of size n+1, and its transpose
  std::vector<float> hb(2 * MW);
  for (int n = 0; n <= P; n++) {
    const int N = n + 1;
    for (int m = 0; m < N; m++)
      for (int k = 0; k < N; k++) {
        const double cm = m == 0 ? std::sqrt(1.0 / N) : std::sqrt(2.0 / N);
        const double v = cm * std::cos(M_PI * (2 * k + 1) * m / (2.0 * N));
        hb[moff(n) + m * N + k] = (float)v;
        hb[MW + moff(n) + k * N + m] = (float)v;
      }
  }
  CK(hipMemcpyToSymbol(HIP_SYMBOL(cB), hb.data(), hb.size() * sizeof(float)));
  const Pack pk = make_pack();
  std::vector<float> ha(2 * 3 * 16 * 16, 0.f);
  for (int t = 0; t < 2; t++)
    for (int b = 0; b < 3; b++)
      for (int rp = 0; rp < 16; rp++) {
        const int r = (rp >> 2) + 4 * (rp & 3);  // stored row 4g+i holds logical row g+4i
        if (pk.coef[b][r] < 0) continue;
        const int n = pk.deg[b][r];
        const int m = pk.coef[b][r] - cidx(n, 0);
        for (int k = 0; k < 16; k++) {
          if (pk.deg[b][k] != n) continue;
          const int kk = pk.coef[b][k] - cidx(n, 0);
          ha[((t * 3 + b) * 16 + rp) * 16 + k] = hb[t * MW + moff(n) + m * (n + 1) + kk];
        }
      }
  CK(hipMemcpyToSymbol(HIP_SYMBOL(cA), ha.data(), ha.size() * sizeof(float)));
  CK(hipMemcpyToSymbol(HIP_SYMBOL(cCoef), pk.coef, sizeof(pk.coef)));
  float *ov = nullptr, *om = nullptr;
  CK(hipMalloc(&ov, npairs * sizeof(float)));
  CK(hipMalloc(&om, npairs * sizeof(float)));
  const dim3 gv((npairs + 255) / 256), gm((npairs / (8 * TILES) * 64 + 255) / 256);
  hipEvent_t e0, e1, e2;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  CK(hipEventCreate(&e2));
  float tv = 1e30f, tm = 1e30f;
  for (int rep = 0; rep < 3; rep++) {  // best of three (the first includes warm-up)
    CK(hipEventRecord(e0));
    hipLaunchKernelGGL(k_valu, gv, dim3(256), 0, 0, npairs, apps, ov);
    CK(hipEventRecord(e1));
    hipLaunchKernelGGL(k_mfma, gm, dim3(256), 0, 0, npairs, apps, om);
    CK(hipEventRecord(e2));
    CK(hipEventSynchronize(e2));
    float a = 0.f, b = 0.f;
    CK(hipEventElapsedTime(&a, e0, e1));
    CK(hipEventElapsedTime(&b, e1, e2));
    tv = std::fmin(tv, a);
    tm = std::fmin(tm, b);
  }
  std::vector<float> hv(npairs), hm(npairs);
  CK(hipMemcpy(hv.data(), ov, npairs * sizeof(float), hipMemcpyDeviceToHost));
  CK(hipMemcpy(hm.data(), om, npairs * sizeof(float), hipMemcpyDeviceToHost));
  double worst = 0.0, scale = 0.0;
  for (uint32_t j = 0; j < npairs; j++) scale = std::fmax(scale, std::fabs((double)hv[j]));
  for (uint32_t j = 0; j < npairs; j++)
    worst = std::fmax(worst, std::fabs((double)hv[j] - (double)hm[j]) / scale);
  const double pa = (double)npairs * apps;
  std::printf(
      "{\"p\": %d, \"pairs\": %u, \"applications\": %d, \"valu_ms\": %.4f, \"mfma_ms\": %.4f, "
      "\"valu_ns_per_pair_app\": %.5f, \"mfma_ns_per_pair_app\": %.5f, \"mfma_over_valu\": %.4f, "
      "\"useful_macs_per_pair_app\": %d, \"max_rel_diff\": %.3e}\n",
      P, npairs, apps, tv, tm, tv * 1e6 / pa, tm * 1e6 / pa, tm / tv, 2 * MW, worst);
  CK(hipFree(ov));
  CK(hipFree(om));
  return worst < 1e-4 ? 0 : 3;
}
