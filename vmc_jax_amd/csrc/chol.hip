// Cholesky factorisation S = L L^T of real symmetric positive definite
// matrices (row-major, lower triangle) and the solve L L^T x = b.
//
// Right-looking, 64-column panels.  Per panel one launch factors the 64 x 64
// diagonal block in LDS (every workgroup redundantly, so no workgroup waits
// for another) and solves the rows below against it, 64 rows per workgroup;
// one launch subtracts the panel's outer product from the trailing lower
// triangle with FP64 MFMA tiles.  Two matrices go through the same launches
// (blockIdx.y), which costs next to nothing for these latency-bound sizes.
// A pivot that is not positive (or NaN) records its 1-based index in info;
// the remaining arithmetic then runs on NaN without any data-dependent
// addressing.  Only the lower triangle is read; the strict upper triangle
// of each diagonal block keeps its input values, the rest of the upper
// triangle is updated but never read.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"

namespace jv {

namespace {

constexpr int CB = 64;  // panel width

typedef double v4d __attribute__((ext_vector_type(4)));

// panel k0: diagonal block (blockIdx.x == 0) and row blocks below it.  256
// threads.  The diagonal block is factored right-looking in LDS (per column:
// pivot and scaling, then the rank-one update of the trailing triangle spread
// over all threads).  A row block is solved with four threads per row, thread
// q of a row holding columns q, q + 4, ... in registers: per column the owner
// finishes x_j, the row's four threads exchange it by shuffle and update
// their later columns.
__global__ void __launch_bounds__(256) chol_panel(double* A0, double* A1, int p, int k0, int* info) {
  double* A = blockIdx.y ? A1 : A0;
  __shared__ double Ld[CB][CB + 1];
  const int tid = threadIdx.x;
  const int b = min(CB, p - k0);
  for (int e = tid; e < CB * CB; e += 256) {
    const int r = e / CB, c = e % CB;
    if (r < b && c <= r) Ld[r][c] = A[(size_t)(k0 + r) * p + k0 + c];
  }
  __syncthreads();
  for (int j = 0; j < b; ++j) {
    const double dj = Ld[j][j];
    const double sj = sqrt(dj);
    if (blockIdx.x == 0 && tid == 0 && !(dj > 0.0) && info[blockIdx.y] == 0) info[blockIdx.y] = k0 + j + 1;
    const int m = b - j - 1;
    __syncthreads();
    if (tid < m) Ld[j + 1 + tid][j] /= sj;
    if (tid == 0) Ld[j][j] = sj;
    __syncthreads();
    for (int e = tid; e < m * m; e += 256) {
      const int t = j + 1 + e / m, c = j + 1 + e % m;
      if (c <= t) Ld[t][c] -= Ld[t][j] * Ld[c][j];
    }
    __syncthreads();
  }
  if (blockIdx.x == 0) {
    for (int e = tid; e < CB * CB; e += 256) {
      const int r = e / CB, c = e % CB;
      if (r < b && c <= r) A[(size_t)(k0 + r) * p + k0 + c] = Ld[r][c];
    }
    return;
  }
  // rows below (only when b == CB): x L_kk^T = a
  const int r = tid / 4, q = tid % 4;
  const int i = k0 + CB * blockIdx.x + r;
  const bool live = i < p;
  double* row = A + (size_t)(live ? i : k0) * p + k0;
  double x[CB / 4];
#pragma unroll
  for (int k = 0; k < CB / 4; ++k) x[k] = live ? row[q + 4 * k] : 0.0;
  const int base = (threadIdx.x % 64) & ~3;
#pragma unroll
  for (int j = 0; j < CB; ++j) {
    if (q == (j & 3)) x[j >> 2] /= Ld[j][j];
    const double xj = __shfl(x[j >> 2], base + (j & 3), 64);
#pragma unroll
    for (int k = 0; k < CB / 4; ++k)
      if (q + 4 * k > j) x[k] -= xj * Ld[q + 4 * k][j];
  }
  if (live) {
#pragma unroll
    for (int k = 0; k < CB / 4; ++k) row[q + 4 * k] = x[k];
  }
}

// A[i][j] -= sum_l L[i][k0 + l] L[j][k0 + l] over the lower tiles of the
// trailing matrix (rows / columns >= k0 + CB).  256 threads: four waves of
// 32 x 32, FP64 MFMA 16x16x4 (A operand A[row = lane % 16][k = lane / 16],
// B operand B[k = lane / 16][col = lane % 16], result element r of lane l at
// C[4 r + l / 16][l % 16]).
__global__ void __launch_bounds__(256) chol_update(double* A0, double* A1, int p, int k0) {
  double* A = blockIdx.y ? A1 : A0;
  const int k1 = k0 + CB;
  // linear lower-tile index -> (ti, tj), ti >= tj
  int bidx = blockIdx.x, ti = 0;
  while (bidx > ti) {
    bidx -= ti + 1;
    ++ti;
  }
  const int tj = bidx;
  const int i0 = k1 + CB * ti, j0 = k1 + CB * tj;
  __shared__ double As[32][CB + 1], Bs[32][CB + 1];
  const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
  const int wm = (wv / 2) * 32, wn = (wv % 2) * 32;
  v4d acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int c = 0; c < 2; ++c) acc[a][c] = v4d{0, 0, 0, 0};
  for (int kc = 0; kc < CB; kc += 32) {
    for (int r = 0; r < (CB * 32) / 256; ++r) {
      const int e = tid + 256 * r;
      const int m = e / 32, k = e % 32;
      const int gi = i0 + m, gj = j0 + m;
      As[k][m] = gi < p ? A[(size_t)gi * p + k0 + kc + k] : 0.0;
      Bs[k][m] = gj < p ? A[(size_t)gj * p + k0 + kc + k] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 32; ks += 4) {
      double ar[2], br[2];
      for (int a = 0; a < 2; ++a) ar[a] = As[ks + lane / 16][wm + 16 * a + lane % 16];
      for (int c = 0; c < 2; ++c) br[c] = Bs[ks + lane / 16][wn + 16 * c + lane % 16];
      for (int a = 0; a < 2; ++a)
        for (int c = 0; c < 2; ++c) acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[c], acc[a][c], 0, 0, 0);
    }
    __syncthreads();
  }
  for (int a = 0; a < 2; ++a)
    for (int c = 0; c < 2; ++c)
      for (int r = 0; r < 4; ++r) {
        const int gi = i0 + wm + 16 * a + 4 * r + lane / 16;
        const int gj = j0 + wn + 16 * c + lane % 16;
        if (gi < p && gj < p && gi >= gj) A[(size_t)gi * p + gj] -= acc[a][c][r];
      }
}

__global__ void chol_zero_info(int* info) {
  if (threadIdx.x < 2) info[threadIdx.x] = 0;
}

// x = (L L^T)^{-1} b with L the lower triangle of the row-major p x p matrix.
// One workgroup of 1024 threads, left-looking over blocks of 64 unknowns: all
// threads form the block's right-hand side from the unknowns already solved
// (independent loads, several accumulators per thread), then wave 0 solves
// the diagonal block from LDS.  The unknowns live in LDS (XL) or in x.
constexpr int SOLVE_LDS_MAX = 3000;

template <bool XL>
__global__ void __launch_bounds__(1024) chol_solve_kernel(const double* __restrict__ L, int p,
                                                         const double* __restrict__ b, double* x) {
  __shared__ double Ld[CB][CB + 1];
  __shared__ double red[16][CB];
  __shared__ double xl[XL ? SOLVE_LDS_MAX : 1];
  double* xv = XL ? xl : x;
  const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
  const int nblk = (p + CB - 1) / CB;
  // forward: L y = b
  for (int bi = 0; bi < nblk; ++bi) {
    const int kb = bi * CB, nb = min(CB, p - kb);
    for (int e = tid; e < CB * CB; e += 1024) {
      const int r = e / CB, c = e % CB;
      if (r < nb && c <= r) Ld[r][c] = L[(size_t)(kb + r) * p + kb + c];
    }
    // row r = tid / 16 of the block, 16 threads per row over the columns < kb
    {
      const int r = tid / 16, sg = tid % 16;
      double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (r < nb) {
        const double* Lr = L + (size_t)(kb + r) * p;
        int l = sg;
        for (; l + 112 < kb; l += 128) {
#pragma unroll
          for (int u = 0; u < 8; ++u) s[u] += Lr[l + 16 * u] * xv[l + 16 * u];
        }
        for (; l < kb; l += 16) s[0] += Lr[l] * xv[l];
      }
      double t = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
      for (int o = 8; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
      if (sg == 0) red[0][r] = t;
    }
    __syncthreads();
    if (wv == 0) {
      double r = lane < nb ? b[kb + lane] - red[0][lane] : 0.0;
      for (int j = 0; j < nb; ++j) {
        const double yj = __shfl(r, j, 64) / Ld[j][j];
        if (lane == j) r = yj;
        else if (lane > j && lane < nb) r -= Ld[lane][j] * yj;
      }
      if (lane < nb) xv[kb + lane] = r;
    }
    __syncthreads();
  }
  // backward: L^T x = y, blocks from the last; column c of the block gets
  // sum over the solved rows below of L[i][c] x[i] (lanes over c, waves over i)
  for (int bi = nblk - 1; bi >= 0; --bi) {
    const int kb = bi * CB, nb = min(CB, p - kb);
    for (int e = tid; e < CB * CB; e += 1024) {
      const int r = e / CB, c = e % CB;
      if (r < nb && c <= r) Ld[r][c] = L[(size_t)(kb + r) * p + kb + c];
    }
    {
      double s[4] = {0, 0, 0, 0};
      if (lane < nb) {
        const double* Lc = L + kb + lane;
        int i = kb + nb + wv;
        for (; i + 48 < p; i += 64) {
#pragma unroll
          for (int u = 0; u < 4; ++u) s[u] += Lc[(size_t)(i + 16 * u) * p] * xv[i + 16 * u];
        }
        for (; i < p; i += 16) s[0] += Lc[(size_t)i * p] * xv[i];
      }
      red[wv][lane] = (s[0] + s[1]) + (s[2] + s[3]);
    }
    __syncthreads();
    if (wv == 0) {
      double s = 0.0;
      for (int w = 0; w < 16; ++w) s += red[w][lane];
      double r = lane < nb ? xv[kb + lane] - s : 0.0;
      for (int j = nb - 1; j >= 0; --j) {
        const double xj = __shfl(r, j, 64) / Ld[j][j];
        if (lane == j) r = xj;
        else if (lane < j) r -= Ld[j][lane] * xj;
      }
      if (lane < nb) xv[kb + lane] = r;
    }
    __syncthreads();
  }
  if (XL)
    for (int i = tid; i < p; i += 1024) x[i] = xl[i];
}

}  // namespace

// Factors A0 and (when A1 is not null) A1 in place; info[0], info[1] receive
// the first non-positive pivot (1-based) or 0.
int chol_factor2(double* A0, double* A1, int p, int* info, hipStream_t st) {
  const unsigned nmat = A1 ? 2 : 1;
  hipLaunchKernelGGL(chol_zero_info, dim3(1), dim3(64), 0, st, info);
  JV_LAUNCH_CHECK("chol_zero_info");
  for (int k0 = 0; k0 < p; k0 += CB) {
    const int b = std::min(CB, p - k0);
    const int below = p - k0 - b;
    const unsigned gx = 1 + (unsigned)((below + CB - 1) / CB);
    hipLaunchKernelGGL(chol_panel, dim3(gx, nmat), dim3(256), 0, st, A0, A1 ? A1 : A0, p, k0, info);
    JV_LAUNCH_CHECK("chol_panel");
    if (below > 0) {
      const int T = (below + CB - 1) / CB;
      hipLaunchKernelGGL(chol_update, dim3((unsigned)(T * (T + 1) / 2), nmat), dim3(256), 0, st, A0, A1 ? A1 : A0, p,
                         k0);
      JV_LAUNCH_CHECK("chol_update");
    }
  }
  return JV_OK;
}

// x = (L L^T)^{-1} b (x and b may not alias).
int chol_solve(const double* L, int p, const double* b, double* x, hipStream_t st) {
  if (p <= SOLVE_LDS_MAX) hipLaunchKernelGGL(chol_solve_kernel<true>, dim3(1), dim3(1024), 0, st, L, p, b, x);
  else hipLaunchKernelGGL(chol_solve_kernel<false>, dim3(1), dim3(1024), 0, st, L, p, b, x);
  JV_LAUNCH_CHECK("chol_solve_kernel");
  return JV_OK;
}

}  // namespace jv
