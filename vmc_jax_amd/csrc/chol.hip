// Cholesky factorisation S = L L^T of real symmetric positive definite
// matrices (row-major, lower triangle) and the inverse of the factor.
//
// Right-looking, 64-column panels.  Per panel one launch factors the 64 x 64
// diagonal block in LDS (every workgroup redundantly, so no workgroup waits
// for another) and solves the rows below against it, 64 rows per workgroup;
// one launch subtracts the panel's outer product from the trailing lower
// triangle with FP64 MFMA tiles.  Two matrices go through the same launches
// (blockIdx.y), which costs next to nothing for these latency-bound sizes.
// The other workgroups of the panel launch read the diagonal block's input
// values, so workgroup 0 leaves the factored block in a scratch buffer and the
// trailing update (the next launch, which never reads that block) copies it
// into place; a panel launch of one workgroup writes it directly.
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

// Lane `lane` (a compile-time constant after unrolling) of v to every lane,
// as a scalar operand.
__device__ inline double readlane_d(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// Lane S of each group of four lanes to the whole group (DPP quad_perm
// broadcast: a VALU move, no LDS round trip).
template <int S>
__device__ inline double quad_bcast(double v) {
  constexpr int ctrl = S * 0x55;  // quad_perm [S, S, S, S]
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), ctrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), ctrl, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ inline double quad_bcast(double v, int s) {
  switch (s & 3) {
    case 0: return quad_bcast<0>(v);
    case 1: return quad_bcast<1>(v);
    case 2: return quad_bcast<2>(v);
    default: return quad_bcast<3>(v);
  }
}

// panel k0: diagonal block (blockIdx.x == 0) and row blocks below it.  256
// threads.  The diagonal block is factored right-looking by one wave, row t
// in the registers of lane t (per column: pivot and scaling, then the
// rank-one update of the trailing triangle).  A row block is solved with four
// threads per row, thread
// q of a row holding columns q, q + 4, ... in registers: per column the owner
// finishes x_j, the row's four threads exchange it by shuffle and update
// their later columns.
__global__ void __launch_bounds__(256) chol_panel(double* A0, double* A1, int p, int k0, int* info,
                                                  double* __restrict__ dscr) {
  chain_priority();
  double* A = blockIdx.y ? A1 : A0;
  __shared__ double Ld[CB][CB + 1];
  const int tid = threadIdx.x;
  const int b = min(CB, p - k0);
  {  // all 16 loads in flight together (clamped, always valid addresses), then LDS
    double v[(CB * CB) / 256];
#pragma unroll
    for (int u = 0; u < (CB * CB) / 256; ++u) {
      const int e = tid + 256 * u, r = e / CB, c = e % CB;
      const int rr = min(r, b - 1), cc = min(c, rr);
      v[u] = A[(size_t)(k0 + rr) * p + k0 + cc];
    }
#pragma unroll
    for (int u = 0; u < (CB * CB) / 256; ++u) {
      const int e = tid + 256 * u, r = e / CB, c = e % CB;
      Ld[r][c] = (r < b && c <= r) ? v[u] : 0.0;
    }
  }
  __syncthreads();
  // The diagonal block is factored by wave 0 alone, lane t holding row t in
  // registers (columns fully unrolled).  The pivot and the scaled column
  // entries L_cj reach the other lanes by readlane (scalar operands of the
  // update), not through LDS: per column the dependent chain is the pivot,
  // its square root, the quotient and one FMA.  (An LDS exchange of the
  // column took 104k cycles per block; rows in LDS with a runtime loop over
  // the pivots 330k; four waves with a quarter of each row and one barrier
  // per column 112k.)
  __shared__ double rdiag[CB];  // 1 / L_jj for the row solve
  if (tid < 64) {
    const int t = tid;
    int first_bad = 0;  // 1-based column of the first non-positive pivot
    double x[CB];
#pragma unroll
    for (int c = 0; c < CB; ++c) x[c] = (t < b && c <= t) ? Ld[t][c] : 0.0;
#pragma unroll
    for (int j = 0; j < CB; ++j) {
      if (j < b) {
        const double dj = readlane_d(x[j], j);
        const double sj = sqrt(dj);
        if (first_bad == 0 && !(dj > 0.0)) first_bad = j + 1;
        double l = 0.0;
        if (t > j) {
          l = x[j] / sj;
          x[j] = l;
        } else if (t == j) {
          x[j] = sj;
        }
#pragma unroll
        for (int c = j + 1; c < CB; ++c) {
          const double lc = readlane_d(l, c);
          if (c <= t) x[c] -= l * lc;
        }
      }
    }
    double dt = 1.0;
    if (t < b) {
#pragma unroll
      for (int c = 0; c < CB; ++c)
        if (c <= t) {
          Ld[t][c] = x[c];
          if (c == t) dt = x[c];
        }
    }
    rdiag[t] = 1.0 / dt;
    if (blockIdx.x == 0 && t == 0 && first_bad != 0 && info[blockIdx.y] == 0) info[blockIdx.y] = k0 + first_bad;
  }
  __syncthreads();
  if (blockIdx.x == 0) {
    const bool alone = gridDim.x == 1;
    double* dst = alone ? A + (size_t)k0 * p + k0 : dscr + (size_t)blockIdx.y * CB * CB;
    const int ld = alone ? p : CB;
    for (int e = tid; e < CB * CB; e += 256) {
      const int r = e / CB, c = e % CB;
      if (r < b && c <= r) dst[(size_t)r * ld + c] = Ld[r][c];
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
  for (int k = 0; k < CB / 4; ++k) {
    const double v = row[q + 4 * k];  // row points at a valid row even when !live
    x[k] = live ? v : 0.0;
  }
  // x_j = (a_j - sum_{l<j} x_l L_jl) / L_jj by the reciprocal; the owner of
  // column j broadcasts x_j to its row's four lanes by DPP
#pragma unroll
  for (int j = 0; j < CB; ++j) {
    if (q == (j & 3)) x[j >> 2] *= rdiag[j];
    const double xj = quad_bcast(x[j >> 2], j);
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
__global__ void __launch_bounds__(256) chol_update(double* A0, double* A1, int p, int k0,
                                                   const double* __restrict__ dscr) {
  chain_priority();
  double* A = blockIdx.y ? A1 : A0;
  if (blockIdx.x == 0) {  // the factored diagonal block of panel k0 (a full CB x CB block here)
    const double* src = dscr + (size_t)blockIdx.y * CB * CB;
    for (int e = threadIdx.x; e < CB * CB; e += 256) {
      const int r = e / CB, c = e % CB;
      if (c <= r) A[(size_t)(k0 + r) * p + k0 + c] = src[e];
    }
  }
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
    double va[(CB * 32) / 256], vb[(CB * 32) / 256];
#pragma unroll
    for (int r = 0; r < (CB * 32) / 256; ++r) {
      const int e = tid + 256 * r;
      const int m = e / 32, k = e % 32;
      va[r] = A[(size_t)min(i0 + m, p - 1) * p + k0 + kc + k];
      vb[r] = A[(size_t)min(j0 + m, p - 1) * p + k0 + kc + k];
    }
#pragma unroll
    for (int r = 0; r < (CB * 32) / 256; ++r) {
      const int e = tid + 256 * r;
      const int m = e / 32, k = e % 32;
      As[k][m] = i0 + m < p ? va[r] : 0.0;
      Bs[k][m] = j0 + m < p ? vb[r] : 0.0;
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

// W = L^{-1} for the lower-triangular L, right-looking and in place on W = I
// (row-major; the strict upper triangle stays zero).  Block row i: every tile
// W_ij (j <= i) becomes L_ii^{-1} W_ij (inv_row: forward substitution with
// L_ii in LDS, four threads per column, thread q owning rows q, q + 4, ...),
// then every tile below takes W_kj -= L_ki W_ij (inv_update: FP64 MFMA
// tiles, k > i, j <= i).
__global__ void inv_identity(double* __restrict__ W, int64_t p) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < p * p) W[t] = (t / p == t % p) ? 1.0 : 0.0;
}

__global__ void __launch_bounds__(256) inv_row(const double* __restrict__ L, int p, int i0, double* W) {
  chain_priority();
  __shared__ double Ls[CB][CB + 1];
  __shared__ double rdl[CB];  // 1 / L_rr
  const int tid = threadIdx.x;
  const int c0 = CB * blockIdx.x;  // tile column block j <= i
  const int nb = min(CB, p - i0);
  {
    double v[(CB * CB) / 256];
#pragma unroll
    for (int u = 0; u < (CB * CB) / 256; ++u) {
      const int e = tid + 256 * u, r = e / CB, c = e % CB;
      const int rr = min(r, nb - 1), cc = min(c, rr);
      v[u] = L[(size_t)(i0 + rr) * p + i0 + cc];
    }
#pragma unroll
    for (int u = 0; u < (CB * CB) / 256; ++u) {
      const int e = tid + 256 * u, r = e / CB, c = e % CB;
      const double lv = (r < nb && c <= r) ? v[u] : (r == c ? 1.0 : 0.0);
      Ls[r][c] = lv;
      if (r == c) rdl[r] = 1.0 / lv;
    }
  }
  __syncthreads();
  const int col = tid / 4, q = tid % 4;
  const int gc = c0 + col;
  const bool live = gc < p;
  double x[CB / 4];
#pragma unroll
  for (int u = 0; u < CB / 4; ++u) {
    const int gi = i0 + q + 4 * u;
    const double v = W[(size_t)min(gi, p - 1) * p + min(gc, p - 1)];
    x[u] = (live && gi < p) ? v : 0.0;
  }
  // forward substitution by the reciprocal diagonal; the owner of row r
  // broadcasts x_r to its column's four lanes by DPP
#pragma unroll
  for (int r = 0; r < CB; ++r) {
    if (q == (r & 3)) x[r >> 2] *= rdl[r];
    const double xr = quad_bcast(x[r >> 2], r);
#pragma unroll
    for (int u = 0; u < CB / 4; ++u)
      if (q + 4 * u > r) x[u] -= Ls[q + 4 * u][r] * xr;
  }
#pragma unroll
  for (int u = 0; u < CB / 4; ++u) {
    const int gi = i0 + q + 4 * u;
    if (live && gi < p) W[(size_t)gi * p + gc] = x[u];
  }
}

// tiles (k, j): k over the block rows below i0 (blockIdx.x / (i + 1)), j <= i
__global__ void __launch_bounds__(256) inv_update(const double* __restrict__ L, int p, int i0, double* W) {
  chain_priority();
  const int nj = i0 / CB + 1;
  const int kt = blockIdx.x / nj, jt = blockIdx.x % nj;
  const int r0 = i0 + CB * (kt + 1), c0 = CB * jt;
  __shared__ double As[32][CB + 1], Bs[32][CB + 1];
  const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
  const int wm = (wv / 2) * 32, wn = (wv % 2) * 32;
  v4d acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int c = 0; c < 2; ++c) acc[a][c] = v4d{0, 0, 0, 0};
  for (int kc = 0; kc < CB; kc += 32) {
    double va[(CB * 32) / 256], vb[(CB * 32) / 256];
#pragma unroll
    for (int u = 0; u < (CB * 32) / 256; ++u) {
      const int e = tid + 256 * u;
      const int m = e / 32, k = e % 32;  // As[k][m] = L[r0 + m][i0 + kc + k]
      va[u] = L[(size_t)min(r0 + m, p - 1) * p + i0 + kc + k];
      const int n = e % CB, kb = e / CB;  // Bs[kb][n] = W[i0 + kc + kb][c0 + n]
      vb[u] = W[(size_t)min(i0 + kc + kb, p - 1) * p + min(c0 + n, p - 1)];
    }
#pragma unroll
    for (int u = 0; u < (CB * 32) / 256; ++u) {
      const int e = tid + 256 * u;
      const int m = e / 32, k = e % 32;
      As[k][m] = r0 + m < p ? va[u] : 0.0;
      const int n = e % CB, kb = e / CB;
      Bs[kb][n] = (i0 + kc + kb < p && c0 + n < p) ? vb[u] : 0.0;
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
        const int gi = r0 + wm + 16 * a + 4 * r + lane / 16;
        const int gj = c0 + wn + 16 * c + lane % 16;
        if (gi < p && gj < p) W[(size_t)gi * p + gj] -= acc[a][c][r];
      }
}

}  // namespace

// Factors A0 and (when A1 is not null) A1 in place; info[0], info[1] receive
// the first non-positive pivot (1-based) or 0.
int chol_factor2(double* A0, double* A1, int p, int* info, hipStream_t st) {
  const unsigned nmat = A1 ? 2 : 1;
  double* dscr = (double*)workspace(sizeof(double) * 2 * CB * CB, 103);
  if (!dscr) return fail(JV_E_NOMEM, "cholesky: workspace allocation failed");
  hipLaunchKernelGGL(chol_zero_info, dim3(1), dim3(64), 0, st, info);
  JV_LAUNCH_CHECK("chol_zero_info");
  for (int k0 = 0; k0 < p; k0 += CB) {
    const int b = std::min(CB, p - k0);
    const int below = p - k0 - b;
    const unsigned gx = 1 + (unsigned)((below + CB - 1) / CB);
    hipLaunchKernelGGL(chol_panel, dim3(gx, nmat), dim3(256), 0, st, A0, A1 ? A1 : A0, p, k0, info, dscr);
    JV_LAUNCH_CHECK("chol_panel");
    if (below > 0) {
      const int T = (below + CB - 1) / CB;
      hipLaunchKernelGGL(chol_update, dim3((unsigned)(T * (T + 1) / 2), nmat), dim3(256), 0, st, A0, A1 ? A1 : A0, p,
                         k0, dscr);
      JV_LAUNCH_CHECK("chol_update");
    }
  }
  return JV_OK;
}

// W = L^{-1} of the factor L (W and L may not alias).
int chol_inverse_lower(const double* L, int p, double* W, hipStream_t st) {
  const int64_t pp = (int64_t)p * p;
  hipLaunchKernelGGL(inv_identity, dim3((unsigned)((pp + 255) / 256)), dim3(256), 0, st, W, (int64_t)p);
  JV_LAUNCH_CHECK("inv_identity");
  const int nblk = (p + CB - 1) / CB;
  for (int bi = 0; bi < nblk; ++bi) {
    const int i0 = CB * bi;
    hipLaunchKernelGGL(inv_row, dim3((unsigned)(bi + 1)), dim3(256), 0, st, L, p, i0, W);
    JV_LAUNCH_CHECK("inv_row");
    const int below = nblk - bi - 1;
    if (below > 0) {
      hipLaunchKernelGGL(inv_update, dim3((unsigned)(below * (bi + 1))), dim3(256), 0, st, L, p, i0, W);
      JV_LAUNCH_CHECK("inv_update");
    }
  }
  return JV_OK;
}

}  // namespace jv
