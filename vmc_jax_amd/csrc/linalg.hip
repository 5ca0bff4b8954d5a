// Sample statistics, FP64 MFMA Gram / GEMM kernels, rocSOLVER eigendecompositions
// and the MinSR / TDVP solves.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "common.hpp"

namespace jv {

// tridiag.hip
bool trd_supported(int64_t n);
size_t trd_work_doubles(int n);
int trd_eig_solve(double* S, int n, double* ev, double* c, double* Zbuf, double* work, hipStream_t st);
int trd_apply(const double* V, const double* tau, int n, double* x, double* Tbuf, hipStream_t st);
size_t trd_wy_offset(int n);
int trd_eigh_full(double* S, int n, double* ev, hipStream_t st);
bool trdc_supported(int64_t n);
int trdc_eigh_full(double* S, int n, double* ev, hipStream_t st);
int trd_reduce(double* S, int n, double* ev, double* c, double* work, hipStream_t st);
int trd_tridiag_eig(int n, double* ev, double* work, double* Zbuf, hipStream_t st);
int trd_wy_build(const double* V, const double* tau, int n, double* Tbuf, hipStream_t st);
int trd_wy_apply(const double* V, int n, double* x, const double* Tbuf, bool transpose, hipStream_t st);
bool aux_stream(hipStream_t* st, hipEvent_t* e0, hipEvent_t* e1);

rocblas_handle blas_handle(hipStream_t stream);
int gram_pp_f64(const double* D, int64_t n, int64_t p, int cpx, double* S, hipStream_t st, int slot);

typedef double v4d __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ centring
// Column partial sums of w_n * O[n][j] over row chunks (deterministic order).
constexpr int CHUNK_ROWS = 256;
// Row chunk of the transposed GEMV's partial sums: 64 rows give 8192 x 1240
// data 640 workgroups (5 column blocks x 128 chunks; 160 with 256 rows, too
// few to keep the loads of 256 CUs in flight).
constexpr int GEMV_T_ROWS = 64;

// Column partial sums of one CHUNK_ROWS chunk per blockIdx.y: a workgroup of
// COLS_PER_WG columns x ROW_GROUPS row groups (threadIdx.y), each thread summing
// every ROW_GROUPS-th row of the chunk, the groups combined through LDS (four
// times the workgroups of one thread per column: 8192 x 1240 data fill 640).
constexpr int COLS_PER_WG = 64, ROW_GROUPS = 4;

template <typename TIn>
__global__ void __launch_bounds__(COLS_PER_WG * ROW_GROUPS) col_wsum_partial(const TIn* __restrict__ O,
                                                                             const double* __restrict__ w, int64_t n,
                                                                             int64_t p, int cpx,
                                                                             double* __restrict__ part) {
  chain_priority();
  __shared__ double sr[ROW_GROUPS][COLS_PER_WG], si[ROW_GROUPS][COLS_PER_WG];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int64_t j = (int64_t)blockIdx.x * COLS_PER_WG + tx;
  const int64_t r0 = (int64_t)blockIdx.y * CHUNK_ROWS, r1 = std::min<int64_t>(n, r0 + CHUNK_ROWS);
  double ar = 0.0, ai = 0.0;
  if (j < p) {
    if (cpx) {
      for (int64_t r = r0 + ty; r < r1; r += ROW_GROUPS) {
        ar = fma(w[r], O[2 * (r * p + j)], ar);
        ai = fma(w[r], O[2 * (r * p + j) + 1], ai);
      }
    } else {
      for (int64_t r = r0 + ty; r < r1; r += ROW_GROUPS) ar = fma(w[r], O[r * p + j], ar);
    }
  }
  sr[ty][tx] = ar;
  si[ty][tx] = ai;
  __syncthreads();
  if (ty == 0 && j < p) {
    ar = (sr[0][tx] + sr[1][tx]) + (sr[2][tx] + sr[3][tx]);
    if (cpx) {
      ai = (si[0][tx] + si[1][tx]) + (si[2][tx] + si[3][tx]);
      part[2 * (blockIdx.y * p + j)] = ar;
      part[2 * (blockIdx.y * p + j) + 1] = ai;
    } else {
      part[blockIdx.y * p + j] = ar;
    }
  }
}

// col_wsum_partial of a single column (p = 1, e.g. E_loc): one row per
// thread (blockDim.x = CHUNK_ROWS) and a workgroup sum, written to the same
// partial layout part[chunk * len + component].
template <typename TIn>
__global__ void vec_wsum_partial(const TIn* __restrict__ O, const double* __restrict__ w, int64_t n, int cpx,
                                 double* __restrict__ part) {
  chain_priority();
  __shared__ double red[CHUNK_ROWS / WAVE];
  const int64_t r = (int64_t)blockIdx.y * CHUNK_ROWS + threadIdx.x;
  double ar = 0.0, ai = 0.0;
  if (r < n) {
    ar = w[r] * (double)O[cpx ? 2 * r : r];
    if (cpx) ai = w[r] * (double)O[2 * r + 1];
  }
  ar = block_sum(ar, red);
  if (cpx) ai = block_sum(ai, red);
  if (threadIdx.x == 0) {
    part[(int64_t)blockIdx.y * (cpx ? 2 : 1)] = ar;
    if (cpx) part[(int64_t)blockIdx.y * 2 + 1] = ai;
  }
}

__global__ void col_reduce(const double* __restrict__ part, int nchunks, int64_t len, double* __restrict__ out) {
  chain_priority();
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= len) return;
  // four chunks per iteration into independent sums (loads in flight together)
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int c = 0;
  for (; c + 3 < nchunks; c += 4) {
    a0 += part[c * len + j];
    a1 += part[(c + 1) * len + j];
    a2 += part[(c + 2) * len + j];
    a3 += part[(c + 3) * len + j];
  }
  for (; c < nchunks; ++c) a0 += part[c * len + j];
  out[j] = (a0 + a1) + (a2 + a3);
}

// Columns across threads (x), rows strided across blockIdx.y: no division
// per element.
template <typename TIn>
__global__ void center_apply(const TIn* __restrict__ O, const double* __restrict__ w,
                             const double* __restrict__ mean, int64_t n, int64_t p, int cpx,
                             TIn* __restrict__ D) {
  chain_priority();
  const int64_t len = p * (cpx ? 2 : 1);
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= len) return;
  const double m = mean[j];
  for (int64_t r = blockIdx.y; r < n; r += gridDim.y) {
    const int64_t t = r * len + j;
    D[t] = (TIn)(sqrt(w[r]) * ((double)O[t] - m));
  }
}

// Doubles of partial column sums the centring needs: one row of len values
// per chunk of CHUNK_ROWS samples.
inline int64_t center_partials(int64_t n, int64_t p, int cpx) {
  return (n + CHUNK_ROWS - 1) / CHUNK_ROWS * p * (cpx ? 2 : 1);
}

// Centring with the partial sums in `part` (center_partials(n, p, cpx)
// doubles, every one written by col_wsum_partial before col_reduce reads it,
// so nothing needs clearing).
// TIn = float: the mixed form's FP32 observations, mean in FP64, centred data
// rounded to FP32.
template <typename TIn>
int center_with(const TIn* O, const double* w, int64_t n, int64_t p, int cpx, double* mean, TIn* D,
                double* part, hipStream_t st) {
  int nchunks = (int)((n + CHUNK_ROWS - 1) / CHUNK_ROWS);
  int64_t len = p * (cpx ? 2 : 1);
  if (p == 1)
    hipLaunchKernelGGL(vec_wsum_partial<TIn>, dim3(1, nchunks), dim3(CHUNK_ROWS), 0, st, O, w, n, cpx, part);
  else
    hipLaunchKernelGGL(col_wsum_partial<TIn>, dim3((unsigned)((p + COLS_PER_WG - 1) / COLS_PER_WG), nchunks),
                       dim3(COLS_PER_WG, ROW_GROUPS), 0, st, O, w, n, p, cpx, part);
  JV_LAUNCH_CHECK("col_wsum_partial");
  hipLaunchKernelGGL(col_reduce, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, part, nchunks, len, mean);
  JV_LAUNCH_CHECK("col_reduce");
  hipLaunchKernelGGL(center_apply<TIn>, dim3((unsigned)((len + 255) / 256), (unsigned)std::max<int64_t>(1, std::min<int64_t>(n, 1024))),
                     dim3(256), 0, st, O, w, mean, n,
                     p, cpx, D);
  JV_LAUNCH_CHECK("center_apply");
  return JV_OK;
}

int center(const double* O, const double* w, int64_t n, int64_t p, int cpx, double* mean, double* D,
           hipStream_t st, int ws_slot = 2) {
  double* part = (double*)workspace(sizeof(double) * (size_t)center_partials(n, p, cpx), ws_slot);
  if (!part) return fail(JV_E_NOMEM, "jv_center: workspace allocation failed");
  return center_with(O, w, n, p, cpx, mean, D, part, st);
}

// Partial sums of jv_center_noalloc: a fixed array per device, reached through
// its device address (hipGetSymbolAddress), resolved once per device.  Its
// name in host code is only a host-side stand-in and is never passed to a
// kernel or a copy.
constexpr int64_t CENTER_NOALLOC_MAX_PARTIALS = 2 * 1024 * 1024;  // doubles (16 MiB per device)
static __device__ __align__(256) double scratch[CENTER_NOALLOC_MAX_PARTIALS];

constexpr int MAX_SCRATCH_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_SCRATCH_DEVICES] = {};

int center_scratch(double** out) {
  int d = 0;
  JV_HIP(hipGetDevice(&d));
  if (d < 0 || d >= MAX_SCRATCH_DEVICES) return fail(JV_E_ARG, "jv_center_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[d]) {
    void* p = nullptr;
    JV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(JV_E_HIP, "jv_center_noalloc: hipGetSymbolAddress returned NULL");
    g_scratch_ptr[d] = (double*)p;
  }
  *out = g_scratch_ptr[d];
  return JV_OK;
}

// ------------------------------------------------------------------ GEMM (FP64 MFMA)
// C[M][N] = alpha * opA * opB (+ beta C), opA(m, k) = A[m*sam + k*sak],
// opB(k, n) = B[k*sbk + n*sbn] (strides in elements; complex elements are two
// doubles), optional conjugation of either operand.  sym: C is Hermitian;
// only tiles with tm <= tn are computed and mirrored.
struct GemmArgs {
  int64_t M, N, K;
  const double* A;
  int64_t sam, sak;
  int conjA;
  const double* B;
  int64_t sbk, sbn;
  int conjB;
  double* C;
  int64_t ldc;
  int cpx;
  int sym;
  double alpha;
  int64_t kchunk;  // split-K: workgroup row blockIdx.y covers k in [y*kchunk, (y+1)*kchunk)
  double* part;    // split-K partial tiles [gridDim.y][M][N] (null: write C directly)
};

constexpr int BM = 64, BN = 64, BK = 16;

// MFMA 16x16x4 f64 operand / result lane layout (measured on gfx950): A[row =
// lane%16][k = lane/16], B[k = lane/16][col = lane%16], result element j of lane
// l is C[4j + l/16][l%16].
__device__ inline int mf_arow(int lane) { return lane % 16; }
__device__ inline int mf_ak(int lane) { return lane / 16; }
__device__ inline int mf_bcol(int lane) { return lane % 16; }
__device__ inline int mf_bk(int lane) { return lane / 16; }
__device__ inline int mf_crow(int lane, int j) { return 4 * j + lane / 16; }
__device__ inline int mf_ccol(int lane, int j) { return lane % 16; }

template <bool CPX>
__global__ void __launch_bounds__(256) gemm_f64(GemmArgs g) {
  __shared__ double As[2][BK][BM + 1];
  __shared__ double Bs[2][BK][BN + 1];
  int tm, tn;
  const int64_t ntn = (g.N + BN - 1) / BN;
  if (g.sym) {
    // linear block id -> (tm, tn) with tm <= tn
    int64_t b = blockIdx.x, row = 0, rowlen = ntn;
    while (b >= rowlen) { b -= rowlen; ++row; --rowlen; }
    tm = (int)row;
    tn = (int)(row + b);
  } else {
    tm = blockIdx.x / ntn;
    tn = blockIdx.x % ntn;
  }
  const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
  const int wm = (wv / 2) * 32, wn = (wv % 2) * 32;  // wave sub-tile origin
  const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
  v4d accr[2][2], acci[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      accr[a][b] = v4d{0, 0, 0, 0};
      acci[a][b] = v4d{0, 0, 0, 0};
    }
  const double sa_im = g.conjA ? -1.0 : 1.0, sb_im = g.conjB ? -1.0 : 1.0;
  const bool a_k_fast = g.sak == 1, b_k_fast = g.sbk == 1;
  const int64_t kbeg = g.part ? (int64_t)blockIdx.y * g.kchunk : 0;
  const int64_t kend = g.part ? std::min<int64_t>(g.K, kbeg + g.kchunk) : g.K;
  for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
    // global -> LDS (zero padded), each thread moves 4 elements of each tile
    for (int r = 0; r < (BM * BK) / 256; ++r) {
      int e = tid + 256 * r;
      int mm = a_k_fast ? e / BK : e % BM;
      int kk = a_k_fast ? e % BK : e / BM;
      int64_t gm = m0 + mm, gk = k0 + kk;
      double vr = 0.0, vi = 0.0;
      if (gm < g.M && gk < kend) {
        int64_t off = gm * g.sam + gk * g.sak;
        if (CPX) { vr = g.A[2 * off]; vi = sa_im * g.A[2 * off + 1]; }
        else vr = g.A[off];
      }
      As[0][kk][mm] = vr;
      if (CPX) As[1][kk][mm] = vi;
      int nn = b_k_fast ? e / BK : e % BN;
      int kb = b_k_fast ? e % BK : e / BN;
      int64_t gn = n0 + nn, gkb = k0 + kb;
      vr = 0.0; vi = 0.0;
      if (gn < g.N && gkb < kend) {
        int64_t off = gkb * g.sbk + gn * g.sbn;
        if (CPX) { vr = g.B[2 * off]; vi = sb_im * g.B[2 * off + 1]; }
        else vr = g.B[off];
      }
      Bs[0][kb][nn] = vr;
      if (CPX) Bs[1][kb][nn] = vi;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      double ar[2], ai[2], br[2], bi[2];
      for (int a = 0; a < 2; ++a) {
        ar[a] = As[0][ks + mf_ak(lane)][wm + 16 * a + mf_arow(lane)];
        if (CPX) ai[a] = As[1][ks + mf_ak(lane)][wm + 16 * a + mf_arow(lane)];
      }
      for (int b = 0; b < 2; ++b) {
        br[b] = Bs[0][ks + mf_bk(lane)][wn + 16 * b + mf_bcol(lane)];
        if (CPX) bi[b] = Bs[1][ks + mf_bk(lane)][wn + 16 * b + mf_bcol(lane)];
      }
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
          accr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], accr[a][b], 0, 0, 0);
          if (CPX) {
            accr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[a], bi[b], accr[a][b], 0, 0, 0);
            acci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bi[b], acci[a][b], 0, 0, 0);
            acci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], br[b], acci[a][b], 0, 0, 0);
          }
        }
    }
    __syncthreads();
  }
  // epilogue
  if (g.part) {  // split-K partial tile, reduced and mirrored by gemm_reduce
    double* P = g.part + (size_t)blockIdx.y * g.M * g.N * (CPX ? 2 : 1);
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int j = 0; j < 4; ++j) {
          int64_t i = m0 + wm + 16 * a + mf_crow(lane, j);
          int64_t c = n0 + wn + 16 * b + mf_ccol(lane, j);
          if (i >= g.M || c >= g.N) continue;
          if (CPX) {
            P[2 * (i * g.N + c)] = accr[a][b][j];
            P[2 * (i * g.N + c) + 1] = acci[a][b][j];
          } else {
            P[i * g.N + c] = accr[a][b][j];
          }
        }
    return;
  }
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int j = 0; j < 4; ++j) {
        int64_t i = m0 + wm + 16 * a + mf_crow(lane, j);
        int64_t c = n0 + wn + 16 * b + mf_ccol(lane, j);
        if (i >= g.M || c >= g.N) continue;
        double vr = g.alpha * accr[a][b][j];
        double vi = CPX ? g.alpha * acci[a][b][j] : 0.0;
        if (g.sym && tm == tn && i > c) continue;
        if (CPX) {
          g.C[2 * (i * g.ldc + c)] = vr;
          g.C[2 * (i * g.ldc + c) + 1] = vi;
          if (g.sym && i != c) {
            g.C[2 * (c * g.ldc + i)] = vr;
            g.C[2 * (c * g.ldc + i) + 1] = -vi;
          }
        } else {
          g.C[i * g.ldc + c] = vr;
          if (g.sym && i != c) g.C[c * g.ldc + i] = vr;
        }
      }
}

// Sums the split-K partials in split order into C (alpha scaled), mirroring
// Hermitian outputs exactly like the direct epilogue.
template <bool CPX>
__global__ void gemm_reduce(GemmArgs g, int splits) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.M * g.N) return;
  const int64_t i = t / g.N, c = t % g.N;
  if (g.sym && (i / BM > c / BN || (i / BM == c / BN && i > c))) return;
  const size_t stride = (size_t)g.M * g.N * (CPX ? 2 : 1);
  double vr = 0.0, vi = 0.0;
  for (int s = 0; s < splits; ++s) {
    if (CPX) {
      vr += g.part[s * stride + 2 * t];
      vi += g.part[s * stride + 2 * t + 1];
    } else {
      vr += g.part[s * stride + t];
    }
  }
  vr *= g.alpha;
  vi *= g.alpha;
  if (CPX) {
    g.C[2 * (i * g.ldc + c)] = vr;
    g.C[2 * (i * g.ldc + c) + 1] = vi;
    if (g.sym && i != c) {
      g.C[2 * (c * g.ldc + i)] = vr;
      g.C[2 * (c * g.ldc + i) + 1] = -vi;
    }
  } else {
    g.C[i * g.ldc + c] = vr;
    if (g.sym && i != c) g.C[c * g.ldc + i] = vr;
  }
}

// Large plain products go to rocBLAS (Tensile reaches 85-90 % of the FP64
// MFMA peak on them, e.g. the 16384 x 4224 x 4224 SNR projection); the
// hand-written kernel keeps the Hermitian Grams (triangle tiles + split K beat
// a full rocBLAS gemm, which computes both triangles) and small products.  Row-major C = opA opB
// is computed as the column-major C^T = opB^T opA^T.  Returns 1 when the
// strides or conjugations have no BLAS form.
int gemm_blas(const GemmArgs& g, hipStream_t st) {
  const bool cpx = g.cpx != 0;
  rocblas_operation tx, ty;
  int64_t ldx, ldy;
  if (g.sbn == 1 && !g.conjB) {
    tx = rocblas_operation_none;
    ldx = g.sbk;
  } else if (g.sbk == 1) {
    tx = g.conjB ? rocblas_operation_conjugate_transpose : rocblas_operation_transpose;
    ldx = g.sbn;
  } else {
    return 1;
  }
  if (g.sak == 1 && !g.conjA) {
    ty = rocblas_operation_none;
    ldy = g.sam;
  } else if (g.sam == 1) {
    ty = g.conjA ? rocblas_operation_conjugate_transpose : rocblas_operation_transpose;
    ldy = g.sak;
  } else {
    return 1;
  }
  rocblas_handle h = blas_handle(st);
  if (!h) return 1;
  rocblas_status s;
  if (cpx) {
    const rocblas_double_complex al = {g.alpha, 0.0}, be = {0.0, 0.0};
    s = rocblas_zgemm(h, tx, ty, (rocblas_int)g.N, (rocblas_int)g.M, (rocblas_int)g.K, &al,
                      (const rocblas_double_complex*)g.B, (rocblas_int)ldx, (const rocblas_double_complex*)g.A,
                      (rocblas_int)ldy, &be, (rocblas_double_complex*)g.C, (rocblas_int)g.ldc);
  } else {
    const double al = g.alpha, be = 0.0;
    s = rocblas_dgemm(h, tx, ty, (rocblas_int)g.N, (rocblas_int)g.M, (rocblas_int)g.K, &al, g.B, (rocblas_int)ldx,
                      g.A, (rocblas_int)ldy, &be, g.C, (rocblas_int)g.ldc);
  }
  return s == rocblas_status_success ? 0 : -1;
}

int gemm_blas_min_flops_log2() {
  static int v = [] {
    const char* s = getenv("VMC_JAX_AMD_BLAS_MIN_FLOPS_LOG2");
    return s ? atoi(s) : 31;  // 2^31 flops
  }();
  return v;
}

// Split-K partial tiles live in workspace slot part_slot: callers on different
// streams that may run at the same time pass different slots.
int gemm(const GemmArgs& g0, hipStream_t st, bool allow_blas = true, int part_slot = 18) {
  GemmArgs g = g0;
  g.part = nullptr;
  g.kchunk = 0;
  if (g.M <= 0 || g.N <= 0) return JV_OK;
  if (allow_blas) {
    const double flops = 2.0 * (double)g.M * (double)g.N * (double)g.K * (g.cpx ? 4.0 : 1.0);
    const bool ours = g.sym != 0;  // Hermitian Grams: triangle tiles halve the work
    if (!ours && flops >= std::ldexp(1.0, gemm_blas_min_flops_log2())) {
      ScopedTimer tm(g.sym ? "gram" : "gemm", st);
      const int r = gemm_blas(g, st);
      if (r == 0) return JV_OK;
      if (r < 0) return fail(JV_E_SOLVER, "rocblas gemm failed");
    }
  }
  const int64_t ntm = (g.M + BM - 1) / BM, ntn = (g.N + BN - 1) / BN;
  int64_t blocks = g.sym ? ntn * (ntn + 1) / 2 : ntm * ntn;
  if (g.sym && (g.M != g.N)) return fail(JV_E_ARG, "symmetric GEMM needs a square output");
  // split K when the tiles alone cannot fill the chip (~4 workgroups per CU)
  int splits = 1;
  while (blocks * splits < 1024 && g.K / (splits * 2) >= 4 * BK) splits *= 2;
  ScopedTimer tm(g.sym ? "gram" : "gemm", st);
  if (splits > 1) {
    const size_t bytes = sizeof(double) * (size_t)splits * g.M * g.N * (g.cpx ? 2 : 1);
    g.part = (double*)workspace(bytes, part_slot);
    if (!g.part) return fail(JV_E_NOMEM, "gemm: split-K workspace allocation failed");
    g.kchunk = ((g.K + splits - 1) / splits + BK - 1) / BK * BK;
    splits = (int)((g.K + g.kchunk - 1) / g.kchunk);
  }
  const dim3 grid((unsigned)blocks, (unsigned)splits);
  if (g.cpx) hipLaunchKernelGGL(gemm_f64<true>, grid, dim3(256), 0, st, g);
  else hipLaunchKernelGGL(gemm_f64<false>, grid, dim3(256), 0, st, g);
  JV_LAUNCH_CHECK("gemm_f64");
  if (splits > 1) {
    const unsigned rb = (unsigned)((g.M * g.N + 255) / 256);
    if (g.cpx) hipLaunchKernelGGL(gemm_reduce<true>, dim3(rb), dim3(256), 0, st, g, splits);
    else hipLaunchKernelGGL(gemm_reduce<false>, dim3(rb), dim3(256), 0, st, g, splits);
    JV_LAUNCH_CHECK("gemm_reduce");
  }
  return JV_OK;
}

// ------------------------------------------------------------------ GEMV
// y[i] = alpha * sum_j M[i][j] x[j] (op 'N'), one wave per row.
template <bool CPX>
__global__ void gemv_n(const double* __restrict__ M, const double* __restrict__ x, int64_t rows, int64_t cols,
                       int conjM, double alpha, double* __restrict__ y) {
  const int lane = threadIdx.x % 64;
  const int64_t r = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  if (r >= rows) return;
  double ar = 0.0, ai = 0.0;
  for (int64_t j = lane; j < cols; j += 64) {
    if (CPX) {
      double mr = M[2 * (r * cols + j)], mi = M[2 * (r * cols + j) + 1];
      if (conjM) mi = -mi;
      double xr = x[2 * j], xi = x[2 * j + 1];
      ar += mr * xr - mi * xi;
      ai += mr * xi + mi * xr;
    } else {
      ar = fma(M[r * cols + j], x[j], ar);
    }
  }
  ar = wave_sum(ar);
  if (CPX) ai = wave_sum(ai);
  if (lane == 0) {
    if (CPX) { y[2 * r] = alpha * ar; y[2 * r + 1] = alpha * ai; }
    else y[r] = alpha * ar;
  }
}

// y[j] = alpha * sum_i op(M[i][j]) x[i] (op 'T' or 'C'), row-chunk partials.
template <bool CPX, typename TM = double>
__global__ void gemv_t_partial(const TM* __restrict__ M, const double* __restrict__ x, int64_t rows,
                               int64_t cols, int conjM, double* __restrict__ part) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  int64_t r0 = (int64_t)blockIdx.y * GEMV_T_ROWS, r1 = std::min<int64_t>(rows, r0 + GEMV_T_ROWS);
  double ar = 0.0, ai = 0.0;
  if (CPX) {
    for (int64_t r = r0; r < r1; ++r) {
      double mr = M[2 * (r * cols + j)], mi = M[2 * (r * cols + j) + 1];
      if (conjM) mi = -mi;
      double xr = x[2 * r], xi = x[2 * r + 1];
      ar += mr * xr - mi * xi;
      ai += mr * xi + mi * xr;
    }
  } else {
    // four rows per iteration into independent sums, so their loads are in
    // flight together
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int64_t r = r0;
    for (; r + 3 < r1; r += 4) {
      a0 = fma((double)M[r * cols + j], x[r], a0);
      a1 = fma((double)M[(r + 1) * cols + j], x[r + 1], a1);
      a2 = fma((double)M[(r + 2) * cols + j], x[r + 2], a2);
      a3 = fma((double)M[(r + 3) * cols + j], x[r + 3], a3);
    }
    for (; r < r1; ++r) a0 = fma((double)M[r * cols + j], x[r], a0);
    ar = (a0 + a1) + (a2 + a3);
  }
  if (CPX) {
    part[2 * (blockIdx.y * cols + j)] = ar;
    part[2 * (blockIdx.y * cols + j) + 1] = ai;
  } else {
    part[blockIdx.y * cols + j] = ar;
  }
}

__global__ void scale_vec(double* __restrict__ v, int64_t len, double alpha) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) v[i] *= alpha;
}

// y = M^T x for a real FP32 matrix M [rows][cols] (the mixed form's centred
// data), accumulated in FP64.
int gemv_t_f32(const float* M, const double* x, int64_t rows, int64_t cols, double* y, hipStream_t st, int ws_slot) {
  const int nchunks = (int)((rows + GEMV_T_ROWS - 1) / GEMV_T_ROWS);
  double* part = (double*)workspace(sizeof(double) * (size_t)nchunks * cols, ws_slot);
  if (!part) return fail(JV_E_NOMEM, "gemv: workspace allocation failed");
  hipLaunchKernelGGL((gemv_t_partial<false, float>), dim3((unsigned)((cols + 255) / 256), nchunks), dim3(256), 0, st, M,
                     x, rows, cols, 0, part);
  JV_LAUNCH_CHECK("gemv_t_partial");
  hipLaunchKernelGGL(col_reduce, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, st, part, nchunks, cols, y);
  JV_LAUNCH_CHECK("col_reduce");
  return JV_OK;
}

int gemv(bool trans, const double* M, const double* x, int64_t rows, int64_t cols, int cpx, int conjM,
         double alpha, double* y, hipStream_t st, int ws_slot) {
  if (!trans) {
    int64_t blocks = (rows + 3) / 4;
    if (cpx) hipLaunchKernelGGL(gemv_n<true>, dim3((unsigned)blocks), dim3(256), 0, st, M, x, rows, cols, conjM, alpha, y);
    else hipLaunchKernelGGL(gemv_n<false>, dim3((unsigned)blocks), dim3(256), 0, st, M, x, rows, cols, conjM, alpha, y);
    JV_LAUNCH_CHECK("gemv_n");
    return JV_OK;
  }
  int nchunks = (int)((rows + GEMV_T_ROWS - 1) / GEMV_T_ROWS);
  int64_t len = cols * (cpx ? 2 : 1);
  double* part = (double*)workspace(sizeof(double) * (size_t)nchunks * len, ws_slot);
  if (!part) return fail(JV_E_NOMEM, "gemv: workspace allocation failed");
  if (cpx) hipLaunchKernelGGL(gemv_t_partial<true>, dim3((unsigned)((cols + 255) / 256), nchunks), dim3(256), 0, st, M, x, rows, cols, conjM, part);
  else hipLaunchKernelGGL(gemv_t_partial<false>, dim3((unsigned)((cols + 255) / 256), nchunks), dim3(256), 0, st, M, x, rows, cols, conjM, part);
  JV_LAUNCH_CHECK("gemv_t_partial");
  hipLaunchKernelGGL(col_reduce, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, part, nchunks, len, y);
  JV_LAUNCH_CHECK("col_reduce");
  if (alpha != 1.0) {
    hipLaunchKernelGGL(scale_vec, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, y, len, alpha);
    JV_LAUNCH_CHECK("scale_vec");
  }
  return JV_OK;
}

// ------------------------------------------------------------------ eigh
// rocSOLVER works column-major; a row-major symmetric / Hermitian A is seen as
// A^T = conj(A).  On return the row-major buffer holds z_k in row k, where
// conj(A) z_k = lambda_k z_k, i.e. the eigenvectors of A are conj(row k).
int eigh_raw(int64_t n, int cpx, double* A, double* w, hipStream_t st) {
  ScopedTimer tm("eigh", st);
  if (!cpx && trd_supported(n)) return trd_eigh_full(A, (int)n, w, st);
  if (cpx && trdc_supported(n)) return trdc_eigh_full(A, (int)n, w, st);
  rocblas_handle h = blas_handle(st);
  if (!h) return fail(JV_E_SOLVER, "rocblas_create_handle failed");
  double* E = (double*)workspace(sizeof(double) * (size_t)n + 64, 3);
  rocblas_int* info = (rocblas_int*)workspace(64, 4);
  if (!E || !info) return fail(JV_E_NOMEM, "eigh: workspace allocation failed");
  rocblas_status s;
  if (cpx)
    s = rocsolver_zheevd(h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)n,
                         (rocblas_double_complex*)A, (rocblas_int)n, w, E, info);
  else
    s = rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)n, A, (rocblas_int)n, w, E,
                         info);
  if (s != rocblas_status_success) return fail(JV_E_SOLVER, "rocsolver syevd/heevd status " + std::to_string((int)s));
  rocblas_int hinfo = 0;
  JV_HIP(hipMemcpyAsync(&hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, st));
  JV_HIP(hipStreamSynchronize(st));
  if (hinfo != 0) return fail(JV_E_SOLVER, "eigendecomposition did not converge (info " + std::to_string(hinfo) + ")");
  return JV_OK;
}

// out[i][k] = conj(in[k][i])
__global__ void conj_transpose(const double* __restrict__ in, double* __restrict__ out, int64_t n, int cpx) {
  __shared__ double tr[32][33], ti[32][33];
  int64_t bx = (int64_t)blockIdx.x * 32, by = (int64_t)blockIdx.y * 32;
  int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
  for (int yy = ty; yy < 32; yy += 8) {
    int64_t r = by + yy, c = bx + tx;
    if (r < n && c < n) {
      if (cpx) { tr[yy][tx] = in[2 * (r * n + c)]; ti[yy][tx] = in[2 * (r * n + c) + 1]; }
      else tr[yy][tx] = in[r * n + c];
    }
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    int64_t r = bx + yy, c = by + tx;
    if (r < n && c < n) {
      if (cpx) { out[2 * (r * n + c)] = tr[tx][yy]; out[2 * (r * n + c) + 1] = -ti[tx][yy]; }
      else out[r * n + c] = tr[tx][yy];
    }
  }
}

// ------------------------------------------------------------------ pinv apply
// y[k] *= 1/lambda_k when |lambda_k| > rcond * max|lambda| else 0 (jnp.linalg.pinv,
// hermitian=True).  One block.
__global__ void pinv_scale(const double* __restrict__ ev, int64_t n, double rcond, int cpx, double* __restrict__ y,
                           int* __restrict__ rank) {
  __shared__ double red[16];
  double mx = 0.0;
  for (int64_t k = threadIdx.x; k < n; k += blockDim.x) mx = fmax(mx, fabs(ev[k]));
  // block max
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = mx;
  __syncthreads();
  mx = 0.0;
  for (int i = 0; i < (int)blockDim.x / 64; ++i) mx = fmax(mx, red[i]);
  const double cut = rcond * mx;
  int kept = 0;
  for (int64_t k = threadIdx.x; k < n; k += blockDim.x) {
    double l = ev[k];
    double inv = fabs(l) > cut ? 1.0 / l : 0.0;
    kept += fabs(l) > cut;
    if (cpx) { y[2 * k] *= inv; y[2 * k + 1] *= inv; }
    else y[k] *= inv;
  }
  __syncthreads();
  double kk = block_sum((double)kept, red);
  if (threadIdx.x == 0) *rank = (int)kk;
}

__global__ void add_diag(double* __restrict__ T, int64_t n, double shift, int cpx) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) T[(i * n + i) * (cpx ? 2 : 1)] += shift;
}

// Partial sums of squares over a grid-stride slice; sum_finish adds them up.
__global__ void __launch_bounds__(256) sum_sq_partial(const double* __restrict__ A, int64_t len,
                                                      double* __restrict__ part) {
  __shared__ double red[4];
  double a = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x)
    a = fma(A[i], A[i], a);
  a = block_sum(a, red);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

__global__ void __launch_bounds__(256) sum_finish(const double* __restrict__ part, int m, double* __restrict__ out) {
  __shared__ double red[4];
  double a = 0.0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) a += part[i];
  a = block_sum(a, red);
  if (threadIdx.x == 0) *out = a;
}

// Margin of the Cholesky test.  If the factorisation of a symmetric p x p
// matrix A completes in floating point, L L^T = A + dA with
// ||dA||_2 <= p gamma_{p+2} ||A||_2, gamma_k = k u / (1 - k u), u = 2^-53
// (Higham, Accuracy and Stability of Numerical Algorithms, 2nd ed., Thm 10.3
// with || |L| |L^T| ||_2 <= p ||A||_2; gamma_{p+1} there, one rounding more
// here because the entries below a diagonal block are multiplied by the
// rounded reciprocal of the pivot instead of divided by it), so
// lambda_min(A) >= -||dA||_2.  For A = S' - c I (S' = S, or S + a E after a
// deflation) and
// c = F (rcond + d + x) / (1 - d), d = p gamma_{p+2}, F >= ||S'||_F >= lambda_max,
// a completed factorisation gives lambda_min(S') >= c - d (F + c) >=
// (rcond + x) F: every eigenvalue of S' exceeds the pseudo-inverse cutoff
// rcond lambda_max plus the deflation's perturbation bound x F.
__device__ inline double chol_margin(double F, int64_t p, double rcond, double x) {
  const double u = 1.1102230246251565e-16;
  const double g = (double)(p + 2) * u / (1.0 - (double)(p + 2) * u);
  const double d = (double)p * g;
  return F * (rcond + d + x) / (1.0 - d);
}

// Sc = S' - c I and S2 = S' with S' = S + a E, where E has ones in the
// blk x blk diagonal blocks and a = add[0] (no add: a = 0) deflates the
// centring null vector found by null_probe / null_decide (add[2]: its relative
// perturbation bound x); c = chol_margin(||S||_F + ||a E||_F, ...).
__global__ void shifted_copy(const double* __restrict__ S, int64_t p, const double* __restrict__ fro2, double rcond,
                             const double* __restrict__ add, int64_t blk, double* __restrict__ Sc,
                             double* __restrict__ S2) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= p * p) return;
  const double a = add ? add[0] : 0.0;
  const double F = sqrt(*fro2) + fabs(a) * (double)blk * sqrt((double)(p / blk));
  const double c = chol_margin(F, p, rcond, add ? add[2] : 0.0);
  const int64_t i = t / p, j = t % p;
  double v = S[t];
  if (add && i / blk == j / blk) v += a;
  Sc[t] = v - (i == j ? c : 0.0);
  S2[t] = v;
}

// Centred data (row n scaled by sqrt(w_n), sum_n sqrt(w_n) D_n = 0) make
// u0 = sqrt(w) a null vector of T = D D^dagger, so the pseudo-inverse of T
// always cuts one eigenvalue and T itself never passes the Cholesky test.  For
// uniform weights u0 = 1/sqrt(n): part[i] = |(T u0)_i|^2, part[n + i] =
// Re T_ii and part[2n + i] = sum_j |T_ij|^2 (one wave per row).
__global__ void __launch_bounds__(256) null_probe(const double* __restrict__ T, int64_t n, int cpx,
                                                  double* __restrict__ part) {
  const int lane = threadIdx.x % WAVE;
  const int64_t i = (int64_t)blockIdx.x * 4 + threadIdx.x / WAVE;
  if (i >= n) return;
  double sr = 0.0, si = 0.0, sq = 0.0;
  if (cpx) {
    const double* row = T + 2 * i * n;
    for (int64_t j = lane; j < n; j += WAVE) {
      sr += row[2 * j];
      si += row[2 * j + 1];
      sq = fma(row[2 * j], row[2 * j], fma(row[2 * j + 1], row[2 * j + 1], sq));
    }
  } else {
    const double* row = T + i * n;
    for (int64_t j = lane; j < n; j += WAVE) {
      sr += row[j];
      sq = fma(row[j], row[j], sq);
    }
  }
  sr = wave_sum(sr);
  si = wave_sum(si);
  sq = wave_sum(sq);
  if (lane == 0) {
    part[i] = (sr * sr + si * si) / (double)n;
    part[n + i] = cpx ? T[2 * (i * n + i)] : T[i * n + i];
    part[2 * n + i] = sq;
  }
}

// Deflation test of the centring null vector, r = T u0 (u0 = 1/sqrt(n)):
// accepted when ||r|| <= NULL_K n eps ||T||_F (u0 is an eigenvector of T up
// to rounding: T~ = T - r u0^T - u0 r^T + (u0^T r) u0 u0^T has u0 exactly,
// ||T - T~||_2 <= 2 ||r||, and T~ shares all other eigenvalues with T to within
// 2 ||r||) and ||r|| <= rcond mu / 2 (mu = max Re T_ii <= lambda_max, so the
// eigenvalue near u0, at most 2 ||r||, lies below the cutoff rcond lambda_max
// and the reference's pseudo-inverse cuts it).  Then addv = (mu / n, 1, x) with
// x = 2 NULL_K n eps per deflated vector (one real, two for the real form of a
// complex T), else (0, 0, 0).  T + mu u0 u0^T moves the cut eigenvalue up;
// by interlacing, a successful Cholesky test of the deflated matrix with the
// margin x shows every other eigenvalue of T above the cutoff.
constexpr double NULL_K = 64.0;
__global__ void __launch_bounds__(1024) null_decide(const double* __restrict__ part, int64_t n, double rcond, int cpx,
                                                    double* __restrict__ addv) {
  __shared__ double rs[1024], rm[1024], rf[1024];
  const int tid = threadIdx.x;
  double s = 0.0, m = 0.0, f = 0.0;
  for (int64_t i = tid; i < n; i += 1024) {
    s += part[i];
    m = fmax(m, part[n + i]);
    f += part[2 * n + i];
  }
  rs[tid] = s;
  rm[tid] = m;
  rf[tid] = f;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (tid < w) {
      rs[tid] += rs[tid + w];
      rm[tid] = fmax(rm[tid], rm[tid + w]);
      rf[tid] += rf[tid + w];
    }
    __syncthreads();
  }
  if (tid == 0) {
    const double eps = 2.220446049250313e-16;
    const double r = sqrt(rs[0]), F = sqrt(rf[0]);
    const bool defl = rm[0] > 0.0 && r <= 0.5 * rcond * rm[0] && r <= NULL_K * (double)n * eps * F;
    addv[0] = defl ? rm[0] / (double)n : 0.0;
    addv[1] = defl ? 1.0 : 0.0;
    addv[2] = defl ? (cpx ? 4.0 : 2.0) * NULL_K * (double)n * eps : 0.0;
  }
}

// b <- b - u0 (u0^T b) for each of the `halves` length-n parts of b (u0 =
// 1/sqrt(n): the mean is removed) when addv[1] is set; rank[0] = n - addv[1],
// rank[1] = 1 (the Cholesky route).
__global__ void __launch_bounds__(1024) project_null(double* __restrict__ b, int64_t n, int halves,
                                                     const double* __restrict__ addv, int* __restrict__ rank) {
  __shared__ double red[1024];
  const int tid = threadIdx.x;
  const bool defl = addv[1] != 0.0;
  if (tid == 0) {
    rank[0] = (int)n - (defl ? 1 : 0);
    rank[1] = 1;
  }
  if (!defl) return;
  for (int h = 0; h < halves; ++h) {
    double* v = b + (int64_t)h * n;
    double sum = 0.0;
    for (int64_t i = tid; i < n; i += 1024) sum += v[i];
    red[tid] = sum;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
      if (tid < w) red[tid] += red[tid + w];
      __syncthreads();
    }
    const double mean = red[0] / (double)n;
    __syncthreads();
    for (int64_t i = tid; i < n; i += 1024) v[i] -= mean;
    __syncthreads();
  }
}

__global__ void set_ints(int* __restrict__ v, int a, int b) {
  if (threadIdx.x == 0) {
    v[0] = a;
    v[1] = b;
  }
}

__global__ void set_identity(double* __restrict__ Q, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n * n) Q[t] = (t / n == t % n) ? 1.0 : 0.0;
}

// Q[off + i][off + j] -= U[i][j] for i, j < m
__global__ void sub_trailing(double* __restrict__ Q, int64_t n, int64_t off, int64_t m, const double* __restrict__ U) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m * m) return;
  const int64_t i = t / m, j = t - i * m;
  Q[(off + i) * n + off + j] -= U[t];
}

__global__ void real_part(const double* __restrict__ in, int64_t n, double* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[2 * i];
}

}  // namespace jv

using namespace jv;

extern "C" {

int jv_center(const double* obs, const double* w, int64_t n, int64_t p, int32_t is_complex, double* mean,
              double* data, void* stream) {
  if (!obs || !w || !mean || !data || n < 1 || p < 1) return fail(JV_E_ARG, "jv_center: bad arguments");
  return center(obs, w, n, p, is_complex, mean, data, (hipStream_t)stream);
}

int jv_center_f32(const float* obs, const double* w, int64_t n, int64_t p, int32_t is_complex, double* mean,
                  float* data, void* stream) {
  if (!obs || !w || !mean || !data || n < 1 || p < 1) return fail(JV_E_ARG, "jv_center_f32: bad arguments");
  double* part = (double*)workspace(sizeof(double) * (size_t)center_partials(n, p, is_complex), 2);
  if (!part) return fail(JV_E_NOMEM, "jv_center_f32: workspace allocation failed");
  return center_with<float>(obs, w, n, p, is_complex, mean, data, part, (hipStream_t)stream);
}

int jv_center_noalloc(const double* obs, const double* w, int64_t n, int64_t p, int32_t is_complex, double* mean,
                      double* data, void* stream) {
  if (!obs || !w || !mean || !data || n < 1 || p < 1) return fail(JV_E_ARG, "jv_center_noalloc: bad arguments");
  const int64_t need = center_partials(n, p, is_complex);
  if (need > CENTER_NOALLOC_MAX_PARTIALS)
    return fail(JV_E_ARG, "jv_center_noalloc: ceil(n / 256) * p * (is_complex ? 2 : 1) = " + std::to_string(need) +
                              " partial sums exceed the fixed scratch of " +
                              std::to_string(CENTER_NOALLOC_MAX_PARTIALS) + " doubles; use jv_center");
  double* part = nullptr;
  int rc = center_scratch(&part);
  if (rc) return rc;
  return center_with(obs, w, n, p, is_complex, mean, data, part, (hipStream_t)stream);
}

int jv_gram_nn(const double* data, int64_t n, int64_t p, int32_t is_complex, double* T, void* stream) {
  if (!data || !T || n < 1 || p < 1) return fail(JV_E_ARG, "jv_gram_nn: bad arguments");
  // T = D D^dagger: opA = D (m = row, k = col), opB(k, n) = conj(D[n][k])
  GemmArgs g{n, n, p, data, p, 1, 0, data, 1, p, 1, T, n, is_complex, 1, 1.0};
  return gemm(g, (hipStream_t)stream);
}

int jv_gram_pp(const double* data, int64_t n, int64_t p, int32_t is_complex, double* S, void* stream) {
  if (!data || !S || n < 1 || p < 1) return fail(JV_E_ARG, "jv_gram_pp: bad arguments");
  // S = D^dagger D: the real Gram of D's interleaved storage (complex data:
  // combined 2 x 2 blocks), 128 x 128 upper tiles on the FP64 matrix cores
  return gram_pp_f64(data, n, p, is_complex, S, (hipStream_t)stream, 18);
}

int jv_eigh(int64_t n, int32_t is_complex, double* A, double* w, void* stream) {
  if (!A || !w || n < 1) return fail(JV_E_ARG, "jv_eigh: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  int rc = eigh_raw(n, is_complex, A, w, st);
  if (rc) return rc;
  size_t bytes = sizeof(double) * (size_t)n * n * (is_complex ? 2 : 1);
  double* tmp = (double*)workspace(bytes, 5);
  if (!tmp) return fail(JV_E_NOMEM, "jv_eigh: workspace allocation failed");
  JV_HIP(hipMemcpyAsync(tmp, A, bytes, hipMemcpyDeviceToDevice, st));
  dim3 grid((unsigned)((n + 31) / 32), (unsigned)((n + 31) / 32));
  hipLaunchKernelGGL(conj_transpose, grid, dim3(256), 0, st, tmp, A, n, is_complex);
  JV_LAUNCH_CHECK("conj_transpose");
  return JV_OK;
}

}  // extern "C"
namespace jv {
static bool minsr_cholesky_enabled();
static int chol_route(const double* S, int64_t p, double rcond, int* flags, bool* ok, double** L, hipStream_t st,
                      const double* add = nullptr, int64_t blk = 1);
int chol_inverse_lower(const double* L, int p, double* W, hipStream_t st);
static int minsr_dual_finish_info(int rc, jv_solve_info* info, int64_t p, double rcond, bool deferred,
                                  hipStream_t st);

// R = [[Re T, -Im T], [Im T, Re T]] (2n x 2n real symmetric) of the complex
// Hermitian T (n x n, interleaved): same eigenvalues as T, each twice.
__global__ void realify_hermitian(const double* __restrict__ T, int64_t n, double* __restrict__ R) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t m = 2 * n;
  if (t >= m * m) return;
  const int64_t i = t / m, j = t % m;
  const int64_t a = i % n, c = j % n;
  const double re = T[2 * (a * n + c)], im = T[2 * (a * n + c) + 1];
  R[t] = (i < n) == (j < n) ? re : (i < n ? -im : im);
}

// b = [Re e; Im e] (cpx) or Re e
__global__ void split_complex(const double* __restrict__ e, int64_t n, int cpx, double* __restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  b[t] = e[2 * t];
  if (cpx) b[n + t] = e[2 * t + 1];
}

// x (interleaved complex) = xr + i xi from [xr; xi]
__global__ void join_complex(const double* __restrict__ y, int64_t n, double* __restrict__ x) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  x[2 * t] = y[t];
  x[2 * t + 1] = y[n + t];
}
}  // namespace jv

extern "C" {

int jv_minsr_solve(const double* data, const double* e, int64_t n, int64_t p, int32_t is_complex, double rcond,
                   double diag_shift, int32_t try_cholesky, double* update, jv_solve_info* info, void* stream) {
  if (!data || !e || !update || n < 1 || p < 1) return fail(JV_E_ARG, "jv_minsr_solve: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const int cpx = is_complex;
  // T = D D^dagger (+ shift I for real-parameter nets, minsr.py:72)
  double* T = (double*)workspace(sizeof(double) * (size_t)n * n * (cpx ? 2 : 1), 6);
  double* ev = (double*)workspace(sizeof(double) * (size_t)n, 7);
  double* y = (double*)workspace(sizeof(double) * (size_t)n * 2, 8);
  double* x = (double*)workspace(sizeof(double) * (size_t)n * 2, 9);
  double* er = (double*)workspace(sizeof(double) * (size_t)n, 10);
  int* rank = (int*)workspace(64, 11);
  if (!T || !ev || !y || !x || !er || !rank) return fail(JV_E_NOMEM, "jv_minsr_solve: workspace allocation failed");
  int rc = jv_gram_nn(data, n, p, cpx, T, stream);
  if (rc) return rc;
  if (!cpx && diag_shift != 0.0) {
    hipLaunchKernelGGL(add_diag, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, T, n, diag_shift, 0);
    JV_LAUNCH_CHECK("add_diag");
  }
  hipLaunchKernelGGL(set_ints, dim3(1), dim3(64), 0, st, rank, 0, 0);  // rank[1]: route taken
  JV_LAUNCH_CHECK("set_ints");
  if (try_cholesky && minsr_cholesky_enabled()) {
    // Cholesky route (as in jv_minsr_solve_dual): T, or its real form R for
    // complex T, has no eigenvalue at or below the cutoff; then
    // x = W^T W b with W = L^-1 (two GEMVs, no n^3 product)
    const int64_t m = cpx ? 2 * n : n;
    // the centring null vector 1/sqrt(n) of uniformly weighted data is deflated
    // (its eigenvalue is cut by the pseudo-inverse, and b is projected off it)
    double* probe = (double*)workspace(sizeof(double) * 3 * (size_t)n, 101);
    double* addv = (double*)workspace(sizeof(double) * 4, 102);
    if (!probe || !addv) return fail(JV_E_NOMEM, "jv_minsr_solve: workspace allocation failed");
    hipLaunchKernelGGL(null_probe, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, T, n, cpx, probe);
    JV_LAUNCH_CHECK("null_probe");
    hipLaunchKernelGGL(null_decide, dim3(1), dim3(1024), 0, st, probe, n, rcond, cpx, addv);
    JV_LAUNCH_CHECK("null_decide");
    double* R = T;
    if (cpx) {
      R = (double*)workspace(sizeof(double) * (size_t)m * m, 99);
      if (!R) return fail(JV_E_NOMEM, "jv_minsr_solve: workspace allocation failed");
      hipLaunchKernelGGL(realify_hermitian, dim3((unsigned)((m * m + 255) / 256)), dim3(256), 0, st, T, n, R);
      JV_LAUNCH_CHECK("realify_hermitian");
    }
    bool ok = false;
    double* L = nullptr;
    if ((rc = chol_route(R, m, rcond, rank, &ok, &L, st, addv, n))) return rc;
    if (ok) {
      double* W = (double*)workspace(sizeof(double) * (size_t)m * m, 95);  // R - c I is no longer needed
      double* b = (double*)workspace(sizeof(double) * 2 * (size_t)m, 100);
      if (!W || !b) return fail(JV_E_NOMEM, "jv_minsr_solve: workspace allocation failed");
      {
        ScopedTimer tm("chol_inverse", st);
        if ((rc = chol_inverse_lower(L, (int)m, W, st))) return rc;
      }
      hipLaunchKernelGGL(split_complex, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, e, n, cpx, b);
      JV_LAUNCH_CHECK("split_complex");
      hipLaunchKernelGGL(project_null, dim3(1), dim3(1024), 0, st, b, n, cpx ? 2 : 1, addv, rank);
      JV_LAUNCH_CHECK("project_null");
      if ((rc = gemv(false, W, b, m, m, 0, 0, 1.0, b + m, st, 12))) return rc;  // W b
      if ((rc = gemv(true, W, b + m, m, m, 0, 0, 1.0, b, st, 12))) return rc;   // W^T W b
      const double* xs = b;
      if (cpx) {
        hipLaunchKernelGGL(join_complex, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, b, n, x);
        JV_LAUNCH_CHECK("join_complex");
        xs = x;
      }
      // update = -D^dagger x (project_null has set the rank)
      if ((rc = gemv(true, data, xs, n, p, cpx, 1, -1.0, update, st, 12))) return rc;
      return minsr_dual_finish_info(JV_OK, info, n, rcond, false, st);
    }
  }
  rc = eigh_raw(n, cpx, T, ev, st);
  if (rc) return rc;
  // T's rows hold z_k (eigenvectors u_k = conj(z_k)); y = R e, scale, x = R^H y
  const double* rhs = e;
  if (!cpx) {  // real-parameter nets: only Re(e) couples to the real data (minsr.py:75-78)
    hipLaunchKernelGGL(real_part, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, e, n, er);
    JV_LAUNCH_CHECK("real_part");
    rhs = er;
  }
  rc = gemv(false, T, rhs, n, n, cpx, 0, 1.0, y, st, 12);
  if (rc) return rc;
  hipLaunchKernelGGL(pinv_scale, dim3(1), dim3(1024), 0, st, ev, n, rcond, cpx, y, rank);
  JV_LAUNCH_CHECK("pinv_scale");
  rc = gemv(true, T, y, n, n, cpx, 1, 1.0, x, st, 12);
  if (rc) return rc;
  // update = -D^dagger x (holomorphic: the caller applies the layout factor)
  rc = gemv(true, data, x, n, p, cpx, 1, -1.0, update, st, 12);
  if (rc) return rc;
  return minsr_dual_finish_info(JV_OK, info, n, rcond, false, st);
}

}  // extern "C"

// ====================================================================== TDVP
namespace jv {

// The same work split as col_wsum_partial.
__global__ void __launch_bounds__(COLS_PER_WG * ROW_GROUPS) col_norm2_partial(const double* __restrict__ D, int64_t n,
                                                                              int64_t p, int cpx,
                                                                              double* __restrict__ part) {
  __shared__ double sa[ROW_GROUPS][COLS_PER_WG];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int64_t j = (int64_t)blockIdx.x * COLS_PER_WG + tx;
  const int64_t r0 = (int64_t)blockIdx.y * CHUNK_ROWS, r1 = std::min<int64_t>(n, r0 + CHUNK_ROWS);
  double a = 0.0;
  if (j < p) {
    for (int64_t r = r0 + ty; r < r1; r += ROW_GROUPS) {
      if (cpx) {
        double x = D[2 * (r * p + j)], y = D[2 * (r * p + j) + 1];
        a = fma(x, x, fma(y, y, a));
      } else {
        double x = D[r * p + j];
        a = fma(x, x, a);
      }
    }
  }
  sa[ty][tx] = a;
  __syncthreads();
  if (ty == 0 && j < p) part[blockIdx.y * p + j] = (sa[0][tx] + sa[1][tx]) + (sa[2][tx] + sa[3][tx]);
}

// col_norm2_partial of a single column: one row per thread and a workgroup sum.
__global__ void vec_norm2_partial(const double* __restrict__ D, int64_t n, int cpx, double* __restrict__ part) {
  __shared__ double red[CHUNK_ROWS / WAVE];
  const int64_t r = (int64_t)blockIdx.y * CHUNK_ROWS + threadIdx.x;
  double a = 0.0;
  if (r < n) {
    if (cpx) {
      const double x = D[2 * r], y = D[2 * r + 1];
      a = fma(x, x, y * y);
    } else {
      a = D[r] * D[r];
    }
  }
  a = block_sum(a, red);
  if (threadIdx.x == 0) part[blockIdx.y] = a;
}

int col_norm2(const double* D, int64_t n, int64_t p, int cpx, double* out, hipStream_t st, int slot) {
  int nchunks = (int)((n + CHUNK_ROWS - 1) / CHUNK_ROWS);
  double* part = (double*)workspace(sizeof(double) * (size_t)nchunks * p, slot);
  if (!part) return fail(JV_E_NOMEM, "col_norm2: workspace allocation failed");
  if (p == 1)
    hipLaunchKernelGGL(vec_norm2_partial, dim3(1, nchunks), dim3(CHUNK_ROWS), 0, st, D, n, cpx, part);
  else
    hipLaunchKernelGGL(col_norm2_partial, dim3((unsigned)((p + COLS_PER_WG - 1) / COLS_PER_WG), nchunks),
                       dim3(COLS_PER_WG, ROW_GROUPS), 0, st, D, n, p, cpx, part);
  JV_LAUNCH_CHECK("col_norm2_partial");
  hipLaunchKernelGGL(col_reduce, dim3((unsigned)((p + 255) / 256)), dim3(256), 0, st, part, nchunks, p, out);
  JV_LAUNCH_CHECK("col_reduce");
  return JV_OK;
}

// F0 = -x * raw (raw = D^dagger e); F = makeReal(F0): 'real' -> Re, 'imag' -> i Im
__global__ void tdvp_force(const double* __restrict__ raw, int64_t pr, double xr, double xi, int imag,
                           double* __restrict__ F0, double* __restrict__ F) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pr) return;
  double a = raw[2 * i], b = raw[2 * i + 1];
  double fr = -(xr * a - xi * b), fi = -(xr * b + xi * a);
  F0[2 * i] = fr;
  F0[2 * i + 1] = fi;
  F[2 * i] = imag ? 0.0 : fr;
  F[2 * i + 1] = imag ? fi : 0.0;
}

// S = makeReal(S0) (+ diag(shift * diag(S)), tdvp.py:145-146).  'real': real
// matrix; 'imag': complex matrix i Im(S0).
__global__ void tdvp_project_S(const double* __restrict__ S0, int64_t pr, int imag, double shift,
                               double* __restrict__ S) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= pr * pr) return;
  int64_t i = t / pr, j = t % pr;
  if (imag) {
    S[2 * t] = 0.0;
    S[2 * t + 1] = S0[2 * t + 1];
  } else {
    double v = S0[2 * t];
    if (i == j && shift > 1e-10) v += shift * v;
    S[t] = v;
  }
}

// Per-sample SNR data Q[n][k] = makeReal(-x conj(D[n][k]) e_n / w_n)
// (stats.py:255-265 covar_data followed by transform, tdvp.py:175-178).
__global__ void tdvp_snr_data(const double* __restrict__ D, const double* __restrict__ e,
                              const double* __restrict__ w, int64_t n, int64_t pr, double xr, double xi, int imag,
                              double* __restrict__ Q) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * pr) return;
  int64_t r = t / pr;
  double dr = D[2 * t], di = -D[2 * t + 1];  // conj(D)
  double er = e[2 * r], ei = e[2 * r + 1];
  double pr_ = dr * er - di * ei, pi_ = dr * ei + di * er;
  double inv = 1.0 / w[r];
  pr_ *= inv;
  pi_ *= inv;
  double qr = -(xr * pr_ - xi * pi_), qi = -(xr * pi_ + xi * pr_);
  if (imag) { Q[2 * t] = 0.0; Q[2 * t + 1] = qi; }
  else Q[t] = qr;
}

__global__ void tdvp_snr(const double* __restrict__ VtF, const double* __restrict__ var, int64_t pr, int cpx,
                         double nsamp, double* __restrict__ snr) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= pr) return;
  double a2 = cpx ? VtF[2 * k] * VtF[2 * k] + VtF[2 * k + 1] * VtF[2 * k + 1] : VtF[k] * VtF[k];
  snr[k] = sqrt(fabs(nsamp * a2 / var[k]));
}

// One block per cutoff candidate c: residual_c = ||(pinvEv ev - 1) VtF|| / ||F||
// with pinvEv = invEv * 1/(1+(max(c, pinvCutoff)/|ev/ev_max|)^6) [* 1/(1+(snrTol/snr)^6)]
// (tdvp.py:193-209).
__device__ inline double tdvp_coef(double ev, double evmax, double cut, double pinv_cutoff, int use_snr,
                                   double snr_tol, double snr) {
  double rel = fabs(ev / evmax);
  double inv = rel > 1e-14 ? 1.0 / ev : 0.0;
  double q = fmax(cut, pinv_cutoff) / rel;
  double q2 = q * q;
  double r = 1.0 / (1.0 + q2 * q2 * q2);
  if (use_snr) {
    double s = snr_tol / snr;
    double s2 = s * s;
    r *= 1.0 / (1.0 + s2 * s2 * s2);
  }
  return inv * r;
}

__global__ void tdvp_residuals(const double* __restrict__ ev, const double* __restrict__ VtF,
                               const double* __restrict__ snr, const double* __restrict__ cuts, int64_t pr, int cpx,
                               double pinv_cutoff, int use_snr, double snr_tol, const double* __restrict__ fnorm,
                               double* __restrict__ res) {
  __shared__ double red[16];
  const double cut = cuts[blockIdx.x];
  const double evmax = ev[pr - 1];
  double acc = 0.0;
  for (int64_t k = threadIdx.x; k < pr; k += blockDim.x) {
    double c = tdvp_coef(ev[k], evmax, cut, pinv_cutoff, use_snr, snr_tol, snr[k]);
    double f = c * ev[k] - 1.0;
    double a2 = cpx ? VtF[2 * k] * VtF[2 * k] + VtF[2 * k + 1] * VtF[2 * k + 1] : VtF[k] * VtF[k];
    acc = fma(f * f, a2, acc);
  }
  double s = block_sum(acc, red);
  if (threadIdx.x == 0) res[blockIdx.x] = sqrt(s) / *fnorm;
}

__global__ void vec_norm(const double* __restrict__ v, int64_t len, double* __restrict__ out) {
  __shared__ double red[16];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < len; i += blockDim.x) acc = fma(v[i], v[i], acc);
  double s = block_sum(acc, red);
  if (threadIdx.x == 0) *out = sqrt(s);
}

// c_k = pinvEv_k * VtF_k for the selected candidate
__global__ void tdvp_coeffs(const double* __restrict__ ev, const double* __restrict__ VtF,
                            const double* __restrict__ snr, int64_t pr, int cpx, double cut, double pinv_cutoff,
                            int use_snr, double snr_tol, double* __restrict__ c) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= pr) return;
  double f = tdvp_coef(ev[k], ev[pr - 1], cut, pinv_cutoff, use_snr, snr_tol, snr[k]);
  if (cpx) { c[2 * k] = f * VtF[2 * k]; c[2 * k + 1] = f * VtF[2 * k + 1]; }
  else c[k] = f * VtF[k];
}

__global__ void to_complex(const double* __restrict__ in, int64_t n, double* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) { out[2 * i] = in[i]; out[2 * i + 1] = 0.0; }
}

// tdvp_error = |1 + (Re(u^T S0 u) - 2 Re(u^T F0)) / Var E|  (tdvp.py:102-104)
__global__ void tdvp_error_k(const double* __restrict__ u, const double* __restrict__ S0u,
                             const double* __restrict__ F0, int64_t pr, const double* __restrict__ evar,
                             double* __restrict__ out) {
  __shared__ double red[16];
  double a = 0.0;
  for (int64_t i = threadIdx.x; i < pr; i += blockDim.x) a += u[i] * (S0u[2 * i] - 2.0 * F0[2 * i]);
  double s = block_sum(a, red);
  if (threadIdx.x == 0) *out = fabs(1.0 + s / *evar);
}

// ---- holomorphic structure of the jVMC layout.  For a holomorphic network the
// flat gradient row is [g, i g] per complex leaf (vqs.py:66-69): column
// col_ig[c] = i * column col_g[c] of the compact gradient G.  Then
//   S0 = O^H O  = [[A, iA], [-iA, A]] (split order), A = G^H G  (Pc x Pc),
//   'real': S = Re S0 = [[Ar, -Ai], [Ai, Ar]], the real form of A: every
//           eigenvalue of A twice, eigenvectors [a; b] and [-b; a] for u = a + i b;
//   'imag': S = i Im S0 = i [[Ai, Ar], [-Ar, Ai]]: eigenvalues +-lambda_k(A),
//           eigenvectors [u; -i u] / sqrt2 (+lambda) and [conj u; i conj u] / sqrt2
//           (-lambda).
// So one Pc x Pc Hermitian eigendecomposition replaces the 2Pc x 2Pc one.

// G[r][c] = D[r][col_g[c]]
__global__ void holo_gather(const double* __restrict__ D, int64_t n, int64_t pr, const int64_t* __restrict__ col_g,
                            int64_t pc, double* __restrict__ G) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * pc) return;
  const int64_t r = t / pc, c = t - r * pc;
  const int64_t src = r * pr + col_g[c];
  G[2 * t] = D[2 * src];
  G[2 * t + 1] = D[2 * src + 1];
}

// S0 (flat order, pr x pr complex) from A (pc x pc complex)
__global__ void holo_expand_S0(const double* __restrict__ A, int64_t pc, const int64_t* __restrict__ col_g,
                               const int64_t* __restrict__ col_ig, int64_t pr, double* __restrict__ S0) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= pc * pc) return;
  const int64_t c = t / pc, d = t - c * pc;
  const double ar = A[2 * t], ai = A[2 * t + 1];
  const int64_t gc = col_g[c], ic = col_ig[c], gd = col_g[d], id = col_ig[d];
  double* p;
  p = S0 + 2 * (gc * pr + gd); p[0] = ar;  p[1] = ai;   // A
  p = S0 + 2 * (gc * pr + id); p[0] = -ai; p[1] = ar;   // i A
  p = S0 + 2 * (ic * pr + gd); p[0] = ai;  p[1] = -ar;  // -i A
  p = S0 + 2 * (ic * pr + id); p[0] = ar;  p[1] = ai;   // A
}

// A <- A + shift * diag(A) on the diagonal ('real' mode, tdvp.py:145-146: the
// shift of S = Re S0 is the shift of A on both halves)
__global__ void holo_shift(double* __restrict__ A, int64_t pc, double shift) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= pc || !(shift > 1e-10)) return;
  A[2 * (c * pc + c)] += shift * A[2 * (c * pc + c)];
}

// Eigenpairs of S from those of A.  R (pc x pc complex) holds conj(u_k) in row
// k (the eigh_raw convention), lam ascending.  Z rows receive conj(v_j) in flat
// order with ev ascending.
__global__ void holo_eig(const double* __restrict__ R, const double* __restrict__ lam, int64_t pc,
                         const int64_t* __restrict__ col_g, const int64_t* __restrict__ col_ig, int64_t pr, int imag,
                         double* __restrict__ Z, double* __restrict__ ev) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= pc * pc) return;
  const int64_t k = t / pc, c = t - k * pc;
  const double rr = R[2 * t], ri = R[2 * t + 1];  // conj(u_k)[c]
  const int64_t gc = col_g[c], ic = col_ig[c];
  if (imag) {
    const double h = 0.70710678118654752440;
    // row pc-1-k: eigenvalue -lam_k, conj(v) = [u; -i u] / sqrt2, u = conj(R_k)
    double* z = Z + 2 * ((pc - 1 - k) * pr);
    z[2 * gc] = h * rr;   z[2 * gc + 1] = -h * ri;
    z[2 * ic] = -h * ri;  z[2 * ic + 1] = -h * rr;
    // row pc+k: eigenvalue +lam_k, conj(v) = [conj u; i conj u] / sqrt2 = [R_k; i R_k] / sqrt2
    z = Z + 2 * ((pc + k) * pr);
    z[2 * gc] = h * rr;   z[2 * gc + 1] = h * ri;
    z[2 * ic] = -h * ri;  z[2 * ic + 1] = h * rr;
    if (c == 0) {
      ev[pc - 1 - k] = -lam[k];
      ev[pc + k] = lam[k];
    }
  } else {
    // rows 2k, 2k+1: [a; b] and [-b; a] with u_k = a + i b = conj(R_k)
    const double a = rr, b = -ri;
    double* z = Z + (2 * k) * pr;
    z[gc] = a;
    z[ic] = b;
    z = Z + (2 * k + 1) * pr;
    z[gc] = -b;
    z[ic] = a;
    if (c == 0) {
      ev[2 * k] = lam[k];
      ev[2 * k + 1] = lam[k];
    }
  }
}

// SNR projection through A's eigenvectors.  With Qc[r][c] = q_g + i q_ig (the
// real SNR data of the pair; 'imag' stores q as the imaginary part) and
// M = Qc R_A^T (R_A rows = conj(u_k)), the projection Q Z^T onto S's
// eigenvectors is, per pair k:
//   'imag': column pc+k = i M / sqrt2, column pc-1-k = i conj(M) / sqrt2;
//   'real': column 2k = Re M, column 2k+1 = Im M.
__global__ void holo_qc(const double* __restrict__ Q, int64_t n, int64_t pr, const int64_t* __restrict__ col_g,
                        const int64_t* __restrict__ col_ig, int64_t pc, int imag, double* __restrict__ Qc) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * pc) return;
  const int64_t r = t / pc, c = t - r * pc;
  const int64_t a = r * pr + col_g[c], b = r * pr + col_ig[c];
  Qc[2 * t] = imag ? Q[2 * a + 1] : Q[a];
  Qc[2 * t + 1] = imag ? Q[2 * b + 1] : Q[b];
}

__global__ void holo_snr_scatter(const double* __restrict__ M, int64_t n, int64_t pc, int imag,
                                 double* __restrict__ R) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * pc) return;
  const int64_t r = t / pc, k = t - r * pc;
  const double mr = M[2 * t], mi = M[2 * t + 1];
  const int64_t pr = 2 * pc;
  if (imag) {
    const double h = 0.70710678118654752440;
    double* p = R + 2 * (r * pr + pc + k);  // i M / sqrt2
    p[0] = -h * mi;
    p[1] = h * mr;
    p = R + 2 * (r * pr + pc - 1 - k);      // i conj(M) / sqrt2
    p[0] = h * mi;
    p[1] = h * mr;
  } else {
    R[r * pr + 2 * k] = mr;
    R[r * pr + 2 * k + 1] = mi;
  }
}

__global__ void sum_vec(const double* __restrict__ v, int64_t n, double* __restrict__ out) {
  __shared__ double red[16];
  double a = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) a += v[i];
  double s = block_sum(a, red);
  if (threadIdx.x == 0) *out = s;
}

}  // namespace jv

extern "C" {

int jv_col_norm2(const double* data, int64_t n, int64_t p, int32_t is_complex, double* out, void* stream) {
  if (!data || !out || n < 1 || p < 1) return fail(JV_E_ARG, "jv_col_norm2: bad arguments");
  return col_norm2(data, n, p, is_complex, out, (hipStream_t)stream, 13);
}

int jv_gemm(int64_t M, int64_t N, int64_t K, const double* A, int64_t sam, int64_t sak, int32_t conjA,
            const double* B, int64_t sbk, int64_t sbn, int32_t conjB, double* C, int64_t ldc, int32_t is_complex,
            double alpha, void* stream) {
  if (!A || !B || !C || M < 0 || N < 0 || K < 1) return fail(JV_E_ARG, "jv_gemm: bad arguments");
  GemmArgs g{M, N, K, A, sam, sak, conjA, B, sbk, sbn, conjB, C, ldc, is_complex, 0, alpha};
  return gemm(g, (hipStream_t)stream);
}

int jv_gemv(int32_t trans, const double* M, const double* x, int64_t rows, int64_t cols, int32_t is_complex,
            int32_t conj, double alpha, double* y, void* stream) {
  if (!M || !x || !y || rows < 1 || cols < 1) return fail(JV_E_ARG, "jv_gemv: bad arguments");
  return gemv(trans != 0, M, x, rows, cols, is_complex, conj, alpha, y, (hipStream_t)stream, 14);
}

}  // extern "C"

namespace jv {
int gram_pp_bf16x3(const void* D, int f32, int64_t n, int64_t p, int cpx, double* S, hipStream_t st, int slot);
// The TDVP solve; with col_g / col_ig (device, pc entries each) the data has
// the holomorphic [g, i g] structure and S is built and diagonalised through
// the Pc x Pc Gram A = G^H G.
static int tdvp_solve_impl(const double* D, const double* e, const double* w, int64_t n, int64_t pr,
                           const double* x, int32_t make_real, double diag_shift, int32_t use_snr,
                           double glob_num_samples, double snr_tol, double pinv_tol, double pinv_cutoff,
                           double* update, double* S, double* F, double* ev, double* snr, jv_solve_info* info,
                           const int64_t* col_g, const int64_t* col_ig, int64_t pc, void* stream,
                           bool mixed_gram = false) {
  if (!D || !e || !w || !x || !update || !S || !F || !ev || !snr || n < 1 || pr < 1)
    return fail(JV_E_ARG, "jv_tdvp_solve: bad arguments");
  const bool holo = col_g != nullptr;
  if (holo && (!col_ig || pc < 1 || 2 * pc != pr))
    return fail(JV_E_ARG, "jv_tdvp_solve_holomorphic: the column maps must split pr = 2 pc columns");
  if (!(pinv_tol < 1.0)) return fail(JV_E_ARG, "pinvTol >= 1 leaves the regulariser undefined (tdvp.py:195-198)");
  hipStream_t st = (hipStream_t)stream;
  const int imag = make_real != 0;
  const double xr = x[0], xi = x[1];
  const size_t P2 = (size_t)pr * pr;
  double* S0 = (double*)workspace(sizeof(double) * 2 * P2, 20);
  double* Z = (double*)workspace(sizeof(double) * (imag ? 2 : 1) * P2, 21);
  double* raw = (double*)workspace(sizeof(double) * 2 * pr, 22);
  double* F0 = (double*)workspace(sizeof(double) * 2 * pr, 23);
  double* VtF = (double*)workspace(sizeof(double) * 2 * pr, 24);
  double* Q = (double*)workspace(sizeof(double) * (imag ? 2 : 1) * (size_t)n * pr, 25);
  double* R = (double*)workspace(sizeof(double) * (imag ? 2 : 1) * (size_t)n * pr, 26);
  double* mean = (double*)workspace(sizeof(double) * 2 * pr, 27);
  double* var = (double*)workspace(sizeof(double) * pr, 28);
  double* scal = (double*)workspace(sizeof(double) * 8, 29);  // fnorm, evar, tdvp_error
  double* cvec = (double*)workspace(sizeof(double) * 2 * pr, 30);
  double* ucpx = (double*)workspace(sizeof(double) * 2 * pr, 31);
  double* S0u = (double*)workspace(sizeof(double) * 2 * pr, 32);
  double* Fr = (double*)workspace(sizeof(double) * pr, 33);
  if (!S0 || !Z || !raw || !F0 || !VtF || !Q || !R || !mean || !var || !scal || !cvec || !ucpx || !S0u || !Fr)
    return fail(JV_E_NOMEM, "jv_tdvp_solve: workspace allocation failed");
  const unsigned bp = (unsigned)((pr + 255) / 256);
  int rc;
  // S0 = D^dagger D, F0 = -x D^dagger e  (tdvp.py:139-142)
  double* A = nullptr;
  if (holo) {
    double* G = (double*)workspace(sizeof(double) * 2 * (size_t)n * pc, 38);
    A = (double*)workspace(sizeof(double) * 2 * (size_t)pc * pc, 39);
    if (!G || !A) return fail(JV_E_NOMEM, "jv_tdvp_solve: workspace allocation failed");
    hipLaunchKernelGGL(holo_gather, dim3((unsigned)((n * pc + 255) / 256)), dim3(256), 0, st, D, n, pr, col_g, pc, G);
    JV_LAUNCH_CHECK("holo_gather");
    // the mixed form takes A from the compensated BF16 Gram (FP64 data in)
    if ((rc = mixed_gram ? gram_pp_bf16x3(G, 0, n, pc, 1, A, st, 19) : jv_gram_pp(G, n, pc, 1, A, stream))) return rc;
    hipLaunchKernelGGL(holo_expand_S0, dim3((unsigned)((pc * pc + 255) / 256)), dim3(256), 0, st, A, pc, col_g,
                       col_ig, pr, S0);
    JV_LAUNCH_CHECK("holo_expand_S0");
  } else {
    if ((rc = jv_gram_pp(D, n, pr, 1, S0, stream))) return rc;
  }
  if ((rc = gemv(true, D, e, n, pr, 1, 1, 1.0, raw, st, 34))) return rc;
  hipLaunchKernelGGL(tdvp_force, dim3(bp), dim3(256), 0, st, raw, pr, xr, xi, imag, F0, F);
  JV_LAUNCH_CHECK("tdvp_force");
  hipLaunchKernelGGL(tdvp_project_S, dim3((unsigned)((P2 + 255) / 256)), dim3(256), 0, st, S0, pr, imag,
                     diag_shift, S);
  JV_LAUNCH_CHECK("tdvp_project_S");
  // eigendecomposition of S (tdvp.py:153-171); Z rows hold z_k
  if (holo) {
    double* lam = (double*)workspace(sizeof(double) * (size_t)pc, 44);
    if (!lam) return fail(JV_E_NOMEM, "jv_tdvp_solve: workspace allocation failed");
    if (!imag) {
      hipLaunchKernelGGL(holo_shift, dim3((unsigned)((pc + 255) / 256)), dim3(256), 0, st, A, pc, diag_shift);
      JV_LAUNCH_CHECK("holo_shift");
    }
    if ((rc = eigh_raw(pc, 1, A, lam, st))) return rc;
    hipLaunchKernelGGL(holo_eig, dim3((unsigned)((pc * pc + 255) / 256)), dim3(256), 0, st, A, lam, pc, col_g,
                       col_ig, pr, imag, Z, ev);
    JV_LAUNCH_CHECK("holo_eig");
  } else {
    JV_HIP(hipMemcpyAsync(Z, S, sizeof(double) * (imag ? 2 : 1) * P2, hipMemcpyDeviceToDevice, st));
    if ((rc = eigh_raw(pr, imag, Z, ev, st))) return rc;
  }
  // VtF_k = v_k^dagger F = sum_i z_k[i] F_i
  if (imag) {
    if ((rc = gemv(false, Z, F, pr, pr, 1, 0, 1.0, VtF, st, 34))) return rc;
  } else {
    hipLaunchKernelGGL(real_part, dim3(bp), dim3(256), 0, st, F, pr, Fr);
    JV_LAUNCH_CHECK("real_part");
    if ((rc = gemv(false, Z, Fr, pr, pr, 0, 0, 1.0, VtF, st, 34))) return rc;
  }
  // SNR (tdvp.py:173-181): rho = Q conj(V) = Q Z^T, weighted variance per column
  hipLaunchKernelGGL(tdvp_snr_data, dim3((unsigned)((n * pr + 255) / 256)), dim3(256), 0, st, D, e, w, n, pr, xr,
                     xi, imag, Q);
  JV_LAUNCH_CHECK("tdvp_snr_data");
  if (holo) {
    double* Qc = (double*)workspace(sizeof(double) * 2 * (size_t)n * pc, 38);
    double* M = (double*)workspace(sizeof(double) * 2 * (size_t)n * pc, 45);
    if (!Qc || !M) return fail(JV_E_NOMEM, "jv_tdvp_solve: workspace allocation failed");
    const unsigned nb = (unsigned)((n * pc + 255) / 256);
    hipLaunchKernelGGL(holo_qc, dim3(nb), dim3(256), 0, st, Q, n, pr, col_g, col_ig, pc, imag, Qc);
    JV_LAUNCH_CHECK("holo_qc");
    GemmArgs g{n, pc, pc, Qc, pc, 1, 0, A, 1, pc, 0, M, pc, 1, 0, 1.0};
    if ((rc = gemm(g, st))) return rc;
    hipLaunchKernelGGL(holo_snr_scatter, dim3(nb), dim3(256), 0, st, M, n, pc, imag, R);
    JV_LAUNCH_CHECK("holo_snr_scatter");
  } else {
    GemmArgs g{n, pr, pr, Q, pr, 1, 0, Z, 1, pr, 0, R, pr, imag, 0, 1.0};
    if ((rc = gemm(g, st))) return rc;
  }
  if ((rc = center(R, w, n, pr, imag, mean, R, st))) return rc;
  if ((rc = col_norm2(R, n, pr, imag, var, st, 35))) return rc;
  hipLaunchKernelGGL(tdvp_snr, dim3(bp), dim3(256), 0, st, VtF, var, pr, imag, glob_num_samples, snr);
  JV_LAUNCH_CHECK("tdvp_snr");
  // ||F||, Var E = sum |e|^2
  if (imag) {
    hipLaunchKernelGGL(vec_norm, dim3(1), dim3(1024), 0, st, F, 2 * pr, scal);
  } else {
    hipLaunchKernelGGL(vec_norm, dim3(1), dim3(1024), 0, st, Fr, pr, scal);
  }
  JV_LAUNCH_CHECK("vec_norm");
  if ((rc = col_norm2(e, n, 1, 1, scal + 1, st, 36))) return rc;
  // regulariser candidates (host computes the same cutoff sequence as tdvp.py:196-199)
  std::vector<double> cuts;
  double cutoff = 1e-2;
  do {
    cutoff *= 0.8;
    cuts.push_back(cutoff);
  } while (cutoff > pinv_cutoff);
  const int K = (int)cuts.size();
  double* dcuts = (double*)workspace(sizeof(double) * 2 * K, 37);
  if (!dcuts) return fail(JV_E_NOMEM, "jv_tdvp_solve: workspace allocation failed");
  double* dres = dcuts + K;
  JV_HIP(hipMemcpyAsync(dcuts, cuts.data(), sizeof(double) * K, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(tdvp_residuals, dim3(K), dim3(256), 0, st, ev, VtF, snr, dcuts, pr, imag, pinv_cutoff, use_snr,
                     snr_tol, scal, dres);
  JV_LAUNCH_CHECK("tdvp_residuals");
  std::vector<double> res(K);
  JV_HIP(hipMemcpyAsync(res.data(), dres, sizeof(double) * K, hipMemcpyDeviceToHost, st));
  JV_HIP(hipStreamSynchronize(st));
  int ks = K - 1;
  for (int k = 0; k < K; ++k)
    if (!(res[k] > pinv_tol)) { ks = k; break; }
  hipLaunchKernelGGL(tdvp_coeffs, dim3(bp), dim3(256), 0, st, ev, VtF, snr, pr, imag, cuts[ks], pinv_cutoff, use_snr,
                     snr_tol, cvec);
  JV_LAUNCH_CHECK("tdvp_coeffs");
  // update = Re(V c): 'real' -> Z^T c ; 'imag' -> Re(Z^H c)
  if (imag) {
    if ((rc = gemv(true, Z, cvec, pr, pr, 1, 1, 1.0, ucpx, st, 34))) return rc;
    hipLaunchKernelGGL(real_part, dim3(bp), dim3(256), 0, st, ucpx, pr, update);
    JV_LAUNCH_CHECK("real_part");
  } else {
    if ((rc = gemv(true, Z, cvec, pr, pr, 0, 0, 1.0, update, st, 34))) return rc;
  }
  // tdvp_error from the unprojected S0, F0
  hipLaunchKernelGGL(to_complex, dim3(bp), dim3(256), 0, st, update, pr, ucpx);
  JV_LAUNCH_CHECK("to_complex");
  if ((rc = gemv(false, S0, ucpx, pr, pr, 1, 0, 1.0, S0u, st, 34))) return rc;
  hipLaunchKernelGGL(tdvp_error_k, dim3(1), dim3(1024), 0, st, update, S0u, F0, pr, scal + 1, scal + 2);
  JV_LAUNCH_CHECK("tdvp_error_k");
  double hs[3];
  JV_HIP(hipMemcpyAsync(hs, scal, sizeof(double) * 3, hipMemcpyDeviceToHost, st));
  JV_HIP(hipStreamSynchronize(st));
  if (info) {
    info->residual = res[ks];
    info->pinv_cutoff = std::max(cuts[ks], pinv_cutoff);
    info->iterations = ks + 1;
    info->rank = 0;
    info->e_var = hs[1];
    info->tdvp_error = hs[2];
  }
  return JV_OK;
}
}  // namespace jv

extern "C" {

int jv_tdvp_solve(const double* D, const double* e, const double* w, int64_t n, int64_t pr, const double* x,
                  int32_t make_real, double diag_shift, int32_t use_snr, double glob_num_samples, double snr_tol,
                  double pinv_tol, double pinv_cutoff, double* update, double* S, double* F, double* ev,
                  double* snr, jv_solve_info* info, void* stream) {
  return tdvp_solve_impl(D, e, w, n, pr, x, make_real, diag_shift, use_snr, glob_num_samples, snr_tol, pinv_tol,
                         pinv_cutoff, update, S, F, ev, snr, info, nullptr, nullptr, 0, stream);
}

int jv_tdvp_solve_holomorphic(const double* D, const double* e, const double* w, int64_t n, int64_t pr,
                              const int64_t* col_g, const int64_t* col_ig, int64_t pc, const double* x,
                              int32_t make_real, double diag_shift, int32_t use_snr, double glob_num_samples,
                              double snr_tol, double pinv_tol, double pinv_cutoff, double* update, double* S,
                              double* F, double* ev, double* snr, jv_solve_info* info, void* stream) {
  if (!col_g || !col_ig) return fail(JV_E_ARG, "jv_tdvp_solve_holomorphic: null column maps");
  return tdvp_solve_impl(D, e, w, n, pr, x, make_real, diag_shift, use_snr, glob_num_samples, snr_tol, pinv_tol,
                         pinv_cutoff, update, S, F, ev, snr, info, col_g, col_ig, pc, stream);
}

int jv_tdvp_solve_holomorphic_mixed(const double* D, const double* e, const double* w, int64_t n, int64_t pr,
                                    const int64_t* col_g, const int64_t* col_ig, int64_t pc, const double* x,
                                    int32_t make_real, double diag_shift, int32_t use_snr, double glob_num_samples,
                                    double snr_tol, double pinv_tol, double pinv_cutoff, double* update, double* S,
                                    double* F, double* ev, double* snr, jv_solve_info* info, void* stream) {
  if (!col_g || !col_ig) return fail(JV_E_ARG, "jv_tdvp_solve_holomorphic_mixed: null column maps");
  return tdvp_solve_impl(D, e, w, n, pr, x, make_real, diag_shift, use_snr, glob_num_samples, snr_tol, pinv_tol,
                         pinv_cutoff, update, S, F, ev, snr, info, col_g, col_ig, pc, stream, true);
}


}  // extern "C"

// ------------------------------------------------------------------ subsets and cross-validation
namespace jv {
// nrm = sum_k w[start + k step]; wout[k] = w[start + k step] / nrm; scal[0] = nrm
// (one workgroup, fixed summation order).
__global__ void subset_weights(const double* __restrict__ w, int64_t start, int64_t step, int64_t count,
                               double* __restrict__ wout, double* __restrict__ scal) {
  __shared__ double red[16];
  __shared__ double tot;
  double a = 0.0;
  for (int64_t k = threadIdx.x; k < count; k += blockDim.x) a += w[start + k * step];
  double s = block_sum(a, red);
  if (threadIdx.x == 0) { tot = s; scal[0] = s; }
  __syncthreads();
  for (int64_t k = threadIdx.x; k < count; k += blockDim.x) wout[k] = w[start + k * step] / tot;
}

// part[chunk][j] = sum over the chunk's rows of sqrt(w'_k) data[r_k][j] / sqrt(nrm)
__global__ void subset_mean_partial(const double* __restrict__ D, const double* __restrict__ wn,
                                    const double* __restrict__ scal, int64_t start, int64_t step, int64_t count,
                                    int64_t len, double* __restrict__ part) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= len) return;
  const double rs = 1.0 / sqrt(scal[0]);
  int64_t k0 = (int64_t)blockIdx.y * CHUNK_ROWS, k1 = std::min<int64_t>(count, k0 + CHUNK_ROWS);
  double a = 0.0;
  for (int64_t k = k0; k < k1; ++k) a = fma(sqrt(wn[k]), D[(start + k * step) * len + j] * rs, a);
  part[blockIdx.y * len + j] = a;
}

__global__ void subset_apply(const double* __restrict__ D, const double* __restrict__ wn,
                             const double* __restrict__ scal, const double* __restrict__ mean0,
                             double* __restrict__ mean1, int64_t start, int64_t step, int64_t count, int64_t len,
                             double* __restrict__ out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count * len) return;
  int64_t k = t / len, j = t - k * len;
  const double d = D[(start + k * step) * len + j] / sqrt(scal[0]);
  out[t] = d + sqrt(wn[k]) * (mean0[j] - (mean1[j] + mean0[j]));
}

__global__ void add_vec(double* __restrict__ y, const double* __restrict__ x, int64_t len) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) y[i] += x[i];
}

// out[0] = ||S2 u - F2||^2, out[1] = ||F2||^2, out[2] = Re(u^T S0 u) - 2 Re(u^T F0) with
// S0 u given (complex), S2 u = makeReal(S0 u) + shift' diag(S2) u, F2 = makeReal(F0).
__global__ void cross_validate_k(const double* __restrict__ u, const double* __restrict__ S0u,
                                 const double* __restrict__ F0, const double* __restrict__ dg, int64_t pr, int imag,
                                 double shift, double* __restrict__ out) {
  __shared__ double red[16];
  double rr = 0.0, ff = 0.0, q = 0.0;
  for (int64_t i = threadIdx.x; i < pr; i += blockDim.x) {
    const double sr = S0u[2 * i], si = S0u[2 * i + 1], fr = F0[2 * i], fi = F0[2 * i + 1];
    q += u[i] * (sr - 2.0 * fr);
    if (imag) {  // i Im(S0 u) - i Im(F0): diag(S) = i Im(diag S0) = 0
      const double d = si - fi;
      rr = fma(d, d, rr);
      ff = fma(fi, fi, ff);
    } else {
      const double d = sr + (shift > 1e-10 ? shift * dg[i] * u[i] : 0.0) - fr;
      rr = fma(d, d, rr);
      ff = fma(fr, fr, ff);
    }
  }
  rr = block_sum(rr, red);
  ff = block_sum(ff, red);
  q = block_sum(q, red);
  if (threadIdx.x == 0) { out[0] = rr; out[1] = ff; out[2] = q; }
}
}  // namespace jv

extern "C" {

int jv_subset(const double* data, const double* w, const double* mean, int64_t n, int64_t p, int32_t is_complex,
              int64_t start, int64_t step, int64_t count, double* out_data, double* out_w, double* out_mean,
              void* stream) {
  using namespace jv;
  if (!data || !w || !mean || !out_data || !out_w || !out_mean || p < 1 || step < 1 || start < 0 || count < 1 ||
      start + (count - 1) * step >= n)
    return fail(JV_E_ARG, "jv_subset: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const int64_t len = p * (is_complex ? 2 : 1);
  const int nchunks = (int)((count + CHUNK_ROWS - 1) / CHUNK_ROWS);
  double* scal = (double*)workspace(sizeof(double) * 8, 50);
  double* part = (double*)workspace(sizeof(double) * (size_t)nchunks * len, 51);
  if (!scal || !part) return fail(JV_E_NOMEM, "jv_subset: workspace allocation failed");
  hipLaunchKernelGGL(subset_weights, dim3(1), dim3(1024), 0, st, w, start, step, count, out_w, scal);
  JV_LAUNCH_CHECK("subset_weights");
  hipLaunchKernelGGL(subset_mean_partial, dim3((unsigned)((len + 255) / 256), nchunks), dim3(256), 0, st, data,
                     out_w, scal, start, step, count, len, part);
  JV_LAUNCH_CHECK("subset_mean_partial");
  // out_mean = sum of the partials (the shift from the old mean); the old mean is added after the data pass
  hipLaunchKernelGGL(col_reduce, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, part, nchunks, len, out_mean);
  JV_LAUNCH_CHECK("col_reduce");
  hipLaunchKernelGGL(subset_apply, dim3((unsigned)((count * len + 255) / 256)), dim3(256), 0, st, data, out_w, scal,
                     mean, out_mean, start, step, count, len, out_data);
  JV_LAUNCH_CHECK("subset_apply");
  hipLaunchKernelGGL(add_vec, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, out_mean, mean, len);
  JV_LAUNCH_CHECK("add_vec");
  return JV_OK;
}

int jv_tdvp_cross_validate(const double* D, const double* e, int64_t n, int64_t pr, const double* x,
                           int32_t make_real, double diag_shift, const double* update, double* out, void* stream) {
  using namespace jv;
  if (!D || !e || !x || !update || !out || n < 1 || pr < 1)
    return fail(JV_E_ARG, "jv_tdvp_cross_validate: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const int imag = make_real != 0;
  double* ucpx = (double*)workspace(sizeof(double) * 2 * pr, 52);
  double* y = (double*)workspace(sizeof(double) * 2 * n, 53);
  double* S0u = (double*)workspace(sizeof(double) * 2 * pr, 54);
  double* raw = (double*)workspace(sizeof(double) * 2 * pr, 55);
  double* F0 = (double*)workspace(sizeof(double) * 2 * pr, 56);
  double* F = (double*)workspace(sizeof(double) * 2 * pr, 57);
  double* dg = (double*)workspace(sizeof(double) * pr, 58);
  double* scal = (double*)workspace(sizeof(double) * 4, 59);
  if (!ucpx || !y || !S0u || !raw || !F0 || !F || !dg || !scal)
    return fail(JV_E_NOMEM, "jv_tdvp_cross_validate: workspace allocation failed");
  const unsigned bp = (unsigned)((pr + 255) / 256);
  int rc;
  hipLaunchKernelGGL(to_complex, dim3(bp), dim3(256), 0, st, update, pr, ucpx);
  JV_LAUNCH_CHECK("to_complex");
  // S0 u = D^H (D u) with S0 = D^H D (stats.py:235-245), F0 = -x D^H e (tdvp.py:139-140)
  if ((rc = gemv(false, D, ucpx, n, pr, 1, 0, 1.0, y, st, 60))) return rc;
  if ((rc = gemv(true, D, y, n, pr, 1, 1, 1.0, S0u, st, 60))) return rc;
  if ((rc = gemv(true, D, e, n, pr, 1, 1, 1.0, raw, st, 60))) return rc;
  hipLaunchKernelGGL(tdvp_force, dim3(bp), dim3(256), 0, st, raw, pr, x[0], x[1], imag, F0, F);
  JV_LAUNCH_CHECK("tdvp_force");
  if (!imag && diag_shift > 1e-10) {  // diag(Re S0) = column norms of D (tdvp.py:145-146)
    if ((rc = col_norm2(D, n, pr, 1, dg, st, 61))) return rc;
  }
  hipLaunchKernelGGL(cross_validate_k, dim3(1), dim3(1024), 0, st, update, S0u, F0, dg, pr, imag, diag_shift, scal);
  JV_LAUNCH_CHECK("cross_validate_k");
  double h[3];
  JV_HIP(hipMemcpyAsync(h, scal, sizeof(h), hipMemcpyDeviceToHost, st));
  JV_HIP(hipStreamSynchronize(st));
  out[0] = sqrt(h[0]) / sqrt(h[1]);
  out[1] = h[2];
  return JV_OK;
}

}  // extern "C"

// ====================================================================== MinSR, dual form
namespace jv {
__global__ void neg_copy(const double* __restrict__ in, int64_t len, double* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) out[i] = -in[i];
}
// Storage and Gram arithmetic of the dual solve's data: FP64 data with the FP64
// MFMA Gram (the reference's precision), or the mixed form's FP32 data with
// the compensated BF16 Gram.
// S_given: the Gram (with the shift) formed beforehand by dual_gram into a
// caller's buffer, consumed by the solve.
struct DualData {
  const void* ptr;
  bool f32;
  double* S_given = nullptr;
};
static int minsr_dual(DualData data, const double* e, const double* e_raw, const double* w,
                      hipEvent_t e_ready, int64_t n, int64_t p, int32_t is_complex, double rcond, double diag_shift,
                      bool try_chol, double* update, hipStream_t st);
static int dual_gram(DualData data, int64_t n, int64_t p, int cpx, double diag_shift, double* S, hipStream_t st);
int gram_pp_bf16x3(const void* D, int f32, int64_t n, int64_t p, int cpx, double* S, hipStream_t st, int slot);
static int minsr_dual_finish_info(int rc, jv_solve_info* info, int64_t p, double rcond, bool deferred,
                                  hipStream_t st);
static int minsr_dual_post_info(int rc, jv_solve_info* info, double rcond, bool deferred, hipStream_t st);
}  // namespace jv

extern "C" {

int jv_minsr_solve_dual(const double* data, const double* e, int64_t n, int64_t p, int32_t is_complex,
                        double rcond, double diag_shift, int32_t try_cholesky, double* update, jv_solve_info* info,
                        void* stream) {
  if (!data || !e || !update || n < 1 || p < 1) return fail(JV_E_ARG, "jv_minsr_solve_dual: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  return minsr_dual_finish_info(minsr_dual({data, false}, e, nullptr, nullptr, nullptr, n, p, is_complex, rcond,
                                           diag_shift, try_cholesky != 0, update, st),
                                info, p, rcond, false, st);
}

int jv_minsr_solve_dual_deferred(const double* data, const double* e_raw, const double* w, void* e_ready, int64_t n,
                                 int64_t p, int32_t is_complex, double rcond, double diag_shift, int32_t try_cholesky,
                                 double* update, jv_solve_info* info, void* stream) {
  if (!data || !e_raw || !w || !update || n < 1 || p < 1)
    return fail(JV_E_ARG, "jv_minsr_solve_dual_deferred: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  return minsr_dual_finish_info(minsr_dual({data, false}, nullptr, e_raw, w, (hipEvent_t)e_ready, n, p, is_complex,
                                           rcond, diag_shift, try_cholesky != 0, update, st),
                                info, p, rcond, true, st);
}

int jv_minsr_solve_dual_deferred_mixed(const float* data, const double* e_raw, const double* w, void* e_ready,
                                       int64_t n, int64_t p, int32_t is_complex, double rcond, double diag_shift,
                                       int32_t try_cholesky, double* update, jv_solve_info* info, void* stream) {
  if (!data || !e_raw || !w || !update || n < 1 || p < 1)
    return fail(JV_E_ARG, "jv_minsr_solve_dual_deferred_mixed: bad arguments");
  if (is_complex) return fail(JV_E_ARG, "jv_minsr_solve_dual_deferred_mixed: real data only");
  hipStream_t st = (hipStream_t)stream;
  return minsr_dual_finish_info(minsr_dual({data, true}, nullptr, e_raw, w, (hipEvent_t)e_ready, n, p, 0, rcond,
                                           diag_shift, try_cholesky != 0, update, st),
                                info, p, rcond, true, st);
}

int jv_minsr_dual_gram(const void* data, int32_t f32, int64_t n, int64_t p, int32_t is_complex, double diag_shift,
                       double* S, void* stream) {
  if (!data || !S || n < 1 || p < 1) return fail(JV_E_ARG, "jv_minsr_dual_gram: bad arguments");
  if (f32 && is_complex) return fail(JV_E_ARG, "jv_minsr_dual_gram: FP32 data is real only");
  return dual_gram({data, f32 != 0}, n, p, is_complex, diag_shift, S, (hipStream_t)stream);
}

int jv_minsr_solve_dual_deferred_gram(const void* data, int32_t f32, double* S, const double* e_raw,
                                      const double* w, void* e_ready, int64_t n, int64_t p, int32_t is_complex,
                                      double rcond, double diag_shift, int32_t try_cholesky, double* update,
                                      jv_solve_info* info, void* stream) {
  if (!data || !S || !e_raw || !w || !update || n < 1 || p < 1)
    return fail(JV_E_ARG, "jv_minsr_solve_dual_deferred_gram: bad arguments");
  if (f32 && is_complex) return fail(JV_E_ARG, "jv_minsr_solve_dual_deferred_gram: FP32 data is real only");
  hipStream_t st = (hipStream_t)stream;
  return minsr_dual_finish_info(minsr_dual({data, f32 != 0, S}, nullptr, e_raw, w, (hipEvent_t)e_ready, n, p,
                                           is_complex, rcond, diag_shift, try_cholesky != 0, update, st),
                                info, p, rcond, true, st);
}

int jv_minsr_solve_dual_deferred_gram_async(const void* data, int32_t f32, double* S, const double* e_raw,
                                            const double* w, void* e_ready, int64_t n, int64_t p,
                                            int32_t is_complex, double rcond, double diag_shift,
                                            int32_t try_cholesky, double* update, jv_solve_info* info,
                                            void* stream) {
  if (!data || !S || !e_raw || !w || !update || !info || n < 1 || p < 1)
    return fail(JV_E_ARG, "jv_minsr_solve_dual_deferred_gram_async: bad arguments");
  if (f32 && is_complex) return fail(JV_E_ARG, "jv_minsr_solve_dual_deferred_gram_async: FP32 data is real only");
  hipStream_t st = (hipStream_t)stream;
  return minsr_dual_post_info(minsr_dual({data, f32 != 0, S}, nullptr, e_raw, w, (hipEvent_t)e_ready, n, p,
                                         is_complex, rcond, diag_shift, try_cholesky != 0, update, st),
                              info, rcond, true, st);
}

}  // extern "C"

namespace jv {
// Q = H_0 ... H_{n-2} (reflectors in the rows of V, compact-WY T factors in
// Tbuf) formed explicitly: blocks applied last to first to the identity, each
// touching only the trailing (n - k0 - 1) square, as three products
// W = V_b^T Q, W2 = T_b W, U = V_b W2 and Q -= U.  Own GEMM kernels only (no
// rocBLAS handle) with split-K scratch in slot 94, so it may run beside the
// tridiagonal eigensolver (rocSOLVER's or the GEMMs of dc.hip) on another
// stream.
static int trd_form_q(const double* V, int n, const double* Tbuf, double* Q, double* W, double* W2, double* U,
                      hipStream_t st) {
  const int64_t nn = (int64_t)n * n;
  hipLaunchKernelGGL(set_identity, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, Q, (int64_t)n);
  JV_LAUNCH_CHECK("set_identity");
  const int nref = n - 1, wb = 32, nblk = (nref + wb - 1) / wb;
  for (int b = nblk - 1; b >= 0; --b) {
    const int64_t k0 = (int64_t)b * wb, nb = std::min<int64_t>(wb, nref - k0);
    const int64_t off = k0 + 1, m = n - off;
    const double* Vb = V + k0 * n + off;  // V_b^T(p, i) = Vb[p * n + i]
    const double* Tb = Tbuf + (size_t)b * wb * wb;
    int rc;
    GemmArgs g1{nb, m, m, Vb, n, 1, 0, Q + off * n + off, n, 1, 0, W, m, 0, 0, 1.0};
    if ((rc = gemm(g1, st, false, 94))) return rc;
    GemmArgs g2{nb, m, nb, Tb, wb, 1, 0, W, m, 1, 0, W2, m, 0, 0, 1.0};
    if ((rc = gemm(g2, st, false, 94))) return rc;
    GemmArgs g3{m, m, nb, Vb, 1, n, 0, W2, m, 1, 0, U, m, 0, 0, 1.0};
    if ((rc = gemm(g3, st, false, 94))) return rc;
    hipLaunchKernelGGL(sub_trailing, dim3((unsigned)((m * m + 255) / 256)), dim3(256), 0, st, Q, (int64_t)n, off, m, U);
    JV_LAUNCH_CHECK("sub_trailing");
  }
  return JV_OK;
}

// Cholesky route of the real dual solve (VMC_JAX_AMD_MINSR_CHOLESKY=0 turns it
// off).  With c = ||S||_F (rcond + 8 p eps) >= rcond lambda_max(S), a
// successful Cholesky factorisation of S - c I shows (Sylvester's inertia) that
// every eigenvalue of S exceeds the pseudo-inverse cutoff rcond lambda_max: the
// pseudo-inverse is the inverse, and S x = b by Cholesky gives the same update
// as the eigendecomposition.  S - c I and a copy of S are factored together
// (S itself stays intact for the eigendecomposition); *ok reports whether the
// first factorisation succeeded, *L is then the factor of S.  flags[2] and
// flags[3] receive the two factorisations' info; reads flags[2] on the host.
static bool minsr_cholesky_enabled() {
  const char* s = getenv("VMC_JAX_AMD_MINSR_CHOLESKY");
  return !(s && std::string(s) == "0");
}

int chol_factor2(double* A0, double* A1, int p, int* info, hipStream_t st);
int chol_inverse_lower(const double* L, int p, double* W, hipStream_t st);

static int chol_route(const double* S, int64_t p, double rcond, int* flags, bool* ok, double** L, hipStream_t st,
                      const double* add, int64_t blk) {
  *ok = false;
  const size_t pp = (size_t)p * p;
  double* Sc = (double*)workspace(sizeof(double) * pp, 95);
  double* S2 = (double*)workspace(sizeof(double) * pp, 97);
  double* part = (double*)workspace(sizeof(double) * 257, 96);
  if (!Sc || !S2 || !part) return fail(JV_E_NOMEM, "jv_minsr_solve_dual: workspace allocation failed");
  {
    ScopedTimer tm("cholesky", st);
    hipLaunchKernelGGL(sum_sq_partial, dim3(256), dim3(256), 0, st, S, (int64_t)pp, part);
    JV_LAUNCH_CHECK("sum_sq_partial");
    hipLaunchKernelGGL(sum_finish, dim3(1), dim3(256), 0, st, part, 256, part + 256);
    JV_LAUNCH_CHECK("sum_finish");
    hipLaunchKernelGGL(shifted_copy, dim3((unsigned)((pp + 255) / 256)), dim3(256), 0, st, S, p, part + 256, rcond, add,
                       blk, Sc, S2);
    JV_LAUNCH_CHECK("shifted_copy");
    // small matrices: both factorisations in the same latency-bound launches;
    // large ones: the proof first, so that a refused attempt costs one
    const bool together = p < 2048;
    int rc = chol_factor2(Sc, together ? S2 : nullptr, (int)p, flags + 2, st);
    if (rc) return rc;
  }
  int hinfo = -1;
  JV_HIP(hipMemcpyAsync(&hinfo, flags + 2, sizeof(int), hipMemcpyDeviceToHost, st));
  JV_HIP(hipStreamSynchronize(st));
  if (hinfo != 0) return JV_OK;  // an eigenvalue at or below the cutoff is possible
  if (p >= 2048) {
    ScopedTimer tm("cholesky", st);
    int rc = chol_factor2(S2, nullptr, (int)p, flags + 3, st);
    if (rc) return rc;
  }
  *L = S2;
  *ok = true;
  return JV_OK;
}

// Dual MinSR solve.  Either e (centred) is given, or e_raw with the weights w
// and an event after which e_raw is complete (NULL: ready in stream order); e_raw
// is then centred here (SampledObs, stats.py:197-205).  On the real tridiagonal
// path the Householder reduction does not need b = D^T e: the producer of e_raw
// overlaps the Gram and the reduction, and the centring, b and Q^T b run on the
// auxiliary stream during dstedc.
// S = D^dagger D (+ shift): the same nonzero spectrum as T = D D^dagger (+ shift).
static int dual_gram(DualData data, int64_t n, int64_t p, int cpx, double diag_shift, double* S, hipStream_t st) {
  int rc = data.f32 ? gram_pp_bf16x3(data.ptr, 1, n, p, cpx, S, st, 19)
                    : jv_gram_pp((const double*)data.ptr, n, p, cpx, S, (void*)st);
  if (rc) return rc;
  if (!cpx && diag_shift != 0.0) {
    hipLaunchKernelGGL(add_diag, dim3((unsigned)((p + 255) / 256)), dim3(256), 0, st, S, p, diag_shift, 0);
    JV_LAUNCH_CHECK("add_diag");
  }
  return JV_OK;
}

static int minsr_dual(DualData data, const double* e, const double* e_raw, const double* w,
                      hipEvent_t e_ready, int64_t n, int64_t p, int32_t is_complex, double rcond, double diag_shift,
                      bool try_chol, double* update, hipStream_t st) {
  const int cpx = is_complex;
  double* S = data.S_given ? data.S_given : (double*)workspace(sizeof(double) * (size_t)p * p * (cpx ? 2 : 1), 6);
  double* ev = (double*)workspace(sizeof(double) * (size_t)p, 7);
  double* y = (double*)workspace(sizeof(double) * (size_t)p * 2, 8);
  double* x = (double*)workspace(sizeof(double) * (size_t)p * 2, 9);
  double* er = (double*)workspace(sizeof(double) * (size_t)n, 10);
  int* rank = (int*)workspace(64, 11);
  double* b = (double*)workspace(sizeof(double) * (size_t)p * 2, 15);
  if (!S || !ev || !y || !x || !er || !rank || !b) return fail(JV_E_NOMEM, "jv_minsr_solve_dual: workspace allocation failed");
  int rc = data.S_given ? JV_OK : dual_gram(data, n, p, cpx, diag_shift, S, st);
  if (rc) return rc;
  const bool deferred = e == nullptr;
  // deferred: E_loc mean (2 doubles) and variance, read back with the solve info
  double* emean = deferred ? (double*)workspace(sizeof(double) * 4, 50) : nullptr;
  double* ec = deferred ? (double*)workspace(sizeof(double) * 2 * (size_t)n, 51) : nullptr;
  if (deferred && (!emean || !ec)) return fail(JV_E_NOMEM, "jv_minsr_solve_dual: workspace allocation failed");
  // centred e (when deferred) and b = D^dagger e on stream s
  auto rhs_b = [&](hipStream_t s) -> int {
    const double* ee = e;
    if (deferred) {
      if (e_ready) JV_HIP(hipStreamWaitEvent(s, e_ready, 0));
      int r = center(e_raw, w, n, 1, 1, emean, ec, s, 48);
      if (r) return r;
      if ((r = col_norm2(ec, n, 1, 1, emean + 2, s, 98))) return r;  // Var E_loc (stats.py:248-252)
      ee = ec;
    }
    const double* rhs = ee;
    if (!cpx) {
      hipLaunchKernelGGL(real_part, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ee, n, er);
      JV_LAUNCH_CHECK("real_part");
      rhs = er;
    }
    if (data.f32) return gemv_t_f32((const float*)data.ptr, rhs, n, p, b, s, deferred ? 49 : 12);  // b = D^T e
    return gemv(true, (const double*)data.ptr, rhs, n, p, cpx, 1, 1.0, b, s, deferred ? 49 : 12);  // b = D^dagger e
  };
  hipLaunchKernelGGL(set_ints, dim3(1), dim3(64), 0, st, rank, 0, 0);  // rank[1]: route taken
  JV_LAUNCH_CHECK("set_ints");
  if (!cpx && try_chol && minsr_cholesky_enabled()) {
    bool ok = false;
    double* L = nullptr;
    if ((rc = chol_route(S, p, rcond, rank, &ok, &L, st))) return rc;
    if (ok) {
      // S^-1 = W^T W with W = L^-1 (into S), formed while E_loc may still run;
      // after b only a GEMV remains
      {
        ScopedTimer tm("chol_inverse", st);
        double* Wi = (double*)workspace(sizeof(double) * (size_t)p * p, 95);  // S - c I is no longer needed
        if (!Wi) return fail(JV_E_NOMEM, "jv_minsr_solve_dual: workspace allocation failed");
        if ((rc = chol_inverse_lower(L, (int)p, Wi, st))) return rc;
        GemmArgs g{p, p, p, Wi, 1, p, 0, Wi, p, 1, 0, S, p, 0, 1, 1.0};
        if ((rc = gemm(g, st))) return rc;
      }
      // b = D^T e on this stream (deferred: after E_loc's event; the
      // factorisation and inverse are done long before E_loc in practice)
      if ((rc = rhs_b(st))) return rc;
      if ((rc = gemv(false, S, b, p, p, 0, 0, 1.0, x, st, 12))) return rc;
      hipLaunchKernelGGL(set_ints, dim3(1), dim3(64), 0, st, rank, (int)p, 1);
      JV_LAUNCH_CHECK("set_ints");
      hipLaunchKernelGGL(neg_copy, dim3((unsigned)((p + 255) / 256)), dim3(256), 0, st, x, p, update);
      JV_LAUNCH_CHECK("neg_copy");
      return JV_OK;
    }
    // otherwise the eigendecomposition below
  }
  if (deferred && !cpx && trd_supported(p)) {
    hipStream_t s2;
    hipEvent_t ev_red, ev_rhs;
    if (!aux_stream(&s2, &ev_red, &ev_rhs)) return fail(JV_E_HIP, "jv_minsr_solve_dual: auxiliary stream creation failed");
    double* Z = (double*)workspace(sizeof(double) * (size_t)p * p, 16);
    double* tw = (double*)workspace(sizeof(double) * trd_work_doubles((int)p), 17);
    if (!Z || !tw) return fail(JV_E_NOMEM, "jv_minsr_solve_dual: workspace allocation failed");
    double* Tbuf = tw + trd_wy_offset((int)p);
    const size_t pp = (size_t)p * p;
    double* Q = (double*)workspace(sizeof(double) * pp, 52);
    double* Wq = (double*)workspace(sizeof(double) * 2 * 32 * (size_t)p, 53);
    double* Uq = (double*)workspace(sizeof(double) * pp, 54);
    if (!Q || !Wq || !Uq) return fail(JV_E_NOMEM, "jv_minsr_solve_dual: workspace allocation failed");
    {
      ScopedTimer tm("eigh", st);
      if ((rc = trd_reduce(S, (int)p, ev, nullptr, tw, st))) return rc;
      JV_HIP(hipEventRecord(ev_red, st));
      // auxiliary stream, after the reduction (so D and the reflectors are
      // complete) and after e_raw: the explicit Q and x = Q^T D^T e, both
      // overlapping dstedc below
      JV_HIP(hipStreamWaitEvent(s2, ev_red, 0));
      if ((rc = trd_wy_build(S, tw + p, (int)p, Tbuf, s2))) return rc;
      if ((rc = trd_form_q(S, (int)p, Tbuf, Q, Wq, Wq + 32 * (size_t)p, Uq, s2))) return rc;
      if ((rc = rhs_b(s2))) return rc;
      if ((rc = gemv(true, Q, b, p, p, 0, 0, 1.0, x, s2, 49))) return rc;
      JV_HIP(hipEventRecord(ev_rhs, s2));
      if ((rc = trd_tridiag_eig((int)p, ev, tw, Z, st))) return rc;
    }
    JV_HIP(hipStreamWaitEvent(st, ev_rhs, 0));
    if ((rc = gemv(false, Z, x, p, p, 0, 0, 1.0, y, st, 12))) return rc;
    hipLaunchKernelGGL(pinv_scale, dim3(1), dim3(1024), 0, st, ev, p, rcond, 0, y, rank);
    JV_LAUNCH_CHECK("pinv_scale");
    if ((rc = gemv(true, Z, y, p, p, 0, 1, 1.0, x, st, 12))) return rc;
    {
      ScopedTimer tm("trd_apply", st);
      if ((rc = gemv(false, Q, x, p, p, 0, 0, 1.0, y, st, 12))) return rc;  // y = Q x
    }
    int64_t len = p;
    hipLaunchKernelGGL(neg_copy, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, y, len, update);
    JV_LAUNCH_CHECK("neg_copy");
    return JV_OK;
  }
  if ((rc = rhs_b(st))) return rc;
  if (!cpx && trd_supported(p)) {
    // S = Q Z diag(ev) Z^T Q^T: Householder tridiagonalisation (Q^T b formed on
    // the way), divide and conquer on T, then x = Q Z pinv(ev) Z^T Q^T b
    double* Z = (double*)workspace(sizeof(double) * (size_t)p * p, 16);
    double* tw = (double*)workspace(sizeof(double) * trd_work_doubles((int)p), 17);
    if (!Z || !tw) return fail(JV_E_NOMEM, "jv_minsr_solve_dual: workspace allocation failed");
    JV_HIP(hipMemcpyAsync(x, b, sizeof(double) * (size_t)p, hipMemcpyDeviceToDevice, st));
    {
      ScopedTimer tm("eigh", st);
      if ((rc = trd_eig_solve(S, (int)p, ev, x, Z, tw, st))) return rc;
    }
    if ((rc = gemv(false, Z, x, p, p, 0, 0, 1.0, y, st, 12))) return rc;
    hipLaunchKernelGGL(pinv_scale, dim3(1), dim3(1024), 0, st, ev, p, rcond, 0, y, rank);
    JV_LAUNCH_CHECK("pinv_scale");
    if ((rc = gemv(true, Z, y, p, p, 0, 1, 1.0, x, st, 12))) return rc;
    if ((rc = trd_apply(S, tw + p, (int)p, x, tw + trd_wy_offset((int)p), st))) return rc;
  } else {
    if ((rc = eigh_raw(p, cpx, S, ev, st))) return rc;
    if ((rc = gemv(false, S, b, p, p, cpx, 0, 1.0, y, st, 12))) return rc;
    hipLaunchKernelGGL(pinv_scale, dim3(1), dim3(1024), 0, st, ev, p, rcond, cpx, y, rank);
    JV_LAUNCH_CHECK("pinv_scale");
    if ((rc = gemv(true, S, y, p, p, cpx, 1, 1.0, x, st, 12))) return rc;
  }
  int64_t len = p * (cpx ? 2 : 1);
  hipLaunchKernelGGL(neg_copy, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, x, len, update);
  JV_LAUNCH_CHECK("neg_copy");
  return JV_OK;
}

// Reads the pseudo-inverse rank back into info (after a successful solve).
// With deferred, also the E_loc mean and variance formed on the auxiliary
// stream (the main stream has waited for it).
static int minsr_dual_finish_info(int rc, jv_solve_info* info, int64_t p, double rcond, bool deferred,
                                  hipStream_t st) {
  if (rc || !info) return rc;
  int* rank = (int*)workspace(64, 11);
  if (!rank) return fail(JV_E_NOMEM, "jv_minsr_solve_dual: workspace allocation failed");
  int hr[4] = {0, 0, 0, 0};
  double em[3] = {0.0, 0.0, 0.0};
  JV_HIP(hipMemcpyAsync(hr, rank, sizeof(hr), hipMemcpyDeviceToHost, st));
  if (deferred) {
    const double* emean = (const double*)workspace(sizeof(double) * 4, 50);
    if (!emean) return fail(JV_E_NOMEM, "jv_minsr_solve_dual: workspace allocation failed");
    JV_HIP(hipMemcpyAsync(em, emean, sizeof(em), hipMemcpyDeviceToHost, st));
  }
  JV_HIP(hipStreamSynchronize(st));
  info->e_mean[0] = em[0];
  info->e_mean[1] = em[1];
  info->e_var = em[2];
  if (hr[1] == 1 && hr[3] != 0)
    return fail(JV_E_SOLVER, "jv_minsr_solve_dual: Cholesky factorisation failed (info " + std::to_string(hr[3]) + ")");
  info->rank = hr[0];
  info->iterations = hr[1];
  info->residual = 0.0;
  info->pinv_cutoff = rcond;
  (void)p;
  return JV_OK;
}

// The outcome record of a non-blocking solve, written on the device from the
// rank / route slot and (deferred) the E_loc mean and variance slot; a
// Cholesky factorisation that failed on the accepted route is reported as
// iterations = -1 - its info.
__global__ void write_solve_info(const int* __restrict__ hr, const double* __restrict__ em, double rcond,
                                 jv_solve_info* __restrict__ rec) {
  if (threadIdx.x != 0) return;
  rec->residual = 0.0;
  rec->pinv_cutoff = rcond;
  rec->iterations = (hr[1] == 1 && hr[3] != 0) ? -1 - hr[3] : hr[1];
  rec->rank = hr[0];
  rec->e_var = em ? em[2] : 0.0;
  rec->tdvp_error = 0.0;
  rec->e_mean[0] = em ? em[0] : 0.0;
  rec->e_mean[1] = em ? em[1] : 0.0;
}

// minsr_dual_finish_info without the host synchronisation: the record is
// formed on the device and copied to info in stream order.
static int minsr_dual_post_info(int rc, jv_solve_info* info, double rcond, bool deferred, hipStream_t st) {
  if (rc || !info) return rc;
  const int* rank = (const int*)workspace(64, 11);
  const double* emean = deferred ? (const double*)workspace(sizeof(double) * 4, 50) : nullptr;
  jv_solve_info* rec = (jv_solve_info*)workspace(sizeof(jv_solve_info), 121);
  if (!rank || (deferred && !emean) || !rec) return fail(JV_E_NOMEM, "jv_minsr_solve_dual: workspace allocation failed");
  hipLaunchKernelGGL(write_solve_info, dim3(1), dim3(64), 0, st, rank, emean, rcond, rec);
  JV_LAUNCH_CHECK("write_solve_info");
  JV_HIP(hipMemcpyAsync(info, rec, sizeof(jv_solve_info), hipMemcpyDeviceToHost, st));
  return JV_OK;
}
}  // namespace jv
