// Mixed-precision Gram matrices on the BF16 matrix cores (gfx950
// v_mfma_f32_16x16x32_bf16) with compensated products.
//
// Every input element x (FP32 or FP64) is split into x_hi = bf16(x) and
// x_lo = bf16(x - x_hi), so x = x_hi + x_lo + O(2^-18 |x|); a product is taken
// as x_hi y_hi + x_hi y_lo + x_lo y_hi (three BF16 MFMAs, FP32 accumulation),
// dropping x_lo y_lo = O(2^-18 |x y|).  Partial Grams over slices of the sample
// axis are summed in FP64, and the result is written in FP64 for the FP64
// solvers.
//
// R = X^T X for a real row-major X [k][q] (k = samples): for real data X = D
// gives S = D^T D (stats.py:235-245, the dual MinSR Gram of minsr.py:67-78);
// for complex data X is D's interleaved (re, im) storage viewed as [k][2p], and
// the Hermitian A = D^H D follows from the 2 x 2 blocks of R:
//   Re A_ab = R[2a][2b] + R[2a+1][2b+1],  Im A_ab = R[2a][2b+1] - R[2a+1][2b].
#include <algorithm>
#include <cmath>
#include <string>

#include "common.hpp"

namespace jv {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {
constexpr int GT = 128;        // output tile (rows and columns of R)
constexpr int GK = 64;         // samples per LDS stage
constexpr int GROW = GK + 8;   // LDS row pitch in bf16 (144 B: conflict-free 16-byte reads)
constexpr int GPLANE = GT * GROW;

// Upper-triangle tile t -> (tm, tn), tm <= tn, row by row.
__device__ inline void upper_tile(int t, int nt, int& tm, int& tn) {
  int r = 0, len = nt;
  while (t >= len) {
    t -= len;
    ++r;
    --len;
  }
  tm = r;
  tn = r + t;
}

__device__ inline void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}
__device__ inline void split_bf16(double x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)(float)x;
  lo = (__bf16)(float)(x - (double)(float)hi);
}

// Four consecutive columns X[k][col .. col + 3] (col % 4 == 0): one 16-byte
// (FP32) or two 16-byte (FP64) loads when VEC (q % 4 == 0, so a group of four
// lies entirely inside or outside [0, q)), else four element loads.
template <typename T, bool VEC>
__device__ inline void load4(const T* __restrict__ X, int64_t q, int64_t k, int64_t col, T v[4]) {
  if (VEC) {
    if constexpr (sizeof(T) == 4) {
      const float4 x = *(const float4*)(X + k * q + col);
      v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
    } else {
      const double2 x = *(const double2*)(X + k * q + col);
      const double2 y = *(const double2*)(X + k * q + col + 2);
      v[0] = x.x; v[1] = x.y; v[2] = y.x; v[3] = y.y;
    }
  } else {
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = X[k * q + (col + c < q ? col + c : 0)];
  }
}

// Loads the 8 x 4 block X[k0 + 8 kg + r][c0 + 4 mg + c] (r < 8, c < 4) into v,
// zero outside [0, kend) x [0, q).  Addresses are clamped so every load is
// issued unconditionally.
template <typename T, bool VEC>
__device__ inline void load_block(const T* __restrict__ X, int64_t q, int64_t k0, int64_t kend, int64_t c0, int mg,
                                  int kg, T v[8][4]) {
  const int64_t col = c0 + 4 * mg;
  const int64_t colc = col < q ? col : 0;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int64_t k = k0 + 8 * kg + r;
    const bool kin = k < kend;
    T x[4];
    load4<T, VEC>(X, q, kin ? k : 0, colc, x);
#pragma unroll
    for (int c = 0; c < 4; ++c) v[r][c] = (kin && col + c < q) ? x[c] : (T)0;
  }
}

// Stores the split block into the [col][k] planes (hi, lo) of one operand.
template <typename T>
__device__ inline void store_block(__bf16* hi, __bf16* lo, int mg, int kg, const T v[8][4]) {
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    bf16x8 h, l;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      __bf16 a, b;
      split_bf16(v[r][c], a, b);
      h[r] = a;
      l[r] = b;
    }
    const int off = (4 * mg + c) * GROW + 8 * kg;
    *(bf16x8*)(hi + off) = h;
    *(bf16x8*)(lo + off) = l;
  }
}
}  // namespace

// One upper tile (tm <= tn) of R = X^T X over the sample slice blockIdx.y,
// written as FP32 to part[(blockIdx.y * ntiles + tile) * GT * GT + r * GT + c].
// 4 waves, each owning a 64 x 64 quarter as 4 x 4 MFMA blocks of 16 x 16.
template <typename T, bool VEC>
__global__ void __launch_bounds__(256) gram3_tn(const T* __restrict__ X, int64_t k_total, int64_t q, int64_t kchunk,
                                                int ntile_side, float* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char gram_lds[];
  __bf16* Ahi = (__bf16*)gram_lds;
  __bf16* Alo = Ahi + GPLANE;
  __bf16* Bhi = Alo + GPLANE;
  __bf16* Blo = Bhi + GPLANE;
  int tm, tn;
  upper_tile(blockIdx.x, ntile_side, tm, tn);
  const int ntiles = ntile_side * (ntile_side + 1) / 2;
  const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
  const int wm = (wv / 2) * 64, wn = (wv % 2) * 64;
  const int mg = tid % 32, kg = tid / 32;
  const int64_t m0 = (int64_t)tm * GT, n0 = (int64_t)tn * GT;
  const int64_t kbeg = (int64_t)blockIdx.y * kchunk;
  const int64_t kend = std::min<int64_t>(k_total, kbeg + kchunk);
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  T va[8][4], vb[8][4];
  if (kbeg < kend) {
    load_block<T, VEC>(X, q, kbeg, kend, m0, mg, kg, va);
    load_block<T, VEC>(X, q, kbeg, kend, n0, mg, kg, vb);
  }
  const int fr = lane % 16, fk = 8 * (lane / 16);
  for (int64_t k0 = kbeg; k0 < kend; k0 += GK) {
    __syncthreads();
    store_block(Ahi, Alo, mg, kg, va);
    store_block(Bhi, Blo, mg, kg, vb);
    __syncthreads();
    if (k0 + GK < kend) {
      load_block<T, VEC>(X, q, k0 + GK, kend, m0, mg, kg, va);
      load_block<T, VEC>(X, q, k0 + GK, kend, n0, mg, kg, vb);
    }
#pragma unroll
    for (int s = 0; s < GK; s += 32) {
      bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int off = (wm + 16 * i + fr) * GROW + s + fk;
        ah[i] = *(const bf16x8*)(Ahi + off);
        al[i] = *(const bf16x8*)(Alo + off);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int off = (wn + 16 * j + fr) * GROW + s + fk;
        bh[j] = *(const bf16x8*)(Bhi + off);
        bl[j] = *(const bf16x8*)(Blo + off);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
        }
    }
  }
  // result element v of lane l in block (i, j): row 4 (l / 16) + v, column l % 16
  float* P = part + ((size_t)blockIdx.y * ntiles + blockIdx.x) * GT * GT;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = wm + 16 * i + 4 * (lane / 16) + v, c = wn + 16 * j + lane % 16;
        P[r * GT + c] = acc[i][j][v];
      }
}

// S (FP64, row-major, ld = p) from the tile partials: real S[i][j] = R[i][j];
// complex (cpx) S = A with the 2 x 2 combination above.  One workgroup per
// 32 x 32 block of an upper tile: partial sums read along rows (coalesced),
// added in slice order in FP64 and staged in LDS, so both the block and its
// mirror (transpose, conjugate for complex data) are written along rows.  Only
// entries on or above the diagonal of R are used, and the lower triangle of S
// is their exact mirror.
__global__ void __launch_bounds__(256) gram3_reduce(const float* __restrict__ part, int splits, int ntile_side,
                                                    int64_t p, int cpx, double* __restrict__ S) {
  __shared__ double M[32][33];
  int tm, tn;
  upper_tile(blockIdx.x / 16, ntile_side, tm, tn);
  const int sa = (blockIdx.x % 16) / 4, sb = blockIdx.x % 4;
  if (tm == tn && sa > sb) return;  // below the diagonal of a diagonal tile: the mirror of (sb, sa)
  const int ntiles = ntile_side * (ntile_side + 1) / 2;
  const int tile = blockIdx.x / 16;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
  const int64_t q = p * (cpx ? 2 : 1);
  const int64_t r0 = (int64_t)tm * GT + 32 * sa, c0 = (int64_t)tn * GT + 32 * sb;  // R indices of the block
  const size_t stride = (size_t)ntiles * GT * GT;
  for (int k = 0; k < 4; ++k) {
    const int i = ty + 8 * k;
    const size_t off = (size_t)tile * GT * GT + (size_t)(32 * sa + i) * GT + (32 * sb + tx);
    double v = 0.0;
    for (int sp = 0; sp < splits; ++sp) v += (double)part[sp * stride + off];
    M[i][tx] = v;
  }
  __syncthreads();
  if (!cpx) {
    for (int k = 0; k < 4; ++k) {
      const int i = ty + 8 * k;
      const int64_t r = r0 + i, c = c0 + tx;
      if (r < p && c < p && r <= c) S[r * p + c] = M[i][tx];
      // mirror: S[c0 + i][r0 + tx] = R[r0 + tx][c0 + i]
      const int64_t mr = c0 + i, mc = r0 + tx;
      if (mr < p && mc < p && mc < mr) S[mr * p + mc] = M[tx][i];
    }
    return;
  }
  // complex: 16 x 16 outputs of the block, A = r0 / 2 + a, B = c0 / 2 + b
  {
    const int a = threadIdx.x / 16, b = threadIdx.x % 16;
    const int64_t A = r0 / 2 + a, B = c0 / 2 + b;
    if (2 * A < q && 2 * B < q && A <= B) {
      const double re = M[2 * a][2 * b] + M[2 * a + 1][2 * b + 1];
      const double im = A == B ? 0.0 : M[2 * a][2 * b + 1] - M[2 * a + 1][2 * b];
      S[2 * (A * p + B)] = re;
      S[2 * (A * p + B) + 1] = im;
    }
  }
  {
    const int a = threadIdx.x % 16, b = threadIdx.x / 16;  // mirror S[B][A], consecutive threads along A
    const int64_t A = r0 / 2 + a, B = c0 / 2 + b;
    if (2 * A < q && 2 * B < q && A < B) {
      const double re = M[2 * a][2 * b] + M[2 * a + 1][2 * b + 1];
      const double im = M[2 * a][2 * b + 1] - M[2 * a + 1][2 * b];
      S[2 * (B * p + A)] = re;
      S[2 * (B * p + A) + 1] = -im;
    }
  }
  (void)q;
}

int set_lds_limit(const void* fn, size_t bytes);

// S = D^dagger D from FP32 (f32 != 0) or FP64 data D [n][p] (complex:
// interleaved), result FP64 [p][p] (complex interleaved).  Partial tiles live
// in workspace slot `slot`.
int gram_pp_bf16x3(const void* D, int f32, int64_t n, int64_t p, int cpx, double* S, hipStream_t st, int slot) {
  const int64_t q = p * (cpx ? 2 : 1);
  const int64_t side = (q + GT - 1) / GT;
  if (side > 4096) return fail(JV_E_ARG, "gram_pp_bf16x3: matrix too large");
  const int nts = (int)side;
  const int64_t ntiles = side * (side + 1) / 2;
  // slices of the sample axis: enough workgroups for two per CU, slices of at
  // least 512 samples
  int64_t splits = 1;
  while (ntiles * splits < 512 && (n + splits * 2 - 1) / (splits * 2) >= 512) splits *= 2;
  int64_t kchunk = ((n + splits - 1) / splits + GK - 1) / GK * GK;
  splits = (n + kchunk - 1) / kchunk;
  if (splits < 1) splits = 1;
  float* part = (float*)workspace(sizeof(float) * (size_t)splits * ntiles * GT * GT, slot);
  if (!part) return fail(JV_E_NOMEM, "gram_pp_bf16x3: workspace allocation failed");
  const size_t lds = 4 * (size_t)GPLANE * sizeof(__bf16);
  const dim3 grid((unsigned)ntiles, (unsigned)splits);
  ScopedTimer tm("gram", st);
  const bool vec = q % 4 == 0 && ((uintptr_t)D % 16) == 0;
#define JV_G3(Tv, Vv)                                                                              \
  {                                                                                                \
    int rc = set_lds_limit((const void*)gram3_tn<Tv, Vv>, lds);                                    \
    if (rc) return rc;                                                                             \
    hipLaunchKernelGGL((gram3_tn<Tv, Vv>), grid, dim3(256), lds, st, (const Tv*)D, n, q, kchunk, nts, part); \
  }
  if (f32) {
    if (vec) JV_G3(float, true) else JV_G3(float, false)
  } else {
    if (vec) JV_G3(double, true) else JV_G3(double, false)
  }
#undef JV_G3
  JV_LAUNCH_CHECK("gram3_tn");
  hipLaunchKernelGGL(gram3_reduce, dim3((unsigned)(ntiles * 16)), dim3(256), 0, st, part, (int)splits, nts, p, cpx, S);
  JV_LAUNCH_CHECK("gram3_reduce");
  return JV_OK;
}

}  // namespace jv

extern "C" {

int jv_gram_pp_mixed(const void* data, int32_t data_f32, int64_t n, int64_t p, int32_t is_complex, double* S,
                     void* stream) {
  if (!data || !S || n < 1 || p < 1) return jv::fail(JV_E_ARG, "jv_gram_pp_mixed: bad arguments");
  return jv::gram_pp_bf16x3(data, data_f32, n, p, is_complex, S, (hipStream_t)stream, 19);
}

}  // extern "C"
