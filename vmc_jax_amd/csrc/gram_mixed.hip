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

// Workgroups are dispatched to the 8 XCDs round-robin by linear id, and each
// XCD has its own L2.  Workgroup w of nwg takes the (slice, tile) pair at
// position start(w % 8) + w / 8 of the slice-major order, so every XCD works
// through one contiguous run of pairs: whole sample slices when there are at
// least 8 of them (each slice's column panels are then fetched by one XCD),
// else a run of tiles of one slice.  Within a slice the tiles are ordered by
// square super-blocks of XSB x XSB tiles, so a run touches few column panels.
// Returns the tile (tm <= tn), its row-major upper-triangle index and the slice.
constexpr int XCDS = 8;
constexpr int XSB = 8;
__device__ inline void xcd_pair(int w, int nwg, int nt, int& tm, int& tn, int& tile, int& slice) {
  const int x = w % XCDS, j = w / XCDS;
  const int base = nwg / XCDS, rem = nwg % XCDS;
  const int q = x * base + (x < rem ? x : rem) + j;
  const int ntiles = nt * (nt + 1) / 2;
  slice = q / ntiles;
  int k = q - slice * ntiles;
  const int nsb = (nt + XSB - 1) / XSB;
  tm = tn = tile = 0;
  for (int I = 0; I < nsb; ++I) {
    const int r0 = I * XSB, r1 = min(nt, r0 + XSB);
    for (int J = I; J < nsb; ++J) {
      const int c0 = J * XSB, c1 = min(nt, c0 + XSB);
      const int cnt = I == J ? (r1 - r0) * (r1 - r0 + 1) / 2 : (r1 - r0) * (c1 - c0);
      if (k < cnt) {
        if (I == J) {
          upper_tile(k, r1 - r0, tm, tn);
          tm += r0;
          tn += r0;
        } else {
          tm = r0 + k / (c1 - c0);
          tn = c0 + k % (c1 - c0);
        }
        tile = tm * nt - tm * (tm - 1) / 2 + (tn - tm);
        return;
      }
      k -= cnt;
    }
  }
}

__device__ inline void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}
__device__ inline void split_bf16(double x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)(float)x;
  lo = (__bf16)(float)(x - (double)(float)hi);
}

}  // namespace

// ------------------------------------------------------------------ FP64 Gram
// R = X^T X in FP64 on the FP64 matrix cores (v_mfma_f64_16x16x4f64): the
// same 128 x 128 upper tiles and sample slices as the compensated Gram, 32
// samples per LDS stage kept in X's own [sample][column] order (row pitch 130
// doubles), two 16-byte loads per thread and row group issued one stage ahead
// from clamped addresses.  A lane's fragment is A[row l % 16][k l / 16] and
// B[k l / 16][column l % 16]; result element v of lane l is
// C[4 v + l / 16][l % 16] (the FP64 instruction's row order, unlike the BF16
// one's C[4 (l / 16) + v][l % 16]).
constexpr int G6K = 32;
constexpr int G6ROW = GT + 2;

template <bool VEC>
__global__ void __launch_bounds__(256) gram64_tn(const double* __restrict__ X, int64_t k_total, int64_t q,
                                                 int64_t kchunk, int ntile_side, double* __restrict__ part) {
  chain_priority();
  __shared__ double As[G6K][G6ROW];
  __shared__ double Bs[G6K][G6ROW];
  int tm, tn, tile, slice;
  xcd_pair(blockIdx.x, gridDim.x, ntile_side, tm, tn, tile, slice);
  const int ntiles = ntile_side * (ntile_side + 1) / 2;
  const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
  const int wm = (wv / 2) * 64, wn = (wv % 2) * 64;
  const int cp = tid % 64, rg = tid / 64;  // column pair, row group: rows rg + 4 j
  const int64_t m0 = (int64_t)tm * GT, n0 = (int64_t)tn * GT;
  const int64_t kbeg = (int64_t)slice * kchunk;
  const int64_t kend = std::min<int64_t>(k_total, kbeg + kchunk);
  typedef double f64x4 __attribute__((ext_vector_type(4)));
  f64x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
  double2 va[8], vb[8];
  const int64_t ca = m0 + 2 * cp, cb = n0 + 2 * cp;
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int64_t k = k0 + rg + 4 * j;
      const bool kin = k < kend;
      const int64_t kc = kin ? k : 0;
      double2 a, b;
      if (VEC) {  // q even: a column pair lies inside or outside [0, q) as a whole, 16-byte aligned
        a = *(const double2*)(X + kc * q + (ca < q ? ca : 0));
        b = *(const double2*)(X + kc * q + (cb < q ? cb : 0));
      } else {
        a.x = X[kc * q + (ca < q ? ca : 0)];
        a.y = X[kc * q + (ca + 1 < q ? ca + 1 : 0)];
        b.x = X[kc * q + (cb < q ? cb : 0)];
        b.y = X[kc * q + (cb + 1 < q ? cb + 1 : 0)];
      }
      va[j] = double2{(kin && ca < q) ? a.x : 0.0, (kin && ca + 1 < q) ? a.y : 0.0};
      vb[j] = double2{(kin && cb < q) ? b.x : 0.0, (kin && cb + 1 < q) ? b.y : 0.0};
    }
  };
  if (kbeg < kend) load(kbeg);
  const int fr = lane % 16, fk = lane / 16;
  for (int64_t k0 = kbeg; k0 < kend; k0 += G6K) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      As[rg + 4 * j][2 * cp] = va[j].x;
      As[rg + 4 * j][2 * cp + 1] = va[j].y;
      Bs[rg + 4 * j][2 * cp] = vb[j].x;
      Bs[rg + 4 * j][2 * cp + 1] = vb[j].y;
    }
    __syncthreads();
    if (k0 + G6K < kend) load(k0 + G6K);
#pragma unroll
    for (int ks = 0; ks < G6K; ks += 4) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[ks + fk][wm + 16 * i + fr];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[ks + fk][wn + 16 * j + fr];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }
  double* P = part + ((size_t)slice * ntiles + tile) * GT * GT;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = wm + 16 * i + 4 * v + lane / 16, c = wn + 16 * j + lane % 16;
        P[r * GT + c] = acc[i][j][v];
      }
}

// Split pass: X [k][q] (FP32 or FP64) -> Y, bf16, blocked for the Gram's
// stage loads: Y[((kb * qp + c) * 2 + h) * GK + j] = (h ? lo : hi) of
// X[kb * GK + j][c], zero for samples >= k_total or columns >= q (qp = the
// column count rounded up to whole tiles).  One thread per 8 samples of one
// column; consecutive threads take consecutive columns (coalesced reads).
template <typename T>
__global__ void __launch_bounds__(256) gram3_split(const T* __restrict__ X, int64_t k_total, int64_t q, int64_t qp,
                                                   int64_t nkb, __bf16* __restrict__ Y) {
  chain_priority();
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nkb * qp * (GK / 8)) return;
  const int64_t c = t % qp, rest = t / qp;
  const int g = (int)(rest % (GK / 8));
  const int64_t kb = rest / (GK / 8);
  bf16x8 h, l;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int64_t k = kb * GK + 8 * g + r;
    const bool in = k < k_total && c < q;
    const T x = X[(in ? k : 0) * q + (in ? c : 0)];
    __bf16 a, b;
    split_bf16(in ? x : (T)0, a, b);
    h[r] = a;
    l[r] = b;
  }
  __bf16* dst = Y + ((kb * qp + c) * 2) * GK + 8 * g;
  *(bf16x8*)dst = h;
  *(bf16x8*)(dst + GK) = l;
}

// gram3_tn on pre-split data: a stage's 128 columns x 2 planes x 64 samples
// of one operand are 32 KiB contiguous in Y, read as 8 coalesced 16-byte
// loads per thread and stored to the LDS planes without conversion.
__global__ void __launch_bounds__(256) gram3_pre(const __bf16* __restrict__ Y, int64_t nkb, int64_t qp, int64_t kbchunk,
                                                 int ntile_side, float* __restrict__ part) {
  chain_priority();
  extern __shared__ __align__(16) unsigned char gram_lds[];
  __bf16* Ap = (__bf16*)gram_lds;          // [2][GT][GROW]: hi then lo
  __bf16* Bp = Ap + 2 * GPLANE;
  int tm, tn, tile, slice;
  xcd_pair(blockIdx.x, gridDim.x, ntile_side, tm, tn, tile, slice);
  const int ntiles = ntile_side * (ntile_side + 1) / 2;
  const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
  const int wm = (wv / 2) * 64, wn = (wv % 2) * 64;
  const int64_t kb0 = (int64_t)slice * kbchunk;
  const int64_t kb1 = std::min<int64_t>(nkb, kb0 + kbchunk);
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  // chunk u (0 .. 2047) of a stage: column u / 16, plane (u / 8) % 2, samples 8 (u % 8) .. + 7
  bf16x8 va[8], vb[8];
  auto load = [&](int64_t kb) {
    const bf16x8* A = (const bf16x8*)(Y + (kb * qp + (int64_t)tm * GT) * 2 * GK);
    const bf16x8* B = (const bf16x8*)(Y + (kb * qp + (int64_t)tn * GT) * 2 * GK);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      va[j] = A[tid + 256 * j];
      vb[j] = B[tid + 256 * j];
    }
  };
  if (kb0 < kb1) load(kb0);
  const int fr = lane % 16, fk = 8 * (lane / 16);
  for (int64_t kb = kb0; kb < kb1; ++kb) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int u = tid + 256 * j;
      const int off = ((u / 8) % 2) * GPLANE + (u / 16) * GROW + 8 * (u % 8);
      *(bf16x8*)(Ap + off) = va[j];
      *(bf16x8*)(Bp + off) = vb[j];
    }
    __syncthreads();
    if (kb + 1 < kb1) load(kb + 1);
#pragma unroll
    for (int s = 0; s < GK; s += 32) {
      bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int off = (wm + 16 * i + fr) * GROW + s + fk;
        ah[i] = *(const bf16x8*)(Ap + off);
        al[i] = *(const bf16x8*)(Ap + GPLANE + off);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int off = (wn + 16 * j + fr) * GROW + s + fk;
        bh[j] = *(const bf16x8*)(Bp + off);
        bl[j] = *(const bf16x8*)(Bp + GPLANE + off);
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
  float* P = part + ((size_t)slice * ntiles + tile) * GT * GT;
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
template <typename TP>
__global__ void __launch_bounds__(256) gram3_reduce(const TP* __restrict__ part, int splits, int ntile_side,
                                                    int64_t p, int cpx, double* __restrict__ S) {
  chain_priority();
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
  // slices of the sample axis (whole 64-sample stages): enough workgroups for
  // two per CU, slices of at least 512 samples
  const int64_t nkb = (n + GK - 1) / GK, qp = side * GT;
  int64_t splits = 1;
  while (ntiles * splits < 512 && nkb / (splits * 2) >= 8) splits *= 2;
  const int64_t kbchunk = (nkb + splits - 1) / splits;
  splits = (nkb + kbchunk - 1) / kbchunk;
  float* part = (float*)workspace(sizeof(float) * (size_t)splits * ntiles * GT * GT, slot);
  __bf16* Y = (__bf16*)workspace(sizeof(__bf16) * (size_t)nkb * qp * 2 * GK, 120);
  if (!part || !Y) return fail(JV_E_NOMEM, "gram_pp_bf16x3: workspace allocation failed");
  const size_t lds = 4 * (size_t)GPLANE * sizeof(__bf16);
  ScopedTimer tm("gram", st);
  const int64_t nthr = nkb * qp * (GK / 8);
  if (f32)
    hipLaunchKernelGGL(gram3_split<float>, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, st, (const float*)D, n,
                       q, qp, nkb, Y);
  else
    hipLaunchKernelGGL(gram3_split<double>, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, st, (const double*)D,
                       n, q, qp, nkb, Y);
  JV_LAUNCH_CHECK("gram3_split");
  int rc = set_lds_limit((const void*)gram3_pre, lds);
  if (rc) return rc;
  hipLaunchKernelGGL(gram3_pre, dim3((unsigned)(ntiles * splits)), dim3(256), lds, st, Y, nkb, qp, kbchunk, nts, part);
  JV_LAUNCH_CHECK("gram3_pre");
  hipLaunchKernelGGL(gram3_reduce<float>, dim3((unsigned)(ntiles * 16)), dim3(256), 0, st, part, (int)splits, nts, p,
                     cpx, S);
  JV_LAUNCH_CHECK("gram3_reduce");
  return JV_OK;
}

// S = D^dagger D in FP64 from FP64 data D [n][p] (complex: interleaved), the
// symmetric / Hermitian Gram of jv_gram_pp.  Partial tiles in workspace slot
// `slot`.
int gram_pp_f64(const double* D, int64_t n, int64_t p, int cpx, double* S, hipStream_t st, int slot) {
  const int64_t q = p * (cpx ? 2 : 1);
  const int64_t side = (q + GT - 1) / GT;
  if (side > 4096) return fail(JV_E_ARG, "gram_pp_f64: matrix too large");
  const int nts = (int)side;
  const int64_t ntiles = side * (side + 1) / 2;
  const int64_t nkb = (n + G6K - 1) / G6K;
  int64_t splits = 1;
  while (ntiles * splits < 512 && nkb / (splits * 2) >= 16) splits *= 2;
  const int64_t kchunk = (nkb + splits - 1) / splits * G6K;
  splits = (n + kchunk - 1) / kchunk;
  double* part = (double*)workspace(sizeof(double) * (size_t)splits * ntiles * GT * GT, slot);
  if (!part) return fail(JV_E_NOMEM, "gram_pp_f64: workspace allocation failed");
  ScopedTimer tm("gram", st);
  if (q % 2 == 0 && ((uintptr_t)D % 16) == 0)
    hipLaunchKernelGGL(gram64_tn<true>, dim3((unsigned)(ntiles * splits)), dim3(256), 0, st, D, n, q, kchunk, nts,
                       part);
  else
    hipLaunchKernelGGL(gram64_tn<false>, dim3((unsigned)(ntiles * splits)), dim3(256), 0, st, D, n, q, kchunk, nts,
                       part);
  JV_LAUNCH_CHECK("gram64_tn");
  hipLaunchKernelGGL(gram3_reduce<double>, dim3((unsigned)(ntiles * 16)), dim3(256), 0, st, part, (int)splits, nts,
                     p, cpx, S);
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
