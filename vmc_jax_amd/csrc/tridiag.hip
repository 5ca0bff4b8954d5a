// Householder tridiagonalisation of a real symmetric matrix, one launch per
// column, and the application of its orthogonal factor to vectors: the
// eigendecomposition behind the pseudo-inverse of the MinSR Gram matrix
// (minsr.py:67-78, jnp.linalg.pinv(hermitian=True) = eigh + relative cutoff).
//
// A = Q T Q^T, Q = H_0 H_1 ... H_{n-2}, H_k = I - tau_k v_k v_k^T with
// v_k[k+1] = 1 and v_k zero above k+1 (LAPACK dsytrd 'L' conventions: the same
// reflector and sign rules as dlarfg, beta = -sign(alpha) ||x||).
//
// Step k needs y = A^(k-1) v_k (the trailing matrix after k-1 updates) from
// every column, so each step is one launch (the launch boundary is the grid
// synchronisation).  Workgroup g owns columns j = g (mod G); the rank-2
// update of step k-1 is deferred into step k's pass over the owned columns,
// which also forms y_j = column_j . v_k and
//   u_j = A^(k-1)[j][k+1] - tau_k y_j.
// Column k+1 after update k is then x = u + gamma v_k with the scalar
// gamma = 2 alpha' - tau_k y_{k+1}, alpha' = tau_k^2 / 2 (y . v_k), so every
// workgroup rebuilds x, v_{k+1} and w_k = tau_k y - alpha' v_k from the
// published y, u and the per-workgroup partial dots alone: no workgroup ever
// waits for another inside a launch.  One extra workgroup applies H_k to a
// right-hand side as the reflectors appear (c = Q^T b).
#include <hip/hip_runtime.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>

#include "common.hpp"

namespace jv {

rocblas_handle blas_handle(hipStream_t stream);

namespace {

constexpr int TRD_T = 512;
constexpr int TRD_W = TRD_T / WAVE;
constexpr int TRD_R = 9;  // n-vector elements per thread (n <= TRD_T * TRD_R)
constexpr int TRD_CP = 16;  // column rows per lane per load batch

struct TrdArgs {
  double* M;    // n x n symmetric (row k is overwritten by reflector k once step k has run)
  double* d;    // diagonal of T [n]
  double* e;    // off-diagonal of T [n-1]
  double* tau;  // [n]
  double* y;    // [2][n]
  double* u;    // [2][n]
  double* pd;   // [2][G]
  double* c;    // right-hand side receiving Q^T (may be null)
  double* v0;   // reflector 0 until step 1 copies it into row 0 of M
  int n, G;
};

// Block sum in a fixed order (identical in every workgroup for identical
// inputs); red holds TRD_W doubles.
__device__ inline double trd_block_sum(double v, double* red) {
  v = wave_sum(v);
  const int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < TRD_W; ++i) s += red[i];
  return s;
}

// Step k of the reduction (k = 0 .. n-1; step n-1 only finishes d[n-1]).
// Global reads are batched per thread (all loads of a batch are issued before
// any is used): each step is latency bound, not bandwidth bound.
__global__ void __launch_bounds__(TRD_T) trd_step(TrdArgs a, int k) {
  chain_priority();
  extern __shared__ double sh[];
  const int n = a.n, G = a.G, tid = threadIdx.x, g = blockIdx.x;
  const bool qwg = g == G;
  // first owned column >= k+1
  int j0 = g + ((k + 1 - g + G - 1) / G) * G;
  if (g >= k + 1) j0 = g;
  const bool owns = !qwg && j0 < n;
  if (!owns && g != 0 && !qwg) {
    if (threadIdx.x == 0) a.pd[(k & 1) * G + g] = 0.0;
    return;
  }
  if (qwg && (a.c == nullptr || k > n - 2)) return;
  const int w = tid / WAVE, lane = tid % WAVE;
  const int jw = j0 + w * G;
  double* vk = sh;          // v_k
  double* vp = sh + n;      // v_{k-1}
  double* wp = sh + 2 * n;  // w_{k-1}
  double* x = sh + 3 * n;   // column k after update k-1
  double* red = sh + 4 * n; // 2 * TRD_W
  // ---- prologue: rebuild column k, its reflector and w_{k-1}
  if (k == 0) {
    double xr[TRD_R];
#pragma unroll
    for (int r = 0; r < TRD_R; ++r) {
      const int i = tid + r * TRD_T;
      xr[r] = i < n ? a.M[i] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < TRD_R; ++r) {
      const int i = tid + r * TRD_T;
      if (i < n) {
        x[i] = xr[r];
        vp[i] = 0.0;
        wp[i] = 0.0;
      }
    }
  } else {
    const int pb = (k - 1) & 1;
    const double* yv = a.y + (size_t)pb * n;
    const double* uv = a.u + (size_t)pb * n;
    const double* vrow = k == 1 ? a.v0 : a.M + (size_t)(k - 1) * n;
    double yr[TRD_R], ur[TRD_R], vr[TRD_R];
#pragma unroll
    for (int r = 0; r < TRD_R; ++r) {
      const int i = tid + r * TRD_T;
      const bool in = i >= k && i < n;
      yr[r] = in ? yv[i] : 0.0;
      ur[r] = in ? uv[i] : 0.0;
      vr[r] = in ? vrow[i] : 0.0;
    }
    double pdv = 0.0;
    for (int q = tid; q < G; q += TRD_T) pdv += a.pd[pb * G + q];
    const double tp = a.tau[k - 1];
    const double yk = yv[k];
    const double sdot = trd_block_sum(pdv, red);
    const double ap = 0.5 * tp * tp * sdot;
    const double gam = 2.0 * ap - tp * yk;
#pragma unroll
    for (int r = 0; r < TRD_R; ++r) {
      const int i = tid + r * TRD_T;
      if (i < n) {
        vp[i] = vr[r];
        wp[i] = tp * yr[r] - ap * vr[r];
        x[i] = ur[r] + gam * vr[r];
      }
    }
  }
  __syncthreads();
  double tk = 0.0;
  if (k <= n - 2) {
    double part = 0.0;
#pragma unroll
    for (int r = 0; r < TRD_R; ++r) {
      const int i = tid + r * TRD_T;
      if (i >= k + 2 && i < n) part += x[i] * x[i];
    }
    const double sigma = trd_block_sum(part, red + TRD_W);
    const double alpha = x[k + 1];
    double beta = alpha, scale = 0.0;
    if (sigma != 0.0) {
      beta = -copysign(sqrt(alpha * alpha + sigma), alpha);
      tk = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    for (int i = tid; i < n; i += TRD_T) vk[i] = i < k + 1 ? 0.0 : (i == k + 1 ? 1.0 : x[i] * scale);
    if (g == 0 && tid == 0) {
      a.e[k] = beta;
      a.tau[k] = tk;
    }
  } else {
    for (int i = tid; i < n; i += TRD_T) vk[i] = 0.0;
  }
  if (g == 0 && tid == 0) a.d[k] = x[k];
  __syncthreads();
  if (k > n - 2) return;
  if (g == 0) {
    // reflector k into row k of M: no step >= k reads column k, except that
    // step 0's prologue reads column 0, so reflector 0 waits in v0 until step 1
    double* row = k == 0 ? a.v0 : a.M + (size_t)k * n;
    for (int i = tid; i < n; i += TRD_T) row[i] = vk[i];
    if (k == 1)
      for (int i = tid; i < n; i += TRD_T) a.M[i] = a.v0[i];
  }
  if (qwg) {  // c <- H_k c
    if (tk != 0.0) {
      double cr[TRD_R];
      double part = 0.0;
#pragma unroll
      for (int r = 0; r < TRD_R; ++r) {
        const int i = tid + r * TRD_T;
        cr[r] = (i > k && i < n) ? a.c[i] : 0.0;
      }
#pragma unroll
      for (int r = 0; r < TRD_R; ++r) {
        const int i = tid + r * TRD_T;
        if (i > k && i < n) part += vk[i] * cr[r];
      }
      const double dot = trd_block_sum(part, red) * tk;
#pragma unroll
      for (int r = 0; r < TRD_R; ++r) {
        const int i = tid + r * TRD_T;
        if (i > k && i < n) a.c[i] = cr[r] - dot * vk[i];
      }
    }
    return;
  }
  if (!owns) {
    if (tid == 0) a.pd[(k & 1) * G + g] = 0.0;
    return;
  }
  // ---- owned columns j >= k+1: apply update k-1 to rows >= k+1, y_j, u_j
  const int cb = k & 1;
  double ppart = 0.0;
  for (int j = jw; j < n; j += TRD_W * G) {
    double* col = a.M + (size_t)j * n;
    const double vj = vp[j], wj = wp[j];
    double acc = 0.0, cj = 0.0;
    for (int base = k + 1; base < n; base += TRD_CP * WAVE) {
      double cv[TRD_CP];
#pragma unroll
      for (int t = 0; t < TRD_CP; ++t) {
        const int i = base + t * WAVE + lane;
        cv[t] = i < n ? col[i] : 0.0;
      }
#pragma unroll
      for (int t = 0; t < TRD_CP; ++t) {
        const int i = base + t * WAVE + lane;
        if (i < n) {
          double v = cv[t];
          if (k >= 1) v -= vp[i] * wj + wp[i] * vj;
          col[i] = v;
          acc += v * vk[i];
          if (i == k + 1) cj = v;
        }
      }
    }
    const double yj = wave_sum(acc);
    if (lane == 0) {
      a.y[(size_t)cb * n + j] = yj;
      a.u[(size_t)cb * n + j] = cj - tk * yj;
      ppart += yj * vk[j];
    }
  }
  __syncthreads();
  if (lane == 0) red[w] = ppart;
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int i = 0; i < TRD_W; ++i) s += red[i];
    a.pd[cb * G + g] = s;
  }
}

// ---- complex Hermitian version (LAPACK zhetrd 'L' conventions: H_k = I - tau_k
// v_k v_k^H, beta real, so T is real symmetric tridiagonal).  Row j of M holds
// A[j][:]; the rank-2 update of step k-1 is A - v w^H - w v^H with
// w = tau y - alpha' v, alpha' = |tau|^2 / 2 Re(y^H v), y = A v; column k+1
// after update k is x = u + gamma v with u_j = A[j][k+1] - tau y_j and
// gamma = 2 alpha' - conj(tau y_{k+1}).  Step 0 reads column 0 as conj(row 0).
constexpr int TRDC_R = 5;   // complex n-vector elements per thread
constexpr int TRDC_CP = 8;  // row elements per lane per load batch

struct TrdcArgs {
  double* M;    // n x n Hermitian, complex interleaved (row k <- reflector k after step k)
  double* d;    // [n]
  double* e;    // [n-1]
  double* tau;  // complex [n]
  double* y;    // complex [2][n]
  double* u;    // complex [2][n]
  double* pd;   // [2][G]
  double* v0;   // complex [n]: reflector 0 until step 1 copies it into row 0
  int n, G;
};

__device__ inline cd cmul(cd a, cd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline cd cmulc(cd a, cd b) { return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im}; }  // a conj(b)

// R: complex n-vector elements per thread (n <= TRD_T * R).  VG = false keeps
// v_{k-1}, w_{k-1} and the column in LDS next to v_k (64 n bytes); VG = true
// (large n) keeps only v_k in LDS (16 n bytes: the column is built in its
// place) and the column pass rebuilds v_{k-1} (row k-1 of M) and
// w_{k-1} = tau_{k-1} y_{k-1} - alpha' v_{k-1} from global memory.
template <int R, bool VG>
__global__ void __launch_bounds__(TRD_T) trd_step_c(TrdcArgs a, int k) {
  chain_priority();
  extern __shared__ double sh[];
  const int n = a.n, G = a.G, tid = threadIdx.x, g = blockIdx.x;
  int j0 = g + ((k + 1 - g + G - 1) / G) * G;
  if (g >= k + 1) j0 = g;
  const bool owns = j0 < n;
  if (!owns && g != 0) {
    if (threadIdx.x == 0) a.pd[(k & 1) * G + g] = 0.0;
    return;
  }
  const int w = tid / WAVE, lane = tid % WAVE;
  const int jw = j0 + w * G;
  cd* vk = (cd*)sh;
  cd* vp = VG ? nullptr : vk + n;
  cd* wp = VG ? nullptr : vk + 2 * n;
  cd* x = VG ? vk : vk + 3 * n;
  double* red = (double*)(VG ? vk + n : vk + 4 * n);  // 2 * TRD_W
  const cd* Mc = (const cd*)a.M;
  cd tp = {0.0, 0.0};  // tau_{k-1}
  double ap = 0.0;     // alpha'_{k-1}
  if (k == 0) {
    cd xr[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = tid + r * TRD_T;
      xr[r] = i < n ? Mc[i] : cd{0.0, 0.0};
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = tid + r * TRD_T;
      if (i < n) {
        x[i] = cconj(xr[r]);
        if (!VG) {
          vp[i] = {0.0, 0.0};
          wp[i] = {0.0, 0.0};
        }
      }
    }
  } else {
    const int pb = (k - 1) & 1;
    const cd* yv = (const cd*)a.y + (size_t)pb * n;
    const cd* uv = (const cd*)a.u + (size_t)pb * n;
    const cd* vrow = (const cd*)(k == 1 ? a.v0 : a.M + 2 * (size_t)(k - 1) * n);
    cd yr[R], ur[R], vr[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = tid + r * TRD_T;
      const bool in = i >= k && i < n;
      yr[r] = in ? yv[i] : cd{0.0, 0.0};
      ur[r] = in ? uv[i] : cd{0.0, 0.0};
      vr[r] = in ? vrow[i] : cd{0.0, 0.0};
    }
    double pdv = 0.0;
    for (int q = tid; q < G; q += TRD_T) pdv += a.pd[pb * G + q];
    tp = {a.tau[2 * (k - 1)], a.tau[2 * (k - 1) + 1]};
    const cd yk = yv[k];
    const double sdot = trd_block_sum(pdv, red);
    ap = 0.5 * cabs2(tp) * sdot;
    const cd ty = cmul(tp, yk);
    const cd gam = {2.0 * ap - ty.re, ty.im};  // 2 alpha' - conj(tau y_k)
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = tid + r * TRD_T;
      if (i < n) {
        if (!VG) {
          vp[i] = vr[r];
          const cd t = cmul(tp, yr[r]);
          wp[i] = {t.re - ap * vr[r].re, t.im - ap * vr[r].im};
        }
        const cd gv = cmul(gam, vr[r]);
        x[i] = {ur[r].re + gv.re, ur[r].im + gv.im};
      }
    }
  }
  __syncthreads();
  cd tk = {0.0, 0.0};
  const double xkk = x[k].re;
  double sigma = 0.0;
  cd alpha = {0.0, 0.0};
  if (k <= n - 2) {
    double part = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = tid + r * TRD_T;
      if (i >= k + 2 && i < n) part += cabs2(x[i]);
    }
    sigma = trd_block_sum(part, red + TRD_W);
    alpha = x[k + 1];
  }
  if (VG) __syncthreads();  // every read of x above precedes v_k overwriting it
  if (k <= n - 2) {
    double beta = alpha.re;
    cd scale = {0.0, 0.0};
    if (sigma != 0.0 || alpha.im != 0.0) {
      beta = -copysign(sqrt(cabs2(alpha) + sigma), alpha.re);
      tk = {(beta - alpha.re) / beta, -alpha.im / beta};
      const cd dn = {alpha.re - beta, alpha.im};
      const double inv = 1.0 / cabs2(dn);
      scale = {dn.re * inv, -dn.im * inv};
    }
    for (int i = tid; i < n; i += TRD_T)
      vk[i] = i < k + 1 ? cd{0.0, 0.0} : (i == k + 1 ? cd{1.0, 0.0} : cmul(x[i], scale));
    if (g == 0 && tid == 0) {
      a.e[k] = beta;
      a.tau[2 * k] = tk.re;
      a.tau[2 * k + 1] = tk.im;
    }
  } else {
    for (int i = tid; i < n; i += TRD_T) vk[i] = {0.0, 0.0};
  }
  if (g == 0 && tid == 0) a.d[k] = xkk;
  __syncthreads();
  if (k > n - 2) return;
  if (g == 0) {
    cd* row = (cd*)(k == 0 ? a.v0 : a.M + 2 * (size_t)k * n);
    for (int i = tid; i < n; i += TRD_T) row[i] = vk[i];
    if (k == 1)
      for (int i = tid; i < n; i += TRD_T) ((cd*)a.M)[i] = ((const cd*)a.v0)[i];
  }
  if (!owns) {
    if (tid == 0) a.pd[(k & 1) * G + g] = 0.0;
    return;
  }
  const int cb = k & 1;
  double ppart = 0.0;
  // VG: v_{k-1} is row k-1 of M (reflector 0 is still in v0 at k = 1; no row
  // read here is written in this step) and y_{k-1} is the other y buffer
  const cd* vg = (const cd*)(k == 1 ? a.v0 : a.M + 2 * (size_t)(k >= 1 ? k - 1 : 0) * n);
  const cd* yg = (const cd*)a.y + (size_t)((k + 1) & 1) * n;
  auto w_of = [&](cd v, cd y) {
    const cd t = cmul(tp, y);
    return cd{t.re - ap * v.re, t.im - ap * v.im};
  };
  for (int j = jw; j < n; j += TRD_W * G) {
    cd* row = (cd*)(a.M + 2 * (size_t)j * n);
    cd vj, wj;
    if (VG) {
      vj = k >= 1 ? vg[j] : cd{0.0, 0.0};
      wj = k >= 1 ? w_of(vj, yg[j]) : cd{0.0, 0.0};
    } else {
      vj = vp[j];
      wj = wp[j];
    }
    double accr = 0.0, acci = 0.0;
    cd cj = {0.0, 0.0};
    for (int base = k + 1; base < n; base += TRDC_CP * WAVE) {
      cd cv[TRDC_CP], gv[VG ? TRDC_CP : 1], gy[VG ? TRDC_CP : 1];
#pragma unroll
      for (int t = 0; t < TRDC_CP; ++t) {
        const int i = base + t * WAVE + lane;
        cv[t] = i < n ? row[i] : cd{0.0, 0.0};
        if (VG && k >= 1) {
          const int ic = min(i, n - 1);
          gv[VG ? t : 0] = vg[ic];
          gy[VG ? t : 0] = yg[ic];
        }
      }
#pragma unroll
      for (int t = 0; t < TRDC_CP; ++t) {
        const int i = base + t * WAVE + lane;
        if (i < n) {
          cd v = cv[t];
          if (k >= 1) {
            cd vpi, wpi;
            if (VG) {
              vpi = gv[VG ? t : 0];
              wpi = w_of(vpi, gy[VG ? t : 0]);
            } else {
              vpi = vp[i];
              wpi = wp[i];
            }
            const cd p1 = cmulc(vj, wpi), p2 = cmulc(wj, vpi);
            v = {v.re - p1.re - p2.re, v.im - p1.im - p2.im};
          }
          row[i] = v;
          const cd q = cmul(v, vk[i]);
          accr += q.re;
          acci += q.im;
          if (i == k + 1) cj = v;
        }
      }
    }
    const cd yj = {wave_sum(accr), wave_sum(acci)};
    if (lane == 0) {
      ((cd*)a.y)[(size_t)cb * n + j] = yj;
      const cd ty = cmul(tk, yj);
      ((cd*)a.u)[(size_t)cb * n + j] = {cj.re - ty.re, cj.im - ty.im};
      ppart += yj.re * vk[j].re + yj.im * vk[j].im;  // Re(conj(y_j) v_j)
    }
  }
  __syncthreads();
  if (lane == 0) red[w] = ppart;
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int i = 0; i < TRD_W; ++i) s += red[i];
    a.pd[cb * G + g] = s;
  }
}

// Real eigenvectors of T (column-major n x n) -> complex (imaginary part 0).
__global__ void trdc_expand(const double* __restrict__ Z, int64_t len, double* __restrict__ Zc) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= len) return;
  Zc[2 * t] = Z[t];
  Zc[2 * t + 1] = 0.0;
}

// out = conj(in) elementwise (complex interleaved)
__global__ void trdc_conj(const double* __restrict__ in, int64_t len, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= len) return;
  out[2 * t] = in[2 * t];
  out[2 * t + 1] = -in[2 * t + 1];
}

// ---- compact WY form of reflector blocks: H_k0 ... H_k1 = I - V T V^T with
// V = [v_k0 .. v_k1] and T upper triangular (LAPACK dlarft, forward, columnwise)
constexpr int WY_B = 32;

// T of reflector block blockIdx.x (reflectors blockIdx.x*WY_B ..), from the
// Gram Y = V^T V accumulated over 256-row chunks staged in LDS.
__global__ void __launch_bounds__(256) trd_wy_t(const double* __restrict__ V, const double* __restrict__ tau,
                                                int n, int nref, double* __restrict__ Tout) {
  chain_priority();
  __shared__ double Vs[WY_B][257];
  __shared__ double Y[WY_B][WY_B + 1];
  __shared__ double T[WY_B][WY_B + 1];
  const int tid = threadIdx.x;
  const int k0 = blockIdx.x * WY_B, nb = min(WY_B, nref - k0);
  // pairs (p, q), p <= q, two per thread (528 pairs for WY_B = 32)
  double acc[3] = {0.0, 0.0, 0.0};
  int pp[3], qq[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    int t = tid + r * 256, p = 0;
    while (p < WY_B && t >= WY_B - p) {
      t -= WY_B - p;
      ++p;
    }
    pp[r] = p < WY_B ? p : -1;
    qq[r] = p < WY_B ? p + t : -1;
  }
  for (int c0 = k0 + 1; c0 < n; c0 += 256) {
    __syncthreads();
    double st[WY_B];
    const int ic = min(c0 + tid, n - 1);
#pragma unroll
    for (int p = 0; p < WY_B; ++p) st[p] = V[(size_t)min(k0 + p, n - 1) * n + ic];
#pragma unroll
    for (int p = 0; p < WY_B; ++p) Vs[p][tid] = (p < nb && c0 + tid < n) ? st[p] : 0.0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      if (pp[r] < 0) continue;
      double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
      for (int i = 0; i < 256; i += 4) {
        a0 = fma(Vs[pp[r]][i], Vs[qq[r]][i], a0);
        a1 = fma(Vs[pp[r]][i + 1], Vs[qq[r]][i + 1], a1);
        a2 = fma(Vs[pp[r]][i + 2], Vs[qq[r]][i + 2], a2);
        a3 = fma(Vs[pp[r]][i + 3], Vs[qq[r]][i + 3], a3);
      }
      acc[r] += (a0 + a1) + (a2 + a3);
    }
  }
#pragma unroll
  for (int r = 0; r < 3; ++r)
    if (pp[r] >= 0) {
      Y[pp[r]][qq[r]] = acc[r];
      Y[qq[r]][pp[r]] = acc[r];
    }
  __syncthreads();
  if (tid < WAVE) {
    const int j = tid;
    for (int i = 0; i < WY_B; ++i) {
      const double ti = i < nb ? tau[k0 + i] : 0.0;
      double v = 0.0;
      if (j < i) {
        for (int l = j; l < i; ++l) v += T[j][l] * Y[l][i];
        v *= -ti;
      }
      __builtin_amdgcn_wave_barrier();
      if (j < WY_B) T[j][i] = j < i ? v : (j == i ? ti : 0.0);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  __syncthreads();
  for (int e = tid; e < WY_B * WY_B; e += 256) Tout[(size_t)blockIdx.x * WY_B * WY_B + e] = T[e / WY_B][e % WY_B];
}

// x <- Q x (TR = false: blocks last to first, x -= V (T (V^T x))) or
// x <- Q^T x (TR = true: blocks first to last with T^T).  The reflector loop
// is unrolled to WY_B so each thread's loads of a block are issued together
// (one memory latency per phase, not one per reflector).
template <bool TR>
__global__ void __launch_bounds__(TRD_T) trd_apply_wy(const double* __restrict__ V, const double* __restrict__ Tg,
                                                      int n, int nref, double* __restrict__ x) {
  chain_priority();
  __shared__ double part[TRD_W][WY_B];
  __shared__ double wv[WY_B], tv[WY_B];
  const int tid = threadIdx.x, w = tid / WAVE, lane = tid % WAVE;
  double xr[TRD_R];
#pragma unroll
  for (int r = 0; r < TRD_R; ++r) {
    const int i = tid + r * TRD_T;
    xr[r] = i < n ? x[i] : 0.0;
  }
  const int nblk = (nref + WY_B - 1) / WY_B;
  for (int bi = 0; bi < nblk; ++bi) {
    const int b = TR ? bi : nblk - 1 - bi;
    const int k0 = b * WY_B, nb = min(WY_B, nref - k0);
    // w = V^T x
    double acc[WY_B];
#pragma unroll
    for (int p = 0; p < WY_B; ++p) acc[p] = 0.0;
    // loads use clamped (always valid) indices instead of branches, so the
    // WY_B loads of one r are in flight together
#pragma unroll
    for (int r = 0; r < TRD_R; ++r) {
      if (r * TRD_T >= n) break;
      const int i = tid + r * TRD_T;
      const int ic = i < n ? i : n - 1;
      const double xv = i < n ? xr[r] : 0.0;
#pragma unroll
      for (int p = 0; p < WY_B; ++p) acc[p] = fma(V[(size_t)min(k0 + p, n - 1) * n + ic], xv, acc[p]);
    }
#pragma unroll
    for (int p = 0; p < WY_B; ++p) {
      const double t = wave_sum(acc[p]);
      if (lane == 0) part[w][p] = t;
    }
    __syncthreads();
    if (tid < nb) {
      double t = 0.0;
#pragma unroll
      for (int q = 0; q < TRD_W; ++q) t += part[q][tid];
      wv[tid] = t;
    }
    __syncthreads();
    // t = T w or T^T w (zero past the block's last reflector, whose rows are
    // never weighted: acc[p >= nb] is ignored and tv[p >= nb] = 0)
    if (tid >= nb && tid < WY_B) tv[tid] = 0.0;
    if (tid < nb) {
      const double* Tb = Tg + (size_t)b * WY_B * WY_B;
      double tr[WY_B];
#pragma unroll
      for (int l = 0; l < WY_B; ++l) tr[l] = TR ? Tb[(size_t)l * WY_B + tid] : Tb[(size_t)tid * WY_B + l];
      double t = 0.0;
#pragma unroll
      for (int l = 0; l < WY_B; ++l)
        if ((TR ? l <= tid : l >= tid) && l < nb) t = fma(tr[l], wv[l], t);
      tv[tid] = t;
    }
    __syncthreads();
    // x -= V t
#pragma unroll
    for (int r = 0; r < TRD_R; ++r) {
      if (r * TRD_T >= n) break;
      const int i = tid + r * TRD_T;
      const int ic = i < n ? i : n - 1;
      double t = 0.0;
#pragma unroll
      for (int p = 0; p < WY_B; ++p) t = fma(V[(size_t)min(k0 + p, n - 1) * n + ic], tv[p], t);
      xr[r] -= i < n ? t : 0.0;
    }
  }
#pragma unroll
  for (int r = 0; r < TRD_R; ++r) {
    const int i = tid + r * TRD_T;
    if (i < n) x[i] = xr[r];
  }
}

}  // namespace

// Largest matrix: an n-vector is TRD_R elements per thread.
constexpr int TRD_MAX_N = TRD_T * TRD_R;

int trd_groups(int n) {
  static int v = [] {
    const char* s = getenv("VMC_JAX_AMD_TRIDIAG_GROUPS");
    return s ? atoi(s) : 0;
  }();
  return v > 0 ? std::min(std::min(v, 1024), std::max(1, n / 2)) : std::max(1, std::min(256, n / 4));
}

int trd_min_n() {
  static int v = [] {
    const char* s = getenv("VMC_JAX_AMD_TRIDIAG_MIN_N");
    return s ? atoi(s) : 128;
  }();
  return v;
}

bool trd_supported(int64_t n) { return n >= trd_min_n() && n >= 3 && n <= TRD_MAX_N; }

// Complex: 4 complex n-vectors in LDS and TRDC_R elements per thread.
constexpr int TRDC_MAX_N = 2496;   // four complex n-vectors in LDS
constexpr int TRDC_R_BIG = 7;      // large-n variant: v_k alone in LDS
// Up to the largest tridiagonal of the own divide and conquer.  At n = 2600
// the whole eigendecomposition takes 76 ms against rocSOLVER zheevd's 98 ms;
// at n = 4096 the 4096 launches (37 us each) make it 229 ms against 212 ms, so
// larger matrices stay with zheevd.
constexpr int TRDC_BIG_MAX_N = 3276;
static_assert(TRDC_BIG_MAX_N <= TRD_T * TRDC_R_BIG, "n-vector elements per thread");
bool trdc_supported(int64_t n) { return n >= trd_min_n() && n >= 3 && n <= TRDC_BIG_MAX_N; }

// Eigen-solve of the real symmetric n x n matrix S (destroyed) restricted to
// what the pseudo-inverse needs: ev = eigenvalues (ascending), c (length n,
// in/out) = Z^T Q^T c, with Z (n x n, row j = eigenvector j of T) left in Zbuf,
// and the reflectors in S for trd_apply.  Workspaces come from the caller.
int trd_reduce(double* S, int n, double* ev, double* c, double* work, hipStream_t st) {
  const int G = trd_groups(n);
  TrdArgs a;
  a.M = S;
  a.d = ev;
  a.e = work;
  a.tau = work + n;
  a.y = work + 2 * (size_t)n;
  a.u = work + 4 * (size_t)n;
  a.v0 = work + 6 * (size_t)n;
  a.pd = work + 7 * (size_t)n;
  a.c = c;
  a.n = n;
  a.G = G;
  const size_t lds = sizeof(double) * (4 * (size_t)n + 2 * TRD_W);
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)trd_step, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute(trd_step)");
  }
  {
    ScopedTimer tm("tridiag", st);
    for (int k = 0; k < n; ++k) {
      hipLaunchKernelGGL(trd_step, dim3((unsigned)(G + 1)), dim3(TRD_T), lds, st, a, k);
    }
    JV_LAUNCH_CHECK("trd_step");
  }
  return JV_OK;
}

// Eigenvalues (ev: the diagonal of T on entry) and eigenvectors of the
// tridiagonal T left by trd_reduce in work (rocSOLVER divide and conquer):
// column-major Z, so row j of the row-major buffer is eigenvector j.
int dc_tridiag_eig(int n, double* d, const double* e, double* Z, hipStream_t st);

bool dc_supported(int64_t n);

// Tridiagonal eigensolver: the divide and conquer of dc.hip (stream ordered),
// or rocSOLVER's dstedc (which destroys e and needs a host read of its
// convergence flag) with VMC_JAX_AMD_STEDC=rocsolver.  The variable is read on
// every call.
bool stedc_own(int n) {
  const char* s = getenv("VMC_JAX_AMD_STEDC");
  return !(s && std::string(s) == "rocsolver") && dc_supported(n);
}

int stedc(int n, double* ev, double* e, double* Z, rocblas_int* info, hipStream_t st) {
  if (stedc_own(n)) return dc_tridiag_eig(n, ev, e, Z, st);
  rocblas_handle h = blas_handle(st);
  if (!h) return fail(JV_E_SOLVER, "rocblas_create_handle failed");
  {
    ScopedTimer tm("stedc", st);
    rocblas_status s = rocsolver_dstedc(h, rocblas_evect_tridiagonal, n, ev, e, Z, n, info);
    if (s != rocblas_status_success) return fail(JV_E_SOLVER, "rocsolver_dstedc status " + std::to_string((int)s));
  }
  rocblas_int hinfo = 0;
  JV_HIP(hipMemcpyAsync(&hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, st));
  JV_HIP(hipStreamSynchronize(st));
  if (hinfo != 0) return fail(JV_E_SOLVER, "tridiagonal eigensolver did not converge (info " + std::to_string(hinfo) + ")");
  return JV_OK;
}

int trd_tridiag_eig(int n, double* ev, double* work, double* Zbuf, hipStream_t st) {
  const int G = trd_groups(n);
  return stedc(n, ev, work, Zbuf, (rocblas_int*)(work + 7 * (size_t)n + 2 * G + 8), st);
}

int trd_eig_solve(double* S, int n, double* ev, double* c, double* Zbuf, double* work, hipStream_t st) {
  int rc = trd_reduce(S, n, ev, c, work, st);
  return rc ? rc : trd_tridiag_eig(n, ev, work, Zbuf, st);
}

size_t trd_work_doubles(int n);

// Full eigendecomposition of the real symmetric S (row-major, destroyed):
// ev ascending, S overwritten with the eigenvectors as rows (the same
// convention as the rocSOLVER path).  The reflectors sit where LAPACK's dsytrd
// ('L') stores them (v_k(k+2:n) in column k, tau_k), so rocSOLVER's dormtr
// forms Q Z from the tridiagonal eigenvectors Z.
int trd_eigh_full(double* S, int n, double* ev, hipStream_t st) {
  double* Z = (double*)workspace(sizeof(double) * (size_t)n * n, 16);
  double* tw = (double*)workspace(sizeof(double) * trd_work_doubles(n), 17);
  if (!Z || !tw) return fail(JV_E_NOMEM, "eigh: workspace allocation failed");
  int rc = trd_eig_solve(S, n, ev, nullptr, Z, tw, st);
  if (rc) return rc;
  rocblas_handle h = blas_handle(st);
  if (!h) return fail(JV_E_SOLVER, "rocblas_create_handle failed");
  {
    ScopedTimer tm("ormtr", st);
    rocblas_status s = rocsolver_dormtr(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, n, n, S, n,
                                        tw + n, Z, n);
    if (s != rocblas_status_success) return fail(JV_E_SOLVER, "rocsolver_dormtr status " + std::to_string((int)s));
  }
  JV_HIP(hipMemcpyAsync(S, Z, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToDevice, st));
  return JV_OK;
}

// Full eigendecomposition of the complex Hermitian S (row-major, destroyed):
// ev ascending, row k of S <- conj(v_k) (the convention of the rocSOLVER path,
// which decomposes the column-major view conj(S)).  Householder reduction to a
// real tridiagonal T (trd_step_c), rocSOLVER dstedc for T, zunmtr for Q Z.
int trdc_eigh_full(double* S, int n, double* ev, hipStream_t st) {
  const int G = trd_groups(n);
  const size_t wd = 13 * (size_t)n + 2 * (size_t)G + 64;
  double* work = (double*)workspace(sizeof(double) * wd, 46);
  double* Z = (double*)workspace(sizeof(double) * (size_t)n * n, 16);
  double* Zc = (double*)workspace(sizeof(double) * 2 * (size_t)n * n, 47);
  if (!work || !Z || !Zc) return fail(JV_E_NOMEM, "eigh: workspace allocation failed");
  TrdcArgs a;
  a.M = S;
  a.d = ev;
  a.e = work;
  a.tau = work + n;
  a.y = work + 3 * (size_t)n;
  a.u = work + 7 * (size_t)n;
  a.v0 = work + 11 * (size_t)n;
  a.pd = work + 13 * (size_t)n;
  a.n = n;
  a.G = G;
  const bool big = n > TRDC_MAX_N || n > TRD_T * TRDC_R;
  const void* fn = big ? (const void*)trd_step_c<TRDC_R_BIG, true> : (const void*)trd_step_c<TRDC_R, false>;
  const size_t lds = sizeof(double) * ((big ? 2 : 8) * (size_t)n + 2 * TRD_W);
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute(trd_step_c)");
  }
  {
    ScopedTimer tm("tridiag", st);
    for (int k = 0; k < n; ++k) {
      if (big)
        hipLaunchKernelGGL((trd_step_c<TRDC_R_BIG, true>), dim3((unsigned)G), dim3(TRD_T), lds, st, a, k);
      else
        hipLaunchKernelGGL((trd_step_c<TRDC_R, false>), dim3((unsigned)G), dim3(TRD_T), lds, st, a, k);
    }
    JV_LAUNCH_CHECK("trd_step_c");
  }
  int rc = stedc(n, ev, a.e, Z, (rocblas_int*)(work + 13 * (size_t)n + 2 * G + 8), st);
  if (rc) return rc;
  rocblas_handle h = blas_handle(st);
  if (!h) return fail(JV_E_SOLVER, "rocblas_create_handle failed");
  const int64_t nn = (int64_t)n * n;
  hipLaunchKernelGGL(trdc_expand, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, Z, nn, Zc);
  JV_LAUNCH_CHECK("trdc_expand");
  {
    ScopedTimer tm("ormtr", st);
    rocblas_status s = rocsolver_zunmtr(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, n, n,
                                        (rocblas_double_complex*)S, n, (rocblas_double_complex*)a.tau,
                                        (rocblas_double_complex*)Zc, n);
    if (s != rocblas_status_success) return fail(JV_E_SOLVER, "rocsolver_zunmtr status " + std::to_string((int)s));
  }
  hipLaunchKernelGGL(trdc_conj, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, Zc, nn, S);
  JV_LAUNCH_CHECK("trdc_conj");
  return JV_OK;
}

int trd_wy_blocks(int n) { return (n - 1 + WY_B - 1) / WY_B; }

// x <- Q x with the reflectors in the rows of V and tau; Tbuf holds
// trd_wy_blocks(n) * WY_B^2 doubles.
int trd_wy_build(const double* V, const double* tau, int n, double* Tbuf, hipStream_t st) {
  const int nref = n - 1, nblk = trd_wy_blocks(n);
  hipLaunchKernelGGL(trd_wy_t, dim3((unsigned)nblk), dim3(256), 0, st, V, tau, n, nref, Tbuf);
  JV_LAUNCH_CHECK("trd_wy_t");
  return JV_OK;
}

// x <- Q x (transpose false) or Q^T x with the T factors of trd_wy_build.
int trd_wy_apply(const double* V, int n, double* x, const double* Tbuf, bool transpose, hipStream_t st) {
  if (transpose)
    hipLaunchKernelGGL(trd_apply_wy<true>, dim3(1), dim3(TRD_T), 0, st, V, Tbuf, n, n - 1, x);
  else
    hipLaunchKernelGGL(trd_apply_wy<false>, dim3(1), dim3(TRD_T), 0, st, V, Tbuf, n, n - 1, x);
  JV_LAUNCH_CHECK("trd_apply_wy");
  return JV_OK;
}

int trd_apply(const double* V, const double* tau, int n, double* x, double* Tbuf, hipStream_t st) {
  ScopedTimer tm("trd_apply", st);
  int rc = trd_wy_build(V, tau, n, Tbuf, st);
  return rc ? rc : trd_wy_apply(V, n, x, Tbuf, false, st);
}

size_t trd_work_doubles(int n) {
  return 7 * (size_t)n + 2 * (size_t)trd_groups(n) + 64 + (size_t)trd_wy_blocks(n) * WY_B * WY_B;
}

// Offset (doubles) of the WY T factors inside the trd workspace.
size_t trd_wy_offset(int n) { return 7 * (size_t)n + 2 * (size_t)trd_groups(n) + 64; }

}  // namespace jv
