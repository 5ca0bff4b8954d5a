// Divide and conquer for the eigenvalues and eigenvectors of a real symmetric
// tridiagonal matrix T (the task of LAPACK's dstedc, here done on the device
// without host synchronisation).
//
// T is split recursively at the middle: T = diag(T1', T2') + |b| v v^T with
// v = (e_last, sign(b) e_first) and the two diagonal entries next to the split
// reduced by |b|.  Leaves (<= DC_LEAF rows) are solved by implicit QL with
// Wilkinson shifts, one wavefront each.  A merge of two solved halves is the
// rank-one problem D + rho z z^T (D their eigenvalues, z the last row of the
// left and the first row of the right eigenvectors / sqrt2, rho = 2 |b|):
//  * deflation (Cuppen / Dongarra-Sorensen): rho |z_i| <= tol, or a Givens
//    rotation of two columns with close eigenvalues that zeroes one z entry;
//  * the secular equation 1 + rho sum z_i^2 / (d_i - lambda) = 0 for each
//    remaining root, by bisection in the coordinate shifted to the nearer pole
//    (so d_i - lambda_j keeps full relative accuracy), one wavefront per root;
//  * eigenvectors from recomputed weights z^_i (Gu and Eisenstat), which keeps
//    them orthogonal: z^_i^2 = (lambda_i - d_i) / rho * prod_{j != i}
//    (lambda_j - d_i) / (d_j - d_i), u_j = (z^_i / (d_i - lambda_j))_i / norm;
//  * the merged eigenvectors [Q1 0; 0 Q2] P G U, one batched GEMM per level.
// Merges run level by level, all merges of a level in the same launches; the
// eigenvector matrix is updated in place.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "common.hpp"

namespace jv {


namespace {

constexpr int DC_LEAF = 32;
// the merge preparation keeps 20 bytes per row of the merged node in LDS
constexpr int DC_MAX_N = 65536 / 20;
// flags kept in the merged column table next to the column (< 2^29)
constexpr int DC_NEG = 1 << 30, DC_ROT = 1 << 29, DC_COL = DC_ROT - 1;

// Node i at depth k of the tree covers rows [i n / 2^k, (i + 1) n / 2^k)
// (floors) and splits at (2 i + 1) n / 2^(k + 1): children are nodes 2i and
// 2i + 1 at depth k + 1, and every node at depth D has at most ceil(n / 2^D)
// rows, so all leaves sit at one depth.
struct DcNode {
  int start, n1, n2;
};

__host__ __device__ inline DcNode dc_node(int n, int k, int i) {
  const int64_t a = ((int64_t)i * n) >> k, m = ((int64_t)(2 * i + 1) * n) >> (k + 1),
                z = ((int64_t)(i + 1) * n) >> k;
  return DcNode{(int)a, (int)(m - a), (int)(z - m)};
}

// leaf depth: the smallest D with ceil(n / 2^D) <= leaf, leaf = DC_LEAF, or
// DC_LEAF / 2 below n = 512 (the sequential QL chain of a leaf grows as its
// size squared; at small n an extra merge level costs less than it saves)
inline int dc_depth(int n) {
  const int leaf = n < 512 ? DC_LEAF / 2 : DC_LEAF;
  int D = 0;
  while (((int64_t)n + (int64_t(1) << D) - 1) >> D > leaf) ++D;
  return D;
}

// per-merge scratch, node-local at offset start of each n-array
struct DcBufs {
  double* dd;    // non-deflated poles, ascending [k]
  double* zz;    // their z [k]
  double* dD;    // deflated eigenvalues [n - k]
  double* lam;   // roots [k]
  double* tau;   // root - pole[org] [k]
  double* zh;    // Gu-Eisenstat weights [k]
  double* rc;    // rotation c, s pairs [2 * nrot]
  double* rho;   // per node: rho (at start)
  int* colN;     // global column of each non-deflated entry [k]
  int* colD;     // global column of each deflated entry [n - k]
  int* org;      // pole index of each root [k]
  int* rot;      // rotation column pairs [2 * nrot]
  int* cnt;      // per node at start: k, nrot
  int* rank;     // final position of each output entry [n]
};

// Sort key of an eigenvalue: NaN orders as +inf (ties by position), so every
// ranking below is a permutation whatever the values, and the merge positions
// derived from it stay distinct.
__device__ inline double dc_key(double v) { return v != v ? __builtin_inf() : v; }

// ---------------------------------------------------------------- leaves
// One wave per leaf: implicit QL with Wilkinson shifts (the scalar recurrence
// runs redundantly in every lane, lane r owns row r of the leaf's Z).
// dg and w may alias: every lane reads its entry of dg before the first
// barrier, w is written after the last one.
__global__ void __launch_bounds__(64) dc_leaves(int D, const double* dg,
                                               const double* __restrict__ eg, int n, double* __restrict__ Q,
                                               double* w) {
  chain_priority();
  __shared__ double d[DC_LEAF], ee[DC_LEAF], Zs[DC_LEAF][DC_LEAF + 1];
  __shared__ int rk[DC_LEAF];
  const int s0 = (int)(((int64_t)blockIdx.x * n) >> D);
  const int m = (int)((((int64_t)blockIdx.x + 1) * n) >> D) - s0, lane = threadIdx.x;
  if (lane < m) {
    d[lane] = dg[s0 + lane];
    ee[lane] = lane < m - 1 ? eg[s0 + lane] : 0.0;
    for (int c = 0; c < m; ++c) Zs[lane][c] = c == lane ? 1.0 : 0.0;
  }
  __syncthreads();
  const double eps = 2.220446049250313e-16;
  for (int l = 0; l < m; ++l) {
    for (int iter = 0; iter < 100; ++iter) {
      int mm = l;
      for (; mm < m - 1; ++mm) {
        const double dd = fabs(d[mm]) + fabs(d[mm + 1]);
        if (fabs(ee[mm]) <= eps * dd) break;
      }
      if (mm == l) break;
      double g = (d[l + 1] - d[l]) / (2.0 * ee[l]);
      double r = hypot(g, 1.0);
      g = d[mm] - d[l] + ee[l] / (g + copysign(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      bool under = false;
      for (int i = mm - 1; i >= l; --i) {
        const double f = s * ee[i], b = c * ee[i];
        r = hypot(f, g);
        __syncthreads();
        if (lane == 0) ee[i + 1] = r;
        __syncthreads();
        if (r == 0.0) {
          __syncthreads();
          if (lane == 0) {
            d[i + 1] -= p;
            ee[mm] = 0.0;
          }
          __syncthreads();
          under = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        const double di1 = g + p;
        g = c * r - b;
        __syncthreads();
        if (lane == 0) d[i + 1] = di1;
        if (lane < m) {
          const double zi = Zs[lane][i], zi1 = Zs[lane][i + 1];
          Zs[lane][i + 1] = s * zi + c * zi1;
          Zs[lane][i] = c * zi - s * zi1;
        }
        __syncthreads();
      }
      if (under) continue;
      __syncthreads();
      if (lane == 0) {
        d[l] -= p;
        ee[l] = g;
        ee[mm] = 0.0;
      }
      __syncthreads();
    }
  }
  // ascending order: rank of each eigenvalue (ties by index)
  if (lane < m) {
    int r = 0;
    const double x = d[lane], kx = dc_key(x);
    for (int j = 0; j < m; ++j) {
      const double kj = dc_key(d[j]);
      r += (kj < kx || (kj == kx && j < lane)) ? 1 : 0;
    }
    rk[lane] = r;
    w[s0 + r] = x;
  }
  __syncthreads();
  if (lane < m)
    for (int c = 0; c < m; ++c) Q[(size_t)(s0 + rk[c]) * n + s0 + lane] = Zs[lane][c];
}

__global__ void dc_zero(double* __restrict__ p, int64_t len) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < len) p[t] = 0.0;
}

// d[s-1] -= |e[s-1]|, d[s] -= |e[s-1]| at every split s: the inner leaf
// boundaries j n / 2^D, 0 < j < 2^D
__global__ void dc_split_adjust(int n, int D, double* __restrict__ d, const double* __restrict__ e) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (j >= (1 << D)) return;
  const int s = (int)(((int64_t)j * n) >> D);
  const double b = fabs(e[s - 1]);
  d[s - 1] -= b;
  d[s] -= b;
}

// ---------------------------------------------------------------- merge: preparation
__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// One wave per node.  The lanes build z and merge the children's ascending
// eigenvalue lists into LDS (each entry's merged position by binary search in
// the other list; ties keep the left child first); lane 0 then runs the
// deflation pass over LDS and streams its results out.  Dynamic LDS: 20 bytes
// per row of the largest node.
__global__ void __launch_bounds__(64) dc_prep(int lvk, const double* __restrict__ e, int n,
                                             const double* __restrict__ Q, const double* __restrict__ w, DcBufs b,
                                             double* __restrict__ zt) {
  chain_priority();
  extern __shared__ double lds[];
  const DcNode nd = dc_node(n, lvk, blockIdx.x);
  const int s0 = nd.start, n1 = nd.n1, n2 = nd.n2, nn = nd.n1 + nd.n2, lane = threadIdx.x;
  double* ldt = lds;          // merged eigenvalues
  double* lzs = lds + nn;     // merged z
  int* lct = (int*)(lds + 2 * nn);  // merged global columns
  const double beta = e[s0 + n1 - 1];
  const double sg = beta >= 0.0 ? 1.0 : -1.0;
  const double h = 0.70710678118654752440;
  // z in the children's order: last row of the left block, sign * first row of the right block
  for (int c = lane; c < nn; c += WAVE) {
    const int row = c < n1 ? s0 + n1 - 1 : s0 + n1;
    zt[s0 + c] = (c < n1 ? h : sg * h) * Q[(size_t)(s0 + c) * n + row];
  }
  // every entry of the merged table holds a column of this node even if the
  // positions below were not a permutation
  for (int c = lane; c < nn; c += WAVE) {
    ldt[c] = 0.0;
    lzs[c] = 0.0;
    lct[c] = s0 + c;
  }
  __syncthreads();
  const double* wl = w + s0;
  const double* wr = w + s0 + n1;
  double dmax = 0.0, zmax = 0.0;
  for (int c = lane; c < nn; c += WAVE) {
    const double x = w[s0 + c], kx = dc_key(x);
    int pos;
    if (c < n1) {  // left entry: right entries strictly below x precede it
      int lo = 0, hi = n2;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (dc_key(wr[mid]) < kx) lo = mid + 1;
        else hi = mid;
      }
      pos = c + lo;
    } else {  // right entry: left entries <= x precede it
      int lo = 0, hi = n1;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (dc_key(wl[mid]) <= kx) lo = mid + 1;
        else hi = mid;
      }
      pos = (c - n1) + lo;
    }
    const double z = zt[s0 + c];
    ldt[pos] = x;
    lzs[pos] = z;
    lct[pos] = s0 + c;
    dmax = fmax(dmax, fabs(x));
    zmax = fmax(zmax, fabs(z));
  }
  dmax = wave_max(dmax);
  zmax = wave_max(zmax);
  __syncthreads();
  const double rho = 2.0 * fabs(beta);
  const double tol = 8.0 * 2.220446049250313e-16 * fmax(dmax, zmax);
  // Flags in the column table, computed by all lanes over contiguous chunks:
  // DC_NEG for a negligible coupling, DC_ROT where the pair (previous
  // non-negligible entry, t) passes the rotation test with the values before
  // any rotation.  The sequential pass below reuses DC_ROT unless the
  // previous entry was itself just rotated.
  {
    const int C = (nn + WAVE - 1) / WAVE;
    const int c0 = min(nn, lane * C), c1 = min(nn, c0 + C);
    int last = -1;
    for (int t = c0; t < c1; ++t) {
      if (rho * fabs(lzs[t]) <= tol) lct[t] |= DC_NEG;
      else last = t;
    }
    int incl = last;
    for (int o = 1; o < WAVE; o <<= 1) {
      const int v = __shfl_up(incl, o, 64);
      if (lane >= o) incl = max(incl, v);
    }
    int q = __shfl_up(incl, 1, 64);
    if (lane == 0) q = -1;
    for (int t = c0; t < c1; ++t) {
      if (lct[t] & DC_NEG) continue;
      if (q >= 0) {
        const double zp = lzs[q], zi = lzs[t];
        const double tz = hypot(zp, zi);
        const double c = zi / tz, s = -zp / tz;
        if (fabs((ldt[t] - ldt[q]) * c * s) <= tol) lct[t] |= DC_ROT;
      }
      q = t;
    }
  }
  __syncthreads();
  if (lane != 0) return;
  // Deflation in ascending order.  The kept poles stay ascending: a rotation
  // replaces the last kept pole d_p by d_t' in [d_p, d_t], and kept neighbours
  // differ by more than 2 tol, far above the rounding of d_t'.
  int k = 0, ndf = 0, nrot = 0, prev = -1;
  bool prev_rotated = false;
  for (int t = 0; t < nn; ++t) {
    const int f = lct[t];
    const int col = f & DC_COL;
    if (f & DC_NEG) {  // negligible coupling
      b.dD[s0 + ndf] = ldt[t];
      b.colD[s0 + ndf] = col;
      ++ndf;
      continue;
    }
    bool rotated = false;
    if (prev >= 0 && (prev_rotated || (f & DC_ROT))) {
      const double zp = lzs[prev], zi = lzs[t];
      const double tz = hypot(zp, zi);
      const double c = zi / tz, s = -zp / tz;
      const double dp = ldt[prev], di = ldt[t];
      if (fabs((di - dp) * c * s) <= tol) {
        // rotate the pair so that z[prev] vanishes: prev deflates, t takes its place
        const int pcol = lct[prev] & DC_COL;
        b.rot[2 * (s0 + nrot)] = pcol;
        b.rot[2 * (s0 + nrot) + 1] = col;
        b.rc[2 * (s0 + nrot)] = c;
        b.rc[2 * (s0 + nrot) + 1] = s;
        ++nrot;
        lzs[t] = tz;
        ldt[t] = dp * s * s + di * c * c;
        b.dD[s0 + ndf] = dp * c * c + di * s * s;
        b.colD[s0 + ndf] = pcol;
        ++ndf;
        --k;
        rotated = true;
      }
    }
    b.dd[s0 + k] = ldt[t];
    b.zz[s0 + k] = lzs[t];
    b.colN[s0 + k] = col;
    ++k;
    prev = t;
    prev_rotated = rotated;
  }
  b.cnt[2 * s0] = k;
  b.cnt[2 * s0 + 1] = nrot;
  b.rho[s0] = rho;
}

// Givens rotations of eigenvector columns, in deflation order; thread = row.
__global__ void dc_rotate(int lvk, int n, const DcBufs b, double* __restrict__ Q) {
  chain_priority();
  const DcNode nd = dc_node(n, lvk, blockIdx.y);
  const int nn = nd.n1 + nd.n2, s0 = nd.start;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nn) return;
  const int nrot = b.cnt[2 * s0 + 1];
  const int row = s0 + r;
  for (int q = 0; q < nrot; ++q) {
    const int ca = b.rot[2 * (s0 + q)], cb = b.rot[2 * (s0 + q) + 1];
    const double c = b.rc[2 * (s0 + q)], s = b.rc[2 * (s0 + q) + 1];
    double* qa = Q + (size_t)ca * n + row;
    double* qb = Q + (size_t)cb * n + row;
    const double a = *qa, bb = *qb;
    *qa = c * a + s * bb;
    *qb = -s * a + c * bb;
  }
}

// ---------------------------------------------------------------- secular equation
// One wave per root j of node y: bisection on tau in the coordinate of the
// nearer pole, f(tau) = 1 + rho sum z_i^2 / ((d_i - d_org) - tau).
__global__ void __launch_bounds__(256) dc_secular(int lvk, int n, DcBufs b) {
  chain_priority();
  extern __shared__ double sec[];  // poles [k], squared couplings [k]
  const DcNode nd = dc_node(n, lvk, blockIdx.y);
  const int s0 = nd.start;
  const int k = b.cnt[2 * s0];
  if ((int)(blockIdx.x * (blockDim.x / WAVE)) >= k) return;  // uniform over the workgroup
  double* dd = sec;
  double* z2 = sec + k;
  for (int t = threadIdx.x; t < k; t += blockDim.x) {
    dd[t] = b.dd[s0 + t];
    const double z = b.zz[s0 + t];
    z2[t] = z * z;
  }
  __syncthreads();
  const int j = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
  const int lane = threadIdx.x % WAVE;
  if (j >= k) return;
  const double rho = b.rho[s0];
  auto fsum = [&](int org, double t) {
    double a = 0.0;
    const double dorg = dd[org];
    for (int i = lane; i < k; i += WAVE) a += z2[i] / ((dd[i] - dorg) - t);
    return 1.0 + rho * wave_sum(a);
  };
  const double lo = dd[j];
  double hi;
  if (j < k - 1) {
    hi = dd[j + 1];
  } else {
    double a = 0.0;
    for (int i = lane; i < k; i += WAVE) a += z2[i];
    hi = lo + rho * wave_sum(a);
  }
  const double mid = 0.5 * (lo + hi);
  int org;
  double ta, tb;
  if (j < k - 1 && fsum(j, mid - lo) < 0.0) {  // root right of the midpoint: origin at the upper pole
    org = j + 1;
    ta = mid - hi;
    tb = 0.0;
  } else {
    org = j;
    ta = 0.0;
    tb = j < k - 1 ? mid - lo : hi - lo;
  }
  for (int it = 0; it < 300; ++it) {
    const double t = 0.5 * (ta + tb);
    if (t == ta || t == tb) break;
    if (fsum(org, t) < 0.0) ta = t;
    else tb = t;
  }
  if (lane == 0) {
    const double t = 0.5 * (ta + tb);
    b.org[s0 + j] = org;
    b.tau[s0 + j] = t;
    b.lam[s0 + j] = dd[org] + t;
  }
}

// Gu-Eisenstat weights; thread = pole i.  The roots' data (pole, origin
// pole, tau) pass through LDS in chunks of 1024, in root order.
__global__ void __launch_bounds__(256) dc_zhat(int lvk, int n, DcBufs b) {
  chain_priority();
  __shared__ double sd[1024], so[1024], st[1024];
  const DcNode nd = dc_node(n, lvk, blockIdx.y);
  const int s0 = nd.start;
  const int k = b.cnt[2 * s0];
  if ((int)(blockIdx.x * blockDim.x) >= k) return;  // uniform over the workgroup
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < k;
  const double rho = b.rho[s0];
  const double* dd = b.dd + s0;
  const double di = live ? dd[i] : 0.0;
  double v = live ? -((di - dd[b.org[s0 + i]]) - b.tau[s0 + i]) / rho : 0.0;  // (lambda_i - d_i) / rho
  for (int c0 = 0; c0 < k; c0 += 1024) {
    const int m = min(1024, k - c0);
    for (int t = threadIdx.x; t < m; t += blockDim.x) {
      sd[t] = dd[c0 + t];
      so[t] = dd[b.org[s0 + c0 + t]];
      st[t] = b.tau[s0 + c0 + t];
    }
    __syncthreads();
    if (live)
      for (int u = 0; u < m; ++u)
        if (c0 + u != i) v *= -((di - so[u]) - st[u]) / (sd[u] - di);
    __syncthreads();
  }
  if (live) b.zh[s0 + i] = copysign(sqrt(fabs(v)), b.zz[s0 + i]);
}

// Eigenvector j of the rank-one problem into column j of the node's U block
// (rows / columns past k zero); the node's G block gets the non-deflated
// columns of Q in pole order, then the deflated ones.  One wave per column.
__global__ void __launch_bounds__(256) dc_vectors(int lvk, int n, const DcBufs b,
                                                 const double* __restrict__ Q, double* __restrict__ U,
                                                 double* __restrict__ G) {
  chain_priority();
  const DcNode nd = dc_node(n, lvk, blockIdx.y);
  const int s0 = nd.start, nn = nd.n1 + nd.n2;
  const int j = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
  const int lane = threadIdx.x % WAVE;
  if (j >= nn) return;
  const int k = b.cnt[2 * s0];
  double* Uc = U + (size_t)(s0 + j) * n + s0;
  if (j < k) {
    const double* dd = b.dd + s0;
    const double dorg = dd[b.org[s0 + j]], tj = b.tau[s0 + j];
    double a = 0.0;
    for (int i = lane; i < k; i += WAVE) {
      const double u = b.zh[s0 + i] / ((dd[i] - dorg) - tj);
      a += u * u;
    }
    const double inv = 1.0 / sqrt(wave_sum(a));
    for (int i = lane; i < nn; i += WAVE)
      Uc[i] = i < k ? b.zh[s0 + i] / ((dd[i] - dorg) - tj) * inv : 0.0;
  } else {
    for (int i = lane; i < nn; i += WAVE) Uc[i] = 0.0;
  }
  const int src = j < k ? b.colN[s0 + j] : b.colD[s0 + (j - k)];
  const double* qs = Q + (size_t)src * n + s0;
  double* gc = G + (size_t)(s0 + j) * n + s0;
  for (int i = lane; i < nn; i += WAVE) gc[i] = qs[i];
}

// Merged eigenvectors of every node of a level in one launch (blockIdx.y =
// node): C(r, j) = sum_{i < k} G(r, i) U(i, j) for the k non-deflated columns
// (column-major blocks at offset start (n + 1) inside the n x n buffers;
// columns j >= k are the deflated ones, taken from G as they are).  Tiles of
// C^T (64 j x 64 r) with FP64 MFMA 16x16x4 over K chunks of 32, four waves
// of 32 x 32 (A operand A[row = lane % 16][k = lane / 16], B operand
// B[k = lane / 16][col = lane % 16], result element q of lane l at
// C[4 q + l / 16][l % 16]); tiles and chunks beyond k are skipped.
__global__ void __launch_bounds__(256) dc_merge_gemm(int lvk, int n, DcBufs b, const double* __restrict__ G,
                                                    const double* __restrict__ U, double* __restrict__ C) {
  chain_priority();
  const DcNode nd = dc_node(n, lvk, blockIdx.y);
  const int s0 = nd.start, m = nd.n1 + nd.n2;
  const int k = b.cnt[2 * s0];
  const int T = (m + 63) / 64;
  const int j0 = 64 * (blockIdx.x / T), r0 = 64 * (blockIdx.x % T);
  if (blockIdx.x >= (unsigned)(T * T) || j0 >= k) return;  // uniform over the workgroup
  __shared__ double As[32][65], Bs[32][65];
  const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
  const int wm = (wv / 2) * 32, wn = (wv % 2) * 32;
  typedef double v4d __attribute__((ext_vector_type(4)));
  v4d acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int c = 0; c < 2; ++c) acc[a][c] = v4d{0, 0, 0, 0};
  const size_t base = (size_t)s0 * n + s0;
  for (int i0 = 0; i0 < k; i0 += 32) {
    double va[8], vb[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = tid + 256 * u;
      const int jj = e / 32, kk = e % 32;  // As[kk][jj] = U(i0 + kk, j0 + jj)
      va[u] = U[base + (size_t)min(j0 + jj, m - 1) * n + min(i0 + kk, m - 1)];
      const int kb = e / 64, rr = e % 64;  // Bs[kb][rr] = G(r0 + rr, i0 + kb)
      vb[u] = G[base + (size_t)min(i0 + kb, m - 1) * n + min(r0 + rr, m - 1)];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = tid + 256 * u;
      const int jj = e / 32, kk = e % 32;
      As[kk][jj] = (j0 + jj < k && i0 + kk < k) ? va[u] : 0.0;
      const int kb = e / 64, rr = e % 64;
      Bs[kb][rr] = (i0 + kb < k && r0 + rr < m) ? vb[u] : 0.0;
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
      for (int q = 0; q < 4; ++q) {
        const int j = j0 + wm + 16 * a + 4 * q + lane / 16;
        const int r = r0 + wn + 16 * c + lane % 16;
        if (j < k && r < m) C[base + (size_t)j * n + r] = acc[a][c][q];
      }
}

// Final order of a node: the roots (entries t < k) and deflated values
// (entries t >= k) sorted by (NaN-safe key, entry index) with a bitonic
// network in LDS, padded to a power of two with +inf keys and indices >= nn.
// Dynamic LDS: 12 bytes per padded entry.
__global__ void __launch_bounds__(256) dc_rank(int lvk, int n, DcBufs b, double* __restrict__ w) {
  chain_priority();
  extern __shared__ double lv[];
  const DcNode nd = dc_node(n, lvk, blockIdx.x);
  const int s0 = nd.start, nn = nd.n1 + nd.n2;
  const int k = b.cnt[2 * s0];
  int P2 = 1;
  while (P2 < nn) P2 <<= 1;
  double* key = lv;
  int* idx = (int*)(lv + P2);
  for (int t = threadIdx.x; t < P2; t += blockDim.x) {
    key[t] = t < nn ? dc_key(t < k ? b.lam[s0 + t] : b.dD[s0 + (t - k)]) : __builtin_inf();
    idx[t] = t;
  }
  __syncthreads();
  for (int size = 2; size <= P2; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < (P2 >> 1); t += blockDim.x) {
        const int i = 2 * stride * (t / stride) + (t % stride), j = i + stride;
        const bool asc = (i & size) == 0;
        const double ki = key[i], kj = key[j];
        const int ii = idx[i], ij = idx[j];
        const bool gt = ki > kj || (ki == kj && ii > ij);
        if (gt == asc) {
          key[i] = kj;
          key[j] = ki;
          idx[i] = ij;
          idx[j] = ii;
        }
      }
      __syncthreads();
    }
  for (int r = threadIdx.x; r < nn; r += blockDim.x) {
    const int t = idx[r];
    b.rank[s0 + t] = r;
    w[s0 + r] = t < k ? b.lam[s0 + t] : b.dD[s0 + (t - k)];
  }
}

// Q columns of node y <- entry t (GEMM output column t < k, or G column t
// >= k) at position rank[t]; thread = row.
__global__ void dc_scatter(int lvk, int n, const DcBufs b, const double* __restrict__ C,
                           const double* __restrict__ G, double* __restrict__ Q) {
  chain_priority();
  const DcNode nd = dc_node(n, lvk, blockIdx.z);
  const int s0 = nd.start, nn = nd.n1 + nd.n2;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = blockIdx.y;
  if (r >= nn || t >= nn) return;
  const int k = b.cnt[2 * s0];
  const double* src = (t < k ? C : G) + (size_t)(s0 + t) * n + s0;
  Q[(size_t)(s0 + b.rank[s0 + t]) * n + s0 + r] = src[r];
}

}  // namespace

bool dc_supported(int64_t n) { return n >= 1 && n <= DC_MAX_N; }

// Eigenvalues (d: diagonal on entry, eigenvalues ascending on exit) and
// eigenvectors (Z: column-major n x n, column j = eigenvector j) of the
// symmetric tridiagonal matrix with diagonal d and off-diagonal e[n-1] (e is
// not modified).  Stream ordered, no host synchronisation once the tree of a
// size has been uploaded.  Workspace slots 90-93.
int dc_tridiag_eig(int n, double* d, const double* e, double* Z, hipStream_t st) {
  if (n < 1) return JV_OK;
  if (!dc_supported(n)) return fail(JV_E_ARG, "tridiagonal eigensolver: n = " + std::to_string(n) + " exceeds " +
                                                 std::to_string(DC_MAX_N));
  const int D = dc_depth(n);
  const size_t nn = (size_t)n * n;
  // G (gathered eigenvectors), U (rank-one eigenvectors), C (their product)
  double* big = D > 0 ? (double*)workspace(sizeof(double) * 3 * nn, 90) : nullptr;
  // doubles: dd, zz, dD, lam, tau, zh, rho, z, rotations (2n)
  double* dv = (double*)workspace(sizeof(double) * 10 * (size_t)n + 64, 91);
  // ints: colN, colD, org, rank, rotation pairs (2n), counts (2n)
  int* iv = (int*)workspace(sizeof(int) * 8 * (size_t)n + 64, 92);
  if ((D > 0 && !big) || !dv || !iv) return fail(JV_E_NOMEM, "tridiagonal eigensolver: workspace allocation failed");
  double* G = big;
  double* U = big + nn;
  double* C = big + 2 * nn;
  DcBufs b;
  b.dd = dv;
  b.zz = dv + n;
  b.dD = dv + 2 * (size_t)n;
  b.lam = dv + 3 * (size_t)n;
  b.tau = dv + 4 * (size_t)n;
  b.zh = dv + 5 * (size_t)n;
  b.rho = dv + 6 * (size_t)n;
  double* zt = dv + 7 * (size_t)n;  // z in the children's order
  b.rc = dv + 8 * (size_t)n;        // 2n
  b.colN = iv;
  b.colD = iv + n;
  b.org = iv + 2 * (size_t)n;
  b.rank = iv + 3 * (size_t)n;
  b.rot = iv + 4 * (size_t)n;       // 2n
  b.cnt = iv + 6 * (size_t)n;       // 2n
  ScopedTimer tm("stedc", st);
  if (D > 0) {
    const int ns = (1 << D) - 1;
    hipLaunchKernelGGL(dc_split_adjust, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, st, n, D, d, e);
    JV_LAUNCH_CHECK("dc_split_adjust");
  }
  hipLaunchKernelGGL(dc_zero, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, Z, (int64_t)nn);
  JV_LAUNCH_CHECK("dc_zero");
  hipLaunchKernelGGL(dc_leaves, dim3(1u << D), dim3(64), 0, st, D, d, e, n, Z, d);
  JV_LAUNCH_CHECK("dc_leaves");
  for (int k = D - 1; k >= 0; --k) {
    const unsigned nodes = 1u << k;
    const int mx = (int)(((int64_t)n + (int64_t(1) << k) - 1) >> k);  // largest node at depth k
    hipLaunchKernelGGL(dc_prep, dim3(nodes), dim3(64), (size_t)mx * 20, st, k, e, n, Z, d, b, zt);
    JV_LAUNCH_CHECK("dc_prep");
    hipLaunchKernelGGL(dc_rotate, dim3((unsigned)((mx + 255) / 256), nodes), dim3(256), 0, st, k, n, b, Z);
    JV_LAUNCH_CHECK("dc_rotate");
    hipLaunchKernelGGL(dc_secular, dim3((unsigned)((mx + 3) / 4), nodes), dim3(256), (size_t)mx * 16, st, k, n, b);
    JV_LAUNCH_CHECK("dc_secular");
    hipLaunchKernelGGL(dc_zhat, dim3((unsigned)((mx + 255) / 256), nodes), dim3(256), 0, st, k, n, b);
    JV_LAUNCH_CHECK("dc_zhat");
    hipLaunchKernelGGL(dc_vectors, dim3((unsigned)((mx + 3) / 4), nodes), dim3(256), 0, st, k, n, b, Z, U, G);
    JV_LAUNCH_CHECK("dc_vectors");
    // C block = G block U block for every node of the level
    {
      const unsigned T = (unsigned)((mx + 63) / 64);
      hipLaunchKernelGGL(dc_merge_gemm, dim3(T * T, nodes), dim3(256), 0, st, k, n, b, G, U, C);
      JV_LAUNCH_CHECK("dc_merge_gemm");
    }
    int p2 = 1;
    while (p2 < mx) p2 <<= 1;
    hipLaunchKernelGGL(dc_rank, dim3(nodes), dim3(256), (size_t)p2 * 12, st, k, n, b, d);
    JV_LAUNCH_CHECK("dc_rank");
    hipLaunchKernelGGL(dc_scatter, dim3((unsigned)((mx + 255) / 256), (unsigned)mx, nodes), dim3(256), 0, st, k, n,
                       b, C, G, Z);
    JV_LAUNCH_CHECK("dc_scatter");
  }
  return JV_OK;
}

}  // namespace jv

using namespace jv;

extern "C" int jv_stedc(int64_t n, double* d, const double* e, double* z, void* stream) {
  if (n < 0 || (n > 0 && (!d || !z)) || (n > 1 && !e)) return fail(JV_E_ARG, "jv_stedc: bad arguments");
  if (n == 0) return JV_OK;
  if (!dc_supported(n)) return fail(JV_E_ARG, "jv_stedc: n above " + std::to_string(DC_MAX_N));
  return dc_tridiag_eig((int)n, d, e, z, (hipStream_t)stream);
}
