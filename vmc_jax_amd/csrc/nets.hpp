// Network geometry and device routines shared by the forward / O_k, local
// estimator and Metropolis kernels.
#pragma once
#include "common.hpp"

namespace jv {

// Wave-local LDS ordering point: all LDS writes of this wave before the call are
// visible to all lanes of the wave after it.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------------ RBM
struct RbmGeom {
  int L, H;
  int cpx;          // complex parameters (CpxRBM)
  int64_t ob, ow, oa;  // flat offsets of hidden bias, kernel, visible bias (-1 if absent)
};

// theta_h = b_h + sum_i sigma_i W_ih for one hidden unit (rbm.py:32-38).
__device__ inline cd rbm_theta(const RbmGeom& g, const double* P, const int32_t* s, int h) {
  cd th = {0.0, 0.0};
  if (g.ob >= 0) {
    th.re = P[g.ob + h];
    if (g.cpx) th.im = P[g.ob + g.H + h];
  }
  const double* Wr = P + g.ow;
  const double* Wi = P + g.ow + (int64_t)g.L * g.H;
  for (int i = 0; i < g.L; ++i) {
    double sg = (double)(2 * s[i] - 1);
    th.re = fma(sg, Wr[(int64_t)i * g.H + h], th.re);
    if (g.cpx) th.im = fma(sg, Wi[(int64_t)i * g.H + h], th.im);
  }
  return th;
}

// ------------------------------------------------------------------ CNN
struct CnnGeom {
  int n;            // sites
  int Lx, Ly;       // extents (Ly = 1 for 1D)
  int Fx, Fy, FF;   // filter
  int nl;           // layers
  int C[JV_MAX_LAYERS];     // output channels
  int Cin[JV_MAX_LAYERS];   // input channels
  int aoff[JV_MAX_LAYERS];  // offset (doubles) of layer-l block inside an activation set
  int atot;                 // doubles per activation set (sum n*C[l])
  int64_t koff[JV_MAX_LAYERS];  // flat offsets of kernels
  int64_t boff[JV_MAX_LAYERS];  // flat offsets of biases (-1 if absent)
  int P;            // real parameter count
  double inv_nrm;   // 1 / sqrt(n * C[nl-1])  (cnn.py:169-171)
};

// Neighbour tables: nbr[p*FF + k] = input site read by output p at kernel
// offset k = (ky, kx) row-major: ((i+ky) mod Lx, (j+kx) mod Ly)  (cnn.py:160-167,
// circular pad (0, F-1) then VALID correlation).  inv[q*FF + k] = the output p
// with nbr[p*FF + k] == q.
__device__ inline void cnn_build_tables(const CnnGeom& g, short* nbr, short* inv, int tid, int nthreads) {
  for (int t = tid; t < g.n * g.FF; t += nthreads) {
    int p = t / g.FF, k = t % g.FF;
    int i = p / g.Ly, j = p % g.Ly;
    int ky = k / g.Fy, kx = k % g.Fy;
    int qi = (i + ky) % g.Lx, qj = (j + kx) % g.Ly;
    nbr[t] = (short)(qi * g.Ly + qj);
    int pi = (i - ky) % g.Lx; if (pi < 0) pi += g.Lx;
    int pj = (j - kx) % g.Ly; if (pj < 0) pj += g.Ly;
    inv[t] = (short)(pi * g.Ly + pj);
  }
}

// One output of layer l: z = b + sum_{k, ci} X[nbr(p,k), ci] K[k, ci, co].
// X is the layer input: x0 (one channel) for l == 0, else layer l-1 activations.
template <typename T>
__device__ inline T cnn_preact(const CnnGeom& g, const T* P, int l, const T* X, const short* nbr, int p, int co) {
  const int cin = g.Cin[l], cout = g.C[l];
  const T* K = P + g.koff[l];
  T z = g.boff[l] >= 0 ? P[g.boff[l] + co] : (T)0;
  for (int k = 0; k < g.FF; ++k) {
    const T* x = X + (int)nbr[p * g.FF + k] * cin;
    const T* w = K + (k * cin) * cout + co;
    for (int ci = 0; ci < cin; ++ci) z = fma(x[ci], w[ci * cout], z);
  }
  return z;
}

__device__ inline double act_tanh(double z) { return tanh(z); }
__device__ inline float act_tanh(float z) { return tanh_f32(z); }

// Same, reading the changed values of a local update: input position q is taken
// from Xnew when mask[q] is set, else from Xold.
__device__ inline double cnn_preact_masked(const CnnGeom& g, const double* P, int l,
                                          const double* Xold, const double* Xnew,
                                          const unsigned char* mask, const short* nbr, int p, int co) {
  const int cin = g.Cin[l], cout = g.C[l];
  const double* K = P + g.koff[l];
  double z = g.boff[l] >= 0 ? P[g.boff[l] + co] : 0.0;
  for (int k = 0; k < g.FF; ++k) {
    int q = nbr[p * g.FF + k];
    const double* x = (mask[q] ? Xnew : Xold) + q * cin;
    const double* w = K + (k * cin) * cout + co;
    for (int ci = 0; ci < cin; ++ci) z = fma(x[ci], w[ci * cout], z);
  }
  return z;
}

// Full forward of one configuration by a group of `nt` threads (thread index
// `tid`), activations into A (an activation set), optional tanh(z) into D.
// Caller synchronises the group before reading A; `sync` is a group barrier.
template <typename T, typename Sync>
__device__ inline void cnn_forward_group(const CnnGeom& g, const T* P, const T* x0, T* A, T* D, const short* nbr,
                                         int tid, int nt, Sync sync) {
  for (int l = 0; l < g.nl; ++l) {
    const T* X = l == 0 ? x0 : A + g.aoff[l - 1];
    T* Al = A + g.aoff[l];
    T* Dl = D ? D + g.aoff[l] : nullptr;
    int cout = g.C[l];
    for (int t = tid; t < g.n * cout; t += nt) {
      int p = t / cout, co = t % cout;
      T z = cnn_preact(g, P, l, X, nbr, p, co);
      Al[t] = log_cosh(z);
      if (Dl) Dl[t] = act_tanh(z);
    }
    sync();
  }
}

// Deterministic compaction of mask[0..n) into list (by one wave); returns count
// to every lane of the wave.
__device__ inline int wave_compact(const unsigned char* mask, int n, short* list, int lane) {
  int cnt = 0;
  for (int base = 0; base < n; base += WAVE) {
    int p = base + lane;
    bool on = p < n && mask[p];
    uint64_t bal = __ballot(on);
    if (on) {
      int pre = __popcll(bal & ((1ull << lane) - 1ull));
      list[cnt + pre] = (short)p;
    }
    cnt += __popcll(bal);
  }
  return cnt;
}

// Per-wave scratch of a receptive-field-limited update.
struct LocalScratch {
  unsigned char* mask;  // (nl + 1) * n bytes: mask[0] = changed sites, mask[l+1] = affected outputs of layer l
  short* list;          // n entries
  double* xnew;         // n doubles: new input values at changed sites
  double* Anew;         // one activation set: new values at affected outputs
};

// log psi(s') - log psi(s) for s' differing from s at `nchg` sites, computed by
// one wave from the base activations A of s (x0: base input).  Only outputs
// whose receptive field contains a changed site are recomputed; all other
// outputs are unchanged exactly.  Returns the difference to all lanes; new
// activations are left in sc.Anew (valid where the layer mask is set; the last
// layer only when store_last).
__device__ inline double cnn_local_delta(const CnnGeom& g, const double* P, const double* x0,
                                         const double* A, const short* nbr, int nchg,
                                         const int* sites, const int* vals, LocalScratch sc, int lane,
                                         bool store_last) {
  const int n = g.n;
  for (int p = lane; p < n; p += WAVE) sc.mask[p] = 0;
  wave_sync();
  if (lane < nchg) {
    sc.mask[sites[lane]] = 1;
    sc.xnew[sites[lane]] = (double)(2 * vals[lane] - 1);
  }
  wave_sync();
  double acc = 0.0;
  for (int l = 0; l < g.nl; ++l) {
    const unsigned char* min = sc.mask + l * n;
    unsigned char* mout = sc.mask + (l + 1) * n;
    for (int p = lane; p < n; p += WAVE) {
      unsigned char on = 0;
      for (int k = 0; k < g.FF; ++k) on |= min[nbr[p * g.FF + k]];
      mout[p] = on;
    }
    wave_sync();
    int cnt = wave_compact(mout, n, sc.list, lane);
    wave_sync();
    const double* Xold = l == 0 ? x0 : A + g.aoff[l - 1];
    const double* Xnew = l == 0 ? sc.xnew : sc.Anew + g.aoff[l - 1];
    const double* Aold = A + g.aoff[l];
    double* An = sc.Anew + g.aoff[l];
    int cout = g.C[l];
    bool last = l == g.nl - 1;
    for (int t = lane; t < cnt * cout; t += WAVE) {
      int p = sc.list[t / cout], co = t % cout;
      double z = cnn_preact_masked(g, P, l, Xold, Xnew, min, nbr, p, co);
      double a = log_cosh(z);
      if (!last || store_last) An[p * cout + co] = a;
      if (last) acc += a - Aold[p * cout + co];
    }
    wave_sync();
  }
  return wave_sum(acc) * g.inv_nrm;
}

}  // namespace jv

namespace jv {

// Weights are read through the constant address space: their indices are
// wave-uniform, so these become scalar loads served by the scalar cache.
typedef const __attribute__((address_space(4))) double cdbl;
__device__ inline const cdbl* as_const(const double* p) { return (const cdbl*)p; }

// ------------------------------------------------------------------ position-major CNN
// One lane computes all output channels of one position.  The weight index
// depends only on loop counters, so weight loads are wave-uniform (served from
// the scalar cache); the accumulation order per channel equals cnn_preact's, so
// both give bit-identical pre-activations.
// Activations of these kernels are channel-planar, X[c * n + pos], so lanes
// that handle consecutive positions read consecutive LDS words (no bank
// conflicts when every lane reads the same channel).
template <int CM>
__device__ inline void cnn_preact_all(const CnnGeom& g, const double* __restrict__ Pg, int l, const double* X,
                                      const short* nbr, int p, double* z) {
  const int cin = g.Cin[l], cout = g.C[l];
  const cdbl* K = as_const(Pg + g.koff[l]);
  const int64_t bo = g.boff[l];
  const cdbl* B = as_const(Pg + (bo >= 0 ? bo : 0));
#pragma unroll
  for (int c = 0; c < CM; ++c) z[c] = (c < cout && bo >= 0) ? B[c] : 0.0;
  for (int k = 0; k < g.FF; ++k) {
    const int q = nbr[p * g.FF + k];
    const cdbl* w = K + k * cin * cout;
#pragma unroll
    for (int ci = 0; ci < CM; ++ci) {
      if (ci < cin) {
        const double xv = X[ci * g.n + q];
#pragma unroll
        for (int c = 0; c < CM; ++c)
          if (c < cout) z[c] = fma(xv, w[ci * cout + c], z[c]);
      }
    }
  }
}

// Forward pass, position-parallel, channel-planar activations: A + aoff[l]
// holds layer l as [c][pos]; D (tanh z) likewise when non-null.
template <int CM, typename Sync>
__device__ inline void cnn_forward_pos(const CnnGeom& g, const double* __restrict__ Pg, const double* x0, double* A,
                                       double* D, const short* nbr, int tid, int nt, Sync sync) {
  double z[CM];
  for (int l = 0; l < g.nl; ++l) {
    const double* X = l == 0 ? x0 : A + g.aoff[l - 1];
    double* Al = A + g.aoff[l];
    double* Dl = D ? D + g.aoff[l] : nullptr;
    const int cout = g.C[l];
    for (int p = tid; p < g.n; p += nt) {
      cnn_preact_all<CM>(g, Pg, l, X, nbr, p, z);
#pragma unroll
      for (int c = 0; c < CM; ++c)
        if (c < cout) {
          Al[c * g.n + p] = log_cosh(z[c]);
          if (Dl) Dl[c * g.n + p] = tanh(z[c]);
        }
    }
    sync();
  }
}

// Receptive-field windows of a single flipped site s = (si, sj).  The outputs
// of layer l that depend on s are ((si - a) mod Lx, (sj - b) mod Ly) for
// 0 <= a < ex(l), 0 <= b < ey(l), ex(l) = 1 + (l+1)(Fx-1); window index a*ey + b.
// Valid while ex(nl-1) <= Lx and ey(nl-1) <= Ly (no wrap-around aliasing).
// LDS of the planar CNN forward / gradient kernel (cnn_fwd_grad_p): slots of
// kernel-gradient partial sums (items of cm values), one per thread or one per
// (k, ci) pair when the pairs outnumber the threads, plus the planes.
__host__ __device__ inline int cnn_p_items(const CnnGeom& g, int cm, int nt) { return nt > g.FF * cm ? nt : g.FF * cm; }
__host__ __device__ inline size_t cnn_p_lds(const CnnGeom& g, int cm, int nt, int tsize) {
  const size_t t = (size_t)g.n + 2 * (size_t)g.nl * cm * g.n + (size_t)cnn_p_items(g, cm, nt) * cm;
  return 8 * 8 + (size_t)tsize * t + 2 * 2 * (size_t)g.n * g.FF;
}

__device__ __host__ inline int win_ex(const CnnGeom& g, int l) { return l < 0 ? 1 : 1 + (l + 1) * (g.Fx - 1); }
__device__ __host__ inline int win_ey(const CnnGeom& g, int l) { return l < 0 ? 1 : 1 + (l + 1) * (g.Fy - 1); }
__device__ __host__ inline bool win_ok(const CnnGeom& g) {
  // the window kernels unroll filters up to 3x3
  return g.Fx <= 3 && g.Fy <= 3 && win_ex(g, g.nl - 1) <= g.Lx && win_ey(g, g.nl - 1) <= g.Ly;
}

// Pre-activations of window element (a, b) of layer l after flipping site
// (si, sj).  Inputs inside the layer-(l-1) window come from Wprev (planar,
// [ci][a*ey + b]); for l == 0 the only changed input is xflip at the site.
template <int CM>
__device__ inline void cnn_preact_win(const CnnGeom& g, const double* __restrict__ Pg, int l, const double* Xold,
                                      const double* Wprev, double xflip, int si, int sj, int a, int b, double* z) {
  const int cin = g.Cin[l], cout = g.C[l];
  const cdbl* K = as_const(Pg + g.koff[l]);
  const int64_t bo = g.boff[l];
  const cdbl* B = as_const(Pg + (bo >= 0 ? bo : 0));
  const int pex = win_ex(g, l - 1), pey = win_ey(g, l - 1);
#pragma unroll
  for (int c = 0; c < CM; ++c) z[c] = (c < cout && bo >= 0) ? B[c] : 0.0;
  for (int ky = 0; ky < g.Fx; ++ky) {
    const int aq = a - ky;
    int qi = si - aq;
    qi = qi < 0 ? qi + g.Lx : (qi >= g.Lx ? qi - g.Lx : qi);
    for (int kx = 0; kx < g.Fy; ++kx) {
      const int bq = b - kx;
      int qj = sj - bq;
      qj = qj < 0 ? qj + g.Ly : (qj >= g.Ly ? qj - g.Ly : qj);
      const bool in = aq >= 0 && aq < pex && bq >= 0 && bq < pey;
      const int k = ky * g.Fy + kx;
      const cdbl* w = K + k * cin * cout;
      if (l == 0) {
        const double xv = in ? xflip : Xold[qi * g.Ly + qj];
#pragma unroll
        for (int c = 0; c < CM; ++c)
          if (c < cout) z[c] = fma(xv, w[c], z[c]);
      } else {
        // planar: base [ci][pos] (stride n), window [ci][widx] (stride pex*pey)
        const double* x = in ? Wprev + (aq * pey + bq) : Xold + (qi * g.Ly + qj);
        const int xs = in ? pex * pey : g.n;
#pragma unroll
        for (int ci = 0; ci < CM; ++ci) {
          if (ci < cin) {
            const double xv = x[ci * xs];
#pragma unroll
            for (int c = 0; c < CM; ++c)
              if (c < cout) z[c] = fma(xv, w[ci * cout + c], z[c]);
          }
        }
      }
    }
  }
}

// Site index of window element (a, b) of a flip at (si, sj).
__device__ inline int win_pos(const CnnGeom& g, int si, int sj, int a, int b) {
  int pi = si - a;
  pi = pi < 0 ? pi + g.Lx : pi;
  int pj = sj - b;
  pj = pj < 0 ? pj + g.Ly : pj;
  return pi * g.Ly + pj;
}

}  // namespace jv
