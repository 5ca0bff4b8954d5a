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
__device__ inline double cnn_preact(const CnnGeom& g, const double* P, int l, const double* X,
                                   const short* nbr, int p, int co) {
  const int cin = g.Cin[l], cout = g.C[l];
  const double* K = P + g.koff[l];
  double z = g.boff[l] >= 0 ? P[g.boff[l] + co] : 0.0;
  for (int k = 0; k < g.FF; ++k) {
    const double* x = X + (int)nbr[p * g.FF + k] * cin;
    const double* w = K + (k * cin) * cout + co;
    for (int ci = 0; ci < cin; ++ci) z = fma(x[ci], w[ci * cout], z);
  }
  return z;
}

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
template <typename Sync>
__device__ inline void cnn_forward_group(const CnnGeom& g, const double* P, const double* x0,
                                         double* A, double* D, const short* nbr, int tid, int nt,
                                         Sync sync) {
  for (int l = 0; l < g.nl; ++l) {
    const double* X = l == 0 ? x0 : A + g.aoff[l - 1];
    double* Al = A + g.aoff[l];
    double* Dl = D ? D + g.aoff[l] : nullptr;
    int cout = g.C[l];
    for (int t = tid; t < g.n * cout; t += nt) {
      int p = t / cout, co = t % cout;
      double z = cnn_preact(g, P, l, X, nbr, p, co);
      Al[t] = log_cosh(z);
      if (Dl) Dl[t] = tanh(z);
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
