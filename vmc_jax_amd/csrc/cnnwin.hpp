// Receptive-field window machinery of the CNN local-energy and Metropolis
// kernels (operators / proposals that change one site per string or step).
//
// Flipping site s = (si, sj) changes only the outputs of layer l inside the
// window ((si - a) mod Lx, (sj - b) mod Ly), 0 <= a < ex(l), 0 <= b < ey(l),
// ex(l) = 1 + (l+1)(Fx-1) (window index a*ey + b).  The kernels recompute those
// outputs from the stored activations of the unflipped configuration and sum
// the change of the last layer: the same log psi(s') as a full forward pass of
// the reference network (cnn.py:160-171) with a fraction of the work.
//
// Weights are read from a packed copy with every layer padded to CM channels
// (zeros): a padded output channel has pre-activation +0 in every
// configuration, so its change is exactly zero, and a padded input channel
// multiplies a zero weight, so the accumulation of the real channels is the
// same sequence of fused multiply-adds as the unpadded network's.
#pragma once
#include "nets.hpp"

namespace jv {

// LDS view of a generic pointer that points into the workgroup's LDS, so the
// loads below compile to ds_read instead of flat loads.
typedef const __attribute__((address_space(3))) double ldbl;
__device__ inline ldbl* as_lds(const double* p) { return (ldbl*)p; }
typedef const __attribute__((address_space(3))) float ldflt;
__device__ inline ldflt* as_lds(const float* p) { return (ldflt*)p; }
template <typename T>
using lds_t = const __attribute__((address_space(3))) T;
template <typename T>
using lds_w = __attribute__((address_space(3))) T;
// Wave-uniform weights through the constant address space (scalar loads).
template <typename T>
using cst_t = const __attribute__((address_space(4))) T;

// Layer l of the packed weights: kernel ((k*G + gi)*cinp + ci)*CG + cg for
// output channel c = gi*CG + cg (CG = CM/G, cinp = 1 for the first layer and
// CM after), then the bias [c].  G groups the output channels so that a wave
// owning group gi reads one contiguous block per kernel offset.
struct CnnPack {
  int cm, groups;
  int64_t koff[JV_MAX_LAYERS];
  int64_t boff[JV_MAX_LAYERS];
  int64_t total;
};

inline CnnPack make_cnn_pack(const CnnGeom& g, int cm, int groups) {
  CnnPack p;
  p.cm = cm;
  p.groups = groups;
  int64_t off = 0;
  for (int l = 0; l < g.nl; ++l) {
    const int cinp = l == 0 ? 1 : cm;
    p.koff[l] = off;
    off += (int64_t)g.FF * cinp * cm;
    p.boff[l] = off;
    off += cm;
  }
  p.total = off;
  return p;
}

// Writes the packed weights of params (jVMC flat layout) to out[pk.total]
// (FP64, or rounded to FP32 for the mixed-precision kernels).
template <typename T>
__global__ void cnn_pack(CnnGeom g, CnnPack pk, const double* __restrict__ Pg, T* __restrict__ out);

// Launches cnn_pack into workspace slot `slot`; returns the packed buffer or
// nullptr (error set).
template <typename T>
T* pack_cnn_weights(const CnnGeom& g, const CnnPack& pk, const double* params, hipStream_t st, int slot);

// Largest channel count rounded up to the kernels' instantiations (0: none).
int cnn_cmax(const CnnGeom& g);

// Activations of one configuration with the packed weights (G channel
// groups), channel-planar:
// layer l occupies A + l*lstride as CM planes of cstride doubles (positions
// [0, n)).  All CM planes are written (padded channels hold log_cosh(0));
// callers sum only the real channels of the last layer.
// With Zs, the pre-activations of layers l >= 1 are kept as well, at
// Zs + (l - 1) * lstride in the same planar layout.
template <int CM, int G, typename T, typename Sync>
__device__ inline void win_forward(const CnnGeom& g, const cst_t<T>* __restrict__ Wp, const CnnPack& pk,
                                   const T* x0, T* A, int lstride, int cstride, const short* nbr,
                                   int tid, int nt, Sync sync, T* Zs = nullptr) {
  for (int l = 0; l < g.nl; ++l) {
    const cst_t<T>* K = Wp + pk.koff[l];
    const cst_t<T>* B = Wp + pk.boff[l];
    lds_t<T>* X = (lds_t<T>*)(l == 0 ? x0 : A + (l - 1) * lstride);
    lds_t<T>* X0 = (lds_t<T>*)x0;
    T* Al = A + l * lstride;
    for (int p = tid; p < g.n; p += nt) {
      T z[CM];
#pragma unroll
      for (int c = 0; c < CM; ++c) z[c] = B[c];
      for (int k = 0; k < g.FF; ++k) {
        const int q = nbr[p * g.FF + k];
        constexpr int CG = CM / G;
        if (l == 0) {
          const T xv = X0[q];
          const cst_t<T>* w = K + k * CM;
#pragma unroll
          for (int c = 0; c < CM; ++c) z[c] = fma(xv, w[(c / CG) * CG + c % CG], z[c]);
        } else {
          const cst_t<T>* w = K + k * CM * CM;
#pragma unroll
          for (int ci = 0; ci < CM; ++ci) {
            const T xv = X[ci * cstride + q];
#pragma unroll
            for (int c = 0; c < CM; ++c) z[c] = fma(xv, w[((c / CG) * CM + ci) * CG + c % CG], z[c]);
          }
        }
      }
#pragma unroll
      for (int c = 0; c < CM; ++c) Al[c * cstride + p] = log_cosh(z[c]);
      if (Zs != nullptr && l >= 1) {
#pragma unroll
        for (int c = 0; c < CM; ++c) Zs[(l - 1) * lstride + c * cstride + p] = z[c];
      }
    }
    sync();
  }
}

// Position index of window element (a, b) of a flip at (si, sj).
__device__ inline int win_site(const CnnGeom& g, int si, int sj, int a, int b) {
  int pi = si - a;
  pi = pi < 0 ? pi + g.Lx : pi;
  int pj = sj - b;
  pj = pj < 0 ? pj + g.Ly : pj;
  return pi * g.Ly + pj;
}

// Pre-activations (all CM channels, packed weights with G = 1) of window
// element (a, b) of layer l >= 1 after flipping (si, sj).  Inputs inside the
// layer-(l-1) window come from Xw ([ci][pwsz] planar), the others from the
// stored activations Xb ([ci][n] planar).
template <int CM, typename T>
__device__ inline void win_preact(const CnnGeom& g, const cst_t<T>* __restrict__ K, const cst_t<T>* __restrict__ B,
                                  const T* Xb, const T* Xw, int pex, int pey, int si, int sj, int a,
                                  int b, T* z) {
  const int pwsz = pex * pey;
#pragma unroll
  for (int c = 0; c < CM; ++c) z[c] = B[c];
  for (int ky = 0; ky < g.Fx; ++ky) {
    const int aq = a - ky;
    int qi = si - aq;
    qi = qi < 0 ? qi + g.Lx : (qi >= g.Lx ? qi - g.Lx : qi);
    const bool rin = aq >= 0 && aq < pex;
    for (int kx = 0; kx < g.Fy; ++kx) {
      const int bq = b - kx;
      int qj = sj - bq;
      qj = qj < 0 ? qj + g.Ly : (qj >= g.Ly ? qj - g.Ly : qj);
      const bool in = rin && bq >= 0 && bq < pey;
      lds_t<T>* x = (lds_t<T>*)(in ? Xw + (aq * pey + bq) : Xb + (qi * g.Ly + qj));
      const int xs = in ? pwsz : g.n;
      const cst_t<T>* w = K + (ky * g.Fy + kx) * CM * CM;
#pragma unroll
      for (int ci = 0; ci < CM; ++ci) {
        const T xv = x[ci * xs];
#pragma unroll
        for (int c = 0; c < CM; ++c) z[c] = fma(xv, w[ci * CM + c], z[c]);
      }
    }
  }
}

// Layer-0 version: one input channel, the flipped site reads xflip.
template <int CM, typename T>
__device__ inline void win_preact0(const CnnGeom& g, const cst_t<T>* __restrict__ K, const cst_t<T>* __restrict__ B,
                                   const T* x0, T xflip, int site, int si, int sj, int a, int b,
                                   T* z) {
#pragma unroll
  for (int c = 0; c < CM; ++c) z[c] = B[c];
  for (int ky = 0; ky < g.Fx; ++ky) {
    int qi = si - a + ky;
    qi = qi < 0 ? qi + g.Lx : (qi >= g.Lx ? qi - g.Lx : qi);
    for (int kx = 0; kx < g.Fy; ++kx) {
      int qj = sj - b + kx;
      qj = qj < 0 ? qj + g.Ly : (qj >= g.Ly ? qj - g.Ly : qj);
      const int q = qi * g.Ly + qj;
      const T xv = q == site ? xflip : ((lds_t<T>*)x0)[q];
      const cst_t<T>* w = K + (ky * g.Fy + kx) * CM;
#pragma unroll
      for (int c = 0; c < CM; ++c) z[c] = fma(xv, w[c], z[c]);
    }
  }
}

// 3x3-filter versions of win_preact / win_preact0 for layer LC of a network
// whose windows are (1 + 2 (l + 1))^2: the same inputs and the same order of
// fused multiply-adds (kernel row, kernel column, input channel, output
// channel), with the window shapes as constants and the nine input locations
// computed before any load.
__device__ inline int wrap_once(int q, int L) { return q < 0 ? q + L : (q >= L ? q - L : q); }

template <int CM, int LC, typename T>
__device__ inline void win_preact3(const CnnGeom& g, const cst_t<T>* __restrict__ K, const cst_t<T>* __restrict__ B,
                                   const T* Xb, const T* Xw, int si, int sj, int a, int b, T* z) {
  // Inputs inside the previous layer's window come from Xw, the others from
  // the stored layer at row si - a + ky, column sj - b + kx (periodic; the
  // window fits the lattice, so one period at most): the rows are wrapped
  // once per kernel row, the three columns once, and the two candidates are
  // element offsets from Xb selected without a branch.
  constexpr int pex = 1 + 2 * LC, pey = pex, pwsz = pex * pey;
  const int dw = (int)(Xw - Xb);
  int cl[3];
  cl[0] = wrap_once(sj - b, g.Ly);
  cl[1] = cl[0] + 1 == g.Ly ? 0 : cl[0] + 1;
  cl[2] = cl[1] + 1 == g.Ly ? 0 : cl[1] + 1;
  int row = wrap_once(si - a, g.Lx) * g.Ly;
#pragma unroll
  for (int c = 0; c < CM; ++c) z[c] = B[c];
#pragma unroll 1
  for (int ky = 0; ky < 3; ++ky) {
    const int aq = a - ky;
    const bool rin = (unsigned)aq < (unsigned)pex;
#pragma unroll
    for (int kx = 0; kx < 3; ++kx) {
      const int bq = b - kx;
      const bool in = rin && (unsigned)bq < (unsigned)pey;
      const int off = in ? dw + (aq * pey + bq) : row + cl[kx];
      const int xs = in ? pwsz : g.n;
      lds_t<T>* x = (lds_t<T>*)(Xb + off);
      const cst_t<T>* w = K + (ky * 3 + kx) * CM * CM;
#pragma unroll
      for (int ci = 0; ci < CM; ++ci) {
        const T xv = x[ci * xs];
#pragma unroll
        for (int c = 0; c < CM; ++c) z[c] = fma(xv, w[ci * CM + c], z[c]);
      }
    }
    row += g.Ly;
    if (row == g.n) row = 0;
  }
}

// Change of the pre-activations of window element (a, b) of layer LC >= 1 of
// a 3x3 network after one flip: z += sum over the kernel offsets whose input
// lies inside the layer-(LC-1) window of W[k] dX, with dX the change of that
// layer's activations (window buffer D: element e, channel ci of flip f at
// D[(ci * pwsz + e) * fs + f]).  Inputs outside the window did not change and
// are skipped; the offsets taken depend on (a, b) alone, so lanes that share
// a window element share the branch.
template <int CM, int LC, typename T>
__device__ inline void win_delta3(const cst_t<T>* __restrict__ K, const T* D, int fs, int f, int a, int b, T* z) {
  constexpr int pex = 1 + 2 * LC, pwsz = pex * pex;  // the layer-(LC-1) window
#pragma unroll 1
  for (int k = 0; k < 9; ++k) {
    const int ky = k / 3, kx = k - 3 * (k / 3);
    const int aq = a - ky, bq = b - kx;
    if (aq >= 0 && aq < pex && bq >= 0 && bq < pex) {
      lds_t<T>* x = (lds_t<T>*)(D + (aq * pex + bq) * fs + f);
      const cst_t<T>* w = K + k * CM * CM;
#pragma unroll
      for (int ci = 0; ci < CM; ++ci) {
        const T xv = x[ci * pwsz * fs];
#pragma unroll
        for (int c = 0; c < CM; ++c) z[c] = fma(xv, w[ci * CM + c], z[c]);
      }
    }
  }
}

template <int CM, typename T>
__device__ inline void win_preact03(const CnnGeom& g, const cst_t<T>* __restrict__ K, const cst_t<T>* __restrict__ B,
                                    const T* x0, T xflip, int site, int si, int sj, int a, int b, T* z) {
#pragma unroll
  for (int c = 0; c < CM; ++c) z[c] = B[c];
  T xv[9];
#pragma unroll
  for (int ky = 0; ky < 3; ++ky) {
    const int qi = wrap_once(si - a + ky, g.Lx);
#pragma unroll
    for (int kx = 0; kx < 3; ++kx) {
      const int q = qi * g.Ly + wrap_once(sj - b + kx, g.Ly);
      const T v = ((lds_t<T>*)x0)[q];
      xv[ky * 3 + kx] = q == site ? xflip : v;
    }
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const cst_t<T>* w = K + k * CM;
#pragma unroll
    for (int c = 0; c < CM; ++c) z[c] = fma(xv[k], w[c], z[c]);
  }
}

}  // namespace jv
