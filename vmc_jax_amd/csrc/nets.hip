// Network descriptors, forward pass and fused per-sample log-derivatives O_k.
#include <algorithm>
#include <string>
#include <vector>

#include "cnnwin.hpp"
#include "rbm.hpp"

namespace jv {

// ------------------------------------------------------------------ host geometry
bool make_rbm_geom(const jv_net* net, RbmGeom& g) {
  if (!net || (net->kind != JV_NET_CPXRBM && net->kind != JV_NET_RBM)) return false;
  g.L = net->num_sites;
  g.H = net->num_hidden;
  g.cpx = net->kind == JV_NET_CPXRBM;
  g.ob = net->off_hidden_bias;
  g.ow = net->off_kernel;
  g.oa = net->off_visible_bias;
  return true;
}

bool make_cnn_geom(const jv_net* net, CnnGeom& g) {
  if (!net || net->kind != JV_NET_CNN) return false;
  g.Lx = net->shape[0];
  g.Ly = net->ndim == 2 ? net->shape[1] : 1;
  g.n = g.Lx * g.Ly;
  g.Fx = net->filter[0];
  g.Fy = net->ndim == 2 ? net->filter[1] : 1;
  g.FF = g.Fx * g.Fy;
  g.nl = net->num_layers;
  int off = 0;
  for (int l = 0; l < g.nl; ++l) {
    g.C[l] = net->channels[l];
    g.Cin[l] = l == 0 ? 1 : net->channels[l - 1];
    g.aoff[l] = off;
    off += g.n * g.C[l];
    g.koff[l] = net->off_conv_kernel[l];
    g.boff[l] = net->off_conv_bias[l];
  }
  g.atot = off;
  g.P = (int)net->num_real_params;
  g.inv_nrm = 1.0 / sqrt((double)g.n * g.C[g.nl - 1]);
  return true;
}

// LDS bytes of the CNN forward / gradient kernel (tsize: bytes of the
// arithmetic type, 8 for FP64, 4 for the mixed form).
size_t cnn_fwd_lds(const CnnGeom& g, bool grad, int tsize = 8) {
  size_t d = (size_t)g.P + g.n + (size_t)g.atot * (grad ? 2 : 1);
  size_t s = 2 * (size_t)g.n * g.FF;
  return 8 * 8 + d * tsize + s * 2 + 16;
}

int set_lds_limit(const void* fn, size_t bytes) {
  if (bytes > 160 * 1024) return fail(JV_E_NOMEM, "network needs " + std::to_string(bytes) + " bytes of LDS per workgroup (limit 160 KiB)");
  if (bytes > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute(MaxDynamicSharedMemorySize)");
  }
  return JV_OK;
}

// ------------------------------------------------------------------ RBM kernel
// One wavefront per sample (rbm_fwd_grad_sample in rbm.hpp).
__global__ void __launch_bounds__(256) rbm_fwd_grad(RbmGeom g, const double* __restrict__ P,
                                                    const int32_t* __restrict__ cfg, int64_t n,
                                                    int layout, double* __restrict__ grad,
                                                    double* __restrict__ logpsi) {
  const int lane = threadIdx.x % WAVE;
  const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / WAVE);
  const int64_t rowlen = rbm_grad_row_doubles(g, layout);
  for (int64_t s = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE; s < n; s += wstride)
    rbm_fwd_grad_sample(g, P, cfg + s * g.L, layout, grad ? grad + s * rowlen : nullptr,
                        logpsi ? logpsi + 2 * s : nullptr, lane);
}

// ------------------------------------------------------------------ CNN kernel
// One workgroup per sample (grid-stride): forward into LDS, then reverse pass
// for dlogpsi/dtheta (cnn.py:131-171).  T = double: the FP64 form; T = float:
// the mixed form (FP32 activations and O, log psi summed in FP64).
template <typename T>
__global__ void __launch_bounds__(256) cnn_fwd_grad(CnnGeom g, const double* __restrict__ Pg,
                                                    const int32_t* __restrict__ cfg, int64_t n,
                                                    int layout, T* __restrict__ grad,
                                                    double* __restrict__ logpsi) {
  extern __shared__ double lds[];
  const bool want_grad = grad != nullptr;
  double* red = lds;
  T* P = (T*)(red + 8);
  T* x0 = P + g.P;
  T* A = x0 + g.n;
  T* D = want_grad ? A + g.atot : nullptr;
  short* nbr = (short*)(A + g.atot * (want_grad ? 2 : 1));
  short* inv = nbr + g.n * g.FF;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int i = tid; i < g.P; i += nt) P[i] = (T)Pg[i];
  cnn_build_tables(g, nbr, inv, tid, nt);
  __syncthreads();
  auto sync = [] { __syncthreads(); };
  for (int64_t s = blockIdx.x; s < n; s += gridDim.x) {
    const int32_t* c = cfg + s * g.n;
    for (int i = tid; i < g.n; i += nt) x0[i] = (T)(2 * c[i] - 1);
    __syncthreads();
    cnn_forward_group<T>(g, P, x0, A, D, nbr, tid, nt, sync);
    const int L1 = g.nl - 1;
    double part = 0.0;
    for (int t = tid; t < g.n * g.C[L1]; t += nt) part += (double)A[g.aoff[L1] + t];
    double lp = block_sum(part, red) * g.inv_nrm;
    if (logpsi && tid == 0) {
      logpsi[2 * s] = lp;
      logpsi[2 * s + 1] = 0.0;
    }
    if (!want_grad) continue;
    // reverse pass; D_l is overwritten by delta_l = d out / d z_l
    T* row = grad + s * (int64_t)g.P * (layout == 0 ? 1 : 2);
    auto put = [&](int64_t col, T v) {
      if (layout == 0) row[col] = v;
      else { row[2 * col] = v; row[2 * col + 1] = (T)0; }
    };
    for (int t = tid; t < g.n * g.C[L1]; t += nt) D[g.aoff[L1] + t] *= (T)g.inv_nrm;
    __syncthreads();
    for (int l = L1; l >= 0; --l) {
      const int cin = g.Cin[l], cout = g.C[l];
      const T* dl = D + g.aoff[l];
      const T* X = l == 0 ? x0 : A + g.aoff[l - 1];
      // kernel gradient: sum_p delta[p, co] X[nbr(p, k), ci]
      for (int t = tid; t < g.FF * cin * cout; t += nt) {
        int co = t % cout, kc = t / cout, ci = kc % cin, k = kc / cin;
        T acc = 0;
        for (int p = 0; p < g.n; ++p) acc = fma(dl[p * cout + co], X[nbr[p * g.FF + k] * cin + ci], acc);
        put(g.koff[l] + t, acc);
      }
      if (g.boff[l] >= 0) {
        for (int co = tid; co < cout; co += nt) {
          T acc = 0;
          for (int p = 0; p < g.n; ++p) acc += dl[p * cout + co];
          put(g.boff[l] + co, acc);
        }
      }
      if (l > 0) {
        // delta_{l-1}[q, ci] = tanh(z_{l-1}[q, ci]) * sum_{k, co} delta_l[inv(q, k), co] K[k, ci, co]
        const T* K = P + g.koff[l];
        T* dp = D + g.aoff[l - 1];
        for (int t = tid; t < g.n * cin; t += nt) {
          int q = t / cin, ci = t % cin;
          T acc = 0;
          for (int k = 0; k < g.FF; ++k) {
            const T* d = dl + (int)inv[q * g.FF + k] * cout;
            const T* w = K + (k * cin + ci) * cout;
            for (int co = 0; co < cout; ++co) acc = fma(d[co], w[co], acc);
          }
          dp[t] *= acc;
        }
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------ CNN kernel, planar and register-blocked
// The same forward pass and gradients for networks whose channels fit CM
// (padded with zero channels) and whose planes fit LDS: activations planar
// (A_l: CM planes of n values), tanh(z) and then the back-propagated deltas
// interleaved per position (D_l[p][c]), weights packed as for the window
// kernels and read as wave-uniform scalar operands.  Per thread:
//  * forward: one position, all CM output channels (win_forward's order);
//  * kernel gradient dW[k][ci][co] = sum_p delta[p][co] x[ci][nbr(p, k)]: one
//    (k, ci) pair and one slice of the positions, all CM output channels
//    (one gather of x and one contiguous read of delta[p] per term), slices
//    summed in slice order;
//  * delta_{l-1}[q][ci] = tanh(z)[q][ci] sum_{k, co} delta_l[inv(q, k)][co]
//    K[k][ci][co]: one position, all CM input channels, in the reference
//    kernel's order (k, then co) for every ci.

template <int CM, typename T>
__global__ void __launch_bounds__(256) cnn_fwd_grad_p(CnnGeom g, const T* __restrict__ Wpk, CnnPack pk,
                                                      const int32_t* __restrict__ cfg, int64_t n, int layout,
                                                      T* __restrict__ grad, double* __restrict__ logpsi) {
  chain_priority();
  extern __shared__ double lds[];
  const bool want_grad = grad != nullptr;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int L1 = g.nl - 1, nn = g.n;
  double* red = lds;
  T* x0 = (T*)(red + 8);
  T* A = x0 + nn;                          // A + l * CM * n: planes [c][p]
  T* D = A + (size_t)g.nl * CM * nn;       // D + l * n * CM: [p][c]
  T* part = D + (size_t)g.nl * CM * nn;    // [nt][CM]
  short* nbr = (short*)(part + (size_t)cnn_p_items(g, CM, nt) * CM);
  short* inv = nbr + nn * g.FF;
  const cst_t<T>* Wp = (const cst_t<T>*)Wpk;
  cnn_build_tables(g, nbr, inv, tid, nt);
  __syncthreads();
  for (int64_t s = blockIdx.x; s < n; s += gridDim.x) {
    const int32_t* c = cfg + s * nn;
    for (int i = tid; i < nn; i += nt) x0[i] = (T)(2 * c[i] - 1);
    __syncthreads();
    for (int l = 0; l < g.nl; ++l) {
      const cst_t<T>* K = Wp + pk.koff[l];
      const cst_t<T>* B = Wp + pk.boff[l];
      lds_t<T>* X = (lds_t<T>*)(l == 0 ? x0 : A + (size_t)(l - 1) * CM * nn);
      T* Al = A + (size_t)l * CM * nn;
      T* Dl = D + (size_t)l * CM * nn;
      for (int p = tid; p < nn; p += nt) {
        T z[CM];
#pragma unroll
        for (int co = 0; co < CM; ++co) z[co] = B[co];
        for (int k = 0; k < g.FF; ++k) {
          const int q = nbr[p * g.FF + k];
          if (l == 0) {
            const T xv = X[q];
            const cst_t<T>* w = K + k * CM;
#pragma unroll
            for (int co = 0; co < CM; ++co) z[co] = fma(xv, w[co], z[co]);
          } else {
            const cst_t<T>* w = K + k * CM * CM;
#pragma unroll
            for (int ci = 0; ci < CM; ++ci) {
              const T xv = X[ci * nn + q];
#pragma unroll
              for (int co = 0; co < CM; ++co) z[co] = fma(xv, w[ci * CM + co], z[co]);
            }
          }
        }
        if (want_grad) {
#pragma unroll
          for (int co = 0; co < CM; ++co) {
            T th;
            Al[co * nn + p] = log_cosh_tanh(z[co], th);
            Dl[p * CM + co] = th;
          }
        } else {
#pragma unroll
          for (int co = 0; co < CM; ++co) Al[co * nn + p] = log_cosh(z[co]);
        }
      }
      __syncthreads();
    }
    double pl = 0.0;
    for (int t = tid; t < g.C[L1] * nn; t += nt) pl += (double)A[(size_t)L1 * CM * nn + t];
    const double lp = block_sum(pl, red) * g.inv_nrm;
    if (logpsi && tid == 0) {
      logpsi[2 * s] = lp;
      logpsi[2 * s + 1] = 0.0;
    }
    if (!want_grad) continue;
    T* row = grad + s * (int64_t)g.P * (layout == 0 ? 1 : 2);
    auto put = [&](int64_t col, T v) {
      if (layout == 0) row[col] = v;
      else {
        row[2 * col] = v;
        row[2 * col + 1] = (T)0;
      }
    };
    {
      T* Dl = D + (size_t)L1 * CM * nn;
      for (int t = tid; t < nn * CM; t += nt) Dl[t] *= (T)g.inv_nrm;
    }
    __syncthreads();
    for (int l = L1; l >= 0; --l) {
      const int cinr = g.Cin[l], coutr = g.C[l];
      lds_t<T>* X = (lds_t<T>*)(l == 0 ? x0 : A + (size_t)(l - 1) * CM * nn);
      lds_t<T>* dl = (lds_t<T>*)(D + (size_t)l * CM * nn);
      const int pairs = g.FF * cinr;
      const int ps = max(1, min(nn, nt / pairs));  // position slices
      // kernel-gradient partial sums: item (slice, pair), all CM output channels
      for (int it = tid; it < pairs * ps; it += nt) {
        const int sl = it / pairs, pr = it - sl * pairs;
        const int k = pr / cinr, ci = pr - k * cinr;
        const int p0 = (int)((int64_t)sl * nn / ps), p1 = (int)((int64_t)(sl + 1) * nn / ps);
        T acc[CM];
#pragma unroll
        for (int co = 0; co < CM; ++co) acc[co] = (T)0;
        for (int p = p0; p < p1; ++p) {
          const T xv = X[ci * nn + nbr[p * g.FF + k]];
#pragma unroll
          for (int co = 0; co < CM; ++co) acc[co] = fma(dl[p * CM + co], xv, acc[co]);
        }
#pragma unroll
        for (int co = 0; co < CM; ++co) part[it * CM + co] = acc[co];
      }
      // deltas of the layer below (every position, all CM input channels)
      if (l > 0) {
        const cst_t<T>* K = Wp + pk.koff[l];
        T* dp = D + (size_t)(l - 1) * CM * nn;
        for (int q = tid; q < nn; q += nt) {
          T acc[CM];
#pragma unroll
          for (int ci = 0; ci < CM; ++ci) acc[ci] = (T)0;
          for (int k = 0; k < g.FF; ++k) {
            lds_t<T>* d = dl + (int)inv[q * g.FF + k] * CM;
            const cst_t<T>* w = K + k * CM * CM;
#pragma unroll
            for (int co = 0; co < CM; ++co) {
              const T dv = d[co];
#pragma unroll
              for (int ci = 0; ci < CM; ++ci) acc[ci] = fma(dv, w[ci * CM + co], acc[ci]);
            }
          }
          // written after every thread has read delta_l of this layer: dp is
          // another layer's buffer, read by no thread in this phase
#pragma unroll
          for (int ci = 0; ci < CM; ++ci) dp[q * CM + ci] *= acc[ci];
        }
      }
      __syncthreads();
      // sum the slices in slice order; bias gradients over all positions
      for (int t = tid; t < pairs * coutr; t += nt) {
        const int pr = t / coutr, co = t - pr * coutr;
        T v = part[pr * CM + co];
        for (int sl = 1; sl < ps; ++sl) v += part[(sl * pairs + pr) * CM + co];
        put(g.koff[l] + t, v);
      }
      if (g.boff[l] >= 0) {
        for (int co = tid; co < coutr; co += nt) {
          T acc = 0;
          for (int p = 0; p < nn; ++p) acc += dl[p * CM + co];
          put(g.boff[l] + co, acc);
        }
      }
      __syncthreads();
    }
  }
}

// Launches cnn_fwd_grad_p when the padded channel count has an instantiation
// and its LDS fits (*done = true); *done = false when it does not apply.
// Returns the status (packed weights in workspace slot `slot`).
template <typename T>
int cnn_p_launch(const CnnGeom& g, const double* params, const int32_t* configs, int64_t n, int layout, T* grad,
                 double* logpsi, int64_t blocks, hipStream_t st, int slot, bool* done);

// ------------------------------------------------------------------ packed CNN weights
template <typename T>
__global__ void cnn_pack(CnnGeom g, CnnPack pk, const double* __restrict__ Pg, T* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= pk.total) return;
  int l = 0;
  while (l + 1 < g.nl && t >= pk.koff[l + 1]) ++l;
  const int cm = pk.cm, cg_n = cm / pk.groups;
  const int cinp = l == 0 ? 1 : cm;
  double v = 0.0;
  if (t >= pk.boff[l]) {
    const int c = (int)(t - pk.boff[l]);
    if (c < g.C[l] && g.boff[l] >= 0) v = Pg[g.boff[l] + c];
  } else {
    int64_t r = t - pk.koff[l];
    const int cg = (int)(r % cg_n);
    r /= cg_n;
    const int ci = (int)(r % cinp);
    r /= cinp;
    const int gi = (int)(r % pk.groups);
    const int k = (int)(r / pk.groups);
    const int c = gi * cg_n + cg;
    if (ci < g.Cin[l] && c < g.C[l]) v = Pg[g.koff[l] + ((int64_t)k * g.Cin[l] + ci) * g.C[l] + c];
  }
  out[t] = (T)v;
}

template <typename T>
T* pack_cnn_weights(const CnnGeom& g, const CnnPack& pk, const double* params, hipStream_t st, int slot) {
  T* out = (T*)workspace(sizeof(T) * (size_t)pk.total, slot);
  if (!out) {
    fail(JV_E_NOMEM, "packed CNN weights: workspace allocation failed");
    return nullptr;
  }
  hipLaunchKernelGGL(cnn_pack<T>, dim3((unsigned)((pk.total + 255) / 256)), dim3(256), 0, st, g, pk, params, out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    hip_fail(e, "cnn_pack launch");
    return nullptr;
  }
  return out;
}

template double* pack_cnn_weights<double>(const CnnGeom&, const CnnPack&, const double*, hipStream_t, int);
template float* pack_cnn_weights<float>(const CnnGeom&, const CnnPack&, const double*, hipStream_t, int);

int cnn_cmax(const CnnGeom& g) {
  int m = 0;
  for (int l = 0; l < g.nl; ++l) m = std::max(m, g.C[l]);
  if (m <= 4) return 4;
  if (m <= 8) return 8;
  if (m <= 16) return 16;
  if (m <= 32) return 32;
  return 0;
}

template <typename T>
int cnn_p_launch(const CnnGeom& g, const double* params, const int32_t* configs, int64_t n, int layout, T* grad,
                 double* logpsi, int64_t blocks, hipStream_t st, int slot, bool* done) {
  *done = false;
  const int cm = cnn_cmax(g);
  if (cm == 0) return JV_OK;
  const size_t lds = cnn_p_lds(g, cm, 256, (int)sizeof(T));
  if (lds > 160 * 1024) return JV_OK;
  const CnnPack pk = make_cnn_pack(g, cm, 1);
  const T* wpk = pack_cnn_weights<T>(g, pk, params, st, slot);
  if (!wpk) return JV_E_NOMEM;
  int rc = JV_OK;
#define JV_CNNP(CMv)                                                                                      \
  if (cm == CMv) {                                                                                        \
    rc = set_lds_limit((const void*)cnn_fwd_grad_p<CMv, T>, lds);                                         \
    if (rc) return rc;                                                                                    \
    hipLaunchKernelGGL((cnn_fwd_grad_p<CMv, T>), dim3((unsigned)blocks), dim3(256), lds, st, g, wpk, pk,   \
                       configs, n, layout, grad, logpsi);                                                 \
    JV_LAUNCH_CHECK("cnn_fwd_grad_p");                                                                    \
    *done = true;                                                                                         \
  }
  JV_CNNP(4) JV_CNNP(8) JV_CNNP(16) JV_CNNP(32)
#undef JV_CNNP
  return rc;
}
template int cnn_p_launch<double>(const CnnGeom&, const double*, const int32_t*, int64_t, int, double*, double*,
                                  int64_t, hipStream_t, int, bool*);
template int cnn_p_launch<float>(const CnnGeom&, const double*, const int32_t*, int64_t, int, float*, double*,
                                 int64_t, hipStream_t, int, bool*);

}  // namespace jv

using namespace jv;

extern "C" {

int jv_net_init(jv_net* net) {
  if (!net) return fail(JV_E_ARG, "jv_net_init: null descriptor");
  if (net->ndim < 1 || net->ndim > 2) return fail(JV_E_ARG, "ndim must be 1 or 2");
  if (net->precision != JV_PREC_FP64 && net->precision != JV_PREC_MIXED)
    return fail(JV_E_ARG, "precision must be JV_PREC_FP64 or JV_PREC_MIXED");
  if (net->ndim == 1) net->shape[1] = 1;
  if (net->shape[0] < 1 || net->shape[1] < 1) return fail(JV_E_ARG, "lattice extents must be positive");
  net->num_sites = net->shape[0] * net->shape[1];
  net->off_hidden_bias = net->off_kernel = net->off_visible_bias = -1;
  for (int l = 0; l < JV_MAX_LAYERS; ++l) net->off_conv_kernel[l] = net->off_conv_bias[l] = -1;
  if (net->kind == JV_NET_CPXRBM || net->kind == JV_NET_RBM) {
    const int L = net->num_sites, H = net->num_hidden;
    if (H < 1) return fail(JV_E_ARG, "num_hidden must be positive");
    if (net->kind == JV_NET_RBM && net->visible_bias) return fail(JV_E_ARG, "the real RBM has no visible bias");
    const int mult = net->kind == JV_NET_CPXRBM ? 2 : 1;
    int64_t off = 0, np = 0;
    // tree_flatten order: Dense_0/bias, Dense_0/kernel, visible_bias
    if (net->hidden_bias) { net->off_hidden_bias = off; off += (int64_t)mult * H; np += H; }
    net->off_kernel = off; off += (int64_t)mult * L * H; np += (int64_t)L * H;
    if (net->visible_bias) { net->off_visible_bias = off; off += (int64_t)mult * L; np += L; }
    net->is_complex = net->kind == JV_NET_CPXRBM;
    net->num_params = np;
    net->num_real_params = off;
    return JV_OK;
  }
  if (net->kind == JV_NET_CNN) {
    const int nl = net->num_layers;
    if (nl < 1 || nl > JV_MAX_LAYERS) return fail(JV_E_ARG, "num_layers must be in [1, 8]");
    if (net->ndim == 1) net->filter[1] = 1;
    if (net->filter[0] < 1 || net->filter[1] < 1) return fail(JV_E_ARG, "filter extents must be positive");
    if (net->num_sites > 32767) return fail(JV_E_ARG, "CNN lattices are limited to 32767 sites");
    const int FF = net->filter[0] * net->filter[1];
    // entries (name, leaf, size) sorted as JAX sorts dict keys
    struct E { std::string name; std::string leaf; int layer; int64_t size; };
    std::vector<E> es;
    for (int l = 0; l < nl; ++l) {
      if (net->channels[l] < 1) return fail(JV_E_ARG, "channels must be positive");
      int cin = l == 0 ? 1 : net->channels[l - 1];
      std::string nm = "Conv_" + std::to_string(l);
      if (net->layer_bias[l]) es.push_back({nm, "bias", l, net->channels[l]});
      es.push_back({nm, "kernel", l, (int64_t)FF * cin * net->channels[l]});
    }
    std::sort(es.begin(), es.end(), [](const E& a, const E& b) {
      return a.name != b.name ? a.name < b.name : a.leaf < b.leaf;
    });
    int64_t off = 0;
    for (auto& e : es) {
      if (e.leaf == "bias") net->off_conv_bias[e.layer] = off;
      else net->off_conv_kernel[e.layer] = off;
      off += e.size;
    }
    net->is_complex = 0;
    net->num_params = off;
    net->num_real_params = off;
    CnnGeom g;
    make_cnn_geom(net, g);
    if (cnn_fwd_lds(g, true) > 160 * 1024)
      return fail(JV_E_NOMEM, "CNN too large for one workgroup's LDS");
    return JV_OK;
  }
  return fail(JV_E_ARG, "unknown network kind");
}

int jv_net_gradients(const jv_net* net, const double* params, const int32_t* configs, int64_t n,
                     int32_t layout, double* grad, double* logpsi, void* stream) {
  if (!net || !params || !configs || n < 0) return fail(JV_E_ARG, "jv_net_gradients: bad arguments");
  if (n == 0) return JV_OK;
  if (layout != 0 && layout != 1) return fail(JV_E_ARG, "layout must be 0 or 1");
  hipStream_t st = (hipStream_t)stream;
  RbmGeom rg;
  CnnGeom cg;
  if (make_rbm_geom(net, rg)) {
    int64_t blocks = std::min<int64_t>((n + 3) / 4, 4096);
    ScopedTimer tm(grad ? "net_grad" : "net_forward", st);
    hipLaunchKernelGGL(rbm_fwd_grad, dim3((unsigned)blocks), dim3(256), 0, st, rg, params, configs, n,
                       layout, grad, logpsi);
    JV_LAUNCH_CHECK("rbm_fwd_grad");
    return JV_OK;
  }
  if (make_cnn_geom(net, cg)) {
    int64_t blocks = std::min<int64_t>(n, 8192);
    ScopedTimer tm(grad ? "net_grad" : "net_forward", st);
    bool done = false;
    int prc = cnn_p_launch<double>(cg, params, configs, n, layout, grad, logpsi, blocks, st, 48, &done);
    if (prc || done) return prc;
    size_t lds = cnn_fwd_lds(cg, grad != nullptr);
    int rc = set_lds_limit((const void*)cnn_fwd_grad<double>, lds);
    if (rc) return rc;
    hipLaunchKernelGGL(cnn_fwd_grad<double>, dim3((unsigned)blocks), dim3(256), lds, st, cg, params, configs, n,
                       layout, grad, logpsi);
    JV_LAUNCH_CHECK("cnn_fwd_grad");
    return JV_OK;
  }
  return fail(JV_E_ARG, "jv_net_gradients: unknown network kind (call jv_net_init first)");
}

int jv_net_gradients_f32(const jv_net* net, const double* params, const int32_t* configs, int64_t n, float* grad,
                         double* logpsi, void* stream) {
  if (!net || !params || !configs || !grad || n < 0) return fail(JV_E_ARG, "jv_net_gradients_f32: bad arguments");
  if (n == 0) return JV_OK;
  hipStream_t st = (hipStream_t)stream;
  CnnGeom cg;
  if (!make_cnn_geom(net, cg)) return fail(JV_E_ARG, "jv_net_gradients_f32: only the CNN has an FP32 form");
  const int64_t blocks = std::min<int64_t>(n, 8192);
  ScopedTimer tm("net_grad", st);
  bool done = false;
  int prc = cnn_p_launch<float>(cg, params, configs, n, 0, grad, logpsi, blocks, st, 49, &done);
  if (prc || done) return prc;
  const size_t lds = cnn_fwd_lds(cg, true, 4);
  int rc = set_lds_limit((const void*)cnn_fwd_grad<float>, lds);
  if (rc) return rc;
  hipLaunchKernelGGL(cnn_fwd_grad<float>, dim3((unsigned)blocks), dim3(256), lds, st, cg, params, configs, n, 0, grad,
                     logpsi);
  JV_LAUNCH_CHECK("cnn_fwd_grad");
  return JV_OK;
}

int jv_net_forward(const jv_net* net, const double* params, const int32_t* configs, int64_t n,
                   double* logpsi, void* stream) {
  if (!logpsi) return fail(JV_E_ARG, "jv_net_forward: null output");
  return jv_net_gradients(net, params, configs, n, 0, nullptr, logpsi, stream);
}

}  // extern "C"
