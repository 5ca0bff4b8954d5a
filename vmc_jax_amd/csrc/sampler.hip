// Metropolis sampling with walkers and log-psi caches resident in LDS, and the
// sample weights of the Metropolis and exact samplers.
#include <algorithm>

#include "nets.hpp"

namespace jv {

bool make_rbm_geom(const jv_net* net, RbmGeom& g);
bool make_cnn_geom(const jv_net* net, CnnGeom& g);
int set_lds_limit(const void* fn, size_t bytes);

__global__ void zero_i64(int64_t* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ------------------------------------------------------------------ RBM chains
// One wavefront per chain.  theta_h and Re log cosh(theta_h) are cached in LDS
// and updated on acceptance; a proposal costs one log-cosh per hidden unit
// instead of the full network re-evaluation of sampler.py:335.  theta is
// recomputed from scratch whenever a sample is emitted.
__global__ void __launch_bounds__(256) mc_rbm(RbmGeom g, const double* __restrict__ P, jv_mc_args a,
                                              int32_t* __restrict__ walkers, int32_t* __restrict__ configs,
                                              int64_t* __restrict__ counters) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
  const int chain = blockIdx.x * nw + w;
  if (chain >= a.num_chains) return;
  const size_t per = (size_t)3 * g.H + (g.L + 1) / 2;
  double* base = lds + (size_t)w * per;
  cd* th = (cd*)base;
  double* lc = base + 2 * g.H;
  int32_t* s = (int32_t*)(base + 3 * g.H);
  const double* Wr = P + g.ow;
  const double* Wi = P + g.ow + (int64_t)g.L * g.H;
  for (int i = lane; i < g.L; i += WAVE) s[i] = walkers[(int64_t)chain * g.L + i];
  wave_sync();
  auto refresh = [&]() {
    for (int h = lane; h < g.H; h += WAVE) {
      cd t = rbm_theta(g, P, s, h);
      th[h] = t;
      lc[h] = g.cpx ? re_log_cosh(t) : log_cosh(t.re);
    }
  };
  refresh();
  int64_t nacc = 0;
  const int total = a.therm_steps + a.samples_per_chain * a.sweep_steps;
  for (int step = 0; step < total; ++step) {
    int site;
    double u;
    mc_draw(a.seed, (uint32_t)step, (uint32_t)chain, a.call_index, a.stream, g.L, site, u);
    const double d = -2.0 * (double)(2 * s[site] - 1);
    double dl = 0.0;
    for (int h = lane; h < g.H; h += WAVE) {
      const int64_t wi = (int64_t)site * g.H + h;
      cd t = {fma(d, Wr[wi], th[h].re), g.cpx ? fma(d, Wi[wi], th[h].im) : 0.0};
      double l = g.cpx ? re_log_cosh(t) : log_cosh(t.re);
      dl += l - lc[h];
    }
    dl = wave_sum(dl);
    if (g.oa >= 0) dl = fma(d, P[g.oa + site], dl);
    const bool acc = u < exp(a.mu * dl);  // random.bernoulli(key, P), sampler.py:336-340
    if (acc) {
      for (int h = lane; h < g.H; h += WAVE) {
        const int64_t wi = (int64_t)site * g.H + h;
        cd t = {fma(d, Wr[wi], th[h].re), g.cpx ? fma(d, Wi[wi], th[h].im) : 0.0};
        th[h] = t;
        lc[h] = g.cpx ? re_log_cosh(t) : log_cosh(t.re);
      }
      wave_sync();
      if (lane == 0) s[site] ^= 1;
      ++nacc;
    }
    wave_sync();
    const int k = step + 1 - a.therm_steps;
    if (k > 0 && k % a.sweep_steps == 0) {
      const int64_t t = k / a.sweep_steps - 1;
      int32_t* out = configs + (t * a.num_chains + chain) * g.L;
      for (int i = lane; i < g.L; i += WAVE) out[i] = s[i];
      refresh();
      wave_sync();
    }
  }
  for (int i = lane; i < g.L; i += WAVE) walkers[(int64_t)chain * g.L + i] = s[i];
  if (lane == 0) {
    atomicAdd((unsigned long long*)&counters[0], (unsigned long long)total);
    atomicAdd((unsigned long long*)&counters[1], (unsigned long long)nacc);
  }
}

// ------------------------------------------------------------------ CNN chains
// One wavefront per chain; the chain's activations of every layer live in LDS.
// A proposal recomputes only the outputs whose receptive field contains the
// flipped site; on acceptance those outputs are committed, so the cached
// activations always equal a fresh forward pass of the walker.
struct McCnnLayout {
  size_t shared, per_chain;
};

__host__ __device__ inline McCnnLayout mc_cnn_layout(const CnnGeom& g) {
  McCnnLayout L;
  L.shared = 8 * (size_t)g.P + 2 * 2 * (size_t)g.n * g.FF;
  L.shared = (L.shared + 15) & ~(size_t)15;
  // x0 (n doubles), A (atot), xnew (n), Anew (atot), cfg (n int), list (n short), mask ((nl+1) n bytes)
  L.per_chain = 8 * (2 * (size_t)g.n + 2 * (size_t)g.atot) + 4 * (size_t)g.n + 2 * (size_t)g.n +
                (size_t)(g.nl + 1) * g.n;
  L.per_chain = (L.per_chain + 15) & ~(size_t)15;
  return L;
}

__global__ void __launch_bounds__(256) mc_cnn(CnnGeom g, const double* __restrict__ Pg, jv_mc_args a,
                                              int32_t* __restrict__ walkers, int32_t* __restrict__ configs,
                                              int64_t* __restrict__ counters) {
  extern __shared__ double lds[];
  const McCnnLayout lay = mc_cnn_layout(g);
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid % WAVE, w = tid / WAVE, nw = nt / WAVE;
  double* P = lds;
  short* nbr = (short*)(P + g.P);
  short* inv = nbr + g.n * g.FF;
  for (int i = tid; i < g.P; i += nt) P[i] = Pg[i];
  cnn_build_tables(g, nbr, inv, tid, nt);
  __syncthreads();
  const int chain = blockIdx.x * nw + w;
  if (chain >= a.num_chains) return;
  char* cb = (char*)lds + lay.shared + (size_t)w * lay.per_chain;
  double* x0 = (double*)cb;
  double* A = x0 + g.n;
  LocalScratch sc;
  sc.xnew = A + g.atot;
  sc.Anew = sc.xnew + g.n;
  int32_t* s = (int32_t*)(sc.Anew + g.atot);
  sc.list = (short*)(s + g.n);
  sc.mask = (unsigned char*)(sc.list + g.n);
  for (int i = lane; i < g.n; i += WAVE) {
    s[i] = walkers[(int64_t)chain * g.n + i];
    x0[i] = (double)(2 * s[i] - 1);
  }
  wave_sync();
  cnn_forward_group(g, P, x0, A, nullptr, nbr, lane, WAVE, [] { wave_sync(); });
  int64_t nacc = 0;
  const int total = a.therm_steps + a.samples_per_chain * a.sweep_steps;
  for (int step = 0; step < total; ++step) {
    int site;
    double u;
    mc_draw(a.seed, (uint32_t)step, (uint32_t)chain, a.call_index, a.stream, g.n, site, u);
    int sites[1] = {site};
    int vals[1] = {1 - s[site]};
    double dl = cnn_local_delta(g, P, x0, A, nbr, 1, sites, vals, sc, lane, true);
    const bool acc = u < exp(a.mu * dl);
    if (acc) {
      for (int l = 0; l < g.nl; ++l) {
        const unsigned char* m = sc.mask + (l + 1) * g.n;
        const int cout = g.C[l];
        double* Al = A + g.aoff[l];
        const double* An = sc.Anew + g.aoff[l];
        for (int t = lane; t < g.n * cout; t += WAVE)
          if (m[t / cout]) Al[t] = An[t];
      }
      if (lane == 0) {
        s[site] = vals[0];
        x0[site] = (double)(2 * vals[0] - 1);
      }
      ++nacc;
    }
    wave_sync();
    const int k = step + 1 - a.therm_steps;
    if (k > 0 && k % a.sweep_steps == 0) {
      const int64_t t = k / a.sweep_steps - 1;
      int32_t* out = configs + (t * a.num_chains + chain) * g.n;
      for (int i = lane; i < g.n; i += WAVE) out[i] = s[i];
    }
  }
  for (int i = lane; i < g.n; i += WAVE) walkers[(int64_t)chain * g.n + i] = s[i];
  if (lane == 0) {
    atomicAdd((unsigned long long*)&counters[0], (unsigned long long)total);
    atomicAdd((unsigned long long*)&counters[1], (unsigned long long)nacc);
  }
}

// ------------------------------------------------------------------ CNN chains, single-flip windows
// One wavefront per chain; lane j computes all channels of window position j of
// each layer (closed-form receptive field of the flipped site), so a proposal
// costs one pass over the 9 / 25 / 49 window positions of the 3x3 net.
__host__ __device__ inline size_t mc_win_chain_doubles(const CnnGeom& g) {
  size_t d = (size_t)g.n + g.atot + (g.n + 1) / 2;  // x0, A, walker (int32)
  for (int l = 0; l < g.nl; ++l) d += (size_t)win_ex(g, l) * win_ey(g, l) * g.C[l];
  return d;
}
__host__ __device__ inline size_t mc_win_shared_doubles(const CnnGeom& g) {
  return ((size_t)g.n * g.FF + 3) / 4;  // nbr table (shorts)
}

template <int CM>
__global__ void __launch_bounds__(256) mc_cnn_win(CnnGeom g, const double* __restrict__ Pg, jv_mc_args a,
                                                  int32_t* __restrict__ walkers, int32_t* __restrict__ configs,
                                                  int64_t* __restrict__ counters) {
  extern __shared__ double lds[];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid % WAVE, w = tid / WAVE, nw = nt / WAVE;
  short* nbr = (short*)lds;
  for (int t = tid; t < g.n * g.FF; t += nt) {
    int p = t / g.FF, k = t % g.FF;
    int i = p / g.Ly, j = p % g.Ly, ky = k / g.Fy, kx = k % g.Fy;
    nbr[t] = (short)(((i + ky) % g.Lx) * g.Ly + (j + kx) % g.Ly);
  }
  __syncthreads();
  const int chain = blockIdx.x * nw + w;
  if (chain >= a.num_chains) return;
  double* x0 = lds + mc_win_shared_doubles(g) + (size_t)w * mc_win_chain_doubles(g);
  double* A = x0 + g.n;
  double* W[JV_MAX_LAYERS];
  double* wp = A + g.atot;
  for (int l = 0; l < g.nl; ++l) {
    W[l] = wp;
    wp += (size_t)win_ex(g, l) * win_ey(g, l) * g.C[l];
  }
  int32_t* s = (int32_t*)wp;
  for (int i = lane; i < g.n; i += WAVE) {
    s[i] = walkers[(int64_t)chain * g.n + i];
    x0[i] = (double)(2 * s[i] - 1);
  }
  wave_sync();
  cnn_forward_pos<CM>(g, Pg, x0, A, nullptr, nbr, lane, WAVE, [] { wave_sync(); });
  const int L1 = g.nl - 1;
  int64_t nacc = 0;
  double z[CM];
  const int total = a.therm_steps + a.samples_per_chain * a.sweep_steps;
  for (int step = 0; step < total; ++step) {
    int site;
    double u;
    mc_draw(a.seed, (uint32_t)step, (uint32_t)chain, a.call_index, a.stream, g.n, site, u);
    const int si = site / g.Ly, sj = site % g.Ly;
    const int nv = 1 - s[site];
    const double xflip = (double)(2 * nv - 1);
    double diff = 0.0;
    for (int l = 0; l < g.nl; ++l) {
      const int ey = win_ey(g, l), wsz = win_ex(g, l) * ey, cout = g.C[l];
      const double* Xold = l == 0 ? x0 : A + g.aoff[l - 1];
      for (int j = lane; j < wsz; j += WAVE) {
        const int aa = j / ey, bb = j % ey;
        cnn_preact_win<CM>(g, Pg, l, Xold, l == 0 ? nullptr : W[l - 1], xflip, si, sj, aa, bb, z);
        double* Wo = W[l] + j;  // planar [c][j]
        const double* Ab = A + g.aoff[l] + win_pos(g, si, sj, aa, bb);
#pragma unroll
        for (int c = 0; c < CM; ++c)
          if (c < cout) {
            const double v = log_cosh(z[c]);
            Wo[c * wsz] = v;
            if (l == L1) diff += v - Ab[c * g.n];
          }
      }
      wave_sync();
    }
    const double dl = wave_sum(diff) * g.inv_nrm;
    const bool acc = u < exp(a.mu * dl);  // random.bernoulli(key, P), sampler.py:336-340
    if (acc) {
      for (int l = 0; l < g.nl; ++l) {
        const int ey = win_ey(g, l), wsz = win_ex(g, l) * ey, cout = g.C[l];
        for (int t = lane; t < wsz * cout; t += WAVE) {
          const int c = t / wsz, j = t % wsz;
          A[g.aoff[l] + c * g.n + win_pos(g, si, sj, j / ey, j % ey)] = W[l][t];
        }
      }
      if (lane == 0) {
        s[site] = nv;
        x0[site] = xflip;
      }
      ++nacc;
    }
    wave_sync();
    const int k = step + 1 - a.therm_steps;
    if (k > 0 && k % a.sweep_steps == 0) {
      const int64_t t = k / a.sweep_steps - 1;
      int32_t* out = configs + (t * a.num_chains + chain) * g.n;
      for (int i = lane; i < g.n; i += WAVE) out[i] = s[i];
    }
  }
  for (int i = lane; i < g.n; i += WAVE) walkers[(int64_t)chain * g.n + i] = s[i];
  if (lane == 0) {
    atomicAdd((unsigned long long*)&counters[0], (unsigned long long)total);
    atomicAdd((unsigned long long*)&counters[1], (unsigned long long)nacc);
  }
}

int cnn_cmax(const CnnGeom& g);

// ------------------------------------------------------------------ weights
// p_i = exp(scale * Re(logpsi_i) - shift); one block, deterministic order.
__global__ void __launch_bounds__(1024) weights_sum(const double* __restrict__ lp, int64_t n, double scale,
                                                    double shift, double* __restrict__ p,
                                                    double* __restrict__ sum_out) {
  __shared__ double red[16];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    double v = exp(scale * lp[2 * i] - shift);
    p[i] = v;
    acc += v;
  }
  double s = block_sum(acc, red);
  if (threadIdx.x == 0) *sum_out = s;
}

__global__ void weights_norm(double* __restrict__ p, int64_t n, const double* __restrict__ sum) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = p[i] / *sum;
}

int launch_weights(const double* logpsi, int64_t n, double scale, double shift, double* p, double* norm_out,
                   hipStream_t st) {
  double* sum = (double*)workspace(sizeof(double) * 2, 1);
  if (!sum) return fail(JV_E_NOMEM, "weights: workspace allocation failed");
  hipLaunchKernelGGL(weights_sum, dim3(1), dim3(1024), 0, st, logpsi, n, scale, shift, p, sum);
  JV_LAUNCH_CHECK("weights_sum");
  hipLaunchKernelGGL(weights_norm, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, sum);
  JV_LAUNCH_CHECK("weights_norm");
  if (norm_out) {
    JV_HIP(hipMemcpyAsync(norm_out, sum, sizeof(double), hipMemcpyDeviceToHost, st));
    JV_HIP(hipStreamSynchronize(st));
  }
  return JV_OK;
}

}  // namespace jv

using namespace jv;

extern "C" {

int jv_mc_sample(const jv_net* net, const double* params, const jv_mc_args* args, int32_t* walkers,
                 int32_t* configs, double* logpsi, int64_t* counters, void* stream) {
  if (!net || !params || !args || !walkers || !counters) return fail(JV_E_ARG, "jv_mc_sample: bad arguments");
  const jv_mc_args a = *args;
  if (a.num_chains < 1 || a.sweep_steps < 1 || a.therm_steps < 0 || a.samples_per_chain < 0)
    return fail(JV_E_ARG, "jv_mc_sample: bad chain / sweep counts");
  if (a.samples_per_chain > 0 && !configs) return fail(JV_E_ARG, "jv_mc_sample: null configs");
  if (a.mu < 0.0 || a.mu > 2.0) return fail(JV_E_ARG, "mu must be in the range [0, 2]");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(zero_i64, dim3(1), dim3(64), 0, st, counters, 2);
  JV_LAUNCH_CHECK("zero_i64");
  RbmGeom rg;
  CnnGeom cg;
  if (make_rbm_geom(net, rg)) {
    const int nw = 4;
    size_t lds = nw * 8 * ((size_t)3 * rg.H + (rg.L + 1) / 2);
    int rc = set_lds_limit((const void*)mc_rbm, lds);
    if (rc) return rc;
    ScopedTimer tm("mc_sweep", st);
    hipLaunchKernelGGL(mc_rbm, dim3((unsigned)((a.num_chains + nw - 1) / nw)), dim3(64 * nw), lds, st, rg,
                       params, a, walkers, configs, counters);
    JV_LAUNCH_CHECK("mc_rbm");
  } else if (make_cnn_geom(net, cg) && win_ok(cg) && cnn_cmax(cg) > 0) {
    const int nw = a.num_chains <= 1024 ? 1 : 4;
    size_t lds = 8 * (mc_win_shared_doubles(cg) + nw * mc_win_chain_doubles(cg));
    const unsigned blocks = (unsigned)((a.num_chains + nw - 1) / nw);
    ScopedTimer tm("mc_sweep", st);
    int rc = JV_OK;
    switch (cnn_cmax(cg)) {
#define JV_MCW(CMv)                                                                                     \
  case CMv:                                                                                             \
    rc = set_lds_limit((const void*)mc_cnn_win<CMv>, lds);                                              \
    if (rc) return rc;                                                                                  \
    hipLaunchKernelGGL(mc_cnn_win<CMv>, dim3(blocks), dim3(64 * nw), lds, st, cg, params, a, walkers,       \
                       configs, counters);                                                              \
    break;
      JV_MCW(4) JV_MCW(8) JV_MCW(16) JV_MCW(32)
#undef JV_MCW
    }
    JV_LAUNCH_CHECK("mc_cnn_win");
  } else if (make_cnn_geom(net, cg)) {
    McCnnLayout lay = mc_cnn_layout(cg);
    int nw = 4;
    while (nw > 1 && lay.shared + nw * lay.per_chain > 160 * 1024) nw /= 2;
    size_t lds = lay.shared + nw * lay.per_chain;
    int rc = set_lds_limit((const void*)mc_cnn, lds);
    if (rc) return rc;
    ScopedTimer tm("mc_sweep", st);
    hipLaunchKernelGGL(mc_cnn, dim3((unsigned)((a.num_chains + nw - 1) / nw)), dim3(64 * nw), lds, st, cg,
                       params, a, walkers, configs, counters);
    JV_LAUNCH_CHECK("mc_cnn");
  } else {
    return fail(JV_E_ARG, "jv_mc_sample: unknown network kind");
  }
  // log psi of the emitted samples from a fresh forward pass (sampler.py:293-296)
  int64_t ns = (int64_t)a.samples_per_chain * a.num_chains;
  if (logpsi && ns > 0) return jv_net_forward(net, params, configs, ns, logpsi, stream);
  return JV_OK;
}

int jv_exact_weights(const double* logpsi, int64_t n, double last_norm, double log_prob_factor, double* p,
                     double* norm_out, void* stream) {
  if (!logpsi || !p || n < 1 || log_prob_factor <= 0) return fail(JV_E_ARG, "jv_exact_weights: bad arguments");
  return launch_weights(logpsi, n, 1.0 / log_prob_factor, last_norm / log_prob_factor, p, norm_out,
                        (hipStream_t)stream);
}

int jv_mc_weights(const double* logpsi, int64_t n, double mu, double log_prob_factor, double* p, void* stream) {
  if (!logpsi || !p || n < 1 || log_prob_factor <= 0) return fail(JV_E_ARG, "jv_mc_weights: bad arguments");
  return launch_weights(logpsi, n, 1.0 / log_prob_factor - mu, 0.0, p, nullptr, (hipStream_t)stream);
}

}  // extern "C"
