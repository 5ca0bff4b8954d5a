// Symmetrised networks (jVMC's SymNet, sym_wrapper.py:29-83) on the device.
//
// log psi(s) = avgFun({log psi_net(g s)}_g, factors) over the orbit of lattice
// symmetries.  The orbit is applied as a signed site map per element (see
// jv_orbit), the inner network runs unchanged on the |orbit| transformed rows of
// each sample, and a one-wave-per-sample kernel forms the average and the
// weights w_g = d avg / d c_g.  Gradients are sum_g w_g O_net(g s), combined
// column-parallel with the orbit summed in a fixed order (deterministic).
//
// Work is chunked over samples so the transformed rows and their inner
// gradients stay within fixed workspace budgets.
#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "nets.hpp"

namespace jv {

namespace {
// workspace slots of this file.  jv_sym_eloc may run on a side stream while
// jv_sym_gradients runs on the main stream, so the forward pass inside
// jv_sym_eloc has rows / log-psi slots of its own (WS_EROWS, WS_ELP).
constexpr int WS_ROWS = 70, WS_LP = 71, WS_EROWS = 72, WS_ELP = 73, WS_SP = 74, WS_ME = 75, WS_CNT = 76,
              WS_LPSP = 77, WS_GI = 78, WS_W = 79, WS_GLO = 80, WS_MC_LP = 81, WS_MC_PROP = 82, WS_MC_U = 83,
              WS_MC_NACC = 84, WS_MC_LPP = 85;
constexpr size_t ROW_BYTES_MAX = size_t(256) << 20;  // transformed configurations per chunk
constexpr size_t GRAD_BYTES_MAX = size_t(1) << 30;   // inner gradient rows per chunk
}  // namespace

// Row r = s * G + g of out holds g applied to sample s.
__global__ void __launch_bounds__(256) sym_expand(const int32_t* __restrict__ cfg, int64_t ns,
                                                  const int32_t* __restrict__ map, int G, int L,
                                                  int32_t* __restrict__ out) {
  const int64_t tot = ns * G * L;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / L;
    const int i = (int)(t - r * L);
    const int64_t s = r / G;
    const int g = (int)(r - s * G);
    const int m = map[(int64_t)g * L + i];
    const int32_t* c = cfg + s * L;
    out[t] = m >= 0 ? c[m] : 1 - c[-1 - m];
  }
}

__device__ inline double wave_max(double v) {
  for (int o = WAVE / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// One wave per sample: out[s] = avg_g c[s*G + g], W[s][g] = d out / d c (complex).
// cpx: the coefficients or the factors are complex (jax promotes both).
__global__ void __launch_bounds__(256) sym_avg(const double* __restrict__ lp, const double* __restrict__ fac,
                                               int64_t ns, int G, int avg, int cpx, double* __restrict__ out,
                                               double* __restrict__ W) {
  const int lane = threadIdx.x % WAVE;
  const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / WAVE);
  for (int64_t s = blockIdx.x * (int64_t)(blockDim.x / WAVE) + threadIdx.x / WAVE; s < ns; s += wstride) {
    const double* a = lp + 2 * s * G;
    double* w = W ? W + 2 * s * G : nullptr;
    double ore = 0.0, oim = 0.0;
    if (avg == JV_AVG_EXP) {
      // logsumexp(a, b): a_g -> -inf where b_g = 0, shift by the largest real part
      double amax = -INFINITY;
      for (int g = lane; g < G; g += WAVE)
        if (fac[2 * g] != 0.0 || fac[2 * g + 1] != 0.0) amax = fmax(amax, a[2 * g]);
      amax = wave_max(amax);
      if (!isfinite(amax)) amax = 0.0;
      double sr = 0.0, si = 0.0;
      for (int g = lane; g < G; g += WAVE) {
        const cd b = {fac[2 * g], fac[2 * g + 1]};
        if (b.re == 0.0 && b.im == 0.0) continue;
        const double e = exp(a[2 * g] - amax);
        const cd t = cpx ? cd{e * cos(a[2 * g + 1]), e * sin(a[2 * g + 1])} * b : cd{e * b.re, 0.0};
        sr += t.re;
        si += t.im;
      }
      sr = wave_sum(sr);
      si = wave_sum(si);
      if (cpx) {
        ore = log(hypot(sr, si)) + amax;
        oim = atan2(si, sr);
      } else {
        ore = sr < 0.0 ? NAN : log(fabs(sr)) + amax;
      }
      if (w) {
        const double n2 = sr * sr + si * si;
        for (int g = lane; g < G; g += WAVE) {
          const cd b = {fac[2 * g], fac[2 * g + 1]};
          cd t = {0.0, 0.0};
          if (b.re != 0.0 || b.im != 0.0) {
            const double e = exp(a[2 * g] - amax);
            t = cpx ? cd{e * cos(a[2 * g + 1]), e * sin(a[2 * g + 1])} * b : cd{e * b.re, 0.0};
          }
          cd q;
          if (cpx) q = {(t.re * sr + t.im * si) / n2, (t.im * sr - t.re * si) / n2};  // t / S
          else q = {sr < 0.0 ? 0.0 : t.re / sr, 0.0};
          w[2 * g] = q.re;
          w[2 * g + 1] = q.im;
        }
      }
    } else if (avg == JV_AVG_LOG) {
      double sr = 0.0, si = 0.0;
      for (int g = lane; g < G; g += WAVE) {
        sr += a[2 * g];
        si += a[2 * g + 1];
      }
      ore = wave_sum(sr) / G;
      oim = cpx ? wave_sum(si) / G : 0.0;
      if (w)
        for (int g = lane; g < G; g += WAVE) {
          w[2 * g] = 1.0 / G;
          w[2 * g + 1] = 0.0;
        }
    } else {  // JV_AVG_SEP
      double se = 0.0, sc = 0.0, ss = 0.0;
      for (int g = lane; g < G; g += WAVE) {
        se += exp(2.0 * a[2 * g]);
        sc += cos(a[2 * g + 1]);
        ss += sin(a[2 * g + 1]);
      }
      se = wave_sum(se);
      sc = wave_sum(sc);
      ss = wave_sum(ss);
      ore = 0.5 * log(se / G);
      oim = atan2(ss / G, sc / G);
      if (w)
        for (int g = lane; g < G; g += WAVE) {
          w[2 * g] = exp(2.0 * a[2 * g]) / se;
          w[2 * g + 1] = 0.0;
        }
    }
    if (lane == 0) {
      out[2 * s] = ore;
      out[2 * s + 1] = oim;
    }
  }
}

// out[s][k] = sum_g W[s][g] in[s*G + g][k] over columns k < K (complex W).
template <bool IN_CPX, bool OUT_CPX>
__global__ void __launch_bounds__(256) sym_combine(const double* __restrict__ in, const double* __restrict__ W,
                                                   int64_t ns, int G, int64_t K, double* __restrict__ out) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= K) return;
  for (int64_t s = blockIdx.y; s < ns; s += gridDim.y) {
    double ar = 0.0, ai = 0.0;
    for (int g = 0; g < G; ++g) {
      const int64_t r = s * G + g;
      const double wr = W[2 * r], wi = W[2 * r + 1];
      if (IN_CPX) {
        const double xr = in[2 * (r * K + k)], xi = in[2 * (r * K + k) + 1];
        ar = fma(wr, xr, fma(-wi, xi, ar));
        ai = fma(wr, xi, fma(wi, xr, ai));
      } else {
        const double x = in[r * K + k];
        ar = fma(wr, x, ar);
        ai = fma(wi, x, ai);
      }
    }
    if (OUT_CPX) {
      out[2 * (s * K + k)] = ar;
      out[2 * (s * K + k) + 1] = ai;
    } else {
      out[s * K + k] = ar;
    }
  }
}

// O_loc[s] = sum_j m[s][j] exp(lp'[s][j] - lp[s]) over all nops slots (base.py:162-164).
__global__ void __launch_bounds__(256) sym_oloc(const double* __restrict__ me, const double* __restrict__ lpsp,
                                                const double* __restrict__ lps, int64_t ns, int nops,
                                                double* __restrict__ out) {
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const cd y = {lps[2 * s], lps[2 * s + 1]};
  double ar = 0.0, ai = 0.0;
  for (int j = 0; j < nops; ++j) {
    const int64_t q = s * nops + j;
    const cd m = {me[2 * q], me[2 * q + 1]};
    const cd t = m * cexp(cd{lpsp[2 * q], lpsp[2 * q + 1]} - y);
    ar += t.re;
    ai += t.im;
  }
  out[2 * s] = ar;
  out[2 * s + 1] = ai;
}

// ---------------------------------------------------------------- Metropolis
// Proposal of step `step` for every chain: the walker with one site flipped,
// drawn with the same Philox counters as the single-network chain kernels.
__global__ void __launch_bounds__(256) sym_mc_propose(const int32_t* __restrict__ walkers, int C, int L,
                                                      jv_mc_args a, int step, int32_t* __restrict__ prop,
                                                      double* __restrict__ u) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)C * L) return;
  const int c = (int)(t / L), i = (int)(t % L);
  int site;
  double uu;
  mc_draw(a.seed, (uint32_t)step, (uint32_t)c, a.call_index, a.stream, L, site, uu);
  const int32_t v = walkers[t];
  prop[t] = i == site ? (v ^ 1) : v;
  if (i == 0) u[c] = uu;
}

// Accept / reject (sampler.py:334-348) and emission of the walker after every
// sweep_steps proposals past thermalisation.  lacc holds mu Re log psi.
__global__ void __launch_bounds__(64) sym_mc_accept(int32_t* __restrict__ walkers, const int32_t* __restrict__ prop,
                                                    double* __restrict__ lacc, const double* __restrict__ lpp,
                                                    const double* __restrict__ u, int C, int L, jv_mc_args a,
                                                    int step, int32_t* __restrict__ configs,
                                                    int64_t* __restrict__ nacc) {
  const int c = blockIdx.x;
  if (c >= C) return;
  const double lnew = a.mu * lpp[2 * c];
  const bool acc = u[c] < exp(lnew - lacc[c]);
  int32_t* wk = walkers + (int64_t)c * L;
  if (acc) {
    for (int i = threadIdx.x; i < L; i += blockDim.x) wk[i] = prop[(int64_t)c * L + i];
    if (threadIdx.x == 0) {
      lacc[c] = lnew;
      nacc[c] += 1;
    }
  }
  const int k = step + 1 - a.therm_steps;
  if (k > 0 && k % a.sweep_steps == 0) {
    const int64_t t = k / a.sweep_steps - 1;
    int32_t* out = configs + (t * C + c) * L;
    for (int i = threadIdx.x; i < L; i += blockDim.x) out[i] = acc ? prop[(int64_t)c * L + i] : wk[i];
  }
}

__global__ void sym_mc_init(const double* __restrict__ lp, int C, double mu, double* __restrict__ lacc,
                            int64_t* __restrict__ nacc) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  lacc[c] = mu * lp[2 * c];
  nacc[c] = 0;
}

// counters = (proposed, accepted), the accepted count summed in chain order.
__global__ void __launch_bounds__(64) sym_mc_count(const int64_t* __restrict__ nacc, int C, int64_t proposed,
                                                   int64_t* __restrict__ counters) {
  if (threadIdx.x != 0) return;
  int64_t s = 0;
  for (int c = 0; c < C; ++c) s += nacc[c];
  counters[0] = proposed;
  counters[1] = s;
}

namespace {

int check_orbit(const jv_net* net, const jv_orbit* o, const char* who) {
  if (!net || !o) return fail(JV_E_ARG, std::string(who) + ": null descriptor");
  if (o->num_elements < 1 || !o->map || !o->factor)
    return fail(JV_E_ARG, std::string(who) + ": empty orbit");
  if (o->num_sites != net->num_sites)
    return fail(JV_E_ARG, std::string(who) + ": orbit acts on a different number of sites than the network");
  if (o->avg < JV_AVG_EXP || o->avg > JV_AVG_SEP) return fail(JV_E_ARG, std::string(who) + ": unknown avg");
  return JV_OK;
}

// Complex coefficient / weight arithmetic (jax promotes real network outputs
// when the factors are complex).
bool sym_cpx(const jv_net* net, const jv_orbit* o) {
  return net->is_complex || (o->avg == JV_AVG_EXP && o->complex_factor);
}

int64_t rows_per_chunk(const jv_net* net, const jv_orbit* o, size_t extra_row_bytes) {
  const size_t rb = sizeof(int32_t) * (size_t)net->num_sites;
  int64_t rows = (int64_t)(ROW_BYTES_MAX / rb);
  if (extra_row_bytes) rows = std::min<int64_t>(rows, (int64_t)(GRAD_BYTES_MAX / extra_row_bytes));
  return std::max<int64_t>(1, rows / o->num_elements);  // samples per chunk
}

int launch_avg(const double* lp, const jv_orbit* o, int64_t ns, int cpx, double* out, double* W, hipStream_t st) {
  const int64_t blocks = std::min<int64_t>((ns + 3) / 4, 8192);
  hipLaunchKernelGGL(sym_avg, dim3((unsigned)blocks), dim3(256), 0, st, lp, o->factor, ns, o->num_elements, o->avg,
                     cpx, out, W);
  JV_LAUNCH_CHECK("sym_avg");
  return JV_OK;
}

int launch_expand(const int32_t* cfg, int64_t ns, const jv_orbit* o, int L, int32_t* rows, hipStream_t st) {
  const int64_t tot = ns * o->num_elements * L;
  const int64_t blocks = std::min<int64_t>((tot + 255) / 256, 1 << 16);
  hipLaunchKernelGGL(sym_expand, dim3((unsigned)blocks), dim3(256), 0, st, cfg, ns, o->map, o->num_elements, L, rows);
  JV_LAUNCH_CHECK("sym_expand");
  return JV_OK;
}

// Symmetrised log-amplitudes with the expanded rows and per-element log-psi
// in workspace slots ws_rows / ws_lp.
int sym_forward(const jv_net* net, const double* params, const jv_orbit* orbit, const int32_t* configs, int64_t n,
                double* logpsi, void* stream, int ws_rows, int ws_lp) {
  int rc = check_orbit(net, orbit, "jv_sym_forward");
  if (rc) return rc;
  if (!params || !configs || !logpsi || n < 0) return fail(JV_E_ARG, "jv_sym_forward: bad arguments");
  if (n == 0) return JV_OK;
  hipStream_t st = (hipStream_t)stream;
  const int L = net->num_sites, G = orbit->num_elements;
  const int64_t cs = std::min<int64_t>(n, rows_per_chunk(net, orbit, 0));
  int32_t* rows = (int32_t*)workspace(sizeof(int32_t) * (size_t)cs * G * L, ws_rows);
  double* lp = (double*)workspace(sizeof(double) * 2 * (size_t)cs * G, ws_lp);
  if (!rows || !lp) return fail(JV_E_NOMEM, "jv_sym_forward: workspace allocation failed");
  const int cpx = sym_cpx(net, orbit);
  for (int64_t s0 = 0; s0 < n; s0 += cs) {
    const int64_t ns = std::min(cs, n - s0);
    if ((rc = launch_expand(configs + s0 * L, ns, orbit, L, rows, st))) return rc;
    if ((rc = jv_net_forward(net, params, rows, ns * G, lp, stream))) return rc;
    if ((rc = launch_avg(lp, orbit, ns, cpx, logpsi + 2 * s0, nullptr, st))) return rc;
  }
  return JV_OK;
}

}  // namespace
}  // namespace jv

using namespace jv;

extern "C" {

int jv_sym_forward(const jv_net* net, const double* params, const jv_orbit* orbit, const int32_t* configs,
                   int64_t n, double* logpsi, void* stream) {
  return sym_forward(net, params, orbit, configs, n, logpsi, stream, WS_ROWS, WS_LP);
}

int jv_sym_gradients(const jv_net* net, const double* params, const jv_orbit* orbit, const int32_t* configs,
                     int64_t n, int32_t layout, double* grad, double* logpsi, void* stream) {
  int rc = check_orbit(net, orbit, "jv_sym_gradients");
  if (rc) return rc;
  if (!params || !configs || !grad || n < 0) return fail(JV_E_ARG, "jv_sym_gradients: bad arguments");
  if (layout != 0 && layout != 1) return fail(JV_E_ARG, "layout must be 0 or 1");
  if (orbit->avg == JV_AVG_SEP && net->is_complex)
    return fail(JV_E_ARG, "jv_sym_gradients: avgFun_Coefficients_Sep makes a complex-parameter network "
                          "non-holomorphic; its gradients are not implemented");
  if (n == 0) return JV_OK;
  hipStream_t st = (hipStream_t)stream;
  const int L = net->num_sites, G = orbit->num_elements;
  const int cpx = sym_cpx(net, orbit);
  // inner rows: layout 1 complex [num_real_params]; layout 0 complex (CpxRBM) or real [num_params]
  const int64_t K = layout == 1 ? net->num_real_params : net->num_params;
  const bool in_cpx = layout == 1 || net->is_complex;
  const bool out_cpx = layout == 1 || net->is_complex || cpx;
  const size_t in_row = sizeof(double) * (in_cpx ? 2 : 1) * (size_t)K;
  const int64_t cs = std::min<int64_t>(n, rows_per_chunk(net, orbit, in_row));
  int32_t* rows = (int32_t*)workspace(sizeof(int32_t) * (size_t)cs * G * L, WS_ROWS);
  double* lp = (double*)workspace(sizeof(double) * 2 * (size_t)cs * G, WS_LP);
  double* gi = (double*)workspace(in_row * (size_t)cs * G, WS_GI);
  double* W = (double*)workspace(sizeof(double) * 2 * (size_t)cs * G, WS_W);
  double* lo = logpsi ? nullptr : (double*)workspace(sizeof(double) * 2 * (size_t)cs, WS_GLO);
  if (!rows || !lp || !gi || !W || (!logpsi && !lo)) return fail(JV_E_NOMEM, "jv_sym_gradients: workspace allocation failed");
  for (int64_t s0 = 0; s0 < n; s0 += cs) {
    const int64_t ns = std::min(cs, n - s0);
    if ((rc = launch_expand(configs + s0 * L, ns, orbit, L, rows, st))) return rc;
    if ((rc = jv_net_gradients(net, params, rows, ns * G, layout, gi, lp, stream))) return rc;
    if ((rc = launch_avg(lp, orbit, ns, cpx, logpsi ? logpsi + 2 * s0 : lo, W, st))) return rc;
    const size_t out_row = (out_cpx ? 2 : 1) * (size_t)K;
    dim3 grid((unsigned)((K + 255) / 256), (unsigned)std::min<int64_t>(ns, 65535));
    double* o = grad + (size_t)s0 * out_row;
    ScopedTimer tm("sym_combine", st);
    if (in_cpx)
      hipLaunchKernelGGL((sym_combine<true, true>), grid, dim3(256), 0, st, gi, W, ns, G, K, o);
    else if (out_cpx)
      hipLaunchKernelGGL((sym_combine<false, true>), grid, dim3(256), 0, st, gi, W, ns, G, K, o);
    else
      hipLaunchKernelGGL((sym_combine<false, false>), grid, dim3(256), 0, st, gi, W, ns, G, K, o);
    JV_LAUNCH_CHECK("sym_combine");
  }
  return JV_OK;
}

int jv_sym_eloc(const jv_net* net, const double* params, const jv_orbit* orbit, const jv_optable* op,
                const double* prefactor, const int32_t* configs, const double* logpsi_s, int64_t n, double* out,
                void* stream) {
  int rc = check_orbit(net, orbit, "jv_sym_eloc");
  if (rc) return rc;
  if (!op || !params || !configs || !logpsi_s || !out || n < 0) return fail(JV_E_ARG, "jv_sym_eloc: bad arguments");
  if (n == 0) return JV_OK;
  if (op->num_ops < 1) return fail(JV_E_ARG, "jv_sym_eloc: empty operator");
  hipStream_t st = (hipStream_t)stream;
  const int L = net->num_sites, nops = op->num_ops;
  // s' of cs samples are cs * nops configurations, each expanded |orbit| times by jv_sym_forward
  const int64_t cs = std::max<int64_t>(1, std::min<int64_t>(n, rows_per_chunk(net, orbit, 0) * 4 / nops));
  int32_t* sp = (int32_t*)workspace(sizeof(int32_t) * (size_t)cs * nops * L, WS_SP);
  double* me = (double*)workspace(sizeof(double) * 2 * (size_t)cs * nops, WS_ME);
  int32_t* cnt = (int32_t*)workspace(sizeof(int32_t) * (size_t)cs, WS_CNT);
  double* lpsp = (double*)workspace(sizeof(double) * 2 * (size_t)cs * nops, WS_LPSP);
  if (!sp || !me || !cnt || !lpsp) return fail(JV_E_NOMEM, "jv_sym_eloc: workspace allocation failed");
  ScopedTimer tm("sym_eloc", st);
  for (int64_t s0 = 0; s0 < n; s0 += cs) {
    const int64_t ns = std::min(cs, n - s0);
    if ((rc = jv_get_s_primes(op, prefactor, configs + s0 * L, ns, L, sp, me, cnt, stream))) return rc;
    if ((rc = sym_forward(net, params, orbit, sp, ns * nops, lpsp, stream, WS_EROWS, WS_ELP))) return rc;
    hipLaunchKernelGGL(sym_oloc, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, st, me, lpsp,
                       logpsi_s + 2 * s0, ns, nops, out + 2 * s0);
    JV_LAUNCH_CHECK("sym_oloc");
  }
  return JV_OK;
}

int jv_sym_mc_sample(const jv_net* net, const double* params, const jv_orbit* orbit, const jv_mc_args* args,
                     int32_t* walkers, int32_t* configs, double* logpsi, int64_t* counters, void* stream) {
  int rc = check_orbit(net, orbit, "jv_sym_mc_sample");
  if (rc) return rc;
  if (!params || !args || !walkers || !counters) return fail(JV_E_ARG, "jv_sym_mc_sample: bad arguments");
  const jv_mc_args a = *args;
  if (a.num_chains < 1 || a.sweep_steps < 1 || a.therm_steps < 0 || a.samples_per_chain < 0)
    return fail(JV_E_ARG, "jv_sym_mc_sample: bad chain / sweep counts");
  if (a.samples_per_chain > 0 && !configs) return fail(JV_E_ARG, "jv_sym_mc_sample: null configs");
  if (a.mu < 0.0 || a.mu > 2.0) return fail(JV_E_ARG, "mu must be in the range [0, 2]");
  hipStream_t st = (hipStream_t)stream;
  const int C = a.num_chains, L = net->num_sites;
  double* lacc = (double*)workspace(sizeof(double) * (size_t)C, WS_MC_LP);
  int32_t* prop = (int32_t*)workspace(sizeof(int32_t) * (size_t)C * L, WS_MC_PROP);
  double* u = (double*)workspace(sizeof(double) * (size_t)C, WS_MC_U);
  int64_t* nacc = (int64_t*)workspace(sizeof(int64_t) * (size_t)C, WS_MC_NACC);
  double* lpp = (double*)workspace(sizeof(double) * 2 * (size_t)C, WS_MC_LPP);
  if (!lacc || !prop || !u || !nacc || !lpp) return fail(JV_E_NOMEM, "jv_sym_mc_sample: workspace allocation failed");
  ScopedTimer tm("sym_mc", st);
  if ((rc = jv_sym_forward(net, params, orbit, walkers, C, lpp, stream))) return rc;
  hipLaunchKernelGGL(sym_mc_init, dim3((unsigned)((C + 255) / 256)), dim3(256), 0, st, lpp, C, a.mu, lacc, nacc);
  JV_LAUNCH_CHECK("sym_mc_init");
  const int64_t total = (int64_t)a.therm_steps + (int64_t)a.samples_per_chain * a.sweep_steps;
  if (total > INT32_MAX) return fail(JV_E_ARG, "jv_sym_mc_sample: too many steps");
  const unsigned pblocks = (unsigned)(((int64_t)C * L + 255) / 256);
  for (int step = 0; step < (int)total; ++step) {
    hipLaunchKernelGGL(sym_mc_propose, dim3(pblocks), dim3(256), 0, st, walkers, C, L, a, step, prop, u);
    JV_LAUNCH_CHECK("sym_mc_propose");
    if ((rc = jv_sym_forward(net, params, orbit, prop, C, lpp, stream))) return rc;
    hipLaunchKernelGGL(sym_mc_accept, dim3((unsigned)C), dim3(64), 0, st, walkers, prop, lacc, lpp, u, C, L, a,
                       step, configs, nacc);
    JV_LAUNCH_CHECK("sym_mc_accept");
  }
  hipLaunchKernelGGL(sym_mc_count, dim3(1), dim3(64), 0, st, nacc, C, total * C, counters);
  JV_LAUNCH_CHECK("sym_mc_count");
  const int64_t ns = (int64_t)a.samples_per_chain * C;
  if (logpsi && ns > 0) return jv_sym_forward(net, params, orbit, configs, ns, logpsi, stream);
  return JV_OK;
}

}  // extern "C"
