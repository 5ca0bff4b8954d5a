// RBM device routines shared by the single-network kernels and the batched
// (many instances per launch) kernels: one Metropolis chain, one local
// estimator sample and one forward + O_k row, each run by one wavefront.
#pragma once
#include "nets.hpp"
#include "ops.hpp"

namespace jv {

// LDS doubles of one chain: theta (H complex), Re log cosh theta (H), walker (L int32).
__host__ __device__ inline size_t mc_rbm_lds_doubles(const RbmGeom& g) { return (size_t)3 * g.H + (g.L + 1) / 2; }

// Metropolis chain `chain` of MCSampler._get_samples_mcmc / _sweep
// (sampler.py:264-367) for an RBM.  theta_h and Re log cosh(theta_h) are
// cached in LDS and updated on acceptance; a proposal costs one log-cosh per
// hidden unit instead of the full network re-evaluation of sampler.py:335.
// theta is recomputed from scratch whenever a sample is emitted.  Returns the
// number of accepted proposals (valid in every lane).
__device__ inline int64_t mc_rbm_chain(const RbmGeom& g, const double* __restrict__ P, const jv_mc_args& a,
                                       int32_t* __restrict__ walkers, int32_t* __restrict__ configs, int chain,
                                       int lane, double* base) {
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
  return nacc;
}

// O_loc of one sample (base.py:162-192) for an RBM: theta and log cosh(theta)
// stay in LDS (th: 2H complex of the wave); each string with |m| > 1e-6 costs
// one complex log-cosh per hidden unit.  lps_s / out_s: complex log psi(s)
// given by the caller and the result.
__device__ inline void eloc_rbm_sample(const RbmGeom& g, const double* __restrict__ P, const OpDev& op,
                                       const int32_t* c, const double* lps_s, double* out_s, cd* th, int lane) {
  cd* lc = th + g.H;
  const double* Wr = P + g.ow;
  const double* Wi = P + g.ow + (int64_t)g.L * g.H;
  double lr = 0.0, li = 0.0;
  for (int h = lane; h < g.H; h += WAVE) {
    cd t = rbm_theta(g, P, c, h);
    cd l = g.cpx ? log_cosh(t) : cd{log_cosh(t.re), 0.0};
    th[h] = t;
    lc[h] = l;
    lr += l.re;
    li += l.im;
  }
  if (g.oa >= 0) {
    for (int i = lane; i < g.L; i += WAVE) {
      double sg = (double)(2 * c[i] - 1);
      lr = fma(sg, P[g.oa + i], lr);
      li = fma(sg, P[g.oa + g.L + i], li);
    }
  }
  lr = wave_sum(lr);
  li = wave_sum(li);
  // log psi_internal(s) - log psi_s given by the caller (base.py:164)
  cd off = {lr - lps_s[0], li - lps_s[1]};
  wave_sync();
  cd dsum = diag_sum(op, c);
  cd acc = {0.0, 0.0};
  int sites[MAXLEN], vals[MAXLEN], nch;
  for (int o = 0; o < op.num_ops; ++o) {
    cd m = merged_m(op, o, apply_string(op, o, c, sites, vals, nch), dsum);
    if (!nonzero(m)) continue;
    cd dl = {0.0, 0.0};
    if (nch > 0) {
      for (int h = lane; h < g.H; h += WAVE) {
        cd t = th[h];
        for (int j = 0; j < nch; ++j) {
          double d = 2.0 * (double)(vals[j] - c[sites[j]]);
          int64_t wi = (int64_t)sites[j] * g.H + h;
          t.re = fma(d, Wr[wi], t.re);
          if (g.cpx) t.im = fma(d, Wi[wi], t.im);
        }
        cd l = g.cpx ? log_cosh(t) : cd{log_cosh(t.re), 0.0};
        dl.re += l.re - lc[h].re;
        dl.im += l.im - lc[h].im;
      }
      dl.re = wave_sum(dl.re);
      dl.im = wave_sum(dl.im);
      if (g.oa >= 0) {
        for (int j = 0; j < nch; ++j) {
          double d = 2.0 * (double)(vals[j] - c[sites[j]]);
          dl.re = fma(d, P[g.oa + sites[j]], dl.re);
          dl.im = fma(d, P[g.oa + g.L + sites[j]], dl.im);
        }
      }
    }
    acc = acc + m * cexp(dl + off);
  }
  if (lane == 0) {
    out_s[0] = acc.re;
    out_s[1] = acc.im;
  }
  wave_sync();
}

// Doubles per gradient row: compact (complex or real) or jVMC layout (complex).
__host__ __device__ inline int64_t rbm_grad_row_doubles(const RbmGeom& g, int layout) {
  const int64_t Pc = (g.ob >= 0 ? g.H : 0) + (int64_t)g.L * g.H + (g.oa >= 0 ? g.L : 0);
  return layout == 0 ? (g.cpx ? 2 * Pc : Pc) : 2 * (g.cpx ? 2 * Pc : Pc);
}

// log psi and O_k of one configuration (rbm.py:30-38, 80-88; vqs.py:46-69);
// lanes own hidden units h = lane + 64 c.  row (may be null): the gradient row
// in layout 0 (compact) or 1 (jVMC [g, i g] per complex leaf); logpsi_s (may
// be null): complex log psi.
__device__ inline void rbm_fwd_grad_sample(const RbmGeom& g, const double* __restrict__ P, const int32_t* c,
                                           int layout, double* row, double* logpsi_s, int lane) {
  // column writers: compact (complex or real) or jVMC layout (complex; [g, i g] per complex leaf)
  auto put = [&](int64_t leaf_off_flat, int64_t leaf_off_compact, int64_t leaf_size, int64_t j, cd v) {
    if (layout == 0) {
      if (g.cpx) {
        row[2 * (leaf_off_compact + j)] = v.re;
        row[2 * (leaf_off_compact + j) + 1] = v.im;
      } else {
        row[leaf_off_compact + j] = v.re;
      }
    } else {
      int64_t col = leaf_off_flat + j;
      row[2 * col] = v.re;
      row[2 * col + 1] = v.im;
      if (g.cpx) {  // i * g
        int64_t col2 = leaf_off_flat + leaf_size + j;
        row[2 * col2] = -v.im;
        row[2 * col2 + 1] = v.re;
      }
    }
  };
  const int64_t cb = 0, cw = (g.ob >= 0 ? g.H : 0), ca = cw + (int64_t)g.L * g.H;
  double accr = 0.0, acci = 0.0;
  for (int h = lane; h < g.H; h += WAVE) {
    cd th = rbm_theta(g, P, c, h);
    if (g.cpx) {
      cd lc = log_cosh(th);
      accr += lc.re;
      acci += lc.im;
    } else {
      accr += log_cosh(th.re);
    }
    if (row) {
      cd t = g.cpx ? ctanh(th) : cd{tanh(th.re), 0.0};
      if (g.ob >= 0) put(g.ob, cb, g.H, h, t);
      for (int i = 0; i < g.L; ++i) {
        double sg = (double)(2 * c[i] - 1);
        put(g.ow, cw, (int64_t)g.L * g.H, (int64_t)i * g.H + h, {sg * t.re, sg * t.im});
      }
    }
  }
  if (g.oa >= 0) {
    for (int i = lane; i < g.L; i += WAVE) {
      double sg = (double)(2 * c[i] - 1);
      accr = fma(sg, P[g.oa + i], accr);
      acci = fma(sg, P[g.oa + g.L + i], acci);
      if (row) put(g.oa, ca, g.L, i, {sg, 0.0});
    }
  }
  accr = wave_sum(accr);
  acci = wave_sum(acci);
  if (logpsi_s && lane == 0) {
    logpsi_s[0] = accr;
    logpsi_s[1] = acci;
  }
}

}  // namespace jv
