// Operator strings: s' enumeration with nonzero compaction, and the fused local
// estimator O_loc(s) = sum_j m_j psi(s'_j)/psi(s) with log psi(s') obtained by
// incremental (RBM) or receptive-field-limited (CNN) re-evaluation.
#include <algorithm>

#include "nets.hpp"

namespace jv {

bool make_rbm_geom(const jv_net* net, RbmGeom& g);
bool make_cnn_geom(const jv_net* net, CnnGeom& g);
int set_lds_limit(const void* fn, size_t bytes);

constexpr int MAXLEN = 8;  // longest operator string handled in-kernel

struct OpDev {
  int num_ops, max_len, num_diag;
  int L;  // sites per configuration
  const int32_t* idx;
  const int32_t* map;
  const double* matel;
  const int32_t* diag;
  const double* pref;
};

// Applies string o to s (branch_free.py:456-478, factors in stored order) and
// returns m with the sites whose final value differs from s.
__device__ inline cd apply_string(const OpDev& op, int o, const int32_t* s, int* sites, int* vals,
                                  int& nch) {
  cd m = {op.pref[2 * o], op.pref[2 * o + 1]};
  int ts[MAXLEN], tv[MAXLEN];
  int nt = 0;
  for (int k = 0; k < op.max_len; ++k) {
    const int base = o * op.max_len + k;
    // indices are clamped like JAX gathers clamp (site to [0, L), state to [0, 1])
    const int i = min(max(op.idx[base], 0), op.L - 1);
    int slot = -1;
    for (int t = 0; t < nt; ++t)
      if (ts[t] == i) slot = t;
    const int cur = min(max(slot >= 0 ? tv[slot] : s[i], 0), 1);
    cd f = {op.matel[2 * (2 * base + cur)], op.matel[2 * (2 * base + cur) + 1]};
    m = m * f;
    const int nv = min(max(op.map[2 * base + cur], 0), 1);
    if (slot >= 0) tv[slot] = nv;
    else { ts[nt] = i; tv[nt] = nv; ++nt; }
  }
  nch = 0;
  for (int t = 0; t < nt; ++t)
    if (tv[t] != s[ts[t]]) { sites[nch] = ts[t]; vals[nch] = tv[t]; ++nch; }
  return m;
}

__device__ inline bool is_diag(const OpDev& op, int o) {
  for (int d = 0; d < op.num_diag; ++d)
    if (op.diag[d] == o) return true;
  return false;
}

// Matrix element of slot o after the diagonal merge (branch_free.py:483-485):
// with more than one diagonal string, slot diag[0] carries their sum and the
// other diagonal slots are zero.
__device__ inline cd merged_m(const OpDev& op, int o, cd m_raw, cd diag_sum) {
  if (op.num_diag > 1 && is_diag(op, o)) return o == op.diag[0] ? diag_sum : cd{0.0, 0.0};
  return m_raw;
}

__device__ inline cd diag_sum(const OpDev& op, const int32_t* s) {
  cd acc = {0.0, 0.0};
  int sites[MAXLEN], vals[MAXLEN], nch;
  for (int d = 0; d < op.num_diag; ++d) acc = acc + apply_string(op, op.diag[d], s, sites, vals, nch);
  return acc;
}

__device__ inline bool nonzero(cd m) { return sqrt(cabs2(m)) > 1e-6; }

// ------------------------------------------------------------------ s' enumeration
// One thread per sample: merged matrix elements, compaction order (base.py:91-114).
__global__ void sprime_compact(OpDev op, const int32_t* __restrict__ cfg, int64_t n, int L,
                               double* __restrict__ matel, int32_t* __restrict__ slot_op,
                               int32_t* __restrict__ count) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int32_t* c = cfg + s * L;
  cd dsum = diag_sum(op, c);
  int sites[MAXLEN], vals[MAXLEN], nch;
  int cnt = 0;
  double* mrow = matel + 2 * s * op.num_ops;
  int32_t* orow = slot_op + s * op.num_ops;
  for (int o = 0; o < op.num_ops; ++o) {
    cd m = merged_m(op, o, apply_string(op, o, c, sites, vals, nch), dsum);
    if (nonzero(m)) {
      mrow[2 * cnt] = m.re;
      mrow[2 * cnt + 1] = m.im;
      orow[cnt] = o;
      ++cnt;
    }
  }
  for (int j = cnt; j < op.num_ops; ++j) {
    mrow[2 * j] = 0.0;
    mrow[2 * j + 1] = 0.0;
    orow[j] = op.num_ops - 1;
  }
  count[s] = cnt;
}

// One thread per (sample, slot): write s' of the slot's string.
__global__ void sprime_write(OpDev op, const int32_t* __restrict__ cfg, int64_t n, int L,
                             const int32_t* __restrict__ slot_op, int32_t* __restrict__ sp) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * op.num_ops) return;
  int64_t s = t / op.num_ops;
  const int32_t* c = cfg + s * L;
  int sites[MAXLEN], vals[MAXLEN], nch;
  apply_string(op, slot_op[t], c, sites, vals, nch);
  int32_t* out = sp + t * L;
  for (int i = 0; i < L; ++i) out[i] = c[i];
  for (int j = 0; j < nch; ++j) out[sites[j]] = vals[j];
}

// ------------------------------------------------------------------ RBM O_loc
// One wavefront per sample.  theta and log cosh(theta) stay in LDS; each string
// with |m| > 1e-6 costs one complex log-cosh per hidden unit.
__global__ void __launch_bounds__(256) eloc_rbm(RbmGeom g, const double* __restrict__ P, OpDev op,
                                                const int32_t* __restrict__ cfg,
                                                const double* __restrict__ lps, int64_t n,
                                                double* __restrict__ out) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
  const int nw = blockDim.x / WAVE;
  cd* th = (cd*)lds + (size_t)w * 2 * g.H;
  cd* lc = th + g.H;
  const double* Wr = P + g.ow;
  const double* Wi = P + g.ow + (int64_t)g.L * g.H;
  for (int64_t s = (int64_t)blockIdx.x * nw + w; s < n; s += (int64_t)gridDim.x * nw) {
    const int32_t* c = cfg + s * g.L;
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
    cd off = {lr - lps[2 * s], li - lps[2 * s + 1]};
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
      out[2 * s] = acc.re;
      out[2 * s + 1] = acc.im;
    }
    wave_sync();
  }
}

// ------------------------------------------------------------------ CNN O_loc
// One workgroup per sample: block-wide forward of s into LDS, then each wave
// takes strings round-robin and re-evaluates only the receptive field of the
// changed sites.
struct EllocCnnLayout {
  size_t shared, per_wave;
};

__host__ __device__ inline EllocCnnLayout eloc_cnn_layout(const CnnGeom& g, bool keep_last) {
  EllocCnnLayout L;
  // shared: P, x0, A (doubles); cfg (int); red (8 doubles); nbr + inv (shorts)
  L.shared = 8 * ((size_t)g.P + g.n + g.atot + 8) + 4 * (size_t)g.n + 2 * 2 * (size_t)g.n * g.FF;
  L.shared = (L.shared + 15) & ~(size_t)15;
  int anew = keep_last ? g.atot : g.aoff[g.nl - 1];
  // per wave: xnew (n doubles) + Anew (doubles) + list (n shorts) + mask ((nl+1) n bytes)
  L.per_wave = 8 * ((size_t)g.n + (anew > 0 ? anew : 1)) + 2 * (size_t)g.n + (size_t)(g.nl + 1) * g.n;
  L.per_wave = (L.per_wave + 15) & ~(size_t)15;
  return L;
}

__device__ inline LocalScratch cnn_scratch(const CnnGeom& g, char* base, bool keep_last) {
  LocalScratch sc;
  int anew = keep_last ? g.atot : g.aoff[g.nl - 1];
  sc.xnew = (double*)base;
  sc.Anew = sc.xnew + g.n;
  sc.list = (short*)(sc.Anew + (anew > 0 ? anew : 1));
  sc.mask = (unsigned char*)(sc.list + g.n);
  return sc;
}

__global__ void __launch_bounds__(512) eloc_cnn(CnnGeom g, const double* __restrict__ Pg, OpDev op,
                                                const int32_t* __restrict__ cfg,
                                                const double* __restrict__ lps, int64_t n,
                                                double* __restrict__ out) {
  extern __shared__ double lds[];
  const EllocCnnLayout lay = eloc_cnn_layout(g, false);
  double* P = lds;
  double* x0 = P + g.P;
  double* A = x0 + g.n;
  double* red = A + g.atot;
  int32_t* cs = (int32_t*)(red + 8);
  short* nbr = (short*)(cs + g.n);
  short* inv = nbr + g.n * g.FF;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid % WAVE, w = tid / WAVE, nw = nt / WAVE;
  LocalScratch sc = cnn_scratch(g, (char*)lds + lay.shared + (size_t)w * lay.per_wave, false);
  for (int i = tid; i < g.P; i += nt) P[i] = Pg[i];
  cnn_build_tables(g, nbr, inv, tid, nt);
  __syncthreads();
  auto sync = [] { __syncthreads(); };
  for (int64_t s = blockIdx.x; s < n; s += gridDim.x) {
    const int32_t* c = cfg + s * g.n;
    for (int i = tid; i < g.n; i += nt) {
      cs[i] = c[i];
      x0[i] = (double)(2 * c[i] - 1);
    }
    __syncthreads();
    cnn_forward_group(g, P, x0, A, nullptr, nbr, tid, nt, sync);
    const int L1 = g.nl - 1;
    double part = 0.0;
    for (int t = tid; t < g.n * g.C[L1]; t += nt) part += A[g.aoff[L1] + t];
    double lp = block_sum(part, red) * g.inv_nrm;
    cd off = {lp - lps[2 * s], -lps[2 * s + 1]};
    cd dsum = diag_sum(op, cs);
    cd acc = {0.0, 0.0};
    int sites[MAXLEN], vals[MAXLEN], nch;
    for (int o = w; o < op.num_ops; o += nw) {
      cd m = merged_m(op, o, apply_string(op, o, cs, sites, vals, nch), dsum);
      if (!nonzero(m)) continue;
      double d = 0.0;
      if (nch > 0) d = cnn_local_delta(g, P, x0, A, nbr, nch, sites, vals, sc, lane, false);
      acc = acc + m * cexp({d + off.re, off.im});
    }
    // deterministic cross-wave sum in wave order
    __syncthreads();
    if (lane == 0) {
      red[2 * w] = acc.re;
      red[2 * w + 1] = acc.im;
    }
    __syncthreads();
    if (tid == 0) {
      cd tot = {0.0, 0.0};
      for (int i = 0; i < nw; ++i) tot = tot + cd{red[2 * i], red[2 * i + 1]};
      out[2 * s] = tot.re;
      out[2 * s + 1] = tot.im;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ CNN O_loc, single-flip windows
// One workgroup per sample.  After a position-major forward pass of s into LDS,
// every string is classified (m, changed sites); strings that change one site
// are processed in chunks of `fch` flips, layer by layer: all threads compute
// window outputs of all flips of the chunk (one output position, all channels
// per thread), so each layer costs one barrier for the whole chunk.  The last
// layer's output changes are reduced per flip in a fixed order, and the
// sample's O_loc is summed over strings in string order.
struct EloWinLayout {
  size_t bytes;
  int fch;
};

__host__ __device__ inline size_t elo_win_bytes(const CnnGeom& g, int nops, int fch) {
  size_t d = (size_t)g.n + g.atot + 8 + 4 * (size_t)nops;  // x0, A, red, m + contrib (complex)
  for (int l = 0; l < g.nl - 1; ++l) d += (size_t)fch * win_ex(g, l) * win_ey(g, l) * g.C[l];
  d += (size_t)fch * win_ex(g, g.nl - 1) * win_ey(g, g.nl - 1);  // last-layer partial sums
  size_t i32 = (size_t)g.n + nops + fch;                          // cs, code, flip list
  size_t sh = (size_t)g.n * g.FF;                                 // nbr
  return 8 * d + 4 * i32 + 2 * sh + 64;
}

template <int CM>
__global__ void __launch_bounds__(256) eloc_cnn_win(CnnGeom g, const double* __restrict__ Pg, OpDev op,
                                                    const int32_t* __restrict__ cfg,
                                                    const double* __restrict__ lps, int64_t n, int fch,
                                                    double* __restrict__ out) {
  extern __shared__ double lds[];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int nops = op.num_ops;
  double* x0 = lds;
  double* A = x0 + g.n;
  double* red = A + g.atot;
  double* mstr = red + 8;            // complex [nops]
  double* contrib = mstr + 2 * nops; // complex [nops]
  double* W[JV_MAX_LAYERS];
  double* wp = contrib + 2 * nops;
  for (int l = 0; l < g.nl - 1; ++l) {
    W[l] = wp;
    wp += (size_t)fch * win_ex(g, l) * win_ey(g, l) * g.C[l];
  }
  double* part = wp;
  wp += (size_t)fch * win_ex(g, g.nl - 1) * win_ey(g, g.nl - 1);
  int32_t* cs = (int32_t*)wp;
  int32_t* code = cs + g.n;     // per string: -1 none changed, -2 several, else 2*site + new value
  int32_t* flips = code + nops; // string index of each queued flip
  short* nbr = (short*)(flips + fch);
  for (int t = tid; t < g.n * g.FF; t += nt) {
    int p = t / g.FF, k = t % g.FF;
    int i = p / g.Ly, j = p % g.Ly, ky = k / g.Fy, kx = k % g.Fy;
    nbr[t] = (short)(((i + ky) % g.Lx) * g.Ly + (j + kx) % g.Ly);
  }
  __syncthreads();
  auto sync = [] { __syncthreads(); };
  const int L1 = g.nl - 1;
  const int wlast = win_ex(g, L1) * win_ey(g, L1);
  for (int64_t s = blockIdx.x; s < n; s += gridDim.x) {
    const int32_t* c = cfg + s * g.n;
    for (int i = tid; i < g.n; i += nt) {
      cs[i] = c[i];
      x0[i] = (double)(2 * c[i] - 1);
    }
    __syncthreads();
    cnn_forward_pos<CM>(g, Pg, x0, A, nullptr, nbr, tid, nt, sync);
    double pl = 0.0;
    for (int t = tid; t < g.n * g.C[L1]; t += nt) pl += A[g.aoff[L1] + t];
    const double lp = block_sum(pl, red) * g.inv_nrm;
    const cd off = {lp - lps[2 * s], -lps[2 * s + 1]};
    // classify strings
    for (int o = tid; o < nops; o += nt) {
      int sites[MAXLEN], vals[MAXLEN], nch;
      cd m = apply_string(op, o, cs, sites, vals, nch);
      mstr[2 * o] = m.re;
      mstr[2 * o + 1] = m.im;
      code[o] = nch == 0 ? -1 : (nch == 1 ? 2 * sites[0] + vals[0] : -2);
      contrib[2 * o] = 0.0;
      contrib[2 * o + 1] = 0.0;
    }
    __syncthreads();
    if (tid == 0 && op.num_diag > 1) {  // diagonal merge (branch_free.py:483-485)
      cd sum = {0.0, 0.0};
      for (int d = 0; d < op.num_diag; ++d) sum = sum + cd{mstr[2 * op.diag[d]], mstr[2 * op.diag[d] + 1]};
      for (int d = 1; d < op.num_diag; ++d) { mstr[2 * op.diag[d]] = 0.0; mstr[2 * op.diag[d] + 1] = 0.0; }
      mstr[2 * op.diag[0]] = sum.re;
      mstr[2 * op.diag[0] + 1] = sum.im;
    }
    __syncthreads();
    // strings without a changed site contribute m * exp(off)
    for (int o = tid; o < nops; o += nt) {
      cd m = {mstr[2 * o], mstr[2 * o + 1]};
      if (code[o] == -1 && nonzero(m)) {
        cd r = m * cexp(off);
        contrib[2 * o] = r.re;
        contrib[2 * o + 1] = r.im;
      }
    }
    // single flips, chunk by chunk (the queue is rebuilt in string order by thread 0)
    int next = 0;
    while (true) {
      __syncthreads();
      if (tid == 0) {
        int nf = 0;
        while (next < nops && nf < fch) {
          cd m = {mstr[2 * next], mstr[2 * next + 1]};
          if (code[next] >= 0 && nonzero(m)) flips[nf++] = next;
          ++next;
        }
        red[0] = (double)nf;
        red[1] = (double)next;
      }
      __syncthreads();
      const int nf = (int)red[0];
      next = (int)red[1];
      if (nf == 0) break;
      double z[CM];
      for (int l = 0; l < g.nl; ++l) {
        const int ex = win_ex(g, l), ey = win_ey(g, l), wsz = ex * ey, cout = g.C[l];
        const int pwsz = win_ex(g, l - 1) * win_ey(g, l - 1);
        const double* Xold = l == 0 ? x0 : A + g.aoff[l - 1];
        for (int t = tid; t < nf * wsz; t += nt) {
          const int f = t / wsz, j = t % wsz;
          const int cd_ = code[flips[f]];
          const int site = cd_ >> 1, nv = cd_ & 1;
          const int si = site / g.Ly, sj = site % g.Ly;
          const int a = j / ey, b = j % ey;
          const double* Wprev = l == 0 ? nullptr : W[l - 1] + (size_t)f * pwsz * g.C[l - 1];
          cnn_preact_win<CM>(g, Pg, l, Xold, Wprev, (double)(2 * nv - 1), si, sj, a, b, z);
          if (l < L1) {
            double* Wo = W[l] + (size_t)f * wsz * cout + j;  // planar [c][j] per flip
#pragma unroll
            for (int cc = 0; cc < CM; ++cc)
              if (cc < cout) Wo[cc * wsz] = log_cosh(z[cc]);
          } else {
            const double* Ab = A + g.aoff[l] + win_pos(g, si, sj, a, b);
            double acc = 0.0;
#pragma unroll
            for (int cc = 0; cc < CM; ++cc)
              if (cc < cout) acc += log_cosh(z[cc]) - Ab[cc * g.n];
            part[f * wlast + j] = acc;
          }
        }
        __syncthreads();
      }
      for (int f = tid; f < nf; f += nt) {
        double d = 0.0;
        for (int j = 0; j < wlast; ++j) d += part[f * wlast + j];
        const int o = flips[f];
        cd r = cd{mstr[2 * o], mstr[2 * o + 1]} * cexp({d * g.inv_nrm + off.re, off.im});
        contrib[2 * o] = r.re;
        contrib[2 * o + 1] = r.im;
      }
    }
    __syncthreads();
    if (tid == 0) {
      cd tot = {0.0, 0.0};
      for (int o = 0; o < nops; ++o) {
        tot = tot + cd{contrib[2 * o], contrib[2 * o + 1]};
        // a string changing several sites violates the caller's max_sites_changed bound
        if (code[o] == -2 && nonzero(cd{mstr[2 * o], mstr[2 * o + 1]})) tot = {NAN, NAN};
      }
      out[2 * s] = tot.re;
      out[2 * s + 1] = tot.im;
    }
    __syncthreads();
  }
}

OpDev make_opdev(const jv_optable* op, const double* pref, int L) {
  OpDev d;
  d.L = L;
  d.num_ops = op->num_ops;
  d.max_len = op->max_len;
  d.num_diag = op->num_diag;
  d.idx = op->idx;
  d.map = op->map;
  d.matel = op->matel;
  d.diag = op->diag;
  d.pref = pref;
  return d;
}

int cnn_cmax(const CnnGeom& g) {
  int m = 0;
  for (int l = 0; l < g.nl; ++l) m = std::max(m, g.C[l]);
  if (m <= 4) return 4;
  if (m <= 8) return 8;
  if (m <= 16) return 16;
  if (m <= 32) return 32;
  return 0;
}

int check_op(const jv_optable* op, const double* pref) {
  if (!op || !pref) return fail(JV_E_ARG, "null operator table or prefactors");
  if (op->l_dim != 2) return fail(JV_E_ARG, "only local dimension 2 is supported");
  if (op->num_ops < 1) return fail(JV_E_ARG, "operator has no strings");
  if (op->max_len < 1 || op->max_len > MAXLEN) return fail(JV_E_ARG, "operator strings longer than 8 factors are not supported");
  if (op->num_diag > 0 && !op->diag) return fail(JV_E_ARG, "null diagonal list");
  return JV_OK;
}

}  // namespace jv

using namespace jv;

extern "C" {

int jv_get_s_primes(const jv_optable* op, const double* prefactor, const int32_t* configs, int64_t n,
                    int32_t num_sites, int32_t* sp, double* matel, int32_t* count, void* stream) {
  int rc = check_op(op, prefactor);
  if (rc) return rc;
  if (!configs || !sp || !matel || !count || n < 0 || num_sites < 1) return fail(JV_E_ARG, "jv_get_s_primes: bad arguments");
  if (n == 0) return JV_OK;
  hipStream_t st = (hipStream_t)stream;
  OpDev d = make_opdev(op, prefactor, num_sites);
  int32_t* slot = (int32_t*)workspace(sizeof(int32_t) * (size_t)n * op->num_ops, 0);
  if (!slot) return fail(JV_E_NOMEM, "jv_get_s_primes: workspace allocation failed");
  hipLaunchKernelGGL(sprime_compact, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, d, configs, n,
                     num_sites, matel, slot, count);
  JV_LAUNCH_CHECK("sprime_compact");
  int64_t tot = n * op->num_ops;
  hipLaunchKernelGGL(sprime_write, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, d, configs, n,
                     num_sites, slot, sp);
  JV_LAUNCH_CHECK("sprime_write");
  return JV_OK;
}

int jv_eloc(const jv_net* net, const double* params, const jv_optable* op, const double* prefactor,
            const int32_t* configs, const double* logpsi_s, int64_t n, double* out, void* stream) {
  int rc = check_op(op, prefactor);
  const bool op_single_site = op->max_sites_changed == 1;
  if (rc) return rc;
  if (!net || !params || !configs || !logpsi_s || !out || n < 0) return fail(JV_E_ARG, "jv_eloc: bad arguments");
  if (n == 0) return JV_OK;
  hipStream_t st = (hipStream_t)stream;
  OpDev d = make_opdev(op, prefactor, net->num_sites);
  RbmGeom rg;
  CnnGeom cg;
  if (make_rbm_geom(net, rg)) {
    size_t lds = 4 * 2 * sizeof(cd) * (size_t)rg.H;
    rc = set_lds_limit((const void*)eloc_rbm, lds);
    if (rc) return rc;
    int64_t blocks = std::min<int64_t>((n + 3) / 4, 8192);
    ScopedTimer tm("eloc", st);
    hipLaunchKernelGGL(eloc_rbm, dim3((unsigned)blocks), dim3(256), lds, st, rg, params, d, configs, logpsi_s,
                       n, out);
    JV_LAUNCH_CHECK("eloc_rbm");
    return JV_OK;
  }
  if (make_cnn_geom(net, cg) && op_single_site && win_ok(cg) && cnn_cmax(cg) > 0) {
    int fch = 32;
    while (fch > 1 && elo_win_bytes(cg, op->num_ops, fch) > 78 * 1024) fch /= 2;
    size_t lds = elo_win_bytes(cg, op->num_ops, fch);
    int64_t blocks = std::min<int64_t>(n, 16384);
    ScopedTimer tm("eloc", st);
    switch (cnn_cmax(cg)) {
#define JV_ELO(CMv)                                                                                     \
  case CMv:                                                                                             \
    rc = set_lds_limit((const void*)eloc_cnn_win<CMv>, lds);                                            \
    if (rc) return rc;                                                                                  \
    hipLaunchKernelGGL(eloc_cnn_win<CMv>, dim3((unsigned)blocks), dim3(256), lds, st, cg, params, d, configs, \
                       logpsi_s, n, fch, out);                                                          \
    break;
      JV_ELO(4) JV_ELO(8) JV_ELO(16) JV_ELO(32)
#undef JV_ELO
    }
    JV_LAUNCH_CHECK("eloc_cnn_win");
    return JV_OK;
  }
  if (make_cnn_geom(net, cg)) {
    EllocCnnLayout lay = eloc_cnn_layout(cg, false);
    int waves = 8;
    while (waves > 1 && lay.shared + waves * lay.per_wave > 160 * 1024) waves /= 2;
    size_t lds = lay.shared + waves * lay.per_wave;
    rc = set_lds_limit((const void*)eloc_cnn, lds);
    if (rc) return rc;
    int64_t blocks = std::min<int64_t>(n, 16384);
    ScopedTimer tm("eloc", st);
    hipLaunchKernelGGL(eloc_cnn, dim3((unsigned)blocks), dim3(64 * waves), lds, st, cg, params, d, configs,
                       logpsi_s, n, out);
    JV_LAUNCH_CHECK("eloc_cnn");
    return JV_OK;
  }
  return fail(JV_E_ARG, "jv_eloc: unknown network kind");
}

}  // extern "C"
