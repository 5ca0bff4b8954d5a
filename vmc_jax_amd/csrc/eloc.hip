// Operator strings: s' enumeration with nonzero compaction, and the fused local
// estimator O_loc(s) = sum_j m_j psi(s'_j)/psi(s) with log psi(s') obtained by
// incremental (RBM) or receptive-field-limited (CNN) re-evaluation.
#include <algorithm>

#include <type_traits>

#include "cnnwin.hpp"
#include "ops.hpp"
#include "rbm.hpp"

namespace jv {

bool make_rbm_geom(const jv_net* net, RbmGeom& g);
bool make_cnn_geom(const jv_net* net, CnnGeom& g);
int set_lds_limit(const void* fn, size_t bytes);

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
// One wavefront per sample (eloc_rbm_sample in rbm.hpp).
__global__ void __launch_bounds__(256) eloc_rbm(RbmGeom g, const double* __restrict__ P, OpDev op,
                                                const int32_t* __restrict__ cfg,
                                                const double* __restrict__ lps, int64_t n,
                                                double* __restrict__ out) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
  const int nw = blockDim.x / WAVE;
  cd* th = (cd*)lds + (size_t)w * 2 * g.H;
  for (int64_t s = (int64_t)blockIdx.x * nw + w; s < n; s += (int64_t)gridDim.x * nw)
    eloc_rbm_sample(g, P, op, cfg + s * g.L, lps + 2 * s, out + 2 * s, th, lane);
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
    cnn_forward_group<double>(g, P, x0, A, nullptr, nbr, tid, nt, sync);
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

// Applies string o to s (branch_free.py:456-478, factors in stored order) in
// registers and classifies it: code = -1 when no site changes, -2 when several
// change, else 2*site + new value.
__device__ inline cd classify_string(const OpDev& op, int o, const int32_t* s, int& code) {
  cd m = {op.pref[2 * o], op.pref[2 * o + 1]};
  int ts[MAXLEN], tv[MAXLEN];
  int nt = 0;
#pragma unroll
  for (int k = 0; k < MAXLEN; ++k) {
    ts[k] = -1;
    tv[k] = 0;
  }
#pragma unroll
  for (int k = 0; k < MAXLEN; ++k) {
    if (k < op.max_len) {
      const int base = o * op.max_len + k;
      // indices are clamped like JAX gathers clamp (site to [0, L), state to [0, 1])
      const int i = min(max(op.idx[base], 0), op.L - 1);
      int slot = -1;
#pragma unroll
      for (int t = 0; t < MAXLEN; ++t)
        if (t < nt && ts[t] == i) slot = t;
      int cur = s[i];
#pragma unroll
      for (int t = 0; t < MAXLEN; ++t)
        if (t == slot) cur = tv[t];
      cur = min(max(cur, 0), 1);
      const cd f = {op.matel[2 * (2 * base + cur)], op.matel[2 * (2 * base + cur) + 1]};
      m = m * f;
      const int nv = min(max(op.map[2 * base + cur], 0), 1);
#pragma unroll
      for (int t = 0; t < MAXLEN; ++t) {
        if (slot >= 0 && t == slot) tv[t] = nv;
        if (slot < 0 && t == nt) {
          ts[t] = i;
          tv[t] = nv;
        }
      }
      if (slot < 0) ++nt;
    }
  }
  int nch = 0, first = -1;
#pragma unroll
  for (int t = 0; t < MAXLEN; ++t)
    if (t < nt && tv[t] != s[ts[t]]) {
      if (nch == 0) first = 2 * ts[t] + tv[t];
      ++nch;
    }
  code = nch == 0 ? -1 : (nch == 1 ? first : -2);
  return m;
}

// ------------------------------------------------------------------ CNN O_loc, single-flip windows
// One workgroup per sample (grid-stride).  A forward pass of s with the packed
// weights leaves every layer's activations in LDS; every string is classified
// in registers (m, changed sites); strings changing one site are compacted in
// string order and processed in chunks of `fch` flips, layer by layer: each
// thread computes all CM channels of one window element of one flip, so a
// layer costs one barrier per chunk.  Each flip's last-layer change is reduced
// by one wave; the sample's O_loc is a block reduction over strings.
// LDS: FP64 string data, reductions and per-flip partial sums; activations and
// window buffers in the arithmetic type (tsize bytes: 8 FP64, 4 mixed FP32).
// delta (incremental windows): the pre-activations of layers >= 1 are kept
// as well.
__host__ __device__ inline size_t elo_win_bytes(const CnnGeom& g, int cm, int nops, int fch, int tsize, bool delta) {
  const size_t d = 4 * (size_t)nops + 16 + (size_t)fch * win_ex(g, g.nl - 1) * win_ey(g, g.nl - 1);
  size_t t = (size_t)g.n + (size_t)(delta ? 2 * g.nl - 1 : g.nl) * cm * g.n;
  for (int l = 0; l < g.nl - 1; ++l) t += (size_t)fch * cm * win_ex(g, l) * win_ey(g, l);
  const size_t i32 = (size_t)g.n + 3 * (size_t)nops + 16;
  const size_t sh = (size_t)g.n * g.FF;
  return 8 * d + (size_t)tsize * t + 4 * i32 + 2 * sh;
}

// Flips per chunk: the fewest estimated VALU cycles per flip among the sizes
// whose LDS fits the workgroups per CU below next to a reserve (at least one
// flip); the launch also asks for at least a (wgs + 1)-th of the LDS, so no
// further workgroup of this kernel fills the reserve.  Defaults, measured on
// the 10x10 CNN step, where the kernel runs beside MinSR's solve:
//  - FP64: window recomputation, two workgroups per CU, no reserve (the solve
//    then starts beside it: 128 steps/s; with incremental windows 122, with
//    three workgroups per CU 91, with four 75);
//  - mixed: incremental windows, two workgroups per CU and a 34 KiB reserve,
//    room for one 33 KiB workgroup of the Cholesky kernels that run beside it
//    (202 steps/s; three workgroups per CU 184, no reserve and three
//    workgroups 150).
// Overrides: VMC_JAX_AMD_ELOC_DELTA, VMC_JAX_AMD_ELOC_WGS_PER_CU and
// VMC_JAX_AMD_ELOC_RESERVE_KB (FP64), the same names with _MIXED (mixed form);
// the first is read at every launch.
inline int elo_env_int(const char* name, int dflt) {
  const char* s = getenv(name);
  return s ? atoi(s) : dflt;
}
inline bool elo_delta(int tsize) {
  return elo_env_int(tsize == 4 ? "VMC_JAX_AMD_ELOC_DELTA_MIXED" : "VMC_JAX_AMD_ELOC_DELTA", tsize == 4 ? 1 : 0) != 0;
}
inline int elo_wgs_per_cu(bool delta, int tsize) {
  (void)delta;
  static const int env = elo_env_int("VMC_JAX_AMD_ELOC_WGS_PER_CU", 0);
  static const int env_m = elo_env_int("VMC_JAX_AMD_ELOC_WGS_PER_CU_MIXED", 0);
  const int e = tsize == 4 ? env_m : env;
  return e > 0 ? (e > 8 ? 8 : e) : 2;
}
// Threads per workgroup (256 default; 320, 384 or 512 through
// VMC_JAX_AMD_ELOC_THREADS and VMC_JAX_AMD_ELOC_THREADS_MIXED).  On the 10x10
// CNN 512 threads make the kernel alone 9-12 % faster (FP64 4.01 -> 3.66 ms,
// mixed 2.09 -> 1.84) but, beside MinSR's solve, leave its Cholesky kernels
// too few registers on each SIMD (FP64 128 -> 118.5 steps/s, mixed 203 ->
// 181.5; 384 threads: 122.7 / 190.9).
inline int elo_threads(int tsize) {
  const int t = elo_env_int(tsize == 4 ? "VMC_JAX_AMD_ELOC_THREADS_MIXED" : "VMC_JAX_AMD_ELOC_THREADS", 256);
  return t == 512 || t == 384 || t == 320 ? t : 256;
}
inline size_t elo_reserve(int tsize) {
  static const int env = elo_env_int("VMC_JAX_AMD_ELOC_RESERVE_KB", -1);
  static const int env_m = elo_env_int("VMC_JAX_AMD_ELOC_RESERVE_KB_MIXED", -1);
  const int e = tsize == 4 ? env_m : env;
  return e >= 0 ? (size_t)e * 1024 : (tsize == 4 ? (size_t)34 * 1024 : 0);
}

// With delta, a window element of layer l >= 1 takes on average
// 9 ex(l-1)^2 / ex(l)^2 of the nine kernel offsets (the inputs inside the
// previous layer's window).  VMC_JAX_AMD_ELOC_CHUNK (FP64) and
// VMC_JAX_AMD_ELOC_CHUNK_MIXED force the chunk size.
inline int elo_win_chunk(const CnnGeom& g, int cm, int nops, int threads, int tsize, bool delta) {
  static const int forced_d = elo_env_int("VMC_JAX_AMD_ELOC_CHUNK", 0);
  static const int forced_m = elo_env_int("VMC_JAX_AMD_ELOC_CHUNK_MIXED", 0);
  const int forced = tsize == 4 ? forced_m : forced_d;
  if (forced > 0 && elo_win_bytes(g, cm, nops, forced, tsize, delta) <= (size_t)160 * 1024) return forced;
  int best = 1;
  double best_cost = 1e300;
  const size_t cap = ((size_t)160 * 1024 - elo_reserve(tsize)) / elo_wgs_per_cu(delta, tsize);
  for (int f = 1; f <= 64; ++f) {
    if (f > 1 && elo_win_bytes(g, cm, nops, f, tsize, delta) > cap) break;
    double cost = 0.0;
    for (int l = 0; l < g.nl; ++l) {
      const int wsz = win_ex(g, l) * win_ey(g, l);
      double taps = g.FF;
      if (delta && l >= 1) taps = 9.0 * (win_ex(g, l - 1) * win_ey(g, l - 1)) / wsz;
      const double item = taps * (l == 0 ? 1 : cm) * cm + 45.0 * cm;
      cost += (double)((f * wsz + threads - 1) / threads) * item;
    }
    cost /= f;
    if (cost < best_cost - 1e-9) {
      best_cost = cost;
      best = f;
    }
  }
  return best;
}

// T = double: the FP64 form; T = float: the mixed-precision form (FP32
// activations and window recomputation, log psi, the per-flip changes and
// O_loc accumulated in FP64).  F3: 3x3 filters and at most three layers, so
// window shapes are compile-time constants (index arithmetic by constants,
// unrolled kernel offsets).  Without DL each window element is recomputed
// from its inputs in the order of a forward pass; with DL (incremental
// windows) layers >= 1 are updated from the stored pre-activations and the
// changes of the inputs inside the previous window (on average 3.2 and 4.6 of
// the nine kernel offsets for layers 1 and 2 of a three-layer network), with
// the flip index fastest so that a wave's lanes share window elements.  On
// the 10x10 CNN the incremental form is faster alone (FP64 3.59 against 4.16
// ms, FP32 2.04 against 2.63 ms); which form wins beside MinSR's solve is in
// the notes on the chunk size above.
template <int CM, typename T, bool F3, bool DL>
__global__ void __launch_bounds__(512) eloc_cnn_win(CnnGeom g, const T* __restrict__ Wpk, CnnPack pk, OpDev op,
                                                    const int32_t* __restrict__ cfg,
                                                    const double* __restrict__ lps, int64_t n, int fch,
                                                    double* __restrict__ out) {
  extern __shared__ double lds[];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid % WAVE, wv = tid / WAVE, nw = nt / WAVE;
  const int nops = op.num_ops, L1 = g.nl - 1;
  const int lsz = CM * g.n;  // values per stored layer
  const int wlast = win_ex(g, L1) * win_ey(g, L1);
  constexpr bool DELTA = F3 && DL;
  double* mre = lds;
  double* mim = mre + nops;
  double* cre = mim + nops;
  double* cim = cre + nops;
  double* red = cim + nops;
  double* part = red + 16;  // per flip [wlast]
  T* x0 = (T*)(part + (size_t)fch * wlast);
  T* A = x0 + g.n;  // layer l as [c][n] at A + l*lsz
  // DELTA: pre-activations of layers l >= 1 at Zs + (l-1)*lsz
  T* Zs = DELTA ? A + (size_t)g.nl * lsz : nullptr;
  T* W[JV_MAX_LAYERS];
  T* wp = A + (size_t)(DELTA ? 2 * g.nl - 1 : g.nl) * lsz;
  for (int l = 0; l < L1; ++l) {
    W[l] = wp;  // per flip [c][wsz]; DELTA: activation changes as [c][wsz][fch]
    wp += (size_t)fch * CM * win_ex(g, l) * win_ey(g, l);
  }
  int32_t* cs = (int32_t*)wp;
  int32_t* code = cs + g.n;
  int32_t* flips = code + nops;
  int32_t* finfo = flips + nops;  // per flip: (row << 16) | (column << 1) | new value
  int32_t* wcnt = finfo + nops;
  short* nbr = (short*)(wcnt + 16);
  const cst_t<T>* Wp = (const cst_t<T>*)Wpk;
  for (int t = tid; t < g.n * g.FF; t += nt) {
    const int p = t / g.FF, k = t % g.FF;
    const int i = p / g.Ly, j = p % g.Ly, ky = k / g.Fy, kx = k % g.Fy;
    nbr[t] = (short)(((i + ky) % g.Lx) * g.Ly + (j + kx) % g.Ly);
  }
  auto sync = [] { __syncthreads(); };
  for (int64_t s = blockIdx.x; s < n; s += gridDim.x) {
    const int32_t* c = cfg + s * g.n;
    for (int i = tid; i < g.n; i += nt) {
      cs[i] = c[i];
      x0[i] = (T)(2 * c[i] - 1);
    }
    __syncthreads();
    win_forward<CM, 1, T>(g, Wp, pk, x0, A, lsz, g.n, nbr, tid, nt, sync, Zs);
    double pl = 0.0;
    for (int t = tid; t < g.C[L1] * g.n; t += nt) pl += (double)A[(size_t)L1 * lsz + t];
    const double lp = block_sum(pl, red) * g.inv_nrm;
    const double off_re = lp - lps[2 * s], off_im = -lps[2 * s + 1];
    const double ph_c = cos(off_im), ph_s = sin(off_im);
    for (int o = tid; o < nops; o += nt) {
      int cd_;
      const cd m = classify_string(op, o, cs, cd_);
      mre[o] = m.re;
      mim[o] = m.im;
      code[o] = cd_;
      cre[o] = 0.0;
      cim[o] = 0.0;
    }
    __syncthreads();
    if (op.num_diag > 1) {  // diagonal merge (branch_free.py:483-485)
      double dr = 0.0, di = 0.0;
      for (int d = tid; d < op.num_diag; d += nt) {
        dr += mre[op.diag[d]];
        di += mim[op.diag[d]];
      }
      dr = block_sum(dr, red);
      di = block_sum(di, red);
      for (int d = tid; d < op.num_diag; d += nt) {
        mre[op.diag[d]] = d == 0 ? dr : 0.0;
        mim[op.diag[d]] = d == 0 ? di : 0.0;
      }
      __syncthreads();
    }
    // strings changing no site contribute m exp(off); single flips are listed in string order
    int nflip = 0;
    for (int base = 0; base < nops; base += nt) {
      const int o = base + tid;
      bool flag = false;
      if (o < nops) {
        const double ar = mre[o], ai = mim[o];
        const bool nz = nonzero(cd{ar, ai});
        if (code[o] == -1 && nz) {
          const double e = exp(off_re);
          cre[o] = ar * (e * ph_c) - ai * (e * ph_s);
          cim[o] = ar * (e * ph_s) + ai * (e * ph_c);
        }
        flag = code[o] >= 0 && nz;
      }
      const uint64_t mask = __ballot(flag);
      const int pre = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
      if (lane == 0) wcnt[wv] = __popcll(mask);
      __syncthreads();
      int before = 0, tot = 0;
      for (int i = 0; i < nw; ++i) {
        before += i < wv ? wcnt[i] : 0;
        tot += wcnt[i];
      }
      if (flag) flips[nflip + before + pre] = o;
      nflip += tot;
      __syncthreads();
    }
    for (int i = tid; i < nflip; i += nt) {
      const int cv = code[flips[i]];
      const int site = cv >> 1, si = site / g.Ly;
      finfo[i] = (si << 16) | ((site - si * g.Ly) << 1) | (cv & 1);
    }
    __syncthreads();
    for (int f0 = 0; f0 < nflip; f0 += fch) {
      const int nf = min(fch, nflip - f0);
      // one layer of the chunk; LC >= 0: layer LC of a 3x3 network (shapes known)
      auto layer = [&](auto LC, int lrt) {
        constexpr int lc = decltype(LC)::value;
        const int l = lc >= 0 ? lc : lrt;
        const int ey = lc >= 0 ? 1 + 2 * (lc + 1) : win_ey(g, l);
        const int wsz = lc >= 0 ? ey * ey : win_ex(g, l) * ey;
        const int pex = lc >= 0 ? 1 + 2 * lc : win_ex(g, l - 1), pey = lc >= 0 ? pex : win_ey(g, l - 1);
        const cst_t<T>* K = Wp + pk.koff[l];
        const cst_t<T>* B = Wp + pk.boff[l];
        for (int t = tid; t < nf * wsz; t += nt) {
          // DELTA: the flip index runs fastest, so the lanes of a wave share
          // few window elements (and with them the kernel offsets win_delta3 takes)
          const int f = DELTA ? t % nf : t / wsz, j = DELTA ? t / nf : t - f * wsz;
          const int fi = finfo[f0 + f];
          const int si = fi >> 16, sj = (fi >> 1) & 0x7fff;
          const int site = si * g.Ly + sj;
          const int a = j / ey, b = j - a * ey;
          T z[CM];
          if constexpr (DELTA && lc >= 0) {
            // layer 0 from the spins; layers >= 1 incrementally: the stored
            // pre-activation plus the weighted changes of the inputs
            if constexpr (lc == 0) {
              win_preact03<CM, T>(g, K, B, x0, (T)(2 * (fi & 1) - 1), site, si, sj, a, b, z);
            } else {
              lds_t<T>* zo = (lds_t<T>*)(Zs + (size_t)(l - 1) * lsz + win_site(g, si, sj, a, b));
#pragma unroll
              for (int cc = 0; cc < CM; ++cc) z[cc] = zo[cc * g.n];
              win_delta3<CM, lc, T>(K, W[l - 1], fch, f, a, b, z);
            }
            lds_t<T>* Ab = (lds_t<T>*)(A + (size_t)l * lsz + win_site(g, si, sj, a, b));
            if (l < L1) {
              lds_w<T>* Wo = (lds_w<T>*)(W[l] + (size_t)j * fch + f);
#pragma unroll
              for (int cc = 0; cc < CM; ++cc) Wo[(size_t)cc * wsz * fch] = log_cosh(z[cc]) - Ab[cc * g.n];
            } else {
              double acc = 0.0;
              if constexpr (std::is_same<T, double>::value) {
                acc = log_cosh_change<CM>(z, Ab, g.n);
              } else {
#pragma unroll
                for (int cc = 0; cc < CM; ++cc) acc += (double)(log_cosh(z[cc]) - Ab[cc * g.n]);
              }
              part[j * fch + f] = acc;
            }
            continue;
          } else if constexpr (lc >= 0) {
            if constexpr (lc == 0)
              win_preact03<CM, T>(g, K, B, x0, (T)(2 * (fi & 1) - 1), site, si, sj, a, b, z);
            else
              win_preact3<CM, lc, T>(g, K, B, A + (size_t)(l - 1) * lsz, W[l - 1] + (size_t)f * CM * pex * pey, si,
                                     sj, a, b, z);
          } else {
            if (l == 0)
              win_preact0<CM, T>(g, K, B, x0, (T)(2 * (fi & 1) - 1), site, si, sj, a, b, z);
            else
              win_preact<CM, T>(g, K, B, A + (size_t)(l - 1) * lsz, W[l - 1] + (size_t)f * CM * pex * pey, pex,
                                pey, si, sj, a, b, z);
          }
          if (l < L1) {
            lds_w<T>* Wo = (lds_w<T>*)(W[l] + (size_t)f * CM * wsz + j);
#pragma unroll
            for (int cc = 0; cc < CM; ++cc) Wo[cc * wsz] = log_cosh(z[cc]);
          } else {
            lds_t<T>* Ab = (lds_t<T>*)(A + (size_t)L1 * lsz + win_site(g, si, sj, a, b));
            double acc = 0.0;
            if constexpr (std::is_same<T, double>::value) {
              acc = log_cosh_change<CM>(z, Ab, g.n);
            } else {
#pragma unroll
              for (int cc = 0; cc < CM; ++cc) acc += (double)(log_cosh(z[cc]) - Ab[cc * g.n]);
            }
            part[f * wlast + j] = acc;
          }
        }
        __syncthreads();
      };
      if constexpr (F3) {
        layer(std::integral_constant<int, 0>(), 0);
        if (g.nl > 1) layer(std::integral_constant<int, 1>(), 1);
        if (g.nl > 2) layer(std::integral_constant<int, 2>(), 2);
      } else {
        for (int l = 0; l < g.nl; ++l) layer(std::integral_constant<int, -1>(), l);
      }
      for (int f = wv; f < nf; f += nw) {
        double d = 0.0;
        for (int j = lane; j < wlast; j += WAVE) d += part[DELTA ? j * fch + f : f * wlast + j];
        d = wave_sum(d);
        if (lane == 0) {
          const int o = flips[f0 + f];
          const double e = exp(d * g.inv_nrm + off_re);
          const double ar = mre[o], ai = mim[o];
          cre[o] = ar * (e * ph_c) - ai * (e * ph_s);
          cim[o] = ar * (e * ph_s) + ai * (e * ph_c);
        }
      }
    }
    __syncthreads();
    double tr = 0.0, ti = 0.0, bad = 0.0;
    for (int o = tid; o < nops; o += nt) {
      tr += cre[o];
      ti += cim[o];
      // a string changing several sites violates the caller's max_sites_changed bound
      if (code[o] == -2 && nonzero(cd{mre[o], mim[o]})) bad = 1.0;
    }
    tr = block_sum(tr, red);
    ti = block_sum(ti, red);
    bad = block_sum(bad, red);
    if (tid == 0) {
      out[2 * s] = bad > 0.0 ? NAN : tr;
      out[2 * s + 1] = bad > 0.0 ? NAN : ti;
    }
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
  d.fermi = op->fermi;
  return d;
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
  if (rc) return rc;
  // the window kernels classify strings without Jordan-Wigner signs
  const bool op_single_site = op->max_sites_changed == 1 && !op->fermi;
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
    const int cm = cnn_cmax(cg);
    const bool mixed = net->precision == JV_PREC_MIXED;
    const int ts = mixed ? 4 : 8;
    const bool f3 = cg.Fx == 3 && cg.Fy == 3 && cg.nl <= 3;
    const bool delta = f3 && elo_delta(ts);  // the kernel's DELTA
    const int thr = elo_threads(ts);
    const int fch = elo_win_chunk(cg, cm, op->num_ops, thr, ts, delta);
    // at least a (wgs + 1)-th of the LDS, so that no further workgroup of this
    // kernel takes the reserve
    const size_t lds = std::max(elo_win_bytes(cg, cm, op->num_ops, fch, ts, delta),
                                (size_t)160 * 1024 / (elo_wgs_per_cu(delta, ts) + 1) + 512);
    if (lds <= 160 * 1024) {
      const CnnPack pk = make_cnn_pack(cg, cm, 1);
      ScopedTimer tm("eloc", st);
      // one workgroup per sample (grid-stride beyond 2^20; capping the grid to
      // leave CUs to the solve beside it was measured slower)
      const int64_t blocks = std::min<int64_t>(n, 1 << 20);
#define JV_ELO(CMv, Tv, slot)                                                                            \
  case CMv: {                                                                                            \
    const Tv* wpk = pack_cnn_weights<Tv>(cg, pk, params, st, slot);                                      \
    if (!wpk) return JV_E_NOMEM;                                                                         \
    if (delta) {                                                                                         \
      rc = set_lds_limit((const void*)eloc_cnn_win<CMv, Tv, true, true>, lds);                           \
      if (rc) return rc;                                                                                 \
      hipLaunchKernelGGL((eloc_cnn_win<CMv, Tv, true, true>), dim3((unsigned)blocks), dim3(thr), lds, st, cg, wpk, \
                         pk, d, configs, logpsi_s, n, fch, out);                                         \
    } else if (f3) {                                                                                     \
      rc = set_lds_limit((const void*)eloc_cnn_win<CMv, Tv, true, false>, lds);                          \
      if (rc) return rc;                                                                                 \
      hipLaunchKernelGGL((eloc_cnn_win<CMv, Tv, true, false>), dim3((unsigned)blocks), dim3(thr), lds, st, cg, \
                         wpk, pk, d, configs, logpsi_s, n, fch, out);                                    \
    } else {                                                                                             \
      rc = set_lds_limit((const void*)eloc_cnn_win<CMv, Tv, false, false>, lds);                         \
      if (rc) return rc;                                                                                 \
      hipLaunchKernelGGL((eloc_cnn_win<CMv, Tv, false, false>), dim3((unsigned)blocks), dim3(thr), lds, st, cg, \
                         wpk, pk, d, configs, logpsi_s, n, fch, out);                                    \
    }                                                                                                    \
    break;                                                                                               \
  }
      if (mixed) {
        switch (cm) { JV_ELO(4, float, 42) JV_ELO(8, float, 42) JV_ELO(16, float, 42) JV_ELO(32, float, 42) }
      } else {
        switch (cm) { JV_ELO(4, double, 40) JV_ELO(8, double, 40) JV_ELO(16, double, 40) JV_ELO(32, double, 40) }
      }
#undef JV_ELO
      JV_LAUNCH_CHECK("eloc_cnn_win");
      return JV_OK;
    }
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
