// Metropolis sampling with walkers and log-psi caches resident in LDS, and the
// sample weights of the Metropolis and exact samplers.
#include <algorithm>

#include "cnnwin.hpp"
#include "rbm.hpp"

namespace jv {

bool make_rbm_geom(const jv_net* net, RbmGeom& g);
bool make_cnn_geom(const jv_net* net, CnnGeom& g);
int set_lds_limit(const void* fn, size_t bytes);

__global__ void zero_i64(int64_t* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ------------------------------------------------------------------ RBM chains
// One wavefront per chain (mc_rbm_chain in rbm.hpp).
__global__ void __launch_bounds__(256) mc_rbm(RbmGeom g, const double* __restrict__ P, jv_mc_args a,
                                              int32_t* __restrict__ walkers, int32_t* __restrict__ configs,
                                              int64_t* __restrict__ counters) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
  const int chain = blockIdx.x * nw + w;
  if (chain >= a.num_chains) return;
  const int64_t nacc = mc_rbm_chain(g, P, a, walkers, configs, chain, lane, lds + (size_t)w * mc_rbm_lds_doubles(g));
  if (lane == 0) {
    const int total = a.therm_steps + a.samples_per_chain * a.sweep_steps;
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
  cnn_forward_group<double>(g, P, x0, A, nullptr, nbr, lane, WAVE, [] { wave_sync(); });
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
// One workgroup of G waves per chain.  Wave w owns output channels
// [w*CG, (w+1)*CG) of every layer (CG = CM/G), lane j the window element j, so
// the weights a wave reads are wave-uniform (scalar loads).  Activations live in
// per-layer buffers U + l*CM*CS: CM planes of CS doubles, [0, n) the chain's
// current activations and [n, n + wsz_l) the proposed window.  Before each
// proposal the workgroup tabulates, for every window element and kernel offset,
// where its input lies (a stored position, or n + window index), so the
// convolution reads inputs through one table lookup and immediate offsets.
// Per proposal: one barrier per layer, then the accept / commit / next-table
// phase.  The acceptance draw of step t is a pure function of (seed, t, chain,
// call, stream); each wave computes 64 draws at a time, one per lane.
__host__ __device__ inline int mc_win_tab_len(const CnnGeom& g) {
  int t = 0;
  for (int l = 0; l < g.nl; ++l) t += win_ex(g, l) * win_ey(g, l) * g.FF;
  return t;
}
__host__ __device__ inline int mc_win_pos_len(const CnnGeom& g) {
  int t = 0;
  for (int l = 0; l < g.nl; ++l) t += win_ex(g, l) * win_ey(g, l);
  return t;
}
// LDS: the FP64 reduction slots, then the planes U and x0 in the arithmetic
// type (tsize bytes: 8 FP64, 4 mixed FP32), the tables and the neighbours.
__host__ __device__ inline size_t mc_win_bytes(const CnnGeom& g, int cm, int cs, int tsize = 8) {
  const size_t t = (size_t)g.nl * cm * cs + g.n;  // U, x0
  const size_t i32 = (size_t)g.n + 2 * (size_t)mc_win_tab_len(g) + 2 * (size_t)mc_win_pos_len(g);
  return 8 * 16 + (size_t)tsize * t + 4 * i32 + 2 * (size_t)g.n * g.FF;
}

// Input-location table of the proposal at `site` (all threads of the block).
__device__ inline void mc_win_tables(const CnnGeom& g, int site, int* tab, int* pos, int tid, int nt) {
  const int si = site / g.Ly, sj = site - (site / g.Ly) * g.Ly;
  int toff = 0, poff = 0;
  for (int l = 0; l < g.nl; ++l) {
    const int ey = win_ey(g, l), wsz = win_ex(g, l) * ey;
    const int pex = win_ex(g, l - 1), pey = win_ey(g, l - 1);
    for (int t = tid; t < wsz * g.FF; t += nt) {
      const int j = t / g.FF, k = t - j * g.FF;
      const int a = j / ey, b = j - a * ey;
      const int ky = k / g.Fy, kx = k - ky * g.Fy;
      const int aq = a - ky, bq = b - kx;
      int qi = si - aq;
      qi = qi < 0 ? qi + g.Lx : (qi >= g.Lx ? qi - g.Lx : qi);
      int qj = sj - bq;
      qj = qj < 0 ? qj + g.Ly : (qj >= g.Ly ? qj - g.Ly : qj);
      const bool in = l > 0 && aq >= 0 && aq < pex && bq >= 0 && bq < pey;
      tab[toff + t] = in ? g.n + aq * pey + bq : qi * g.Ly + qj;
    }
    for (int j = tid; j < wsz; j += nt) pos[poff + j] = win_site(g, si, sj, j / ey, j - (j / ey) * ey);
    toff += wsz * g.FF;
    poff += wsz;
  }
}

// Plane stride of the window buffers: the smallest instantiated size holding
// the n positions and the largest window (0 when none does).
inline int mc_win_cs(const CnnGeom& g) {
  int need = 0;
  for (int l = 0; l < g.nl; ++l) need = std::max(need, g.n + win_ex(g, l) * win_ey(g, l));
  for (int cs : {128, 256, 512, 1024})
    if (need <= cs) return cs;
  return 0;
}

// T = double: the FP64 form (acceptance decisions bit-exact against the
// oracle); T = float: the mixed-precision form (FP32 window recomputation,
// the change of log psi summed in FP64 before the acceptance test).
template <int CM, int G, int CS, typename T>
__global__ void __launch_bounds__(64 * G) mc_cnn_win(CnnGeom g, const T* __restrict__ Wpk, CnnPack pk,
                                                     jv_mc_args a, int32_t* __restrict__ walkers,
                                                     int32_t* __restrict__ configs,
                                                     int64_t* __restrict__ counters) {
  constexpr int CG = CM / G;
  extern __shared__ double lds[];
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid % WAVE;
  const int w = __builtin_amdgcn_readfirstlane(tid / WAVE);
  const int chain = blockIdx.x;
  const int L1 = g.nl - 1;
  const int tlen = mc_win_tab_len(g), plen = mc_win_pos_len(g);
  double* red = lds;
  T* U = (T*)(red + 16);
  T* x0 = U + (size_t)g.nl * CM * CS;
  int32_t* s = (int32_t*)(x0 + g.n);
  int32_t* tabs = s + g.n;       // [2][tlen]
  int32_t* poss = tabs + 2 * tlen;  // [2][plen]
  short* nbr = (short*)(poss + 2 * plen);
  const cst_t<T>* Wp = (const cst_t<T>*)Wpk;
  for (int t = tid; t < g.n * g.FF; t += nt) {
    const int p = t / g.FF, k = t % g.FF;
    const int i = p / g.Ly, j = p % g.Ly, ky = k / g.Fy, kx = k % g.Fy;
    nbr[t] = (short)(((i + ky) % g.Lx) * g.Ly + (j + kx) % g.Ly);
  }
  for (int i = tid; i < g.n; i += nt) {
    s[i] = walkers[(int64_t)chain * g.n + i];
    x0[i] = (T)(2 * s[i] - 1);
  }
  __syncthreads();
  win_forward<CM, G, T>(g, Wp, pk, x0, U, CM * CS, CS, nbr, tid, nt, [] { __syncthreads(); });
  const int total = a.therm_steps + a.samples_per_chain * a.sweep_steps;
  // per-lane batch of 64 draws
  int dsite = 0;
  double du = 0.0;
  auto batch = [&](int base) {
    if (base + lane < total)
      mc_draw(a.seed, (uint32_t)(base + lane), (uint32_t)chain, a.call_index, a.stream, g.n, dsite, du);
  };
  batch(0);
  int site = __builtin_amdgcn_readlane(dsite, 0);
  if (total > 0) mc_win_tables(g, site, tabs, poss, tid, nt);
  __syncthreads();
  int64_t nacc = 0;
  for (int step = 0; step < total; ++step) {
    const int cur = step & 1;
    const int* tab = tabs + cur * tlen;
    const int* pos = poss + cur * plen;
    const int r = step & 63;
    const int ulo = __builtin_amdgcn_readlane((int)__double2loint(du), r);
    const int uhi = __builtin_amdgcn_readlane((int)__double2hiint(du), r);
    const double u = __hiloint2double(uhi, ulo);
    double diff = 0.0;
    int toff = 0, poff = 0;
    for (int l = 0; l < g.nl; ++l) {
      const int wsz = win_ex(g, l) * win_ey(g, l);
      const cst_t<T>* K = Wp + pk.koff[l];
      const cst_t<T>* B = Wp + pk.boff[l] + w * CG;
      const T* X = l == 0 ? x0 : U + (size_t)(l - 1) * CM * CS;
      T* Ul = U + (size_t)l * CM * CS + (size_t)w * CG * CS;
      for (int j = lane; j < wsz; j += WAVE) {
        T z[CG];
#pragma unroll
        for (int c = 0; c < CG; ++c) z[c] = B[c];
        const int* tj = tab + toff + j * g.FF;
        if (l == 0) {
          for (int k = 0; k < g.FF; ++k) {
            const int q = tj[k];
            const T xv = q == site ? -x0[q] : x0[q];
            const cst_t<T>* wk = K + (k * G + w) * CG;
#pragma unroll
            for (int c = 0; c < CG; ++c) z[c] = fma(xv, wk[c], z[c]);
          }
        } else {
          for (int k = 0; k < g.FF; ++k) {
            const T* xq = X + tj[k];
            const cst_t<T>* wk = K + (size_t)(k * G + w) * CM * CG;
#pragma unroll
            for (int ci = 0; ci < CM; ++ci) {
              const T xv = xq[ci * CS];
#pragma unroll
              for (int c = 0; c < CG; ++c) z[c] = fma(xv, wk[ci * CG + c], z[c]);
            }
          }
        }
        if (l < L1) {
#pragma unroll
          for (int c = 0; c < CG; ++c) Ul[c * CS + g.n + j] = log_cosh(z[c]);
        } else {
          const int p = pos[poff + j];
#pragma unroll
          for (int c = 0; c < CG; ++c) {
            const T v = log_cosh(z[c]);
            Ul[c * CS + g.n + j] = v;
            diff += (double)(v - Ul[c * CS + p]);
          }
        }
      }
      toff += wsz * g.FF;
      poff += wsz;
      if (l < L1) __syncthreads();
    }
    diff = wave_sum(diff);
    if (lane == 0) red[w] = diff;
    __syncthreads();
    double dsum = 0.0;
#pragma unroll
    for (int i = 0; i < G; ++i) dsum += red[i];
    const double dl = dsum * g.inv_nrm;
    const bool acc = u < exp(a.mu * dl);  // random.bernoulli(key, P), sampler.py:336-340
    if (acc) {
      int t = tid;
      int po = 0;
      for (int l = 0; l < g.nl; ++l) {
        const int wsz = win_ex(g, l) * win_ey(g, l);
        for (; t < CM * wsz; t += nt) {
          const int c = t / wsz, j = t - c * wsz;
          T* Uc = U + ((size_t)l * CM + c) * CS;
          Uc[pos[po + j]] = Uc[g.n + j];
        }
        t -= CM * wsz;
        po += wsz;
      }
      if (tid == 0) {
        s[site] = 1 - s[site];
        x0[site] = -x0[site];
      }
      ++nacc;
    }
    if (step + 1 < total) {
      if (((step + 1) & 63) == 0) batch(step + 1);
      site = __builtin_amdgcn_readlane(dsite, (step + 1) & 63);
      mc_win_tables(g, site, tabs + (cur ^ 1) * tlen, poss + (cur ^ 1) * plen, tid, nt);
    }
    __syncthreads();
    const int k = step + 1 - a.therm_steps;
    if (k > 0 && k % a.sweep_steps == 0) {
      const int64_t t = k / a.sweep_steps - 1;
      int32_t* outc = configs + (t * a.num_chains + chain) * g.n;
      for (int i = tid; i < g.n; i += nt) outc[i] = s[i];
    }
  }
  for (int i = tid; i < g.n; i += nt) walkers[(int64_t)chain * g.n + i] = s[i];
  if (tid == 0) {
    atomicAdd((unsigned long long*)&counters[0], (unsigned long long)total);
    atomicAdd((unsigned long long*)&counters[1], (unsigned long long)nacc);
  }
}



// ------------------------------------------------------------------ CNN chains, 3x3 filters, windows <= 64
// The same chain, proposal order, arithmetic and acceptance as mc_cnn_win for
// 3x3 filters, at most three layers and windows of at most 64 elements (one
// window element per lane in every layer).  What a proposal at site s reads
// differs from mc_cnn_win only in how the input locations are found: a
// window element's nine inputs are either inside the previous layer's window
// (a fixed slot n + aq * pey + bq) or at the stored position (si - a + ky,
// sj - b + kx) (periodic) of tap (ky, kx); which of the two, and the slot,
// depend on the lane and the layer only.  Every lane keeps its 9 slots per
// layer in registers; per proposal and layer it wraps three rows and three
// columns once, so each input address is one add and one select, and all
// input loads of a layer are issued at once, with no table pass and no table
// reads between the draws.
constexpr int MCW3_MAXL = 3;

__host__ __device__ inline bool mc_win3_ok(const CnnGeom& g) {
  return g.Fx == 3 && g.Fy == 3 && g.nl >= 1 && g.nl <= MCW3_MAXL && win_ex(g, g.nl - 1) <= g.Lx &&
         win_ey(g, g.nl - 1) <= g.Ly && win_ex(g, g.nl - 1) * win_ey(g, g.nl - 1) <= WAVE;
}
// LDS: the FP64 reduction slots, the planes U and x0 in the arithmetic type,
// the configuration and the neighbours of the first forward pass.
__host__ __device__ inline size_t mc_win3_bytes(const CnnGeom& g, int cm, int cs, int tsize = 8) {
  const size_t t = (size_t)g.nl * cm * cs + g.n;
  return 8 * 16 + (size_t)tsize * t + 4 * (size_t)g.n + 2 * (size_t)g.n * g.FF;
}
// Wrap a row (column) index shifted by at most one period.
__device__ inline int wrap1(int q, int L) { return q < 0 ? q + L : (q >= L ? q - L : q); }

// logpsi (optional): log psi of every emitted configuration, summed from the
// maintained activations (which equal a fresh forward pass of the current
// configuration exactly: every window is recomputed with the forward's
// operations in the forward's order) in the order of the forward kernels'
// own sum, so the values equal jv_net_forward's.
template <int CM, int G, int CS, typename T>
__global__ void __launch_bounds__(64 * G) mc_cnn_win3(CnnGeom g, const T* __restrict__ Wpk, CnnPack pk,
                                                      jv_mc_args a, int32_t* __restrict__ walkers,
                                                      int32_t* __restrict__ configs,
                                                      int64_t* __restrict__ counters, double* __restrict__ logpsi) {
  constexpr int CG = CM / G;
  extern __shared__ double lds[];
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid % WAVE;
  const int w = __builtin_amdgcn_readfirstlane(tid / WAVE);
  const int chain = blockIdx.x;
  const int L1 = g.nl - 1;
  double* red = lds;
  T* U = (T*)(red + 16);
  T* x0 = U + (size_t)g.nl * CM * CS;
  int32_t* s = (int32_t*)(x0 + g.n);
  short* nbr = (short*)(s + g.n);
  const cst_t<T>* Wp = (const cst_t<T>*)Wpk;
  for (int t = tid; t < g.n * g.FF; t += nt) {
    const int p = t / g.FF, k = t % g.FF;
    const int i = p / g.Ly, j = p % g.Ly, ky = k / g.Fy, kx = k % g.Fy;
    nbr[t] = (short)(((i + ky) % g.Lx) * g.Ly + (j + kx) % g.Ly);
  }
  for (int i = tid; i < g.n; i += nt) {
    s[i] = walkers[(int64_t)chain * g.n + i];
    x0[i] = (T)(2 * s[i] - 1);
  }
  // per layer: this lane's window element (a, b) and its nine input slots
  // (>= 0: the slot in the previous layer's window; -1: the stored layer)
  int slot[MCW3_MAXL][9];
  int wa[MCW3_MAXL], wb[MCW3_MAXL];
#pragma unroll
  for (int l = 0; l < MCW3_MAXL; ++l) {
    const int ey = 1 + 2 * (l + 1), pex = l == 0 ? 1 : 1 + 2 * l, pey = pex;
    const int ja = lane / ey, jb = lane - (lane / ey) * ey;
    wa[l] = ja;
    wb[l] = jb;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int aq = ja - k / 3, bq = jb - k % 3;
      const bool in = l > 0 && aq >= 0 && aq < pex && bq >= 0 && bq < pey;
      slot[l][k] = in ? g.n + aq * pey + bq : -1;
    }
  }
  __syncthreads();
  win_forward<CM, G, T>(g, Wp, pk, x0, U, CM * CS, CS, nbr, tid, nt, [] { __syncthreads(); });
  const int total = a.therm_steps + a.samples_per_chain * a.sweep_steps;
  int dsite = 0;
  double du = 0.0;
  auto batch = [&](int base) {
    if (base + lane < total)
      mc_draw(a.seed, (uint32_t)(base + lane), (uint32_t)chain, a.call_index, a.stream, g.n, dsite, du);
  };
  batch(0);
  int site = __builtin_amdgcn_readlane(dsite, 0);
  __syncthreads();
  int64_t nacc = 0;
  for (int step = 0; step < total; ++step) {
    const int r = step & 63;
    const int ulo = __builtin_amdgcn_readlane((int)__double2loint(du), r);
    const int uhi = __builtin_amdgcn_readlane((int)__double2hiint(du), r);
    const double u = __hiloint2double(uhi, ulo);
    const int si = site / g.Ly, sj = site - (site / g.Ly) * g.Ly;
    double diff = 0.0;
#pragma unroll
    for (int l = 0; l < MCW3_MAXL; ++l) {
      if (l < g.nl) {
        const int ey = 1 + 2 * (l + 1), wsz = ey * ey;
        const cst_t<T>* K = Wp + pk.koff[l];
        const cst_t<T>* B = Wp + pk.boff[l] + w * CG;
        T* Ul = U + (size_t)l * CM * CS + (size_t)w * CG * CS;
        if (lane < wsz) {
          // rows si - a + ky and columns sj - b + kx of the stored layer
          // (a < Lx: the window fits the lattice, so one period at most)
          int rw[3], cl[3];
          rw[0] = wrap1(si - wa[l], g.Lx);
          cl[0] = wrap1(sj - wb[l], g.Ly);
#pragma unroll
          for (int r = 1; r < 3; ++r) {
            rw[r] = rw[r - 1] + 1 == g.Lx ? 0 : rw[r - 1] + 1;
            cl[r] = cl[r - 1] + 1 == g.Ly ? 0 : cl[r - 1] + 1;
          }
#pragma unroll
          for (int r = 0; r < 3; ++r) rw[r] *= g.Ly;
          int q[9];
#pragma unroll
          for (int k = 0; k < 9; ++k) q[k] = slot[l][k] >= 0 ? slot[l][k] : rw[k / 3] + cl[k % 3];
          T z[CG];
#pragma unroll
          for (int c = 0; c < CG; ++c) z[c] = B[c];
          if (l == 0) {
            lds_t<T>* X0 = (lds_t<T>*)x0;
            T xv[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) xv[k] = X0[q[k]];
#pragma unroll
            for (int k = 0; k < 9; ++k) xv[k] = q[k] == site ? -xv[k] : xv[k];
#pragma unroll
            for (int k = 0; k < 9; ++k) {
              const cst_t<T>* wk = K + (k * G + w) * CG;
#pragma unroll
              for (int c = 0; c < CG; ++c) z[c] = fma(xv[k], wk[c], z[c]);
            }
          } else {
            lds_t<T>* X = (lds_t<T>*)(U + (size_t)(l - 1) * CM * CS);
#pragma unroll
            for (int k = 0; k < 9; ++k) {
              T xv[CM];
#pragma unroll
              for (int ci = 0; ci < CM; ++ci) xv[ci] = X[q[k] + ci * CS];
              const cst_t<T>* wk = K + (size_t)(k * G + w) * CM * CG;
#pragma unroll
              for (int ci = 0; ci < CM; ++ci)
#pragma unroll
                for (int c = 0; c < CG; ++c) z[c] = fma(xv[ci], wk[ci * CG + c], z[c]);
            }
          }
          if (l < L1) {
#pragma unroll
            for (int c = 0; c < CG; ++c) Ul[c * CS + g.n + lane] = log_cosh(z[c]);
          } else {
            const int p = rw[0] + cl[0];  // the element's own position
#pragma unroll
            for (int c = 0; c < CG; ++c) {
              const T v = log_cosh(z[c]);
              Ul[c * CS + g.n + lane] = v;
              diff += (double)(v - Ul[c * CS + p]);
            }
          }
        }
        if (l < L1) __syncthreads();
      }
    }
    diff = wave_sum(diff);
    if (lane == 0) red[w] = diff;
    __syncthreads();
    double dsum = 0.0;
#pragma unroll
    for (int i = 0; i < G; ++i) dsum += red[i];
    const double dl = dsum * g.inv_nrm;
    const bool acc = u < exp(a.mu * dl);  // random.bernoulli(key, P), sampler.py:336-340
    if (acc) {
      // wave w commits its own channels of every layer's window
#pragma unroll
      for (int l = 0; l < MCW3_MAXL; ++l) {
        const int ey = 1 + 2 * (l + 1);
        if (l < g.nl && lane < ey * ey) {
          const int p = wrap1(si - wa[l], g.Lx) * g.Ly + wrap1(sj - wb[l], g.Ly);
          T* Ul = U + (size_t)l * CM * CS + (size_t)w * CG * CS;
#pragma unroll
          for (int c = 0; c < CG; ++c) Ul[c * CS + p] = Ul[c * CS + g.n + lane];
        }
      }
      if (tid == 0) {
        s[site] = 1 - s[site];
        x0[site] = -x0[site];
      }
      ++nacc;
    }
    if (step + 1 < total) {
      if (((step + 1) & 63) == 0) batch(step + 1);
      site = __builtin_amdgcn_readlane(dsite, (step + 1) & 63);
    }
    __syncthreads();
    const int k = step + 1 - a.therm_steps;
    if (k > 0 && k % a.sweep_steps == 0) {
      const int64_t t = k / a.sweep_steps - 1;
      int32_t* outc = configs + (t * a.num_chains + chain) * g.n;
      for (int i = tid; i < g.n; i += nt) outc[i] = s[i];
      if (logpsi) {
        double pl = 0.0;
        for (int e = tid; e < g.C[L1] * g.n; e += nt) {
          const int c = e / g.n, p = e - c * g.n;
          pl += (double)U[((size_t)L1 * CM + c) * CS + p];
        }
        const double lp = block_sum(pl, red) * g.inv_nrm;
        if (tid == 0) {
          logpsi[2 * (t * a.num_chains + chain)] = lp;
          logpsi[2 * (t * a.num_chains + chain) + 1] = 0.0;
        }
      }
    }
  }
  for (int i = tid; i < g.n; i += nt) walkers[(int64_t)chain * g.n + i] = s[i];
  if (tid == 0) {
    atomicAdd((unsigned long long*)&counters[0], (unsigned long long)total);
    atomicAdd((unsigned long long*)&counters[1], (unsigned long long)nacc);
  }
}

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
  bool fused_lp = false;  // log psi of the emitted samples written by the sampling kernel
  if (make_rbm_geom(net, rg)) {
    const int nw = 4;
    size_t lds = nw * 8 * mc_rbm_lds_doubles(rg);
    int rc = set_lds_limit((const void*)mc_rbm, lds);
    if (rc) return rc;
    ScopedTimer tm("mc_sweep", st);
    hipLaunchKernelGGL(mc_rbm, dim3((unsigned)((a.num_chains + nw - 1) / nw)), dim3(64 * nw), lds, st, rg,
                       params, a, walkers, configs, counters);
    JV_LAUNCH_CHECK("mc_rbm");
  } else if (make_cnn_geom(net, cg) && win_ok(cg) && cnn_cmax(cg) > 0 && mc_win_cs(cg) > 0 &&
             mc_win_bytes(cg, cnn_cmax(cg), mc_win_cs(cg), net->precision == JV_PREC_MIXED ? 4 : 8) <= 160 * 1024) {
    const bool mixed = net->precision == JV_PREC_MIXED;
    const int cm = cnn_cmax(cg), cs = mc_win_cs(cg);
    // waves per chain: each owns cm / groups output channels
    // (one channel per wave, eight waves per chain, measured slower with both
    // kernels: 3.69 against 2.99 ms and 3.50 against 1.93 ms per step)
    const int groups = cm >= 8 ? 4 : 2;
    // 3x3 filters with windows of at most 64 elements: input codes in registers
    const bool w3 = mc_win3_ok(cg);
    // the emitted samples' log psi from the chains' own activations: equal to
    // the FP64 forward kernel's when the workgroup sums with its 256 threads
    // (the mixed form keeps the FP64 forward for them)
    fused_lp = w3 && !mixed && groups == 4 && logpsi != nullptr && cnn_p_lds(cg, cm, 256, 8) <= 160 * 1024;
    const size_t lds = w3 ? mc_win3_bytes(cg, cm, cs, mixed ? 4 : 8) : mc_win_bytes(cg, cm, cs, mixed ? 4 : 8);
    const CnnPack pk = make_cnn_pack(cg, cm, groups);
    ScopedTimer tm("mc_sweep", st);
    const void* wpk = mixed ? (const void*)pack_cnn_weights<float>(cg, pk, params, st, 43)
                            : (const void*)pack_cnn_weights<double>(cg, pk, params, st, 41);
    if (!wpk) return JV_E_NOMEM;
    int rc = JV_OK;
    const void* fn = nullptr;
#define JV_MCW(CMv, Gv, CSv, Tv)                                                                        \
  if (cm == CMv && groups == Gv && cs == CSv && (sizeof(Tv) == 4) == mixed) {                          \
    if (w3) {                                                                                           \
      fn = (const void*)mc_cnn_win3<CMv, Gv, CSv, Tv>;                                                  \
      rc = set_lds_limit(fn, lds);                                                                      \
      if (rc) return rc;                                                                                \
      hipLaunchKernelGGL((mc_cnn_win3<CMv, Gv, CSv, Tv>), dim3((unsigned)a.num_chains), dim3(64 * Gv), lds, \
                         st, cg, (const Tv*)wpk, pk, a, walkers, configs, counters,                     \
                         fused_lp ? logpsi : nullptr);                                                 \
    } else {                                                                                            \
      fn = (const void*)mc_cnn_win<CMv, Gv, CSv, Tv>;                                                   \
      rc = set_lds_limit(fn, lds);                                                                      \
      if (rc) return rc;                                                                                \
      hipLaunchKernelGGL((mc_cnn_win<CMv, Gv, CSv, Tv>), dim3((unsigned)a.num_chains), dim3(64 * Gv), lds, \
                         st, cg, (const Tv*)wpk, pk, a, walkers, configs, counters);                    \
    }                                                                                                   \
  }
#define JV_MCW_CS(CMv, Gv)                                                                              \
  JV_MCW(CMv, Gv, 128, double) JV_MCW(CMv, Gv, 256, double) JV_MCW(CMv, Gv, 512, double)               \
  JV_MCW(CMv, Gv, 1024, double) JV_MCW(CMv, Gv, 128, float) JV_MCW(CMv, Gv, 256, float)                \
  JV_MCW(CMv, Gv, 512, float) JV_MCW(CMv, Gv, 1024, float)
    JV_MCW_CS(4, 2) JV_MCW_CS(8, 4) JV_MCW_CS(16, 4) JV_MCW_CS(32, 4)
#undef JV_MCW_CS
#undef JV_MCW
    if (!fn) return fail(JV_E_ARG, "jv_mc_sample: no window kernel for this channel count");
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
  if (logpsi && ns > 0 && !fused_lp) return jv_net_forward(net, params, configs, ns, logpsi, stream);
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
