// Batched field sweep: many independent RBM states (different lattice sizes,
// hidden-unit counts and operators) advanced by one launch per phase.  A
// device-resident instance table maps every workgroup of a launch to one
// instance; the kernels then run the same per-chain / per-sample routines as
// the single-network kernels (rbm.hpp), so each instance's samples, O_loc and
// O_k are those of jv_mc_sample / jv_eloc / jv_net_gradients on its own.
//
// The Philox call index of every instance lives in the table and is advanced
// on the device after each sampling pass, so a captured step (hipGraph)
// replays with fresh draws.  Nothing in jv_batch_step allocates, synchronises
// or issues a memset, so it can be captured.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "rbm.hpp"

namespace jv {

bool make_rbm_geom(const jv_net* net, RbmGeom& g);
int set_lds_limit(const void* fn, size_t bytes);

struct BatchInst {
  RbmGeom g;
  const double* P;
  OpDev op;
  jv_mc_args a;
  int32_t* walkers;
  int32_t* configs;
  double* logpsi;
  int64_t* counters;
  double* eloc;
  double* grad;
  int32_t layout;
  int64_t n;       // samples_per_chain * num_chains
  int32_t wg[3];   // first workgroup of this instance in the sample / gradient / O_loc launches
};

}  // namespace jv

struct jv_batch {
  std::vector<jv::BatchInst> host;
  jv::BatchInst* dev = nullptr;
  int32_t count = 0;
  int32_t waves[3] = {4, 4, 4};      // waves per workgroup per phase
  int32_t per_wave[3] = {1, 1, 1};   // samples per wave (gradient and O_loc phases)
  int32_t total_wg[3] = {0, 0, 0};
  size_t lds[3] = {0, 0, 0};
};

struct jv_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace jv {

// Instance owning workgroup `b` of launch `ph` (binary search over the
// ascending first-workgroup indices).
__device__ inline int batch_find(const BatchInst* __restrict__ T, int count, int ph, int b) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (T[mid].wg[ph] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(512) batch_sample(const BatchInst* __restrict__ T, int count, size_t lds_per_wave) {
  extern __shared__ double lds[];
  const int i = batch_find(T, count, 0, blockIdx.x);
  const BatchInst& I = T[i];
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
  const int chain = (blockIdx.x - I.wg[0]) * nw + w;
  if (chain >= I.a.num_chains) return;
  const int64_t nacc = mc_rbm_chain(I.g, I.P, I.a, I.walkers, I.configs, chain, lane, lds + (size_t)w * lds_per_wave);
  if (lane == 0) {
    const int total = I.a.therm_steps + I.a.samples_per_chain * I.a.sweep_steps;
    atomicAdd((unsigned long long*)&I.counters[0], (unsigned long long)total);
    atomicAdd((unsigned long long*)&I.counters[1], (unsigned long long)nacc);
  }
}

__global__ void __launch_bounds__(512) batch_grad(const BatchInst* __restrict__ T, int count, int per_wave) {
  const int i = batch_find(T, count, 1, blockIdx.x);
  const BatchInst& I = T[i];
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
  const int64_t rowlen = rbm_grad_row_doubles(I.g, I.layout);
  const int64_t s0 = (int64_t)(blockIdx.x - I.wg[1]) * nw * per_wave;
  for (int k = 0; k < per_wave; ++k) {
    const int64_t s = s0 + (int64_t)k * nw + w;
    if (s >= I.n) break;
    rbm_fwd_grad_sample(I.g, I.P, I.configs + s * I.g.L, I.layout, I.grad ? I.grad + s * rowlen : nullptr,
                        I.logpsi + 2 * s, lane);
  }
}

__global__ void __launch_bounds__(512) batch_eloc(const BatchInst* __restrict__ T, int count, int per_wave,
                                                  size_t lds_per_wave) {
  extern __shared__ double lds[];
  const int i = batch_find(T, count, 2, blockIdx.x);
  const BatchInst& I = T[i];
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
  cd* th = (cd*)(lds + (size_t)w * lds_per_wave);
  const int64_t s0 = (int64_t)(blockIdx.x - I.wg[2]) * nw * per_wave;
  for (int k = 0; k < per_wave; ++k) {
    const int64_t s = s0 + (int64_t)k * nw + w;
    if (s >= I.n) break;
    eloc_rbm_sample(I.g, I.P, I.op, I.configs + s * I.g.L, I.logpsi + 2 * s, I.eloc + 2 * s, th, lane);
  }
}

__global__ void batch_begin(BatchInst* __restrict__ T, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) {
    T[i].counters[0] = 0;
    T[i].counters[1] = 0;
  }
}

// After a sampling pass: the next pass draws from a fresh Philox counter block
// (the per-call advance of MCSampler, sampler.py:288-291).
__global__ void batch_advance(BatchInst* __restrict__ T, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) T[i].a.call_index += 1;
}

OpDev make_opdev(const jv_optable* op, const double* pref, int L);
int check_op(const jv_optable* op, const double* pref);

// Workgroup layout of every phase from the current launch shapes; uploads the
// table (synchronous: never called inside a capture).
static int batch_layout(jv_batch* b) {
  size_t mc_lds = 8, el_lds = 8;
  for (auto& I : b->host) {
    mc_lds = std::max(mc_lds, 8 * mc_rbm_lds_doubles(I.g));
    el_lds = std::max(el_lds, (size_t)2 * sizeof(cd) * I.g.H);
  }
  int tot[3] = {0, 0, 0};
  for (auto& I : b->host) {
    I.wg[0] = tot[0];
    tot[0] += (I.a.num_chains + b->waves[0] - 1) / b->waves[0];
    for (int ph = 1; ph < 3; ++ph) {
      I.wg[ph] = tot[ph];
      const int64_t per = (int64_t)b->waves[ph] * b->per_wave[ph];
      tot[ph] += (int)((I.n + per - 1) / per);
    }
  }
  for (int ph = 0; ph < 3; ++ph) b->total_wg[ph] = tot[ph];
  b->lds[0] = mc_lds;
  b->lds[1] = 0;
  b->lds[2] = el_lds;
  if (b->waves[0] * mc_lds > 160 * 1024 || b->waves[2] * el_lds > 160 * 1024)
    return fail(JV_E_NOMEM, "jv_batch: the launch shape needs more than 160 KiB of LDS per workgroup");
  JV_HIP(hipMemcpy(b->dev, b->host.data(), sizeof(BatchInst) * b->host.size(), hipMemcpyHostToDevice));
  int rc = set_lds_limit((const void*)batch_sample, b->waves[0] * mc_lds);
  if (!rc) rc = set_lds_limit((const void*)batch_eloc, b->waves[2] * el_lds);
  return rc;
}

// Event timers are left out of captured launches (a captured event record
// would need an external event node).
static bool capturing(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

static int batch_launch(jv_batch* b, int ph, hipStream_t st) {
  const int nw = b->waves[ph];
  const bool cap = capturing(st);
  switch (ph) {
    case JV_BATCH_SAMPLE: {
      ScopedTimer tm(cap ? nullptr : "batch_sample", st);
      hipLaunchKernelGGL(batch_sample, dim3((unsigned)b->total_wg[0]), dim3(64 * nw), nw * b->lds[0], st, b->dev,
                         b->count, b->lds[0] / 8);
      JV_LAUNCH_CHECK("batch_sample");
      hipLaunchKernelGGL(batch_advance, dim3((unsigned)((b->count + 63) / 64)), dim3(64), 0, st, b->dev, b->count);
      JV_LAUNCH_CHECK("batch_advance");
      return JV_OK;
    }
    case JV_BATCH_GRAD: {
      ScopedTimer tm(cap ? nullptr : "batch_grad", st);
      hipLaunchKernelGGL(batch_grad, dim3((unsigned)b->total_wg[1]), dim3(64 * nw), 0, st, b->dev, b->count,
                         b->per_wave[1]);
      JV_LAUNCH_CHECK("batch_grad");
      return JV_OK;
    }
    case JV_BATCH_ELOC: {
      ScopedTimer tm(cap ? nullptr : "batch_eloc", st);
      hipLaunchKernelGGL(batch_eloc, dim3((unsigned)b->total_wg[2]), dim3(64 * nw), nw * b->lds[2], st, b->dev,
                         b->count, b->per_wave[2], b->lds[2] / 8);
      JV_LAUNCH_CHECK("batch_eloc");
      return JV_OK;
    }
  }
  return fail(JV_E_ARG, "jv_batch: unknown phase");
}

}  // namespace jv

using namespace jv;

extern "C" {

int jv_batch_create(const jv_batch_item* items, int32_t count, jv_batch** out) {
  if (!items || count < 1 || !out) return fail(JV_E_ARG, "jv_batch_create: bad arguments");
  *out = nullptr;
  jv_batch* b = new jv_batch();
  b->count = count;
  b->host.resize(count);
  for (int32_t i = 0; i < count; ++i) {
    const jv_batch_item& it = items[i];
    BatchInst& I = b->host[i];
    int rc = check_op(&it.op, it.prefactor);
    if (!rc && !make_rbm_geom(&it.net, I.g)) rc = fail(JV_E_ARG, "jv_batch_create: item networks must be RBMs");
    if (!rc && (!it.params || !it.walkers || !it.configs || !it.logpsi || !it.counters || !it.eloc))
      rc = fail(JV_E_ARG, "jv_batch_create: null buffer in item " + std::to_string(i));
    if (!rc && (it.grad_layout != 0 && it.grad_layout != 1)) rc = fail(JV_E_ARG, "jv_batch_create: bad grad_layout");
    const jv_mc_args& a = it.mc;
    if (!rc && (a.num_chains < 1 || a.sweep_steps < 1 || a.therm_steps < 0 || a.samples_per_chain < 1))
      rc = fail(JV_E_ARG, "jv_batch_create: bad chain / sweep counts in item " + std::to_string(i));
    if (!rc && (a.mu < 0.0 || a.mu > 2.0)) rc = fail(JV_E_ARG, "mu must be in the range [0, 2]");
    if (rc) {
      delete b;
      return rc;
    }
    I.P = it.params;
    I.op = make_opdev(&it.op, it.prefactor, it.net.num_sites);
    I.a = a;
    I.walkers = it.walkers;
    I.configs = it.configs;
    I.logpsi = it.logpsi;
    I.counters = it.counters;
    I.eloc = it.eloc;
    I.grad = it.grad;
    I.layout = it.grad_layout;
    I.n = (int64_t)a.samples_per_chain * a.num_chains;
  }
  if (hipMalloc((void**)&b->dev, sizeof(BatchInst) * count) != hipSuccess) {
    delete b;
    return fail(JV_E_NOMEM, "jv_batch_create: instance table allocation failed");
  }
  int rc = batch_layout(b);
  if (rc) {
    (void)hipFree(b->dev);
    delete b;
    return rc;
  }
  *out = b;
  return JV_OK;
}

int jv_batch_destroy(jv_batch* b) {
  if (!b) return JV_OK;
  if (b->dev) (void)hipFree(b->dev);
  delete b;
  return JV_OK;
}

int jv_batch_set_shape(jv_batch* b, int32_t phase, int32_t waves, int32_t samples_per_wave) {
  if (!b || phase < 0 || phase > 2) return fail(JV_E_ARG, "jv_batch_set_shape: bad arguments");
  if (waves != 1 && waves != 2 && waves != 4 && waves != 8)
    return fail(JV_E_ARG, "jv_batch_set_shape: waves must be 1, 2, 4 or 8");
  if (samples_per_wave < 1 || samples_per_wave > 64)
    return fail(JV_E_ARG, "jv_batch_set_shape: samples_per_wave must be in [1, 64]");
  const int32_t ow = b->waves[phase], os = b->per_wave[phase];
  b->waves[phase] = waves;
  b->per_wave[phase] = phase == JV_BATCH_SAMPLE ? 1 : samples_per_wave;
  int rc = batch_layout(b);
  if (rc) {
    b->waves[phase] = ow;
    b->per_wave[phase] = os;
    (void)batch_layout(b);
  }
  return rc;
}

int jv_batch_phase(jv_batch* b, int32_t phase, void* stream) {
  if (!b) return fail(JV_E_ARG, "jv_batch_phase: null batch");
  return batch_launch(b, phase, (hipStream_t)stream);
}

int jv_batch_step(jv_batch* b, void* stream) {
  if (!b) return fail(JV_E_ARG, "jv_batch_step: null batch");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(batch_begin, dim3((unsigned)((b->count + 63) / 64)), dim3(64), 0, st, b->dev, b->count);
  JV_LAUNCH_CHECK("batch_begin");
  int rc = batch_launch(b, JV_BATCH_SAMPLE, st);
  if (!rc) rc = batch_launch(b, JV_BATCH_GRAD, st);
  if (!rc) rc = batch_launch(b, JV_BATCH_ELOC, st);
  return rc;
}

int jv_graph_capture(jv_batch* b, void* stream, jv_graph** out) {
  if (!b || !out || !stream) return fail(JV_E_ARG, "jv_graph_capture: bad arguments (a non-default stream is needed)");
  *out = nullptr;
  hipStream_t st = (hipStream_t)stream;
  JV_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const int rc = jv_batch_step(b, stream);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  hipGraphExec_t x = nullptr;
  const hipError_t e2 = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e2 != hipSuccess) {
    (void)hipGraphDestroy(g);
    return hip_fail(e2, "hipGraphInstantiate");
  }
  jv_graph* gr = new jv_graph();
  gr->graph = g;
  gr->exec = x;
  *out = gr;
  return JV_OK;
}

int jv_graph_launch(jv_graph* g, void* stream) {
  if (!g) return fail(JV_E_ARG, "jv_graph_launch: null graph");
  JV_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
  return JV_OK;
}

int jv_graph_destroy(jv_graph* g) {
  if (!g) return JV_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return JV_OK;
}

}  // extern "C"
