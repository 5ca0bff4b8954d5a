// Error plumbing, per-device workspace and solver handles for libjvmc_hip.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <sys/syscall.h>
#include <unistd.h>

#include <mutex>
#include <vector>

#include "common.hpp"

namespace jv {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  return JV_E_HIP;
}

namespace {
struct DeviceState {
  std::vector<void*> ptr;
  std::vector<size_t> size;
  rocblas_handle blas = nullptr;
  hipStream_t aux = nullptr;             // second stream for work that overlaps the caller's
  hipEvent_t ev[2] = {nullptr, nullptr};
};
// Workspaces and solver handles are per host thread and device, so threads
// that drive different streams (one per job) never share scratch memory.
struct ThreadState {
  std::vector<DeviceState> dev;
  ~ThreadState();
};
std::mutex g_mu;                      // guards the registry below
std::vector<ThreadState*> g_threads;  // live thread states, for jv_shutdown
thread_local ThreadState t_state;
thread_local bool t_registered = false;

void release(DeviceState& s, int d) {
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(d);
  (void)hipDeviceSynchronize();
  for (void* p : s.ptr)
    if (p) (void)hipFree(p);
  s.ptr.clear();
  s.size.clear();
  if (s.blas) rocblas_destroy_handle(s.blas);
  s.blas = nullptr;
  for (auto& e : s.ev)
    if (e) (void)hipEventDestroy(e), e = nullptr;
  if (s.aux) (void)hipStreamDestroy(s.aux);
  s.aux = nullptr;
  (void)hipSetDevice(cur);
}

// A worker thread that ends returns its workspaces and handles.  The main
// thread's thread_local state is destroyed only inside exit(), when the HIP
// runtime and the libraries it loaded (a profiler's included) may already be
// finalising: it makes no HIP call then and leaves the memory to the process
// teardown (jv_shutdown releases it explicitly while the runtime is up).
ThreadState::~ThreadState() {
  std::lock_guard<std::mutex> lk(g_mu);
  const bool main_thread = (pid_t)syscall(SYS_gettid) == getpid();
  if (!main_thread)
    for (size_t d = 0; d < dev.size(); ++d)
      if (!dev[d].ptr.empty() || dev[d].blas || dev[d].aux) release(dev[d], (int)d);
  for (size_t i = 0; i < g_threads.size(); ++i)
    if (g_threads[i] == this) { g_threads.erase(g_threads.begin() + i); break; }
}

// rocBLAS loads its kernel libraries, and the HIP runtime rocSOLVER's code
// objects, on first use.  Threads that make their first solver calls at the
// same moment raced inside those loads; they now happen once per device,
// before any handle is handed out: rocblas_initialize() plus one tiny real and
// one tiny complex eigendecomposition on a private stream.
constexpr int MAX_DEVICES = 64;
std::once_flag g_warm_once[MAX_DEVICES];
int g_warm_rc[MAX_DEVICES];

void warm_libraries(int d) {
  rocblas_initialize();
  int rc = 0;
  rocblas_handle h = nullptr;
  hipStream_t st = nullptr;
  double* buf = nullptr;
  constexpr int n = 16;
  const size_t bytes = sizeof(double) * (2 * n * n + 4 * n) + sizeof(rocblas_int);
  if (rocblas_create_handle(&h) != rocblas_status_success) rc = JV_E_SOLVER;
  if (!rc && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) rc = JV_E_HIP;
  if (!rc && hipMalloc((void**)&buf, bytes) != hipSuccess) rc = JV_E_NOMEM;
  if (!rc) {
    rocblas_set_stream(h, st);
    (void)hipMemsetAsync(buf, 0, bytes, st);  // A = 0: a valid symmetric / Hermitian input
    double* A = buf;
    double* w = buf + 2 * n * n;
    double* E = w + n;
    rocblas_int* info = (rocblas_int*)(E + 2 * n);
    if (rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, n, A, n, w, E, info) !=
            rocblas_status_success ||
        rocsolver_zheevd(h, rocblas_evect_original, rocblas_fill_upper, n, (rocblas_double_complex*)A, n, w, E,
                         info) != rocblas_status_success)
      rc = JV_E_SOLVER;
    if (hipStreamSynchronize(st) != hipSuccess && !rc) rc = JV_E_HIP;
  }
  if (buf) (void)hipFree(buf);
  if (st) (void)hipStreamDestroy(st);
  if (h) rocblas_destroy_handle(h);
  g_warm_rc[d] = rc;
}

int warm(int d) {
  if (d < 0 || d >= MAX_DEVICES) return JV_E_ARG;
  std::call_once(g_warm_once[d], warm_libraries, d);
  return g_warm_rc[d];
}

DeviceState& state() {
  if (!t_registered) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_threads.push_back(&t_state);
    t_registered = true;
  }
  int d = 0;
  (void)hipGetDevice(&d);
  if ((int)t_state.dev.size() <= d) t_state.dev.resize(d + 1);
  return t_state.dev[d];
}
}  // namespace

// Grow-only scratch buffer `slot` of the calling thread on the current device.
// Replacing a buffer synchronises the device first so in-flight kernels never
// see it vanish.
void* workspace(size_t bytes, int slot) {
  DeviceState& s = state();
  if ((int)s.ptr.size() <= slot) {
    s.ptr.resize(slot + 1, nullptr);
    s.size.resize(slot + 1, 0);
  }
  if (s.size[slot] < bytes) {
    if (s.ptr[slot]) {
      (void)hipDeviceSynchronize();
      (void)hipFree(s.ptr[slot]);
      s.ptr[slot] = nullptr;
      s.size[slot] = 0;
    }
    size_t b = bytes < 256 ? 256 : bytes;
    if (hipMalloc(&s.ptr[slot], b) != hipSuccess) {
      s.ptr[slot] = nullptr;
      return nullptr;
    }
    s.size[slot] = b;
  }
  return s.ptr[slot];
}

// The calling thread's auxiliary stream on the current device (default
// priority, non-blocking) and two timing-free events for joining it.
bool aux_stream(hipStream_t* st, hipEvent_t* e0, hipEvent_t* e1) {
  DeviceState& s = state();
  if (!s.aux) {
    hipStream_t a = nullptr;
    hipEvent_t e[2] = {nullptr, nullptr};
    bool ok = hipStreamCreateWithFlags(&a, hipStreamNonBlocking) == hipSuccess;
    for (auto& x : e) ok = ok && hipEventCreateWithFlags(&x, hipEventDisableTiming) == hipSuccess;
    if (!ok) {  // leave nothing half made
      for (auto& x : e)
        if (x) (void)hipEventDestroy(x);
      if (a) (void)hipStreamDestroy(a);
      return false;
    }
    s.aux = a;
    s.ev[0] = e[0];
    s.ev[1] = e[1];
  }
  *st = s.aux;
  *e0 = s.ev[0];
  *e1 = s.ev[1];
  return true;
}

rocblas_handle blas_handle(hipStream_t stream) {
  DeviceState& s = state();
  if (!s.blas) {
    int d = 0;
    (void)hipGetDevice(&d);
    if (warm(d) != JV_OK) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    if (rocblas_create_handle(&s.blas) != rocblas_status_success) return nullptr;
    // deterministic kernels only (no atomic split-K): results must not depend on scheduling
    rocblas_set_atomics_mode(s.blas, rocblas_atomics_not_allowed);
  }
  rocblas_set_stream(s.blas, stream);
  return s.blas;
}

}  // namespace jv

extern "C" {

int jv_version(void) { return 100; }

const char* jv_last_error(void) { return jv::g_last_error.c_str(); }

int jv_set_device(int device) {
  JV_HIP(hipSetDevice(device));
  return JV_OK;
}

int jv_init(int device) {
  JV_HIP(hipSetDevice(device));
  int rc = jv::warm(device);
  if (rc) return jv::fail(rc, "jv_init: loading the rocBLAS / rocSOLVER kernels failed");
  return JV_OK;
}

int jv_workspace_bytes(int64_t* bytes) {
  if (!bytes) return jv::fail(JV_E_ARG, "jv_workspace_bytes: null output");
  int d = 0;
  JV_HIP(hipGetDevice(&d));
  size_t total = 0;
  if ((int)jv::t_state.dev.size() > d) {
    const auto& s = jv::t_state.dev[d];
    for (size_t b : s.size) total += b;
    size_t lib = 0;  // the solver library handle's own device memory
    if (s.blas && rocblas_get_device_memory_size(s.blas, &lib) == rocblas_status_success) total += lib;
  }
  *bytes = (int64_t)total;
  return JV_OK;
}

int jv_shutdown(void) {
  // releases the calling thread's workspaces and handles on every device
  for (size_t d = 0; d < jv::t_state.dev.size(); ++d) {
    auto& s = jv::t_state.dev[d];
    if (!s.ptr.empty() || s.blas || s.aux) jv::release(s, (int)d);
  }
  return JV_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ kernel timers
// Optional HIP-event timing of named kernels on the stream they run on.
#include <map>

namespace jv {
namespace {
struct TimerRec {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
};
std::mutex g_tmu;
bool g_timing = false;
std::map<std::string, TimerRec> g_timers;
}  // namespace

bool timing_enabled() { return g_timing; }

void* timer_begin(const char* name, hipStream_t st) {
  if (!g_timing || !name) return nullptr;
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return nullptr;
  if (hipEventCreate(&b) != hipSuccess) { (void)hipEventDestroy(a); return nullptr; }
  (void)hipEventRecord(a, st);
  std::lock_guard<std::mutex> lk(g_tmu);
  auto& r = g_timers[name];
  r.ev.push_back({a, b});
  return (void*)b;
}

void timer_end(void* tok, hipStream_t st) {
  if (tok) (void)hipEventRecord((hipEvent_t)tok, st);
}

}  // namespace jv

extern "C" {

int jv_timer_enable(int on) {
  jv::g_timing = on != 0;
  return JV_OK;
}

int jv_timer_reset(void) {
  std::lock_guard<std::mutex> lk(jv::g_tmu);
  for (auto& kv : jv::g_timers)
    for (auto& p : kv.second.ev) {
      (void)hipEventSynchronize(p.second);
      (void)hipEventDestroy(p.first);
      (void)hipEventDestroy(p.second);
    }
  jv::g_timers.clear();
  return JV_OK;
}

int jv_timer_read(const char* name, double* total_ms, int64_t* count) {
  if (!name || !total_ms || !count) return jv::fail(JV_E_ARG, "jv_timer_read: bad arguments");
  std::lock_guard<std::mutex> lk(jv::g_tmu);
  *total_ms = 0.0;
  *count = 0;
  auto it = jv::g_timers.find(name);
  if (it == jv::g_timers.end()) return JV_OK;
  for (auto& p : it->second.ev) {
    JV_HIP(hipEventSynchronize(p.second));
    float ms = 0.f;
    JV_HIP(hipEventElapsedTime(&ms, p.first, p.second));
    *total_ms += ms;
    *count += 1;
  }
  return JV_OK;
}

}  // extern "C"
