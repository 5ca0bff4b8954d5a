// Shared device helpers for libjvmc_hip: error plumbing, FP64 complex
// arithmetic, log-cosh, Philox4x32-10 and wavefront reductions (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/jvmc_hip.h"

namespace jv {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define JV_HIP(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return jv::hip_fail(e_, #call); \
  } while (0)

#define JV_LAUNCH_CHECK(what)                          \
  do {                                                 \
    hipError_t e_ = hipGetLastError();                 \
    if (e_ != hipSuccess) return jv::hip_fail(e_, what); \
  } while (0)

// Per-device scratch arena (grown on demand, released by jv_shutdown).
void* workspace(size_t bytes, int slot);

// Optional event timing of named kernels (jv_timer_enable / jv_timer_read).
void* timer_begin(const char* name, hipStream_t st);
void timer_end(void* tok, hipStream_t st);
struct ScopedTimer {
  void* tok;
  hipStream_t st;
  ScopedTimer(const char* name, hipStream_t s) : tok(timer_begin(name, s)), st(s) {}
  ~ScopedTimer() { timer_end(tok, st); }
};

constexpr int WAVE = 64;

// ------------------------------------------------------------------ complex
struct cd {
  double re, im;
};
__host__ __device__ inline cd mk(double r, double i) { return {r, i}; }
__host__ __device__ inline cd operator+(cd a, cd b) { return {a.re + b.re, a.im + b.im}; }
__host__ __device__ inline cd operator-(cd a, cd b) { return {a.re - b.re, a.im - b.im}; }
__host__ __device__ inline cd operator*(cd a, cd b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ inline cd operator*(double s, cd a) { return {s * a.re, s * a.im}; }
__host__ __device__ inline cd cconj(cd a) { return {a.re, -a.im}; }
__host__ __device__ inline double cabs2(cd a) { return a.re * a.re + a.im * a.im; }

__device__ inline cd cexp(cd z) {
  double s, c;
  sincos(z.im, &s, &c);
  double e = exp(z.re);
  return {e * c, e * s};
}

__device__ inline cd cdiv(cd a, cd b) {
  // Smith's algorithm
  if (fabs(b.re) >= fabs(b.im)) {
    double r = b.im / b.re, d = b.re + b.im * r;
    return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  double r = b.re / b.im, d = b.re * r + b.im;
  return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}

// log(1 + w) for |w| <= 1 (argument of the log-cosh tail).
__device__ inline cd clog1p(cd w) {
  double x = w.re, y = w.im;
  // |1+w|^2 - 1 = 2x + x^2 + y^2, evaluated without cancellation for small w
  double t = fma(x, x, fma(y, y, 2.0 * x));
  return {0.5 * log1p(t), atan2(y, 1.0 + x)};
}

constexpr double LN2 = 0.69314718055994530942;

// log(cosh(x)) = x*sgn + log1p(exp(-2 x sgn)) - log 2, sgn from the sign bit of
// Re x (activation_functions.py:19-22).
__device__ inline cd log_cosh(cd x) {
  double sg = signbit(x.re) ? -1.0 : 1.0;
  cd y = {x.re * sg, x.im * sg};
  cd e = cexp({-2.0 * y.re, -2.0 * y.im});
  cd l = clog1p(e);
  return {y.re + l.re - LN2, y.im + l.im};
}
// Real part only: Re log cosh(x) = a + 0.5 log1p(2 e^{-2a} cos 2b + e^{-4a}) - log 2
__device__ inline double re_log_cosh(cd x) {
  double a = fabs(x.re);
  double b = x.im;
  double e2 = exp(-2.0 * a);
  double t = fma(2.0 * e2, cos(2.0 * b), e2 * e2);
  return a + 0.5 * log1p(t) - LN2;
}
// Real log cosh with the reference formula |x| - log 2 + log1p(t),
// t = exp(-2|x|) in (0, 1] (activation_functions.py:19-22), written out so the
// hot loops run ~45 FP64 operations instead of the library's exp plus
// double-double log1p:
//   exp: y = -2 min(|x|, 40) = n ln2 + r, |r| <= ln2/2, e^r by a degree-11
//        polynomial (fit error 3e-18), t = 2^n e^r;
//   log1p(t) = 2 atanh(s), s = t / (2 + t) in [0, 1/3],
//        atanh(s) = s + s^3 R(s^2), R a degree-9 polynomial on [0, 1/9]
//        (fit error 5e-17); the quotient uses the hardware reciprocal with
//        two Newton steps and a residual correction (2 + t lies in [2, 3]).
// Clamping |x| at 40 changes log1p(t) by < 1e-34.  The absolute error equals
// that of the reference formula in double (3.1e-16 on |x| < 3); the
// coefficients come from fit_log_cosh.py.
constexpr double LC_EXP[12] = {2.5110037605963777e-08, 2.7632639639041029e-07, 2.755724091857897e-06,
                               2.4801485482328494e-05, 0.00019841269890047113, 0.0013888888952314775,
                               0.0083333333333196011,  0.041666666666488099,   0.1666666666666668,
                               0.50000000000000189,    1.0,                    1.0};
constexpr double LC_ATANH[10] = {0.080820228059717655, 0.044003692430723371, 0.060005364047953563,
                                 0.066570721048274675, 0.076927850324793115, 0.090908947180664293,
                                 0.11111111358713932,  0.14285714283554379,  0.20000000000007292,
                                 0.33333333333333331};
// e^{-2 min(a, 40)} for a >= 0: the exponential step of log_cosh.
__host__ __device__ inline double exp_m2(double a) {
  const double y = -2.0 * fmin(a, 40.0);
  const double n = rint(y * 1.4426950408889634);
  double r = fma(-n, 6.93147180559945286227e-01, y);
  r = fma(-n, 2.31904681384629955842e-17, r);
  double p = LC_EXP[0];
#pragma unroll
  for (int k = 1; k < 12; ++k) p = fma(p, r, LC_EXP[k]);
  return ldexp(p, (int)n);
}
__host__ __device__ inline double log_cosh(double x) {
  const double a = fabs(x);
  const double t = exp_m2(a);
  const double d = 2.0 + t;
#if defined(__HIP_DEVICE_COMPILE__)
  double q = __builtin_amdgcn_rcp(d);
  q = fma(q, fma(-d, q, 1.0), q);
  q = fma(q, fma(-d, q, 1.0), q);
  double s = t * q;
  s = fma(fma(-d, s, t), q, s);
#else
  const double s = t / d;
#endif
  const double z = s * s;
  double w = LC_ATANH[0];
#pragma unroll
  for (int k = 1; k < 10; ++k) w = fma(w, z, LC_ATANH[k]);
  return (a - LN2) + 2.0 * fma(s * z, w, s);
}
// Natural logarithm of a finite P >= 1: P = 2^e m with m in [sqrt(1/2),
// sqrt(2)), log m = 2 atanh(s), s = (m - 1) / (m + 1), |s| <= 0.1716, by the
// atanh polynomial of log_cosh (fitted on s^2 <= 1/9) and the quotient by the
// hardware reciprocal with two Newton steps and a residual correction
// (m + 1 lies in [1.7, 2.5]).  Absolute error a few 1e-16 for P <= 2^8.
__device__ inline double log_ge1(double P) {
  int e;
  double m = frexp(P, &e);  // [0.5, 1)
  if (m < 0.70710678118654752) {
    m *= 2.0;
    e -= 1;
  }
  const double num = m - 1.0, d = m + 1.0;
  double q = __builtin_amdgcn_rcp(d);
  q = fma(q, fma(-d, q, 1.0), q);
  q = fma(q, fma(-d, q, 1.0), q);
  double s = num * q;
  s = fma(fma(-d, s, num), q, s);
  const double z = s * s;
  double w = LC_ATANH[0];
#pragma unroll
  for (int k = 1; k < 10; ++k) w = fma(w, z, LC_ATANH[k]);
  const double de = (double)e;
  return fma(de, 6.93147180559945286227e-01, fma(de, 2.31904681384629955842e-17, 2.0 * fma(s * z, w, s)));
}

// Sum over the CM channels of one window element of log cosh(z_c) - A_c
// (A_c = stored log cosh of the unflipped configuration, at stride cs):
// sum (|z_c| - log 2 - A_c) + log prod (1 + e^{-2|z_c|}), one logarithm for
// the element instead of one log1p per channel.  Each factor lies in (1, 2],
// so the product (<= 2^CM) neither overflows nor loses range; the absolute
// error (a few 1e-16 per factor) matches the per-channel form's.
template <int CM, typename A>
__device__ inline double log_cosh_change(const double* z, A* Ab, int cs) {
  double prod = 1.0, lin = 0.0;
#pragma unroll
  for (int c = 0; c < CM; ++c) {
    const double a = fabs(z[c]);
    prod *= 1.0 + exp_m2(a);
    lin += (a - LN2) - Ab[c * cs];
  }
  return lin + log_ge1(prod);
}

// log cosh(x) and tanh(x) from one exponential: with y = -2 min(|x|, 40) =
// n ln2 + r and e^r = 1 + r P(r) (the polynomial of log_cosh before its last
// step), e^y - 1 = r P(r) when n = 0 (no cancellation for small |x|) and
// 2^n e^r - 1 otherwise; tanh |x| = -(e^y - 1) / (2 + (e^y - 1)), the
// quotient by the hardware reciprocal with two Newton steps and a residual
// correction (2 + (e^y - 1) lies in [1, 2]).  For |x| >= 40 tanh rounds to +-1.
__device__ inline double log_cosh_tanh(double x, double& th) {
  const double a = fabs(x);
  const double y = -2.0 * fmin(a, 40.0);
  const double n = rint(y * 1.4426950408889634);
  double r = fma(-n, 6.93147180559945286227e-01, y);
  r = fma(-n, 2.31904681384629955842e-17, r);
  double p = LC_EXP[0];
#pragma unroll
  for (int k = 1; k < 11; ++k) p = fma(p, r, LC_EXP[k]);
  const double em1r = r * p;             // e^r - 1
  const double er = fma(p, r, LC_EXP[11]);  // e^r
  const double t = ldexp(er, (int)n);
  const double em1 = n == 0.0 ? em1r : t - 1.0;
  // tanh
  const double dt = 2.0 + em1;
  double qt = __builtin_amdgcn_rcp(dt);
  qt = fma(qt, fma(-dt, qt, 1.0), qt);
  qt = fma(qt, fma(-dt, qt, 1.0), qt);
  double ht = -em1 * qt;
  ht = fma(fma(dt, -ht, -em1), qt, ht);
  th = x < 0.0 ? -ht : ht;
  // log cosh, as log_cosh
  const double d = 2.0 + t;
  double q = __builtin_amdgcn_rcp(d);
  q = fma(q, fma(-d, q, 1.0), q);
  q = fma(q, fma(-d, q, 1.0), q);
  double s = t * q;
  s = fma(fma(-d, s, t), q, s);
  const double z = s * s;
  double w = LC_ATANH[0];
#pragma unroll
  for (int k = 1; k < 10; ++k) w = fma(w, z, LC_ATANH[k]);
  return (a - LN2) + 2.0 * fma(s * z, w, s);
}

// FP32 log cosh of the mixed-precision network kernels:
// |x| - ln 2 + ln(1 + e^{-2|x|}) with the hardware base-2 exponential and
// logarithm (v_exp_f32 / v_log_f32); absolute error a few 1e-8 for |x| < 10,
// e^{-2|x|} underflows to 0 beyond |x| ~ 60, where the tail is below FP32
// resolution anyway.
__device__ inline float log_cosh(float x) {
  const float a = fabsf(x);
  const float t = __builtin_amdgcn_exp2f(-2.8853900817779268f * a);  // e^{-2a} = 2^{-2a log2 e}
  return (a - 0.6931471805599453f) + 0.6931471805599453f * __builtin_amdgcn_logf(1.0f + t);
}
// FP32 tanh from the same exponential: sign(x) (1 - t) / (1 + t), t = e^{-2|x|}.
__device__ inline float tanh_f32(float x) {
  const float t = __builtin_amdgcn_exp2f(-2.8853900817779268f * fabsf(x));
  const float r = (1.0f - t) / (1.0f + t);
  return x < 0.0f ? -r : r;
}
__device__ inline float log_cosh_tanh(float x, float& th) {
  th = tanh_f32(x);
  return log_cosh(x);
}

// tanh = d log_cosh / dx, via exp(-2|Re x|) so it never overflows.
__device__ inline cd ctanh(cd x) {
  double sg = signbit(x.re) ? -1.0 : 1.0;
  cd e = cexp({-2.0 * sg * x.re, -2.0 * sg * x.im});
  cd t = cdiv({1.0 - e.re, -e.im}, {1.0 + e.re, e.im});
  return {sg * t.re, sg * t.im};
}

// ------------------------------------------------------------------ Philox4x32-10
// Salmon et al., SC'11.  Same counter layout as the oracle.
__device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Site in [0, n) and a 53-bit uniform in [0, 1) for (step, chain, call, stream).
// Raises the issuing wave's arbitration priority.  Kernels on the solve's
// dependency chain (centring, Gram, Cholesky factor and inverse, O_k, the
// Householder reduction and the tridiagonal divide and conquer) run beside
// the compute-bound local-energy kernel on another stream; with the higher
// priority their few latency-bound waves are issued first when both are
// ready, so the chain is not slowed by the waves it shares CUs with.
__device__ inline void chain_priority() { __builtin_amdgcn_s_setprio(2); }

__device__ inline void mc_draw(uint64_t seed, uint32_t step, uint32_t chain, uint32_t call,
                               uint32_t stream, int n, int& site, double& u) {
  uint32_t c[4] = {step, chain, call, stream};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  site = (int)(((uint64_t)c[0] * (uint64_t)n) >> 32);
  double a = (double)(c[1] >> 5), b = (double)(c[2] >> 6);
  u = (a * 67108864.0 + b) * (1.0 / 9007199254740992.0);
}

// ------------------------------------------------------------------ reductions
// Wave-wide sum with DPP lane moves (no LDS round trips): a Hillis-Steele
// scan inside each 16-lane row (row_shr 1, 2, 4, 8), then row_bcast 15 / 31
// carry the row totals to lane 63, whose value every lane receives.  The
// association order is fixed, so the result is deterministic and uniform.
template <int CTRL, int ROWS>
__device__ inline double dpp_shift(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, false);
  return __hiloint2double(hi, lo);
}
__device__ inline double wave_sum(double v) {
  v += dpp_shift<0x111, 0xf>(v);  // row_shr:1
  v += dpp_shift<0x112, 0xf>(v);  // row_shr:2
  v += dpp_shift<0x114, 0xf>(v);  // row_shr:4
  v += dpp_shift<0x118, 0xf>(v);  // row_shr:8
  v += dpp_shift<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_shift<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

// Block-wide sum; red must hold blockDim.x / 64 doubles.  All threads get the sum.
__device__ inline double block_sum(double v, double* red) {
  v = wave_sum(v);
  int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  int nw = blockDim.x / WAVE;
  for (int i = 0; i < nw; ++i) s += red[i];
  __syncthreads();
  return s;
}

}  // namespace jv
