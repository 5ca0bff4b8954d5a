// Operator strings on the device: the compiled tables of branch_free.py:348-414
// applied to one configuration (branch_free.py:443-487), the diagonal merge and
// the nonzero rule of base.py:91-160.
#pragma once
#include "common.hpp"

namespace jv {

constexpr int MAXLEN = 8;  // longest operator string handled in-kernel

struct OpDev {
  int num_ops, max_len, num_diag;
  int L;  // sites per configuration
  const int32_t* idx;
  const int32_t* map;
  const double* matel;
  const int32_t* diag;
  const double* pref;
  const int32_t* fermi;  // [num_ops][max_len] or nullptr
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
    if (op.fermi && op.fermi[base]) {
      // Jordan-Wigner sign: parity of the occupied sites after i in the
      // configuration this factor acts on (branch_free.py:463-466)
      int occ = 0;
      for (int j = i + 1; j < op.L; ++j) occ += s[j];
      for (int t = 0; t < nt; ++t)
        if (ts[t] > i) occ += tv[t] - s[ts[t]];
      if (occ & 1) m = {-m.re, -m.im};
    }
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

}  // namespace jv
