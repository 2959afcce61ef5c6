// gate_plan.cpp -- host-side gate classification (no device code).
//
// The reference applies every gate with the same dense 2^k x 2^k loop
// (src/braket/snuqs/csrc/operation/gate_operations_impl_cuda.cu:25-121).  Here
// the matrix entries decide which kernel runs and which amplitudes it touches;
// the arithmetic on a touched amplitude is the same row x column sum.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <unordered_map>

#include "sq_common.hpp"

namespace sq {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}
const std::string &last_error() { return g_last_error; }

std::vector<double> dagger(const double *matrix, int k) {
  const int d = 1 << k;
  std::vector<double> out(2 * d * d);
  for (int r = 0; r < d; ++r)
    for (int c = 0; c < d; ++c) {
      out[2 * (r * d + c)] = matrix[2 * (c * d + r)];
      out[2 * (r * d + c) + 1] = -matrix[2 * (c * d + r) + 1];
    }
  return out;
}

static bool is_one(cd z) { return z.real() == 1.0 && z.imag() == 0.0; }
static bool is_zero(cd z) { return z.real() == 0.0 && z.imag() == 0.0; }

// Remove local bit j from index x (bits above j shift down).
static inline int squeeze(int x, int j) {
  return (x & ((1 << j) - 1)) | ((x >> (j + 1)) << j);
}

int make_plan(const double *matrix, const uint32_t *targets, int k,
              int num_qubits, int nbits, GatePlan *plan, int flags) {
  if (k < 0 || k > kMaxK)
    return fail(SQ_ERR_UNSUPPORTED,
                "gates on more than 5 targets are not supported (k=" +
                    std::to_string(k) + ")");
  int bits[kMaxK];
  for (int j = 0; j < k; ++j) {
    if ((int)targets[j] >= num_qubits)
      return fail(SQ_ERR_INVALID, "target " + std::to_string(targets[j]) +
                                      " out of range for " +
                                      std::to_string(num_qubits) + " qubits");
    bits[j] = num_qubits - 1 - (int)targets[j];
    if (bits[j] >= nbits)
      return fail(SQ_ERR_INVALID,
                  "target " + std::to_string(targets[j]) +
                      " has a stride beyond the state (num_qubits=" +
                      std::to_string(num_qubits) +
                      ", state holds 2^" + std::to_string(nbits) + ")");
    for (int i = 0; i < j; ++i)
      if (targets[i] == targets[j])
        return fail(SQ_ERR_INVALID, "duplicate target " +
                                        std::to_string(targets[j]));
  }
  GatePlan p;
  // Canonical order: ascending bit position.  sorted[j] = reference target
  // index of the j-th lowest bit.
  int order[kMaxK];
  for (int j = 0; j < k; ++j) order[j] = j;
  std::sort(order, order + k, [&](int a, int b) { return bits[a] < bits[b]; });
  const int d = 1 << k;
  // map reference local index -> canonical local index
  std::vector<int> to_can(d);
  for (int r = 0; r < d; ++r) {
    int c = 0;
    for (int j = 0; j < k; ++j) {
      int ref_j = order[j];  // reference target index for canonical bit j
      int bit = (r >> (k - 1 - ref_j)) & 1;
      c |= bit << j;
    }
    to_can[r] = c;
  }
  std::vector<cd> M(d * d);
  for (int r = 0; r < d; ++r)
    for (int c = 0; c < d; ++c)
      M[to_can[r] * d + to_can[c]] =
          cd(matrix[2 * (r * d + c)], matrix[2 * (r * d + c) + 1]);
  std::vector<int> tb(k);
  for (int j = 0; j < k; ++j) tb[j] = bits[order[j]];

  // Control extraction.
  const bool general = (flags & kPlanDense) != 0;
  int kk = k;
  bool changed = !general;
  while (changed && kk > 0) {
    changed = false;
    const int dd = 1 << kk;
    for (int j = 0; j < kk && !changed; ++j) {
      for (int val = 1; val >= 0 && !changed; --val) {
        // matrix must be identity on the subspace where bit j != val and
        // must not mix the two subspaces
        bool ok = true;
        for (int r = 0; r < dd && ok; ++r)
          for (int c = 0; c < dd && ok; ++c) {
            int rb = (r >> j) & 1, cb = (c >> j) & 1;
            cd z = M[r * dd + c];
            if (rb != cb) {
              if (!is_zero(z)) ok = false;
            } else if (rb != val) {
              if (r == c ? !is_one(z) : !is_zero(z)) ok = false;
            }
          }
        if (!ok) continue;
        const int nd = dd / 2;
        std::vector<cd> R(nd * nd);
        for (int r = 0; r < dd; ++r)
          for (int c = 0; c < dd; ++c)
            if (((r >> j) & 1) == val && ((c >> j) & 1) == val)
              R[squeeze(r, j) * nd + squeeze(c, j)] = M[r * dd + c];
        p.fixed_mask |= 1ull << tb[j];
        if (val) p.fixed_val |= 1ull << tb[j];
        tb.erase(tb.begin() + j);
        M.swap(R);
        --kk;
        changed = true;
      }
    }
  }
  p.kk = kk;
  for (int j = 0; j < kk; ++j) p.tbit[j] = tb[j];
  const int dd = 1 << kk;
  const double sub = 1.0 / double(1ull << __builtin_popcountll(p.fixed_mask));

  if (kk == 0) {
    p.m = {M[0]};
    p.kind = is_one(M[0]) && !general ? Kind::NOOP : Kind::PHASE;
    p.touched = p.kind == Kind::NOOP ? 0.0 : sub;
    *plan = std::move(p);
    return SQ_OK;
  }
  bool diag = !general, mono = !general;
  for (int r = 0; r < dd; ++r) {
    int nz = 0;
    for (int c = 0; c < dd; ++c) {
      if (!is_zero(M[r * dd + c])) {
        ++nz;
        if (r != c) diag = false;
      }
    }
    if (nz != 1) mono = false;
  }
  if (mono) {  // also require one nonzero per column
    std::vector<int> cnt(dd, 0);
    for (int r = 0; r < dd; ++r)
      for (int c = 0; c < dd; ++c)
        if (!is_zero(M[r * dd + c])) cnt[c]++;
    for (int c = 0; c < dd; ++c)
      if (cnt[c] != 1) mono = false;
  }
  if (diag) {
    p.kind = Kind::DIAG;
    p.m.resize(dd);
    int changed_n = 0;
    for (int r = 0; r < dd; ++r) {
      p.m[r] = M[r * dd + r];
      if (!is_one(p.m[r])) ++changed_n;
    }
    p.touched = sub * changed_n / dd;
  } else if (mono && kk <= 3) {  // larger permutations run as dense blocks
    p.kind = Kind::PERM;
    p.m.resize(dd);
    int changed_n = 0;
    for (int r = 0; r < dd; ++r)
      for (int c = 0; c < dd; ++c)
        if (!is_zero(M[r * dd + c])) {
          p.perm[r] = c;
          p.m[r] = M[r * dd + c];
          if (c != r || !is_one(p.m[r])) {
            p.moved |= 1u << r;
            ++changed_n;
          }
        }
    p.touched = sub * changed_n / dd;
  } else {
    p.kind = Kind::DENSE;
    p.m = M;
    p.touched = sub;
    // A 4-target block on a large state runs as a 5-target block with an
    // identity factor on the lowest free bit (I (x) M in that bit's place):
    // the same amplitudes change, and the 5-target MFMA kernel's staged
    // tiles sweep HBM faster than the 4-target kernel does.  Needs bit 0
    // free of controls (the complex64 tiles' 16 B accesses) and enough free
    // positions left for full tiles.
    const int nfixed = __builtin_popcountll(p.fixed_mask);
    if (kk == 4 && !(flags & kPlanExact) && !(p.fixed_mask & 1) && nbits - nfixed - 5 >= 10) {
      uint64_t used = p.fixed_mask;
      for (int j = 0; j < kk; ++j) used |= 1ull << tb[j];
      int nb = 0;
      while ((used >> nb) & 1) ++nb;
      int j0 = 0;  // local position of the new bit (tbit stays ascending)
      while (j0 < kk && tb[j0] < nb) ++j0;
      const int d5 = 32;
      std::vector<cd> W(d5 * d5, cd(0.0, 0.0));
      for (int r = 0; r < d5; ++r)
        for (int c = 0; c < d5; ++c)
          if (((r >> j0) & 1) == ((c >> j0) & 1)) W[r * d5 + c] = M[squeeze(r, j0) * dd + squeeze(c, j0)];
      tb.insert(tb.begin() + j0, nb);
      p.kk = 5;
      for (int j = 0; j < 5; ++j) p.tbit[j] = tb[j];
      p.m.swap(W);
    }
  }
  *plan = std::move(p);
  return SQ_OK;
}

}  // namespace sq
