// sq_common.hpp -- shared definitions of the MI355X state-vector engine.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include "snuqs_hip.h"

namespace sq {

using cd = std::complex<double>;

// Thread-local error string behind sq_last_error(); fail() records and returns.
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define SQ_HIP(call)                                                       \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess)                                                  \
      return ::sq::fail(SQ_ERR_HIP, std::string(#call) + ": " +            \
                                        hipGetErrorString(e_));           \
  } while (0)

#define SQ_TRY(call)          \
  do {                        \
    int rc_ = (call);         \
    if (rc_ != SQ_OK) return rc_; \
  } while (0)

inline size_t itemsize(sq_dtype_t dt) { return dt == SQ_COMPLEX64 ? 8 : 16; }
inline bool valid_dtype(int dt) { return dt == SQ_COMPLEX64 || dt == SQ_COMPLEX128; }
inline bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }
inline int log2u(uint64_t x) { return 63 - __builtin_clzll(x); }

constexpr int kMaxK = 5;  // largest dense block (32 x 32)

// ---------------------------------------------------------------------------
// Gate plan: what one gate does to the amplitude array, in bit-position space
// (bit b of the amplitude index; qubit q of an n-qubit space is bit n-1-q).
//
// A gate on k targets is first rewritten in canonical order (local index bit j
// <-> j-th lowest target bit), then every target on which the matrix acts as
// |0><0| (x) I + |1><1| (x) U (or the negated form) is turned into a fixed bit
// (a control), and the remaining matrix is classified:
//   PHASE : 0 targets left, multiply the fixed-bit subspace by m[0]
//   DIAG  : diagonal on kk targets
//   PERM  : one nonzero per row/column: out[r] = m[r] * in[perm[r]]
//   DENSE : anything else
// NOOP when the gate is the identity on every touched subspace.
// ---------------------------------------------------------------------------
enum class Kind : int { NOOP = 0, PHASE = 1, DIAG = 2, PERM = 3, DENSE = 4 };

struct GatePlan {
  Kind kind = Kind::NOOP;
  int kk = 0;                 // targets left after control extraction
  int tbit[kMaxK] = {0};      // their bit positions, ascending
  uint64_t fixed_mask = 0;    // control bits
  uint64_t fixed_val = 0;     // required values of the control bits
  std::vector<cd> m;          // DENSE: 2^kk x 2^kk row-major; DIAG/PERM: 2^kk
                              // entries; PHASE: 1 entry
  int perm[1 << kMaxK] = {0}; // PERM: source local index of each output
  uint32_t moved = 0;         // PERM: local indices whose value changes
  double touched = 0.0;       // fraction of amplitudes whose value changes
  // Launch shape of the 5-target MFMA kernels chosen for this gate by
  // sq_program_tune; a negative field defers to the process-wide tuning.
  struct Shape {
    int blocks_per_cu = -1;  // grid cap / 256 CUs
    int order = -1;          // 0: tiles dealt cyclically, 1: contiguous per block, 2: cyclic over the bit-reversed tile index
    int pipe = -1;           // prefetch the next tile
    int wide = -1;           // complex64: 64-group tiles (512 B runs)
  } shape;
};

// Build the plan of `matrix` (row-major, reference local order: targets[0] is
// the MSB) on `targets` of an nbits-bit index space.  Returns SQ_OK or
// SQ_ERR_INVALID (duplicate / out-of-range targets, k > kMaxK).
// kPlanDense: no control extraction and no diagonal / permutation
// classification: a DENSE plan on all k targets (PHASE for k = 0), the form
// every matrix of a gate shares whatever its parameters.
constexpr int kPlanDense = 1;
// kPlanExact: keep a dense plan on exactly its own targets (no widening of a
// 4-target block to 5 targets for the MFMA kernels).
constexpr int kPlanExact = 2;
int make_plan(const double *matrix, const uint32_t *targets, int k,
              int num_qubits, int nbits, GatePlan *plan, int flags = 0);

// Conjugate transpose of a row-major 2^k x 2^k complex128 matrix.
std::vector<double> dagger(const double *matrix, int k);

// Per-device scratch memory owned by the library (grow-only).
struct Workspace {
  void *dev = nullptr;
  size_t dev_bytes = 0;
  void *host = nullptr;  // pinned
  size_t host_bytes = 0;
};

}  // namespace sq
