// kernels.hpp -- device-side building blocks shared by the kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "sq_common.hpp"

namespace sq {

template <typename R>
struct CV;
template <>
struct CV<float> {
  using T = float2;
};
template <>
struct CV<double> {
  using T = double2;
};

constexpr int kBlock = 256;  // threads per workgroup for the sweep kernels

// Insert zero bits at ascending positions (lowmask[j] = 2^p_j - 1) into a free
// index, then set the fixed (control) bits.
struct IdxMap {
  uint64_t lowmask[8];
  uint64_t fixed_val;
  int nins;
};

__device__ __forceinline__ uint64_t deposit(uint64_t f, const IdxMap &m) {
  for (int j = 0; j < m.nins; ++j) {
    uint64_t lo = f & m.lowmask[j];
    f = ((f ^ lo) << 1) | lo;
  }
  return f | m.fixed_val;
}

// Host: build the IdxMap that inserts `ins_mask` bits (ascending).
inline IdxMap make_idxmap(uint64_t ins_mask, uint64_t fixed_val) {
  IdxMap m{};
  m.nins = 0;
  for (int b = 0; b < 64; ++b)
    if ((ins_mask >> b) & 1) m.lowmask[m.nins++] = (1ull << b) - 1;
  m.fixed_val = fixed_val;
  return m;
}

// Launch one plan on a state.  `dev_matrix` is the device copy of the real
// block form of a DENSE plan with kk >= 4 (see fused_real_block); unused
// otherwise.
int launch_plan(const GatePlan &p, void *sv, sq_dtype_t dt, uint64_t nelems,
                hipStream_t stream, const void *dev_matrix);

// Real block form [[Re, -Im], [Im, Re]] of the canonical 2^kk x 2^kk matrix,
// interleaved so that real row 2r+c / column 2o+c' correspond to output
// amplitude r component c and input amplitude o component c', stored
// column-major (element (i, kcol) at kcol * D + i), D = 2^(kk+1).  Returned
// as doubles; the c64 launcher converts.
std::vector<double> fused_real_block(const GatePlan &p);

int launch_init(void *sv, sq_dtype_t dt, uint64_t nelems, bool basis,
                hipStream_t stream);

// Plain 16-byte-vector copy (bandwidth ceiling for the roofline).
int launch_copy(void *dst, const void *src, size_t bytes, hipStream_t stream);

}  // namespace sq
