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

// A dispatch holds fewer than 2^32 work-items; sweep kernels are grid-stride
// loops over at most kMaxBlocks workgroups.
constexpr uint64_t kMaxBlocks = 1ull << 20;
inline dim3 sweep_grid(uint64_t work, uint64_t per_block) {
  uint64_t b = (work + per_block - 1) / per_block;
  if (b > kMaxBlocks) b = kMaxBlocks;
  if (b == 0) b = 1;
  return dim3((unsigned)b);
}

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

// Launch-shape parameters chosen by the autotuner (sq_tune_set).
struct Tuning {
  int mfma_blocks_per_cu[2] = {256, 256};  // grid cap of the MFMA kernels [c64, c128]
  int mfma5_pipeline[2] = {1, 1};        // 5-target MFMA: prefetch the next tile
};
Tuning &tuning();

// Host: build the IdxMap that inserts `ins_mask` bits (ascending).
inline IdxMap make_idxmap(uint64_t ins_mask, uint64_t fixed_val) {
  IdxMap m{};
  m.nins = 0;
  for (int b = 0; b < 64; ++b)
    if ((ins_mask >> b) & 1) m.lowmask[m.nins++] = (1ull << b) - 1;
  m.fixed_val = fixed_val;
  return m;
}

// Launch one plan on a state.  `dev_matrix` is the device copy of a DENSE
// plan with kk >= 4 in the layout of fused_block_3m; unused otherwise.
int launch_plan(const GatePlan &p, void *sv, sq_dtype_t dt, uint64_t nelems,
                hipStream_t stream, const void *dev_matrix);

// Device layout of a DENSE plan with kk >= 4: entry (o, r) of the canonical
// 2^kk x 2^kk matrix m[r][o] stored as {Re, Im, Re + Im, 0} at index
// o * 2^kk + r.  Returned as doubles; the complex64 path converts to float.
std::vector<double> fused_block_3m(const GatePlan &p);

int launch_init(void *sv, sq_dtype_t dt, uint64_t nelems, bool basis,
                hipStream_t stream);

// Plain 16-byte-vector copy (bandwidth ceiling for the roofline).
int launch_copy(void *dst, const void *src, size_t bytes, hipStream_t stream);

}  // namespace sq
