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
constexpr uint64_t kMaxBlocks = 1ull << 22;
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

// Non-temporal global access (streaming: the sweep kernels touch each
// amplitude once per launch).  Loads and stores are both marked: on MI355X
// an in-place streaming sweep moves ~10% more bytes per second than with
// ordinary accesses, while marking only one side gains nothing.
typedef float nt_f2 __attribute__((ext_vector_type(2)));
typedef float nt_f4 __attribute__((ext_vector_type(4)));
typedef double nt_d2 __attribute__((ext_vector_type(2)));
typedef unsigned int nt_u4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float ld_nt(const float *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ double ld_nt(const double *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float2 ld_nt(const float2 *p) {
  const nt_f2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f2 *>(p));
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ double2 ld_nt(const double2 *p) {
  const nt_d2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_d2 *>(p));
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ float4 ld_nt(const float4 *p) {
  const nt_f4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f4 *>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint4 ld_nt(const uint4 *p) {
  const nt_u4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_u4 *>(p));
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st_nt(float *p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_nt(double *p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_nt(float2 *p, float2 v) {
  const nt_f2 w = {v.x, v.y};
  __builtin_nontemporal_store(w, reinterpret_cast<nt_f2 *>(p));
}
__device__ __forceinline__ void st_nt(double2 *p, double2 v) {
  const nt_d2 w = {v.x, v.y};
  __builtin_nontemporal_store(w, reinterpret_cast<nt_d2 *>(p));
}
__device__ __forceinline__ void st_nt(float4 *p, float4 v) {
  const nt_f4 w = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(w, reinterpret_cast<nt_f4 *>(p));
}
__device__ __forceinline__ void st_nt(uint4 *p, uint4 v) {
  const nt_u4 w = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(w, reinterpret_cast<nt_u4 *>(p));
}

// A 2x2 complex matrix in the state's precision, row-major; y = m x on the
// amplitude pair (x0, x1) whose target bit is 0 / 1.
template <typename R>
struct Mat2 {
  R r[4], i[4];
};

template <typename T, typename R>
__device__ __forceinline__ void mat2_apply(const Mat2<R> &m, const T &x0, const T &x1, T &y0,
                                           T &y1) {
  y0.x = m.r[0] * x0.x - m.i[0] * x0.y + m.r[1] * x1.x - m.i[1] * x1.y;
  y0.y = m.r[0] * x0.y + m.i[0] * x0.x + m.r[1] * x1.y + m.i[1] * x1.x;
  y1.x = m.r[2] * x0.x - m.i[2] * x0.y + m.r[3] * x1.x - m.i[3] * x1.y;
  y1.y = m.r[2] * x0.y + m.i[2] * x0.x + m.r[3] * x1.y + m.i[3] * x1.x;
}

// Sum over the 64 lanes of a wave (every lane gets the total).
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Launch-shape parameters chosen by the autotuner (sq_tune_set).
struct Tuning {
  int mfma_blocks_per_cu[2] = {256, 256};  // grid cap of the MFMA kernels [c64, c128]
  int mfma5_pipeline[2] = {1, 1};        // 5-target MFMA: prefetch the next tile
  int mfma5_order[2] = {0, 0};           // 5-target MFMA: 0 cyclic tiles, 1 blocked, 2 scattered
  int mfma5_wide = 0;                    // complex64 5-target MFMA: 64-group tiles
  int mfma_tiles_per_wave = 4;           // 4-5 target MFMA, untuned programs and batches: tiles per wave
  int qaoa_mid_scatter = 0;              // QAOA window passes above bit 11: bit-reversed chunk order
  int tune_scatter = 1;                  // sq_program_tune also tries tile order 2
  int adjoint_layers = 1;                // adjoint sweeps: layer passes (12 / 8 bits a pass)
  int adjoint_wide_windows = 1;          // layer passes: 6-bit windows with 512 B runs
  int adjoint_fuse_runs = 3;             // diagonal runs folded into layer low passes (1 fwd, 2 rev)
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
// Circuit batches on the 5-target MFMA kernels: the bytes of one table
// entry; the tiles of a plan on those kernels (0 when it does not run on
// them); the entry of a DENSE plan on 5 targets of a state of nelems
// amplitudes at amplitude offset psi, its waves walking tiles_per_wave tiles
// each, its matrix appended to g3m (4 reals per element, the layout of the
// 4-5 target kernels), false when the plan does not run on those kernels;
// one launch over a table of `count` entries.
size_t cbatch_mfma_entry_bytes();
uint64_t cbatch_mfma_tiles(const GatePlan &p, bool c64, uint64_t nelems);
bool cbatch_mfma_entry(const GatePlan &p, bool c64, uint64_t nelems, uint64_t psi,
                       uint64_t first_block, uint64_t tiles_per_wave,
                       std::vector<double> *g3m, void *entry, uint64_t *nblocks);
// Tiles per wave of a launch of the 4-5 target MFMA kernels over total_tiles
// tiles (untuned programs and batches; the mfma_tiles_per_wave knob, lowered
// to keep 2048 workgroups or more).
uint64_t mfma_wave_tiles(uint64_t total_tiles);
int launch_cbatch_mfma(sq_dtype_t dt, void *states, const void *g3m, const void *tab, int count,
                       uint64_t blocks, hipStream_t st);

int launch_plan(const GatePlan &p, void *sv, sq_dtype_t dt, uint64_t nelems,
                hipStream_t stream, const void *dev_matrix);

// Device layout of a DENSE plan with kk >= 4: entry (o, r) of the canonical
// 2^kk x 2^kk matrix m[r][o] stored as {Re, Im, Re + Im, 0} at index
// o * 2^kk + r.  Returned as doubles; the complex64 path converts to float.
std::vector<double> fused_block_3m(const GatePlan &p);

int launch_init(void *sv, sq_dtype_t dt, uint64_t nelems, bool basis,
                hipStream_t stream);
// |0...0> in each of nsets consecutive states of nelems amplitudes.
int launch_init_batch(void *sv, sq_dtype_t dt, uint64_t nelems, uint32_t nsets,
                      hipStream_t stream);

// Parameter batches (one launch per gate over nsets states of nelems
// amplitudes, state y at sv + y * nelems).  plans[y] is state y's plan of the
// gate; all share their structure.  batch_gate_table appends the per-state
// kernel arguments of the gate's launch (nothing for 4-5 target blocks,
// whose matrices live in per-state block images gstride elements apart).
int batch_gate_table(const GatePlan *const *plans, uint32_t nsets, sq_dtype_t dt,
                     uint64_t nelems, std::vector<char> *tab);
int launch_batch_gate(const GatePlan &p, sq_dtype_t dt, uint64_t nelems, uint32_t nsets,
                      void *sv, const void *tab, const void *gdev, uint64_t gstride,
                      hipStream_t stream);

// Plain 16-byte-vector copy (bandwidth ceiling for the roofline).
int launch_copy(void *dst, const void *src, size_t bytes, hipStream_t stream);

}  // namespace sq
