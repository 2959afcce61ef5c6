// chunk_layout.hpp -- register / LDS layouts of 4096-amplitude chunks, shared
// by the kernels that apply whole layers of one-qubit gates in one pass over
// a state (batched QAOA, adjoint layer passes).
#pragma once

#include "kernels.hpp"

namespace sq {

constexpr int kQL = 12;            // log2 chunk size
constexpr int kQChunk = 1 << kQL;  // amplitudes per chunk

// ---- full 4096-amplitude chunks (n >= 12): register phases ----------------
// Thread t holds 16 amplitudes of the chunk in one of three layouts, each
// putting 4 index bits in the register index r:
//   L2: j = t + 256 r (bits 8-11 in r; the global load / store layout, coalesced)
//   L1: j = (t & 15) | r << 4 | (t >> 4) << 8      (bits 4-7 in r)
//   L0: j = r | t << 4                             (bits 0-3 in r)
// RX on the r bits is applied in registers; layouts change through LDS with
// slot swz(j) = j + j / 16 (one pad slot per 16), which makes every slot a
// per-thread base plus a multiple of r -- 17 r (L0, L1: the pads shift the
// 16-element rows) or 272 r (L2) -- so the addresses fold into instruction
// offsets, and keeps the three patterns free of bank conflicts for 8 B and
// 16 B elements (L2 with 8 B elements: one 2-way pair per 32 lanes).
constexpr int kQSlots = kQChunk + kQChunk / 16;
__device__ __forceinline__ int qswz(int j) { return j + (j >> 4); }
__device__ __forceinline__ int qlay(int layout, int t, int r) {
  return layout == 2 ? t + 256 * r
         : layout == 1 ? (t & 15) | (r << 4) | ((t >> 4) << 8)
                       : r | (t << 4);
}

// qswz(qlay(layout, t, r)) in closed form, base(t) + step * r (t < 256):
//   L2: t + t/16 + 272 r;  L1: (t & 15) + 272 (t / 16) + 17 r;  L0: 17 t + r,
// so a thread keeps one LDS address per layout and the 16 slots are
// instruction offsets.
__device__ __forceinline__ int qslot_base(int layout, int t) {
  return layout == 2 ? t + (t >> 4) : layout == 1 ? (t & 15) + 272 * (t >> 4) : 17 * t;
}
__host__ __device__ constexpr int qslot_step(int layout) {
  return layout == 2 ? 272 : layout == 1 ? 17 : 1;
}

// Own 16 slots out in layout `from`, back in layout `to` (one barrier; the
// caller separates successive uses of X by a barrier).
template <typename T>
__device__ __forceinline__ void qtranspose(T (&v)[16], T *X, int t, int from, int to) {
  T *w = X + qslot_base(from, t);
  const int ws = qslot_step(from);
#pragma unroll
  for (int r = 0; r < 16; ++r) w[ws * r] = v[r];
  __syncthreads();
  const T *q = X + qslot_base(to, t);
  const int qs = qslot_step(to);
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = q[qs * r];
}

// Mid-pass geometry: W high bits h0 .. h0+W-1 beside C contiguous low bits.
// A workgroup holds 2^12 amplitudes whose local index j maps to the global
// index bit by bit: local bits 0..C-1 -> global bits 0..C-1 (2^C consecutive
// amplitudes), local bits C..C+W-1 -> h0..h0+W-1 (the bits rotated), local
// bits C+W..11 -> global bits C..11-W (carried unchanged); the chunk number
// fills the other n - 12 bits.
__host__ __device__ __forceinline__ int qmid_pos(int k, int W, int h0, int C) {
  return k < C ? k : k < C + W ? h0 + (k - C) : k - W;
}

// Chunk access.  32-bit offsets: raw buffer loads / stores on a resource at
// the chunk's base (wave-uniform), the thread's part of the byte offset as
// the VGPR offset and the register's part (uniform) as the SGPR offset, so a
// thread keeps one offset register for all 16 amplitudes of a state.  The
// caller keeps every byte it touches below 2^31 from the base (well inside
// the resource's 2^32 - 1 byte range).  64-bit offsets: flat global
// addresses.  Both non-temporal, as the other sweeps.
constexpr int kBufNT = 2;                // cache policy: non-temporal
constexpr int kBufDword3 = 0x00020000;  // gfx9 raw buffer resource word 3

template <typename T, typename O>
struct ChunkIO;

template <typename T>
struct ChunkIO<T, uint32_t> {
  __amdgpu_buffer_rsrc_t rs;
  __device__ __forceinline__ explicit ChunkIO(T *base)
      : rs(__builtin_amdgcn_make_buffer_rsrc(base, 0, -1, kBufDword3)) {}
  __device__ __forceinline__ T load(uint32_t vo, uint32_t so) const {
    if constexpr (sizeof(T) == 8)
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rs, (int)vo, (int)so, kBufNT));
    else
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b128(rs, (int)vo, (int)so, kBufNT));
  }
  __device__ __forceinline__ void store(uint32_t vo, uint32_t so, const T &v) const {
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    if constexpr (sizeof(T) == 8)
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), rs, (int)vo, (int)so, kBufNT);
    else
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, v), rs, (int)vo, (int)so, kBufNT);
  }
};

template <typename T>
struct ChunkIO<T, uint64_t> {
  char *p;
  __device__ __forceinline__ explicit ChunkIO(T *base) : p(reinterpret_cast<char *>(base)) {}
  __device__ __forceinline__ T load(uint64_t vo, uint64_t so) const {
    return ld_nt(reinterpret_cast<const T *>(p + (vo | so)));
  }
  __device__ __forceinline__ void store(uint64_t vo, uint64_t so, const T &v) const {
    st_nt(reinterpret_cast<T *>(p + (vo | so)), v);
  }
};

}  // namespace sq
