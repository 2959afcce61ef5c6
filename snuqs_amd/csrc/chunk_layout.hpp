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

// W = min(8, n - h0) high bits h0 .. h0+W-1 -----------------------------------
// A workgroup holds 2^12 amplitudes whose local index j maps to the global
// index bit by bit: local bits 0-3 -> global bits 0-3 (16 consecutive
// amplitudes, 128 B of complex64), local bits 4..3+W -> h0..h0+W-1 (the bits
// rotated), local bits 4+W..11 -> global bits 4..11-W (carried unchanged);
// the chunk number fills the other n - 12 bits.  Loads use layout L2 (local
// bits 8-11 in the register index), then one LDS transpose to L1 (bits 4-7),
// and the stores go out in L1.  Eight rotated bits per pass, against the 4-5
// a lane can hold in registers without LDS, halves the passes over high bits.
__device__ __forceinline__ int qmid_pos(int k, int W, int h0) {
  return k < 4 ? k : k < 4 + W ? h0 + (k - 4) : k - W;
}

}  // namespace sq
