// host_tier.cpp -- states larger than HBM: the state lives in pinned host
// memory and each pass streams it through device buffers.
//
// Replaces the reference's hybrid and CPU-offload tiers
// (src/braket/snuqs/simulation.py:286-351, transpile.py:184-298), which keep
// the state in host memory, copy GPU-sized slices in, apply the slice-local
// gates and copy them back, with Swap gates inserted to bring other qubits
// into the local range.  Here no data is permuted: a pass names which host
// bits are local (the lowest k bits plus any m - k others), and the slice for
// one assignment v of the remaining (global) bits is gathered from 2^(m-k)
// contiguous chunks of 2^k amplitudes into a compact 2^m-amplitude device
// buffer (device bit j = host bit local_bits[j]), the pass's program (gates
// rewritten to device bits) runs on it, and the chunks are scattered back.
// Copies in, compute and copies out run on three streams over a ring of
// device buffers, so PCIe traffic in both directions overlaps the kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "snuqs_hip.h"
#include "sq_common.hpp"

namespace sq {

// v's low bits deposited, in order, on the set bits of mask.
static inline uint64_t deposit_bits(uint64_t v, uint64_t mask) {
  uint64_t r = 0;
  for (uint64_t m = mask; m && v; m &= m - 1, v >>= 1)
    if (v & 1) r |= m & (~m + 1);
  return r;
}

struct PassStreams {
  hipStream_t in = nullptr, comp = nullptr, out = nullptr;
  std::vector<hipEvent_t> ev_in, ev_comp, ev_out;
  ~PassStreams() {
    for (auto *v : {&ev_in, &ev_comp, &ev_out})
      for (hipEvent_t e : *v) (void)hipEventDestroy(e);
    for (hipStream_t s : {in, comp, out})
      if (s) (void)hipStreamDestroy(s);
  }
};

}  // namespace sq

using namespace sq;

extern "C" {

int sq_host_initialize_basis_z(void *host_state, sq_dtype_t dt, uint64_t n, uint32_t threads) {
  if (!host_state) return fail(SQ_ERR_INVALID, "null host state");
  if (!valid_dtype(dt)) return fail(SQ_ERR_INVALID, "bad dtype");
  if (!is_pow2(n)) return fail(SQ_ERR_INVALID, "num_elems must be a power of two");
  const size_t bytes = n * itemsize(dt);
  const uint32_t nt = std::max<uint32_t>(1, std::min<uint32_t>(threads ? threads : 1, 64));
  const size_t per = (bytes / nt + 4095) / 4096 * 4096;
  std::vector<std::thread> pool;
  for (uint32_t t = 0; t < nt; ++t) {
    const size_t b0 = std::min(bytes, per * t), b1 = std::min(bytes, per * (t + 1));
    if (b1 > b0) pool.emplace_back([=] { std::memset((char *)host_state + b0, 0, b1 - b0); });
  }
  for (auto &th : pool) th.join();
  if (dt == SQ_COMPLEX64) ((float *)host_state)[0] = 1.0f;
  else ((double *)host_state)[0] = 1.0;
  return SQ_OK;
}

int sq_host_tier_pass(void *program, void *host_state, sq_dtype_t dt, uint64_t n,
                      uint32_t num_local, const uint32_t *local_bits, void *const *buffers,
                      uint32_t num_buffers, int zero_slices, void *stream) {
  if (!program || !host_state || !local_bits || !buffers)
    return fail(SQ_ERR_INVALID, "null argument");
  if (!valid_dtype(dt)) return fail(SQ_ERR_INVALID, "bad dtype");
  if (!is_pow2(n)) return fail(SQ_ERR_INVALID, "num_elems must be a power of two");
  const int nbits = log2u(n);
  const int m = (int)num_local;
  if (m < 1 || m > nbits) return fail(SQ_ERR_INVALID, "num_local must be in [1, log2(num_elems)]");
  if (num_buffers < 1 || num_buffers > 16) return fail(SQ_ERR_INVALID, "1 to 16 device buffers");
  uint64_t lmask = 0;
  for (int j = 0; j < m; ++j) {
    if ((int)local_bits[j] >= nbits) return fail(SQ_ERR_INVALID, "local bit beyond the state");
    if (lmask >> local_bits[j] & 1) return fail(SQ_ERR_INVALID, "repeated local bit");
    lmask |= 1ull << local_bits[j];
  }
  for (uint32_t b = 0; b < num_buffers; ++b)
    if (!buffers[b]) return fail(SQ_ERR_INVALID, "null device buffer");
  // contiguous low prefix: device bits 0..k-1 are host bits 0..k-1
  int k = 0;
  while (k < m && (int)local_bits[k] == k) ++k;
  if (k == 0) return fail(SQ_ERR_INVALID, "local_bits must start with host bit 0");
  const size_t isz = itemsize(dt);
  const uint64_t chunk = 1ull << k, nchunks = 1ull << (m - k);
  const uint64_t nslices = 1ull << (nbits - m);
  const uint64_t gmask = (n - 1) & ~lmask;
  // host offset (amplitudes) of chunk c: its bits on host bits local_bits[k..m-1]
  std::vector<uint64_t> coff(nchunks);
  for (uint64_t c = 0; c < nchunks; ++c) {
    uint64_t o = 0;
    for (int i = 0; i < m - k; ++i)
      if (c >> i & 1) o |= 1ull << local_bits[k + i];
    coff[c] = o;
  }
  PassStreams ps;
  SQ_HIP(hipStreamCreateWithFlags(&ps.in, hipStreamNonBlocking));
  SQ_HIP(hipStreamCreateWithFlags(&ps.comp, hipStreamNonBlocking));
  SQ_HIP(hipStreamCreateWithFlags(&ps.out, hipStreamNonBlocking));
  for (auto *v : {&ps.ev_in, &ps.ev_comp, &ps.ev_out})
    for (uint32_t b = 0; b < num_buffers; ++b) {
      hipEvent_t e;
      SQ_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      v->push_back(e);
    }
  // work already queued on `stream` (e.g. host-state initialisation done on
  // the device side, or an earlier pass on that stream) comes first
  hipEvent_t start;
  SQ_HIP(hipEventCreateWithFlags(&start, hipEventDisableTiming));
  SQ_HIP(hipEventRecord(start, (hipStream_t)stream));
  SQ_HIP(hipStreamWaitEvent(ps.in, start, 0));
  (void)hipEventDestroy(start);
  char *host = (char *)host_state;
  std::vector<bool> used(num_buffers, false);
  uint32_t b = 0;
  for (uint64_t v = 0; v < nslices; ++v) {
    // slices other than 0 of |0...0> are zero and stay zero under any gates
    if (zero_slices && v != 0) continue;
    char *buf = (char *)buffers[b];
    if (used[b]) SQ_HIP(hipStreamWaitEvent(ps.in, ps.ev_out[b], 0));  // written back
    const uint64_t base = deposit_bits(v, gmask);
    for (uint64_t c = 0; c < nchunks; ++c)
      SQ_HIP(hipMemcpyAsync(buf + (c << k) * isz, host + (base | coff[c]) * isz, chunk * isz,
                            hipMemcpyHostToDevice, ps.in));
    SQ_HIP(hipEventRecord(ps.ev_in[b], ps.in));
    SQ_HIP(hipStreamWaitEvent(ps.comp, ps.ev_in[b], 0));
    SQ_TRY(sq_program_run(program, buf, 1ull << m, ps.comp));
    SQ_HIP(hipEventRecord(ps.ev_comp[b], ps.comp));
    SQ_HIP(hipStreamWaitEvent(ps.out, ps.ev_comp[b], 0));
    for (uint64_t c = 0; c < nchunks; ++c)
      SQ_HIP(hipMemcpyAsync(host + (base | coff[c]) * isz, buf + (c << k) * isz, chunk * isz,
                            hipMemcpyDeviceToHost, ps.out));
    SQ_HIP(hipEventRecord(ps.ev_out[b], ps.out));
    used[b] = true;
    b = (b + 1) % num_buffers;
  }
  SQ_HIP(hipStreamSynchronize(ps.out));
  SQ_HIP(hipStreamSynchronize(ps.comp));
  SQ_HIP(hipStreamSynchronize(ps.in));
  return SQ_OK;
}

}  // extern "C"
