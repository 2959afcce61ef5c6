// kernels_gate.hip -- gate-application kernels for gfx950 (MI355X).
//
// Replaces cu::applyGate and its four kernels
// (src/braket/snuqs/csrc/operation/gate_operations_impl_cuda.cu:13-175), which
// launch one thread per amplitude index and idle 50-87.5% of the lanes.  Here
// every lane owns one group of touched amplitudes: the launch covers the free
// index space only (targets and controls removed), consecutive lanes get
// consecutive free indices so each wavefront load is one contiguous run, and
// only the amplitudes the gate changes are read and written.
//
//   k_phase  : PHASE plans (GPhase, Z/S/T/PhaseShift, CZ, CPhaseShift*, ...)
//   k_diag   : DIAG plans (RotZ, ZZ, ...)
//   k_perm   : PERM plans (X, Y, CX, CY, Swap, ISwap, PSwap, CCNot, CSwap)
//   k_dense  : DENSE plans on 1-3 targets (H, V, RotX/Y, U, XX, YY, XY, Unitary)
//   k_mfma4_*, k_mfma5_* : DENSE plans on 4 / 5 targets (fused blocks) on the
//              matrix cores
//   k_dense_lds : DENSE plans on 4-5 targets for small states (VALU)
#include <vector>

#include "kernels.hpp"

namespace sq {

// ---------------------------------------------------------------------------
// initialisation: initialization_impl_cuda.cu:9-25 (initializeZero /
// initializeBasis_Z)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_init(T *__restrict__ sv, uint64_t n, int basis) {
  for (uint64_t i0 = (uint64_t)blockIdx.x * (kBlock * 4) + threadIdx.x; i0 < n;
       i0 += (uint64_t)gridDim.x * (kBlock * 4)) {
    uint64_t i = i0;
#pragma unroll
    for (int u = 0; u < 4; ++u, i += kBlock) {
      if (i < n) {
        T z;
        z.x = (basis && i == 0) ? 1 : 0;
        z.y = 0;
        sv[i] = z;
      }
    }
  }
}

int launch_init(void *sv, sq_dtype_t dt, uint64_t n, bool basis,
                hipStream_t stream) {
  const dim3 blocks = sweep_grid(n, kBlock * 4);
  if (dt == SQ_COMPLEX64)
    hipLaunchKernelGGL(k_init<float2>, blocks, dim3(kBlock), 0, stream,
                       (float2 *)sv, n, (int)basis);
  else
    hipLaunchKernelGGL(k_init<double2>, blocks, dim3(kBlock), 0, stream,
                       (double2 *)sv, n, (int)basis);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

Tuning &tuning() {
  static Tuning t;
  return t;
}

// Plain copy, 16 bytes per lane per access, 4 accesses in flight per lane.
__global__ void __launch_bounds__(kBlock)
    k_copy(uint4 *__restrict__ dst, const uint4 *__restrict__ src, uint64_t n16) {
  for (uint64_t i = (uint64_t)blockIdx.x * (kBlock * 4) + threadIdx.x; i < n16;
       i += (uint64_t)gridDim.x * (kBlock * 4)) {
    uint4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (i + (uint64_t)u * kBlock < n16) v[u] = src[i + (uint64_t)u * kBlock];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (i + (uint64_t)u * kBlock < n16) dst[i + (uint64_t)u * kBlock] = v[u];
  }
}

int launch_copy(void *dst, const void *src, size_t bytes, hipStream_t stream) {
  const uint64_t n16 = bytes / 16;
  if (n16 == 0) return SQ_OK;
  hipLaunchKernelGGL(k_copy, sweep_grid(n16, kBlock), dim3(kBlock), 0, stream,
                     (uint4 *)dst, (const uint4 *)src, n16);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

// ---------------------------------------------------------------------------
// PHASE: multiply the amplitudes whose control bits match by one scalar.
// ---------------------------------------------------------------------------
template <typename R, int U>
__global__ void __launch_bounds__(kBlock)
    k_phase(typename CV<R>::T *__restrict__ sv, uint64_t nfree, IdxMap im,
            R pr, R pi) {
  using T = typename CV<R>::T;
  for (uint64_t f0 = (uint64_t)blockIdx.x * (kBlock * U) + threadIdx.x; f0 < nfree;
       f0 += (uint64_t)gridDim.x * (kBlock * U)) {
    T v[U];
    uint64_t idx[U];
  #pragma unroll
    for (int u = 0; u < U; ++u) {
      uint64_t f = f0 + (uint64_t)u * kBlock;
      idx[u] = deposit(f, im);
      if (f < nfree) v[u] = sv[idx[u]];
    }
  #pragma unroll
    for (int u = 0; u < U; ++u) {
      if (f0 + (uint64_t)u * kBlock < nfree) {
        T o;
        o.x = pr * v[u].x - pi * v[u].y;
        o.y = pr * v[u].y + pi * v[u].x;
        sv[idx[u]] = o;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// DIAG: amplitude i (with matching control bits) times d[local index of i].
// ---------------------------------------------------------------------------
template <typename R>
struct DiagArgs {
  R dr[1 << kMaxK], di[1 << kMaxK];
  int tbit[kMaxK];
  int kk;
};

template <typename R, int U>
__global__ void __launch_bounds__(kBlock)
    k_diag(typename CV<R>::T *__restrict__ sv, uint64_t nfree, IdxMap im,
           DiagArgs<R> a) {
  using T = typename CV<R>::T;
  __shared__ R sdr[1 << kMaxK], sdi[1 << kMaxK];
  if (threadIdx.x < (1u << a.kk)) {
    sdr[threadIdx.x] = a.dr[threadIdx.x];
    sdi[threadIdx.x] = a.di[threadIdx.x];
  }
  __syncthreads();
  for (uint64_t f0 = (uint64_t)blockIdx.x * (kBlock * U) + threadIdx.x; f0 < nfree;
       f0 += (uint64_t)gridDim.x * (kBlock * U)) {
    T v[U];
    uint64_t idx[U];
  #pragma unroll
    for (int u = 0; u < U; ++u) {
      uint64_t f = f0 + (uint64_t)u * kBlock;
      idx[u] = deposit(f, im);
      if (f < nfree) v[u] = sv[idx[u]];
    }
  #pragma unroll
    for (int u = 0; u < U; ++u) {
      if (f0 + (uint64_t)u * kBlock < nfree) {
        int r = 0;
        for (int j = 0; j < a.kk; ++j) r |= (int)((idx[u] >> a.tbit[j]) & 1) << j;
        const R pr = sdr[r], pi = sdi[r];
        T o;
        o.x = pr * v[u].x - pi * v[u].y;
        o.y = pr * v[u].y + pi * v[u].x;
        sv[idx[u]] = o;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// PERM: out[r] = ph[r] * in[perm[r]] for the moved local indices r.
// src[r] / dst[r] are the amplitude offsets of perm[r] and r.
// ---------------------------------------------------------------------------
template <typename R>
struct PermArgs {
  uint64_t src[8], dst[8];
  R pr[8], pi[8];
  int count;  // number of moved local indices (<= 8)
};

template <typename R, int U>
__global__ void __launch_bounds__(kBlock)
    k_perm(typename CV<R>::T *__restrict__ sv, uint64_t nfree, IdxMap im,
           PermArgs<R> a) {
  using T = typename CV<R>::T;
  for (uint64_t f0 = (uint64_t)blockIdx.x * (kBlock * U) + threadIdx.x; f0 < nfree;
       f0 += (uint64_t)gridDim.x * (kBlock * U)) {
    T v[U][8];
    uint64_t base[U];
  #pragma unroll
    for (int u = 0; u < U; ++u) {
      uint64_t f = f0 + (uint64_t)u * kBlock;
      base[u] = deposit(f, im);
      if (f < nfree) {
  #pragma unroll
        for (int r = 0; r < 8; ++r)
          if (r < a.count) v[u][r] = sv[base[u] + a.src[r]];
      }
    }
  #pragma unroll
    for (int u = 0; u < U; ++u) {
      if (f0 + (uint64_t)u * kBlock < nfree) {
  #pragma unroll
        for (int r = 0; r < 8; ++r) {
          if (r < a.count) {
            T o;
            o.x = a.pr[r] * v[u][r].x - a.pi[r] * v[u][r].y;
            o.y = a.pr[r] * v[u][r].y + a.pi[r] * v[u][r].x;
            sv[base[u] + a.dst[r]] = o;
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// DENSE, 1-3 targets: matrix in kernel arguments (scalar registers).
// ---------------------------------------------------------------------------
template <typename R, int K>
struct DenseArgs {
  R mr[1 << (2 * K)], mi[1 << (2 * K)];
  uint64_t off[1 << K];
};

template <typename R, int K, int U>
__global__ void __launch_bounds__(kBlock)
    k_dense(typename CV<R>::T *__restrict__ sv, uint64_t nfree, IdxMap im,
            DenseArgs<R, K> a) {
  using T = typename CV<R>::T;
  constexpr int D = 1 << K;
  for (uint64_t f0 = (uint64_t)blockIdx.x * (kBlock * U) + threadIdx.x; f0 < nfree;
       f0 += (uint64_t)gridDim.x * (kBlock * U)) {
    T x[U][D];
    uint64_t base[U];
  #pragma unroll
    for (int u = 0; u < U; ++u) {
      uint64_t f = f0 + (uint64_t)u * kBlock;
      base[u] = deposit(f, im);
      if (f < nfree) {
  #pragma unroll
        for (int c = 0; c < D; ++c) x[u][c] = sv[base[u] + a.off[c]];
      }
    }
  #pragma unroll
    for (int u = 0; u < U; ++u) {
      if (f0 + (uint64_t)u * kBlock < nfree) {
  #pragma unroll
        for (int r = 0; r < D; ++r) {
          R yr = a.mr[r * D] * x[u][0].x - a.mi[r * D] * x[u][0].y;
          R yi = a.mr[r * D] * x[u][0].y + a.mi[r * D] * x[u][0].x;
  #pragma unroll
          for (int c = 1; c < D; ++c) {
            yr += a.mr[r * D + c] * x[u][c].x - a.mi[r * D + c] * x[u][c].y;
            yi += a.mr[r * D + c] * x[u][c].y + a.mi[r * D + c] * x[u][c].x;
          }
          T o;
          o.x = yr;
          o.y = yi;
          sv[base[u] + a.off[r]] = o;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Dense 4-5 target matrices live in device memory in one layout shared by the
// kernels below: entry (o, r) = {Re m[r][o], Im m[r][o], Re + Im, 0} at
// index o * 2^K + r (column-major in the complex matrix; see fused_block_3m).
// ---------------------------------------------------------------------------
template <typename R>
struct Q4;
template <>
struct Q4<float> {
  using T = float4;
};
template <>
struct Q4<double> {
  using T = HIP_vector_type<double, 4>;
};

struct OffArgs {
  uint64_t o[1 << kMaxK];
};

// DENSE, 4-5 targets, VALU: matrix staged in LDS (broadcast reads).  Used for
// states too small to fill MFMA tiles.
template <typename R, int K>
__global__ void __launch_bounds__(kBlock)
    k_dense_lds(typename CV<R>::T *__restrict__ sv, uint64_t nfree, IdxMap im,
                OffArgs off, const typename Q4<R>::T *__restrict__ G) {
  using T = typename CV<R>::T;
  constexpr int D = 1 << K;
  __shared__ R gr[D * D], gi[D * D];  // [o][r]
  for (int e = threadIdx.x; e < D * D; e += kBlock) {
    gr[e] = G[e].x;
    gi[e] = G[e].y;
  }
  __syncthreads();
  for (uint64_t f = (uint64_t)blockIdx.x * kBlock + threadIdx.x; f < nfree;
       f += (uint64_t)gridDim.x * kBlock) {
    const uint64_t base = deposit(f, im);
    T x[D];
#pragma unroll
    for (int c = 0; c < D; ++c) x[c] = sv[base + off.o[c]];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      R yr = 0, yi = 0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        const R mr = gr[c * D + r], mi = gi[c * D + r];
        yr += mr * x[c].x - mi * x[c].y;
        yi += mr * x[c].y + mi * x[c].x;
      }
      T o;
      o.x = yr;
      o.y = yi;
      sv[base + off.o[r]] = o;
    }
  }
}

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// DENSE, 4 targets on the matrix cores, real block form: the 16x16 complex G
// is the 32x32 real matrix Gr = [[Re, -Im], [Im, Re]] interleaved so real row
// 2r+c / column 2o+c' are amplitude r component c / amplitude o component c'
// (staged in LDS column-major from the shared layout).  complex64:
// v_mfma_f32_32x32x2f32, 32 groups per wave tile; complex128:
// v_mfma_f64_16x16x4f64, 16 groups per wave tile.  Lane l takes one float
// (double) of one amplitude per MFMA step.
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ void stage_real_block16(R *As, const typename Q4<R>::T *G) {
  constexpr int RD = 32;
  for (int e = threadIdx.x; e < RD * RD; e += kBlock) {
    const int kcol = e / RD, i = e % RD;
    const int r = i >> 1, c = i & 1, o = kcol >> 1, cp = kcol & 1;
    const auto z = G[o * 16 + r];
    As[e] = (c == cp) ? z.x : (c ? z.y : -z.y);
  }
}

__global__ void __launch_bounds__(kBlock)
    k_mfma4_c64(float2 *__restrict__ sv, uint64_t ntiles, IdxMap im, OffArgs off,
                const float4 *__restrict__ G) {
  constexpr int D = 32, NS = 16;
  __shared__ float As[D * D];
  stage_real_block16<float>(As, G);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 31, half = lane >> 5;
  float a[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) a[s] = As[(2 * s + half) * D + col];
  for (uint64_t tile = (uint64_t)blockIdx.x * 4 + wave; tile < ntiles;
       tile += (uint64_t)gridDim.x * 4) {
    const uint64_t base = deposit(tile * 32 + col, im);
    const float *src = reinterpret_cast<const float *>(sv + base) + half;
    float b[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) b[s] = src[2 * off.o[s]];
    floatx16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.f;
#pragma unroll
    for (int s = 0; s < NS; ++s)
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[s], b[s], acc, 0, 0, 0);
    // C layout: lane l, register v -> column l%32, row 8(v/4) + 4(l/32) + v%4.
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = 4 * q;  // + 2*half, +1
      const uint64_t o0 = half ? off.o[r + 2] : off.o[r];
      const uint64_t o1 = half ? off.o[r + 3] : off.o[r + 1];
      sv[base + o0] = make_float2(acc[4 * q + 0], acc[4 * q + 1]);
      sv[base + o1] = make_float2(acc[4 * q + 2], acc[4 * q + 3]);
    }
  }
}

__global__ void __launch_bounds__(kBlock)
    k_mfma4_c128(double2 *__restrict__ sv, uint64_t ntiles, IdxMap im, OffArgs off,
                 const typename Q4<double>::T *__restrict__ G) {
  constexpr int D = 32, NT = 2, NS = 8;
  __shared__ double As[D * D];
  stage_real_block16<double>(As, G);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, q = lane >> 4;
  for (uint64_t tile = (uint64_t)blockIdx.x * 4 + wave; tile < ntiles;
       tile += (uint64_t)gridDim.x * 4) {
    const uint64_t base = deposit(tile * 16 + col, im);
    const double *src = reinterpret_cast<const double *>(sv + base) + (q & 1);
    double b[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const uint64_t o = (q >> 1) ? off.o[2 * s + 1] : off.o[2 * s];
      b[s] = src[2 * o];
    }
    doublex4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[t][v] = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const double av = As[(4 * s + q) * D + 16 * t + col];
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b[s], acc[t], 0, 0, 0);
      }
    // C layout (v_mfma_f64_16x16x4f64): lane l, register v -> column l%16,
    // row 4v + l/16: component q%2 of output amplitude 8t + 2v + q/2.
    double *dst = reinterpret_cast<double *>(sv + base) + (q & 1);
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = 8 * t + 2 * v;
        const uint64_t o = (q >> 1) ? off.o[r + 1] : off.o[r];
        dst[2 * o] = acc[t][v];
      }
  }
}

// ---------------------------------------------------------------------------
// DENSE, 5 targets on the matrix cores: three real GEMMs per complex product
// (T1 = Gr Xr, T2 = Gi Xi, T3 = (Gr + Gi)(Xr + Xi); Yr = T1 - T2,
// Yi = T3 - T1 - T2), i.e. 48 instead of 64 MFMAs per wave tile, with whole
// complex amplitudes as operands: the B operand of a step is lane l's
// amplitude o of group l%32 (l%16 for complex128), loaded as one float2
// (double2), and each output register pair is one amplitude of one group.
// The next tile's amplitudes are loaded before the current tile's MFMAs, so
// every wave keeps a tile of loads in flight while its matrix cores run.
// A operands {Re, Im, Re + Im} are read from LDS per step (ds_read_b128).
// ---------------------------------------------------------------------------
template <bool PIPE>
__global__ void __launch_bounds__(kBlock)
    k_mfma5_c64(float2 *__restrict__ sv, uint64_t ntiles, IdxMap im, OffArgs off,
                const float4 *__restrict__ G) {
  constexpr int D = 32, NS = 16;
  __shared__ float4 As[D * D];  // [o][r]
  for (int e = threadIdx.x; e < D * D; e += kBlock) As[e] = G[e];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 31, half = lane >> 5;
  uint64_t tile = (uint64_t)blockIdx.x * 4 + wave;
  const uint64_t step = (uint64_t)gridDim.x * 4;
  if (tile >= ntiles) return;
  // Lane half h reads input amplitude o = 2s + h (local bit 0 = h) and writes
  // output amplitudes r = 8(v/4) + 4h + v%4 (local bit 2 = h): both are a
  // uniform offset plus a per-lane constant.
  const uint64_t lin = half ? off.o[1] : 0;   // stride of local bit 0
  const uint64_t lout = half ? off.o[4] : 0;  // stride of local bit 2
  float2 cur[NS], nxt[NS];
  uint64_t base = deposit(tile * 32 + col, im);
#pragma unroll
  for (int s = 0; s < NS; ++s) cur[s] = sv[base + lin + off.o[2 * s]];
  while (true) {
    const uint64_t ntile = tile + step;
    const bool more = ntile < ntiles;
    uint64_t nbase = 0;
    if (PIPE && more) {
      nbase = deposit(ntile * 32 + col, im);
#pragma unroll
      for (int s = 0; s < NS; ++s) nxt[s] = sv[nbase + lin + off.o[2 * s]];
    }
    floatx16 t1, t2, t3;
#pragma unroll
    for (int v = 0; v < 16; ++v) t1[v] = t2[v] = t3[v] = 0.f;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const float4 a = As[(2 * s + half) * D + col];
      t1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, cur[s].x, t1, 0, 0, 0);
      t2 = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, cur[s].y, t2, 0, 0, 0);
      t3 = __builtin_amdgcn_mfma_f32_32x32x2f32(a.z, cur[s].x + cur[s].y, t3, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 16; ++v)
      sv[base + lout + off.o[8 * (v / 4) + (v % 4)]] =
          make_float2(t1[v] - t2[v], t3[v] - t1[v] - t2[v]);
    if (!more) break;
    tile = ntile;
    if (PIPE) {
      base = nbase;
#pragma unroll
      for (int s = 0; s < NS; ++s) cur[s] = nxt[s];
    } else {
      base = deposit(tile * 32 + col, im);
#pragma unroll
      for (int s = 0; s < NS; ++s) cur[s] = sv[base + lin + off.o[2 * s]];
    }
  }
}

template <bool PIPE>
__global__ void __launch_bounds__(kBlock)
    k_mfma5_c128(double2 *__restrict__ sv, uint64_t ntiles, IdxMap im, OffArgs off,
                 const typename Q4<double>::T *__restrict__ G) {
  using D4 = typename Q4<double>::T;
  constexpr int D = 32, NS = 8, NT = 2;
  __shared__ D4 As[D * D];  // [o][r], 32 KB
  for (int e = threadIdx.x; e < D * D; e += kBlock) As[e] = G[e];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, q = lane >> 4;
  uint64_t tile = (uint64_t)blockIdx.x * 4 + wave;
  const uint64_t step = (uint64_t)gridDim.x * 4;
  if (tile >= ntiles) return;
  // Lane quarter q reads input amplitude o = 4s + q and writes output
  // amplitudes r = 16t + 4v + q: q is local bits 0-1 in both, so one per-lane
  // constant plus uniform offsets addresses everything.
  const uint64_t lq = ((q & 1) ? off.o[1] : 0) + ((q & 2) ? off.o[2] : 0);
  double2 cur[NS], nxt[NS];
  uint64_t base = deposit(tile * 16 + col, im);
#pragma unroll
  for (int s = 0; s < NS; ++s) cur[s] = sv[base + lq + off.o[4 * s]];
  while (true) {
    const uint64_t ntile = tile + step;
    const bool more = ntile < ntiles;
    uint64_t nbase = 0;
    if (PIPE && more) {
      nbase = deposit(ntile * 16 + col, im);
#pragma unroll
      for (int s = 0; s < NS; ++s) nxt[s] = sv[nbase + lq + off.o[4 * s]];
    }
    doublex4 t1[NT], t2[NT], t3[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) t1[t][v] = t2[t][v] = t3[t][v] = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const double xs = cur[s].x + cur[s].y;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const D4 a = As[(4 * s + q) * D + 16 * t + col];
        t1[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, cur[s].x, t1[t], 0, 0, 0);
        t2[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, cur[s].y, t2[t], 0, 0, 0);
        t3[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.z, xs, t3[t], 0, 0, 0);
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        sv[base + lq + off.o[16 * t + 4 * v]] =
            make_double2(t1[t][v] - t2[t][v], t3[t][v] - t1[t][v] - t2[t][v]);
    if (!more) break;
    tile = ntile;
    if (PIPE) {
      base = nbase;
#pragma unroll
      for (int s = 0; s < NS; ++s) cur[s] = nxt[s];
    } else {
      base = deposit(tile * 16 + col, im);
#pragma unroll
      for (int s = 0; s < NS; ++s) cur[s] = sv[base + lq + off.o[4 * s]];
    }
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
std::vector<double> fused_block_3m(const GatePlan &p) {
  const int d = 1 << p.kk;
  std::vector<double> g(4 * d * d);
  for (int r = 0; r < d; ++r)
    for (int o = 0; o < d; ++o) {
      const cd z = p.m[r * d + o];
      double *e = &g[4 * (o * d + r)];
      e[0] = z.real();
      e[1] = z.imag();
      e[2] = z.real() + z.imag();
      e[3] = 0.0;
    }
  return g;
}

static uint64_t local_offset(const GatePlan &p, int r) {
  uint64_t o = 0;
  for (int j = 0; j < p.kk; ++j)
    if ((r >> j) & 1) o += 1ull << p.tbit[j];
  return o;
}

static uint64_t target_mask(const GatePlan &p) {
  uint64_t m = 0;
  for (int j = 0; j < p.kk; ++j) m |= 1ull << p.tbit[j];
  return m;
}

static inline dim3 grid_for(uint64_t work, int per_block) {
  return sweep_grid(work, (uint64_t)per_block);
}

template <typename R>
static int launch_plan_t(const GatePlan &p, typename CV<R>::T *sv,
                         uint64_t nelems, hipStream_t st, const R *gdev) {
  const int nfixed = __builtin_popcountll(p.fixed_mask);
  switch (p.kind) {
    case Kind::NOOP:
      return SQ_OK;
    case Kind::PHASE: {
      constexpr int U = 4;
      const uint64_t nfree = nelems >> nfixed;
      IdxMap im = make_idxmap(p.fixed_mask, p.fixed_val);
      hipLaunchKernelGGL((k_phase<R, U>), grid_for(nfree, kBlock * U),
                         dim3(kBlock), 0, st, sv, nfree, im, (R)p.m[0].real(),
                         (R)p.m[0].imag());
      break;
    }
    case Kind::DIAG: {
      constexpr int U = 4;
      const uint64_t nfree = nelems >> nfixed;
      IdxMap im = make_idxmap(p.fixed_mask, p.fixed_val);
      DiagArgs<R> a{};
      for (int r = 0; r < (1 << p.kk); ++r) {
        a.dr[r] = (R)p.m[r].real();
        a.di[r] = (R)p.m[r].imag();
      }
      for (int j = 0; j < p.kk; ++j) a.tbit[j] = p.tbit[j];
      a.kk = p.kk;
      hipLaunchKernelGGL((k_diag<R, U>), grid_for(nfree, kBlock * U),
                         dim3(kBlock), 0, st, sv, nfree, im, a);
      break;
    }
    case Kind::PERM: {
      if (p.kk > 3) return fail(SQ_ERR_UNSUPPORTED, "internal: PERM plan on >3 targets");
      constexpr int U = 2;
      const uint64_t nfree = nelems >> (nfixed + p.kk);
      IdxMap im = make_idxmap(p.fixed_mask | target_mask(p), p.fixed_val);
      PermArgs<R> a{};
      a.count = 0;
      for (int r = 0; r < (1 << p.kk); ++r) {
        if (!((p.moved >> r) & 1)) continue;
        a.src[a.count] = local_offset(p, p.perm[r]);
        a.dst[a.count] = local_offset(p, r);
        a.pr[a.count] = (R)p.m[r].real();
        a.pi[a.count] = (R)p.m[r].imag();
        ++a.count;
      }
      hipLaunchKernelGGL((k_perm<R, U>), grid_for(nfree, kBlock * U),
                         dim3(kBlock), 0, st, sv, nfree, im, a);
      break;
    }
    case Kind::DENSE: {
      const uint64_t nfree = nelems >> (nfixed + p.kk);
      IdxMap im = make_idxmap(p.fixed_mask | target_mask(p), p.fixed_val);
      const int d = 1 << p.kk;
      if (p.kk <= 3) {
#define SQ_DENSE_CASE(KV, UV)                                               \
  case KV: {                                                                \
    DenseArgs<R, KV> a{};                                                   \
    for (int i = 0; i < d * d; ++i) {                                       \
      a.mr[i] = (R)p.m[i].real();                                           \
      a.mi[i] = (R)p.m[i].imag();                                           \
    }                                                                       \
    for (int r = 0; r < d; ++r) a.off[r] = local_offset(p, r);              \
    hipLaunchKernelGGL((k_dense<R, KV, UV>), grid_for(nfree, kBlock * UV),  \
                       dim3(kBlock), 0, st, sv, nfree, im, a);              \
    break;                                                                  \
  }
        switch (p.kk) {
          SQ_DENSE_CASE(1, 4)
          SQ_DENSE_CASE(2, 2)
          SQ_DENSE_CASE(3, 1)
          default:
            break;
        }
#undef SQ_DENSE_CASE
      } else {
        if (gdev == nullptr)
          return fail(SQ_ERR_INVALID, "internal: dense 4-5 target plan without device matrix");
        OffArgs off{};
        for (int r = 0; r < d; ++r) off.o[r] = local_offset(p, r);
        constexpr bool c64 = sizeof(R) == 4;
        const uint64_t tile_groups = c64 ? 32 : 16;
        const bool mfma = nfree >= tile_groups && (nfree % tile_groups) == 0;
        if (mfma) {
          const uint64_t ntiles = nfree / tile_groups;
          uint64_t blocks = (ntiles + 3) / 4;
          const int di = c64 ? 0 : 1;
          const uint64_t cap = 256ull * (uint64_t)tuning().mfma_blocks_per_cu[di];
          if (blocks > cap) blocks = cap;
          const bool pipe = tuning().mfma5_pipeline[di] != 0;
          using Q = typename Q4<R>::T;
          if constexpr (c64) {
            if (p.kk == 4)
              hipLaunchKernelGGL(k_mfma4_c64, dim3(blocks), dim3(kBlock), 0, st,
                                 (float2 *)sv, ntiles, im, off, (const Q *)gdev);
            else
              hipLaunchKernelGGL(pipe ? k_mfma5_c64<true> : k_mfma5_c64<false>, dim3(blocks),
                                 dim3(kBlock), 0, st, (float2 *)sv, ntiles, im, off,
                                 (const Q *)gdev);
          } else {
            if (p.kk == 4)
              hipLaunchKernelGGL(k_mfma4_c128, dim3(blocks), dim3(kBlock), 0, st,
                                 (double2 *)sv, ntiles, im, off, (const Q *)gdev);
            else
              hipLaunchKernelGGL(pipe ? k_mfma5_c128<true> : k_mfma5_c128<false>,
                                 dim3(blocks), dim3(kBlock), 0, st, (double2 *)sv, ntiles,
                                 im, off, (const Q *)gdev);
          }
        } else {
          using Q = typename Q4<R>::T;
          if (p.kk == 4)
            hipLaunchKernelGGL((k_dense_lds<R, 4>), grid_for(nfree, kBlock), dim3(kBlock),
                               0, st, sv, nfree, im, off, (const Q *)gdev);
          else
            hipLaunchKernelGGL((k_dense_lds<R, 5>), grid_for(nfree, kBlock), dim3(kBlock),
                               0, st, sv, nfree, im, off, (const Q *)gdev);
        }
      }
      break;
    }
  }
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

int launch_plan(const GatePlan &p, void *sv, sq_dtype_t dt, uint64_t nelems,
                hipStream_t stream, const void *dev_matrix) {
  if (dt == SQ_COMPLEX64)
    return launch_plan_t<float>(p, (float2 *)sv, nelems, stream,
                                (const float *)dev_matrix);
  return launch_plan_t<double>(p, (double2 *)sv, nelems, stream,
                               (const double *)dev_matrix);
}

}  // namespace sq
