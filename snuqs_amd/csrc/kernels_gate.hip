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
#include <cstring>
#include <vector>

#include "kernels.hpp"

namespace sq {

// ---------------------------------------------------------------------------
// initialisation: initialization_impl_cuda.cu:9-25 (initializeZero /
// initializeBasis_Z)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_init(T *__restrict__ sv, uint64_t n, int basis, uint64_t bmask) {
  for (uint64_t i0 = (uint64_t)blockIdx.x * (kBlock * 4) + threadIdx.x; i0 < n;
       i0 += (uint64_t)gridDim.x * (kBlock * 4)) {
    uint64_t i = i0;
#pragma unroll
    for (int u = 0; u < 4; ++u, i += kBlock) {
      if (i < n) {
        T z;
        z.x = (basis && (i & bmask) == 0) ? 1 : 0;  // bmask: all ones, or 2^n - 1 per state
        z.y = 0;
        st_nt(sv + i, z);
      }
    }
  }
}

static int launch_init_mask(void *sv, sq_dtype_t dt, uint64_t n, bool basis, uint64_t bmask,
                            hipStream_t stream) {
  const dim3 blocks = sweep_grid(n, kBlock * 4);
  if (dt == SQ_COMPLEX64)
    hipLaunchKernelGGL(k_init<float2>, blocks, dim3(kBlock), 0, stream,
                       (float2 *)sv, n, (int)basis, bmask);
  else
    hipLaunchKernelGGL(k_init<double2>, blocks, dim3(kBlock), 0, stream,
                       (double2 *)sv, n, (int)basis, bmask);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

int launch_init(void *sv, sq_dtype_t dt, uint64_t n, bool basis, hipStream_t stream) {
  return launch_init_mask(sv, dt, n, basis, ~0ull, stream);
}

int launch_init_batch(void *sv, sq_dtype_t dt, uint64_t nelems, uint32_t nsets,
                      hipStream_t stream) {
  return launch_init_mask(sv, dt, nelems * nsets, true, nelems - 1, stream);
}

Tuning &tuning() {
  static Tuning t;
  return t;
}

// Plain copy, 16 bytes per lane, one access per lane per grid pass
// (non-temporal both ways: the fastest streaming form measured on MI355X).
__global__ void __launch_bounds__(kBlock)
    k_copy(uint4 *__restrict__ dst, const uint4 *__restrict__ src, uint64_t n16) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n16;
       i += (uint64_t)gridDim.x * kBlock)
    st_nt(dst + i, ld_nt(src + i));
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
__device__ __forceinline__ void
    k_phase_body(typename CV<R>::T *__restrict__ sv, uint64_t nfree, IdxMap im,
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
      if (f < nfree) v[u] = ld_nt(sv + idx[u]);
    }
  #pragma unroll
    for (int u = 0; u < U; ++u) {
      if (f0 + (uint64_t)u * kBlock < nfree) {
        T o;
        o.x = pr * v[u].x - pi * v[u].y;
        o.y = pr * v[u].y + pi * v[u].x;
        st_nt(sv + idx[u], o);
      }
    }
  }
}

template <typename R, int U>
__global__ void __launch_bounds__(kBlock)
    k_phase(typename CV<R>::T *__restrict__ sv, uint64_t nfree, IdxMap im, R pr, R pi) {
  k_phase_body<R, U>(sv, nfree, im, pr, pi);
}
// Batched forms (kb_*): one launch over a batch of states, blockIdx.y = the
// state's index y in the batch; state y starts at sv + y * sstride and its
// gate arguments are tab[y] (or its block matrix G + y * gstride).
template <typename R, int U>
__global__ void __launch_bounds__(kBlock)
    kb_phase(typename CV<R>::T *__restrict__ sv, uint64_t sstride, uint64_t nfree, IdxMap im,
             const typename CV<R>::T *__restrict__ tab) {
  const typename CV<R>::T ph = tab[blockIdx.y];
  k_phase_body<R, U>(sv + blockIdx.y * sstride, nfree, im, ph.x, ph.y);
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

// A: DiagArgs by value (kernel argument) or a const reference into a batch's
// per-state table (read in place: its entries are indexed at run time, and a
// copy would sit in scratch).
template <typename R, int U, typename A = DiagArgs<R>>
__device__ __forceinline__ void
    k_diag_body(typename CV<R>::T *__restrict__ sv, uint64_t nfree, IdxMap im, A a) {
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
      if (f < nfree) v[u] = ld_nt(sv + idx[u]);
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
        st_nt(sv + idx[u], o);
      }
    }
  }
}

template <typename R, int U>
__global__ void __launch_bounds__(kBlock)
    k_diag(typename CV<R>::T *__restrict__ sv, uint64_t nfree, IdxMap im, DiagArgs<R> a) {
  k_diag_body<R, U>(sv, nfree, im, a);
}
template <typename R, int U>
__global__ void __launch_bounds__(kBlock)
    kb_diag(typename CV<R>::T *__restrict__ sv, uint64_t sstride, uint64_t nfree, IdxMap im,
            const DiagArgs<R> *__restrict__ tab) {
  k_diag_body<R, U, const DiagArgs<R> &>(sv + blockIdx.y * sstride, nfree, im, tab[blockIdx.y]);
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
__device__ __forceinline__ void
    k_perm_body(typename CV<R>::T *__restrict__ sv, uint64_t nfree, IdxMap im,
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
          if (r < a.count) v[u][r] = ld_nt(sv + base[u] + a.src[r]);
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
            st_nt(sv + base[u] + a.dst[r], o);
          }
        }
      }
    }
  }
}

template <typename R, int U>
__global__ void __launch_bounds__(kBlock)
    k_perm(typename CV<R>::T *__restrict__ sv, uint64_t nfree, IdxMap im, PermArgs<R> a) {
  k_perm_body<R, U>(sv, nfree, im, a);
}
template <typename R, int U>
__global__ void __launch_bounds__(kBlock)
    kb_perm(typename CV<R>::T *__restrict__ sv, uint64_t sstride, uint64_t nfree, IdxMap im,
            const PermArgs<R> *__restrict__ tab) {
  k_perm_body<R, U>(sv + blockIdx.y * sstride, nfree, im, tab[blockIdx.y]);
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
__device__ __forceinline__ void
    k_dense_body(typename CV<R>::T *__restrict__ sv, uint64_t nfree, IdxMap im,
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
        for (int c = 0; c < D; ++c) x[u][c] = ld_nt(sv + base[u] + a.off[c]);
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
          st_nt(sv + base[u] + a.off[r], o);
        }
      }
    }
  }
}

template <typename R, int K, int U>
__global__ void __launch_bounds__(kBlock)
    k_dense(typename CV<R>::T *__restrict__ sv, uint64_t nfree, IdxMap im, DenseArgs<R, K> a) {
  k_dense_body<R, K, U>(sv, nfree, im, a);
}
template <typename R, int K, int U>
__global__ void __launch_bounds__(kBlock)
    kb_dense(typename CV<R>::T *__restrict__ sv, uint64_t sstride, uint64_t nfree, IdxMap im,
             const DenseArgs<R, K> *__restrict__ tab) {
  k_dense_body<R, K, U>(sv + blockIdx.y * sstride, nfree, im, tab[blockIdx.y]);
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
__device__ __forceinline__ void
    k_dense_lds_body(typename CV<R>::T *__restrict__ sv, uint64_t nfree, IdxMap im,
                OffArgs off, const typename Q4<R>::T *__restrict__ G) {
  using T = typename CV<R>::T;
  constexpr int D = 1 << K;
  __shared__ R gr[D * D], gi[D * D];  // [o][r]
  __shared__ uint64_t offs[D];
  for (int e = threadIdx.x; e < D * D; e += kBlock) {
    gr[e] = G[e].x;
    gi[e] = G[e].y;
  }
  if (threadIdx.x < D) offs[threadIdx.x] = off.o[threadIdx.x];
  __syncthreads();
  for (uint64_t f = (uint64_t)blockIdx.x * kBlock + threadIdx.x; f < nfree;
       f += (uint64_t)gridDim.x * kBlock) {
    const uint64_t base = deposit(f, im);
    T x[D];
#pragma unroll
    for (int c = 0; c < D; ++c) x[c] = sv[base + offs[c]];
    // rows one at a time (not unrolled): the D inputs stay in registers and
    // the row's matrix entries are LDS broadcasts
#pragma unroll 1
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
      sv[base + offs[r]] = o;
    }
  }
}

template <typename R, int K>
__global__ void __launch_bounds__(kBlock)
    k_dense_lds(typename CV<R>::T *__restrict__ sv, uint64_t nfree, IdxMap im, OffArgs off,
                const typename Q4<R>::T *__restrict__ G) {
  k_dense_lds_body<R, K>(sv, nfree, im, off, G);
}
template <typename R, int K>
__global__ void __launch_bounds__(kBlock)
    kb_dense_lds(typename CV<R>::T *__restrict__ sv, uint64_t sstride, uint64_t nfree,
                 IdxMap im, OffArgs off, const typename Q4<R>::T *__restrict__ G,
                 uint64_t gstride) {
  k_dense_lds_body<R, K>(sv + blockIdx.y * sstride, nfree, im, off, G + blockIdx.y * gstride);
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

__device__ __forceinline__ void
    k_mfma4_c64_body(float2 *__restrict__ sv, uint64_t ntiles, IdxMap im, OffArgs off,
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
    for (int s = 0; s < NS; ++s) b[s] = ld_nt(src + 2 * off.o[s]);
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
      st_nt(sv + base + o0, make_float2(acc[4 * q + 0], acc[4 * q + 1]));
      st_nt(sv + base + o1, make_float2(acc[4 * q + 2], acc[4 * q + 3]));
    }
  }
}

__global__ void __launch_bounds__(kBlock)
    k_mfma4_c64(float2 *__restrict__ sv, uint64_t ntiles, IdxMap im, OffArgs off,
                const float4 *__restrict__ G) {
  k_mfma4_c64_body(sv, ntiles, im, off, G);
}
__global__ void __launch_bounds__(kBlock)
    kb_mfma4_c64(float2 *__restrict__ sv, uint64_t sstride, uint64_t ntiles, IdxMap im,
                 OffArgs off, const float4 *__restrict__ G, uint64_t gstride) {
  k_mfma4_c64_body(sv + blockIdx.y * sstride, ntiles, im, off, G + blockIdx.y * gstride);
}


__device__ __forceinline__ void
    k_mfma4_c128_body(double2 *__restrict__ sv, uint64_t ntiles, IdxMap im, OffArgs off,
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
      b[s] = ld_nt(src + 2 * o);
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
        st_nt(dst + 2 * o, acc[t][v]);
      }
  }
}

__global__ void __launch_bounds__(kBlock)
    k_mfma4_c128(double2 *__restrict__ sv, uint64_t ntiles, IdxMap im, OffArgs off,
                 const typename Q4<double>::T *__restrict__ G) {
  k_mfma4_c128_body(sv, ntiles, im, off, G);
}
__global__ void __launch_bounds__(kBlock)
    kb_mfma4_c128(double2 *__restrict__ sv, uint64_t sstride, uint64_t ntiles, IdxMap im,
                  OffArgs off, const typename Q4<double>::T *__restrict__ G, uint64_t gstride) {
  k_mfma4_c128_body(sv + blockIdx.y * sstride, ntiles, im, off, G + blockIdx.y * gstride);
}


// ---------------------------------------------------------------------------
// DENSE, 5 targets on the matrix cores: three real GEMMs per complex product
// (T1 = Gr Xr, T2 = Gi Xi, T3 = (Gr + Gi)(Xr + Xi); Yr = T1 - T2,
// Yi = T3 - T1 - T2), i.e. 48 instead of 64 MFMAs per wave tile, with whole
// complex amplitudes as operands.
//
// A wave tile is G groups (consecutive free indices: G = 32 for complex64,
// 16 for complex128) x the 32 amplitudes of the block, i.e. the amplitudes
// whose index bits at the 5 targets and the lowest log2(G) free positions
// vary.  Those "tile bits" sorted by position give the tile's amplitudes in
// address order j; with no control below them the lowest tile bits are the
// lowest address bits, so every run of consecutive j is contiguous in memory
// (at least 32 complex64 / 16 complex128 amplitudes = 256 B).  The tile is
// moved through LDS: lane L loads and stores 16 B at j = 2L + 128k
// (complex64: two amplitudes) or j = L + 64k (complex128), so each global
// access is a full run of 16 lanes x 16 B whatever the targets; the MFMA
// operands are then read from LDS in (group, amplitude) order.  The LDS slot
// of element j is linear over GF(2) in j's bits (per-bit weights XORed; an
// XOR swizzle keeps both access orders free of bank conflicts), so lanes
// need one address constant plus uniform per-k terms.  Loads of the next
// tile are issued before the current tile's MFMAs (PIPE), and a wave's first
// tile is requested before the workgroup stages the block matrix in LDS, so
// the two latencies overlap.
// ---------------------------------------------------------------------------
// Tiles of wave `wave` of this workgroup, by tile order ORD:
//   0 cyclic    tile 4b + w, then + 4 x grid (concurrent waves on neighbouring
//               tiles);
//   1 blocked   workgroup b owns the contiguous range [b per, (b + 1) per),
//               per a multiple of 4, its waves interleaved;
//   2 scattered the cyclic walk over the bit-reversed tile index (ntiles a
//               power of two): concurrent waves on tiles far apart in the
//               free index space.
// Which order keeps HBM busiest depends on the target bits (a tile's runs
// that differ only in high address bits compete for the same DRAM banks;
// spreading the concurrent tiles helps those), so sq_program_tune picks per
// gate.
struct TileRange {
  uint64_t first, end, step;
  int rbits;  // ORD 2: log2(ntiles), else 0
};
// Workgroup wg of nwg (the launch's grid, or a circuit batch entry's share).
template <int ORD>
__device__ __forceinline__ TileRange tile_range(uint64_t ntiles, int wave, uint64_t wg,
                                                uint64_t nwg) {
  TileRange r;
  r.rbits = 0;
  if (ORD == 1) {
    const uint64_t per = ((ntiles + nwg - 1) / nwg + 3) / 4 * 4;
    r.first = wg * per + wave;
    const uint64_t e = (wg + 1) * per;
    r.end = e < ntiles ? e : ntiles;
    r.step = 4;
  } else {
    r.first = wg * 4 + wave;
    r.end = ntiles;
    r.step = nwg * 4;
    if (ORD == 2) r.rbits = __builtin_ctzll(ntiles);
  }
  return r;
}
template <int ORD>
__device__ __forceinline__ TileRange tile_range(uint64_t ntiles, int wave) {
  return tile_range<ORD>(ntiles, wave, blockIdx.x, gridDim.x);
}
// The tile a walk position stands for.
template <int ORD>
__device__ __forceinline__ uint64_t tile_at(uint64_t t, const TileRange &r) {
  if (ORD == 2) return r.rbits ? __builtin_bitreverse64(t) >> (64 - r.rbits) : t;
  return t;
}

struct StageArgs {
  uint64_t a[10];  // address stride (amplitudes) of tile bit i
  uint32_t w[10];  // LDS slot weight of tile bit i (slots = XOR of weights)
};

// Keep LDS writes and reads of one wave in program order across lanes.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// complex64: element (group c, amplitude r) at slot 64(r>>1) + 2c + (r&1),
// swizzled by XOR ((r>>1) & 15) << 1.
__host__ __device__ constexpr uint32_t swz_c64(uint32_t x) {
  return x ^ (((x >> 6) & 15u) << 1);
}
// complex128: element (c, r) at slot 16r + c, swizzled by XOR (r & 15).
__host__ __device__ constexpr uint32_t swz_c128(uint32_t x) { return x ^ ((x >> 4) & 15u); }

// IM / SA: IdxMap and StageArgs by value (kernel arguments), or const
// references into a circuit batch's entry table (read in place: the index
// map is indexed at run time).
template <bool PIPE, int ORD, typename IM = IdxMap, typename SA = StageArgs>
__device__ __forceinline__ void
    k_mfma5_c64_body(float2 *__restrict__ sv, uint64_t ntiles, IM im, SA sa,
                const float4 *__restrict__ G, uint64_t wg, uint64_t nwg) {
  constexpr int D = 32, NS = 16;
  __shared__ float4 As[D * D];      // [o][r], 16 KB
  __shared__ float2 Xs[4][D * 32];  // one tile per wave, 32 KB
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 31, half = lane >> 5;
  float2 *X = Xs[wave];
  TileRange tr = tile_range<ORD>(ntiles, wave, wg, nwg);
  uint64_t tile = tr.first;
  const bool has = tile < tr.end;
  // j = 2 lane + 128 k + e: lane bits are tile bits 1-6, k bits tile bits 7-9
  uint64_t la = 0;
  uint32_t lw = 0;
#pragma unroll
  for (int i = 0; i < 6; ++i)
    if ((lane >> i) & 1) {
      la += sa.a[i + 1];
      lw ^= sa.w[i + 1];
    }
#define SQ_KA(k) \
  ((((k) & 1) ? sa.a[7] : 0) + (((k) & 2) ? sa.a[8] : 0) + (((k) & 4) ? sa.a[9] : 0))
#define SQ_KW(k) \
  ((((k) & 1) ? sa.w[7] : 0u) ^ (((k) & 2) ? sa.w[8] : 0u) ^ (((k) & 4) ? sa.w[9] : 0u))
  const uint32_t w0 = sa.w[0];
  float4 cur[8], nxt[8];
  uint64_t base = 0;
  if (has) {
    base = deposit(tile_at<ORD>(tile, tr) * 32, im) + la;
#pragma unroll
    for (int k = 0; k < 8; ++k) cur[k] = ld_nt(reinterpret_cast<const float4 *>(sv + base + SQ_KA(k)));
  }
  // the block matrix is staged while the first tile's loads are in flight
  for (int e = threadIdx.x; e < D * D; e += kBlock) As[e] = G[e];
  __syncthreads();
  if (!has) return;
  while (true) {
    const uint64_t ntile = tile + tr.step;
    const bool more = ntile < tr.end;
    uint64_t nbase = 0;
    if (PIPE && more) {
      nbase = deposit(tile_at<ORD>(ntile, tr) * 32, im) + la;
#pragma unroll
      for (int k = 0; k < 8; ++k)
        nxt[k] = ld_nt(reinterpret_cast<const float4 *>(sv + nbase + SQ_KA(k)));
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      X[lw ^ SQ_KW(k)] = make_float2(cur[k].x, cur[k].y);
      X[lw ^ SQ_KW(k) ^ w0] = make_float2(cur[k].z, cur[k].w);
    }
    wave_lds_sync();
    // B operand of step s: lane (col, half) takes amplitude 2s + half of group col
    float2 b[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) b[s] = X[64 * s + 2 * (col ^ s) + half];
    floatx16 t1, t2, t3;
#pragma unroll
    for (int v = 0; v < 16; ++v) t1[v] = t2[v] = t3[v] = 0.f;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const float4 a = As[(2 * s + half) * D + col];
      t1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, b[s].x, t1, 0, 0, 0);
      t2 = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, b[s].y, t2, 0, 0, 0);
      t3 = __builtin_amdgcn_mfma_f32_32x32x2f32(a.z, b[s].x + b[s].y, t3, 0, 0, 0);
    }
    wave_lds_sync();
    // C layout: lane l, register v -> group l%32, amplitude 8(v/4) + 4(l/32) + v%4
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int r = 8 * (v / 4) + 4 * half + (v % 4), rs = r >> 1;
      X[64 * rs + 2 * (col ^ (rs & 15)) + (r & 1)] =
          make_float2(t1[v] - t2[v], t3[v] - t1[v] - t2[v]);
    }
    wave_lds_sync();
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float2 y0 = X[lw ^ SQ_KW(k)], y1 = X[lw ^ SQ_KW(k) ^ w0];
      st_nt(reinterpret_cast<float4 *>(sv + base + SQ_KA(k)), make_float4(y0.x, y0.y, y1.x, y1.y));
    }
    wave_lds_sync();
    if (!more) break;
    tile = ntile;
    if constexpr (PIPE) {
      base = nbase;
#pragma unroll
      for (int k = 0; k < 8; ++k) cur[k] = nxt[k];
    } else {
      base = deposit(tile_at<ORD>(tile, tr) * 32, im) + la;
#pragma unroll
      for (int k = 0; k < 8; ++k)
        cur[k] = ld_nt(reinterpret_cast<const float4 *>(sv + base + SQ_KA(k)));
    }
  }
#undef SQ_KA
#undef SQ_KW
}

template <bool PIPE, int ORD>
__global__ void __launch_bounds__(kBlock)
    k_mfma5_c64(float2 *__restrict__ sv, uint64_t ntiles, IdxMap im, StageArgs sa,
                const float4 *__restrict__ G) {
  k_mfma5_c64_body<PIPE, ORD>(sv, ntiles, im, sa, G, blockIdx.x, gridDim.x);
}
template <bool PIPE, int ORD>
__global__ void __launch_bounds__(kBlock)
    kb_mfma5_c64(float2 *__restrict__ sv, uint64_t sstride, uint64_t ntiles, IdxMap im,
                 StageArgs sa, const float4 *__restrict__ G, uint64_t gstride) {
  k_mfma5_c64_body<PIPE, ORD>(sv + blockIdx.y * sstride, ntiles, im, sa, G + blockIdx.y * gstride,
                              blockIdx.x, gridDim.x);
}


// complex64, wide tile: 64 groups (the lowest 6 free positions) x the 32
// amplitudes of the block per wave, so every global access is 32 lanes x 16 B
// = 512 B runs (half as many runs per tile as the 32-group tile).  The tile
// sits in LDS as two 32-group halves (slot bit 10 = free position 5), each
// multiplied exactly as in k_mfma5_c64 (same operands, same MFMA order, so
// the results are bitwise those of the narrow tile).  80 KB of LDS per
// workgroup: two workgroups per CU.
struct StageArgsW {
  uint64_t a[11];  // address stride (amplitudes) of tile bit i
  uint32_t w[11];  // LDS slot weight of tile bit i
};

template <bool PIPE, int ORD>
__global__ void __launch_bounds__(kBlock)
    k_mfma5w_c64(float2 *__restrict__ sv, uint64_t ntiles, IdxMap im, StageArgsW sa,
                 const float4 *__restrict__ G) {
  constexpr int D = 32, NS = 16, NK = 16;
  __shared__ float4 As[D * D];          // [o][r], 16 KB
  __shared__ float2 Xs[4][2 * D * 32];  // one tile per wave, 64 KB
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 31, half = lane >> 5;
  float2 *X = Xs[wave];
  TileRange tr = tile_range<ORD>(ntiles, wave);
  uint64_t tile = tr.first;
  const bool has = tile < tr.end;
  // j = 2 lane + 128 k + e: lane bits are tile bits 1-6, k bits tile bits 7-10
  uint64_t la = 0;
  uint32_t lw = 0;
#pragma unroll
  for (int i = 0; i < 6; ++i)
    if ((lane >> i) & 1) {
      la += sa.a[i + 1];
      lw ^= sa.w[i + 1];
    }
#define SQ_KA(k)                                                                          \
  ((((k) & 1) ? sa.a[7] : 0) + (((k) & 2) ? sa.a[8] : 0) + (((k) & 4) ? sa.a[9] : 0) + \
   (((k) & 8) ? sa.a[10] : 0))
#define SQ_KW(k)                                                                              \
  ((((k) & 1) ? sa.w[7] : 0u) ^ (((k) & 2) ? sa.w[8] : 0u) ^ (((k) & 4) ? sa.w[9] : 0u) ^ \
   (((k) & 8) ? sa.w[10] : 0u))
  const uint32_t w0 = sa.w[0];
  float4 cur[NK], nxt[NK];
  uint64_t base = 0;
  if (has) {
    base = deposit(tile_at<ORD>(tile, tr) * 64, im) + la;
#pragma unroll
    for (int k = 0; k < NK; ++k)
      cur[k] = ld_nt(reinterpret_cast<const float4 *>(sv + base + SQ_KA(k)));
  }
  // the block matrix is staged while the first tile's loads are in flight
  for (int e = threadIdx.x; e < D * D; e += kBlock) As[e] = G[e];
  __syncthreads();
  if (!has) return;
  while (true) {
    const uint64_t ntile = tile + tr.step;
    const bool more = ntile < tr.end;
    uint64_t nbase = 0;
    if (PIPE && more) {
      nbase = deposit(tile_at<ORD>(ntile, tr) * 64, im) + la;
#pragma unroll
      for (int k = 0; k < NK; ++k)
        nxt[k] = ld_nt(reinterpret_cast<const float4 *>(sv + nbase + SQ_KA(k)));
    }
#pragma unroll
    for (int k = 0; k < NK; ++k) {
      X[lw ^ SQ_KW(k)] = make_float2(cur[k].x, cur[k].y);
      X[lw ^ SQ_KW(k) ^ w0] = make_float2(cur[k].z, cur[k].w);
    }
    wave_lds_sync();
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
      float2 *Xh = X + 1024 * h;
      float2 b[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) b[s] = Xh[64 * s + 2 * (col ^ s) + half];
      floatx16 t1, t2, t3;
#pragma unroll
      for (int v = 0; v < 16; ++v) t1[v] = t2[v] = t3[v] = 0.f;
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const float4 a = As[(2 * s + half) * D + col];
        t1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, b[s].x, t1, 0, 0, 0);
        t2 = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, b[s].y, t2, 0, 0, 0);
        t3 = __builtin_amdgcn_mfma_f32_32x32x2f32(a.z, b[s].x + b[s].y, t3, 0, 0, 0);
      }
      wave_lds_sync();
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int r = 8 * (v / 4) + 4 * half + (v % 4), rs = r >> 1;
        Xh[64 * rs + 2 * (col ^ (rs & 15)) + (r & 1)] =
            make_float2(t1[v] - t2[v], t3[v] - t1[v] - t2[v]);
      }
    }
    wave_lds_sync();
#pragma unroll
    for (int k = 0; k < NK; ++k) {
      const float2 y0 = X[lw ^ SQ_KW(k)], y1 = X[lw ^ SQ_KW(k) ^ w0];
      st_nt(reinterpret_cast<float4 *>(sv + base + SQ_KA(k)), make_float4(y0.x, y0.y, y1.x, y1.y));
    }
    wave_lds_sync();
    if (!more) break;
    tile = ntile;
    if constexpr (PIPE) {
      base = nbase;
#pragma unroll
      for (int k = 0; k < NK; ++k) cur[k] = nxt[k];
    } else {
      base = deposit(tile_at<ORD>(tile, tr) * 64, im) + la;
#pragma unroll
      for (int k = 0; k < NK; ++k)
        cur[k] = ld_nt(reinterpret_cast<const float4 *>(sv + base + SQ_KA(k)));
    }
  }
#undef SQ_KA
#undef SQ_KW
}

template <bool PIPE, int ORD, typename IM = IdxMap, typename SA = StageArgs>
__device__ __forceinline__ void
    k_mfma5_c128_body(double2 *__restrict__ sv, uint64_t ntiles, IM im, SA sa,
                 const typename Q4<double>::T *__restrict__ G, uint64_t wg, uint64_t nwg) {
  using D4 = typename Q4<double>::T;
  constexpr int D = 32, NS = 8, NT = 2;
  __shared__ D4 As[D * D];           // [o][r], 32 KB
  __shared__ double2 Xs[4][D * 16];  // one tile per wave, 32 KB
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, q = lane >> 4;
  double2 *X = Xs[wave];
  TileRange tr = tile_range<ORD>(ntiles, wave, wg, nwg);
  uint64_t tile = tr.first;
  const bool has = tile < tr.end;
  // j = lane + 64 k: lane bits are tile bits 0-5, k bits tile bits 6-8
  uint64_t la = 0;
  uint32_t lw = 0;
#pragma unroll
  for (int i = 0; i < 6; ++i)
    if ((lane >> i) & 1) {
      la += sa.a[i];
      lw ^= sa.w[i];
    }
#define SQ_KA(k) \
  ((((k) & 1) ? sa.a[6] : 0) + (((k) & 2) ? sa.a[7] : 0) + (((k) & 4) ? sa.a[8] : 0))
#define SQ_KW(k) \
  ((((k) & 1) ? sa.w[6] : 0u) ^ (((k) & 2) ? sa.w[7] : 0u) ^ (((k) & 4) ? sa.w[8] : 0u))
  double2 cur[8], nxt[8];
  uint64_t base = 0;
  if (has) {
    base = deposit(tile_at<ORD>(tile, tr) * 16, im) + la;
#pragma unroll
    for (int k = 0; k < 8; ++k) cur[k] = ld_nt(sv + base + SQ_KA(k));
  }
  // the block matrix is staged while the first tile's loads are in flight
  for (int e = threadIdx.x; e < D * D; e += kBlock) As[e] = G[e];
  __syncthreads();
  if (!has) return;
  while (true) {
    const uint64_t ntile = tile + tr.step;
    const bool more = ntile < tr.end;
    uint64_t nbase = 0;
    if (PIPE && more) {
      nbase = deposit(tile_at<ORD>(ntile, tr) * 16, im) + la;
#pragma unroll
      for (int k = 0; k < 8; ++k) nxt[k] = ld_nt(sv + nbase + SQ_KA(k));
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) X[lw ^ SQ_KW(k)] = make_double2(cur[k].x, cur[k].y);
    wave_lds_sync();
    // B operand of step s: lane (col, q) takes amplitude 4s + q of group col
    double2 b[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int o = 4 * s + q;
      b[s] = X[16 * o + (col ^ (o & 15))];
    }
    doublex4 t1[NT], t2[NT], t3[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) t1[t][v] = t2[t][v] = t3[t][v] = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const double xs = b[s].x + b[s].y;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const D4 a = As[(4 * s + q) * D + 16 * t + col];
        t1[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b[s].x, t1[t], 0, 0, 0);
        t2[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b[s].y, t2[t], 0, 0, 0);
        t3[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.z, xs, t3[t], 0, 0, 0);
      }
    }
    wave_lds_sync();
    // C layout (v_mfma_f64_16x16x4f64): lane l, register v of block t ->
    // group l%16, amplitude 16t + 4v + l/16
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = 16 * t + 4 * v + q;
        X[16 * r + (col ^ (r & 15))] =
            make_double2(t1[t][v] - t2[t][v], t3[t][v] - t1[t][v] - t2[t][v]);
      }
    wave_lds_sync();
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const double2 y = X[lw ^ SQ_KW(k)];
      st_nt(sv + base + SQ_KA(k), make_double2(y.x, y.y));
    }
    wave_lds_sync();
    if (!more) break;
    tile = ntile;
    if constexpr (PIPE) {
      base = nbase;
#pragma unroll
      for (int k = 0; k < 8; ++k) cur[k] = nxt[k];
    } else {
      base = deposit(tile_at<ORD>(tile, tr) * 16, im) + la;
#pragma unroll
      for (int k = 0; k < 8; ++k) cur[k] = ld_nt(sv + base + SQ_KA(k));
    }
  }
#undef SQ_KA
#undef SQ_KW
}

template <bool PIPE, int ORD>
__global__ void __launch_bounds__(kBlock)
    k_mfma5_c128(double2 *__restrict__ sv, uint64_t ntiles, IdxMap im, StageArgs sa,
                 const typename Q4<double>::T *__restrict__ G) {
  k_mfma5_c128_body<PIPE, ORD>(sv, ntiles, im, sa, G, blockIdx.x, gridDim.x);
}
template <bool PIPE, int ORD>
__global__ void __launch_bounds__(kBlock)
    kb_mfma5_c128(double2 *__restrict__ sv, uint64_t sstride, uint64_t ntiles, IdxMap im,
                  StageArgs sa, const typename Q4<double>::T *__restrict__ G, uint64_t gstride) {
  k_mfma5_c128_body<PIPE, ORD>(sv + blockIdx.y * sstride, ntiles, im, sa,
                               G + blockIdx.y * gstride, blockIdx.x, gridDim.x);
}


// ---------------------------------------------------------------------------
// DENSE (1-3 targets) and PERM gates whose targets or controls sit on the
// lowest address bits.  Lane-per-group kernels would read 8-16 B per lane at
// a stride there (a group's amplitudes are neighbours), which wastes most of
// every cache line each load instruction touches.  Instead a wave moves a
// tile of 2^(k+6) amplitudes -- the k targets plus the six lowest other bit
// positions (controls or free) -- through LDS in address order with 16 B per
// lane (complex64: two amplitudes; every run of consecutive tile amplitudes
// is contiguous in memory, at least 64 amplitudes long), then each of up to
// 64 lanes applies the gate to one group in LDS.  Amplitudes whose control
// bits do not match pass through unchanged.  LDS slots are tile-local indices
// (address order), so the global<->LDS copies are contiguous.
// ---------------------------------------------------------------------------
struct SmallTile {
  uint64_t a[9];       // address stride (amplitudes) of tile bit i
  uint64_t fval;       // values of the controls outside the tile
  uint32_t cw[6];      // compute phase: tile-local weight of lane bit b
  uint32_t cfix;       // tile-local index bits of the controls inside the tile
  uint32_t off[8];     // tile-local offset of target-local index r
  int nact;            // lanes with a group in the compute phase
  int shift;           // tile positions 0..shift-1: inserted by a shift
};

// Tile base: t's bits spread over the positions outside the tile and the
// outside controls (zeros at those), shifted up past the low tile prefix.
__device__ __forceinline__ uint64_t small_tile_base(uint64_t t, const IdxMap &im,
                                                    const SmallTile &st) {
  return (deposit(t, im) << st.shift) | st.fval;
}

template <typename R, int TB>
__device__ __forceinline__ void small_tile_load(const typename CV<R>::T *sv, uint64_t base,
                                                typename CV<R>::T *X, int lane,
                                                const SmallTile &st) {
  if constexpr (sizeof(R) == 4) {
    // chunk ch = lane + 64 k holds tile amplitudes 2ch, 2ch + 1
    constexpr int CH = 1 << (TB - 7);
    uint64_t la = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i)
      if ((lane >> i) & 1) la += st.a[i + 1];
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      uint64_t ka = 0;
#pragma unroll
      for (int b = 0; b < TB - 7; ++b)
        if ((k >> b) & 1) ka += st.a[7 + b];
      const float4 v = ld_nt(reinterpret_cast<const float4 *>(sv + base + la + ka));
      *reinterpret_cast<float4 *>(X + 2 * (lane + 64 * k)) = v;
    }
  } else {
    constexpr int CH = 1 << (TB - 6);
    uint64_t la = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i)
      if ((lane >> i) & 1) la += st.a[i];
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      uint64_t ka = 0;
#pragma unroll
      for (int b = 0; b < TB - 6; ++b)
        if ((k >> b) & 1) ka += st.a[6 + b];
      const double2 v = ld_nt(sv + base + la + ka);
      X[lane + 64 * k] = make_double2(v.x, v.y);
    }
  }
}

template <typename R, int TB>
__device__ __forceinline__ void small_tile_store(typename CV<R>::T *sv, uint64_t base,
                                                 const typename CV<R>::T *X, int lane,
                                                 const SmallTile &st) {
  if constexpr (sizeof(R) == 4) {
    constexpr int CH = 1 << (TB - 7);
    uint64_t la = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i)
      if ((lane >> i) & 1) la += st.a[i + 1];
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      uint64_t ka = 0;
#pragma unroll
      for (int b = 0; b < TB - 7; ++b)
        if ((k >> b) & 1) ka += st.a[7 + b];
      const float4 v = *reinterpret_cast<const float4 *>(X + 2 * (lane + 64 * k));
      st_nt(reinterpret_cast<float4 *>(sv + base + la + ka), v);
    }
  } else {
    constexpr int CH = 1 << (TB - 6);
    uint64_t la = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i)
      if ((lane >> i) & 1) la += st.a[i];
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      uint64_t ka = 0;
#pragma unroll
      for (int b = 0; b < TB - 6; ++b)
        if ((k >> b) & 1) ka += st.a[6 + b];
      const double2 v = X[lane + 64 * k];
      st_nt(sv + base + la + ka, make_double2(v.x, v.y));
    }
  }
}

// Tile-local index of lane's group (compute phase).
__device__ __forceinline__ uint32_t small_group(int lane, const SmallTile &st) {
  uint32_t c = st.cfix;
#pragma unroll
  for (int b = 0; b < 6; ++b)
    if ((lane >> b) & 1) c |= st.cw[b];
  return c;
}

template <typename R, int K>
__device__ __forceinline__ void
    k_small_dense_body(typename CV<R>::T *__restrict__ sv, uint64_t ntiles, IdxMap im,
                  SmallTile st, DenseArgs<R, K> a) {
  using T = typename CV<R>::T;
  constexpr int TB = K + 6, D = 1 << K;
  __shared__ T Xs[4][1 << TB];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T *X = Xs[wave];
  const TileRange tr = tile_range<0>(ntiles, wave);
  const uint32_t cb = small_group(lane, st);
  for (uint64_t t = tr.first; t < tr.end; t += tr.step) {
    const uint64_t base = small_tile_base(t, im, st);
    small_tile_load<R, TB>(sv, base, X, lane, st);
    wave_lds_sync();
    if (lane < st.nact) {
      T x[D];
#pragma unroll
      for (int c = 0; c < D; ++c) x[c] = X[cb + st.off[c]];
#pragma unroll
      for (int r = 0; r < D; ++r) {
        R yr = a.mr[r * D] * x[0].x - a.mi[r * D] * x[0].y;
        R yi = a.mr[r * D] * x[0].y + a.mi[r * D] * x[0].x;
#pragma unroll
        for (int c = 1; c < D; ++c) {
          yr += a.mr[r * D + c] * x[c].x - a.mi[r * D + c] * x[c].y;
          yi += a.mr[r * D + c] * x[c].y + a.mi[r * D + c] * x[c].x;
        }
        T o;
        o.x = yr;
        o.y = yi;
        X[cb + st.off[r]] = o;
      }
    }
    wave_lds_sync();
    small_tile_store<R, TB>(sv, base, X, lane, st);
    wave_lds_sync();
  }
}

template <typename R, int K>
__global__ void __launch_bounds__(kBlock)
    k_small_dense(typename CV<R>::T *__restrict__ sv, uint64_t ntiles, IdxMap im, SmallTile st,
                  DenseArgs<R, K> a) {
  k_small_dense_body<R, K>(sv, ntiles, im, st, a);
}
template <typename R, int K>
__global__ void __launch_bounds__(kBlock)
    kb_small_dense(typename CV<R>::T *__restrict__ sv, uint64_t sstride, uint64_t ntiles,
                   IdxMap im, SmallTile st, const DenseArgs<R, K> *__restrict__ tab) {
  k_small_dense_body<R, K>(sv + blockIdx.y * sstride, ntiles, im, st, tab[blockIdx.y]);
}


// PERM on the tile: out[dst[i]] = ph[i] * in[src[i]] for the moved local
// indices; src/dst are tile-local offsets from the group's index.
template <typename R>
struct SmallPermArgs {
  uint32_t src[8], dst[8];
  R pr[8], pi[8];
  int count;
};

template <typename R, int K>
__device__ __forceinline__ void
    k_small_perm_body(typename CV<R>::T *__restrict__ sv, uint64_t ntiles, IdxMap im,
                 SmallTile st, SmallPermArgs<R> a) {
  using T = typename CV<R>::T;
  constexpr int TB = K + 6, D = 1 << K;
  __shared__ T Xs[4][1 << TB];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T *X = Xs[wave];
  const TileRange tr = tile_range<0>(ntiles, wave);
  const uint32_t cb = small_group(lane, st);
  for (uint64_t t = tr.first; t < tr.end; t += tr.step) {
    const uint64_t base = small_tile_base(t, im, st);
    small_tile_load<R, TB>(sv, base, X, lane, st);
    wave_lds_sync();
    if (lane < st.nact) {
      T v[D];
#pragma unroll
      for (int i = 0; i < D; ++i)
        if (i < a.count) v[i] = X[cb + a.src[i]];
#pragma unroll
      for (int i = 0; i < D; ++i) {
        if (i < a.count) {
          T o;
          o.x = a.pr[i] * v[i].x - a.pi[i] * v[i].y;
          o.y = a.pr[i] * v[i].y + a.pi[i] * v[i].x;
          X[cb + a.dst[i]] = o;
        }
      }
    }
    wave_lds_sync();
    small_tile_store<R, TB>(sv, base, X, lane, st);
    wave_lds_sync();
  }
}

template <typename R, int K>
__global__ void __launch_bounds__(kBlock)
    k_small_perm(typename CV<R>::T *__restrict__ sv, uint64_t ntiles, IdxMap im, SmallTile st,
                 SmallPermArgs<R> a) {
  k_small_perm_body<R, K>(sv, ntiles, im, st, a);
}
template <typename R, int K>
__global__ void __launch_bounds__(kBlock)
    kb_small_perm(typename CV<R>::T *__restrict__ sv, uint64_t sstride, uint64_t ntiles,
                  IdxMap im, SmallTile st, const SmallPermArgs<R> *__restrict__ tab) {
  k_small_perm_body<R, K>(sv + blockIdx.y * sstride, ntiles, im, st, tab[blockIdx.y]);
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

// Tile geometry of the 5-target MFMA kernels: the 5 targets and the lowest
// `nfb` free positions (5 for complex64, 4 for complex128) in ascending order,
// each with its address stride and its LDS slot weight (see StageArgs).
// False when a control sits below the tile (complex64 then lacks the
// contiguous amplitude pairs its 16 B accesses need).
static bool make_stage(const GatePlan &p, bool c64, StageArgs *sa) {
  const int nfb = c64 ? 5 : 4;
  const uint64_t tmask = target_mask(p);
  int i = 0, nf = 0;
  for (int b = 0; b < 64 && i < nfb + p.kk; ++b) {
    const uint64_t bit = 1ull << b;
    if (p.fixed_mask & bit) continue;
    uint32_t slot;
    if (tmask & bit) {
      int q = 0;
      while (p.tbit[q] != b) ++q;
      slot = c64 ? (q == 0 ? 1u : 1u << (q + 5)) : 16u << q;
    } else {
      if (nf == nfb) continue;
      slot = c64 ? 2u << nf : 1u << nf;
      ++nf;
    }
    sa->a[i] = bit;
    sa->w[i] = c64 ? swz_c64(slot) : swz_c128(slot);
    ++i;
  }
  return i == nfb + p.kk && (!c64 || sa->a[0] == 1);
}

// Geometry of the wide complex64 tile: the 5 targets and the lowest 6 free
// positions; free position 5 selects the LDS half (slot weight 1024).
static bool make_stage_wide(const GatePlan &p, StageArgsW *sa) {
  const uint64_t tmask = target_mask(p);
  int i = 0, nf = 0;
  for (int b = 0; b < 64 && i < 6 + p.kk; ++b) {
    const uint64_t bit = 1ull << b;
    if (p.fixed_mask & bit) continue;
    uint32_t slot;
    if (tmask & bit) {
      int q = 0;
      while (p.tbit[q] != b) ++q;
      slot = q == 0 ? 1u : 1u << (q + 5);
    } else {
      if (nf == 6) continue;
      slot = nf < 5 ? 2u << nf : 1024u;
      ++nf;
    }
    sa->a[i] = bit;
    sa->w[i] = swz_c64(slot);
    ++i;
  }
  return p.kk == 5 && i == 11 && sa->a[0] == 1;
}

// LDS-tiled path of a <= 3-target DENSE / PERM plan: taken when a target or
// control sits below bit 4 (complex64) / 3 (complex128), where lane-strided
// accesses would touch runs shorter than 128 B.  Tile = the targets plus the
// lowest kk + 6 - kk other positions (controls there stay in the tile and
// pass through).  False when the state is too small for it.
static bool make_small_tile(const GatePlan &p, uint64_t nelems, bool c64, SmallTile *st,
                            IdxMap *im, uint64_t *ntiles) {
  if (p.kk < 1 || p.kk > 3 || nelems == 0 || (nelems & (nelems - 1))) return false;
  const int nbits = __builtin_ctzll(nelems);
  const uint64_t tmask = target_mask(p);
  const int lowest = __builtin_ctzll(tmask | p.fixed_mask);
  if (lowest >= (c64 ? 4 : 3)) return false;
  const int TB = p.kk + 6;
  uint64_t tile = tmask;
  for (int b = 0; b < nbits && __builtin_popcountll(tile) < TB; ++b) tile |= 1ull << b;
  if (__builtin_popcountll(tile) != TB) return false;
  const uint64_t fin = p.fixed_mask & tile, fout = p.fixed_mask & ~tile;
  const uint64_t ins = tile | fout;
  const int nins = __builtin_popcountll(ins);
  if (nbits < nins + 2) return false;  // at least 4 tiles
  int m = 0;
  while ((ins >> m) & 1) ++m;
  if (__builtin_popcountll(ins >> m) > 8) return false;
  *st = SmallTile{};
  int idx[64];
  int i = 0, f = 0;
  for (int b = 0; b < 64; ++b) {
    if (!((tile >> b) & 1)) continue;
    idx[b] = i;
    st->a[i] = 1ull << b;
    if (!((tmask >> b) & 1)) {
      if ((fin >> b) & 1) {
        if ((p.fixed_val >> b) & 1) st->cfix |= 1u << i;
      } else {
        st->cw[f++] = 1u << i;
      }
    }
    ++i;
  }
  st->nact = 1 << f;
  for (int r = 0; r < (1 << p.kk); ++r) {
    uint32_t o = 0;
    for (int q = 0; q < p.kk; ++q)
      if ((r >> q) & 1) o |= 1u << idx[p.tbit[q]];
    st->off[r] = o;
  }
  st->shift = m;
  st->fval = p.fixed_val & fout;
  *im = make_idxmap(ins >> m, 0);
  *ntiles = nelems >> nins;
  return true;
}

static inline dim3 grid_for(uint64_t work, int per_block) {
  return sweep_grid(work, (uint64_t)per_block);
}

// Lane-per-group dense sweeps: one group per lane per grid pass over at most
// 2^20 workgroups (4 passes at 2^30 groups).  Measured on MI355X this beats
// several groups per lane per pass for the dense kernels (e.g. H 6.6 vs
// 6.3 TB/s), while the diagonal, phase and permutation kernels keep 4 / 2
// elements per lane (one per pass made them slower).
constexpr uint64_t kSweepBlocks = 1ull << 20;
static inline dim3 grid_sweep1(uint64_t work) {
  uint64_t b = (work + kBlock - 1) / kBlock;
  if (b > kSweepBlocks) b = kSweepBlocks;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

// The 5-target kernel of a launch shape: prefetch or not, tile order 0-2.
template <typename K>
static K mfma5_kernel(K p0, K p1, K p2, K n0, K n1, K n2, bool pipe, int ord) {
  if (pipe) return ord == 1 ? p1 : ord == 2 ? p2 : p0;
  return ord == 1 ? n1 : ord == 2 ? n2 : n0;
}
// The scattered order walks a bit-reversed tile index: ntiles must be a power
// of two (it is for every full state; otherwise the cyclic order runs).
static int scatter_ok(int ord, uint64_t ntiles) {
  return ord == 2 && (ntiles & (ntiles - 1)) != 0 ? 0 : ord;
}

// Kernel arguments of a plan (shared by single-state and batched launches).
template <typename R>
static DiagArgs<R> diag_args(const GatePlan &p) {
  DiagArgs<R> a{};
  for (int r = 0; r < (1 << p.kk); ++r) {
    a.dr[r] = (R)p.m[r].real();
    a.di[r] = (R)p.m[r].imag();
  }
  for (int j = 0; j < p.kk; ++j) a.tbit[j] = p.tbit[j];
  a.kk = p.kk;
  return a;
}

template <typename R>
static SmallPermArgs<R> small_perm_args(const GatePlan &p, const SmallTile &stile) {
  SmallPermArgs<R> a{};
  for (int r = 0; r < (1 << p.kk); ++r) {
    if (!((p.moved >> r) & 1)) continue;
    a.src[a.count] = stile.off[p.perm[r]];
    a.dst[a.count] = stile.off[r];
    a.pr[a.count] = (R)p.m[r].real();
    a.pi[a.count] = (R)p.m[r].imag();
    ++a.count;
  }
  return a;
}

template <typename R>
static PermArgs<R> perm_args(const GatePlan &p) {
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
  return a;
}

template <typename R, int K>
static DenseArgs<R, K> dense_args(const GatePlan &p) {
  constexpr int d = 1 << K;
  DenseArgs<R, K> a{};
  for (int i = 0; i < d * d; ++i) {
    a.mr[i] = (R)p.m[i].real();
    a.mi[i] = (R)p.m[i].imag();
  }
  for (int r = 0; r < d; ++r) a.off[r] = local_offset(p, r);
  return a;
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
      const DiagArgs<R> a = diag_args<R>(p);
      hipLaunchKernelGGL((k_diag<R, U>), grid_for(nfree, kBlock * U),
                         dim3(kBlock), 0, st, sv, nfree, im, a);
      break;
    }
    case Kind::PERM: {
      if (p.kk > 3) return fail(SQ_ERR_UNSUPPORTED, "internal: PERM plan on >3 targets");
      SmallTile stile;
      IdxMap sim;
      uint64_t ntiles;
      if (make_small_tile(p, nelems, sizeof(R) == 4, &stile, &sim, &ntiles)) {
        const SmallPermArgs<R> a = small_perm_args<R>(p, stile);
        const dim3 grid = sweep_grid(ntiles, 4);
        if (p.kk == 1)
          hipLaunchKernelGGL((k_small_perm<R, 1>), grid, dim3(kBlock), 0, st, sv, ntiles, sim, stile, a);
        else if (p.kk == 2)
          hipLaunchKernelGGL((k_small_perm<R, 2>), grid, dim3(kBlock), 0, st, sv, ntiles, sim, stile, a);
        else
          hipLaunchKernelGGL((k_small_perm<R, 3>), grid, dim3(kBlock), 0, st, sv, ntiles, sim, stile, a);
        break;
      }
      constexpr int U = 2;
      const uint64_t nfree = nelems >> (nfixed + p.kk);
      IdxMap im = make_idxmap(p.fixed_mask | target_mask(p), p.fixed_val);
      const PermArgs<R> a = perm_args<R>(p);
      hipLaunchKernelGGL((k_perm<R, U>), grid_for(nfree, kBlock * U),
                         dim3(kBlock), 0, st, sv, nfree, im, a);
      break;
    }
    case Kind::DENSE: {
      const uint64_t nfree = nelems >> (nfixed + p.kk);
      IdxMap im = make_idxmap(p.fixed_mask | target_mask(p), p.fixed_val);
      const int d = 1 << p.kk;
      if (p.kk <= 3) {
        SmallTile stile;
        IdxMap sim;
        uint64_t ntiles = 0;
        const bool small = make_small_tile(p, nelems, sizeof(R) == 4, &stile, &sim, &ntiles);
#define SQ_DENSE_CASE(KV, UV)                                                   \
  case KV: {                                                                    \
    const DenseArgs<R, KV> a = dense_args<R, KV>(p);                            \
    if (small)                                                                  \
      hipLaunchKernelGGL((k_small_dense<R, KV>), sweep_grid(ntiles, 4),         \
                         dim3(kBlock), 0, st, sv, ntiles, sim, stile, a);       \
    else                                                                        \
      hipLaunchKernelGGL((k_dense<R, KV, UV>), grid_sweep1(nfree),              \
                         dim3(kBlock), 0, st, sv, nfree, im, a);                \
    break;                                                                      \
  }
        switch (p.kk) {
          SQ_DENSE_CASE(1, 1)
          SQ_DENSE_CASE(2, 1)
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
        StageArgs sa{};
        // complex64 tiles move 16 B (two amplitudes) per access: the state must
        // start on a 16 B boundary (a slice at an odd amplitude does not)
        const bool aligned = !c64 || ((uintptr_t)sv & 15) == 0;
        const bool mfma = nfree >= tile_groups && (nfree % tile_groups) == 0 &&
                          (p.kk == 4 || (aligned && make_stage(p, c64, &sa)));
        if (mfma) {
          const uint64_t ntiles = nfree / tile_groups;
          uint64_t blocks = (ntiles + 3) / 4;
          const int di = c64 ? 0 : 1;
          const Tuning &tu = tuning();
          const int bpc = p.shape.blocks_per_cu > 0 ? p.shape.blocks_per_cu
                                                    : tu.mfma_blocks_per_cu[di];
          const bool pipe = (p.shape.pipe >= 0 ? p.shape.pipe : tu.mfma5_pipeline[di]) != 0;
          int ord = p.shape.order >= 0 ? p.shape.order : tu.mfma5_order[di];
          if (ord < 0 || ord > 2) ord = 0;
          const uint64_t cap = 256ull * (uint64_t)bpc;
          // untuned: each wave walks several tiles where the grid keeps 2048
          // workgroups or more (a tuned shape keeps its own grid)
          const uint64_t tpw = p.shape.blocks_per_cu > 0 ? 1 : mfma_wave_tiles(ntiles);
          blocks = (ntiles + 4 * tpw - 1) / (4 * tpw);
          if (blocks > cap) blocks = cap;
          using Q = typename Q4<R>::T;
          if constexpr (c64) {
            const bool wide = (p.shape.wide >= 0 ? p.shape.wide : tu.mfma5_wide) != 0;
            StageArgsW sw{};
            if (p.kk == 4) {
              hipLaunchKernelGGL(k_mfma4_c64, dim3(blocks), dim3(kBlock), 0, st,
                                 (float2 *)sv, ntiles, im, off, (const Q *)gdev);
            } else if (wide && ntiles % 2 == 0 && make_stage_wide(p, &sw)) {
              const uint64_t wt = ntiles / 2;  // 64-group tiles
              const uint64_t wtpw = p.shape.blocks_per_cu > 0 ? 1 : mfma_wave_tiles(wt);
              const uint64_t wb0 = (wt + 4 * wtpw - 1) / (4 * wtpw);
              const uint64_t wb = wb0 < cap ? wb0 : cap;
              auto k = mfma5_kernel(k_mfma5w_c64<true, 0>, k_mfma5w_c64<true, 1>,
                                    k_mfma5w_c64<true, 2>, k_mfma5w_c64<false, 0>,
                                    k_mfma5w_c64<false, 1>, k_mfma5w_c64<false, 2>, pipe,
                                    scatter_ok(ord, wt));
              hipLaunchKernelGGL(k, dim3(wb), dim3(kBlock), 0, st, (float2 *)sv, wt, im, sw,
                                 (const Q *)gdev);
            } else {
              auto k = mfma5_kernel(k_mfma5_c64<true, 0>, k_mfma5_c64<true, 1>,
                                    k_mfma5_c64<true, 2>, k_mfma5_c64<false, 0>,
                                    k_mfma5_c64<false, 1>, k_mfma5_c64<false, 2>, pipe,
                                    scatter_ok(ord, ntiles));
              hipLaunchKernelGGL(k, dim3(blocks), dim3(kBlock), 0, st, (float2 *)sv, ntiles, im,
                                 sa, (const Q *)gdev);
            }
          } else {
            if (p.kk == 4) {
              hipLaunchKernelGGL(k_mfma4_c128, dim3(blocks), dim3(kBlock), 0, st,
                                 (double2 *)sv, ntiles, im, off, (const Q *)gdev);
            } else {
              auto k = mfma5_kernel(k_mfma5_c128<true, 0>, k_mfma5_c128<true, 1>,
                                    k_mfma5_c128<true, 2>, k_mfma5_c128<false, 0>,
                                    k_mfma5_c128<false, 1>, k_mfma5_c128<false, 2>, pipe,
                                    scatter_ok(ord, ntiles));
              hipLaunchKernelGGL(k, dim3(blocks), dim3(kBlock), 0, st, (double2 *)sv, ntiles,
                                 im, sa, (const Q *)gdev);
            }
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

// ---------------------------------------------------------------------------
// Circuit batches (kernels_cbatch.hip): gates of several circuits in one
// launch on the 5-target MFMA kernels, each entry its own state, tiles,
// workgroup range and matrix.
// ---------------------------------------------------------------------------
struct CBMfma {
  uint64_t psi;          // amplitude offset of the circuit's state
  uint64_t ntiles;       // tiles of the gate
  uint64_t first_block;  // first workgroup of the entry in its launch
  uint64_t nblocks;      // its workgroups (the tile loop's grid)
  uint64_t g;            // offset of its matrix (3m layout, 4-real elements)
  IdxMap im;
  StageArgs sa;
};

template <bool C64, bool PIPE>
__global__ void __launch_bounds__(kBlock)
    k_cbatch_mfma5(void *__restrict__ states, const void *__restrict__ g3m,
                   const CBMfma *__restrict__ tab, int count) {
  int lo = 0, hi = count - 1;  // the entry holding this workgroup (uniform)
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tab[mid].first_block <= blockIdx.x) lo = mid; else hi = mid - 1;
  }
  const CBMfma &m = tab[__builtin_amdgcn_readfirstlane(lo)];
  const uint64_t wg = blockIdx.x - m.first_block;
  if constexpr (C64)
    k_mfma5_c64_body<PIPE, 0, const IdxMap &, const StageArgs &>(
        (float2 *)states + m.psi, m.ntiles, m.im, m.sa, (const float4 *)g3m + m.g, wg, m.nblocks);
  else
    k_mfma5_c128_body<PIPE, 0, const IdxMap &, const StageArgs &>(
        (double2 *)states + m.psi, m.ntiles, m.im, m.sa,
        (const typename Q4<double>::T *)g3m + m.g, wg, m.nblocks);
}

size_t cbatch_mfma_entry_bytes() { return sizeof(CBMfma); }

// Tiles per wave of a launch of the 4-5 target MFMA kernels over
// `total_tiles` tiles (untuned programs, parameter and circuit batches): up
// to the mfma_tiles_per_wave knob (a wave that walks several tiles overlaps
// the next tile's loads with this one's products and stages the block matrix
// once for all of them), but never so many that the launch drops below
// kMinWaveTileBlocks workgroups (8 per CU).
constexpr uint64_t kMinWaveTileBlocks = 2048;
uint64_t mfma_wave_tiles(uint64_t total_tiles) {
  uint64_t t = (uint64_t)tuning().mfma_tiles_per_wave;
  while (t > 1 && total_tiles < 4 * t * kMinWaveTileBlocks) t >>= 1;
  return t;
}

uint64_t cbatch_mfma_tiles(const GatePlan &p, bool c64, uint64_t nelems) {
  if (p.kind != Kind::DENSE || p.kk != 5) return 0;
  const uint64_t tile_groups = c64 ? 32 : 16;
  const uint64_t nfree = nelems >> (__builtin_popcountll(p.fixed_mask) + p.kk);
  StageArgs sa{};
  if (nfree < tile_groups || nfree % tile_groups || !make_stage(p, c64, &sa)) return 0;
  return nfree / tile_groups;
}

bool cbatch_mfma_entry(const GatePlan &p, bool c64, uint64_t nelems, uint64_t psi,
                       uint64_t first_block, uint64_t tiles_per_wave,
                       std::vector<double> *g3m, void *entry, uint64_t *nblocks) {
  if (p.kind != Kind::DENSE || p.kk != 5) return false;
  const uint64_t tile_groups = c64 ? 32 : 16;
  const uint64_t nfree = nelems >> (__builtin_popcountll(p.fixed_mask) + p.kk);
  CBMfma e{};
  if (nfree < tile_groups || nfree % tile_groups || !make_stage(p, c64, &e.sa)) return false;
  e.psi = psi;
  e.ntiles = nfree / tile_groups;
  e.first_block = first_block;
  const uint64_t cap = 256ull * (uint64_t)tuning().mfma_blocks_per_cu[c64 ? 0 : 1];
  const uint64_t per = 4 * std::max<uint64_t>(tiles_per_wave, 1);  // tiles per workgroup
  e.nblocks = std::min<uint64_t>((e.ntiles + per - 1) / per, cap);
  e.im = make_idxmap(p.fixed_mask | target_mask(p), p.fixed_val);
  e.g = g3m->size() / 4;
  const std::vector<double> b = fused_block_3m(p);
  g3m->insert(g3m->end(), b.begin(), b.end());
  std::memcpy(entry, &e, sizeof(e));
  *nblocks = e.nblocks;
  return true;
}

int launch_cbatch_mfma(sq_dtype_t dt, void *states, const void *g3m, const void *tab, int count,
                       uint64_t blocks, hipStream_t st) {
  const bool c64 = dt == SQ_COMPLEX64;
  const bool pipe = tuning().mfma5_pipeline[c64 ? 0 : 1] != 0;
  const CBMfma *t = (const CBMfma *)tab;
  const dim3 g((unsigned)blocks), b(kBlock);
  if (c64 && pipe) hipLaunchKernelGGL((k_cbatch_mfma5<true, true>), g, b, 0, st, states, g3m, t, count);
  else if (c64) hipLaunchKernelGGL((k_cbatch_mfma5<true, false>), g, b, 0, st, states, g3m, t, count);
  else if (pipe) hipLaunchKernelGGL((k_cbatch_mfma5<false, true>), g, b, 0, st, states, g3m, t, count);
  else hipLaunchKernelGGL((k_cbatch_mfma5<false, false>), g, b, 0, st, states, g3m, t, count);
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

// ---------------------------------------------------------------------------
// Parameter batches: one launch per gate over `nsets` states of `nelems`
// amplitudes each (state y at sv + y * nelems), every state with its own
// matrix for the gate.  The plans of one gate share their structure (kind,
// target and control bits, permutation), so the kernel, grid and index maps
// come from the first plan and only the numbers differ: they sit in a device
// table with one kernel-argument entry per state (or one block matrix per
// state for 4-5 target blocks), read by the kb_* kernels at blockIdx.y.
// ---------------------------------------------------------------------------
enum class BatchKernel : int { NONE, PHASE, DIAG, PERM, SPERM, DENSE, SDENSE, MFMA4, MFMA5, DLDS };

struct BatchShape {
  BatchKernel k = BatchKernel::NONE;
  IdxMap im{};
  SmallTile stile{};
  StageArgs sa{};
  OffArgs off{};
  uint64_t work = 0;  // free indices (lane kernels) or tiles per state
};

template <typename R>
static BatchShape batch_shape(const GatePlan &p, uint64_t nelems) {
  BatchShape b;
  const int nfixed = __builtin_popcountll(p.fixed_mask);
  constexpr bool c64 = sizeof(R) == 4;
  switch (p.kind) {
    case Kind::NOOP:
      break;
    case Kind::PHASE:
    case Kind::DIAG:
      b.k = p.kind == Kind::PHASE ? BatchKernel::PHASE : BatchKernel::DIAG;
      b.im = make_idxmap(p.fixed_mask, p.fixed_val);
      b.work = nelems >> nfixed;
      break;
    case Kind::PERM:
    case Kind::DENSE:
      if (p.kk <= 3) {
        const bool perm = p.kind == Kind::PERM;
        if (make_small_tile(p, nelems, c64, &b.stile, &b.im, &b.work)) {
          b.k = perm ? BatchKernel::SPERM : BatchKernel::SDENSE;
        } else {
          b.k = perm ? BatchKernel::PERM : BatchKernel::DENSE;
          b.im = make_idxmap(p.fixed_mask | target_mask(p), p.fixed_val);
          b.work = nelems >> (nfixed + p.kk);
        }
      } else {
        const uint64_t nfree = nelems >> (nfixed + p.kk);
        const uint64_t tg = c64 ? 32 : 16;
        b.im = make_idxmap(p.fixed_mask | target_mask(p), p.fixed_val);
        for (int r = 0; r < (1 << p.kk); ++r) b.off.o[r] = local_offset(p, r);
        if (nfree >= tg && nfree % tg == 0 && (p.kk == 4 || make_stage(p, c64, &b.sa))) {
          b.k = p.kk == 4 ? BatchKernel::MFMA4 : BatchKernel::MFMA5;
          b.work = nfree / tg;
        } else {
          b.k = BatchKernel::DLDS;
          b.work = nfree;
        }
      }
      break;
  }
  return b;
}

template <typename R, typename A>
static void put(std::vector<char> *tab, const A &a) {
  const char *c = reinterpret_cast<const char *>(&a);
  tab->insert(tab->end(), c, c + sizeof(A));
}

template <typename R>
static int batch_table_t(const GatePlan *const *plans, uint32_t nsets, uint64_t nelems,
                         std::vector<char> *tab) {
  const BatchShape b = batch_shape<R>(*plans[0], nelems);
  for (uint32_t y = 0; y < nsets; ++y) {
    const GatePlan &p = *plans[y];
    switch (b.k) {
      case BatchKernel::PHASE: {
        typename CV<R>::T ph;
        ph.x = (R)p.m[0].real();
        ph.y = (R)p.m[0].imag();
        put<R>(tab, ph);
        break;
      }
      case BatchKernel::DIAG:
        put<R>(tab, diag_args<R>(p));
        break;
      case BatchKernel::PERM:
        put<R>(tab, perm_args<R>(p));
        break;
      case BatchKernel::SPERM:
        put<R>(tab, small_perm_args<R>(p, b.stile));
        break;
      case BatchKernel::DENSE:
      case BatchKernel::SDENSE:
        if (p.kk == 1) put<R>(tab, dense_args<R, 1>(p));
        else if (p.kk == 2) put<R>(tab, dense_args<R, 2>(p));
        else put<R>(tab, dense_args<R, 3>(p));
        break;
      default:  // blocks live in the batch's block images; no table
        return SQ_OK;
    }
  }
  return SQ_OK;
}

int batch_gate_table(const GatePlan *const *plans, uint32_t nsets, sq_dtype_t dt,
                     uint64_t nelems, std::vector<char> *tab) {
  return dt == SQ_COMPLEX64 ? batch_table_t<float>(plans, nsets, nelems, tab)
                            : batch_table_t<double>(plans, nsets, nelems, tab);
}

template <typename R>
static int launch_batch_t(const GatePlan &p, uint64_t nelems, uint32_t nsets,
                          typename CV<R>::T *sv, const void *tab, const R *gdev,
                          uint64_t gstride, hipStream_t st) {
  using T = typename CV<R>::T;
  using Q = typename Q4<R>::T;
  const BatchShape b = batch_shape<R>(p, nelems);
  if (b.k == BatchKernel::NONE) return SQ_OK;
  if (nsets == 0 || nsets > 65535) return fail(SQ_ERR_INVALID, "batch: 1 to 65535 states");
  const bool blocks = b.k == BatchKernel::MFMA4 || b.k == BatchKernel::MFMA5 ||
                      b.k == BatchKernel::DLDS;
  if (blocks ? gdev == nullptr : tab == nullptr)
    return fail(SQ_ERR_INVALID, "internal: batch gate without its argument table");
  const uint64_t ss = nelems;
  switch (b.k) {
    case BatchKernel::PHASE:
      hipLaunchKernelGGL((kb_phase<R, 4>), dim3(grid_for(b.work, kBlock * 4).x, nsets),
                         dim3(kBlock), 0, st, sv, ss, b.work, b.im, (const T *)tab);
      break;
    case BatchKernel::DIAG:
      hipLaunchKernelGGL((kb_diag<R, 4>), dim3(grid_for(b.work, kBlock * 4).x, nsets),
                         dim3(kBlock), 0, st, sv, ss, b.work, b.im, (const DiagArgs<R> *)tab);
      break;
    case BatchKernel::PERM:
      hipLaunchKernelGGL((kb_perm<R, 2>), dim3(grid_for(b.work, kBlock * 2).x, nsets),
                         dim3(kBlock), 0, st, sv, ss, b.work, b.im, (const PermArgs<R> *)tab);
      break;
    case BatchKernel::SPERM: {
      const dim3 g(sweep_grid(b.work, 4).x, nsets);
      const auto *a = (const SmallPermArgs<R> *)tab;
      if (p.kk == 1)
        hipLaunchKernelGGL((kb_small_perm<R, 1>), g, dim3(kBlock), 0, st, sv, ss, b.work, b.im, b.stile, a);
      else if (p.kk == 2)
        hipLaunchKernelGGL((kb_small_perm<R, 2>), g, dim3(kBlock), 0, st, sv, ss, b.work, b.im, b.stile, a);
      else
        hipLaunchKernelGGL((kb_small_perm<R, 3>), g, dim3(kBlock), 0, st, sv, ss, b.work, b.im, b.stile, a);
      break;
    }
    case BatchKernel::DENSE: {
      const dim3 g(grid_sweep1(b.work).x, nsets);
      if (p.kk == 1)
        hipLaunchKernelGGL((kb_dense<R, 1, 1>), g, dim3(kBlock), 0, st, sv, ss, b.work, b.im,
                           (const DenseArgs<R, 1> *)tab);
      else if (p.kk == 2)
        hipLaunchKernelGGL((kb_dense<R, 2, 1>), g, dim3(kBlock), 0, st, sv, ss, b.work, b.im,
                           (const DenseArgs<R, 2> *)tab);
      else
        hipLaunchKernelGGL((kb_dense<R, 3, 1>), g, dim3(kBlock), 0, st, sv, ss, b.work, b.im,
                           (const DenseArgs<R, 3> *)tab);
      break;
    }
    case BatchKernel::SDENSE: {
      const dim3 g(sweep_grid(b.work, 4).x, nsets);
      if (p.kk == 1)
        hipLaunchKernelGGL((kb_small_dense<R, 1>), g, dim3(kBlock), 0, st, sv, ss, b.work, b.im,
                           b.stile, (const DenseArgs<R, 1> *)tab);
      else if (p.kk == 2)
        hipLaunchKernelGGL((kb_small_dense<R, 2>), g, dim3(kBlock), 0, st, sv, ss, b.work, b.im,
                           b.stile, (const DenseArgs<R, 2> *)tab);
      else
        hipLaunchKernelGGL((kb_small_dense<R, 3>), g, dim3(kBlock), 0, st, sv, ss, b.work, b.im,
                           b.stile, (const DenseArgs<R, 3> *)tab);
      break;
    }
    case BatchKernel::MFMA4:
    case BatchKernel::MFMA5: {
      const int di = sizeof(R) == 4 ? 0 : 1;
      const Tuning &tu = tuning();
      const uint64_t per = 4 * mfma_wave_tiles(b.work * nsets);  // tiles per workgroup
      uint64_t nb = (b.work + per - 1) / per;
      const uint64_t cap = 256ull * (uint64_t)tu.mfma_blocks_per_cu[di];
      if (nb > cap) nb = cap;
      const dim3 g((unsigned)nb, nsets);
      const Q *G = (const Q *)gdev;
      if constexpr (sizeof(R) == 4) {
        if (b.k == BatchKernel::MFMA4)
          hipLaunchKernelGGL(kb_mfma4_c64, g, dim3(kBlock), 0, st, sv, ss, b.work, b.im, b.off, G,
                             gstride);
        else
          hipLaunchKernelGGL((kb_mfma5_c64<true, 0>), g, dim3(kBlock), 0, st, sv, ss, b.work,
                             b.im, b.sa, G, gstride);
      } else {
        if (b.k == BatchKernel::MFMA4)
          hipLaunchKernelGGL(kb_mfma4_c128, g, dim3(kBlock), 0, st, sv, ss, b.work, b.im, b.off,
                             G, gstride);
        else
          hipLaunchKernelGGL((kb_mfma5_c128<true, 0>), g, dim3(kBlock), 0, st, sv, ss, b.work,
                             b.im, b.sa, G, gstride);
      }
      break;
    }
    case BatchKernel::DLDS: {
      const dim3 g(grid_for(b.work, kBlock).x, nsets);
      if (p.kk == 4)
        hipLaunchKernelGGL((kb_dense_lds<R, 4>), g, dim3(kBlock), 0, st, sv, ss, b.work, b.im,
                           b.off, (const Q *)gdev, gstride);
      else
        hipLaunchKernelGGL((kb_dense_lds<R, 5>), g, dim3(kBlock), 0, st, sv, ss, b.work, b.im,
                           b.off, (const Q *)gdev, gstride);
      break;
    }
    case BatchKernel::NONE:
      break;
  }
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

int launch_batch_gate(const GatePlan &p, sq_dtype_t dt, uint64_t nelems, uint32_t nsets,
                      void *sv, const void *tab, const void *gdev, uint64_t gstride,
                      hipStream_t st) {
  if (dt == SQ_COMPLEX64)
    return launch_batch_t<float>(p, nelems, nsets, (float2 *)sv, tab, (const float *)gdev,
                                 gstride, st);
  return launch_batch_t<double>(p, nelems, nsets, (double2 *)sv, tab, (const double *)gdev,
                                gstride, st);
}

}  // namespace sq
