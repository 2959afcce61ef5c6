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
//   k_mfma_* : DENSE plans on 4-5 targets (fused blocks) on the matrix cores
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
  uint64_t i = (uint64_t)blockIdx.x * (kBlock * 4) + threadIdx.x;
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

int launch_init(void *sv, sq_dtype_t dt, uint64_t n, bool basis,
                hipStream_t stream) {
  uint64_t blocks = (n + kBlock * 4 - 1) / (kBlock * 4);
  if (dt == SQ_COMPLEX64)
    hipLaunchKernelGGL(k_init<float2>, dim3(blocks), dim3(kBlock), 0, stream,
                       (float2 *)sv, n, (int)basis);
  else
    hipLaunchKernelGGL(k_init<double2>, dim3(blocks), dim3(kBlock), 0, stream,
                       (double2 *)sv, n, (int)basis);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

// Plain copy, 16 bytes per lane per access, 4 accesses in flight per lane.
__global__ void __launch_bounds__(kBlock)
    k_copy(uint4 *__restrict__ dst, const uint4 *__restrict__ src, uint64_t n16) {
  uint64_t i = (uint64_t)blockIdx.x * (kBlock * 4) + threadIdx.x;
  uint4 v[4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (i + (uint64_t)u * kBlock < n16) v[u] = src[i + (uint64_t)u * kBlock];
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (i + (uint64_t)u * kBlock < n16) dst[i + (uint64_t)u * kBlock] = v[u];
}

int launch_copy(void *dst, const void *src, size_t bytes, hipStream_t stream) {
  const uint64_t n16 = bytes / 16;
  if (n16 == 0) return SQ_OK;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)((n16 + kBlock * 4 - 1) / (kBlock * 4))),
                     dim3(kBlock), 0, stream, (uint4 *)dst, (const uint4 *)src, n16);
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
  const uint64_t f0 = (uint64_t)blockIdx.x * (kBlock * U) + threadIdx.x;
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
  const uint64_t f0 = (uint64_t)blockIdx.x * (kBlock * U) + threadIdx.x;
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
  const uint64_t f0 = (uint64_t)blockIdx.x * (kBlock * U) + threadIdx.x;
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
  const uint64_t f0 = (uint64_t)blockIdx.x * (kBlock * U) + threadIdx.x;
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

// ---------------------------------------------------------------------------
// DENSE, 4-5 targets, VALU: matrix staged in LDS (broadcast reads).  Used for
// states too small to fill MFMA tiles.
// ---------------------------------------------------------------------------
struct OffArgs {
  uint64_t o[1 << kMaxK];
};

template <typename R, int K>
__global__ void __launch_bounds__(kBlock)
    k_dense_lds(typename CV<R>::T *__restrict__ sv, uint64_t nfree, IdxMap im,
                OffArgs off, const R *__restrict__ G) {
  using T = typename CV<R>::T;
  constexpr int D = 1 << K;
  constexpr int RD = 2 * D;  // real block dimension
  __shared__ R g[RD * RD];
  for (int e = threadIdx.x; e < RD * RD; e += kBlock) g[e] = G[e];
  __syncthreads();
  const uint64_t f = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (f >= nfree) return;
  const uint64_t base = deposit(f, im);
  T x[D];
#pragma unroll
  for (int c = 0; c < D; ++c) x[c] = sv[base + off.o[c]];
  // g is column-major real block: element (i, kcol) at kcol*RD + i;
  // Re m[r][c] = g(2r, 2c), Im m[r][c] = g(2r+1, 2c).
#pragma unroll
  for (int r = 0; r < D; ++r) {
    R yr = 0, yi = 0;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      const R mr = g[(2 * c) * RD + 2 * r], mi = g[(2 * c) * RD + 2 * r + 1];
      yr += mr * x[c].x - mi * x[c].y;
      yi += mr * x[c].y + mi * x[c].x;
    }
    T o;
    o.x = yr;
    o.y = yi;
    sv[base + off.o[r]] = o;
  }
}

// ---------------------------------------------------------------------------
// DENSE, 4-5 targets on the matrix cores.
//
// Y = G X for a tile of groups, in real block form: the 2^K x 2^K complex G is
// the D x D real matrix Gr (D = 2^(K+1)), a group's 2^K amplitudes the real
// D-vector (re0, im0, re1, im1, ...).  complex64 uses
// v_mfma_f32_32x32x2f32 with 32 groups per wave tile, complex128 uses
// v_mfma_f64_16x16x4f64 with 16 groups per wave tile.  The B operand of one
// MFMA step is loaded straight from HBM: lane l takes one float (double) of one
// amplitude, so a step's 64 lanes read the re and im words of 32 (16)
// consecutive groups.  The A operand (Gr) is staged in LDS once per workgroup.
// ---------------------------------------------------------------------------
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <int K>
__global__ void __launch_bounds__(kBlock)
    k_mfma_c64(float2 *__restrict__ sv, uint64_t ntiles, IdxMap im,
               OffArgs off, const float *__restrict__ G) {
  constexpr int D = 2 << K;   // 32 or 64
  constexpr int NT = D / 32;  // output row tiles
  constexpr int NS = D / 2;   // MFMA k-steps
  __shared__ float As[D * D];
  for (int e = threadIdx.x; e < D * D; e += kBlock) As[e] = G[e];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 31, half = lane >> 5;
  float a[NT][NS];
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int s = 0; s < NS; ++s) a[t][s] = As[(2 * s + half) * D + 32 * t + col];

  for (uint64_t tile = (uint64_t)blockIdx.x * 4 + wave; tile < ntiles;
       tile += (uint64_t)gridDim.x * 4) {
    const uint64_t base = deposit(tile * 32 + col, im);
    const float *src = reinterpret_cast<const float *>(sv + base) + half;
    float b[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) b[s] = src[2 * off.o[s]];
    floatx16 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[t][v] = 0.f;
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
      for (int t = 0; t < NT; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[t][s], b[s], acc[t], 0,
                                                       0, 0);
    // C layout: lane l, register v -> column l%32, row 8(v/4) + 4(l/32) + v%4.
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = 16 * t + 4 * q;  // + 2*half, +1
        const uint64_t o0 = half ? off.o[r + 2] : off.o[r];
        const uint64_t o1 = half ? off.o[r + 3] : off.o[r + 1];
        sv[base + o0] = make_float2(acc[t][4 * q + 0], acc[t][4 * q + 1]);
        sv[base + o1] = make_float2(acc[t][4 * q + 2], acc[t][4 * q + 3]);
      }
  }
}

template <int K>
__global__ void __launch_bounds__(kBlock)
    k_mfma_c128(double2 *__restrict__ sv, uint64_t ntiles, IdxMap im,
                OffArgs off, const double *__restrict__ G) {
  constexpr int D = 2 << K;   // 32 or 64
  constexpr int NT = D / 16;  // output row tiles
  constexpr int NS = D / 4;   // MFMA k-steps
  __shared__ double As[D * D];
  for (int e = threadIdx.x; e < D * D; e += kBlock) As[e] = G[e];
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
    // row 4v + l/16.  Row 16t + 4v + q is component q%2 of output amplitude
    // 8t + 2v + q/2, so each register is one double of one amplitude; lanes
    // with q = 0,1 (2,3) write the re/im halves of one amplitude per group.
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
// host side
// ---------------------------------------------------------------------------
std::vector<double> fused_real_block(const GatePlan &p) {
  const int d = 1 << p.kk, D = 2 * d;
  std::vector<double> g(D * D);
  for (int r = 0; r < d; ++r)
    for (int o = 0; o < d; ++o) {
      const cd z = p.m[r * d + o];
      // element (i, kcol) at kcol * D + i
      g[(2 * o + 0) * D + 2 * r + 0] = z.real();
      g[(2 * o + 1) * D + 2 * r + 0] = -z.imag();
      g[(2 * o + 0) * D + 2 * r + 1] = z.imag();
      g[(2 * o + 1) * D + 2 * r + 1] = z.real();
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
  return dim3((unsigned)((work + per_block - 1) / per_block));
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
          const uint64_t cap = 256ull * 16;  // persistent: 16 workgroups per CU
          if (blocks > cap) blocks = cap;
          if constexpr (c64) {
            if (p.kk == 4)
              hipLaunchKernelGGL((k_mfma_c64<4>), dim3(blocks), dim3(kBlock), 0, st,
                                 (float2 *)sv, ntiles, im, off, (const float *)gdev);
            else
              hipLaunchKernelGGL((k_mfma_c64<5>), dim3(blocks), dim3(kBlock), 0, st,
                                 (float2 *)sv, ntiles, im, off, (const float *)gdev);
          } else {
            if (p.kk == 4)
              hipLaunchKernelGGL((k_mfma_c128<4>), dim3(blocks), dim3(kBlock), 0, st,
                                 (double2 *)sv, ntiles, im, off, (const double *)gdev);
            else
              hipLaunchKernelGGL((k_mfma_c128<5>), dim3(blocks), dim3(kBlock), 0, st,
                                 (double2 *)sv, ntiles, im, off, (const double *)gdev);
          }
        } else {
          if (p.kk == 4)
            hipLaunchKernelGGL((k_dense_lds<R, 4>), grid_for(nfree, kBlock), dim3(kBlock),
                               0, st, sv, nfree, im, off, gdev);
          else
            hipLaunchKernelGGL((k_dense_lds<R, 5>), grid_for(nfree, kBlock), dim3(kBlock),
                               0, st, sv, nfree, im, off, gdev);
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
