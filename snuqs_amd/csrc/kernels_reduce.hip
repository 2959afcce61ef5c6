// kernels_reduce.hip -- results taken from a state: probabilities, norms,
// inner products, Pauli-sum expectations, amplitude gathers, scaling.
//
// The reference reads results on the host with numpy after a full
// device->host copy (src/braket/snuqs/simulation.py:196-220); it has no
// expectation, norm or inner-product code.  Every reduction here accumulates
// in FP64 (also for complex64 states) with a fixed grid and a fixed
// combination order, so results are bitwise reproducible run to run.
#include <algorithm>

#include "kernels.hpp"
#include "kernels_reduce.hpp"

namespace sq {

__device__ __forceinline__ double2 ld2(const float2 *p, uint64_t i) {
  const float2 v = p[i];
  return make_double2((double)v.x, (double)v.y);
}
__device__ __forceinline__ double2 ld2(const double2 *p, uint64_t i) {
  return p[i];
}
__device__ __forceinline__ double2 ld2_nt(const float2 *p, uint64_t i) {
  const float2 v = ld_nt(p + i);
  return make_double2((double)v.x, (double)v.y);
}
__device__ __forceinline__ double2 ld2_nt(const double2 *p, uint64_t i) { return ld_nt(p + i); }

// Block-wide sum of two doubles per thread; result valid in thread 0.
__device__ __forceinline__ void block_sum2(double &a, double &b) {
  __shared__ double sa[kBlock], sb[kBlock];
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sa[threadIdx.x] += sa[threadIdx.x + w];
      sb[threadIdx.x] += sb[threadIdx.x + w];
    }
    __syncthreads();
  }
  a = sa[0];
  b = sb[0];
}

// ---- probabilities: p[i] = re^2 + im^2 in float64, no contraction --------
template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_probabilities(const T *__restrict__ sv, double *__restrict__ p, uint64_t n) {
#pragma clang fp contract(off)
  for (uint64_t i0 = (uint64_t)blockIdx.x * (kBlock * 4) + threadIdx.x; i0 < n;
       i0 += (uint64_t)gridDim.x * (kBlock * 4)) {
    uint64_t i = i0;
#pragma unroll
    for (int u = 0; u < 4; ++u, i += kBlock) {
      if (i < n) {
        const double2 v = ld2_nt(sv, i);
        st_nt(p + i, v.x * v.x + v.y * v.y);
      }
    }
  }
}

// p[i] = |sv[i]|^2 / norm2, norm2 read from device memory (the FP64 sum of the
// same |sv[i]|^2 by the fixed-order reduction), division correctly rounded.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_probabilities_div(const T *__restrict__ sv, double *__restrict__ p, uint64_t n,
                        const double *__restrict__ norm2) {
#pragma clang fp contract(off)
  const double d = *norm2;
  for (uint64_t i0 = (uint64_t)blockIdx.x * (kBlock * 4) + threadIdx.x; i0 < n;
       i0 += (uint64_t)gridDim.x * (kBlock * 4)) {
    uint64_t i = i0;
#pragma unroll
    for (int u = 0; u < 4; ++u, i += kBlock) {
      if (i < n) {
        const double2 v = ld2_nt(sv, i);
        st_nt(p + i, (v.x * v.x + v.y * v.y) / d);
      }
    }
  }
}

// ---- norm^2 and <a|b>: per-block partial sums -----------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_partial_norm2(const T *__restrict__ sv, uint64_t n, double *partial) {
  double acc = 0.0, dummy = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const double2 v = ld2(sv, i);
    acc += v.x * v.x + v.y * v.y;
  }
  block_sum2(acc, dummy);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_partial_inner(const T *__restrict__ a, const T *__restrict__ b, uint64_t n,
                    double *partial /* [2][gridDim.x] */) {
  double re = 0.0, im = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const double2 x = ld2(a, i), y = ld2(b, i);
    re += x.x * y.x + x.y * y.y;  // conj(x) * y
    im += x.x * y.y - x.y * y.x;
  }
  block_sum2(re, im);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = re;
    partial[gridDim.x + blockIdx.x] = im;
  }
}

// ---- Pauli strings ---------------------------------------------------------
// <P> for P|i> = i^ny (-1)^popcount(i & z) |i ^ x>.  With x != 0 the sum runs
// over pairs {i, i^x} (i has the top bit h of x clear):
//   contribution = s_i (conj(psi_j) psi_i + (-1)^ny conj(psi_i) psi_j),
// so ny even accumulates s_i Re(conj(psi_j) psi_i) and ny odd s_i Im(...);
// the finaliser applies 2 * i^ny.  x == 0 accumulates s_i |psi_i|^2.
struct PauliTerm {
  uint64_t x, z;
  uint64_t hmask;  // 2^h - 1 for the top bit h of x (0 if x == 0)
  int ny;
  int pad;
};

template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_partial_pauli(const T *__restrict__ sv, uint64_t n,
                    const PauliTerm *__restrict__ terms, double *partial) {
  // blockIdx.z: state z of a batch of consecutive n-amplitude states (0 for
  // one state); partial row (z, term) at (z * gridDim.y + y) * gridDim.x
  sv += (uint64_t)blockIdx.z * n;
  const PauliTerm pt = terms[blockIdx.y];
  double acc = 0.0, dummy = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  if (pt.x == 0) {
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
      const double2 v = ld2(sv, i);
      const double w = v.x * v.x + v.y * v.y;
      acc += (__popcll(i & pt.z) & 1) ? -w : w;
    }
  } else {
    const uint64_t half = n >> 1;
    const bool odd = pt.ny & 1;
    for (uint64_t f = (uint64_t)blockIdx.x * kBlock + threadIdx.x; f < half; f += stride) {
      const uint64_t lo = f & pt.hmask;
      const uint64_t i = ((f ^ lo) << 1) | lo;
      const uint64_t j = i ^ pt.x;
      const double2 a = ld2(sv, i), b = ld2(sv, j);
      // conj(b) * a
      const double re = b.x * a.x + b.y * a.y;
      const double im = b.x * a.y - b.y * a.x;
      const double v = odd ? im : re;
      acc += (__popcll(i & pt.z) & 1) ? -v : v;
    }
  }
  block_sum2(acc, dummy);
  if (threadIdx.x == 0)
    partial[((uint64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = acc;
}

// out_j = sum_t c'_t (-1)^popcount((j^x_t) & z_t) in_{j^x_t},  c'_t = c_t i^ny_t
struct PauliApplyTerm {
  uint64_t x, z;
  double cr, ci;
};

// y_j += c_t (-1)^{popc(i & z_t)} in_i, i = j ^ x_t, over the terms in order
template <typename T>
__device__ __forceinline__ void pauli_acc(const T *__restrict__ in, uint64_t j,
                                          const PauliApplyTerm &pt, const double2 &v,
                                          double &yr, double &yi) {
  double cr = pt.cr, ci = pt.ci;
  if (__popcll((j ^ pt.x) & pt.z) & 1) {
    cr = -cr;
    ci = -ci;
  }
  yr += cr * v.x - ci * v.y;
  yi += cr * v.y + ci * v.x;
}

// The gathers of 8 terms go out before their sums (the same order of
// additions as one term at a time, so the result is unchanged).
template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_apply_pauli_sum(const T *__restrict__ in, T *__restrict__ out, uint64_t n,
                      const PauliApplyTerm *__restrict__ terms, int nterms) {
  constexpr int U = 8;
  for (uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x; j < n;
       j += (uint64_t)gridDim.x * kBlock) {
    double yr = 0.0, yi = 0.0;
    int t = 0;
    for (; t + U <= nterms; t += U) {
      double2 v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = ld2(in, j ^ terms[t + u].x);
#pragma unroll
      for (int u = 0; u < U; ++u) pauli_acc(in, j, terms[t + u], v[u], yr, yi);
    }
    for (; t < nterms; ++t) pauli_acc(in, j, terms[t], ld2(in, j ^ terms[t].x), yr, yi);
    T o;
    o.x = yr;
    o.y = yi;
    out[j] = o;
  }
}

// ---- fused adjoint step of a 1-target gate with one parameter ---------------
// At reverse step g (psi = the state after gate g, lam = the co-state after
// it): psi <- G^dagger psi (skipped when psi was just reloaded from a
// checkpoint), partial += Re <lam| dG |psi>, lam <- G^dagger lam.  One pass
// over both states instead of separate undo, copy, derivative and inner
// product sweeps.  Gate arithmetic in the state's precision, the inner
// product in FP64.
template <typename T, typename R>
__global__ void __launch_bounds__(kBlock)
    k_adjoint_step1(T *__restrict__ psi, T *__restrict__ lam, uint64_t npair, int bit,
                    Mat2<R> gi, Mat2<R> dg, int undo_psi, double *__restrict__ partial) {
  double acc = 0.0, dummy = 0.0;
  const uint64_t lowm = (1ull << bit) - 1;
  for (uint64_t f = (uint64_t)blockIdx.x * kBlock + threadIdx.x; f < npair;
       f += (uint64_t)gridDim.x * kBlock) {
    const uint64_t lo = f & lowm;
    const uint64_t i0 = ((f ^ lo) << 1) | lo, i1 = i0 | (1ull << bit);
    const T a0 = ld_nt(psi + i0), a1 = ld_nt(psi + i1);
    const T l0 = ld_nt(lam + i0), l1 = ld_nt(lam + i1);
    T p0 = a0, p1 = a1;
    if (undo_psi) mat2_apply(gi, a0, a1, p0, p1);
    T d0, d1;
    mat2_apply(dg, p0, p1, d0, d1);
    acc += (double)l0.x * d0.x + (double)l0.y * d0.y + (double)l1.x * d1.x +
           (double)l1.y * d1.y;
    T q0, q1;
    mat2_apply(gi, l0, l1, q0, q1);
    if (undo_psi) {
      st_nt(psi + i0, p0);
      st_nt(psi + i1, p1);
    }
    st_nt(lam + i0, q0);
    st_nt(lam + i1, q1);
  }
  block_sum2(acc, dummy);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// ---- reverse step over a run of diagonal gates ---------------------------------
// Each thread takes U amplitudes i = base + t + kBlock u of 1024-amplitude
// chunks, so bits 0-9 of i are fixed per (thread, u) and bits 10+ per chunk.
// Every amplitude of psi and lambda is multiplied by the run's G^dagger phase
// product, formed in FP64 (one rounding per amplitude for complex64):
//   - gates on bits below 10 ("local"): their product is formed once per
//     thread and u, before the chunk loop;
//   - gates on bits 10 and above ("uniform"): one factor per chunk, formed
//     across a wave's lanes (one gate per lane) and a product over the wave;
//   - gates with one bit on each side ("mixed"): per amplitude.
// Derivative terms Re sum_i conj(lam_i) psi_i dgg(i) use w = conj(lam) psi,
// which the run does not change (psi and lambda get the same unit-modulus
// phase), so for a local gate the sum is sum_u Re dgg(t, u) W_u with W_u the
// thread's sum of w over its chunks (one wave sum per gate at the end); a
// uniform gate takes dgg(chunk) times the wave's sum of w in that chunk (one
// wave sum per chunk for all of them, each lane adding its own gate's term:
// slots are distinct per gate); a mixed gate is summed per chunk.
// APPLY = false (with LAM): the derivative terms only, psi and lambda read and
// left unchanged (the run's phases are then applied by another pass).
__device__ __forceinline__ int diag_loc(const DiagRunGate &g, uint64_t i) {
  return g.b1 < 0 ? (int)((i >> g.b0) & 1)
                  : (int)((((i >> g.b0) & 1) << 1) | ((i >> g.b1) & 1));
}

template <typename T, typename R, bool LAM, bool APPLY>
__global__ void __launch_bounds__(kBlock)
    k_adjoint_diag_run(T *__restrict__ psi, T *__restrict__ lam, uint64_t n,
                       const DiagRunGate *__restrict__ gates, int m, int nslots,
                       double *__restrict__ partial) {
  constexpr int U = 4, W = kBlock / 64;
  constexpr int kLocBits = 10;  // log2(kBlock * U)
  __shared__ DiagRunGate sg[kMaxDiagRun];
  __shared__ double srow[kMaxDiagSlots][W];
  {
    const double *src = reinterpret_cast<const double *>(gates);
    double *dst = reinterpret_cast<double *>(sg);
    const int words = m * (int)(sizeof(DiagRunGate) / sizeof(double));
    for (int e = threadIdx.x; e < words; e += kBlock) dst[e] = src[e];
    for (int e = threadIdx.x; e < kMaxDiagSlots * W; e += kBlock) (&srow[0][0])[e] = 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  auto is_local = [](const DiagRunGate &g) {
    return g.b0 < kLocBits && (g.b1 < 0 || g.b1 < kLocBits);
  };
  // the uniform gates (at most 64: one per lane) and the mixed ones
  __shared__ unsigned char uni[kMaxDiagRun], mix[kMaxDiagRun];
  __shared__ int nuni, nmix, uslots;
  if (threadIdx.x == 0) {
    int a = 0, b = 0, any = 0;
    for (int k = 0; k < m; ++k) {
      if (sg[k].uniform) {
        uni[a++] = (unsigned char)k;
        any |= sg[k].slot >= 0;
      } else if (!is_local(sg[k])) {
        mix[b++] = (unsigned char)k;
      }
    }
    nuni = a;
    nmix = b;
    uslots = any;
  }
  __syncthreads();
  // phase product of the local gates at local index t + kBlock u
  double lr[U], li[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    lr[u] = 1.0;
    li[u] = 0.0;
  }
  for (int k = 0; APPLY && k < m; ++k) {
    const DiagRunGate &gk = sg[k];
    if (!is_local(gk)) continue;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int loc = diag_loc(gk, (uint64_t)threadIdx.x + (uint64_t)kBlock * u);
      const double gr = gk.ginv[loc][0], gi = gk.ginv[loc][1];
      const double tr = lr[u] * gr - li[u] * gi;
      li[u] = lr[u] * gi + li[u] * gr;
      lr[u] = tr;
    }
  }
  double war[U], wai[U];  // the thread's sums of w over its chunks, per u
#pragma unroll
  for (int u = 0; u < U; ++u) war[u] = wai[u] = 0.0;
  const uint64_t chunk = (uint64_t)kBlock * U;
  for (uint64_t base = (uint64_t)blockIdx.x * chunk; base < n; base += (uint64_t)gridDim.x * chunk) {
    T a[U], l[U];
    double wr[U], wi[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t i = base + threadIdx.x + (uint64_t)kBlock * u;
      a[u] = ld_nt(psi + i);
      if (LAM) {
        l[u] = ld_nt(lam + i);
        // w = conj(lam) psi
        wr[u] = (double)l[u].x * a[u].x + (double)l[u].y * a[u].y;
        wi[u] = (double)l[u].x * a[u].y - (double)l[u].y * a[u].x;
        war[u] += wr[u];
        wai[u] += wi[u];
      } else {
        wr[u] = wi[u] = 0.0;
      }
    }
    double pr[U], pi[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      pr[u] = lr[u];
      pi[u] = li[u];
    }
    // chunk factor: lane j takes uniform gate j, then a product over the
    // wave (lane 0's value for every lane: one rounding order)
    double cr = 1.0, ci = 0.0;
    int myk = -1, myloc = 0;
    if (lane < nuni) {
      myk = uni[lane];
      myloc = diag_loc(sg[myk], base);
      cr = sg[myk].ginv[myloc][0];
      ci = sg[myk].ginv[myloc][1];
    }
    if (APPLY) {
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const double orr = __shfl_xor(cr, off), oi = __shfl_xor(ci, off);
        const double tr = cr * orr - ci * oi;
        ci = cr * oi + ci * orr;
        cr = tr;
      }
      cr = __shfl(cr, 0);
      ci = __shfl(ci, 0);
    }
    if (LAM && uslots) {  // uniform slot gates: dgg(chunk) times the wave's sum of w
      double tr2 = 0.0, ti2 = 0.0;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        tr2 += wr[u];
        ti2 += wi[u];
      }
      const double swr = wave_sum_d(tr2), swi = wave_sum_d(ti2);
      if (myk >= 0 && sg[myk].slot >= 0)  // slots are distinct per gate
        srow[sg[myk].slot][wave] += swr * sg[myk].dgg[myloc][0] - swi * sg[myk].dgg[myloc][1];
    }
    for (int j = 0; j < nmix; ++j) {
      const DiagRunGate &gk = sg[mix[j]];
      double acc = 0.0;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int loc = diag_loc(gk, base + threadIdx.x + (uint64_t)kBlock * u);
        if (APPLY) {
          const double gr = gk.ginv[loc][0], gi = gk.ginv[loc][1];
          const double tr = pr[u] * gr - pi[u] * gi;
          pi[u] = pr[u] * gi + pi[u] * gr;
          pr[u] = tr;
        }
        if (LAM && gk.slot >= 0) acc += wr[u] * gk.dgg[loc][0] - wi[u] * gk.dgg[loc][1];
      }
      if (LAM && gk.slot >= 0) {  // uniform over the block
        acc = wave_sum_d(acc);
        if (lane == 0) srow[gk.slot][wave] += acc;
      }
    }
#pragma unroll
    for (int u = 0; APPLY && u < U; ++u) {
      const uint64_t i = base + threadIdx.x + (uint64_t)kBlock * u;
      const double t = pr[u] * cr - pi[u] * ci;
      pi[u] = pr[u] * ci + pi[u] * cr;
      pr[u] = t;
      T o;
      o.x = (R)((double)a[u].x * pr[u] - (double)a[u].y * pi[u]);
      o.y = (R)((double)a[u].x * pi[u] + (double)a[u].y * pr[u]);
      st_nt(psi + i, o);
      if (LAM) {
        T q;
        q.x = (R)((double)l[u].x * pr[u] - (double)l[u].y * pi[u]);
        q.y = (R)((double)l[u].x * pi[u] + (double)l[u].y * pr[u]);
        st_nt(lam + i, q);
      }
    }
  }
  if (!LAM) return;
  // local gates' derivative terms from the per-u sums of w
  for (int k = 0; k < m; ++k) {
    const DiagRunGate &gk = sg[k];
    if (gk.slot < 0 || !is_local(gk)) continue;
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int loc = diag_loc(gk, (uint64_t)threadIdx.x + (uint64_t)kBlock * u);
      acc += war[u] * gk.dgg[loc][0] - wai[u] * gk.dgg[loc][1];
    }
    acc = wave_sum_d(acc);
    if (lane == 0) srow[gk.slot][wave] += acc;
  }
  __syncthreads();
  for (int s = threadIdx.x; s < nslots; s += kBlock) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < W; ++w) v += srow[s][w];
    partial[(uint64_t)s * gridDim.x + blockIdx.x] = v;
  }
}

// out[2 slot_param[v]] += alpha * sum_b partial[v][b] for v = 0..nslots-1, in
// order by one block: slots of gates that share a parameter add to the same
// output, so they must not be finalized concurrently.
__global__ void __launch_bounds__(kBlock)
    k_finalize_slots(const double *__restrict__ partial, int nblocks, int nslots,
                     const int *__restrict__ slot_param, double alpha, double *out) {
  if (gridDim.x > 1) {  // slots of distinct parameters: one block each
    const int v = blockIdx.x;
    double acc = 0.0, dummy = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kBlock) acc += partial[(uint64_t)v * nblocks + b];
    block_sum2(acc, dummy);
    if (threadIdx.x == 0) out[2 * slot_param[v]] += alpha * acc;
    return;
  }
  for (int v = 0; v < nslots; ++v) {
    double acc = 0.0, dummy = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kBlock) acc += partial[(uint64_t)v * nblocks + b];
    block_sum2(acc, dummy);
    if (threadIdx.x == 0) out[2 * slot_param[v]] += alpha * acc;
    __syncthreads();  // block_sum2's LDS is reused by the next slot
  }
}

int launch_finalize_slots(const double *partial, int nblocks, int nslots,
                          const int *slot_param_dev, bool distinct, double alpha, double *out,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_finalize_slots, dim3(distinct ? nslots : 1), dim3(kBlock), 0, st,
                     partial, nblocks, nslots, slot_param_dev, alpha, out);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

// ---- reverse step over a group of one-target gates on distinct bits ------------
// A lane holds the 2^K amplitudes of psi and lambda of one group (the K target
// bits vary): for each gate j, the pairs along local bit j give
// Re conj(lam) (A_j psi) (A_j = dG_j G_j^dagger), then every G_j^dagger is
// applied to both.  Gate j is local bit K-1-j (gate 0 the most significant).
template <typename R, int K>
struct GroupArgs {
  Mat2<R> gi[K], dg[K];
  uint64_t off[1 << K];  // address offset of local index r
};

template <typename T, typename R, int K, bool LAM>
__global__ void __launch_bounds__(kBlock)
    k_adjoint_step_group(T *__restrict__ psi, T *__restrict__ lam, uint64_t ngroups, IdxMap im,
                         GroupArgs<R, K> a, double *__restrict__ partial) {
  constexpr int D = 1 << K;
  double acc[K];
#pragma unroll
  for (int j = 0; j < K; ++j) acc[j] = 0.0;
  for (uint64_t f = (uint64_t)blockIdx.x * kBlock + threadIdx.x; f < ngroups;
       f += (uint64_t)gridDim.x * kBlock) {
    const uint64_t base = deposit(f, im);
    T x[D], l[D];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      x[r] = ld_nt(psi + base + a.off[r]);
      if (LAM) l[r] = ld_nt(lam + base + a.off[r]);
    }
    if (!LAM) {  // forward: psi <- (every gi) psi
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const int b = 1 << (K - 1 - j);
#pragma unroll
        for (int r = 0; r < D; ++r) {
          if (r & b) continue;
          T y0, y1;
          mat2_apply(a.gi[j], x[r], x[r | b], y0, y1);
          x[r] = y0;
          x[r | b] = y1;
        }
      }
#pragma unroll
      for (int r = 0; r < D; ++r) st_nt(psi + base + a.off[r], x[r]);
      continue;
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const int b = 1 << (K - 1 - j);
#pragma unroll
      for (int r = 0; r < D; ++r) {
        if (r & b) continue;
        T d0, d1;
        mat2_apply(a.dg[j], x[r], x[r | b], d0, d1);
        acc[j] += (double)l[r].x * d0.x + (double)l[r].y * d0.y + (double)l[r | b].x * d1.x +
                  (double)l[r | b].y * d1.y;
      }
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const int b = 1 << (K - 1 - j);
#pragma unroll
      for (int r = 0; r < D; ++r) {
        if (r & b) continue;
        T y0, y1;
        mat2_apply(a.gi[j], x[r], x[r | b], y0, y1);
        x[r] = y0;
        x[r | b] = y1;
        mat2_apply(a.gi[j], l[r], l[r | b], y0, y1);
        l[r] = y0;
        l[r | b] = y1;
      }
    }
#pragma unroll
    for (int r = 0; r < D; ++r) {
      st_nt(psi + base + a.off[r], x[r]);
      st_nt(lam + base + a.off[r], l[r]);
    }
  }
  if (!LAM) return;
#pragma unroll
  for (int j = 0; j < K; j += 2) {
    double u = acc[j], v = j + 1 < K ? acc[j + 1] : 0.0;
    block_sum2(u, v);
    if (threadIdx.x == 0) {
      partial[(uint64_t)j * gridDim.x + blockIdx.x] = u;
      if (j + 1 < K) partial[(uint64_t)(j + 1) * gridDim.x + blockIdx.x] = v;
    }
    __syncthreads();  // block_sum2's LDS is reused
  }
}

// The same step for one-target gates on distinct bits below 6, i.e. inside
// the 64 consecutive amplitudes a wave holds: lane t owns amplitude i
// (consecutive lanes, consecutive amplitudes: every access coalesced), a gate
// on bit b pairs lanes t and t ^ 2^b, and the partner's value comes from a
// cross-lane shuffle.  Lane-per-group addressing would put a gate's pair in
// one lane but make every load instruction stride over the lines, which
// non-temporal accesses then re-fetch.  n is a multiple of 64, so the
// "i < n" tests are uniform per wave and every lane of a shuffle is active.
template <typename R, int K>
struct LaneArgs {
  Mat2<R> gi[K], dg[K];
  int bit[K];
};

template <typename T, typename R>
__device__ __forceinline__ T lane_xor(const T &v, int mask) {
  T o;
  o.x = __shfl_xor(v.x, mask, 64);
  o.y = __shfl_xor(v.y, mask, 64);
  return o;
}

// y = m[s][s] v + m[s][1-s] p, the row of m that belongs to this lane's bit s
template <typename T, typename R>
__device__ __forceinline__ T lane_row(const Mat2<R> &m, int s, const T &v, const T &p) {
  const R ar = s ? m.r[3] : m.r[0], ai = s ? m.i[3] : m.i[0];
  const R br = s ? m.r[2] : m.r[1], bi = s ? m.i[2] : m.i[1];
  T y;
  y.x = ar * v.x - ai * v.y + br * p.x - bi * p.y;
  y.y = ar * v.y + ai * v.x + br * p.y + bi * p.x;
  return y;
}

template <typename T, typename R, int K, bool LAM>
__global__ void __launch_bounds__(kBlock)
    k_adjoint_step_lanes(T *__restrict__ psi, T *__restrict__ lam, uint64_t n,
                         LaneArgs<R, K> a, double *__restrict__ partial) {
  constexpr int U = 4;  // amplitudes per lane in flight
  const int lane = threadIdx.x & 63;
  double acc[K];
#pragma unroll
  for (int j = 0; j < K; ++j) acc[j] = 0.0;
  for (uint64_t i0 = (uint64_t)blockIdx.x * (kBlock * U) + threadIdx.x; i0 < n;
       i0 += (uint64_t)gridDim.x * (kBlock * U)) {
    T x[U], l[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t i = i0 + (uint64_t)u * kBlock;
      if (i < n) {
        x[u] = ld_nt(psi + i);
        if (LAM) l[u] = ld_nt(lam + i);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t i = i0 + (uint64_t)u * kBlock;
      if (i >= n) continue;
      if (LAM) {
        // every derivative term at the group's starting point (the gates
        // act on distinct qubits and commute), then undo the gates
#pragma unroll
        for (int j = 0; j < K; ++j) {
          const int s = (lane >> a.bit[j]) & 1;
          const T d = lane_row<T, R>(a.dg[j], s, x[u], lane_xor<T, R>(x[u], 1 << a.bit[j]));
          acc[j] += (double)l[u].x * d.x + (double)l[u].y * d.y;
        }
      }
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const int s = (lane >> a.bit[j]) & 1, mask = 1 << a.bit[j];
        x[u] = lane_row<T, R>(a.gi[j], s, x[u], lane_xor<T, R>(x[u], mask));
        if (LAM) l[u] = lane_row<T, R>(a.gi[j], s, l[u], lane_xor<T, R>(l[u], mask));
      }
      st_nt(psi + i, x[u]);
      if (LAM) st_nt(lam + i, l[u]);
    }
  }
  if (!LAM) return;
#pragma unroll
  for (int j = 0; j < K; j += 2) {
    double u = acc[j], v = j + 1 < K ? acc[j + 1] : 0.0;
    block_sum2(u, v);
    if (threadIdx.x == 0) {
      partial[(uint64_t)j * gridDim.x + blockIdx.x] = u;
      if (j + 1 < K) partial[(uint64_t)(j + 1) * gridDim.x + blockIdx.x] = v;
    }
    __syncthreads();  // block_sum2's LDS is reused
  }
}

// ---- marginal distribution over k targets ------------------------------------
// out[r] = sum of |sv[i]|^2 over the i whose target bits spell r (target 0
// the most significant bit of r).  Block b sums outcome r = b / nb over the
// free indices f = b % nb (mod nb); partials are combined by k_finalize in a
// fixed order (deterministic).  Free index -> address: the IdxMap when at
// most 8 target bits are inserted, else a bit-by-bit deposit over the free
// positions.
struct MarginalArgs {
  uint64_t tstride[kMaxMarginal];  // address stride of target j
  uint8_t fpos[64];                // free bit positions, ascending
  IdxMap im;
  int k, nfb, use_im, pad;
};

template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_partial_marginal(const T *__restrict__ sv, uint64_t nfree, MarginalArgs a, int nb,
                       double *__restrict__ partial) {
  const uint64_t r = blockIdx.x / nb, bx = blockIdx.x % nb;
  uint64_t off = 0;
  for (int j = 0; j < a.k; ++j)
    if ((r >> (a.k - 1 - j)) & 1) off += a.tstride[j];
  double acc = 0.0, dummy = 0.0;
  for (uint64_t f = bx * kBlock + threadIdx.x; f < nfree; f += (uint64_t)nb * kBlock) {
    uint64_t i;
    if (a.use_im) {
      i = deposit(f, a.im);
    } else {
      i = 0;
      for (int b = 0; b < a.nfb; ++b)
        if ((f >> b) & 1) i |= 1ull << a.fpos[b];
    }
    const double2 v = ld2(sv, i + off);
    acc += v.x * v.x + v.y * v.y;
  }
  block_sum2(acc, dummy);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// ---- finaliser: out[v] = beta*out[v] + alpha[v] * sum_b partial[v][b] -------
__global__ void __launch_bounds__(kBlock)
    k_finalize(const double *__restrict__ partial, int nblocks,
               const double *__restrict__ alpha, double beta, double *out) {
  const int v = blockIdx.x;
  double acc = 0.0, dummy = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += kBlock) acc += partial[(uint64_t)v * nblocks + b];
  block_sum2(acc, dummy);
  if (threadIdx.x == 0) {
    const double a = alpha ? alpha[v] : 1.0;
    out[v] = (beta == 0.0 ? 0.0 : beta * out[v]) + a * acc;
  }
}

// ---- scale / gather --------------------------------------------------------
// The product is formed in FP64 for both dtypes, so a complex64 amplitude is
// rounded once: renormalising by an FP64 norm then leaves the FP64 sum of
// |a|^2 within rounding noise of 1 instead of the 2^-24 a float factor adds.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_scale(T *__restrict__ sv, uint64_t n, double sr, double si) {
  for (uint64_t i0 = (uint64_t)blockIdx.x * (kBlock * 4) + threadIdx.x; i0 < n;
       i0 += (uint64_t)gridDim.x * (kBlock * 4)) {
    uint64_t i = i0;
#pragma unroll
    for (int u = 0; u < 4; ++u, i += kBlock) {
      if (i < n) {
        const T v = ld_nt(sv + i);
        const double vx = v.x, vy = v.y;
        T o;
        o.x = sr * vx - si * vy;
        o.y = sr * vy + si * vx;
        st_nt(sv + i, o);
      }
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_gather(const T *__restrict__ sv, const uint64_t *__restrict__ idx,
             uint64_t count, double2 *__restrict__ out) {
  for (uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x; j < count;
       j += (uint64_t)gridDim.x * kBlock)
    out[j] = ld2(sv, idx[j]);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static dim3 grid1(uint64_t work, int per_block) { return sweep_grid(work, per_block); }

int launch_probabilities(const void *sv, sq_dtype_t dt, uint64_t n, double *p,
                         hipStream_t st) {
  if (dt == SQ_COMPLEX64)
    hipLaunchKernelGGL(k_probabilities<float2>, grid1(n, kBlock * 4), dim3(kBlock), 0, st,
                       (const float2 *)sv, p, n);
  else
    hipLaunchKernelGGL(k_probabilities<double2>, grid1(n, kBlock * 4), dim3(kBlock), 0, st,
                       (const double2 *)sv, p, n);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

int launch_probabilities_div(const void *sv, sq_dtype_t dt, uint64_t n, double *p,
                             const double *norm2_dev, hipStream_t st) {
  if (dt == SQ_COMPLEX64)
    hipLaunchKernelGGL(k_probabilities_div<float2>, grid1(n, kBlock * 4), dim3(kBlock), 0,
                       st, (const float2 *)sv, p, n, norm2_dev);
  else
    hipLaunchKernelGGL(k_probabilities_div<double2>, grid1(n, kBlock * 4), dim3(kBlock), 0,
                       st, (const double2 *)sv, p, n, norm2_dev);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

int launch_marginal(const void *sv, sq_dtype_t dt, uint64_t n, int nbits, const int *tbit,
                    int k, double *partial, int nb, double *out_dev, hipStream_t st) {
  MarginalArgs a{};
  uint64_t tmask = 0;
  for (int j = 0; j < k; ++j) {
    a.tstride[j] = 1ull << tbit[j];
    tmask |= 1ull << tbit[j];
  }
  a.k = k;
  for (int b = 0; b < nbits; ++b)
    if (!((tmask >> b) & 1)) a.fpos[a.nfb++] = (uint8_t)b;
  a.use_im = k <= 8;
  if (a.use_im) a.im = make_idxmap(tmask, 0);
  const uint64_t nfree = n >> k;
  const uint64_t blocks = (1ull << k) * (uint64_t)nb;
  if (dt == SQ_COMPLEX64)
    hipLaunchKernelGGL(k_partial_marginal<float2>, dim3((unsigned)blocks), dim3(kBlock), 0, st,
                       (const float2 *)sv, nfree, a, nb, partial);
  else
    hipLaunchKernelGGL(k_partial_marginal<double2>, dim3((unsigned)blocks), dim3(kBlock), 0, st,
                       (const double2 *)sv, nfree, a, nb, partial);
  hipLaunchKernelGGL(k_finalize, dim3((unsigned)(1ull << k)), dim3(kBlock), 0, st, partial, nb,
                     nullptr, 0.0, out_dev);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

static int red_blocks(uint64_t work);

int launch_adjoint_step1(void *psi, void *lam, sq_dtype_t dt, uint64_t n, int bit,
                         const double *ginv, const double *dgm, int undo_psi, double *partial,
                         double *out_dev, const double *alpha_dev, hipStream_t st) {
  const uint64_t npair = n >> 1;
  const int nb = red_blocks(npair);
  if (dt == SQ_COMPLEX64) {
    Mat2<float> gi, dg;
    for (int e = 0; e < 4; ++e) {
      gi.r[e] = (float)ginv[2 * e];
      gi.i[e] = (float)ginv[2 * e + 1];
      dg.r[e] = (float)dgm[2 * e];
      dg.i[e] = (float)dgm[2 * e + 1];
    }
    hipLaunchKernelGGL((k_adjoint_step1<float2, float>), dim3(nb), dim3(kBlock), 0, st,
                       (float2 *)psi, (float2 *)lam, npair, bit, gi, dg, undo_psi, partial);
  } else {
    Mat2<double> gi, dg;
    for (int e = 0; e < 4; ++e) {
      gi.r[e] = ginv[2 * e];
      gi.i[e] = ginv[2 * e + 1];
      dg.r[e] = dgm[2 * e];
      dg.i[e] = dgm[2 * e + 1];
    }
    hipLaunchKernelGGL((k_adjoint_step1<double2, double>), dim3(nb), dim3(kBlock), 0, st,
                       (double2 *)psi, (double2 *)lam, npair, bit, gi, dg, undo_psi, partial);
  }
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, st, partial, nb, alpha_dev, 1.0,
                     out_dev);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

int diag_run_blocks(uint64_t n) {
  const uint64_t nb = n / ((uint64_t)kBlock * 4);
  return (int)(nb > 1536 ? 1536 : (nb < 1 ? 1 : nb));
}

int launch_adjoint_diag_run(void *psi, void *lam, sq_dtype_t dt, uint64_t n,
                            const DiagRunGate *gates_dev, int m, int nslots,
                            const int *slot_param_dev, bool distinct, double *partial,
                            double *out_dev, hipStream_t st, bool apply) {
  if (m < 1 || m > kMaxDiagRun || nslots < 0 || nslots > kMaxDiagSlots)
    return fail(SQ_ERR_INVALID, "internal: diagonal run out of range");
  const uint64_t chunk = (uint64_t)kBlock * 4;
  if (n % chunk) return fail(SQ_ERR_INVALID, "internal: diagonal run needs n % 1024 == 0");
  // at most 1536 workgroups (2 rounds of 3 per CU): each then covers several
  // chunks, over which its local phase products and sums of w are shared
  const uint64_t nb = (uint64_t)diag_run_blocks(n);
  if (!lam && nslots) return fail(SQ_ERR_INVALID, "internal: derivative slots need lambda");
  if (!lam && !apply) return fail(SQ_ERR_INVALID, "internal: a forward run applies its phases");
  if (dt == SQ_COMPLEX64) {
    auto k = !lam   ? k_adjoint_diag_run<float2, float, false, true>
             : apply ? k_adjoint_diag_run<float2, float, true, true>
                     : k_adjoint_diag_run<float2, float, true, false>;
    hipLaunchKernelGGL(k, dim3((unsigned)nb), dim3(kBlock), 0, st, (float2 *)psi, (float2 *)lam,
                       n, gates_dev, m, nslots, partial);
  } else {
    auto k = !lam   ? k_adjoint_diag_run<double2, double, false, true>
             : apply ? k_adjoint_diag_run<double2, double, true, true>
                     : k_adjoint_diag_run<double2, double, true, false>;
    hipLaunchKernelGGL(k, dim3((unsigned)nb), dim3(kBlock), 0, st, (double2 *)psi, (double2 *)lam,
                       n, gates_dev, m, nslots, partial);
  }
  if (nslots && out_dev)
    hipLaunchKernelGGL(k_finalize_slots, dim3(distinct ? nslots : 1), dim3(kBlock), 0, st,
                       partial, (int)nb, nslots, slot_param_dev, 2.0, out_dev);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

template <typename T, typename R, int K>
static void launch_lanes_t(void *psi, void *lam, uint64_t n, const int *bits, const double *ginv,
                           const double *dgg, double *partial, int nb, hipStream_t st) {
  LaneArgs<R, K> a{};
  for (int j = 0; j < K; ++j) {
    a.bit[j] = bits[j];
    for (int e = 0; e < 4; ++e) {
      a.gi[j].r[e] = (R)ginv[8 * j + 2 * e];
      a.gi[j].i[e] = (R)ginv[8 * j + 2 * e + 1];
      a.dg[j].r[e] = (R)dgg[8 * j + 2 * e];
      a.dg[j].i[e] = (R)dgg[8 * j + 2 * e + 1];
    }
  }
  if constexpr (K <= 4) {
    if (lam) {
      hipLaunchKernelGGL((k_adjoint_step_lanes<T, R, K, true>), dim3(nb), dim3(kBlock), 0, st,
                         (T *)psi, (T *)lam, n, a, partial);
      return;
    }
  }
  hipLaunchKernelGGL((k_adjoint_step_lanes<T, R, K, false>), dim3(nb), dim3(kBlock), 0, st,
                     (T *)psi, (T *)nullptr, n, a, partial);
}

template <typename T, typename R, int K>
static void launch_group_t(void *psi, void *lam, uint64_t n, const int *bits, const double *ginv,
                           const double *dgg, double *partial, int nb, hipStream_t st) {
  GroupArgs<R, K> a{};
  uint64_t ins = 0;
  for (int j = 0; j < K; ++j) {
    ins |= 1ull << bits[j];
    for (int e = 0; e < 4; ++e) {
      a.gi[j].r[e] = (R)ginv[8 * j + 2 * e];
      a.gi[j].i[e] = (R)ginv[8 * j + 2 * e + 1];
      a.dg[j].r[e] = (R)dgg[8 * j + 2 * e];
      a.dg[j].i[e] = (R)dgg[8 * j + 2 * e + 1];
    }
  }
  for (int r = 0; r < (1 << K); ++r) {
    uint64_t o = 0;
    for (int j = 0; j < K; ++j)
      if ((r >> (K - 1 - j)) & 1) o += 1ull << bits[j];
    a.off[r] = o;
  }
  if constexpr (K <= 4) {  // reverse groups hold psi and lambda: at most 4 gates
    if (lam) {
      hipLaunchKernelGGL((k_adjoint_step_group<T, R, K, true>), dim3(nb), dim3(kBlock), 0, st,
                         (T *)psi, (T *)lam, n >> K, make_idxmap(ins, 0), a, partial);
      return;
    }
  }
  hipLaunchKernelGGL((k_adjoint_step_group<T, R, K, false>), dim3(nb), dim3(kBlock), 0, st,
                     (T *)psi, (T *)nullptr, n >> K, make_idxmap(ins, 0), a, partial);
}

int step_group_blocks(uint64_t n, int k, const int *bits) {
  int hi = -1;
  for (int j = 0; j < k; ++j) hi = std::max(hi, bits[j]);
  return red_blocks(hi < 6 ? n / 4 : n >> k);
}

// Deferred finalisation: block p adds, in row order, the sums of its rows
// (row_off[i], row_nb[i] for i in [param_ptr[p], param_ptr[p + 1])), then
// out[2p] += alpha * total.
__global__ void __launch_bounds__(kBlock)
    k_finalize_rows(const double *__restrict__ partial, const uint64_t *__restrict__ row_off,
                    const int *__restrict__ row_nb, const int *__restrict__ param_ptr,
                    double alpha, double *out) {
  const int p = blockIdx.x;
  double total = 0.0;
  for (int i = param_ptr[p]; i < param_ptr[p + 1]; ++i) {
    const double *row = partial + row_off[i];
    double acc = 0.0, dummy = 0.0;
    for (int b = threadIdx.x; b < row_nb[i]; b += kBlock) acc += row[b];
    block_sum2(acc, dummy);
    total += acc;  // every thread holds the block's sum
    __syncthreads();  // block_sum2's LDS is reused by the next row
  }
  if (threadIdx.x == 0 && param_ptr[p + 1] > param_ptr[p]) out[2 * p] += alpha * total;
}

int launch_finalize_rows(const double *partial, const uint64_t *row_off_dev,
                         const int *row_nb_dev, const int *param_ptr_dev, int nparams,
                         double alpha, double *out, hipStream_t st) {
  if (nparams <= 0) return SQ_OK;
  hipLaunchKernelGGL(k_finalize_rows, dim3(nparams), dim3(kBlock), 0, st, partial, row_off_dev,
                     row_nb_dev, param_ptr_dev, alpha, out);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

int launch_adjoint_step_group(void *psi, void *lam, sq_dtype_t dt, uint64_t n, int k,
                              const int *bits, const double *ginv, const double *dgg,
                              const int *slot_param_dev, bool distinct, double *partial,
                              double *out_dev, hipStream_t st) {
  if (k < 2 || k > (lam ? 4 : 5))
    return fail(SQ_ERR_INVALID, "internal: gate group size out of range");
  // bits all below 6: the lane kernel (cross-lane pairs); bits all above 3:
  // the register kernel (a lane holds each group of 2^k amplitudes)
  int lo = 64, hi = -1;
  for (int j = 0; j < k; ++j) {
    lo = std::min(lo, bits[j]);
    hi = std::max(hi, bits[j]);
  }
  const bool lanes = hi < 6;
  if (lanes && (n < 64 || n % 64 != 0))
    return fail(SQ_ERR_INVALID, "internal: lane group on a state below 64 amplitudes");
  if (!lanes && lo < 4) return fail(SQ_ERR_INVALID, "internal: gate group mixes low and high bits");
  // a forward group (lam == nullptr) only applies the ginv matrices to psi:
  // callers pass the gates themselves there; more blocks, no reduction
  const uint64_t work = lanes ? n / 4 : n >> k;  // lanes: 4 amplitudes per lane
  const uint64_t fb = work / kBlock + 1;
  const int nb = lam ? red_blocks(work) : (int)(fb < (1u << 16) ? fb : (1u << 16));
  const bool c64 = dt == SQ_COMPLEX64;
#define SQ_GROUP(K)                                                                         \
  case K:                                                                                   \
    if (lanes && c64)                                                                       \
      launch_lanes_t<float2, float, K>(psi, lam, n, bits, ginv, dgg, partial, nb, st);      \
    else if (lanes)                                                                         \
      launch_lanes_t<double2, double, K>(psi, lam, n, bits, ginv, dgg, partial, nb, st);    \
    else if (c64)                                                                           \
      launch_group_t<float2, float, K>(psi, lam, n, bits, ginv, dgg, partial, nb, st);      \
    else                                                                                    \
      launch_group_t<double2, double, K>(psi, lam, n, bits, ginv, dgg, partial, nb, st);    \
    break;
  switch (k) {
    SQ_GROUP(2)
    SQ_GROUP(3)
    SQ_GROUP(4)
    SQ_GROUP(5)
  }
#undef SQ_GROUP
  if (lam && out_dev)
    hipLaunchKernelGGL(k_finalize_slots, dim3(distinct ? k : 1), dim3(kBlock), 0, st, partial,
                       nb, k, slot_param_dev, 2.0, out_dev);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

static int red_blocks(uint64_t work) {
  uint64_t b = (work + kBlock * 8 - 1) / (kBlock * 8);
  if (b > kRedBlocks) b = kRedBlocks;
  if (b < 1) b = 1;
  return (int)b;
}

int launch_norm2(const void *sv, sq_dtype_t dt, uint64_t n, double *partial,
                 double *out_dev, hipStream_t st) {
  const int nb = red_blocks(n);
  if (dt == SQ_COMPLEX64)
    hipLaunchKernelGGL(k_partial_norm2<float2>, dim3(nb), dim3(kBlock), 0, st,
                       (const float2 *)sv, n, partial);
  else
    hipLaunchKernelGGL(k_partial_norm2<double2>, dim3(nb), dim3(kBlock), 0, st,
                       (const double2 *)sv, n, partial);
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, st, partial, nb,
                     (const double *)nullptr, 0.0, out_dev);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

int launch_inner(const void *a, const void *b, sq_dtype_t dt, uint64_t n,
                 double *partial, double *out_dev, const double *alpha_dev,
                 double beta, hipStream_t st) {
  const int nb = red_blocks(n);
  if (dt == SQ_COMPLEX64)
    hipLaunchKernelGGL(k_partial_inner<float2>, dim3(nb), dim3(kBlock), 0, st,
                       (const float2 *)a, (const float2 *)b, n, partial);
  else
    hipLaunchKernelGGL(k_partial_inner<double2>, dim3(nb), dim3(kBlock), 0, st,
                       (const double2 *)a, (const double2 *)b, n, partial);
  hipLaunchKernelGGL(k_finalize, dim3(2), dim3(kBlock), 0, st, partial, nb,
                     alpha_dev, beta, out_dev);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

// terms_dev: PauliTerm[nterms] on the device; alpha_dev: per-term factors
// (2*i^ny folded in); out_dev[t] = term value.
int launch_pauli_terms(const void *sv, sq_dtype_t dt, uint64_t n, int nterms,
                       const void *terms_dev, const double *alpha_dev,
                       double *partial, double *out_dev, hipStream_t st) {
  const int nb = red_blocks(n / 2);
  if (dt == SQ_COMPLEX64)
    hipLaunchKernelGGL(k_partial_pauli<float2>, dim3(nb, nterms), dim3(kBlock), 0, st,
                       (const float2 *)sv, n, (const PauliTerm *)terms_dev, partial);
  else
    hipLaunchKernelGGL(k_partial_pauli<double2>, dim3(nb, nterms), dim3(kBlock), 0, st,
                       (const double2 *)sv, n, (const PauliTerm *)terms_dev, partial);
  hipLaunchKernelGGL(k_finalize, dim3(nterms), dim3(kBlock), 0, st, partial, nb,
                     alpha_dev, 0.0, out_dev);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

// Batch: term values of nsets consecutive states; partial holds
// nsets x nterms x red_blocks(n / 2) doubles; out_dev[z * out_stride + t].
__global__ void __launch_bounds__(kBlock)
    k_finalize_rows(const double *__restrict__ partial, int nblocks, int nterms,
                    const double *__restrict__ alpha, double *out, uint64_t out_stride) {
  const int v = blockIdx.x, z = v / nterms, t = v % nterms;
  double acc = 0.0, dummy = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += kBlock) acc += partial[(uint64_t)v * nblocks + b];
  block_sum2(acc, dummy);
  if (threadIdx.x == 0) out[(uint64_t)z * out_stride + t] = alpha[t] * acc;
}

size_t pauli_batch_partial_doubles(uint64_t n, uint32_t nsets, int nterms) {
  return (size_t)red_blocks(n / 2) * nsets * (size_t)nterms;
}

int launch_pauli_terms_batch(const void *sv, sq_dtype_t dt, uint64_t n, uint32_t nsets,
                             int nterms, const void *terms_dev, const double *alpha_dev,
                             double *partial, double *out_dev, uint64_t out_stride,
                             hipStream_t st) {
  if (nsets == 0 || nsets > 65535 || nterms < 1 || nterms > 65535)
    return fail(SQ_ERR_INVALID, "internal: batch expectation shape");
  const int nb = red_blocks(n / 2);
  const dim3 grid(nb, nterms, nsets);
  if (dt == SQ_COMPLEX64)
    hipLaunchKernelGGL(k_partial_pauli<float2>, grid, dim3(kBlock), 0, st,
                       (const float2 *)sv, n, (const PauliTerm *)terms_dev, partial);
  else
    hipLaunchKernelGGL(k_partial_pauli<double2>, grid, dim3(kBlock), 0, st,
                       (const double2 *)sv, n, (const PauliTerm *)terms_dev, partial);
  hipLaunchKernelGGL(k_finalize_rows, dim3(nsets * nterms), dim3(kBlock), 0, st, partial, nb,
                     nterms, alpha_dev, out_dev, out_stride);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

size_t pauli_term_bytes() { return sizeof(PauliTerm); }
size_t pauli_apply_term_bytes() { return sizeof(PauliApplyTerm); }

void fill_pauli_terms(int nterms, const uint64_t *x, const uint64_t *z,
                      const uint32_t *ny, const double *coeffs, void *terms_host,
                      double *alpha_host) {
  PauliTerm *t = (PauliTerm *)terms_host;
  for (int i = 0; i < nterms; ++i) {
    t[i].x = x[i];
    t[i].z = z[i];
    t[i].ny = (int)ny[i];
    t[i].pad = 0;
    t[i].hmask = x[i] ? ((1ull << (63 - __builtin_clzll(x[i]))) - 1) : 0;
    double f;
    if (x[i] == 0) {
      f = 1.0;
    } else if ((ny[i] & 1) == 0) {
      f = ((ny[i] / 2) & 1) ? -2.0 : 2.0;  // 2 * i^ny, ny even
    } else {
      f = (((ny[i] - 1) / 2) & 1) ? 2.0 : -2.0;  // 2i * i^ny, ny odd
    }
    alpha_host[i] = f * (coeffs ? coeffs[i] : 1.0);
  }
}

void fill_pauli_apply_terms(int nterms, const uint64_t *x, const uint64_t *z,
                            const uint32_t *ny, const double *coeffs,
                            void *terms_host) {
  PauliApplyTerm *t = (PauliApplyTerm *)terms_host;
  static const double ipr[4] = {1, 0, -1, 0}, ipi[4] = {0, 1, 0, -1};
  for (int i = 0; i < nterms; ++i) {
    t[i].x = x[i];
    t[i].z = z[i];
    t[i].cr = coeffs[i] * ipr[ny[i] & 3];
    t[i].ci = coeffs[i] * ipi[ny[i] & 3];
  }
}

int launch_apply_pauli_sum(const void *in, void *out, sq_dtype_t dt, uint64_t n,
                           int nterms, const void *terms_dev, hipStream_t st) {
  if (dt == SQ_COMPLEX64)
    hipLaunchKernelGGL(k_apply_pauli_sum<float2>, grid1(n, kBlock), dim3(kBlock), 0, st,
                       (const float2 *)in, (float2 *)out, n,
                       (const PauliApplyTerm *)terms_dev, nterms);
  else
    hipLaunchKernelGGL(k_apply_pauli_sum<double2>, grid1(n, kBlock), dim3(kBlock), 0, st,
                       (const double2 *)in, (double2 *)out, n,
                       (const PauliApplyTerm *)terms_dev, nterms);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

int launch_scale(void *sv, sq_dtype_t dt, uint64_t n, double re, double im,
                 hipStream_t st) {
  if (dt == SQ_COMPLEX64)
    hipLaunchKernelGGL((k_scale<float2>), grid1(n, kBlock * 4), dim3(kBlock), 0, st,
                       (float2 *)sv, n, re, im);
  else
    hipLaunchKernelGGL((k_scale<double2>), grid1(n, kBlock * 4), dim3(kBlock), 0,
                       st, (double2 *)sv, n, re, im);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

int launch_gather(const void *sv, sq_dtype_t dt, const uint64_t *idx_dev,
                  uint64_t count, double *out_dev, hipStream_t st) {
  if (dt == SQ_COMPLEX64)
    hipLaunchKernelGGL(k_gather<float2>, grid1(count, kBlock), dim3(kBlock), 0, st,
                       (const float2 *)sv, idx_dev, count, (double2 *)out_dev);
  else
    hipLaunchKernelGGL(k_gather<double2>, grid1(count, kBlock), dim3(kBlock), 0, st,
                       (const double2 *)sv, idx_dev, count, (double2 *)out_dev);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

}  // namespace sq
