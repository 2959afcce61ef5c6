// kernels_sample.hip -- sampling from |psi|^2 (simulation.py:217-220).
//
// numpy's RandomState.choice(N, p=p, size=k) returns
// searchsorted(cdf, u, side='right') with cdf = cumsum(p) / cumsum(p)[-1],
// where cumsum is the sequential float64 recurrence c_i = fl(c_{i-1} + p_i)
// and u = random_sample(k).  The sampler takes u from the caller.
//
// Fast mode: cdf by a deterministic FP64 parallel scan (block scans + a scan
// of block totals); it differs from the sequential recurrence by rounding.
//
// Exact mode reproduces the recurrence bit for bit.  While c stays inside one
// binade [2^e, 2^(e+1)) every value is an integer multiple of u = 2^(e-52), so
// fl(c + p) = (C + d) u with d = p/u rounded to nearest; the increments d do
// not depend on c except at exact ties (p/u = k + 1/2, broken towards the even
// result, i.e. by the parity of C), which are rare and resolved in order.  An
// epoch is therefore an int64 prefix sum: the chain runs binade by binade, the
// element that leaves the binade is added with one ordinary double add, and
// the fast FP64 scan bounds where each binade ends.
#include <algorithm>
#include <cmath>
#include <vector>

#include "kernels.hpp"
#include "kernels_reduce.hpp"

namespace sq {

constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;  // 4096 elements per block

// ---- generic in-place inclusive scan (double or int64) ---------------------
template <typename T>
__device__ void block_scan_tile(T *data, uint64_t n, uint64_t tile, T carry_in,
                                bool use_carry, T *tile_total) {
  __shared__ T s[kScanTile];
  __shared__ T tsum[kBlock];
  const uint64_t base = tile * kScanTile;
  for (int e = threadIdx.x; e < kScanTile; e += kBlock) {
    const uint64_t i = base + e;
    s[e] = i < n ? data[i] : T(0);
  }
  __syncthreads();
  // sequential scan of this thread's run
  T run = 0;
  const int o = threadIdx.x * kScanItems;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    run = (j == 0) ? s[o] : run + s[o + j];
    s[o + j] = run;
  }
  tsum[threadIdx.x] = run;
  __syncthreads();
  // Hillis-Steele inclusive scan of the thread totals
  for (int w = 1; w < kBlock; w <<= 1) {
    T v = (int)threadIdx.x >= w ? tsum[threadIdx.x - w] : T(0);
    __syncthreads();
    if ((int)threadIdx.x >= w) tsum[threadIdx.x] = tsum[threadIdx.x] + v;
    __syncthreads();
  }
  const T excl = threadIdx.x ? tsum[threadIdx.x - 1] : T(0);
  if (tile_total && threadIdx.x == kBlock - 1) *tile_total = tsum[kBlock - 1];
  if (use_carry || threadIdx.x) {
    const T add = use_carry ? (threadIdx.x ? carry_in + excl : carry_in) : excl;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) s[o + j] = s[o + j] + add;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kScanTile; e += kBlock) {
    const uint64_t i = base + e;
    if (i < n) data[i] = s[e];
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_scan_totals(const T *data, uint64_t n, T *totals) {
  // tile totals only (sum of the tile in scan order)
  __shared__ T s[kScanTile];
  __shared__ T tsum[kBlock];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  for (int e = threadIdx.x; e < kScanTile; e += kBlock) {
    const uint64_t i = base + e;
    s[e] = i < n ? data[i] : T(0);
  }
  __syncthreads();
  T run = 0;
  const int o = threadIdx.x * kScanItems;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) run = (j == 0) ? s[o] : run + s[o + j];
  tsum[threadIdx.x] = run;
  __syncthreads();
  for (int w = 1; w < kBlock; w <<= 1) {
    T v = (int)threadIdx.x >= w ? tsum[threadIdx.x - w] : T(0);
    __syncthreads();
    if ((int)threadIdx.x >= w) tsum[threadIdx.x] = tsum[threadIdx.x] + v;
    __syncthreads();
  }
  if (threadIdx.x == kBlock - 1) totals[blockIdx.x] = tsum[kBlock - 1];
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_scan_apply(T *data, uint64_t n, const T *scanned_totals) {
  const uint64_t tile = blockIdx.x;
  const T carry = tile ? scanned_totals[tile - 1] : T(0);
  block_scan_tile<T>(data, n, tile, carry, tile != 0, nullptr);
}

static uint64_t ntiles(uint64_t n) { return (n + kScanTile - 1) / kScanTile; }

template <typename T>
static size_t scan_scratch(uint64_t n) {
  size_t b = 0;
  while (n > 1) {
    n = ntiles(n);
    b += n * sizeof(T);
    if (n <= 1) break;
  }
  return b + 256;
}

// In-place inclusive scan; scratch holds the per-level tile totals.
template <typename T>
static int scan_inplace(T *data, uint64_t n, T *scratch, hipStream_t st) {
  if (n == 0) return SQ_OK;
  const uint64_t nt = ntiles(n);
  if (nt == 1) {
    hipLaunchKernelGGL(k_scan_apply<T>, dim3(1), dim3(kBlock), 0, st, data, n,
                       (const T *)nullptr);
    SQ_HIP(hipGetLastError());
    return SQ_OK;
  }
  T *totals = scratch;
  hipLaunchKernelGGL(k_scan_totals<T>, dim3((unsigned)nt), dim3(kBlock), 0, st,
                     (const T *)data, n, totals);
  SQ_HIP(hipGetLastError());
  SQ_TRY(scan_inplace<T>(totals, nt, scratch + nt, st));
  hipLaunchKernelGGL(k_scan_apply<T>, dim3((unsigned)nt), dim3(kBlock), 0, st, data,
                     n, (const T *)totals);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

// ---- exact-mode epoch kernels ----------------------------------------------
// Increment of p in units of u (ulp of the current binade): d = nearest
// integer, tie flag when p/u is exactly k + 1/2 (d holds k), overflow flag
// when the element alone reaches the top of the binade.
__global__ void __launch_bounds__(kBlock)
    k_increments(const double *__restrict__ p, uint64_t a, uint64_t len,
                 int shift, int64_t *__restrict__ d, int32_t *__restrict__ tie) {
  for (uint64_t m = (uint64_t)blockIdx.x * kBlock + threadIdx.x; m < len;
       m += (uint64_t)gridDim.x * kBlock) {
    const double q = ldexp(p[a + m], shift);  // exact power-of-two scaling
                                              // (inf on overflow, flagged below)
    int64_t di;
    int32_t t = 0;
    if (q >= 9007199254740992.0) {  // 2^53: leaves the binade on its own
      di = (int64_t)1 << 53;
    } else {
      const double k = floor(q);
      const double fr = q - k;  // exact
      di = (int64_t)k;
      if (fr > 0.5) di += 1;
      else if (fr == 0.5) t = 1;
    }
    d[m] = di;
    tie[m] = t;
  }
}

// Resolve ties in order (one thread).  tie holds the inclusive scan of the
// tie flags, d the inclusive scan of the increments; tiepos the positions.
// adj[r] = number of ties among 0..r that rounded up.
__global__ void k_compact_ties(const int32_t *__restrict__ tie_scan, uint64_t len,
                               uint64_t *tiepos) {
  for (uint64_t m = (uint64_t)blockIdx.x * kBlock + threadIdx.x; m < len;
       m += (uint64_t)gridDim.x * kBlock) {
    const int32_t prev = m ? tie_scan[m - 1] : 0;
    if (tie_scan[m] != prev) tiepos[prev] = m;
  }
}

__global__ void k_resolve_ties(const int64_t *__restrict__ dscan,
                               const uint64_t *__restrict__ tiepos, int ntie,
                               int64_t c0, int32_t *adj) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int32_t up = 0;
  for (int r = 0; r < ntie; ++r) {
    const uint64_t m = tiepos[r];
    // C after element m if the tie rounds down: c0 + dscan[m] + up
    const int64_t down = c0 + dscan[m] + up;
    if (down & 1) ++up;  // halfway between down (odd) and down+1 (even)
    adj[r] = up;
  }
}

// C_m = c0 + dscan[m] + (ties at positions <= m rounded up); first m with
// C_m >= 2^53 is the element that leaves the binade.
__global__ void __launch_bounds__(kBlock)
    k_epoch_values(const int64_t *__restrict__ dscan, uint64_t len, int64_t c0,
                   const uint64_t *__restrict__ tiepos, const int32_t *__restrict__ adj,
                   int ntie, int64_t *__restrict__ cval, unsigned long long *cross) {
  for (uint64_t m = (uint64_t)blockIdx.x * kBlock + threadIdx.x; m < len;
       m += (uint64_t)gridDim.x * kBlock) {
    int lo = 0, hi = ntie;  // number of ties with position <= m
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (tiepos[mid] <= m) lo = mid + 1; else hi = mid;
    }
    const int64_t cm = c0 + dscan[m] + (lo ? adj[lo - 1] : 0);
    cval[m] = cm;
    if (cm >= ((int64_t)1 << 53)) atomicMin(cross, (unsigned long long)m);
  }
}

// cdf[a + m] = C_m * u for m < stop; the crossing element (if any) is added
// with one double add by thread 0 of block 0 afterwards.
__global__ void __launch_bounds__(kBlock)
    k_epoch_write(const int64_t *__restrict__ cval, uint64_t a, uint64_t stop,
                  double u, double *__restrict__ cdf) {
  for (uint64_t m = (uint64_t)blockIdx.x * kBlock + threadIdx.x; m < stop;
       m += (uint64_t)gridDim.x * kBlock)
    cdf[a + m] = (double)cval[m] * u;  // exact: C_m < 2^53
}

__global__ void k_epoch_cross(const double *__restrict__ p, double *cdf, uint64_t pos,
                              double *cval_out) {
  if (threadIdx.x || blockIdx.x) return;
  const double c = (pos == 0) ? p[0] : cdf[pos - 1] + p[pos];
  cdf[pos] = c;
  *cval_out = c;
}

// ---- search ------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
    k_search(const double *__restrict__ cdf, uint64_t n, const double *__restrict__ u,
             uint64_t shots, int64_t *__restrict__ out) {
  const double total = cdf[n - 1];
  for (uint64_t s = (uint64_t)blockIdx.x * kBlock + threadIdx.x; s < shots;
       s += (uint64_t)gridDim.x * kBlock) {
    const double v = u[s];
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
      const uint64_t mid = lo + ((hi - lo) >> 1);
      if (cdf[mid] / total <= v) lo = mid + 1; else hi = mid;
    }
    out[s] = (int64_t)lo;
  }
}

int launch_search(const double *cdf, uint64_t n, const double *u_dev, uint64_t shots,
                  int64_t *out_dev, hipStream_t st) {
  if (shots == 0) return SQ_OK;
  hipLaunchKernelGGL(k_search, sweep_grid(shots, kBlock), dim3(kBlock), 0, st, cdf, n, u_dev,
                     shots, out_dev);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

// ---- host driver -----------------------------------------------------------
size_t sample_scratch_bytes(uint64_t n, bool exact) {
  size_t b = scan_scratch<double>(n) + 4096;
  if (exact) {
    // increments + tie flags + scanned values for one epoch segment (<= n),
    // tie positions and adjustments, scan scratch for int64 and int32
    b += n * (sizeof(int64_t) * 2 + sizeof(int32_t)) + scan_scratch<int64_t>(n) +
         scan_scratch<int32_t>(n) + 4096;
  }
  return b;
}

static inline char *align_up(char *p) {
  return (char *)(((uintptr_t)p + 255) & ~(uintptr_t)255);
}

int sample_cdf(const double *p, uint64_t n, double *cdf, bool exact, void *scratch,
               size_t scratch_bytes, hipStream_t st) {
  if (scratch_bytes < sample_scratch_bytes(n, exact))
    return fail(SQ_ERR_INVALID, "internal: sampler scratch too small");
  char *w = align_up((char *)scratch);
  double *dscr = (double *)w;
  w = align_up(w + scan_scratch<double>(n));
  SQ_HIP(hipMemcpyAsync(cdf, p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  SQ_TRY(scan_inplace<double>(cdf, n, dscr, st));  // fast cdf (c')
  if (!exact) return SQ_OK;

  // Exact mode.  The fast cdf c' bounds the position where each binade ends:
  // |c_m - c'_m| <= (m + 1) * 2^-52 * c'_m covers the rounding of both sums.
  int64_t *dinc = (int64_t *)w;
  w = align_up(w + n * sizeof(int64_t));
  int64_t *cval = (int64_t *)w;
  w = align_up(w + n * sizeof(int64_t));
  int32_t *tie = (int32_t *)w;
  w = align_up(w + n * sizeof(int32_t));
  int64_t *iscr = (int64_t *)w;
  w = align_up(w + scan_scratch<int64_t>(n));
  int32_t *tscr = (int32_t *)w;
  w = align_up(w + scan_scratch<int32_t>(n));

  // Host copies of c' are read by binary search through small D2H reads.
  auto read_d = [&](const double *src, double *dst) -> int {
    SQ_HIP(hipMemcpyAsync(dst, src, sizeof(double), hipMemcpyDeviceToHost, st));
    SQ_HIP(hipStreamSynchronize(st));
    return SQ_OK;
  };
  // first index m >= lo with c'_m > thr (n if none); c' is non-decreasing.
  auto first_above = [&](uint64_t lo, double thr, uint64_t *res) -> int {
    uint64_t a = lo, b = n;
    double v;
    while (a < b) {
      const uint64_t mid = a + (b - a) / 2;
      SQ_TRY(read_d(cdf + mid, &v));
      if (v > thr) b = mid; else a = mid + 1;
    }
    *res = a;
    return SQ_OK;
  };

  // Leading zeros: c_m = 0 exactly until the first positive p (c' is 0 there
  // too), then c = p exactly.
  uint64_t pos;
  SQ_TRY(first_above(0, 0.0, &pos));
  if (pos >= n) return SQ_OK;  // all zero: cdf is all zero as in numpy
  double c;
  {
    // cdf[pos] = 0 + p[pos] = p[pos]
    SQ_HIP(hipMemcpyAsync(cdf + pos, p + pos, sizeof(double), hipMemcpyDeviceToDevice, st));
    SQ_TRY(read_d(p + pos, &c));
  }
  double *dev_c = dscr;  // one double of scratch for the crossing value
  unsigned long long *dev_cross = (unsigned long long *)(dscr + 1);

  while (pos + 1 < n) {
    if (!(c > 0.0) || !std::isfinite(c))
      return fail(SQ_ERR_INVALID, "probabilities must be finite and non-negative");
    int e = std::ilogb(c);
    if (e < -1022) e = -1022;
    const double u = std::ldexp(1.0, e - 52);
    const double top = std::ldexp(1.0, e + 1);
    const int64_t c0 = (int64_t)(c / u);  // exact
    // segment [a, hi): a = pos + 1; hi bounds the crossing from above
    const uint64_t a = pos + 1;
    uint64_t hi;
    {
      const double slack = (double)(n + 1) * std::ldexp(1.0, -52) * 2.0;
      SQ_TRY(first_above(a, top * (1.0 + slack), &hi));
      hi = std::min<uint64_t>(n, hi + 1);
      if (hi <= a) hi = std::min<uint64_t>(n, a + 1);
    }
    const uint64_t len = hi - a;
    const dim3 g = sweep_grid(len, kBlock);
    hipLaunchKernelGGL(k_increments, g, dim3(kBlock), 0, st, p, a, len,
                       52 - e, dinc, tie);
    SQ_HIP(hipGetLastError());
    SQ_TRY(scan_inplace<int64_t>(dinc, len, iscr, st));
    SQ_TRY(scan_inplace<int32_t>(tie, len, tscr, st));
    int32_t ntie = 0;
    SQ_HIP(hipMemcpyAsync(&ntie, tie + len - 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    SQ_HIP(hipStreamSynchronize(st));
    // tie positions and adjustments live in the (now free) scan scratch
    uint64_t *tpos = (uint64_t *)iscr;
    int32_t *adj = (int32_t *)tscr;
    if ((size_t)ntie * sizeof(uint64_t) > scan_scratch<int64_t>(n) ||
        (size_t)ntie * sizeof(int32_t) > scan_scratch<int32_t>(n))
      return fail(SQ_ERR_UNSUPPORTED, "exact sampler: too many rounding ties");
    if (ntie > 0) {
      hipLaunchKernelGGL(k_compact_ties, g, dim3(kBlock), 0, st,
                         (const int32_t *)tie, len, tpos);
      hipLaunchKernelGGL(k_resolve_ties, dim3(1), dim3(64), 0, st,
                         (const int64_t *)dinc, (const uint64_t *)tpos, (int)ntie, c0, adj);
      SQ_HIP(hipGetLastError());
    }
    const unsigned long long none = ~0ull;
    SQ_HIP(hipMemcpyAsync(dev_cross, &none, sizeof(none), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_epoch_values, g, dim3(kBlock), 0, st, (const int64_t *)dinc,
                       len, c0, (const uint64_t *)tpos, (const int32_t *)adj, (int)ntie,
                       cval, dev_cross);
    SQ_HIP(hipGetLastError());
    unsigned long long cross = 0;
    SQ_HIP(hipMemcpyAsync(&cross, dev_cross, sizeof(cross), hipMemcpyDeviceToHost, st));
    SQ_HIP(hipStreamSynchronize(st));
    const uint64_t stop = (cross == none) ? len : (uint64_t)cross;
    if (stop > 0) {
      hipLaunchKernelGGL(k_epoch_write, sweep_grid(stop, kBlock), dim3(kBlock), 0, st,
                         (const int64_t *)cval, a, stop, u, cdf);
      SQ_HIP(hipGetLastError());
    }
    if (cross == none) {
      if (hi >= n) break;  // reached the end inside this binade
      return fail(SQ_ERR_INVALID,
                  "exact sampler: binade bound violated (probabilities not finite?)");
    }
    pos = a + stop;
    hipLaunchKernelGGL(k_epoch_cross, dim3(1), dim3(64), 0, st, p, cdf, pos, dev_c);
    SQ_HIP(hipGetLastError());
    SQ_TRY(read_d(dev_c, &c));
  }
  return SQ_OK;
}

}  // namespace sq
