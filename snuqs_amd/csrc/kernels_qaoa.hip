// kernels_qaoa.hip -- batched QAOA MaxCut: energies and adjoint gradients of
// many instances with uneven qubit counts per launch.
//
// No reference counterpart (the reference has no batching, expectation or
// gradient code).  Per instance: |+>^n, then p layers of
//   cost   U_C(g) = prod_edges ZZ(2g) = exp(-i g S),  S_ii = s(i) = |E| - 2 cut(i)
//   mixer  U_B(b) = prod_q RX(2b)     = exp(-i b sum_q X_q)
// energy E = <psi| C |psi> with C_ii = cut(i) (the number of cut edges), and
//   dE/dg_l = 2 Im <lambda| S |psi>,  dE/db_l = 2 Im <lambda| sum_q X_q |psi>
// evaluated on the adjoint sweep where lambda = C psi is carried backwards.
//
// Work unit: a chunk of 2^L amplitudes (L = min(12, n)); one workgroup per
// (instance, chunk).  Qubits on bits below L ("low") are applied inside the
// chunk in LDS together with the cost layer; the higher bits go in groups of
// up to kQGM = 7 bits per "mid" pass over the state, whose workgroups hold 32
// consecutive amplitudes (256 B of complex64 per access) for each of the 2^7
// values of the group's bits.
// Grouping is exact for the gradients too: every X_q commutes with the
// whole mixer layer, so <lambda| X_q |psi> has the same value anywhere inside
// it and a group's partial can be taken before the group is undone.  Each
// launch runs the same pass for every instance that has it, so one sequence
// of launches serves the whole batch (and is captured in one hipGraph).
// Reductions write one FP64 partial per (pass, chunk) and are summed in a
// fixed order by the finaliser: results are bitwise reproducible.
#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "kernels.hpp"
#include "kernels_qaoa.hpp"
#include "chunk_layout.hpp"

namespace sq {

constexpr int kQMaxEdges = 128;
constexpr int kQGM = 7;            // high bits per mid pass
constexpr int kQC = 5;             // contiguous low bits beside them

struct QInstDev {
  uint64_t psi, lam;  // amplitude offsets into the batch state buffer
  uint64_t slot;      // first partial slot
  int n, nedges, edge_off, lc;  // lc = min(L, n)
  int nchunks, ngroups;         // ngroups = ceil((n - lc) / kQGM) mid passes per layer
  int pad1, pad2;
};

// Launch table entry: instance and its first block index in this launch.
struct QEntry {
  int inst;
  int pad;
  uint64_t first_block;
};

__device__ __forceinline__ int find_entry_global(const QEntry *tab, int count, uint64_t b) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tab[mid].first_block <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// The launch's entry (instance) whose block range holds workgroup b, and that
// instance's index.  With at most kBlock entries the table's first blocks are
// staged in LDS by one load per thread and searched there: one global round
// trip instead of a chain of dependent loads before the chunk's own loads can
// start.  Every thread of the workgroup must call it (it synchronises).
struct QFound {
  int entry, inst;
  uint64_t first;
};
__device__ __forceinline__ QFound find_entry(const QEntry *tab, int count, uint64_t b) {
  __shared__ uint64_t sfb[kBlock];
  __shared__ int sinst[kBlock];
  if (count > kBlock) {
    const int e = find_entry_global(tab, count, b);
    return {e, tab[e].inst, tab[e].first_block};
  }
  if ((int)threadIdx.x < count) {
    sfb[threadIdx.x] = tab[threadIdx.x].first_block;
    sinst[threadIdx.x] = tab[threadIdx.x].inst;
  }
  __syncthreads();
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sfb[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return {lo, sinst[lo], sfb[lo]};
}

// find_entry with its results moved to scalar registers: every lane holds the
// same values, and marking them uniform lets the compiler load the instance's
// descriptor and tables with scalar loads and keep chunk-wide terms in SGPRs.
__device__ __forceinline__ QFound find_entry_u(const QEntry *tab, int count, uint64_t b) {
  const QFound f = find_entry(tab, count, b);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)f.first);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(f.first >> 32));
  return {__builtin_amdgcn_readfirstlane(f.entry), __builtin_amdgcn_readfirstlane(f.inst),
          (uint64_t)hi << 32 | lo};
}

__device__ __forceinline__ int cut_of(uint64_t i, const uint8_t *e, int ne) {
  int c = 0;
  for (int k = 0; k < ne; ++k) c += (int)(((i >> e[2 * k]) ^ (i >> e[2 * k + 1])) & 1);
  return c;
}

template <typename R>
__device__ __forceinline__ void cmul(R &x, R &y, R cr, R ci) {
  const R t = x * cr - y * ci;
  y = x * ci + y * cr;
  x = t;
}

// Block-wide FP64 sum; thread 0 returns it.
__device__ double qblock_sum(double v) {
  __shared__ double s[kBlock];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  const double r = s[0];
  __syncthreads();
  return r;
}

// Slot layout of instance j (nch = nchunks): energy [nch], then per layer l:
// gamma [nch], beta [(1 + ngroups) * nch] (low pass first, then mid groups).
__device__ __forceinline__ uint64_t slot_energy(const QInstDev &q) { return q.slot; }
__device__ __forceinline__ uint64_t slot_gamma(const QInstDev &q, int l) {
  return q.slot + (uint64_t)q.nchunks * (1 + (uint64_t)l * (2 + q.ngroups));
}
__device__ __forceinline__ uint64_t slot_beta(const QInstDev &q, int l, int part) {
  return slot_gamma(q, l) + (uint64_t)q.nchunks * (1 + part);
}

// RX(2b) (dir = +1) or RX(-2b) (dir = -1) on the 4 register-index bits.
template <typename T, typename R>
__device__ __forceinline__ void qrx16(T (&v)[16], R c0, R s0) {
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      if (r & (1 << b)) continue;
      const T x = v[r], y = v[r | (1 << b)];
      v[r].x = c0 * x.x + s0 * y.y;
      v[r].y = c0 * x.y - s0 * y.x;
      v[r | (1 << b)].x = c0 * y.x + s0 * x.y;
      v[r | (1 << b)].y = c0 * y.y - s0 * x.x;
    }
}

// Im(conj(l) a) of one amplitude pair in the state's precision: for complex64
// one FMA after one product (|error| <= 2 ulp of |l||a|, the state's own
// rounding scale), so only the per-pair result is widened to FP64 and the
// amplitudes stay in 32-bit registers.
template <typename T>
__device__ __forceinline__ auto qimcl(const T &l, const T &a) {
  return fma(l.x, a.y, -(l.y * a.x));
}

// sum over the 4 register bits b and pairs (r, r^b) of
// Im(conj(l_r) a_{r^b} + conj(l_{r^b}) a_r)  (= Im <l| X_b |a> over the group),
// accumulated in FP64
template <typename T>
__device__ __forceinline__ double qxpart16(const T (&a)[16], const T (&l)[16]) {
  double s0 = 0.0, s1 = 0.0;
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      if (r & (1 << b)) continue;
      const int u = r | (1 << b);
      s0 += (double)qimcl(l[r], a[u]);
      s1 += (double)qimcl(l[u], a[r]);
    }
  return s0 + s1;
}

// Cut counts of full chunks (n >= 12).  With i = base + j, j = t + 256 r < 2^12,
// split the edges by where their bits lie: both below 12 ("lo"), both at 12 or
// above ("hi"), or one of each ("cross", low bit a, high bit b):
//   cut(i) = lo(j) + hi(base) + sum_cross (j_a xor base_b)
//          = lo(j) + C + sum_{a < 12} j_a w_a,
//   C   = hi(base) + sum_a popc(base & M_a),   w_a = popc(M_a) - 2 popc(base & M_a),
// with M_a the mask of bit a's neighbours at bits >= 12.  lo(j) is a 4096-byte
// table per instance, built at batch creation and stored [t][r] so a thread's 16
// values are one 16-byte load; C and w_a depend on the chunk only and are
// computed once per wave in scalar registers.
struct QCutTab {
  uint32_t m[kQL];  // low bit a: its neighbours at bits >= 12
  uint32_t nh[32];  // high bit b: its neighbours at bits above b
};

struct QCutF {
  int K;                  // C + sum_{a < 8} t_a w_a
  int w8, w9, w10, w11;   // w_a of the register bits (uniform)
  uint4 lo;               // lo(t + 256 r) in byte r & 3 of word r >> 2
};

// `base` and `ii` must be wave-uniform.
__device__ __forceinline__ QCutF qcut_full(const QCutTab *__restrict__ cuts,
                                           const uint4 *__restrict__ lotab, int ii, int n,
                                           uint32_t base, int t) {
  QCutF f;
  f.lo = lotab[(size_t)ii * 256 + t];
  const QCutTab &c = cuts[ii];
  int C = 0, w[kQL];
#pragma unroll
  for (int a = 0; a < kQL; ++a) {
    const int pm = __popc(base & c.m[a]);
    C += pm;
    w[a] = __popc(c.m[a]) - 2 * pm;
  }
  for (int b = kQL; b < n; ++b) C += __popc(c.nh[b] & (((base >> b) & 1) ? ~base : base));
  int K = C;
#pragma unroll
  for (int a = 0; a < 8; ++a) K += ((t >> a) & 1) * w[a];
  f.K = K;
  f.w8 = w[8];
  f.w9 = w[9];
  f.w10 = w[10];
  f.w11 = w[11];
  return f;
}

__device__ __forceinline__ int qcut_r(const QCutF &f, int r) {
  const uint32_t word = r < 4 ? f.lo.x : r < 8 ? f.lo.y : r < 12 ? f.lo.z : f.lo.w;
  return f.K + (int)((word >> (8 * (r & 3))) & 0xffu) + (r & 1) * f.w8 + ((r >> 1) & 1) * f.w9 +
         ((r >> 2) & 1) * f.w10 + ((r >> 3) & 1) * f.w11;
}

// Per (instance, layer): the cost-layer phases exp(-i g s_c), s_c = |E| - 2c for
// c = 0..|E|, then (cos b, sin b) of the mixer, evaluated in FP64 and rounded to
// the state's precision once per run by k_qaoa_phases.
constexpr int kQPh = kQMaxEdges + 2;

template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_qaoa_phases(typename CV<R>::T *__restrict__ ph, const QInstDev *__restrict__ inst,
                  const double *__restrict__ gam, const double *__restrict__ bet, int p) {
  using T = typename CV<R>::T;
  const int ii = blockIdx.x / p, layer = blockIdx.x % p;
  const int ne = inst[ii].nedges;
  const double g = gam[ii * p + layer];
  T *o = ph + (size_t)blockIdx.x * kQPh;
  for (int c = threadIdx.x; c <= ne; c += kBlock) {
    double sn, cs;
    sincos(-g * (double)(ne - 2 * c), &sn, &cs);
    o[c] = T{(R)cs, (R)sn};
  }
  if (threadIdx.x == 0) {
    double sn, cs;
    sincos(bet[ii * p + layer], &sn, &cs);
    o[kQPh - 1] = T{(R)cs, (R)sn};
  }
}

template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_qaoa_low_fwd_full(typename CV<R>::T *__restrict__ st, const QInstDev *__restrict__ inst,
                        const QCutTab *__restrict__ cuts, const uint4 *__restrict__ lotab,
                        const typename CV<R>::T *__restrict__ ph, int p,
                        const QEntry *__restrict__ tab, int count, int layer) {
  using T = typename CV<R>::T;
  __shared__ T X[kQSlots];
  __shared__ R ph_r[kQMaxEdges + 1], ph_i[kQMaxEdges + 1];
  const QFound fe = find_entry_u(tab, count, blockIdx.x);
  const int ii = fe.inst;
  const QInstDev q = inst[ii];
  const uint32_t chunk = (uint32_t)(blockIdx.x - fe.first);
  const uint64_t base = (uint64_t)chunk << kQL;
  const T *phl = ph + (size_t)(ii * p + layer) * kQPh;
  const int t = threadIdx.x;
  // the chunk's loads go out first; the phase table and cut terms overlap them
  T v[16];
  const R amp0 = (R)(1.0 / sqrt((double)(1ull << q.n)));
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    if (layer == 0) {
      v[r].x = amp0;
      v[r].y = 0;
    } else {
      v[r] = ld_nt(st + q.psi + base + t + 256 * r);
    }
  }
  if (t <= q.nedges) {
    const T z = phl[t];
    ph_r[t] = z.x;
    ph_i[t] = z.y;
  }
  const QCutF cc = qcut_full(cuts, lotab, ii, q.n, (uint32_t)base, t);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    R x = v[r].x, y = v[r].y;
    const int c = qcut_r(cc, r);
    cmul(x, y, ph_r[c], ph_i[c]);
    v[r].x = x;
    v[r].y = y;
  }
  const T cb = phl[kQPh - 1];
  const R c0 = cb.x, s0 = cb.y;    // RX(2b) = [[c, -i s], [-i s, c]]
  qrx16(v, c0, s0);                // bits 8-11
  qtranspose(v, X, t, 2, 1);
  qrx16(v, c0, s0);                // bits 4-7
  qtranspose(v, X, t, 1, 0);       // writes the L1 slots this thread just read
  qrx16(v, c0, s0);                // bits 0-3
  qtranspose(v, X, t, 0, 2);
#pragma unroll
  for (int r = 0; r < 16; ++r) st_nt(st + q.psi + base + t + 256 * r, v[r]);
}

// ---- low pass, forward, chunks below 2^12 (n < 12): (init |+>) cost(l), RX on
// every bit, butterflies in LDS --------------------------------------------
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_qaoa_low_fwd(typename CV<R>::T *__restrict__ st, const QInstDev *__restrict__ inst,
                   const uint8_t *__restrict__ edges, const typename CV<R>::T *__restrict__ ph,
                   int p, const QEntry *__restrict__ tab, int count, int layer) {
  using T = typename CV<R>::T;
  __shared__ R xr[kQChunk / 2], xi[kQChunk / 2];  // chunks below 2^12 (n < 12)
  __shared__ R ph_r[kQMaxEdges + 1], ph_i[kQMaxEdges + 1];
  const QFound fe = find_entry_u(tab, count, blockIdx.x);
  const int ii = fe.inst;
  const QInstDev q = inst[ii];
  const uint64_t chunk = blockIdx.x - fe.first;
  const int m = 1 << q.lc;
  const uint64_t base = chunk << q.lc;
  const uint8_t *e = edges + 2 * q.edge_off;
  const T *phl = ph + (size_t)(ii * p + layer) * kQPh;
  if ((int)threadIdx.x <= q.nedges) {
    const T z = phl[threadIdx.x];
    ph_r[threadIdx.x] = z.x;
    ph_i[threadIdx.x] = z.y;
  }
  __syncthreads();
  const R amp0 = (R)(1.0 / sqrt((double)(1ull << q.n)));
  for (int j = threadIdx.x; j < m; j += kBlock) {
    R x, y;
    if (layer == 0) {
      x = amp0;
      y = 0;
    } else {
      const T v = ld_nt(st + q.psi + base + j);
      x = v.x;
      y = v.y;
    }
    const int c = cut_of(base + j, e, q.nedges);
    cmul(x, y, ph_r[c], ph_i[c]);
    xr[j] = x;
    xi[j] = y;
  }
  __syncthreads();
  const T cb = phl[kQPh - 1];
  const R c0 = cb.x, s0 = cb.y;  // RX(2b) = [[c, -i s], [-i s, c]]
  for (int bit = 0; bit < q.lc; ++bit) {
    for (int f = threadIdx.x; f < m / 2; f += kBlock) {
      const int lo = f & ((1 << bit) - 1);
      const int i0 = ((f ^ lo) << 1) | lo, i1 = i0 | (1 << bit);
      const R ar = xr[i0], ai = xi[i0], br = xr[i1], bi = xi[i1];
      xr[i0] = c0 * ar + s0 * bi;
      xi[i0] = c0 * ai - s0 * br;
      xr[i1] = c0 * br + s0 * ai;
      xi[i1] = c0 * bi - s0 * ar;
    }
    __syncthreads();
  }
  for (int j = threadIdx.x; j < m; j += kBlock) {
    T o;
    o.x = xr[j];
    o.y = xi[j];
    st_nt(st + q.psi + base + j, o);
  }
}

// ---- energy: lambda = C psi, partial sum of cut(i) |psi_i|^2 ----------------
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_qaoa_energy(typename CV<R>::T *__restrict__ st, const QInstDev *__restrict__ inst,
                  const uint8_t *__restrict__ edges, const QCutTab *__restrict__ cuts,
                  const uint4 *__restrict__ lotab, const QEntry *__restrict__ tab, int count,
                  double *__restrict__ partial) {
  using T = typename CV<R>::T;
  const QFound fe = find_entry_u(tab, count, blockIdx.x);
  const int ii = fe.inst;
  const QInstDev q = inst[ii];
  const uint64_t chunk = blockIdx.x - fe.first;
  const int m = 1 << q.lc;
  const uint64_t base = chunk << q.lc;
  const uint8_t *e = edges + 2 * q.edge_off;
  double acc = 0.0;
  if (q.lc == kQL) {  // uniform per workgroup
    // all 16 loads of the thread in flight at once (m = 16 x kBlock)
    const int t = threadIdx.x;
    T v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = ld_nt(st + q.psi + base + t + kBlock * r);
    const QCutF cc = qcut_full(cuts, lotab, ii, q.n, (uint32_t)base, t);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int c = qcut_r(cc, r);
      acc += (double)c * ((double)v[r].x * v[r].x + (double)v[r].y * v[r].y);
      T o;
      o.x = (R)c * v[r].x;
      o.y = (R)c * v[r].y;
      st_nt(st + q.lam + base + t + kBlock * r, o);
    }
  } else {
    for (int j = threadIdx.x; j < m; j += kBlock) {
      const T v = ld_nt(st + q.psi + base + j);
      const int c = cut_of(base + j, e, q.nedges);
      acc += (double)c * ((double)v.x * v.x + (double)v.y * v.y);
      T o;
      o.x = (R)c * v.x;
      o.y = (R)c * v.y;
      st_nt(st + q.lam + base + j, o);
    }
  }
  acc = qblock_sum(acc);
  if (threadIdx.x == 0) partial[slot_energy(q) + chunk] = acc;
}

template <typename R>
__global__ void __launch_bounds__(kBlock, sizeof(R) == 4 ? 2 : 1)
    k_qaoa_low_rev_full(typename CV<R>::T *__restrict__ st, const QInstDev *__restrict__ inst,
                        const QCutTab *__restrict__ cuts, const uint4 *__restrict__ lotab,
                        const typename CV<R>::T *__restrict__ ph, int p,
                        const QEntry *__restrict__ tab, int count, int layer,
                        double *__restrict__ partial) {
  using T = typename CV<R>::T;
  __shared__ T X[kQSlots];
  __shared__ R ph_r[kQMaxEdges + 1], ph_i[kQMaxEdges + 1];
  const QFound fe = find_entry_u(tab, count, blockIdx.x);
  const int ii = fe.inst;
  const QInstDev q = inst[ii];
  const uint32_t chunk = (uint32_t)(blockIdx.x - fe.first);
  const uint64_t base = (uint64_t)chunk << kQL;
  const T *phl = ph + (size_t)(ii * p + layer) * kQPh;
  const int t = threadIdx.x;
  // the chunk's loads go out first; the phase table and cut terms overlap them
  T a[16], l[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    a[r] = ld_nt(st + q.psi + base + t + 256 * r);
    l[r] = ld_nt(st + q.lam + base + t + 256 * r);
  }
  if (t <= q.nedges) {  // undo: exp(+i g s), the conjugate phases
    const T z = phl[t];
    ph_r[t] = z.x;
    ph_i[t] = -z.y;
  }
  const QCutF cc = qcut_full(cuts, lotab, ii, q.n, (uint32_t)base, t);
  const T cb = phl[kQPh - 1];
  const R c0 = cb.x, s0 = -cb.y;  // RX(-2b)
  // beta partial (any point inside the mixer layer) and undo, 4 bits at a time
  double accb = qxpart16(a, l);  // bits 8-11
  qrx16(a, c0, s0);
  qrx16(l, c0, s0);
  qtranspose(a, X, t, 2, 1);
  __syncthreads();
  qtranspose(l, X, t, 2, 1);
  accb += qxpart16(a, l);  // bits 4-7
  qrx16(a, c0, s0);
  qrx16(l, c0, s0);
  __syncthreads();
  qtranspose(a, X, t, 1, 0);
  __syncthreads();
  qtranspose(l, X, t, 1, 0);
  accb += qxpart16(a, l);  // bits 0-3
  qrx16(a, c0, s0);
  qrx16(l, c0, s0);
  __syncthreads();
  qtranspose(a, X, t, 0, 2);
  __syncthreads();
  qtranspose(l, X, t, 0, 2);
  // gamma partial Im(conj(lambda) s psi) and undo the cost layer
  double accg = 0.0;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int c = qcut_r(cc, r);
    const double sv = (double)(q.nedges - 2 * c);
    R x = a[r].x, y = a[r].y, u = l[r].x, w = l[r].y;
    accg += sv * ((double)u * y - (double)w * x);
    cmul(x, y, ph_r[c], ph_i[c]);
    cmul(u, w, ph_r[c], ph_i[c]);
    T o, m;
    o.x = x;
    o.y = y;
    m.x = u;
    m.y = w;
    st_nt(st + q.psi + base + t + 256 * r, o);
    st_nt(st + q.lam + base + t + 256 * r, m);
  }
  accb = qblock_sum(accb);
  if (t == 0) partial[slot_beta(q, layer, 0) + chunk] = accb;
  accg = qblock_sum(accg);
  if (t == 0) partial[slot_gamma(q, layer) + chunk] = accg;
}

// ---- mid pass: RX(2b) (forward) or RX(-2b) and the beta partial (reverse) on
// the rotated bits h0 .. h0+W-1 ----------------------------------------------
// Geometry (qmid_pos, chunk_layout.hpp) with C = kQC = 5 contiguous
// low bits and W = min(7, n - h0) rotated bits: loads in layout L2 (local
// bits 8-11 in the register index: rotated bits h0+3..), one LDS transpose to
// L1 (local bits 4-7: bit 4 carried, 5-7 rotated), and back to L2 for the
// stores, so every global access is 256 B runs (32 consecutive complex64
// amplitudes); 7 rotated bits still take the 14 high bits of a 26-qubit
// instance in two passes.
struct QMid {
  uint32_t base;     // global offset of the chunk (uniform)
  uint32_t t2;       // thread part of the L2 offset
  int W;
};

// Chunk of a mid-pass workgroup: its index within the instance's launch, or
// (scatter) that index bit-reversed over the instance's n - 12 chunk bits, so
// concurrent workgroups take chunks far apart instead of neighbouring ones.
__device__ __forceinline__ uint32_t qmid_chunk(uint64_t rel, int n, int scatter) {
  const uint32_t c = (uint32_t)rel;
  return scatter && n > kQL ? __builtin_bitreverse32(c) >> (32 - (n - kQL)) : c;
}
__device__ __forceinline__ QMid qmid_setup(uint32_t chunk, int n, int h0, int t) {
  QMid m;
  m.W = min(kQGM, n - h0);
  const int A = h0 - 12 + m.W;  // chunk bits below h0, at 12-W ..
  m.base = ((chunk & ((1u << A) - 1)) << (12 - m.W)) | ((chunk >> A) << (h0 + m.W));
  m.t2 = (uint32_t)(t & 15);
#pragma unroll
  for (int k = 4; k < 8; ++k) m.t2 |= (uint32_t)((t >> k) & 1) << qmid_pos(k, m.W, h0, kQC);
  return m;
}

// register part of the L2 offset: local bits 8-11 in r
__device__ __forceinline__ uint32_t qmid_r(int r, int W, int h0) {
  uint32_t o = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) o |= (uint32_t)((r >> k) & 1) << qmid_pos(8 + k, W, h0, kQC);
  return o;
}

// RX on the register-index bits set in `mask` (uniform), as qrx16
template <typename T, typename R>
__device__ __forceinline__ void qrx16m(T (&v)[16], R c0, R s0, int mask) {
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    if (!((mask >> b) & 1)) continue;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      if (r & (1 << b)) continue;
      const T x = v[r], y = v[r | (1 << b)];
      v[r].x = c0 * x.x + s0 * y.y;
      v[r].y = c0 * x.y - s0 * y.x;
      v[r | (1 << b)].x = c0 * y.x + s0 * x.y;
      v[r | (1 << b)].y = c0 * y.y - s0 * x.x;
    }
  }
}

// qxpart16 over the register-index bits set in `mask` (uniform)
template <typename T>
__device__ __forceinline__ double qxpart16m(const T (&a)[16], const T (&l)[16], int mask) {
  double s0 = 0.0, s1 = 0.0;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    if (!((mask >> b) & 1)) continue;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      if (r & (1 << b)) continue;
      const int u = r | (1 << b);
      s0 += (double)qimcl(l[r], a[u]);
      s1 += (double)qimcl(l[u], a[r]);
    }
  }
  return s0 + s1;
}

// register-bit masks of the rotated local bits kQC..kQC+W-1: L2 phase (local
// 8-11), L1 phase (local 4-7)
__device__ __forceinline__ int qmid_mask2(int W) {
  return ((((1 << W) - 1) << kQC) >> 8) & 15;
}
__device__ __forceinline__ int qmid_mask1(int W) { return ((((1 << W) - 1) << kQC) >> 4) & 15; }

template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_qaoa_mid_fwd(typename CV<R>::T *__restrict__ st, const QInstDev *__restrict__ inst,
                   const typename CV<R>::T *__restrict__ ph, int p,
                   const QEntry *__restrict__ tab, int count, int layer, int h0, int scatter) {
  using T = typename CV<R>::T;
  __shared__ T X[kQSlots];
  const QFound fe = find_entry_u(tab, count, blockIdx.x);
  const int ii = fe.inst;
  const QInstDev q = inst[ii];
  const int t = threadIdx.x;
  const QMid m = qmid_setup(qmid_chunk(blockIdx.x - fe.first, q.n, scatter), q.n, h0, t);
  T *s = st + q.psi + m.base;
  T v[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = ld_nt(s + (m.t2 | qmid_r(r, m.W, h0)));
  const T cb = ph[(size_t)(ii * p + layer) * kQPh + kQPh - 1];
  const R c0 = cb.x, s0 = cb.y;  // RX(2b) = [[c, -i s], [-i s, c]]
  qrx16m(v, c0, s0, qmid_mask2(m.W));
  qtranspose(v, X, t, 2, 1);
  qrx16m(v, c0, s0, qmid_mask1(m.W));
  qtranspose(v, X, t, 1, 2);  // writes the L1 slots this thread just read
#pragma unroll
  for (int r = 0; r < 16; ++r) st_nt(s + (m.t2 | qmid_r(r, m.W, h0)), v[r]);
}

template <typename R>
__global__ void __launch_bounds__(kBlock, sizeof(R) == 4 ? 2 : 1)
    k_qaoa_mid_rev(typename CV<R>::T *__restrict__ st, const QInstDev *__restrict__ inst,
                   const typename CV<R>::T *__restrict__ ph, int p,
                   const QEntry *__restrict__ tab, int count, int layer, int h0,
                   double *__restrict__ partial, int scatter) {
  using T = typename CV<R>::T;
  __shared__ T X[kQSlots];
  const QFound fe = find_entry_u(tab, count, blockIdx.x);
  const int ii = fe.inst;
  const QInstDev q = inst[ii];
  const int t = threadIdx.x;
  const uint32_t chunk = qmid_chunk(blockIdx.x - fe.first, q.n, scatter);
  const QMid m = qmid_setup(chunk, q.n, h0, t);
  T *sa = st + q.psi + m.base, *sl = st + q.lam + m.base;
  T a[16], l[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const uint32_t o = m.t2 | qmid_r(r, m.W, h0);
    a[r] = ld_nt(sa + o);
    l[r] = ld_nt(sl + o);
  }
  const T cb = ph[(size_t)(ii * p + layer) * kQPh + kQPh - 1];
  const R c0 = cb.x, s0 = -cb.y;  // RX(-2b)
  const int m2 = qmid_mask2(m.W), m1 = qmid_mask1(m.W);
  double acc = qxpart16m(a, l, m2);
  qrx16m(a, c0, s0, m2);
  qrx16m(l, c0, s0, m2);
  qtranspose(a, X, t, 2, 1);
  __syncthreads();
  qtranspose(l, X, t, 2, 1);
  acc += qxpart16m(a, l, m1);
  qrx16m(a, c0, s0, m1);
  qrx16m(l, c0, s0, m1);
  qtranspose(a, X, t, 1, 2);  // writes the L1 slots this thread just read
  __syncthreads();
  qtranspose(l, X, t, 1, 2);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const uint32_t o = m.t2 | qmid_r(r, m.W, h0);
    st_nt(sa + o, a[r]);
    st_nt(sl + o, l[r]);
  }
  acc = qblock_sum(acc);
  if (t == 0) partial[slot_beta(q, layer, 1 + (h0 - kQL) / kQGM) + chunk] = acc;
}

// ---- low pass, reverse, chunks below 2^12 (n < 12): beta partials and undo RX
// on the bits, gamma partial
// and undo cost, on psi and lambda --------------------------------------------
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_qaoa_low_rev(typename CV<R>::T *__restrict__ st, const QInstDev *__restrict__ inst,
                   const uint8_t *__restrict__ edges, const typename CV<R>::T *__restrict__ ph,
                   int p, const QEntry *__restrict__ tab, int count, int layer,
                   double *__restrict__ partial) {
  using T = typename CV<R>::T;
  __shared__ R xr[kQChunk / 2], xi[kQChunk / 2], yr[kQChunk / 2], yi[kQChunk / 2];  // n < 12
  __shared__ R ph_r[kQMaxEdges + 1], ph_i[kQMaxEdges + 1];
  const QFound fe = find_entry_u(tab, count, blockIdx.x);
  const int ii = fe.inst;
  const QInstDev q = inst[ii];
  const uint64_t chunk = blockIdx.x - fe.first;
  const int m = 1 << q.lc;
  const uint64_t base = chunk << q.lc;
  const uint8_t *e = edges + 2 * q.edge_off;
  const T *phl = ph + (size_t)(ii * p + layer) * kQPh;
  if ((int)threadIdx.x <= q.nedges) {  // undo: exp(+i g s), the conjugate phases
    const T z = phl[threadIdx.x];
    ph_r[threadIdx.x] = z.x;
    ph_i[threadIdx.x] = -z.y;
  }
  for (int j = threadIdx.x; j < m; j += kBlock) {
    const T v = ld_nt(st + q.psi + base + j), w = ld_nt(st + q.lam + base + j);
    xr[j] = v.x;
    xi[j] = v.y;
    yr[j] = w.x;
    yi[j] = w.y;
  }
  __syncthreads();
  const T cb = phl[kQPh - 1];
  const R c0 = cb.x, s0 = cb.y;
  double accb = 0.0;
  for (int bit = 0; bit < q.lc; ++bit) {
    for (int f = threadIdx.x; f < m / 2; f += kBlock) {
      const int lo = f & ((1 << bit) - 1);
      const int i0 = ((f ^ lo) << 1) | lo, i1 = i0 | (1 << bit);
      const R ar = xr[i0], ai = xi[i0], br = xr[i1], bi = xi[i1];
      const R lar = yr[i0], lai = yi[i0], lbr = yr[i1], lbi = yi[i1];
      accb += (double)lar * bi - (double)lai * br + (double)lbr * ai - (double)lbi * ar;
      xr[i0] = c0 * ar - s0 * bi;
      xi[i0] = c0 * ai + s0 * br;
      xr[i1] = c0 * br - s0 * ai;
      xi[i1] = c0 * bi + s0 * ar;
      yr[i0] = c0 * lar - s0 * lbi;
      yi[i0] = c0 * lai + s0 * lbr;
      yr[i1] = c0 * lbr - s0 * lai;
      yi[i1] = c0 * lbi + s0 * lar;
    }
    __syncthreads();
  }
  double accg = 0.0;
  for (int j = threadIdx.x; j < m; j += kBlock) {
    const int c = cut_of(base + j, e, q.nedges);
    const double s = (double)(q.nedges - 2 * c);
    R x = xr[j], y = xi[j], u = yr[j], w = yi[j];
    // Im(conj(lambda) s psi)
    accg += s * ((double)u * y - (double)w * x);
    cmul(x, y, ph_r[c], ph_i[c]);
    cmul(u, w, ph_r[c], ph_i[c]);
    T o, l;
    o.x = x;
    o.y = y;
    l.x = u;
    l.y = w;
    st_nt(st + q.psi + base + j, o);
    st_nt(st + q.lam + base + j, l);
  }
  accb = qblock_sum(accb);
  if (threadIdx.x == 0) partial[slot_beta(q, layer, 0) + chunk] = accb;
  accg = qblock_sum(accg);
  if (threadIdx.x == 0) partial[slot_gamma(q, layer) + chunk] = accg;
}

// ---- finaliser: out[j] = {E, dE/dg_0..p-1, dE/db_0..p-1} -------------------
__global__ void __launch_bounds__(kBlock)
    k_qaoa_finalize(const QInstDev *__restrict__ inst, int ninst, int p,
                    const double *__restrict__ partial, double *__restrict__ out) {
  const int j = blockIdx.x;
  if (j >= ninst) return;
  const QInstDev q = inst[j];
  const uint64_t nch = q.nchunks;
  for (int v = 0; v < 1 + 2 * p; ++v) {
    uint64_t first, len;
    double scale;
    if (v == 0) {
      first = slot_energy(q);
      len = nch;
      scale = 1.0;
    } else if (v <= p) {
      first = slot_gamma(q, v - 1);
      len = nch;
      scale = 2.0;
    } else {
      first = slot_beta(q, v - 1 - p, 0);
      len = nch * (uint64_t)(1 + q.ngroups);
      scale = 2.0;
    }
    double acc = 0.0;
    for (uint64_t k = threadIdx.x; k < len; k += kBlock) acc += partial[first + k];
    acc = qblock_sum(acc);
    if (threadIdx.x == 0) out[(uint64_t)j * (1 + 2 * p) + v] = scale * acc;
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct QaoaBatch {
  sq_dtype_t dt;
  int ninst, p, maxn;
  std::vector<QInstDev> host_inst;
  void *states = nullptr;
  QInstDev *inst = nullptr;
  uint8_t *edges = nullptr;
  QCutTab *cuts = nullptr;   // per instance
  uint8_t *lotab = nullptr;  // per instance, 4096 bytes
  void *phase = nullptr;     // per (instance, layer), kQPh amplitudes
  double *params = nullptr;  // gam [ninst*p], bet [ninst*p]
  double *partial = nullptr;
  double *out = nullptr;
  QEntry *tables = nullptr;
  struct Launch {
    int kind, layer, bit;  // kind: 0 low_fwd, 2 energy, 4 low_rev, 5 mid_fwd, 6 mid_rev
    int nb;                // low passes: 1 for the full-chunk kernel
    size_t tab_off;
    int count;
    uint64_t blocks;
  };
  std::vector<Launch> launches;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  double state_bytes = 0, traffic_bytes = 0;
};

static void free_batch(QaoaBatch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (void *ptr : {(void *)b->states, (void *)b->inst, (void *)b->edges, (void *)b->cuts,
                    (void *)b->lotab, (void *)b->phase, (void *)b->params, (void *)b->partial,
                    (void *)b->out, (void *)b->tables})
    if (ptr) (void)hipFree(ptr);
}

int qaoa_create(void **handle, sq_dtype_t dt, uint32_t ninst, const uint32_t *nq,
                const uint32_t *nedges, const uint32_t *edges, uint32_t p) {
  if (ninst == 0 || p == 0) return fail(SQ_ERR_INVALID, "qaoa: empty batch");
  auto B = std::make_unique<QaoaBatch>();
  B->dt = dt;
  B->ninst = (int)ninst;
  B->p = (int)p;
  B->maxn = 0;
  const size_t item = itemsize(dt);
  uint64_t amp = 0, slot = 0;
  size_t eoff = 0;
  std::vector<uint8_t> hedges;
  std::vector<QCutTab> hcuts(ninst);
  std::vector<uint8_t> hlo((size_t)ninst * kQChunk, 0);
  for (uint32_t j = 0; j < ninst; ++j) {
    const int n = (int)nq[j];
    if (n < 2 || n > 30) return fail(SQ_ERR_INVALID, "qaoa: qubit counts must be in [2, 30]");
    if (nedges[j] > (uint32_t)kQMaxEdges) return fail(SQ_ERR_INVALID, "qaoa: too many edges");
    QInstDev q{};
    q.n = n;
    q.lc = std::min(n, kQL);
    q.nchunks = 1 << (n - q.lc);
    q.ngroups = (n - q.lc + kQGM - 1) / kQGM;
    q.nedges = (int)nedges[j];
    q.edge_off = (int)(hedges.size() / 2);
    std::vector<uint32_t> seen(n, 0);  // seen[a]: bit b set if edge {a, b} was given
    for (uint32_t k = 0; k < nedges[j]; ++k) {
      const uint32_t a = edges[2 * (eoff + k)], b = edges[2 * (eoff + k) + 1];
      if (a >= (uint32_t)n || b >= (uint32_t)n || a == b)
        return fail(SQ_ERR_INVALID, "qaoa: bad edge");
      // simple graphs only: the cut tables hold neighbour sets as bit masks
      if ((seen[a] >> b) & 1) return fail(SQ_ERR_INVALID, "qaoa: repeated edge");
      seen[a] |= 1u << b;
      seen[b] |= 1u << a;
      hedges.push_back((uint8_t)(n - 1 - a));  // qubit -> bit position
      hedges.push_back((uint8_t)(n - 1 - b));
    }
    if (q.lc == kQL) {  // cut tables of full chunks (see qcut_full)
      QCutTab &ct = hcuts[j];
      std::vector<std::pair<int, int>> lo_edges;
      for (uint32_t k = 0; k < nedges[j]; ++k) {
        const int u = n - 1 - (int)edges[2 * (eoff + k)], v = n - 1 - (int)edges[2 * (eoff + k) + 1];
        const int x = std::min(u, v), y = std::max(u, v);
        if (y < kQL) lo_edges.emplace_back(x, y);
        else if (x < kQL) ct.m[x] |= 1u << y;
        else ct.nh[x] |= 1u << y;
      }
      uint8_t *lo = hlo.data() + (size_t)j * kQChunk;
      for (int t = 0; t < 256; ++t)
        for (int r = 0; r < 16; ++r) {
          const int jj = t + 256 * r;
          int c = 0;
          for (auto &e : lo_edges) c += ((jj >> e.first) ^ (jj >> e.second)) & 1;
          lo[t * 16 + r] = (uint8_t)c;
        }
    }
    eoff += nedges[j];
    q.psi = amp;
    q.lam = amp + (1ull << n);
    amp += 2ull << n;
    q.slot = slot;
    slot += (uint64_t)q.nchunks * (1 + (uint64_t)p * (2 + q.ngroups));
    B->host_inst.push_back(q);
    B->maxn = std::max(B->maxn, n);
    B->state_bytes += 2.0 * std::ldexp(1.0, n) * item;
  }
  if (hedges.empty()) hedges.push_back(0);
  auto dmalloc = [](void **ptr, size_t bytes) -> int {
    if (hipMalloc(ptr, bytes) != hipSuccess) {
      (void)hipGetLastError();
      *ptr = nullptr;
      return fail(SQ_ERR_OOM, "qaoa: device allocation of " + std::to_string(bytes) + " failed");
    }
    return SQ_OK;
  };
  int rc = SQ_OK;
  if (!rc) rc = dmalloc(&B->states, amp * item);
  if (!rc) rc = dmalloc((void **)&B->inst, ninst * sizeof(QInstDev));
  if (!rc) rc = dmalloc((void **)&B->edges, hedges.size());
  if (!rc) rc = dmalloc((void **)&B->cuts, ninst * sizeof(QCutTab));
  if (!rc) rc = dmalloc((void **)&B->lotab, hlo.size());
  if (!rc) rc = dmalloc(&B->phase, (size_t)ninst * p * kQPh * item);
  if (!rc) rc = dmalloc((void **)&B->params, 2 * ninst * p * sizeof(double));
  if (!rc) rc = dmalloc((void **)&B->partial, std::max<uint64_t>(slot, 1) * sizeof(double));
  if (!rc) rc = dmalloc((void **)&B->out, ninst * (1 + 2 * p) * sizeof(double));
  // launch sequence and tables
  std::vector<QEntry> tabs;
  auto add = [&](int kind, int layer, int bit, int nb, auto pred) {
    QaoaBatch::Launch L{kind, layer, bit, nb, tabs.size(), 0, 0};
    for (int j = 0; j < (int)ninst; ++j) {
      const QInstDev &q = B->host_inst[j];
      if (!pred(q)) continue;
      tabs.push_back(QEntry{j, 0, L.blocks});
      L.blocks += (uint64_t)q.nchunks;
      L.count++;
    }
    if (L.count) B->launches.push_back(L);
  };
  auto all = [](const QInstDev &) { return true; };
  auto full = [](const QInstDev &q) { return q.lc == kQL; };
  auto small = [](const QInstDev &q) { return q.lc < kQL; };
  auto above = [](int h0) { return [h0](const QInstDev &q) { return q.n > h0; }; };
  for (int l = 0; l < (int)p; ++l) {
    add(0, l, 0, 1, full);  // nb = 1: full-chunk kernel
    add(0, l, 0, 0, small);
    for (int h0 = kQL; h0 < B->maxn; h0 += kQGM) add(5, l, h0, 0, above(h0));
  }
  add(2, 0, 0, 0, all);
  for (int l = (int)p - 1; l >= 0; --l) {
    for (int h0 = kQL + ((B->maxn - kQL - 1) / kQGM) * kQGM; h0 >= kQL; h0 -= kQGM)
      add(6, l, h0, 0, above(h0));
    add(4, l, 0, 1, full);
    add(4, l, 0, 0, small);
  }
  if (!rc) rc = dmalloc((void **)&B->tables, tabs.size() * sizeof(QEntry));
  if (rc) {
    free_batch(B.get());
    return rc;
  }
  if (hipMemcpy(B->inst, B->host_inst.data(), ninst * sizeof(QInstDev), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(B->edges, hedges.data(), hedges.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(B->cuts, hcuts.data(), ninst * sizeof(QCutTab), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(B->lotab, hlo.data(), hlo.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(B->tables, tabs.data(), tabs.size() * sizeof(QEntry), hipMemcpyHostToDevice) !=
          hipSuccess) {
    free_batch(B.get());
    return fail(SQ_ERR_HIP, "qaoa: upload failed");
  }
  // HBM traffic of one run (reads + writes of the states)
  for (auto &L : B->launches) {
    const double per = (L.kind == 0 || L.kind == 5) ? 2.0 : (L.kind == 2 ? 3.0 : 4.0);
    for (int k = 0; k < L.count; ++k) {
      const QInstDev &q = B->host_inst[tabs[L.tab_off + k].inst];
      const double amps = std::ldexp(1.0, q.n);
      B->traffic_bytes += per * amps * item * ((L.kind == 0 && L.layer == 0) ? 0.5 : 1.0);
    }
  }
  *handle = B.release();
  return SQ_OK;
}

template <typename R>
static int enqueue_t(QaoaBatch *B, hipStream_t st) {
  using T = typename CV<R>::T;
  T *s = (T *)B->states;
  const T *ph = (const T *)B->phase;
  const uint4 *lo = (const uint4 *)B->lotab;
  const int p = B->p;
  const double *gam = B->params, *bet = B->params + (size_t)B->ninst * p;
  hipLaunchKernelGGL(k_qaoa_phases<R>, dim3(B->ninst * p), dim3(kBlock), 0, st, (T *)B->phase,
                     B->inst, gam, bet, p);
  for (auto &L : B->launches) {
    const QEntry *tab = B->tables + L.tab_off;
    const dim3 g((unsigned)L.blocks), bl(kBlock);
    switch (L.kind) {
      case 0:
        if (L.nb)
          hipLaunchKernelGGL(k_qaoa_low_fwd_full<R>, g, bl, 0, st, s, B->inst, B->cuts, lo, ph, p,
                             tab, L.count, L.layer);
        else
          hipLaunchKernelGGL(k_qaoa_low_fwd<R>, g, bl, 0, st, s, B->inst, B->edges, ph, p, tab,
                             L.count, L.layer);
        break;
      case 2:
        hipLaunchKernelGGL(k_qaoa_energy<R>, g, bl, 0, st, s, B->inst, B->edges, B->cuts, lo, tab,
                           L.count, B->partial);
        break;
      case 5:
        hipLaunchKernelGGL(k_qaoa_mid_fwd<R>, g, bl, 0, st, s, B->inst, ph, p, tab, L.count,
                           L.layer, L.bit, tuning().qaoa_mid_scatter);
        break;
      case 6:
        hipLaunchKernelGGL(k_qaoa_mid_rev<R>, g, bl, 0, st, s, B->inst, ph, p, tab, L.count,
                           L.layer, L.bit, B->partial, tuning().qaoa_mid_scatter);
        break;
      case 4:
        if (L.nb)
          hipLaunchKernelGGL(k_qaoa_low_rev_full<R>, g, bl, 0, st, s, B->inst, B->cuts, lo, ph, p,
                             tab, L.count, L.layer, B->partial);
        else
          hipLaunchKernelGGL(k_qaoa_low_rev<R>, g, bl, 0, st, s, B->inst, B->edges, ph, p, tab,
                             L.count, L.layer, B->partial);
        break;
    }
  }
  hipLaunchKernelGGL(k_qaoa_finalize, dim3(B->ninst), dim3(kBlock), 0, st, B->inst, B->ninst,
                     B->p, B->partial, B->out);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

static int enqueue(QaoaBatch *B, hipStream_t st) {
  return B->dt == SQ_COMPLEX64 ? enqueue_t<float>(B, st) : enqueue_t<double>(B, st);
}

int qaoa_run(void *handle, const double *gammas, const double *betas, int use_graph,
             double *out_host, hipStream_t st) {
  QaoaBatch *B = (QaoaBatch *)handle;
  const size_t np = (size_t)B->ninst * B->p;
  SQ_HIP(hipMemcpyAsync(B->params, gammas, np * sizeof(double), hipMemcpyHostToDevice, st));
  SQ_HIP(hipMemcpyAsync(B->params + np, betas, np * sizeof(double), hipMemcpyHostToDevice, st));
  if (use_graph) {
    if (!B->exec) {
      // capture the launch sequence once; parameters live in device memory
      hipStream_t cs;
      SQ_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
      int rc = SQ_OK;
      if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess)
        rc = fail(SQ_ERR_HIP, "qaoa: hipStreamBeginCapture failed");
      if (!rc) rc = enqueue(B, cs);
      hipGraph_t g = nullptr;
      if (hipStreamEndCapture(cs, &g) != hipSuccess && !rc)
        rc = fail(SQ_ERR_HIP, "qaoa: hipStreamEndCapture failed");
      (void)hipStreamDestroy(cs);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      B->graph = g;
      SQ_HIP(hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0));
    }
    SQ_HIP(hipGraphLaunch(B->exec, st));
  } else {
    SQ_TRY(enqueue(B, st));
  }
  SQ_HIP(hipMemcpyAsync(out_host, B->out, (size_t)B->ninst * (1 + 2 * B->p) * sizeof(double),
                        hipMemcpyDeviceToHost, st));
  SQ_HIP(hipStreamSynchronize(st));
  return SQ_OK;
}

int qaoa_destroy(void *handle) {
  QaoaBatch *B = (QaoaBatch *)handle;
  if (!B) return SQ_OK;
  free_batch(B);
  delete B;
  return SQ_OK;
}

int qaoa_stats(void *handle, uint64_t *launches, double *state_bytes, double *traffic_bytes) {
  QaoaBatch *B = (QaoaBatch *)handle;
  if (launches) *launches = B->launches.size() + 2;  // + phase tables, finaliser
  if (state_bytes) *state_bytes = B->state_bytes;
  if (traffic_bytes) *traffic_bytes = B->traffic_bytes;
  return SQ_OK;
}

}  // namespace sq
