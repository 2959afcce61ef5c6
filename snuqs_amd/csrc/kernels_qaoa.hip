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
// Instances of n >= 12 qubits are swept in window passes.  A workgroup holds
// 4096 amplitudes: the 32 consecutive ones of bits 0-4 for each value of 7
// more bits, the window (bits 5-11 for window 0, then 12-18, 19-25, ...; see
// qmid_pos), so every global access moves 256-byte runs of complex64.  An
// instance has k = 1 + ceil((n - 12) / 7) windows.  The mixer of a layer is
// split over the windows (the X_q commute), and the pass that ends a layer on
// its window also applies the next layer's cost (diagonal: it needs only each
// amplitude's global index) and starts the next layer's mixer on that window
// and on bits 0-4.  A forward sweep is then p (k - 1) + 1 passes instead of
// p k (p + 1 for k = 1).  Its last pass (the "turn") also forms the energy
// and lambda = C psi and undoes its own window on lambda, so it writes lambda
// only: psi in memory is already its value before that pass.  The reverse
// sweep undoes the other passes in reverse order with the gradient partials;
// its last pass (the first forward pass undone) only reads.  Instances below
// 12 qubits take one workgroup per instance and layer, in LDS.
// Grouping is exact for the gradients too: every X_q commutes with the whole
// mixer layer, so <lambda| X_q |psi> has the same value anywhere inside it and
// a window's partial can be taken wherever psi and lambda stand at the same
// point of the layer.  Each launch runs one pass (per instance, from its
// launch-table entry) for every instance that has one, so one sequence of
// launches serves the whole batch (and is captured in one hipGraph).
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
constexpr int kQMaxLayers = 256;
constexpr int kQGM = 7;            // window bits
constexpr int kQC = 5;             // contiguous low bits beside them

struct QInstDev {
  uint64_t psi, lam;  // amplitude offsets into the batch state buffer
  uint64_t slot;      // first partial slot
  int n, nedges, edge_off, lc;  // lc = min(12, n)
  int nchunks, nwin;  // nwin: windows (n >= 12), beta partial sets per layer; 1 below
  int win, pad;       // first window (QWin) of the instance
};

// Window-pass descriptor (launch-table entry): window g, flags, the layer lp
// whose mixer the pass applies on the window bits first ("pre"), and the layer
// lc whose cost it applies next, followed by lc's mixer on the window bits
// ("bnd").  Bits 0-4, held by every pass, take each layer's mixer in one pass
// of that layer (kQLow): with its cost when the instance has at most two
// windows, else in its first window-only pass, which has the registers and
// time to spare.  The reverse pass undoes them in the opposite order.
constexpr uint32_t kQInit = 1;     // forward: start from |+> (nothing read)
constexpr uint32_t kQPre = 2;      // RX(2 b_lp) on the window bits
constexpr uint32_t kQBnd = 4;      // cost(lc), then RX(2 b_lc) on the window bits
constexpr uint32_t kQNoStore = 8;  // reverse: the sweep's last pass, partials only
constexpr uint32_t kQLow = 16;     // bits 0-4 too (with lc if kQBnd, else with lp)
__host__ __device__ constexpr uint32_t qpass(int g, uint32_t flags, int lp, int lc) {
  return (uint32_t)g | flags << 4 | (uint32_t)lp << 12 | (uint32_t)lc << 20;
}

// Launch table entry: instance, its pass, and its first block index in this launch.
struct QEntry {
  int inst;
  uint32_t pass;
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

// The launch's entry (instance) whose block range holds workgroup b, that
// instance's index and pass.  With at most kBlock entries the table's first
// blocks are staged in LDS by one load per thread and searched there: one
// global round trip instead of a chain of dependent loads before the chunk's
// own loads can start.  Every thread of the workgroup must call it (it
// synchronises).
struct QFound {
  int entry, inst;
  uint32_t pass;
  uint64_t first;
};
__device__ __forceinline__ QFound find_entry(const QEntry *tab, int count, uint64_t b) {
  __shared__ uint64_t sfb[kBlock];
  __shared__ int sinst[kBlock];
  __shared__ uint32_t spass[kBlock];
  if (count > kBlock) {
    const int e = find_entry_global(tab, count, b);
    return {e, tab[e].inst, tab[e].pass, tab[e].first_block};
  }
  if ((int)threadIdx.x < count) {
    sfb[threadIdx.x] = tab[threadIdx.x].first_block;
    sinst[threadIdx.x] = tab[threadIdx.x].inst;
    spass[threadIdx.x] = tab[threadIdx.x].pass;
  }
  __syncthreads();
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sfb[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return {lo, sinst[lo], spass[lo], sfb[lo]};
}

// find_entry with its results moved to scalar registers: every lane holds the
// same values, and marking them uniform lets the compiler load the instance's
// descriptor and tables with scalar loads and keep chunk-wide terms in SGPRs.
__device__ __forceinline__ QFound find_entry_u(const QEntry *tab, int count, uint64_t b) {
  const QFound f = find_entry(tab, count, b);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)f.first);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(f.first >> 32));
  return {__builtin_amdgcn_readfirstlane(f.entry), __builtin_amdgcn_readfirstlane(f.inst),
          (uint32_t)__builtin_amdgcn_readfirstlane(f.pass), (uint64_t)hi << 32 | lo};
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
// gamma [nch], beta [nwin * nch] (one set per window).
__device__ __forceinline__ uint64_t slot_energy(const QInstDev &q) { return q.slot; }
__device__ __forceinline__ uint64_t slot_gamma(const QInstDev &q, int l) {
  return q.slot + (uint64_t)q.nchunks * (1 + (uint64_t)l * (1 + q.nwin));
}
__device__ __forceinline__ uint64_t slot_beta(const QInstDev &q, int l, int part) {
  return slot_gamma(q, l) + (uint64_t)q.nchunks * (1 + part);
}

// RX(2b) / cos b on the register-index bits set in `mask` (uniform), with
// t = tan b: per pair (x, y), x' = x - i t y and y' = y - i t x, one FMA per
// component; the inverse takes -t.  Bits outside the mask take t = 0, an
// exact identity: no branches, so the registers keep one assignment.
template <typename T, typename R>
__device__ __forceinline__ void qrx16t(T (&v)[16], R t0, int mask) {
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const R t = ((mask >> b) & 1) ? t0 : R(0), nt = -t;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      if (r & (1 << b)) continue;
      const T x = v[r], y = v[r | (1 << b)];
      v[r].x = fma(t, y.y, x.x);
      v[r].y = fma(nt, y.x, x.y);
      v[r | (1 << b)].x = fma(t, x.y, y.x);
      v[r | (1 << b)].y = fma(nt, x.x, y.y);
    }
  }
}

// c^k for a uniform k (the scale factors of qrx16t)
__device__ __forceinline__ double qpowc(double c, int k) {
  double r = 1.0;
  for (int i = 0; i < k; ++i) r *= c;
  return r;
}

// Im(conj(l) a) of one amplitude pair in the state's precision: for complex64
// one FMA after one product (|error| <= 2 ulp of |l||a|, the state's own
// rounding scale), so only per-pair results are widened to FP64 and the
// amplitudes stay in 32-bit registers.
template <typename T>
__device__ __forceinline__ auto qimcl(const T &l, const T &a) {
  return fma(l.x, a.y, -(l.y * a.x));
}

// sum over the register bits b set in `mask` (uniform) and pairs (r, r^b) of
// Im(conj(l_r) a_{r^b} + conj(l_{r^b}) a_r)  (= Im <l| X_b |a> over the pairs):
// the four products of a pair in one chain of FMAs in the state's precision
// (one more rounding than qimcl, at the same scale), the pairs accumulated in
// FP64
template <typename T>
__device__ __forceinline__ double qxpart16m(const T (&a)[16], const T (&l)[16], int mask) {
  double s = 0.0;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    if (!((mask >> b) & 1)) continue;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      if (r & (1 << b)) continue;
      const int u = r | (1 << b);
      const auto d = fma(l[u].x, a[r].y, fma(-l[u].y, a[r].x, qimcl(l[r], a[u])));
      s += (double)d;
    }
  }
  return s;
}

// Per (instance, layer): the cost-layer phases exp(-i g s_c), s_c = |E| - 2c for
// c = 0..|E|, then the mixer as (cos b', sin b') and (tan b', cos b'),
// evaluated in FP64 and rounded to the state's precision once per run by
// k_qaoa_phases.  b' is b reduced to [-pi/4, pi/4] by a multiple of pi/2: the
// energy is periodic in every b with period pi/2, since RX(pi) on every qubit
// is -i X on every qubit, which flips every bit, leaves every cut, the cost
// layers, the mixers and |+>^n unchanged, and only multiplies the state by
// a global phase.  With |tan b'| <= 1 the window passes apply RX(2b') / cos b'
// = [[1, -i t], [-i t, 1]] (one FMA per output component, growth at most
// sqrt 2 per bit) and fold the cos b' factors into their other products.
constexpr int kQPh = kQMaxEdges + 3;
constexpr int kQMixCS = kQPh - 2;  // (cos b', sin b')
constexpr int kQMixTC = kQPh - 1;  // (tan b', cos b')

template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_qaoa_phases(typename CV<R>::T *__restrict__ ph, const QInstDev *__restrict__ inst,
                  const double *__restrict__ gam, const double *__restrict__ bet, int p) {
  using T = typename CV<R>::T;
  const int ii = blockIdx.x / p, layer = blockIdx.x % p;
  const int ne = inst[ii].nedges;
  const double g = gam[(size_t)ii * p + layer];
  T *o = ph + (size_t)blockIdx.x * kQPh;
  for (int c = threadIdx.x; c <= ne; c += kBlock) {
    double sn, cs;
    sincos(-g * (double)(ne - 2 * c), &sn, &cs);
    o[c] = T{(R)cs, (R)sn};
  }
  if (threadIdx.x == 0) {
    const double b = bet[(size_t)ii * p + layer];
    const double br = b - M_PI_2 * nearbyint(b / M_PI_2);
    double sn, cs;
    sincos(br, &sn, &cs);
    o[kQMixCS] = T{(R)cs, (R)sn};
    o[kQMixTC] = T{(R)(sn / cs), (R)cs};
  }
}

// ---- window passes (n >= 12) -------------------------------------------------
// Geometry (qmid_pos, chunk_layout.hpp) with C = kQC = 5 contiguous low bits
// and W = min(7, n - h0) window bits h0 .. h0+W-1: loads and stores in layout
// L2 (local bits 8-11 in the register index), the cost and the window's local
// bits 5-7 in layout L1 (local bits 4-7), bits 0-3 in layout L0.
struct QMid {
  uint32_t base;  // global offset of the chunk (uniform)
  uint32_t t2;    // thread part of the L2 offset
  int W;
};

// Chunk of a window workgroup: its index within the instance's launch, or
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

// A copy of x that the compiler cannot identify with x: store addresses
// formed from it are recomputed (one VALU op each) instead of being kept live
// in registers since the loads.
__device__ __forceinline__ uint32_t qfresh(uint32_t x) {
  asm volatile("" : "+v"(x));
  return x;
}

// register part of the L2 offset: local bits 8-11 in r
__device__ __forceinline__ uint32_t qmid_r(int r, int W, int h0) {
  uint32_t o = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) o |= (uint32_t)((r >> k) & 1) << qmid_pos(8 + k, W, h0, kQC);
  return o;
}

// register-bit masks of the window's local bits kQC..kQC+W-1: L2 phase (local
// 8-11), L1 phase (local 4-7)
__device__ __forceinline__ int qmid_mask2(int W) {
  return ((((1 << W) - 1) << kQC) >> 8) & 15;
}
__device__ __forceinline__ int qmid_mask1(int W) { return ((((1 << W) - 1) << kQC) >> 4) & 15; }

// Window g of an instance: its first bit h0 and width W, and the cut terms of
// its workgroups (see qcut_win): m[k], the neighbours of local bit k's global
// bit outside the 12 local bits; nh[b], for a bit b outside them, its
// neighbours above b that are outside too (0 for local bits).
struct QWin {
  uint32_t m[kQL];
  uint32_t nh[32];
  int h0, W, pad0, pad1;
};

// Cut counts of a window workgroup in layout L1.  With i = base | (local bits
// placed by qmid_pos), split the edges by where their bits lie: both local
// ("lo", a function of the local index j), both outside ("hi", uniform), or
// one of each (local bit a, outside bit b: j_a xor base_b):
//   cut(i) = lo(j) + C + sum_a j_a w_a,
//   C   = hi(base) + sum_a popc(base & m_a),   w_a = popc(m_a) - 2 popc(base & m_a).
// lo(j) is a 4096-byte table per window stored [t][r] in layout L1, so a
// thread's 16 values are one 16-byte load; C and w_a depend on the workgroup
// only and are computed once per wave in scalar registers.
struct QCut1 {
  int K;                // C + sum over the thread's local bits (0-3, 8-11) of j_a w_a
  int w4, w5, w6, w7;   // w_a of the register bits (uniform)
  uint4 lo;             // lo(j(t, r)) in byte r & 3 of word r >> 2
};

// `c` and `base` must be wave-uniform.
__device__ __forceinline__ QCut1 qcut_win(const QWin &c, const uint4 *__restrict__ lo, int n,
                                          uint32_t base, int t) {
  QCut1 f;
  f.lo = lo[t];
  int C = 0, w[kQL];
#pragma unroll
  for (int a = 0; a < kQL; ++a) {
    const int pm = __popc(base & c.m[a]);
    C += pm;
    w[a] = __popc(c.m[a]) - 2 * pm;
  }
  for (int b = 0; b < n; ++b) C += __popc(c.nh[b] & (((base >> b) & 1) ? ~base : base));
  int K = C;
#pragma unroll
  for (int a = 0; a < 4; ++a) K += ((t >> a) & 1) * w[a] + ((t >> (4 + a)) & 1) * w[8 + a];
  f.K = K;
  f.w4 = w[4];
  f.w5 = w[5];
  f.w6 = w[6];
  f.w7 = w[7];
  return f;
}

__device__ __forceinline__ int qcut_r1(const QCut1 &f, int r) {
  const uint32_t word = r < 4 ? f.lo.x : r < 8 ? f.lo.y : r < 12 ? f.lo.z : f.lo.w;
  return f.K + (int)((word >> (8 * (r & 3))) & 0xffu) + (r & 1) * f.w4 + ((r >> 1) & 1) * f.w5 +
         ((r >> 2) & 1) * f.w6 + ((r >> 3) & 1) * f.w7;
}

// Fields of a window-pass descriptor (qpass).
struct QPassF {
  int g, fl, lp, lc;
};
__device__ __forceinline__ QPassF qpass_fields(uint32_t v) {
  return {(int)(v & 15), (int)((v >> 4) & 255), (int)((v >> 12) & 255), (int)((v >> 20) & 255)};
}

// Forward window pass: [RX(2 b_lp) on the window] [cost(lc), RX(2 b_lc) on the
// window and bits 0-4], on psi.
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_qaoa_win_fwd(typename CV<R>::T *__restrict__ st, const QInstDev *__restrict__ inst,
                   const QWin *__restrict__ wins, const uint4 *__restrict__ wlo,
                   const typename CV<R>::T *__restrict__ ph, int p,
                   const QEntry *__restrict__ tab, int count, int scatter) {
  using T = typename CV<R>::T;
  __shared__ T X[kQSlots];
  __shared__ R ph_r[kQMaxEdges + 1], ph_i[kQMaxEdges + 1];
  const QFound fe = find_entry_u(tab, count, blockIdx.x);
  const int ii = fe.inst;
  const QInstDev q = inst[ii];
  const QPassF P = qpass_fields(fe.pass);
  const int wi = q.win + P.g;
  const int h0 = wins[wi].h0;
  const int t = threadIdx.x;
  const uint32_t chunk = qmid_chunk(blockIdx.x - fe.first, q.n, scatter && P.g > 0);
  const QMid m = qmid_setup(chunk, q.n, h0, t);
  const int m2 = qmid_mask2(m.W), m1 = qmid_mask1(m.W);
  T *s = st + q.psi + m.base;
  // the chunk's loads go out first; the phase table and cut terms overlap them
  T v[16];
  if (P.fl & kQInit) {
    const R amp0 = (R)(1.0 / sqrt((double)(1ull << q.n)));
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      v[r].x = amp0;
      v[r].y = 0;
    }
  } else {
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = ld_nt(s + (m.t2 | qmid_r(r, m.W, h0)));
  }
  // mixers as RX / cos (qrx16t): the cos factors of the pass, c_lp^kp (pre)
  // and c_lc^kc (bnd) for kp, kc rotated bits, go into the cost phases, or
  // into one product at the end
  const bool pre = (P.fl & kQPre) != 0, bnd = (P.fl & kQBnd) != 0;
  const int low = (P.fl & kQLow) != 0;  // bits 0-4: L1 register bit 0, then L0
  const int lowp = low && !bnd, lowc = low && bnd;
  const int kp = pre ? m.W + kQC * lowp : 0, kc = m.W + kQC * lowc;
  T mp{0, 1}, mc{0, 1};  // (tan, cos) of layers lp and lc
  if (pre) mp = ph[((size_t)ii * p + P.lp) * kQPh + kQMixTC];
  QCut1 cc{};
  if (bnd) {
    const T *phl = ph + ((size_t)ii * p + P.lc) * kQPh;
    mc = phl[kQMixTC];
    if (t <= q.nedges) {
      const double sg = qpowc((double)mp.y, kp) * qpowc((double)mc.y, kc);
      const T z = phl[t];
      ph_r[t] = (R)(sg * (double)z.x);
      ph_i[t] = (R)(sg * (double)z.y);
    }
    cc = qcut_win(wins[wi], wlo + (size_t)wi * kBlock, q.n, m.base, t);
  }
  if (pre) {
    qrx16t(v, mp.x, m2);
    qtranspose(v, X, t, 2, 1);
    qrx16t(v, mp.x, m1 | lowp);
  } else if (!(P.fl & kQInit)) {
    qtranspose(v, X, t, 2, 1);
  }  // (|+> is the same in every layout)
  if (bnd) {
    if (P.fl & kQInit) __syncthreads();  // the phase table (no transpose since)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      R x = v[r].x, y = v[r].y;
      const int c = qcut_r1(cc, r);
      cmul(x, y, ph_r[c], ph_i[c]);
      v[r].x = x;
      v[r].y = y;
    }
    qrx16t(v, mc.x, m1 | lowc);
  }
  // (bits 0-3 if kQLow) and back to the store layout; each transpose writes
  // the slots this thread just read
  if (low) {
    qtranspose(v, X, t, 1, 0);
    qrx16t(v, bnd ? mc.x : mp.x, 15);
    qtranspose(v, X, t, 0, 2);
  } else {
    qtranspose(v, X, t, 1, 2);
  }
  if (bnd) {
    qrx16t(v, mc.x, m2);
  } else {
    const R f = (R)qpowc((double)mp.y, kp);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      v[r].x *= f;
      v[r].y *= f;
    }
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) st_nt(s + (m.t2 | qmid_r(r, m.W, h0)), v[r]);
}

// The turn: the forward sweep's last pass [RX(2 b_lp) on the window], then the
// energy partial and lambda = C psi, then the window undone on lambda with the
// beta partial of layer lp; writes lambda only (psi in memory is its value
// before the pass, which the reverse sweep continues from).  The window's bits
// 8-11 take their partial against psi with only those bits rotated (kept in
// registers): there psi and the half-undone lambda stand at the same point.
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_qaoa_win_turn(typename CV<R>::T *__restrict__ st, const QInstDev *__restrict__ inst,
                    const QWin *__restrict__ wins, const uint4 *__restrict__ wlo,
                    const typename CV<R>::T *__restrict__ ph, int p,
                    const QEntry *__restrict__ tab, int count, double *__restrict__ partial,
                    int scatter) {
  using T = typename CV<R>::T;
  __shared__ T X[kQSlots];
  const QFound fe = find_entry_u(tab, count, blockIdx.x);
  const int ii = fe.inst;
  const QInstDev q = inst[ii];
  const QPassF P = qpass_fields(fe.pass);
  const int wi = q.win + P.g;
  const int h0 = wins[wi].h0;
  const int t = threadIdx.x;
  const uint32_t chunk = qmid_chunk(blockIdx.x - fe.first, q.n, scatter && P.g > 0);
  const QMid m = qmid_setup(chunk, q.n, h0, t);
  const int m2 = qmid_mask2(m.W), m1 = qmid_mask1(m.W);
  const bool pre = (P.fl & kQPre) != 0;  // (never kQLow: see qsched_fwd)
  T v[16], b[16], l[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = ld_nt(st + q.psi + m.base + (m.t2 | qmid_r(r, m.W, h0)));
  const QCut1 cc = qcut_win(wins[wi], wlo + (size_t)wi * kBlock, q.n, m.base, t);
  // mixer as RX / cos (qrx16t): psi and lambda carry 1 / c^W from here on, and
  // lambda's undo adds another 1 / c^W, so the energy, the partials and the
  // stored lambda take c^(2W)
  T mp{0, 1};
  if (pre) {
    mp = ph[((size_t)ii * p + P.lp) * kQPh + kQMixTC];
    qrx16t(v, mp.x, m2);
  }
  const double f = qpowc((double)mp.y, 2 * m.W * pre);
#pragma unroll
  for (int r = 0; r < 16; ++r) b[r] = v[r];
  qtranspose(v, X, t, 2, 1);
  if (pre) qrx16t(v, mp.x, m1);
  double acce = 0.0, accb = 0.0;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int c = qcut_r1(cc, r);
    acce += (double)c * ((double)v[r].x * v[r].x + (double)v[r].y * v[r].y);
    l[r].x = (R)c * v[r].x;
    l[r].y = (R)c * v[r].y;
  }
  if (pre) {
    accb += qxpart16m(v, l, m1);
    qrx16t(l, -mp.x, m1);
  }
  qtranspose(l, X, t, 1, 2);  // writes the L1 slots this thread read for psi
  if (pre) {
    accb += qxpart16m(b, l, m2);
    qrx16t(l, -mp.x, m2);
    const R fr = (R)f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      l[r].x *= fr;
      l[r].y *= fr;
    }
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) st_nt(st + q.lam + m.base + (m.t2 | qmid_r(r, m.W, h0)), l[r]);
  acce = qblock_sum(acce * f);
  if (t == 0) partial[slot_energy(q) + chunk] = acce;
  if (pre) {
    accb = qblock_sum(accb * f);
    if (t == 0) partial[slot_beta(q, P.lp, P.g) + chunk] = accb;
  }
}

// Reverse window pass on psi and lambda: the forward pass undone in reverse
// order, each part's partial taken before it is undone (beta of lc over its
// bits, gamma of lc at the cost, beta of lp over its bits).  kQNoStore: the
// sweep's last pass; it stops at the gamma partial.  BND selects the passes
// with kQBnd (their own instantiation: the cost and the second mixer need
// far more registers than a window-only pass).  Complex64: 3 waves per SIMD
// (a few registers spill in the kQBnd instantiation; measured faster than 2
// waves without spills, the LDS transposes' barriers wanting more resident
// workgroups).
template <typename R, bool BND>
__global__ void __launch_bounds__(kBlock, sizeof(R) == 4 ? 3 : 1)
    k_qaoa_win_rev(typename CV<R>::T *__restrict__ st, const QInstDev *__restrict__ inst,
                   const QWin *__restrict__ wins, const uint4 *__restrict__ wlo,
                   const typename CV<R>::T *__restrict__ ph, int p,
                   const QEntry *__restrict__ tab, int count, double *__restrict__ partial,
                   int scatter) {
  using T = typename CV<R>::T;
  __shared__ T X[kQSlots];
  __shared__ R ph_r[kQMaxEdges + 1], ph_i[kQMaxEdges + 1];
  const QFound fe = find_entry_u(tab, count, blockIdx.x);
  const int ii = fe.inst;
  const QInstDev q = inst[ii];
  const QPassF P = qpass_fields(fe.pass);
  const int wi = q.win + P.g;
  const int h0 = wins[wi].h0;
  const int t = threadIdx.x;
  const uint32_t chunk = qmid_chunk(blockIdx.x - fe.first, q.n, scatter && P.g > 0);
  const QMid m = qmid_setup(chunk, q.n, h0, t);
  const int m2 = qmid_mask2(m.W), m1 = qmid_mask1(m.W);
  constexpr bool bnd = BND;
  const bool pre = (P.fl & kQPre) != 0;
  const bool store = !(P.fl & kQNoStore);
  T *sa = st + q.psi + m.base, *sl = st + q.lam + m.base;
  T a[16], l[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const uint32_t o = m.t2 | qmid_r(r, m.W, h0);
    a[r] = ld_nt(sa + o);
    l[r] = ld_nt(sl + o);
  }
  // Mixers undone as RX / cos (qrx16t): psi and lambda carry the same factor
  // u (uniform) at every point, and a partial taken there is divided by u^2.
  // The first mixer undone (lc's if bnd, else lp's) on k1 bits leaves
  // u = 1 / c^k1.  bnd: the cost undo multiplies by c_lc^k1 c_lp^W (folded
  // into its phases), so the pre part starts at u = c_lp^W and ends at 1.
  // Window only: one product by c_lp^k1 before the stores.
  const int low = (P.fl & kQLow) != 0;  // bits 0-4: L0, then L1 register bit 0
  const int w2 = __popc(m2), k1 = m.W + kQC * low;
  T mp{0, 1};  // (tan, cos) of layer lp
  if (pre) mp = ph[((size_t)ii * p + P.lp) * kQPh + kQMixTC];
  const double cp = mp.y;
  double accp = 0.0, accq = 0.0, accg = 0.0;  // beta of lc, beta of lp, gamma of lc
  double c1 = cp;
  R nt = -mp.x;  // RX(-2b) / cos of the first mixer undone
  if (bnd) {
    const T *phl = ph + ((size_t)ii * p + P.lc) * kQPh;
    const T mc = phl[kQMixTC];
    c1 = mc.y;
    nt = -mc.x;
    if (t <= q.nedges) {  // undo: exp(+i g s), the conjugate phases, times the factor
      const double sg = qpowc(c1, k1) * qpowc(cp, m.W * pre);
      const T z = phl[t];
      ph_r[t] = (R)(sg * (double)z.x);
      ph_i[t] = (R)(-sg * (double)z.y);
    }
  }
  double acc1 = qxpart16m(a, l, m2);
  qrx16t(a, nt, m2);
  qrx16t(l, nt, m2);
  if (low) {
    qtranspose(a, X, t, 2, 0);
    __syncthreads();
    qtranspose(l, X, t, 2, 0);
    acc1 += qxpart16m(a, l, 15) * qpowc(c1, 2 * w2);  // bits 0-3
    qrx16t(a, nt, 15);
    qrx16t(l, nt, 15);
    qtranspose(a, X, t, 0, 1);  // writes the L0 slots this thread just read
    __syncthreads();
    qtranspose(l, X, t, 0, 1);
  } else {
    qtranspose(a, X, t, 2, 1);
    __syncthreads();
    qtranspose(l, X, t, 2, 1);
  }
  acc1 += qxpart16m(a, l, m1 | low) * qpowc(c1, 2 * (w2 + 4 * low));
  qrx16t(a, nt, m1 | low);
  qrx16t(l, nt, m1 | low);
  if (bnd) {
    accp = acc1;
    // gamma partial Im(conj(lambda) s psi) and undo the cost layer
    const QCut1 cc = qcut_win(wins[wi], wlo + (size_t)wi * kBlock, q.n, m.base, t);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int c = qcut_r1(cc, r);
      const double sv = (double)(q.nedges - 2 * c);
      R x = a[r].x, y = a[r].y, u = l[r].x, w = l[r].y;
      accg += sv * ((double)u * y - (double)w * x);
      if (store) {
        cmul(x, y, ph_r[c], ph_i[c]);
        cmul(u, w, ph_r[c], ph_i[c]);
        a[r].x = x;
        a[r].y = y;
        l[r].x = u;
        l[r].y = w;
      }
    }
    accg *= qpowc(c1, 2 * k1);
    if (pre) {
      accq += qxpart16m(a, l, m1) / qpowc(cp, 2 * m.W);
      qrx16t(a, -mp.x, m1);
      qrx16t(l, -mp.x, m1);
    }
  } else {
    accq = acc1;
  }
  if (store) {
    qtranspose(a, X, t, 1, 2);  // writes the L1 slots this thread just read
    __syncthreads();
    qtranspose(l, X, t, 1, 2);
    if (bnd && pre) {
      accq += qxpart16m(a, l, m2) / qpowc(cp, 2 * w2);
      qrx16t(a, -mp.x, m2);
      qrx16t(l, -mp.x, m2);
    }
    if (!bnd) {
      const R f = (R)qpowc(cp, k1);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        a[r].x *= f;
        a[r].y *= f;
        l[r].x *= f;
        l[r].y *= f;
      }
    }
    const uint32_t t2 = qfresh(m.t2);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const uint32_t o = t2 | qmid_r(r, m.W, h0);
      st_nt(sa + o, a[r]);
      st_nt(sl + o, l[r]);
    }
  }
  if (bnd) {
    accp = qblock_sum(accp);
    if (t == 0) partial[slot_beta(q, P.lc, P.g) + chunk] = accp;
    accg = qblock_sum(accg);
    if (t == 0) partial[slot_gamma(q, P.lc) + chunk] = accg;
  }
  if (pre) {
    accq = qblock_sum(accq);
    if (t == 0) partial[slot_beta(q, P.lp, P.g) + chunk] = accq;
  }
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
  const T *phl = ph + ((size_t)ii * p + layer) * kQPh;
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
  const T cb = phl[kQMixCS];
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

// ---- energy, instances below 12 qubits: lambda = C psi, partial sum of
// cut(i) |psi_i|^2 (larger instances form both in the turn) -------------------
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_qaoa_energy(typename CV<R>::T *__restrict__ st, const QInstDev *__restrict__ inst,
                  const uint8_t *__restrict__ edges, const QEntry *__restrict__ tab, int count,
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
  for (int j = threadIdx.x; j < m; j += kBlock) {
    const T v = ld_nt(st + q.psi + base + j);
    const int c = cut_of(base + j, e, q.nedges);
    acc += (double)c * ((double)v.x * v.x + (double)v.y * v.y);
    T o;
    o.x = (R)c * v.x;
    o.y = (R)c * v.y;
    st_nt(st + q.lam + base + j, o);
  }
  acc = qblock_sum(acc);
  if (threadIdx.x == 0) partial[slot_energy(q) + chunk] = acc;
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
  const T *phl = ph + ((size_t)ii * p + layer) * kQPh;
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
  const T cb = phl[kQMixCS];
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
      len = nch * (uint64_t)q.nwin;
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
  QWin *wins = nullptr;      // windows of the instances of n >= 12
  uint8_t *wlo = nullptr;    // per window, 4096 bytes (qcut_win)
  void *phase = nullptr;     // per (instance, layer), kQPh amplitudes
  double *params = nullptr;  // gam [ninst*p], bet [ninst*p]
  double *partial = nullptr;
  double *out = nullptr;
  QEntry *tables = nullptr;
  struct Launch {
    // kind: n < 12, per layer: 0 low_fwd, 2 energy, 4 low_rev;
    //       n >= 12: 1 win_fwd, 3 win_turn, 6 win_rev of kQBnd passes, 7 of the others
    int kind, layer;
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
  for (void *ptr : {(void *)b->states, (void *)b->inst, (void *)b->edges, (void *)b->wins,
                    (void *)b->wlo, (void *)b->phase, (void *)b->params, (void *)b->partial,
                    (void *)b->out, (void *)b->tables})
    if (ptr) (void)hipFree(ptr);
}

// Window-pass schedule of an instance of n >= 12 with k windows: forward
// passes s = 0 .. F-1 (F = p (k - 1), or p for k = 1), then the turn; the
// reverse sweep undoes s = F-1 .. 0.  Pass s takes window s mod k, so the k
// passes of a layer (from the one applying its cost to the one applying the
// next cost) see every window once; for k = 1 every pass is a whole layer.
static int qsched_nfwd(int k, int p) { return k == 1 ? p : p * (k - 1); }
static uint32_t qsched_fwd(int k, int s) {
  if (k == 1) return qpass(0, (s == 0 ? kQInit : 0) | kQBnd | kQLow, 0, s);
  const int g = s % k, per = k - 1;
  const uint32_t lb = k <= 2 ? kQLow : 0;  // bits 0-4 with the cost, or (k >= 3)
  if (s == 0) return qpass(g, kQInit | kQBnd | lb, 0, 0);  // in the first window-only pass
  if (s % per == 0) return qpass(g, kQPre | kQBnd | lb, s / per - 1, s / per);
  return qpass(g, kQPre | (s % per == 1 ? kQLow : 0), s / per, 0);
}
static uint32_t qsched_turn(int k, int p) {
  return k == 1 ? qpass(0, 0, 0, 0) : qpass((p * (k - 1)) % k, kQPre, p - 1, 0);
}
static uint32_t qsched_rev(int k, int s) {
  return qsched_fwd(k, s) | (s == 0 ? kQNoStore << 4 : 0u);
}

int qaoa_create(void **handle, sq_dtype_t dt, uint32_t ninst, const uint32_t *nq,
                const uint32_t *nedges, const uint32_t *edges, uint32_t p) {
  if (ninst == 0 || p == 0) return fail(SQ_ERR_INVALID, "qaoa: empty batch");
  if (p > (uint32_t)kQMaxLayers) return fail(SQ_ERR_INVALID, "qaoa: at most 256 layers");
  if ((uint64_t)ninst * p > (uint64_t)INT32_MAX)  // one phase-table workgroup each
    return fail(SQ_ERR_INVALID, "qaoa: too many (instance, layer) pairs");
  auto B = std::make_unique<QaoaBatch>();
  B->dt = dt;
  B->ninst = (int)ninst;
  B->p = (int)p;
  B->maxn = 0;
  const size_t item = itemsize(dt);
  uint64_t amp = 0, slot = 0;
  size_t eoff = 0;
  std::vector<uint8_t> hedges;
  std::vector<QWin> hwins;
  std::vector<uint8_t> hlo;
  for (uint32_t j = 0; j < ninst; ++j) {
    const int n = (int)nq[j];
    if (n < 2 || n > 30) return fail(SQ_ERR_INVALID, "qaoa: qubit counts must be in [2, 30]");
    if (nedges[j] > (uint32_t)kQMaxEdges) return fail(SQ_ERR_INVALID, "qaoa: too many edges");
    QInstDev q{};
    q.n = n;
    q.lc = std::min(n, kQL);
    q.nchunks = 1 << (n - q.lc);
    q.nwin = n >= kQL ? 1 + (n - kQL + kQGM - 1) / kQGM : 1;
    q.win = (int)hwins.size();
    q.nedges = (int)nedges[j];
    q.edge_off = (int)(hedges.size() / 2);
    std::vector<uint32_t> seen(n, 0);  // seen[a]: bit b set if edge {a, b} was given
    std::vector<std::pair<int, int>> eb;  // edges as bit positions, lower first
    for (uint32_t k = 0; k < nedges[j]; ++k) {
      const uint32_t a = edges[2 * (eoff + k)], b = edges[2 * (eoff + k) + 1];
      if (a >= (uint32_t)n || b >= (uint32_t)n || a == b)
        return fail(SQ_ERR_INVALID, "qaoa: bad edge");
      // simple graphs only: the cut tables hold neighbour sets as bit masks
      if ((seen[a] >> b) & 1) return fail(SQ_ERR_INVALID, "qaoa: repeated edge");
      seen[a] |= 1u << b;
      seen[b] |= 1u << a;
      const int u = n - 1 - (int)a, v = n - 1 - (int)b;  // qubit -> bit position
      hedges.push_back((uint8_t)u);
      hedges.push_back((uint8_t)v);
      eb.emplace_back(std::min(u, v), std::max(u, v));
    }
    for (int g = 0; n >= kQL && g < q.nwin; ++g) {  // window tables (see QWin, qcut_win)
      QWin w{};
      w.h0 = g == 0 ? kQC : kQL + kQGM * (g - 1);
      w.W = std::min(kQGM, n - w.h0);
      int loc[32];
      std::fill(loc, loc + 32, -1);
      for (int k = 0; k < kQL; ++k) loc[qmid_pos(k, w.W, w.h0, kQC)] = k;
      std::vector<std::pair<int, int>> lo_edges;
      for (auto &e : eb) {
        const int x = e.first, y = e.second;
        if (loc[x] >= 0 && loc[y] >= 0) lo_edges.emplace_back(loc[x], loc[y]);
        else if (loc[x] >= 0) w.m[loc[x]] |= 1u << y;
        else if (loc[y] >= 0) w.m[loc[y]] |= 1u << x;
        else w.nh[x] |= 1u << y;
      }
      hwins.push_back(w);
      const size_t off = hlo.size();
      hlo.resize(off + kQChunk);
      for (int t = 0; t < 256; ++t)
        for (int r = 0; r < 16; ++r) {
          const int jj = (t & 15) | (r << 4) | ((t >> 4) << 8);  // layout L1
          int c = 0;
          for (auto &e : lo_edges) c += ((jj >> e.first) ^ (jj >> e.second)) & 1;
          hlo[off + (size_t)t * 16 + r] = (uint8_t)c;
        }
    }
    eoff += nedges[j];
    q.psi = amp;
    q.lam = amp + (1ull << n);
    amp += 2ull << n;
    q.slot = slot;
    slot += (uint64_t)q.nchunks * (1 + (uint64_t)p * (1 + q.nwin));
    B->host_inst.push_back(q);
    B->maxn = std::max(B->maxn, n);
    B->state_bytes += 2.0 * std::ldexp(1.0, n) * item;
  }
  if (hedges.empty()) hedges.push_back(0);
  if (hwins.empty()) {
    hwins.emplace_back();
    hlo.resize(kQChunk);
  }
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
  if (!rc) rc = dmalloc((void **)&B->wins, hwins.size() * sizeof(QWin));
  if (!rc) rc = dmalloc((void **)&B->wlo, hlo.size());
  if (!rc) rc = dmalloc(&B->phase, (size_t)ninst * p * kQPh * item);
  if (!rc) rc = dmalloc((void **)&B->params, 2 * ninst * p * sizeof(double));
  if (!rc) rc = dmalloc((void **)&B->partial, std::max<uint64_t>(slot, 1) * sizeof(double));
  if (!rc) rc = dmalloc((void **)&B->out, ninst * (1 + 2 * p) * sizeof(double));
  // launch sequence and tables: one launch per pass index, each instance's
  // entry carrying its own pass
  std::vector<QEntry> tabs;
  auto add = [&](int kind, int layer, auto pass_of) {  // pass_of(j) < 0: not in this launch
    QaoaBatch::Launch L{kind, layer, tabs.size(), 0, 0};
    for (int j = 0; j < (int)ninst; ++j) {
      const int64_t ps = pass_of(j);
      if (ps < 0) continue;
      tabs.push_back(QEntry{j, (uint32_t)ps, L.blocks});
      L.blocks += (uint64_t)B->host_inst[j].nchunks;
      L.count++;
    }
    if (L.count) B->launches.push_back(L);
  };
  const auto &hi = B->host_inst;
  const int P = (int)p;
  auto small = [&](int j) -> int64_t { return hi[j].n < kQL ? 0 : -1; };
  for (int l = 0; l < P; ++l) add(0, l, small);
  add(2, 0, small);
  for (int l = P - 1; l >= 0; --l) add(4, l, small);
  int maxf = 0;
  for (auto &q : hi)
    if (q.n >= kQL) maxf = std::max(maxf, qsched_nfwd(q.nwin, P));
  for (int i = 0; i < maxf; ++i)
    add(1, 0, [&](int j) -> int64_t {
      const QInstDev &q = hi[j];
      return q.n < kQL || i >= qsched_nfwd(q.nwin, P) ? -1 : (int64_t)qsched_fwd(q.nwin, i);
    });
  add(3, 0, [&](int j) -> int64_t {
    return hi[j].n < kQL ? -1 : (int64_t)qsched_turn(hi[j].nwin, P);
  });
  for (int i = 0; i < maxf; ++i)
    for (int bnd = 1; bnd >= 0; --bnd)  // the two reverse kernels (k_qaoa_win_rev<R, BND>)
      add(bnd ? 6 : 7, 0, [&](int j) -> int64_t {
        const QInstDev &q = hi[j];
        const int F = qsched_nfwd(q.nwin, P);
        if (q.n < kQL || i >= F) return -1;
        const uint32_t ps = qsched_rev(q.nwin, F - 1 - i);
        return (((ps >> 4) & kQBnd) != 0) == (bnd != 0) ? (int64_t)ps : -1;
      });
  if (!rc) rc = dmalloc((void **)&B->tables, std::max<size_t>(tabs.size(), 1) * sizeof(QEntry));
  if (rc) {
    free_batch(B.get());
    return rc;
  }
  if (hipMemcpy(B->inst, B->host_inst.data(), ninst * sizeof(QInstDev), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(B->edges, hedges.data(), hedges.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(B->wins, hwins.data(), hwins.size() * sizeof(QWin), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(B->wlo, hlo.data(), hlo.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(B->tables, tabs.data(), tabs.size() * sizeof(QEntry), hipMemcpyHostToDevice) !=
          hipSuccess) {
    free_batch(B.get());
    return fail(SQ_ERR_HIP, "qaoa: upload failed");
  }
  // HBM traffic of one run (reads + writes of the states, in state-sized units)
  for (auto &L : B->launches) {
    for (int k = 0; k < L.count; ++k) {
      const QEntry &e = tabs[L.tab_off + k];
      const uint32_t fl = (e.pass >> 4) & 255;
      double per = 0;
      switch (L.kind) {
        case 0: per = L.layer == 0 ? 1.0 : 2.0; break;
        case 2: per = 3.0; break;
        case 4: per = 4.0; break;
        case 1: per = (fl & kQInit) ? 1.0 : 2.0; break;
        case 3: per = 2.0; break;
        case 6:
        case 7: per = (fl & kQNoStore) ? 2.0 : 4.0; break;
      }
      B->traffic_bytes += per * std::ldexp(1.0, hi[e.inst].n) * item;
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
  const uint4 *wlo = (const uint4 *)B->wlo;
  const int p = B->p;
  const int sc = tuning().qaoa_mid_scatter;
  const double *gam = B->params, *bet = B->params + (size_t)B->ninst * p;
  hipLaunchKernelGGL(k_qaoa_phases<R>, dim3(B->ninst * p), dim3(kBlock), 0, st, (T *)B->phase,
                     B->inst, gam, bet, p);
  for (auto &L : B->launches) {
    const QEntry *tab = B->tables + L.tab_off;
    const dim3 g((unsigned)L.blocks), bl(kBlock);
    switch (L.kind) {
      case 0:
        hipLaunchKernelGGL(k_qaoa_low_fwd<R>, g, bl, 0, st, s, B->inst, B->edges, ph, p, tab,
                           L.count, L.layer);
        break;
      case 2:
        hipLaunchKernelGGL(k_qaoa_energy<R>, g, bl, 0, st, s, B->inst, B->edges, tab, L.count,
                           B->partial);
        break;
      case 4:
        hipLaunchKernelGGL(k_qaoa_low_rev<R>, g, bl, 0, st, s, B->inst, B->edges, ph, p, tab,
                           L.count, L.layer, B->partial);
        break;
      case 1:
        hipLaunchKernelGGL(k_qaoa_win_fwd<R>, g, bl, 0, st, s, B->inst, B->wins, wlo, ph, p, tab,
                           L.count, sc);
        break;
      case 3:
        hipLaunchKernelGGL(k_qaoa_win_turn<R>, g, bl, 0, st, s, B->inst, B->wins, wlo, ph, p,
                           tab, L.count, B->partial, sc);
        break;
      case 6:
        hipLaunchKernelGGL((k_qaoa_win_rev<R, true>), g, bl, 0, st, s, B->inst, B->wins, wlo, ph,
                           p, tab, L.count, B->partial, sc);
        break;
      case 7:
        hipLaunchKernelGGL((k_qaoa_win_rev<R, false>), g, bl, 0, st, s, B->inst, B->wins, wlo, ph,
                           p, tab, L.count, B->partial, sc);
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
