// kernels_layer.hip -- adjoint layer passes: a layer of one-target gates on
// distinct qubits, applied to psi (and lambda) in one pass per 12 bits.
//
// The reference has no gradient code (it simulates circuits and reads
// results); these kernels serve sq_adjoint_gradient's forward and reverse
// sweeps.  A workgroup holds 4096 amplitudes of psi (and of lambda) in
// registers, 16 per thread, whose local index j maps to the global index
// with C = 4 or 6 contiguous low bits: local bits 0..C-1 -> global bits
// 0..C-1, local bits C..C+W-1 -> a window h0..h0+W-1 of higher bits, local
// bits C+W..11 -> global bits C..11-W (W = 0: the identity, the "low" pass
// over bits 0-11).  C = 4 (windows of up to 8 bits) loads 128 B runs (16
// complex64 amplitudes; the batched QAOA window passes use C = 5); C = 6
// (windows of up to 6 bits) loads and stores 512 B runs.  The 16 registers hold
// local bits 8-11 (layout L2, the load layout), 4-7 (L1) or 0-3 (L0); a gate
// whose local bit is in the register index is applied in registers, and the
// layouts change through LDS.  So one pass applies up to 12 gates (all of
// them on bits 0-11, or up to 8 on high bits) where the register-group passes
// apply 4-5.
//
// Reverse (LAM): the gates commute, so every gate's derivative term
//   Re <lam| dG G^dagger |psi>
// may be taken at any point inside the layer; each is taken just before the
// gate is undone on both states (psi <- G^dagger psi, lam <- G^dagger lam).
// Forward: psi <- G psi.  Gate arithmetic in the state's precision, the
// derivative sums in FP64, reduced over a fixed grid in a fixed order.
#include "chunk_layout.hpp"
#include "kernels.hpp"
#include "kernels_reduce.hpp"

namespace sq {

// Per local position k (0..11) of the pass: the gate's matrix (G^dagger in
// the reverse sweep, G forward), its dG G^dagger and the partial row of its
// derivative (-1: none).
template <typename R>
struct LayerArgs {
  Mat2<R> g[12];
  Mat2<R> d[12];
  int slot[12];
  int has;  // positions that hold a gate
};

// Global offsets of a chunk's amplitudes (see the file comment): base is the
// chunk's part (elements), t2 / t1 the thread's part in layouts L2 / L1;
// layer_r gives the register's part (lo = 8: L2, lo = 4: L1).  Within a
// chunk the parts are byte offsets of type O: 32-bit when every offset of the
// pass fits (the loads then take the uniform chunk pointer plus one 32-bit
// offset, shared by psi and lambda), else 64-bit.
template <typename O>
struct LayerGeom {
  uint64_t base;
  O t2, t1;
};

__device__ __forceinline__ int layer_pos(int k, int W, int h0, int C) {
  return k < C ? k : k < C + W ? h0 + (k - C) : k - W;
}

template <typename O>
__device__ __forceinline__ O layer_r(int r, int lo, int W, int h0, int C, int esh) {
  O o = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) o |= (O)((r >> k) & 1) << (layer_pos(lo + k, W, h0, C) + esh);
  return o;
}

template <typename O>
__device__ __forceinline__ LayerGeom<O> layer_geom(uint64_t chunk, int W, int h0, int C, int t,
                                                   int esh) {
  LayerGeom<O> m;
  const int A = h0 - 12 + W;  // chunk bits at 12-W .. h0-1, the rest from h0+W up
  m.base = ((chunk & ((1ull << A) - 1)) << (12 - W)) | ((chunk >> A) << (h0 + W));
  m.t2 = (O)(t & 15) << esh;
  m.t1 = (O)(t & 15) << esh;
#pragma unroll
  for (int k = 4; k < 8; ++k) m.t2 |= (O)((t >> k) & 1) << (layer_pos(k, W, h0, C) + esh);
#pragma unroll
  for (int k = 0; k < 4; ++k)
    m.t1 |= (O)((t >> (4 + k)) & 1) << (layer_pos(8 + k, W, h0, C) + esh);
  return m;
}

// Derivative term of one amplitude pair, Re(conj(l0) d0 + conj(l1) d1):
// complex64 forms the pair's sum in FP32 and adds it to the FP64 total (one
// rounding per pair, the order of the state's own error), complex128 in FP64.
__device__ __forceinline__ double pair_term(const float2 &l0, const float2 &d0, const float2 &l1,
                                            const float2 &d1) {
  return (double)(l0.x * d0.x + l0.y * d0.y + l1.x * d1.x + l1.y * d1.y);
}
__device__ __forceinline__ double pair_term(const double2 &l0, const double2 &d0,
                                            const double2 &l1, const double2 &d1) {
  return l0.x * d0.x + l0.y * d0.y + l1.x * d1.x + l1.y * d1.y;
}

// y = m x for a real matrix m (its imaginary parts are zero): half the
// multiplications of mat2_apply.
template <typename T, typename R>
__device__ __forceinline__ void mat2_apply_real(const Mat2<R> &m, const T &x0, const T &x1,
                                                T &y0, T &y1) {
  y0.x = m.r[0] * x0.x + m.r[1] * x1.x;
  y0.y = m.r[0] * x0.y + m.r[1] * x1.y;
  y1.x = m.r[2] * x0.x + m.r[3] * x1.x;
  y1.y = m.r[2] * x0.y + m.r[3] * x1.y;
}

template <bool REAL, typename T, typename R>
__device__ __forceinline__ void mat2_apply_k(const Mat2<R> &m, const T &x0, const T &x1, T &y0,
                                             T &y1) {
  if (REAL)
    mat2_apply_real(m, x0, x1, y0, y1);
  else
    mat2_apply(m, x0, x1, y0, y1);
}

// The gates at local positions k0..k0+3 (register bits 0-3 of the current
// layout), matrices read from LDS.  Every position is applied, without
// branches: an empty position holds the identity (and a zero derivative
// matrix), which keeps the 16 registers of each state in place instead of
// duplicating them across per-gate branches.  Reverse: derivative terms,
// then G^dagger on both states.  REAL: every matrix of the pass is real
// (rotations about Y, Hadamards, their derivatives).
template <typename T, typename R, bool LAM, bool REAL>
__device__ __forceinline__ void layer_phase(T (&a)[16], T (&l)[16], const Mat2<R> *sg,
                                            const Mat2<R> *sd, int k0, double (&acc)[12]) {
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int k = k0 + b;
    const int bb = 1 << b;
    if (LAM) {
      const Mat2<R> d = sd[k];
      double s = 0.0;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        if (r & bb) continue;
        T d0, d1;
        mat2_apply_k<REAL>(d, a[r], a[r | bb], d0, d1);
        s += pair_term(l[r], d0, l[r | bb], d1);
      }
      acc[k] += s;
    }
    const Mat2<R> g = sg[k];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      if (r & bb) continue;
      T y0, y1;
      mat2_apply_k<REAL>(g, a[r], a[r | bb], y0, y1);
      a[r] = y0;
      a[r | bb] = y1;
      if (LAM) {
        mat2_apply_k<REAL>(g, l[r], l[r | bb], y0, y1);
        l[r] = y0;
        l[r | bb] = y1;
      }
    }
  }
}

// Both states from layout F to layout TO through X.  A thread's slots of one
// layout are the same for writing and reading, so writes in layout F need no
// barrier after reads in layout F (FRONT = false: the last reads of X were in
// layout F); lambda's writes wait until psi's reads in layout TO are done.
template <int F, int TO, bool FRONT, typename T, bool LAM>
__device__ __forceinline__ void layer_xpose(T (&a)[16], T (&l)[16], T *X, int t) {
  if (FRONT) __syncthreads();
  qtranspose(a, X, t, F, TO);
  if (LAM) {
    __syncthreads();
    qtranspose(l, X, t, F, TO);
  }
}

// Multiply the thread's 16 amplitudes (layout L2 of a low pass: index
// chunk << 12 | t | r << 8) by a fused diagonal run's phases: the tables'
// products were formed in FP64 and rounded to the state's precision once.
// Per chunk, the chunk's entry and the mixed gates whose local bit lies in
// the thread part (bits 0-7) make one factor per thread; a mixed gate on a
// register bit (8-11) leaves a pair of entries, selected per register.
template <typename T, typename R, bool LAM>
__device__ __forceinline__ void fuse_phases(T (&a)[16], T (&l)[16], const FuseDesc *fz,
                                            uint64_t chunk, int t) {
  const T *tl = reinterpret_cast<const T *>(fz->tl);
  const T cf = reinterpret_cast<const T *>(fz->tc)[chunk];
  const int nmix = fz->nmix;
  R sr = cf.x, si = cf.y;
  R f0r[kFuseMix], f0i[kFuseMix], f1r[kFuseMix], f1i[kFuseMix];
  int rb[kFuseMix];
#pragma unroll
  for (int k = 0; k < kFuseMix; ++k) {
    rb[k] = -1;
    f0r[k] = f1r[k] = 1;
    f0i[k] = f1i[k] = 0;
    if (k >= nmix) continue;
    const DiagRunGate &g = fz->mix[k];
    // loc = (bit b0) << 1 | (bit b1); one of b0, b1 is below 12
    const bool l0 = g.b0 < 12;
    const int lb = l0 ? g.b0 : g.b1, cb = l0 ? g.b1 : g.b0;
    const int cv = (int)((chunk >> (cb - 12)) & 1);
    const int loc0 = l0 ? cv : cv << 1, loc1 = l0 ? (2 | cv) : ((cv << 1) | 1);
    if (lb < 8) {  // thread part: one factor per thread
      const int loc = ((t >> lb) & 1) ? loc1 : loc0;
      const R gr = (R)g.ginv[loc][0], gi = (R)g.ginv[loc][1];
      const R tr = sr * gr - si * gi;
      si = sr * gi + si * gr;
      sr = tr;
    } else {
      rb[k] = lb - 8;
      f0r[k] = (R)g.ginv[loc0][0];
      f0i[k] = (R)g.ginv[loc0][1];
      f1r[k] = (R)g.ginv[loc1][0];
      f1i[k] = (R)g.ginv[loc1][1];
    }
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const T lf = tl[t + 256 * r];
    R pr = sr * lf.x - si * lf.y, pi = sr * lf.y + si * lf.x;
#pragma unroll
    for (int k = 0; k < kFuseMix; ++k) {
      if (rb[k] < 0) continue;
      const bool one = (r >> rb[k]) & 1;
      const R gr = one ? f1r[k] : f0r[k], gi = one ? f1i[k] : f0i[k];
      const R tr = pr * gr - pi * gi;
      pi = pr * gi + pi * gr;
      pr = tr;
    }
    T o;
    o.x = a[r].x * pr - a[r].y * pi;
    o.y = a[r].x * pi + a[r].y * pr;
    a[r] = o;
    if (LAM) {
      o.x = l[r].x * pr - l[r].y * pi;
      o.y = l[r].x * pi + l[r].y * pr;
      l[r] = o;
    }
  }
}

// PH: the phases the pass runs, bit 2: local positions 8-11 (layout L2),
// bit 1: 4-7 (L1), bit 0: 0-3 (L0).  Layout path: L2 (load) -> [L1] ->
// [L0 -> L2]; stores go out in L1 when that is the last layout, else in L2.
// O: offset type (see LayerGeom).  REAL: see layer_phase.  complex128 asks
// for 2 waves per SIMD (otherwise 1: 256 VGPRs and AGPR spills; 5% faster
// measured on the 24-qubit ansatz gradient); complex64 keeps the compiler's
// 3 (4 would spill to scratch and measured no faster).
// WIDE (C = 6): the L1 phase transposes back to L2 before the stores, so
// they go out as 512 B runs too.  FUSE (low pass, PH = 7): a diagonal run's
// phases (fz) before the layer's gates (reverse) or after them (forward).
template <typename T, typename R, typename O, bool LAM, int PH, bool REAL, bool WIDE, bool FUSE>
__global__ void __launch_bounds__(kBlock, sizeof(R) == 4 ? (FUSE && LAM ? 3 : 1) : 2)
    k_adjoint_layer(T *__restrict__ psi, T *__restrict__ lam, uint64_t nchunks, int W, int h0,
                    int C, LayerArgs<R> A, double *__restrict__ partial,
                    const FuseDesc *__restrict__ fz) {
  static_assert(!FUSE || PH == 7, "a fused run rides on the low pass");
  constexpr bool P2 = PH & 4, P1 = PH & 2, P0 = PH & 1;
  constexpr int NW = kBlock / 64;
  constexpr int esh = sizeof(T) == 8 ? 3 : 4;
  __shared__ T X[kQSlots];
  __shared__ Mat2<R> sg[12], sd[12];
  __shared__ double sred[12][NW];
  const int t = threadIdx.x;
  const int has = A.has;
  int dmask = 0;
#pragma unroll
  for (int k = 0; k < 12; ++k) dmask |= (A.slot[k] >= 0 ? 1 : 0) << k;
  dmask &= has;
  if (t < 24) {  // matrices to LDS: identity / zero at empty positions
    const int k = t % 12;
    const bool dv = t >= 12;
    Mat2<R> m{};
    if (dv ? ((dmask >> k) & 1) : ((has >> k) & 1)) {
      m = dv ? A.d[k] : A.g[k];
    } else if (!dv) {
      m.r[0] = 1;
      m.r[3] = 1;
    }
    (dv ? sd : sg)[k] = m;
  }
  __syncthreads();
  double acc[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = 0.0;
  for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const LayerGeom<O> m = layer_geom<O>(c, W, h0, C, t, esh);
    const ChunkIO<T, O> ia(psi + m.base), il(LAM ? lam + m.base : psi);
    T a[16], l[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const O so = layer_r<O>(r, 8, W, h0, C, esh);
      a[r] = ia.load(m.t2, so);
      if (LAM) l[r] = il.load(m.t2, so);
    }
    if (FUSE && LAM) fuse_phases<T, R, true>(a, l, fz, c, t);
    if (P2) layer_phase<T, R, LAM, REAL>(a, l, sg, sd, 8, acc);
    // the previous chunk's last reads of X: layout L1 (P1 only, not WIDE)
    // or L2 (P0, or WIDE: transposed back)
    constexpr bool FRONT = P1 && !P0 && !WIDE;
    if (P1) {
      layer_xpose<2, 1, FRONT, T, LAM>(a, l, X, t);
      layer_phase<T, R, LAM, REAL>(a, l, sg, sd, 4, acc);
    }
    if (P0) {
      if (P1)
        layer_xpose<1, 0, false, T, LAM>(a, l, X, t);
      else
        layer_xpose<2, 0, false, T, LAM>(a, l, X, t);
      layer_phase<T, R, LAM, REAL>(a, l, sg, sd, 0, acc);
      layer_xpose<0, 2, false, T, LAM>(a, l, X, t);
    }
    if (WIDE && P1 && !P0) layer_xpose<1, 2, false, T, LAM>(a, l, X, t);
    if (FUSE && !LAM) fuse_phases<T, R, false>(a, l, fz, c, t);
    constexpr bool L1OUT = P1 && !P0 && !WIDE;
    const O vo = L1OUT ? m.t1 : m.t2;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const O so = L1OUT ? layer_r<O>(r, 4, W, h0, C, esh) : layer_r<O>(r, 8, W, h0, C, esh);
      ia.store(vo, so, a[r]);
      if (LAM) il.store(vo, so, l[r]);
    }
  }
  if (!LAM) return;
  const int lane = t & 63, wave = t >> 6;
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    if (!((dmask >> k) & 1)) continue;
    const double v = wave_sum_d(acc[k]);
    if (lane == 0) sred[k][wave] = v;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 12; ++k) {  // thread k writes position k's row
    if (t != k || !((dmask >> k) & 1)) continue;
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) v += sred[k][w];
    partial[(uint64_t)A.slot[k] * gridDim.x + blockIdx.x] = v;
  }
}

// Instantiated phase sets: L2 alone (a window whose gates sit on its top 4
// bits: no transpose at all), L1 alone (up to 4 bits at the window's
// bottom), L2 + L1 (windows of 5-8 bits) and all three (the low pass); a
// pass runs the smallest set that holds its gates, empty positions as
// identities.
template <typename T, typename R, typename O, bool LAM, bool REAL>
static void launch_layer_r(void *psi, void *lam, uint64_t nchunks, int W, int h0, int C,
                           const LayerArgs<R> &a, double *partial, int nb, hipStream_t st,
                           const FuseDesc *fz) {
  const int need = (((a.has >> 8) & 15) ? 4 : 0) | (((a.has >> 4) & 15) ? 2 : 0) |
                   ((a.has & 15) ? 1 : 0);
  const bool wide = C == 6;  // only phase sets 2 and 6 store differently
  void (*k)(T *, T *, uint64_t, int, int, int, LayerArgs<R>, double *, const FuseDesc *) =
      fz                 ? k_adjoint_layer<T, R, O, LAM, 7, REAL, false, true>
      : need == 4        ? k_adjoint_layer<T, R, O, LAM, 4, REAL, false, false>
      : (need & ~2) == 0 ? (wide ? k_adjoint_layer<T, R, O, LAM, 2, REAL, true, false>
                                 : k_adjoint_layer<T, R, O, LAM, 2, REAL, false, false>)
      : (need & ~6) == 0 ? (wide ? k_adjoint_layer<T, R, O, LAM, 6, REAL, true, false>
                                 : k_adjoint_layer<T, R, O, LAM, 6, REAL, false, false>)
                         : k_adjoint_layer<T, R, O, LAM, 7, REAL, false, false>;
  hipLaunchKernelGGL(k, dim3(nb), dim3(kBlock), 0, st, (T *)psi, (T *)lam, nchunks, W, h0, C,
                     a, partial, fz);
}

template <typename T, typename R, typename O, bool LAM>
static void launch_layer_o(void *psi, void *lam, uint64_t nchunks, int W, int h0, int C,
                           const LayerArgs<R> &a, double *partial, int nb, hipStream_t st,
                           const FuseDesc *fz) {
  bool real = true;
  for (int k = 0; k < 12; ++k) {
    if (!((a.has >> k) & 1)) continue;
    for (int e = 0; e < 4; ++e)
      real = real && a.g[k].i[e] == 0 && (!LAM || a.slot[k] < 0 || a.d[k].i[e] == 0);
  }
  if (real)
    launch_layer_r<T, R, O, LAM, true>(psi, lam, nchunks, W, h0, C, a, partial, nb, st, fz);
  else
    launch_layer_r<T, R, O, LAM, false>(psi, lam, nchunks, W, h0, C, a, partial, nb, st, fz);
}

// 32-bit byte offsets when every byte of the pass lies below 2^31 from the
// chunk base (well inside the resource's 2^32 - 1 byte range).
template <typename T, typename R, bool LAM>
static void launch_layer_t(void *psi, void *lam, uint64_t nchunks, int W, int h0, int C,
                           const LayerArgs<R> &a, double *partial, int nb, hipStream_t st,
                           const FuseDesc *fz) {
  const int top = W == 0 ? 12 : h0 + W;  // offsets below 2^top elements
  if (top + (sizeof(T) == 8 ? 3 : 4) <= 31)
    launch_layer_o<T, R, uint32_t, LAM>(psi, lam, nchunks, W, h0, C, a, partial, nb, st, fz);
  else
    launch_layer_o<T, R, uint64_t, LAM>(psi, lam, nchunks, W, h0, C, a, partial, nb, st, fz);
}

template <typename R>
static LayerArgs<R> layer_args(int count, const int *pos, const int *slot, const double *gm,
                               const double *dm) {
  LayerArgs<R> a{};
  for (int k = 0; k < 12; ++k) a.slot[k] = -1;
  for (int j = 0; j < count; ++j) {
    const int k = pos[j];
    a.has |= 1 << k;
    a.slot[k] = slot ? slot[j] : -1;
    for (int e = 0; e < 4; ++e) {
      a.g[k].r[e] = (R)gm[8 * j + 2 * e];
      a.g[k].i[e] = (R)gm[8 * j + 2 * e + 1];
      if (dm) {
        a.d[k].r[e] = (R)dm[8 * j + 2 * e];
        a.d[k].i[e] = (R)dm[8 * j + 2 * e + 1];
      }
    }
  }
  return a;
}

int layer_pass_blocks(uint64_t n) {
  const uint64_t nchunks = n >> 12;
  return (int)(nchunks < kLayerMaxBlocks ? nchunks : kLayerMaxBlocks);
}

int launch_adjoint_layer(void *psi, void *lam, sq_dtype_t dt, uint64_t n, int h0, int W, int C,
                         int count, const int *pos, const double *gm, const double *dm,
                         const int *slot, int nslots, const int *slot_param_dev, bool distinct,
                         double *partial, double *out_dev, hipStream_t st,
                         const FuseDesc *fuse_dev) {
  if (n < 4096 || (n & 4095) || count < 1 || count > 12 || (C != 4 && C != 6) || W < 0 ||
      W > 12 - C)
    return fail(SQ_ERR_INVALID, "internal: layer pass shape");
  const int nq = __builtin_ctzll(n);
  if (W == 0 ? h0 != 12 : (h0 < 12 || h0 + W > nq))
    return fail(SQ_ERR_INVALID, "internal: layer pass window");
  int seen = 0;
  for (int j = 0; j < count; ++j) {
    if (pos[j] < 0 || pos[j] > 11 || ((seen >> pos[j]) & 1))
      return fail(SQ_ERR_INVALID, "internal: layer pass position");
    seen |= 1 << pos[j];
    if (lam && slot && (slot[j] < -1 || slot[j] >= nslots))
      return fail(SQ_ERR_INVALID, "internal: layer pass slot");
  }
  if (lam && nslots > 12) return fail(SQ_ERR_INVALID, "internal: layer pass slots");
  if (fuse_dev && (W != 0 || (seen & 0x00F) == 0 || (seen & 0x0F0) == 0 || (seen & 0xF00) == 0))
    return fail(SQ_ERR_INVALID, "internal: a fused run needs a low pass with gates in every layout");
  const uint64_t nchunks = n >> 12;
  const int nb = layer_pass_blocks(n);
  if (dt == SQ_COMPLEX64) {
    const LayerArgs<float> a = layer_args<float>(count, pos, lam ? slot : nullptr, gm,
                                                 lam ? dm : nullptr);
    if (lam)
      launch_layer_t<float2, float, true>(psi, lam, nchunks, W, h0, C, a, partial, nb, st, fuse_dev);
    else
      launch_layer_t<float2, float, false>(psi, nullptr, nchunks, W, h0, C, a, nullptr, nb, st, fuse_dev);
  } else {
    const LayerArgs<double> a = layer_args<double>(count, pos, lam ? slot : nullptr, gm,
                                                   lam ? dm : nullptr);
    if (lam)
      launch_layer_t<double2, double, true>(psi, lam, nchunks, W, h0, C, a, partial, nb, st, fuse_dev);
    else
      launch_layer_t<double2, double, false>(psi, nullptr, nchunks, W, h0, C, a, nullptr, nb, st, fuse_dev);
  }
  SQ_HIP(hipGetLastError());
  if (lam && nslots > 0 && out_dev)
    SQ_TRY(launch_finalize_slots(partial, nb, nslots, slot_param_dev, distinct, 2.0, out_dev, st));
  return SQ_OK;
}

// ---- tables of fused diagonal runs ---------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_fuse_tables(const FuseTableJob *__restrict__ jobs, uint64_t nchunks) {
  using R = decltype(T{}.x);
  const FuseTableJob jb = jobs[blockIdx.y];
  const uint64_t x = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool do_l = x < 4096, do_c = x < nchunks;
  double lr = 1.0, li = 0.0, cr = 1.0, ci = 0.0;
  for (int k = 0; k < jb.m; ++k) {
    const DiagRunGate &g = jb.g[k];
    const bool low0 = g.b0 < 12, low1 = g.b1 < 12;  // b1 < 0: one target
    if (low0 && low1 && do_l) {
      const int loc = g.b1 < 0 ? (int)((x >> g.b0) & 1)
                               : (int)((((x >> g.b0) & 1) << 1) | ((x >> g.b1) & 1));
      const double gr = g.ginv[loc][0], gi = g.ginv[loc][1];
      const double tr = lr * gr - li * gi;
      li = lr * gi + li * gr;
      lr = tr;
    }
    if (!low0 && (g.b1 < 0 || !low1) && do_c) {
      const uint64_t i = x << 12;
      const int loc = g.b1 < 0 ? (int)((i >> g.b0) & 1)
                               : (int)((((i >> g.b0) & 1) << 1) | ((i >> g.b1) & 1));
      const double gr = g.ginv[loc][0], gi = g.ginv[loc][1];
      const double tr = cr * gr - ci * gi;
      ci = cr * gi + ci * gr;
      cr = tr;
    }
  }
  if (do_l) reinterpret_cast<T *>(jb.tl)[x] = T{(R)lr, (R)li};
  if (do_c) reinterpret_cast<T *>(jb.tc)[x] = T{(R)cr, (R)ci};
}

int launch_fuse_tables(const FuseTableJob *jobs_dev, int njobs, sq_dtype_t dt, uint64_t n,
                       hipStream_t st) {
  if (njobs <= 0) return SQ_OK;
  if (njobs > 65535 || n < 4096) return fail(SQ_ERR_INVALID, "internal: fuse tables");
  const uint64_t nchunks = n >> 12;
  const uint64_t width = nchunks > 4096 ? nchunks : 4096;
  const dim3 g((unsigned)((width + kBlock - 1) / kBlock), njobs);
  if (dt == SQ_COMPLEX64)
    hipLaunchKernelGGL(k_fuse_tables<float2>, g, dim3(kBlock), 0, st, jobs_dev, nchunks);
  else
    hipLaunchKernelGGL(k_fuse_tables<double2>, g, dim3(kBlock), 0, st, jobs_dev, nchunks);
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

}  // namespace sq
