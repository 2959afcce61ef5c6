// kernels_cbatch.hip -- circuits of uneven qubit counts and different gate
// lists run as one launch sequence.
//
// No reference counterpart (the reference runs one circuit per simulation and
// has no batching; transpile.py:233-234 is where it would go).  Circuit j has
// n_j qubits and its own list of gates (dense matrices on 1-5 targets, e.g.
// the blocks of a fused circuit); its state lives at its own offset in one
// device buffer.  Launch s applies gate s of every circuit that has one: a
// table gives each gate its state, its workgroup range, its matrix and its
// index map, and a workgroup looks up its gate (uniform) and sweeps its
// share of the gate's groups.  So many small circuits, each of which alone
// would leave most of the GPU idle, fill it together, and a batch of C
// circuits of up to G gates takes at most 2G + 2 launches (captured in one
// hipGraph) instead of the sum of the gate counts.
//
// Gates that run on the 5-target MFMA kernels in a one-circuit program (5
// targets, or 4 widened to 5 on states of 15 qubits or more, with full
// tiles) take the same kernels here through an entry table (one launch per
// step for them); the others take k_cbatch_dense.  There a group is the 2^k
// amplitudes a k-target gate mixes: one lane per group per grid pass,
// consecutive lanes on consecutive groups (coalesced runs along the lowest
// free bits); the matrix is staged in LDS, so every lane of a wave reads the
// same entry (a broadcast).  Arithmetic in the state's precision.
#include <algorithm>
#include <memory>
#include <vector>

#include "kernels.hpp"
#include "kernels_cbatch.hpp"

namespace sq {

// One gate of one circuit within a launch.
struct CBGate {
  uint64_t psi;          // amplitude offset of the circuit's state
  uint64_t nfree;        // groups: 2^(n - k)
  uint64_t first_block;  // first workgroup of the gate in its launch
  uint64_t nblocks;      // its workgroups (grid stride over its groups)
  uint64_t mat;          // element offset of its matrix (plan order, row-major)
  IdxMap im;             // group index -> amplitude index (target bits 0)
  int k, pad0, pad1, pad2;
  int tb[kMaxK];         // amplitude bit of target b (local index bit b)
  int pad3[8 - kMaxK];
};

// The entry whose workgroup range holds workgroup b (first_block ascending).
// With at most kBlock entries their first blocks are staged in LDS by one
// load per thread and searched there.  Every thread must call it.
__device__ __forceinline__ int cb_find(const CBGate *tab, int count, uint64_t b) {
  __shared__ uint64_t sfb[kBlock];
  int lo = 0, hi = count - 1;
  if (count > kBlock) {
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tab[mid].first_block <= b) lo = mid; else hi = mid - 1;
    }
  } else {
    if ((int)threadIdx.x < count) sfb[threadIdx.x] = tab[threadIdx.x].first_block;
    __syncthreads();
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (sfb[mid] <= b) lo = mid; else hi = mid - 1;
    }
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

// y = M x on every group of the gate assigned to this workgroup.  M is
// staged in LDS as (re, im) pairs, row-major (a complex64 row is read two
// entries per 16-byte access), after the workgroup's first groups are
// requested: the two latencies overlap.  A group's amplitude offsets come
// from the target bits (uniform), so no table sits between the entry and
// the loads.  Every thread of the workgroup must call it (it synchronises).
template <typename R, int K>
__device__ __forceinline__ void cb_body(typename CV<R>::T *__restrict__ sv, const CBGate &g,
                                        uint64_t wg, const typename CV<R>::T *__restrict__ mat,
                                        typename CV<R>::T *M) {
  using T = typename CV<R>::T;
  constexpr int D = 1 << K;
  int tb[K];
#pragma unroll
  for (int b = 0; b < K; ++b) tb[b] = g.tb[b];
  auto offc = [&](int c) {
    uint64_t o = 0;
#pragma unroll
    for (int b = 0; b < K; ++b)
      if ((c >> b) & 1) o += 1ull << tb[b];
    return o;
  };
  const uint64_t stride = g.nblocks * kBlock;
  uint64_t f = wg * kBlock + threadIdx.x;
  T x[D];
  uint64_t base = 0;
  if (f < g.nfree) {
    base = deposit(f, g.im);
#pragma unroll
    for (int c = 0; c < D; ++c) x[c] = ld_nt(sv + base + offc(c));
  }
  for (int e = threadIdx.x; e < D * D; e += kBlock) M[e] = mat[e];
  __syncthreads();
  while (f < g.nfree) {
    // 1-2 targets: rows unrolled; 3-5: one row at a time (the D inputs stay
    // in registers, the row's entries are LDS broadcasts; unrolled rows would
    // hold the whole matrix in registers)
#pragma unroll K <= 2 ? D : 1
    for (int r = 0; r < D; ++r) {
      R yr = 0, yi = 0;
      if constexpr (sizeof(R) == 4) {
        const float4 *m4 = reinterpret_cast<const float4 *>(M + r * D);
#pragma unroll
        for (int c = 0; c < D; c += 2) {
          const float4 m = m4[c / 2];
          yr += m.x * x[c].x - m.y * x[c].y;
          yi += m.x * x[c].y + m.y * x[c].x;
          yr += m.z * x[c + 1].x - m.w * x[c + 1].y;
          yi += m.z * x[c + 1].y + m.w * x[c + 1].x;
        }
      } else {
#pragma unroll
        for (int c = 0; c < D; ++c) {
          const T m = M[r * D + c];
          yr += m.x * x[c].x - m.y * x[c].y;
          yi += m.x * x[c].y + m.y * x[c].x;
        }
      }
      T o;
      o.x = yr;
      o.y = yi;
      st_nt(sv + base + offc(r), o);
    }
    f += stride;
    if (f < g.nfree) {
      base = deposit(f, g.im);
#pragma unroll
      for (int c = 0; c < D; ++c) x[c] = ld_nt(sv + base + offc(c));
    }
  }
}

template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_cbatch_dense(typename CV<R>::T *__restrict__ st, const typename CV<R>::T *__restrict__ mats,
                   const CBGate *__restrict__ tab, int count) {
  using T = typename CV<R>::T;
  constexpr int DM = 1 << kMaxK;
  __shared__ __align__(16) T M[DM * DM];
  const int e = cb_find(tab, count, blockIdx.x);
  const CBGate &g = tab[e];
  T *sv = st + g.psi;
  const T *mat = mats + g.mat;
  const uint64_t wg = blockIdx.x - g.first_block;
  switch (g.k) {
    case 1: cb_body<R, 1>(sv, g, wg, mat, M); break;
    case 2: cb_body<R, 2>(sv, g, wg, mat, M); break;
    case 3: cb_body<R, 3>(sv, g, wg, mat, M); break;
    case 4: cb_body<R, 4>(sv, g, wg, mat, M); break;
    default: cb_body<R, 5>(sv, g, wg, mat, M); break;
  }
}

// Every state <- |0...0>: zero the buffer, then amplitude 0 of each state.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_cbatch_zero(typename CV<R>::T *__restrict__ st,
                                                        uint64_t nelems) {
  using T = typename CV<R>::T;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < nelems;
       i += (uint64_t)gridDim.x * kBlock)
    st_nt(st + i, T{0, 0});
}
template <typename R>
__global__ void __launch_bounds__(kBlock) k_cbatch_basis(typename CV<R>::T *__restrict__ st,
                                                         const uint64_t *__restrict__ psi,
                                                         uint32_t ncirc) {
  using T = typename CV<R>::T;
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j < ncirc) st[psi[j]] = T{1, 0};
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
constexpr uint64_t kCBMaxBlocksPerGate = 1ull << 14;  // grid stride beyond

struct CBatch {
  sq_dtype_t dt;
  uint32_t ncirc = 0;
  std::vector<uint32_t> nq;
  std::vector<uint64_t> host_psi;
  uint64_t nelems = 0;
  void *states = nullptr;
  void *mats = nullptr;
  void *g3m = nullptr;      // matrices of the MFMA entries (4 reals per element)
  CBGate *tables = nullptr;
  void *mtables = nullptr;  // MFMA entries (cbatch_mfma_entry)
  uint64_t *psi = nullptr;
  struct Launch {
    int mfma;  // 0: k_cbatch_dense over `tables`, 1: the MFMA kernels over `mtables`
    size_t tab_off;
    int count;
    uint64_t blocks;
  };
  std::vector<Launch> launches;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  double state_bytes = 0, traffic_bytes = 0;
};

static void free_cbatch(CBatch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (void *p : {b->states, b->mats, b->g3m, (void *)b->tables, b->mtables, (void *)b->psi})
    if (p) (void)hipFree(p);
}

int cbatch_create(void **handle, sq_dtype_t dt, uint32_t ncirc, const uint32_t *nq,
                  const uint32_t *ng, const uint32_t *nt, const uint32_t *targets,
                  const double *matrices) {
  if (ncirc == 0) return fail(SQ_ERR_INVALID, "circuit batch: no circuits");
  auto B = std::make_unique<CBatch>();
  B->dt = dt;
  B->ncirc = ncirc;
  const size_t item = itemsize(dt);
  // plans of every gate, circuit by circuit: as a program plans it with
  // kPlanDense (a dense matrix on all its targets, 4-target blocks of large
  // states widened to 5 for the MFMA kernels), and on exactly its own targets
  // (kPlanExact) for k_cbatch_dense
  std::vector<std::vector<GatePlan>> plans(ncirc), wplans(ncirc);
  size_t gi = 0, ti = 0, mi = 0;
  uint32_t maxg = 0;
  uint64_t amp = 0;
  for (uint32_t j = 0; j < ncirc; ++j) {
    if (nq[j] < 1 || nq[j] > 30)
      return fail(SQ_ERR_INVALID, "circuit batch: qubit counts must be in [1, 30]");
    plans[j].resize(ng[j]);
    wplans[j].resize(ng[j]);
    for (uint32_t g = 0; g < ng[j]; ++g, ++gi) {
      const int k = (int)nt[gi];
      if (k < 1 || k > kMaxK || k > (int)nq[j])
        return fail(SQ_ERR_INVALID, "circuit batch: circuit " + std::to_string(j) + ", gate " +
                                        std::to_string(g) + ": 1 to 5 targets, at most the qubit count");
      SQ_TRY(make_plan(matrices + mi, targets + ti, k, (int)nq[j], (int)nq[j], &wplans[j][g],
                       kPlanDense));
      SQ_TRY(make_plan(matrices + mi, targets + ti, k, (int)nq[j], (int)nq[j], &plans[j][g],
                       kPlanDense | kPlanExact));
      if (plans[j][g].kind != Kind::DENSE || plans[j][g].kk != k)
        return fail(SQ_ERR_INVALID, "circuit batch: internal: gate not planned dense");
      ti += (size_t)k;
      mi += 2ull << (2 * k);
    }
    maxg = std::max(maxg, ng[j]);
    B->nq.push_back(nq[j]);
    B->host_psi.push_back(amp);
    amp += 1ull << nq[j];
    B->state_bytes += std::ldexp(1.0, (int)nq[j]) * (double)item;
  }
  B->nelems = amp;
  // matrices in the state's precision, tables per launch
  std::vector<double> hm;  // interleaved re/im, converted below
  std::vector<double> g3;  // MFMA entries' matrices
  std::vector<CBGate> tabs;
  std::vector<uint8_t> mtabs;
  const size_t mbytes = cbatch_mfma_entry_bytes();
  const bool c64 = dt == SQ_COMPLEX64;
  for (uint32_t s = 0; s < maxg; ++s) {
    CBatch::Launch L{0, tabs.size(), 0, 0};
    CBatch::Launch M{1, mtabs.size() / mbytes, 0, 0};
    uint64_t mtiles = 0;  // tiles of the step's MFMA entries (sets their grid)
    for (uint32_t j = 0; j < ncirc; ++j)
      if (s < ng[j]) mtiles += cbatch_mfma_tiles(wplans[j][s], c64, 1ull << nq[j]);
    const uint64_t tpw = mfma_wave_tiles(mtiles);
    for (uint32_t j = 0; j < ncirc; ++j) {
      if (s >= ng[j]) continue;
      B->traffic_bytes += 2.0 * std::ldexp(1.0, (int)nq[j]) * (double)item;
      std::vector<uint8_t> ent(mbytes);
      uint64_t nb = 0;
      if (cbatch_mfma_entry(wplans[j][s], c64, 1ull << nq[j], B->host_psi[j], M.blocks, tpw, &g3,
                            ent.data(), &nb)) {
        mtabs.insert(mtabs.end(), ent.begin(), ent.end());
        M.blocks += nb;
        M.count++;
        continue;
      }
      const GatePlan &p = plans[j][s];
      CBGate e{};
      e.psi = B->host_psi[j];
      e.k = p.kk;
      e.nfree = 1ull << (nq[j] - p.kk);
      uint64_t tmask = 0;
      for (int b = 0; b < p.kk; ++b) {
        tmask |= 1ull << p.tbit[b];
        e.tb[b] = p.tbit[b];
      }
      e.im = make_idxmap(tmask, 0);
      e.mat = hm.size() / 2;
      for (const cd &z : p.m) {
        hm.push_back(z.real());
        hm.push_back(z.imag());
      }
      e.first_block = L.blocks;
      e.nblocks = std::min<uint64_t>((e.nfree + kBlock - 1) / kBlock, kCBMaxBlocksPerGate);
      L.blocks += e.nblocks;
      L.count++;
      tabs.push_back(e);
    }
    if (L.blocks > (uint64_t)INT32_MAX || M.blocks > (uint64_t)INT32_MAX)
      return fail(SQ_ERR_INVALID, "circuit batch: too many circuits for one launch per step");
    if (L.count) B->launches.push_back(L);
    if (M.count) B->launches.push_back(M);
  }
  B->traffic_bytes += (double)amp * (double)item;  // the initial state
  auto dmalloc = [](void **ptr, size_t bytes) -> int {
    if (hipMalloc(ptr, std::max<size_t>(bytes, 16)) != hipSuccess) {
      (void)hipGetLastError();
      *ptr = nullptr;
      return fail(SQ_ERR_OOM, "circuit batch: device allocation of " + std::to_string(bytes) +
                                  " failed");
    }
    return SQ_OK;
  };
  const size_t nm = hm.size() / 2;
  int rc = dmalloc(&B->states, amp * item);
  if (!rc) rc = dmalloc(&B->mats, nm * item);
  if (!rc) rc = dmalloc(&B->g3m, g3.size() * (c64 ? sizeof(float) : sizeof(double)));
  if (!rc) rc = dmalloc((void **)&B->tables, tabs.size() * sizeof(CBGate));
  if (!rc) rc = dmalloc(&B->mtables, mtabs.size());
  if (!rc) rc = dmalloc((void **)&B->psi, ncirc * sizeof(uint64_t));
  if (rc) {
    free_cbatch(B.get());
    return rc;
  }
  bool ok = true;
  if (dt == SQ_COMPLEX64) {
    std::vector<float> hf(hm.begin(), hm.end());
    ok = hipMemcpy(B->mats, hf.data(), hf.size() * sizeof(float), hipMemcpyHostToDevice) ==
         hipSuccess;
  } else {
    ok = hipMemcpy(B->mats, hm.data(), hm.size() * sizeof(double), hipMemcpyHostToDevice) ==
         hipSuccess;
  }
  if (c64) {
    std::vector<float> gf(g3.begin(), g3.end());
    ok = ok && (gf.empty() || hipMemcpy(B->g3m, gf.data(), gf.size() * sizeof(float),
                                        hipMemcpyHostToDevice) == hipSuccess);
  } else {
    ok = ok && (g3.empty() || hipMemcpy(B->g3m, g3.data(), g3.size() * sizeof(double),
                                        hipMemcpyHostToDevice) == hipSuccess);
  }
  ok = ok && (tabs.empty() || hipMemcpy(B->tables, tabs.data(), tabs.size() * sizeof(CBGate),
                                        hipMemcpyHostToDevice) == hipSuccess);
  ok = ok && (mtabs.empty() || hipMemcpy(B->mtables, mtabs.data(), mtabs.size(),
                                         hipMemcpyHostToDevice) == hipSuccess);
  ok = ok && hipMemcpy(B->psi, B->host_psi.data(), ncirc * sizeof(uint64_t),
                       hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) {
    free_cbatch(B.get());
    return fail(SQ_ERR_HIP, "circuit batch: upload failed");
  }
  *handle = B.release();
  return SQ_OK;
}

template <typename R>
static int cb_enqueue_t(CBatch *B, hipStream_t st) {
  using T = typename CV<R>::T;
  T *s = (T *)B->states;
  const uint64_t zb = std::min<uint64_t>((B->nelems + kBlock - 1) / kBlock, 1ull << 16);
  hipLaunchKernelGGL(k_cbatch_zero<R>, dim3((unsigned)zb), dim3(kBlock), 0, st, s, B->nelems);
  hipLaunchKernelGGL(k_cbatch_basis<R>, dim3((B->ncirc + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     st, s, B->psi, B->ncirc);
  const size_t mbytes = cbatch_mfma_entry_bytes();
  for (auto &L : B->launches) {
    if (L.mfma)
      SQ_TRY(launch_cbatch_mfma(B->dt, B->states, B->g3m,
                                (const uint8_t *)B->mtables + L.tab_off * mbytes, L.count,
                                L.blocks, st));
    else
      hipLaunchKernelGGL(k_cbatch_dense<R>, dim3((unsigned)L.blocks), dim3(kBlock), 0, st, s,
                         (const T *)B->mats, B->tables + L.tab_off, L.count);
  }
  SQ_HIP(hipGetLastError());
  return SQ_OK;
}

static int cb_enqueue(CBatch *B, hipStream_t st) {
  return B->dt == SQ_COMPLEX64 ? cb_enqueue_t<float>(B, st) : cb_enqueue_t<double>(B, st);
}

int cbatch_run(void *handle, int use_graph, hipStream_t st) {
  CBatch *B = (CBatch *)handle;
  if (!use_graph) return cb_enqueue(B, st);
  if (!B->exec) {
    hipStream_t cs;
    SQ_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    int rc = SQ_OK;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess)
      rc = fail(SQ_ERR_HIP, "circuit batch: hipStreamBeginCapture failed");
    if (!rc) rc = cb_enqueue(B, cs);
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(cs, &g) != hipSuccess && !rc)
      rc = fail(SQ_ERR_HIP, "circuit batch: hipStreamEndCapture failed");
    (void)hipStreamDestroy(cs);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    B->graph = g;
    SQ_HIP(hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0));
  }
  SQ_HIP(hipGraphLaunch(B->exec, st));
  return SQ_OK;
}

int cbatch_copy_state(void *handle, uint32_t circuit, void *dst, uint64_t nelems,
                      hipStream_t st) {
  CBatch *B = (CBatch *)handle;
  if (circuit >= B->ncirc) return fail(SQ_ERR_INVALID, "circuit batch: no such circuit");
  if (nelems != (1ull << B->nq[circuit]))
    return fail(SQ_ERR_INVALID, "circuit batch: destination has " + std::to_string(nelems) +
                                    " elements, circuit " + std::to_string(circuit) + " has 2^" +
                                    std::to_string(B->nq[circuit]));
  const size_t item = itemsize(B->dt);
  SQ_HIP(hipMemcpyAsync(dst, (const char *)B->states + B->host_psi[circuit] * item,
                        nelems * item, hipMemcpyDeviceToDevice, st));
  return SQ_OK;
}

int cbatch_stats(void *handle, uint64_t *launches, double *state_bytes, double *traffic_bytes) {
  CBatch *B = (CBatch *)handle;
  if (launches) *launches = B->launches.size() + 2;  // + the two initial-state kernels
  if (state_bytes) *state_bytes = B->state_bytes;
  if (traffic_bytes) *traffic_bytes = B->traffic_bytes;
  return SQ_OK;
}

int cbatch_destroy(void *handle) {
  CBatch *B = (CBatch *)handle;
  if (!B) return SQ_OK;
  free_cbatch(B);
  delete B;
  return SQ_OK;
}

}  // namespace sq
