// capi.hip -- the extern "C" boundary of libsnuqs_hip.so (include/snuqs_hip.h).
//
// Reference counterparts (src/braket/snuqs/csrc/):
//   runtime  core/cuda/runtime.cu:4-25          memory  buffer/buffer_cuda.cc:13-20
//   streams  stream/stream.cc:6-34              events  event/event.cc:5-26
//   init     functionals/functionals.cc:30-50   apply   functionals/functionals.cc:15-28
// The reference exits the process on any CUDA error (utils_cuda.h:8-17); here
// every entry point returns a status and records a message instead.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <array>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "kernels.hpp"
#include "kernels_qaoa.hpp"
#include "kernels_cbatch.hpp"
#include "kernels_reduce.hpp"

namespace sq {
const std::string &last_error();

// ---- per-device workspace ---------------------------------------------------
static std::mutex g_ws_mutex[64];
static Workspace g_ws[64];

static int current_device(int *dev) {
  SQ_HIP(hipGetDevice(dev));
  if (*dev < 0 || *dev >= 64) return fail(SQ_ERR_UNSUPPORTED, "device index >= 64");
  return SQ_OK;
}

// Caller holds g_ws_mutex[dev].
static int ws_reserve(int dev, size_t dev_bytes, size_t host_bytes, Workspace **out) {
  Workspace &w = g_ws[dev];
  if (w.dev_bytes < dev_bytes) {
    if (w.dev) SQ_HIP(hipFree(w.dev));
    w.dev = nullptr;
    w.dev_bytes = 0;
    size_t b = std::max(dev_bytes, (size_t)1 << 20);
    if (hipMalloc(&w.dev, b) != hipSuccess) {
      (void)hipGetLastError();
      return fail(SQ_ERR_OOM, "workspace: device allocation of " + std::to_string(b) +
                                  " bytes failed");
    }
    w.dev_bytes = b;
  }
  if (w.host_bytes < host_bytes) {
    if (w.host) SQ_HIP(hipHostFree(w.host));
    w.host = nullptr;
    w.host_bytes = 0;
    size_t b = std::max(host_bytes, (size_t)1 << 16);
    SQ_HIP(hipHostMalloc(&w.host, b, hipHostMallocDefault));
    w.host_bytes = b;
  }
  *out = &w;
  return SQ_OK;
}

static inline char *al(char *p) { return (char *)(((uintptr_t)p + 255) & ~(uintptr_t)255); }

static int check_state(const void *sv, int dt, uint64_t n) {
  if (!valid_dtype(dt)) return fail(SQ_ERR_INVALID, "dtype must be SQ_COMPLEX64 or SQ_COMPLEX128");
  if (!is_pow2(n)) return fail(SQ_ERR_INVALID, "num_elems must be a power of two");
  if (!sv) return fail(SQ_ERR_INVALID, "null state pointer");
  return SQ_OK;
}

}  // namespace sq

using namespace sq;

extern "C" {

const char *sq_last_error(void) { return last_error().c_str(); }
const char *sq_version(void) { return "snuqs-hip 0.1 (gfx950)"; }

// ---- runtime ----------------------------------------------------------------
int sq_mem_info(size_t *free_bytes, size_t *total_bytes) {
  SQ_HIP(hipMemGetInfo(free_bytes, total_bytes));
  return SQ_OK;
}
int sq_device_count(int *count) {
  SQ_HIP(hipGetDeviceCount(count));
  return SQ_OK;
}
int sq_set_device(int device) {
  SQ_HIP(hipSetDevice(device));
  return SQ_OK;
}
int sq_get_device(int *device) {
  SQ_HIP(hipGetDevice(device));
  return SQ_OK;
}
int sq_device_synchronize(void) {
  SQ_HIP(hipDeviceSynchronize());
  return SQ_OK;
}

// ---- memory -----------------------------------------------------------------
int sq_malloc(void **ptr, size_t bytes) {
  if (!ptr) return fail(SQ_ERR_INVALID, "null output pointer");
  *ptr = nullptr;
  if (bytes == 0) return SQ_OK;
  if (hipMalloc(ptr, bytes) != hipSuccess) {
    (void)hipGetLastError();
    *ptr = nullptr;
    return fail(SQ_ERR_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  return SQ_OK;
}
int sq_free(void *ptr) {
  if (ptr) SQ_HIP(hipFree(ptr));
  return SQ_OK;
}
int sq_host_alloc_pinned(void **ptr, size_t bytes) {
  if (!ptr) return fail(SQ_ERR_INVALID, "null output pointer");
  if (hipHostMalloc(ptr, bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    *ptr = nullptr;
    return fail(SQ_ERR_OOM, "hipHostMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  return SQ_OK;
}
int sq_host_free_pinned(void *ptr) {
  if (ptr) SQ_HIP(hipHostFree(ptr));
  return SQ_OK;
}
int sq_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream) {
  SQ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
  return SQ_OK;
}
int sq_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream) {
  SQ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
  return SQ_OK;
}
int sq_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream) {
  SQ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return SQ_OK;
}

int sq_copy(void *dst, const void *src, size_t bytes, void *stream) {
  if ((bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
    return fail(SQ_ERR_INVALID, "sq_copy needs 16-byte aligned pointers and sizes");
  return launch_copy(dst, src, bytes, (hipStream_t)stream);
}

// Tuning keys: "<knob>_c64" / "<knob>_c128".
static int *tune_slot(const std::string &k) {
  static const struct {
    const char *name;
    int *(*slot)();
  } table[] = {
      {"mfma_blocks_per_cu_c64", [] { return &tuning().mfma_blocks_per_cu[0]; }},
      {"mfma_blocks_per_cu_c128", [] { return &tuning().mfma_blocks_per_cu[1]; }},
      {"mfma5_pipeline_c64", [] { return &tuning().mfma5_pipeline[0]; }},
      {"mfma5_pipeline_c128", [] { return &tuning().mfma5_pipeline[1]; }},
      {"mfma5_order_c64", [] { return &tuning().mfma5_order[0]; }},
      {"mfma5_order_c128", [] { return &tuning().mfma5_order[1]; }},
      {"qaoa_mid_scatter", [] { return &tuning().qaoa_mid_scatter; }},
      {"tune_scatter", [] { return &tuning().tune_scatter; }},
      {"adjoint_layers", [] { return &tuning().adjoint_layers; }},
      {"adjoint_wide_windows", [] { return &tuning().adjoint_wide_windows; }},
      {"adjoint_fuse_runs", [] { return &tuning().adjoint_fuse_runs; }},
      {"mfma5_wide_c64", [] { return &tuning().mfma5_wide; }},
      {"mfma_tiles_per_wave", [] { return &tuning().mfma_tiles_per_wave; }},
  };
  for (auto &e : table)
    if (k == e.name) return e.slot();
  return nullptr;
}

int sq_tune_set(const char *key, int value) {
  const std::string k = key ? key : "";
  int *slot = tune_slot(k);
  if (!slot) return fail(SQ_ERR_INVALID, "unknown tuning key '" + k + "'");
  if (k.rfind("mfma_blocks_per_cu", 0) == 0 && (value < 1 || value > 4096))
    return fail(SQ_ERR_INVALID, "mfma_blocks_per_cu must be in [1, 4096]");
  if (k.rfind("mfma5_order", 0) == 0 && (value < 0 || value > 2))
    return fail(SQ_ERR_INVALID, "mfma5_order must be 0 (cyclic), 1 (blocked) or 2 (scattered)");
  const bool flag = k.rfind("mfma5_pipeline", 0) == 0 || k.rfind("mfma5_wide", 0) == 0 ||
                    k == "qaoa_mid_scatter" || k == "tune_scatter" || k == "adjoint_layers" ||
                    k == "adjoint_wide_windows";
  if (k == "mfma_tiles_per_wave" && (value < 1 || value > 64))
    return fail(SQ_ERR_INVALID, "mfma_tiles_per_wave must be in [1, 64]");
  if (k == "adjoint_fuse_runs" && (value < 0 || value > 3))
    return fail(SQ_ERR_INVALID, "adjoint_fuse_runs must be in [0, 3]");
  *slot = flag ? (value ? 1 : 0) : value;
  return SQ_OK;
}

int sq_tune_get(const char *key, int *value) {
  const std::string k = key ? key : "";
  int *slot = tune_slot(k);
  if (!slot) return fail(SQ_ERR_INVALID, "unknown tuning key '" + k + "'");
  *value = *slot;
  return SQ_OK;
}

// ---- streams / events (default priority only) ------------------------------
int sq_stream_create(void **stream, int nonblocking) {
  hipStream_t s;
  if (nonblocking)
    SQ_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  else
    SQ_HIP(hipStreamCreate(&s));
  *stream = (void *)s;
  return SQ_OK;
}
int sq_stream_destroy(void *stream) {
  if (stream) SQ_HIP(hipStreamDestroy((hipStream_t)stream));
  return SQ_OK;
}
int sq_stream_synchronize(void *stream) {
  SQ_HIP(hipStreamSynchronize((hipStream_t)stream));
  return SQ_OK;
}
int sq_stream_wait_event(void *stream, void *event) {
  SQ_HIP(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0));
  return SQ_OK;
}
int sq_event_create(void **event) {
  hipEvent_t e;
  SQ_HIP(hipEventCreate(&e));
  *event = (void *)e;
  return SQ_OK;
}
int sq_event_destroy(void *event) {
  if (event) SQ_HIP(hipEventDestroy((hipEvent_t)event));
  return SQ_OK;
}
int sq_event_record(void *event, void *stream) {
  SQ_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
  return SQ_OK;
}
int sq_event_synchronize(void *event) {
  SQ_HIP(hipEventSynchronize((hipEvent_t)event));
  return SQ_OK;
}
int sq_event_elapsed_ms(float *ms, void *start, void *stop) {
  SQ_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return SQ_OK;
}

// ---- init -------------------------------------------------------------------
int sq_initialize_zero(void *sv, sq_dtype_t dt, uint64_t n, void *stream) {
  SQ_TRY(check_state(sv, dt, n));
  return launch_init(sv, dt, n, false, (hipStream_t)stream);
}
int sq_initialize_basis_z(void *sv, sq_dtype_t dt, uint64_t n, void *stream) {
  SQ_TRY(check_state(sv, dt, n));
  return launch_init(sv, dt, n, true, (hipStream_t)stream);
}

// ---- apply ------------------------------------------------------------------
int sq_apply(void *sv, sq_dtype_t dt, uint64_t n, uint32_t num_qubits,
             const uint32_t *targets, uint32_t k, const double *matrix, void *stream) {
  SQ_TRY(check_state(sv, dt, n));
  if (!matrix || (k && !targets)) return fail(SQ_ERR_INVALID, "null matrix or targets");
  if (num_qubits > 63) return fail(SQ_ERR_INVALID, "num_qubits must be < 64");
  GatePlan p;
  SQ_TRY(make_plan(matrix, targets, (int)k, (int)num_qubits, log2u(n), &p));
  hipStream_t st = (hipStream_t)stream;
  if (p.kind == Kind::DENSE && p.kk >= 4) {
    // 4-5 target dense gates need the matrix in device memory: stage it in the
    // workspace and finish before the workspace is released.
    int dev;
    SQ_TRY(current_device(&dev));
    std::lock_guard<std::mutex> lk(g_ws_mutex[dev]);
    std::vector<double> g = fused_block_3m(p);
    const size_t esz = dt == SQ_COMPLEX64 ? 4 : 8;
    const size_t bytes = g.size() * esz;
    Workspace *w;
    SQ_TRY(ws_reserve(dev, bytes, bytes, &w));
    if (dt == SQ_COMPLEX64) {
      float *h = (float *)w->host;
      for (size_t i = 0; i < g.size(); ++i) h[i] = (float)g[i];
    } else {
      std::memcpy(w->host, g.data(), bytes);
    }
    SQ_HIP(hipMemcpyAsync(w->dev, w->host, bytes, hipMemcpyHostToDevice, st));
    SQ_TRY(launch_plan(p, sv, dt, n, st, w->dev));
    SQ_HIP(hipStreamSynchronize(st));
    return SQ_OK;
  }
  return launch_plan(p, sv, dt, n, st, nullptr);
}

}  // extern "C"

// ---- programs -----------------------------------------------------------------
namespace sq {

struct Program {
  sq_dtype_t dt;
  int nq;
  int device;
  std::vector<GatePlan> fwd, inv;
  std::vector<int> nt;                 // targets per gate
  std::vector<size_t> toff;            // offset of a gate's targets
  std::vector<uint32_t> targets;
  std::vector<double> mats;            // the gates' complex128 matrices (as given)
  std::vector<size_t> moffs;           // offset of a gate's matrix in mats
  std::vector<const void *> fwd_dev, inv_dev;  // real-block matrices (kk >= 4)
  std::vector<size_t> slot;            // byte offset of a gate's block slots, or -1
  void *dev_mats = nullptr;
  size_t dev_cap = 0;                  // bytes of dev_mats
  uint64_t launches = 0;
  double bytes = 0.0;
  // hipGraph of the forward launch sequence on one state (sq_program_run_graph)
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  void *graph_sv = nullptr;
};

static void drop_graph(Program *p) {
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->gexec = nullptr;
  p->graph = nullptr;
  p->graph_sv = nullptr;
}

// Real-block matrices of dense 4-5 target plans (layout of fused_block_3m).
// `off[i]` is plan i's byte offset in the image, or -1 for a plan that keeps
// its matrix in kernel arguments; the image is `total` bytes.
static size_t block_bytes(int kk, sq_dtype_t dt) {
  const size_t D = (size_t)2 << kk;
  return ((D * D * (dt == SQ_COMPLEX64 ? 4 : 8) + 255) / 256) * 256;
}
static bool has_block(const GatePlan &p) { return p.kind == Kind::DENSE && p.kk >= 4; }

// Offsets packed by the plans' own sizes (derivative matrices of one call).
static size_t block_layout(const std::vector<const GatePlan *> &plans, sq_dtype_t dt,
                           std::vector<size_t> *off) {
  off->assign(plans.size(), (size_t)-1);
  size_t total = 0;
  for (size_t i = 0; i < plans.size(); ++i)
    if (has_block(*plans[i])) {
      (*off)[i] = total;
      total += block_bytes(plans[i]->kk, dt);
    }
  return total;
}

// Host image of the blocks at the given offsets.
static void block_image(const std::vector<const GatePlan *> &plans, sq_dtype_t dt,
                        const std::vector<size_t> &off, char *img) {
  for (size_t i = 0; i < plans.size(); ++i) {
    if (off[i] == (size_t)-1 || !has_block(*plans[i])) continue;
    std::vector<double> g = fused_block_3m(*plans[i]);
    if (dt == SQ_COMPLEX64) {
      float *h = (float *)(img + off[i]);
      for (size_t e = 0; e < g.size(); ++e) h[e] = (float)g[e];
    } else {
      std::memcpy(img + off[i], g.data(), g.size() * 8);
    }
  }
}

// Plans of a program's gates from `matrices` (the create-time layout: gate g's
// row-major complex128 matrix at 2 * sum_{h<g} 4^nt[h] doubles), its dagger
// for the reverse sweep, and the totals.  Launch shapes tuned for a gate are
// kept while its plan stays a 5-target block.
static int plan_program(Program *prog, const double *matrices) {
  const size_t ng = prog->nt.size();
  std::vector<GatePlan> fwd(ng), inv(ng);
  uint64_t launches = 0;
  double bytes = 0.0;
  size_t moff = 0;
  prog->moffs.resize(ng);
  for (size_t g = 0; g < ng; ++g) {
    const int k = prog->nt[g];
    prog->moffs[g] = moff;
    const uint32_t *t = prog->targets.data() + prog->toff[g];
    const double *m = matrices + moff;
    SQ_TRY(make_plan(m, t, k, prog->nq, prog->nq, &fwd[g]));
    std::vector<double> md = dagger(m, k);
    SQ_TRY(make_plan(md.data(), t, k, prog->nq, prog->nq, &inv[g]));
    if (g < prog->fwd.size() && has_block(prog->fwd[g]) && has_block(fwd[g]) &&
        fwd[g].kk == prog->fwd[g].kk) {
      fwd[g].shape = prog->fwd[g].shape;
      inv[g].shape = prog->inv[g].shape;
    }
    moff += 2ull << (2 * k);
    if (fwd[g].kind != Kind::NOOP) {
      launches++;
      bytes += 2.0 * fwd[g].touched * std::ldexp(1.0, prog->nq) * itemsize(prog->dt);
    }
  }
  prog->fwd.swap(fwd);
  prog->inv.swap(inv);
  prog->mats.assign(matrices, matrices + moff);
  prog->launches = launches;
  prog->bytes = bytes;
  return SQ_OK;
}

// Host image of a program's blocks in its fixed slots: every gate of 4 or more
// targets owns a 5-target slot for its plan and one for its dagger, so new
// matrices for the same gates never change the layout (sq_program_set_matrices
// allocates nothing).  Sets the per-plan device pointers.
static std::vector<char> program_image(Program *p) {
  const size_t ng = p->fwd.size();
  std::vector<char> img(p->dev_cap);
  std::vector<const GatePlan *> plans;
  std::vector<size_t> off;
  for (size_t g = 0; g < ng; ++g) {
    plans.push_back(&p->fwd[g]);
    off.push_back(p->slot[g]);
  }
  for (size_t g = 0; g < ng; ++g) {
    plans.push_back(&p->inv[g]);
    off.push_back(p->slot[g] == (size_t)-1 ? (size_t)-1 : p->slot[g] + block_bytes(kMaxK, p->dt));
  }
  block_image(plans, p->dt, off, img.data());
  p->fwd_dev.assign(ng, nullptr);
  p->inv_dev.assign(ng, nullptr);
  for (size_t i = 0; i < plans.size(); ++i) {
    if (off[i] == (size_t)-1 || !has_block(*plans[i])) continue;
    const void *d = (const char *)p->dev_mats + off[i];
    if (i < ng) p->fwd_dev[i] = d; else p->inv_dev[i - ng] = d;
  }
  return img;
}

}  // namespace sq

extern "C" {

int sq_program_create(void **program, sq_dtype_t dt, uint32_t nq, uint32_t ng,
                      const uint32_t *num_targets, const uint32_t *targets,
                      const double *matrices) {
  if (!program) return fail(SQ_ERR_INVALID, "null output pointer");
  *program = nullptr;
  if (!valid_dtype(dt)) return fail(SQ_ERR_INVALID, "bad dtype");
  if (nq < 1 || nq > 40) return fail(SQ_ERR_INVALID, "num_qubits must be in [1, 40]");
  auto prog = std::make_unique<Program>();
  prog->dt = dt;
  prog->nq = (int)nq;
  SQ_TRY(current_device(&prog->device));
  size_t toff = 0;
  for (uint32_t g = 0; g < ng; ++g) {
    const int k = (int)num_targets[g];
    if (k > kMaxK) return fail(SQ_ERR_UNSUPPORTED, "gate with more than 5 targets");
    prog->nt.push_back(k);
    prog->toff.push_back(toff);
    prog->targets.insert(prog->targets.end(), targets + toff, targets + toff + k);
    toff += k;
  }
  SQ_TRY(plan_program(prog.get(), matrices));
  for (uint32_t g = 0; g < ng; ++g) {
    prog->slot.push_back(prog->nt[g] >= 4 ? prog->dev_cap : (size_t)-1);
    if (prog->nt[g] >= 4) prog->dev_cap += 2 * block_bytes(kMaxK, dt);
  }
  if (prog->dev_cap) {
    if (hipMalloc(&prog->dev_mats, prog->dev_cap) != hipSuccess) {
      (void)hipGetLastError();
      prog->dev_mats = nullptr;
      return fail(SQ_ERR_OOM, "program matrices: device allocation failed");
    }
  }
  std::vector<char> img = program_image(prog.get());
  if (prog->dev_cap) {
    if (hipMemcpy(prog->dev_mats, img.data(), prog->dev_cap, hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipFree(prog->dev_mats);
      return fail(SQ_ERR_HIP, "program matrices: upload failed");
    }
  }
  *program = prog.release();
  return SQ_OK;
}

int sq_program_set_matrices(void *program, const double *matrices, void *stream) {
  Program *p = (Program *)program;
  if (!p) return fail(SQ_ERR_INVALID, "null program");
  if (!matrices && !p->nt.empty()) return fail(SQ_ERR_INVALID, "null matrices");
  hipStream_t st = (hipStream_t)stream;
  // the blocks may still be read by work queued on `stream`; the new image is
  // copied behind it, and the call returns once the copy is done
  SQ_TRY(plan_program(p, matrices));
  drop_graph(p);  // kernel arguments (1-3 target matrices) changed
  std::vector<char> img = program_image(p);
  if (p->dev_cap) {
    SQ_HIP(hipMemcpyAsync(p->dev_mats, img.data(), p->dev_cap, hipMemcpyHostToDevice, st));
    SQ_HIP(hipStreamSynchronize(st));
  }
  return SQ_OK;
}

int sq_program_destroy(void *program) {
  Program *p = (Program *)program;
  if (!p) return SQ_OK;
  drop_graph(p);
  if (p->dev_mats) SQ_HIP(hipFree(p->dev_mats));
  delete p;
  return SQ_OK;
}

static int gate_class(const GatePlan &p, uint64_t n, sq_dtype_t dt) {
  switch (p.kind) {
    case Kind::NOOP: return 0;
    case Kind::PHASE: return 1;
    case Kind::DIAG: return 2;
    case Kind::PERM: return 3;
    case Kind::DENSE: {
      if (p.kk <= 3) return 4;
      const uint64_t nfree = n >> (p.kk + __builtin_popcountll(p.fixed_mask));
      const uint64_t tg = dt == SQ_COMPLEX64 ? 32 : 16;
      return (nfree >= tg && nfree % tg == 0) ? 5 : 6;
    }
  }
  return 0;
}

int sq_program_gate_info(void *program, uint32_t gate, int *kernel_class, int *num_targets,
                         double *algorithmic_bytes) {
  Program *p = (Program *)program;
  if (!p) return fail(SQ_ERR_INVALID, "null program");
  if (gate >= p->fwd.size()) return fail(SQ_ERR_INVALID, "gate index out of range");
  const GatePlan &g = p->fwd[gate];
  if (kernel_class) *kernel_class = gate_class(g, 1ull << p->nq, p->dt);
  if (num_targets) *num_targets = g.kk;
  if (algorithmic_bytes)
    *algorithmic_bytes = 2.0 * g.touched * std::ldexp(1.0, p->nq) * itemsize(p->dt);
  return SQ_OK;
}

int sq_program_run_timed(void *program, void *sv, uint64_t n, void *stream, float *ms) {
  Program *p = (Program *)program;
  if (!p) return fail(SQ_ERR_INVALID, "null program");
  SQ_TRY(check_state(sv, p->dt, n));
  if (n != (1ull << p->nq)) return fail(SQ_ERR_INVALID, "state size does not match program");
  hipStream_t st = (hipStream_t)stream;
  const size_t ng = p->fwd.size();
  std::vector<hipEvent_t> ev(ng + 1);
  for (auto &e : ev) SQ_HIP(hipEventCreate(&e));
  int rc = SQ_OK;
  for (size_t g = 0; g < ng && rc == SQ_OK; ++g) {
    if (hipEventRecord(ev[g], st) != hipSuccess) rc = fail(SQ_ERR_HIP, "hipEventRecord failed");
    if (rc == SQ_OK) rc = launch_plan(p->fwd[g], sv, p->dt, n, st, p->fwd_dev[g]);
  }
  if (rc == SQ_OK && hipEventRecord(ev[ng], st) != hipSuccess)
    rc = fail(SQ_ERR_HIP, "hipEventRecord failed");
  if (rc == SQ_OK && hipEventSynchronize(ev[ng]) != hipSuccess)
    rc = fail(SQ_ERR_HIP, "hipEventSynchronize failed");
  for (size_t g = 0; g < ng && rc == SQ_OK; ++g) {
    float t = 0.f;
    if (hipEventElapsedTime(&t, ev[g], ev[g + 1]) != hipSuccess)
      rc = fail(SQ_ERR_HIP, "hipEventElapsedTime failed");
    ms[g] = p->fwd[g].kind == Kind::NOOP ? 0.f : t;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  return rc;
}

// Launch shapes chosen by sq_program_tune, keyed by what decides the memory
// pattern: dtype, state size, target bits, control bits and values.
namespace {
struct ShapeKey {
  int dt, nbits;
  uint64_t tmask, fmask, fval;
  bool operator<(const ShapeKey &o) const {
    return std::tie(dt, nbits, tmask, fmask, fval) <
           std::tie(o.dt, o.nbits, o.tmask, o.fmask, o.fval);
  }
};
std::mutex g_shape_mutex;
std::map<ShapeKey, GatePlan::Shape> g_shape_cache;
}  // namespace

// Runs the program once with an event before every gate and adds each gate's
// elapsed time to ms[g].
static int run_timed_into(Program *p, void *sv, uint64_t n, hipStream_t st,
                          std::vector<hipEvent_t> &ev, double *ms) {
  const size_t ng = p->fwd.size();
  for (size_t g = 0; g < ng; ++g) {
    SQ_HIP(hipEventRecord(ev[g], st));
    SQ_TRY(launch_plan(p->fwd[g], sv, p->dt, n, st, p->fwd_dev[g]));
  }
  SQ_HIP(hipEventRecord(ev[ng], st));
  SQ_HIP(hipEventSynchronize(ev[ng]));
  for (size_t g = 0; g < ng; ++g) {
    float t = 0.f;
    SQ_HIP(hipEventElapsedTime(&t, ev[g], ev[g + 1]));
    ms[g] += t;
  }
  return SQ_OK;
}

int sq_program_tune(void *program, void *sv, uint64_t n, int reps, void *stream,
                    uint32_t *num_tuned) {
  Program *p = (Program *)program;
  if (!p) return fail(SQ_ERR_INVALID, "null program");
  SQ_TRY(check_state(sv, p->dt, n));
  if (n != (1ull << p->nq)) return fail(SQ_ERR_INVALID, "state size does not match program");
  if (reps < 1 || reps > 100) return fail(SQ_ERR_INVALID, "reps must be in [1, 100]");
  hipStream_t st = (hipStream_t)stream;
  const size_t ng = p->fwd.size();
  drop_graph(p);  // launch shapes change: a captured graph is stale
  // gates to tune; those whose key is cached take the cached shape
  std::vector<size_t> todo;
  std::vector<ShapeKey> keys(ng);
  uint32_t tuned = 0;
  for (size_t g = 0; g < ng; ++g) {
    GatePlan &f = p->fwd[g];
    if (gate_class(f, n, p->dt) != 5 || f.kk != 5) continue;
    uint64_t tmask = 0;
    for (int j = 0; j < f.kk; ++j) tmask |= 1ull << f.tbit[j];
    keys[g] = ShapeKey{(int)p->dt, p->nq, tmask, f.fixed_mask, f.fixed_val};
    std::lock_guard<std::mutex> lk(g_shape_mutex);
    auto it = g_shape_cache.find(keys[g]);
    if (it != g_shape_cache.end()) {
      f.shape = p->inv[g].shape = it->second;
      ++tuned;
    } else {
      todo.push_back(g);
    }
  }
  if (todo.empty()) {
    if (num_tuned) *num_tuned = tuned;
    return SQ_OK;
  }
  // Each candidate shape is timed inside whole-program runs (what a gate
  // costs depends on the launches around it: timing a gate alone, launch
  // after launch, mispredicts it); every gate then keeps its fastest shape.
  std::vector<GatePlan::Shape> cands;
  for (int order = 0; order < (tuning().tune_scatter ? 3 : 2); ++order) {
    // grid caps of 64-1024 blocks per CU: at 2^20 tiles, 16 down to 1 tile
    // per wave
    for (int bpc : {64, 128, 256, 512, 1024}) cands.push_back({bpc, order, 1, 0});
    for (int bpc : {256, 1024}) cands.push_back({bpc, order, 0, 0});
    if (p->dt == SQ_COMPLEX64) {  // 64-group tiles: longer runs, half the waves per CU
      for (int bpc : {128, 256, 512, 1024}) cands.push_back({bpc, order, 1, 1});
      for (int bpc : {256, 1024}) cands.push_back({bpc, order, 0, 1});
    }
  }
  std::vector<hipEvent_t> ev(ng + 1);
  for (auto &e : ev) SQ_HIP(hipEventCreate(&e));
  std::vector<std::vector<double>> ms(cands.size(), std::vector<double>(ng, 0.0));
  int rc = SQ_OK;
  for (size_t c = 0; c < cands.size() && rc == SQ_OK; ++c) {
    for (size_t g : todo) p->fwd[g].shape = cands[c];
    for (int r = 0; r < reps && rc == SQ_OK; ++r)
      rc = run_timed_into(p, sv, n, st, ev, ms[c].data());
  }
  // Second round: each gate's three fastest candidates are timed again,
  // reps + 1 times each (a run sets every gate to its k-th candidate, so one
  // run re-times different shapes of different gates); the pick uses these
  // times only, which damps the selection noise of the screening round.
  constexpr int kTop = 3;
  std::vector<std::array<size_t, kTop>> top(ng);
  for (size_t g : todo) {
    std::vector<size_t> order(cands.size());
    for (size_t c = 0; c < cands.size(); ++c) order[c] = c;
    std::partial_sort(order.begin(), order.begin() + kTop, order.end(),
                      [&](size_t a, size_t b) { return ms[a][g] < ms[b][g]; });
    for (int k = 0; k < kTop; ++k) top[g][k] = order[k];
  }
  std::vector<std::vector<double>> ms2(kTop, std::vector<double>(ng, 0.0));
  for (int k = 0; k < kTop && rc == SQ_OK; ++k) {
    for (size_t g : todo) p->fwd[g].shape = cands[top[g][k]];
    for (int r = 0; r <= reps && rc == SQ_OK; ++r)
      rc = run_timed_into(p, sv, n, st, ev, ms2[k].data());
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  if (rc != SQ_OK) {
    for (size_t g : todo) p->fwd[g].shape = GatePlan::Shape{};
    return rc;
  }
  std::lock_guard<std::mutex> lk(g_shape_mutex);
  for (size_t g : todo) {
    int best = 0;
    for (int k = 1; k < kTop; ++k)
      if (ms2[k][g] < ms2[best][g]) best = k;
    const GatePlan::Shape &sh = cands[top[g][best]];
    p->fwd[g].shape = p->inv[g].shape = sh;
    g_shape_cache[keys[g]] = sh;
    ++tuned;
  }
  if (num_tuned) *num_tuned = tuned;
  return SQ_OK;
}

int sq_program_run_graph(void *program, void *sv, uint64_t n, void *stream) {
  Program *p = (Program *)program;
  if (!p) return fail(SQ_ERR_INVALID, "null program");
  SQ_TRY(check_state(sv, p->dt, n));
  if (n != (1ull << p->nq)) return fail(SQ_ERR_INVALID, "state size does not match program");
  if (p->graph_sv != sv) {
    // capture on a private stream (thread-local mode: other threads' work is
    // not pulled in); the graph holds only kernel launches
    drop_graph(p);
    hipStream_t cs;
    SQ_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    int rc = SQ_OK;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess)
      rc = fail(SQ_ERR_HIP, "hipStreamBeginCapture failed");
    for (size_t g = 0; g < p->fwd.size() && rc == SQ_OK; ++g)
      rc = launch_plan(p->fwd[g], sv, p->dt, n, cs, p->fwd_dev[g]);
    hipGraph_t gr = nullptr;
    if (hipStreamEndCapture(cs, &gr) != hipSuccess && rc == SQ_OK)
      rc = fail(SQ_ERR_HIP, "hipStreamEndCapture failed");
    (void)hipStreamDestroy(cs);
    if (rc == SQ_OK && hipGraphInstantiate(&p->gexec, gr, nullptr, nullptr, 0) != hipSuccess)
      rc = fail(SQ_ERR_HIP, "hipGraphInstantiate failed");
    if (rc != SQ_OK) {
      if (gr) (void)hipGraphDestroy(gr);
      p->gexec = nullptr;
      return rc;
    }
    p->graph = gr;
    p->graph_sv = sv;
  }
  SQ_HIP(hipGraphLaunch(p->gexec, (hipStream_t)stream));
  return SQ_OK;
}

int sq_program_stats(void *program, uint64_t *num_launches, double *bytes) {
  Program *p = (Program *)program;
  if (!p) return fail(SQ_ERR_INVALID, "null program");
  if (num_launches) *num_launches = p->launches;
  if (bytes) *bytes = p->bytes;
  return SQ_OK;
}

int sq_program_run(void *program, void *sv, uint64_t n, void *stream) {
  Program *p = (Program *)program;
  if (!p) return fail(SQ_ERR_INVALID, "null program");
  SQ_TRY(check_state(sv, p->dt, n));
  if (n != (1ull << p->nq))
    return fail(SQ_ERR_INVALID, "program built for " + std::to_string(p->nq) +
                                    " qubits, state has " + std::to_string(n) + " amplitudes");
  hipStream_t st = (hipStream_t)stream;
  for (size_t g = 0; g < p->fwd.size(); ++g)
    SQ_TRY(launch_plan(p->fwd[g], sv, p->dt, n, st, p->fwd_dev[g]));
  return SQ_OK;
}

// ---- results ----------------------------------------------------------------
int sq_probabilities(const void *sv, sq_dtype_t dt, uint64_t n, double *probs, void *stream) {
  SQ_TRY(check_state(sv, dt, n));
  if (!probs) return fail(SQ_ERR_INVALID, "null output");
  return launch_probabilities(sv, dt, n, probs, (hipStream_t)stream);
}

int sq_probabilities_normalized(const void *sv, sq_dtype_t dt, uint64_t n, double *probs,
                                void *stream) {
  SQ_TRY(check_state(sv, dt, n));
  if (!probs) return fail(SQ_ERR_INVALID, "null output");
  int dev;
  SQ_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mutex[dev]);
  Workspace *w;
  SQ_TRY(ws_reserve(dev, (kRedBlocks + 64) * sizeof(double), 64, &w));
  hipStream_t st = (hipStream_t)stream;
  double *res = (double *)w->dev;
  SQ_TRY(launch_norm2(sv, dt, n, res + 32, res, st));
  SQ_TRY(launch_probabilities_div(sv, dt, n, probs, res, st));
  // the workspace is reused by the next call: finish before the lock drops
  SQ_HIP(hipStreamSynchronize(st));
  return SQ_OK;
}

int sq_marginal_probabilities(const void *sv, sq_dtype_t dt, uint64_t n, uint32_t num_qubits,
                              const uint32_t *targets, uint32_t k, double *out, void *stream) {
  SQ_TRY(check_state(sv, dt, n));
  if (!out || (k && !targets)) return fail(SQ_ERR_INVALID, "null output or targets");
  if (k < 1 || k > (uint32_t)kMaxMarginal)
    return fail(SQ_ERR_INVALID, "marginal over 1.." + std::to_string(kMaxMarginal) +
                                    " targets only (use sq_probabilities for the full distribution)");
  const int nbits = log2u(n);
  int tbit[kMaxMarginal];
  uint64_t seen = 0;
  for (uint32_t j = 0; j < k; ++j) {
    if (targets[j] >= num_qubits) return fail(SQ_ERR_INVALID, "target out of range");
    tbit[j] = (int)num_qubits - 1 - (int)targets[j];
    if (tbit[j] >= nbits) return fail(SQ_ERR_INVALID, "target stride beyond the state");
    if ((seen >> tbit[j]) & 1) return fail(SQ_ERR_INVALID, "duplicate target");
    seen |= 1ull << tbit[j];
  }
  const int nb = std::max(1, std::min(kRedBlocks, 65536 >> k));
  const uint64_t nout = 1ull << k;
  int dev;
  SQ_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mutex[dev]);
  Workspace *w;
  SQ_TRY(ws_reserve(dev, (nout * nb + nout) * sizeof(double), nout * sizeof(double), &w));
  hipStream_t st = (hipStream_t)stream;
  double *res = (double *)w->dev;
  double *partial = res + nout;
  SQ_TRY(launch_marginal(sv, dt, n, nbits, tbit, (int)k, partial, nb, res, st));
  SQ_HIP(hipMemcpyAsync(w->host, res, nout * sizeof(double), hipMemcpyDeviceToHost, st));
  SQ_HIP(hipStreamSynchronize(st));
  std::memcpy(out, w->host, nout * sizeof(double));
  return SQ_OK;
}

int sq_norm2(const void *sv, sq_dtype_t dt, uint64_t n, double *out, void *stream) {
  SQ_TRY(check_state(sv, dt, n));
  int dev;
  SQ_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mutex[dev]);
  Workspace *w;
  SQ_TRY(ws_reserve(dev, (kRedBlocks + 64) * sizeof(double), 64, &w));
  hipStream_t st = (hipStream_t)stream;
  double *res = (double *)w->dev;
  double *partial = res + 32;
  SQ_TRY(launch_norm2(sv, dt, n, partial, res, st));
  SQ_HIP(hipMemcpyAsync(w->host, res, sizeof(double), hipMemcpyDeviceToHost, st));
  SQ_HIP(hipStreamSynchronize(st));
  *out = *(double *)w->host;
  return SQ_OK;
}

int sq_scale(void *sv, sq_dtype_t dt, uint64_t n, double re, double im, void *stream) {
  SQ_TRY(check_state(sv, dt, n));
  return launch_scale(sv, dt, n, re, im, (hipStream_t)stream);
}

int sq_amplitudes(const void *sv, sq_dtype_t dt, uint64_t n, const uint64_t *indices,
                  uint64_t count, double *out, void *stream) {
  SQ_TRY(check_state(sv, dt, n));
  if (count == 0) return SQ_OK;
  for (uint64_t j = 0; j < count; ++j)
    if (indices[j] >= n) return fail(SQ_ERR_INVALID, "amplitude index out of range");
  int dev;
  SQ_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mutex[dev]);
  const size_t ib = count * 8, ob = count * 16;
  Workspace *w;
  SQ_TRY(ws_reserve(dev, ib + ob + 512, ib + ob + 512, &w));
  hipStream_t st = (hipStream_t)stream;
  char *hb = (char *)w->host;
  std::memcpy(hb, indices, ib);
  uint64_t *didx = (uint64_t *)w->dev;
  double *dout = (double *)al((char *)w->dev + ib);
  SQ_HIP(hipMemcpyAsync(didx, hb, ib, hipMemcpyHostToDevice, st));
  SQ_TRY(launch_gather(sv, dt, didx, count, dout, st));
  char *hout = al(hb + ib);
  SQ_HIP(hipMemcpyAsync(hout, dout, ob, hipMemcpyDeviceToHost, st));
  SQ_HIP(hipStreamSynchronize(st));
  std::memcpy(out, hout, ob);
  return SQ_OK;
}

int sq_inner(const void *a, const void *b, sq_dtype_t dt, uint64_t n, double *out,
             void *stream) {
  SQ_TRY(check_state(a, dt, n));
  SQ_TRY(check_state(b, dt, n));
  int dev;
  SQ_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mutex[dev]);
  Workspace *w;
  SQ_TRY(ws_reserve(dev, (2 * kRedBlocks + 64) * sizeof(double), 64, &w));
  hipStream_t st = (hipStream_t)stream;
  double *res = (double *)w->dev;
  double *partial = res + 32;
  SQ_TRY(launch_inner(a, b, dt, n, partial, res, nullptr, 0.0, st));
  SQ_HIP(hipMemcpyAsync(w->host, res, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  SQ_HIP(hipStreamSynchronize(st));
  out[0] = ((double *)w->host)[0];
  out[1] = ((double *)w->host)[1];
  return SQ_OK;
}

}  // extern "C"

namespace sq {

static int check_terms(uint32_t nterms, const uint64_t *x, const uint64_t *z,
                       const uint32_t *ny, uint64_t n) {
  for (uint32_t t = 0; t < nterms; ++t) {
    if (x[t] >= n || z[t] >= n) return fail(SQ_ERR_INVALID, "Pauli mask beyond the state");
    if ((uint32_t)__builtin_popcountll(x[t] & z[t]) != ny[t])
      return fail(SQ_ERR_INVALID, "ny must equal popcount(xmask & zmask)");
  }
  return SQ_OK;
}

// Device layout of an expectation call: partial sums, term table, per-term
// factors, per-term values.
struct ExpectLayout {
  double *partial;
  void *terms;
  double *alpha;
  double *vals;
};
static size_t expect_partial_bytes() { return (size_t)(kRedBlocks * 64 + 64) * sizeof(double); }
static size_t expect_ws_dev(uint32_t nterms) {
  return expect_partial_bytes() + nterms * (pauli_term_bytes() + 2 * sizeof(double)) + 1024;
}
static ExpectLayout expect_layout(Workspace *w, uint32_t nterms) {
  ExpectLayout L;
  char *d = (char *)w->dev;
  L.partial = (double *)d;
  L.terms = al(d + expect_partial_bytes());
  L.alpha = (double *)al((char *)L.terms + nterms * pauli_term_bytes());
  L.vals = (double *)al((char *)L.alpha + nterms * sizeof(double));
  return L;
}

// Per-term values <P_t> (times coeffs[t]) into L.vals.  Caller holds the
// workspace lock and reserved expect_ws_dev(nterms) device bytes.
static int expect_terms_locked(Workspace *w, const void *sv, sq_dtype_t dt, uint64_t n,
                               uint32_t nterms, const uint64_t *x, const uint64_t *z,
                               const uint32_t *ny, const double *coeffs, hipStream_t st) {
  ExpectLayout L = expect_layout(w, nterms);
  const size_t tb = nterms * pauli_term_bytes(), ab = nterms * sizeof(double);
  char *h = (char *)w->host;
  double *halpha = (double *)al(h + tb);
  fill_pauli_terms((int)nterms, x, z, ny, coeffs, h, halpha);
  SQ_HIP(hipMemcpyAsync(L.terms, h, tb, hipMemcpyHostToDevice, st));
  SQ_HIP(hipMemcpyAsync(L.alpha, halpha, ab, hipMemcpyHostToDevice, st));
  // at most 64 terms per launch (partials hold 64 x kRedBlocks values)
  for (uint32_t t0 = 0; t0 < nterms; t0 += 64) {
    const uint32_t c = std::min<uint32_t>(64, nterms - t0);
    SQ_TRY(launch_pauli_terms(sv, dt, n, (int)c,
                              (const char *)L.terms + t0 * pauli_term_bytes(),
                              L.alpha + t0, L.partial, L.vals + t0, st));
  }
  return SQ_OK;
}

}  // namespace sq

extern "C" {

int sq_expect_pauli_sum(const void *sv, sq_dtype_t dt, uint64_t n, uint32_t nterms,
                        const uint64_t *x, const uint64_t *z, const uint32_t *ny,
                        const double *coeffs, double *out, double *term_values,
                        void *stream) {
  SQ_TRY(check_state(sv, dt, n));
  SQ_TRY(check_terms(nterms, x, z, ny, n));
  if (nterms == 0) {
    *out = 0.0;
    return SQ_OK;
  }
  int dev;
  SQ_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mutex[dev]);
  Workspace *w;
  SQ_TRY(ws_reserve(dev, expect_ws_dev(nterms), nterms * 64 + 4096, &w));
  hipStream_t st = (hipStream_t)stream;
  // per-term values without the coefficient, then weight on the host
  std::vector<double> ones(nterms, 1.0);
  SQ_TRY(expect_terms_locked(w, sv, dt, n, nterms, x, z, ny, ones.data(), st));
  std::vector<double> vals(nterms);
  SQ_HIP(hipMemcpyAsync(w->host, expect_layout(w, nterms).vals, nterms * sizeof(double),
                        hipMemcpyDeviceToHost, st));
  SQ_HIP(hipStreamSynchronize(st));
  std::memcpy(vals.data(), w->host, nterms * sizeof(double));
  double acc = 0.0;
  for (uint32_t t = 0; t < nterms; ++t) acc += (coeffs ? coeffs[t] : 1.0) * vals[t];
  *out = acc;
  if (term_values) std::memcpy(term_values, vals.data(), nterms * sizeof(double));
  return SQ_OK;
}

}  // extern "C"

// ---- the same expectation without allocation --------------------------------
// For callers that cannot allocate (code under hipGraph capture, serving
// loops, hosts that forbid hipMalloc): sq_expect_pauli_sum's kernels with
// their partial sums, term table, term factors and term values in a fixed
// device array instead of the grow-only workspace.  Each GPU holds its own
// copy of the array until the process ends.  Every slot is written before it
// is read (partials by their block, values by the finaliser with beta = 0),
// so nothing needs clearing between calls.
namespace sq {
constexpr uint32_t kNoAllocMaxTerms = 256;
constexpr size_t kNoAllocTermBytes = 64;  // room per term; >= sizeof(PauliTerm)
constexpr size_t kNoAllocScratchBytes =
    (size_t)(kRedBlocks * 64 + 64) * sizeof(double) +
    kNoAllocMaxTerms * (kNoAllocTermBytes + 2 * sizeof(double)) + 4 * 256;
}  // namespace sq

static __device__ __align__(256) char scratch[sq::kNoAllocScratchBytes];

namespace sq {
static std::mutex g_scratch_mutex[64];
static char *g_scratch_dev[64];  // scratch's address on each device, once known

// Caller holds g_scratch_mutex[dev] and has dev current.
static int scratch_base(int dev, char **base) {
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(SQ_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    g_scratch_dev[dev] = (char *)p;
  }
  *base = g_scratch_dev[dev];
  return SQ_OK;
}
}  // namespace sq

extern "C" {

int sq_expect_pauli_sum_noalloc(const void *sv, sq_dtype_t dt, uint64_t n, uint32_t nterms,
                                const uint64_t *x, const uint64_t *z, const uint32_t *ny,
                                const double *coeffs, double *out, double *term_values,
                                void *stream) {
  SQ_TRY(check_state(sv, dt, n));
  if (!out) return fail(SQ_ERR_INVALID, "null output");
  if (nterms > kNoAllocMaxTerms)
    return fail(SQ_ERR_INVALID, "expect_pauli_sum_noalloc: at most " +
                                    std::to_string(kNoAllocMaxTerms) + " terms per call");
  if (nterms && (!x || !z || !ny)) return fail(SQ_ERR_INVALID, "null Pauli masks");
  SQ_TRY(check_terms(nterms, x, z, ny, n));
  if (nterms == 0) {
    *out = 0.0;
    return SQ_OK;
  }
  const size_t tb = pauli_term_bytes();
  if (tb > kNoAllocTermBytes) return fail(SQ_ERR_UNSUPPORTED, "internal: Pauli term too large");
  int dev;
  SQ_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mutex[dev]);
  char *base;
  SQ_TRY(scratch_base(dev, &base));
  double *partial = (double *)base;
  char *terms = al(base + expect_partial_bytes());
  double *alpha = (double *)al(terms + nterms * tb);
  double *vals = (double *)al((char *)(alpha + nterms));
  // host staging on the stack (no heap, no pinned allocation); the call
  // synchronises before these go out of scope
  alignas(16) unsigned char hterms[kNoAllocMaxTerms * kNoAllocTermBytes];
  double ones[kNoAllocMaxTerms], halpha[kNoAllocMaxTerms], hvals[kNoAllocMaxTerms];
  for (uint32_t t = 0; t < nterms; ++t) ones[t] = 1.0;
  fill_pauli_terms((int)nterms, x, z, ny, ones, hterms, halpha);
  hipStream_t st = (hipStream_t)stream;
  SQ_HIP(hipMemcpyAsync(terms, hterms, nterms * tb, hipMemcpyHostToDevice, st));
  SQ_HIP(hipMemcpyAsync(alpha, halpha, nterms * sizeof(double), hipMemcpyHostToDevice, st));
  for (uint32_t t0 = 0; t0 < nterms; t0 += 64) {
    const uint32_t c = std::min<uint32_t>(64, nterms - t0);
    SQ_TRY(launch_pauli_terms(sv, dt, n, (int)c, terms + t0 * tb, alpha + t0, partial,
                              vals + t0, st));
  }
  SQ_HIP(hipMemcpyAsync(hvals, vals, nterms * sizeof(double), hipMemcpyDeviceToHost, st));
  SQ_HIP(hipStreamSynchronize(st));
  double acc = 0.0;
  for (uint32_t t = 0; t < nterms; ++t) acc += (coeffs ? coeffs[t] : 1.0) * hvals[t];
  *out = acc;
  if (term_values) std::memcpy(term_values, hvals, nterms * sizeof(double));
  return SQ_OK;
}

int sq_apply_pauli_sum(const void *in, void *out, sq_dtype_t dt, uint64_t n, uint32_t nterms,
                       const uint64_t *x, const uint64_t *z, const uint32_t *ny,
                       const double *coeffs, void *stream) {
  SQ_TRY(check_state(in, dt, n));
  SQ_TRY(check_state(out, dt, n));
  if (in == out) return fail(SQ_ERR_INVALID, "apply_pauli_sum: out must differ from in");
  SQ_TRY(check_terms(nterms, x, z, ny, n));
  if (!coeffs) return fail(SQ_ERR_INVALID, "null coeffs");
  int dev;
  SQ_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mutex[dev]);
  const size_t tb = nterms * pauli_apply_term_bytes() + 256;
  Workspace *w;
  SQ_TRY(ws_reserve(dev, tb, tb, &w));
  hipStream_t st = (hipStream_t)stream;
  fill_pauli_apply_terms((int)nterms, x, z, ny, coeffs, w->host);
  SQ_HIP(hipMemcpyAsync(w->dev, w->host, tb, hipMemcpyHostToDevice, st));
  SQ_TRY(launch_apply_pauli_sum(in, out, dt, n, (int)nterms, w->dev, st));
  SQ_HIP(hipStreamSynchronize(st));
  return SQ_OK;
}

int sq_sample(const double *probs, uint64_t n, const double *uniforms, uint64_t shots,
              int exact, int64_t *out, void *stream) {
  if (!probs || !is_pow2(n)) return fail(SQ_ERR_INVALID, "bad probabilities");
  if (shots == 0) return SQ_OK;
  if (!uniforms || !out) return fail(SQ_ERR_INVALID, "null uniforms or output");
  int dev;
  SQ_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mutex[dev]);
  const size_t cdf_b = n * sizeof(double) + 256;
  const size_t scr_b = sample_scratch_bytes(n, exact != 0) + 256;
  const size_t io_b = shots * 16 + 512;
  Workspace *w;
  SQ_TRY(ws_reserve(dev, cdf_b + scr_b + io_b + 1024, io_b, &w));
  hipStream_t st = (hipStream_t)stream;
  char *d = (char *)w->dev;
  double *cdf = (double *)d;
  char *scr = al(d + cdf_b);
  double *du = (double *)al(scr + scr_b);
  int64_t *dout = (int64_t *)al((char *)du + shots * 8);
  SQ_TRY(sample_cdf(probs, n, cdf, exact != 0, scr, scr_b, st));
  std::memcpy(w->host, uniforms, shots * 8);
  SQ_HIP(hipMemcpyAsync(du, w->host, shots * 8, hipMemcpyHostToDevice, st));
  SQ_TRY(launch_search(cdf, n, du, shots, dout, st));
  SQ_HIP(hipMemcpyAsync(w->host, dout, shots * 8, hipMemcpyDeviceToHost, st));
  SQ_HIP(hipStreamSynchronize(st));
  std::memcpy(out, w->host, shots * 8);
  return SQ_OK;
}

// ---- adjoint ----------------------------------------------------------------
int sq_adjoint_gradient(void *program, uint32_t num_derivs, const uint32_t *deriv_gate,
                        const uint32_t *deriv_param, const double *dmatrices,
                        uint32_t num_params, void *sv, void *work0, void *work1,
                        void *const *checkpoints, uint32_t num_checkpoints,
                        uint64_t n, uint32_t nterms, const uint64_t *x, const uint64_t *z,
                        const uint32_t *ny, const double *coeffs, double *energy,
                        double *grads, void *stream) {
  Program *p = (Program *)program;
  if (!p) return fail(SQ_ERR_INVALID, "null program");
  const sq_dtype_t dt = p->dt;
  SQ_TRY(check_state(sv, dt, n));
  SQ_TRY(check_state(work0, dt, n));
  SQ_TRY(check_state(work1, dt, n));
  if (n != (1ull << p->nq)) return fail(SQ_ERR_INVALID, "state size does not match program");
  if (sv == work0 || sv == work1 || work0 == work1)
    return fail(SQ_ERR_INVALID, "sv, work0 and work1 must be distinct");
  SQ_TRY(check_terms(nterms, x, z, ny, n));
  if (!coeffs && nterms) return fail(SQ_ERR_INVALID, "null coeffs");
  for (uint32_t c = 0; c < num_checkpoints; ++c)
    SQ_TRY(check_state(checkpoints[c], dt, n));
  const uint32_t ng = (uint32_t)p->fwd.size();

  // derivative plans, grouped by gate
  std::vector<GatePlan> dplans(num_derivs);
  std::vector<std::vector<uint32_t>> by_gate(ng);
  std::vector<size_t> doffs(num_derivs);
  size_t moff = 0;
  for (uint32_t d = 0; d < num_derivs; ++d) {
    doffs[d] = moff;
    const uint32_t g = deriv_gate[d];
    if (g >= ng) return fail(SQ_ERR_INVALID, "deriv_gate out of range");
    if (deriv_param[d] >= num_params) return fail(SQ_ERR_INVALID, "deriv_param out of range");
    const int k = p->nt[g];
    SQ_TRY(make_plan(dmatrices + moff, p->targets.data() + p->toff[g], k, p->nq, p->nq,
                     &dplans[d]));
    moff += 2ull << (2 * k);
    by_gate[g].push_back(d);
  }
  // derivative plans of 4-5 target gates keep their blocks in the workspace
  std::vector<const GatePlan *> dptr;
  for (auto &dp : dplans) dptr.push_back(&dp);
  std::vector<size_t> dblk_off;
  const size_t dblk_b = block_layout(dptr, dt, &dblk_off);
  std::vector<const void *> dmat_ptr(num_derivs, nullptr);

  // checkpoint c holds the state before gate bound[c]; once the reverse runs
  // and layers are planned, each bound moves down to the lowest gate of the
  // run or layer holding it, so no pass is split by a checkpoint
  std::vector<uint32_t> bound(num_checkpoints);
  for (uint32_t c = 0; c < num_checkpoints; ++c)
    bound[c] = (uint32_t)(((uint64_t)(c + 1) * ng) / (num_checkpoints + 1));
  auto gbit = [&](uint32_t g) { return p->nq - 1 - (int)p->targets[p->toff[g]]; };
  auto is_diag = [](const double *m, int k) {
    const int d = 1 << k;
    for (int r = 0; r < d; ++r)
      for (int c = 0; c < d; ++c)
        if (r != c && (m[2 * (r * d + c)] != 0.0 || m[2 * (r * d + c) + 1] != 0.0)) return false;
    return true;
  };
  // G^dagger and dG G^dagger of a one-target gate (row-major complex 2x2)
  auto adj_mats = [](const double *G, const double *D, double *gi, double *dg) {
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 2; ++c) {  // G^dagger[r][c] = conj(G[c][r])
        gi[2 * (2 * r + c)] = G[2 * (2 * c + r)];
        gi[2 * (2 * r + c) + 1] = -G[2 * (2 * c + r) + 1];
      }
    if (!D) return;
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 2; ++c) {  // (dG G^dagger)[r][c]
        double re = 0, im = 0;
        for (int t = 0; t < 2; ++t) {
          const double ar = D[2 * (2 * r + t)], ai = D[2 * (2 * r + t) + 1];
          const double br = gi[2 * (2 * t + c)], bi = gi[2 * (2 * t + c) + 1];
          re += ar * br - ai * bi;
          im += ar * bi + ai * br;
        }
        dg[2 * (2 * r + c)] = re;
        dg[2 * (2 * r + c) + 1] = im;
      }
  };

  // Runs of consecutive diagonal gates (1-2 targets, at most one derivative
  // each, diagonal derivative) are undone in one pass in the reverse sweep
  // (launch_adjoint_diag_run).  run_top[g] = index of the run whose highest
  // gate is g.
  std::vector<DiagRunGate> rgates;
  std::vector<int> rslots;
  struct Run {
    uint32_t lo, first, m, first_slot, nslots;
  };
  std::vector<Run> runs;
  std::vector<int> run_top(ng, -1);
  std::vector<char> in_run(ng, 0);
  if ((n % 1024) == 0) {
    auto diag_ok = [&](uint32_t g) {
      const int k = p->nt[g];
      if (k < 1 || k > 2 || by_gate[g].size() > 1) return false;
      if (!is_diag(p->mats.data() + p->moffs[g], k)) return false;
      return by_gate[g].empty() || is_diag(dmatrices + doffs[by_gate[g][0]], k);
    };
    for (int g = (int)ng - 1; g >= 0;) {
      int lo = g, nsl = 0;
      while (lo >= 0 && diag_ok((uint32_t)lo) && g - lo < kMaxDiagRun &&
             nsl + (int)by_gate[lo].size() <= kMaxDiagSlots) {
        nsl += (int)by_gate[lo].size();
        --lo;
      }
      ++lo;  // lowest gate of the run
      if (g - lo + 1 >= 2) {
        Run r{(uint32_t)lo, (uint32_t)rgates.size(), (uint32_t)(g - lo + 1),
              (uint32_t)rslots.size(), 0};
        for (int q = g; q >= lo; --q) {
          in_run[q] = 1;
          const int k = p->nt[q];
          const int d = 1 << k;
          const double *G = p->mats.data() + p->moffs[q];
          DiagRunGate e{};
          e.b0 = p->nq - 1 - (int)p->targets[p->toff[q]];
          e.b1 = k == 2 ? p->nq - 1 - (int)p->targets[p->toff[q] + 1] : -1;
          e.slot = -1;
          e.uniform = (e.b0 >= 10 && (e.b1 < 0 || e.b1 >= 10)) ? 1 : 0;
          for (int l = 0; l < d; ++l) {  // G^dagger diagonal = conj(G[l][l])
            e.ginv[l][0] = G[2 * (l * d + l)];
            e.ginv[l][1] = -G[2 * (l * d + l) + 1];
          }
          if (!by_gate[q].empty()) {
            const uint32_t di = by_gate[q][0];
            const double *D = dmatrices + doffs[di];
            e.slot = (int)r.nslots++;
            rslots.push_back((int)deriv_param[di]);
            for (int l = 0; l < d; ++l) {  // (dG G^dagger)_ll = dG_ll conj(G_ll)
              const double dr = D[2 * (l * d + l)], dim = D[2 * (l * d + l) + 1];
              const double gr = e.ginv[l][0], gim = e.ginv[l][1];
              e.dgg[l][0] = dr * gr - dim * gim;
              e.dgg[l][1] = dr * gim + dim * gr;
            }
          }
          rgates.push_back(e);
        }
        run_top[g] = (int)runs.size();
        runs.push_back(r);
        g = lo - 1;
      } else {
        g = g - 1;
      }
    }
  }

  // Layers: consecutive one-target gates on distinct qubits (at most one
  // derivative each in the reverse sweep), applied by layer passes
  // (launch_adjoint_layer): the gates on bits 0-11 in one pass, higher bits in
  // windows, one pass each: windows of up to 6 bits beside 6 contiguous low
  // bits (512 B runs) where they need no more passes than windows of up to
  // 8 bits beside 4 (128 B runs).  A layer is taken where it needs fewer
  // passes than the register groups below would.
  struct Win {
    int h0, W, C;
    uint64_t own;  // the layer's bits this window applies
  };
  auto windows_of = [&](uint64_t hi, int wmax, int C, std::vector<Win> *w) {
    while (hi) {
      const int b = __builtin_ctzll(hi);
      const int h0 = std::max(12, std::min(b, (int)p->nq - wmax));
      const int W = std::min(wmax, (int)p->nq - h0);
      const uint64_t cover = ((1ull << W) - 1) << h0;
      w->push_back({h0, W, C, hi & cover});
      hi &= ~cover;
    }
  };
  auto layer_windows = [&](uint64_t bits) {
    std::vector<Win> w, w6;
    if (bits & 0xFFFull) w.push_back({12, 0, 4, bits & 0xFFFull});
    w6 = w;
    windows_of(bits & ~0xFFFull, 8, 4, &w);
    windows_of(bits & ~0xFFFull, 6, 6, &w6);
    return tuning().adjoint_wide_windows && w6.size() <= w.size() ? w6 : w;
  };
  struct LPass {
    int h0, W, C, count;
    int pos[12], slot[12];
    double gm[96], dm[96];
    uint32_t first_slot, nslots;
  };
  struct Layer {
    uint32_t lo, m, first_pass, npass;
  };
  // one pass per window: reverse (G^dagger, dG G^dagger, slots) or forward (G)
  auto make_passes = [&](uint32_t lo, uint32_t m, uint64_t used, bool rev,
                         std::vector<LPass> *out) {
    for (const Win &w : layer_windows(used)) {
      LPass ps{};
      ps.h0 = w.h0;
      ps.W = w.W;
      ps.C = w.C;
      ps.first_slot = (uint32_t)rslots.size();
      for (uint32_t q = lo; q < lo + m; ++q) {
        const int b = gbit(q);
        if (!((w.own >> b) & 1)) continue;
        const int j = ps.count++;
        ps.pos[j] = w.W == 0 ? b : w.C + (b - w.h0);
        ps.slot[j] = -1;
        const double *G = p->mats.data() + p->moffs[q];
        if (!rev) {
          for (int e = 0; e < 8; ++e) ps.gm[8 * j + e] = G[e];
          continue;
        }
        const double *D = nullptr;
        if (!by_gate[q].empty()) {
          const uint32_t di = by_gate[q][0];
          D = dmatrices + doffs[di];
          ps.slot[j] = (int)ps.nslots++;
          rslots.push_back((int)deriv_param[di]);
        }
        adj_mats(G, D, ps.gm + 8 * j, ps.dm + 8 * j);
      }
      out->push_back(ps);
    }
  };
  const bool use_layers = tuning().adjoint_layers && n >= 4096 && (n % 4096) == 0;
  std::vector<LPass> lpasses;
  std::vector<Layer> layers;
  std::vector<int> layer_top(ng, -1);
  std::vector<char> in_layer(ng, 0);
  if (use_layers) {
    auto ok = [&](int q) { return !in_run[q] && p->nt[q] == 1 && by_gate[q].size() <= 1; };
    for (int g = (int)ng - 1; g >= 0;) {
      if (!ok(g)) {
        --g;
        continue;
      }
      int lo = g;
      uint64_t used = 0;
      int hi1 = 0, lo1 = 0, none = 0;
      while (lo >= 0 && ok(lo) && !((used >> gbit(lo)) & 1)) {
        used |= 1ull << gbit(lo);
        if (by_gate[lo].empty())
          ++none;
        else if (gbit(lo) >= 4)
          ++hi1;
        else
          ++lo1;
        --lo;
      }
      ++lo;
      const uint32_t m = (uint32_t)(g - lo + 1);
      const int est = (hi1 + 3) / 4 + (lo1 + 3) / 4 + none;  // register groups + single steps
      if (m >= 2 && (int)layer_windows(used).size() < est) {
        Layer L{(uint32_t)lo, m, (uint32_t)lpasses.size(), 0};
        make_passes((uint32_t)lo, m, used, true, &lpasses);
        L.npass = (uint32_t)lpasses.size() - L.first_pass;
        for (int q = lo; q <= g; ++q) in_layer[q] = 1;
        layer_top[g] = (int)layers.size();
        layers.push_back(L);
      }
      g = lo - 1;
    }
  }
  // Diagonal runs folded into the low pass of the layer right below them
  // (launch_adjoint_layer with a FuseDesc).  Reverse: the run's derivative
  // terms by a run pass that leaves both states unchanged, then the layer's
  // low pass undoes the run's phases before the layer's gates and goes first
  // among the layer's passes; forward: the low pass applies the run's phases
  // after the layer's gates and goes last.  Needs a low pass with gates in
  // all three register layouts and at most kFuseMix run gates with one bit
  // on each side of bit 12.
  auto low_pass_of = [&](const std::vector<LPass> &ps, const Layer &L) -> int {
    for (uint32_t i = L.first_pass; i < L.first_pass + L.npass; ++i) {
      if (ps[i].W != 0) continue;
      int seen = 0;
      for (int j = 0; j < ps[i].count; ++j) seen |= 1 << ps[i].pos[j];
      return ((seen & 0xF) && (seen & 0xF0) && (seen & 0xF00)) ? (int)i : -1;
    }
    return -1;
  };
  auto run_mixed = [&](const Run &r) {
    std::vector<DiagRunGate> mix;
    for (uint32_t k = r.first; k < r.first + r.m; ++k) {
      const DiagRunGate &e = rgates[k];
      if (e.b1 >= 0 && (e.b0 < 12) != (e.b1 < 12)) mix.push_back(e);
    }
    return mix;
  };
  std::vector<int> rfuse(runs.size(), -1), rfuse_low(runs.size(), -1);
  if (use_layers && (tuning().adjoint_fuse_runs & 2)) {
    std::vector<int> layer_below(ng + 1, -1);  // layer whose last gate is key - 1
    for (size_t L = 0; L < layers.size(); ++L)
      layer_below[layers[L].lo + layers[L].m] = (int)L;
    for (size_t ri = 0; ri < runs.size(); ++ri) {
      const int L = layer_below[runs[ri].lo];
      if (L < 0) continue;
      const int lp = low_pass_of(lpasses, layers[L]);
      if (lp < 0 || run_mixed(runs[ri]).size() > (size_t)kFuseMix) continue;
      rfuse[ri] = L;
      rfuse_low[ri] = lp;
    }
  }
  // bounds to the lowest gate of their run or layer (of the layer for a run
  // folded into it)
  {
    std::vector<uint32_t> seg_lo(ng);
    for (uint32_t g = 0; g < ng; ++g) seg_lo[g] = g;
    for (size_t ri = 0; ri < runs.size(); ++ri) {
      const Run &r = runs[ri];
      const uint32_t lo = rfuse[ri] >= 0 ? layers[rfuse[ri]].lo : r.lo;
      for (uint32_t q = r.lo; q < r.lo + r.m; ++q) seg_lo[q] = lo;
    }
    for (const Layer &L : layers)
      for (uint32_t q = L.lo; q < L.lo + L.m; ++q) seg_lo[q] = L.lo;
    for (uint32_t c = 0; c < num_checkpoints; ++c) bound[c] = seg_lo[bound[c]];
  }
  std::vector<char> is_bound(ng, 0);
  for (uint32_t c = 0; c < num_checkpoints; ++c) is_bound[bound[c]] = 1;

  // Groups of 2-4 consecutive one-target gates with one derivative each, on
  // distinct amplitude bits, outside the runs and layers and not at a
  // checkpoint, take one reverse pass (launch_adjoint_step_group): bits all
  // above 3 (a lane holds each group of amplitudes) or all below 4 (pairs
  // across the lanes of a wave; states of at least 64 amplitudes).
  // grp_top[g] = group whose highest gate is g.
  struct Group {
    uint32_t lo, k, first_slot;
    int bits[5];
    double ginv[40], dgg[40];
  };
  std::vector<Group> groups;
  std::vector<int> grp_top(ng, -1);
  {
    auto grp_ok = [&](int g) {
      return !in_run[g] && !in_layer[g] && !is_bound[g] && p->nt[g] == 1 &&
             by_gate[g].size() == 1 && (gbit(g) >= 4 || p->nq >= 6);
    };
    for (int g = (int)ng - 1; g >= 0;) {
      int lo = g;
      uint64_t used = 0;
      while (lo >= 0 && g - lo < 4 && grp_ok(lo)) {
        const int b = gbit(lo);
        if ((used >> b) & 1) break;
        if (lo < g && (b >= 4) != (gbit(g) >= 4)) break;  // one kind per group
        used |= 1ull << b;
        --lo;
      }
      ++lo;
      if (g - lo + 1 < 2) {
        --g;
        continue;
      }
      Group gr{};
      gr.lo = (uint32_t)lo;
      gr.k = (uint32_t)(g - lo + 1);
      gr.first_slot = (uint32_t)rslots.size();
      for (int j = 0; j < (int)gr.k; ++j) {
        const int q = g - j;
        const uint32_t di = by_gate[q][0];
        gr.bits[j] = gbit(q);
        rslots.push_back((int)deriv_param[di]);
        adj_mats(p->mats.data() + p->moffs[q], dmatrices + doffs[di], gr.ginv + 8 * j,
                 gr.dgg + 8 * j);
      }
      grp_top[g] = (int)groups.size();
      groups.push_back(gr);
      g = lo - 1;
    }
  }

  // Forward passes never hold a checkpoint bound but as their first gate
  // (the copy is taken before the pass).  Forward runs: the same gates,
  // multiplied by their own diagonals; fwd_start[g] = run starting at g.
  std::vector<int> fwd_start(ng, -1);
  std::vector<Run> fruns;
  std::vector<char> in_frun(ng, 0);
  if ((n % 1024) == 0) {
    auto is_diag_g = [&](uint32_t g) {
      const int k = p->nt[g];
      return k >= 1 && k <= 2 && is_diag(p->mats.data() + p->moffs[g], k);
    };
    for (uint32_t g = 0; g < ng;) {
      uint32_t hi = g;
      while (hi < ng && is_diag_g(hi) && (hi == g || !is_bound[hi]) &&
             hi - g < (uint32_t)kMaxDiagRun)
        ++hi;
      if (hi - g >= 2) {
        Run r{g, (uint32_t)rgates.size(), hi - g, 0, 0};
        for (uint32_t q = g; q < hi; ++q) {
          in_frun[q] = 1;
          const int k = p->nt[q];
          const int d = 1 << k;
          const double *G = p->mats.data() + p->moffs[q];
          DiagRunGate e{};
          e.b0 = p->nq - 1 - (int)p->targets[p->toff[q]];
          e.b1 = k == 2 ? p->nq - 1 - (int)p->targets[p->toff[q] + 1] : -1;
          e.slot = -1;
          e.uniform = (e.b0 >= 10 && (e.b1 < 0 || e.b1 >= 10)) ? 1 : 0;
          for (int l = 0; l < d; ++l) {  // the gate's own diagonal
            e.ginv[l][0] = G[2 * (l * d + l)];
            e.ginv[l][1] = G[2 * (l * d + l) + 1];
          }
          rgates.push_back(e);
        }
        fwd_start[g] = (int)fruns.size();
        fruns.push_back(r);
        g = hi;
      } else {
        ++g;
      }
    }
  }
  // forward layers (as the reverse ones; gates with any number of
  // derivatives), taken where they need fewer passes than forward groups
  std::vector<LPass> fpasses;
  std::vector<Layer> flayers;
  std::vector<int> flayer_start(ng, -1);
  std::vector<char> in_flayer(ng, 0);
  if (use_layers) {
    auto ok = [&](uint32_t q) { return !in_frun[q] && p->nt[q] == 1; };
    for (uint32_t g = 0; g < ng;) {
      if (!ok(g)) {
        ++g;
        continue;
      }
      uint32_t hi = g;
      uint64_t used = 0;
      int nhi = 0, nlo = 0;
      while (hi < ng && ok(hi) && (hi == g || !is_bound[hi]) && !((used >> gbit(hi)) & 1)) {
        used |= 1ull << gbit(hi);
        (gbit(hi) >= 4 ? nhi : nlo)++;
        ++hi;
      }
      const uint32_t m = hi - g;
      const int est = (nhi + 4) / 5 + (nlo + 4) / 5;  // forward groups of up to 5
      if (m >= 2 && (int)layer_windows(used).size() < est) {
        Layer L{g, m, (uint32_t)fpasses.size(), 0};
        make_passes(g, m, used, false, &fpasses);
        L.npass = (uint32_t)fpasses.size() - L.first_pass;
        for (uint32_t q = g; q < hi; ++q) in_flayer[q] = 1;
        flayer_start[g] = (int)flayers.size();
        flayers.push_back(L);
      }
      g = hi;
    }
  }
  // forward: a run right after a layer (no checkpoint bound at its start)
  std::vector<int> ffuse(flayers.size(), -1), ffuse_low(flayers.size(), -1);
  if (use_layers && (tuning().adjoint_fuse_runs & 1)) {
    for (size_t fi = 0; fi < flayers.size(); ++fi) {
      const uint32_t next = flayers[fi].lo + flayers[fi].m;
      if (next >= ng || fwd_start[next] < 0 || is_bound[next]) continue;
      const int lp = low_pass_of(fpasses, flayers[fi]);
      if (lp < 0 || run_mixed(fruns[fwd_start[next]]).size() > (size_t)kFuseMix) continue;
      ffuse[fi] = fwd_start[next];
      ffuse_low[fi] = lp;
    }
  }
  // forward groups: 2-5 consecutive one-target gates on distinct bits, all
  // above 3 or all below 4 as in the reverse groups (outside forward runs and
  // layers), applied in one pass (5: psi alone leaves the registers for 32
  // amplitudes per lane; the reverse holds psi and lambda and stops at 4)
  std::vector<Group> fgroups;
  std::vector<int> fgrp_start(ng, -1);
  {
    auto fok = [&](uint32_t g) {
      return !in_frun[g] && !in_flayer[g] && p->nt[g] == 1 && (gbit(g) >= 4 || p->nq >= 6);
    };
    for (uint32_t g = 0; g < ng;) {
      uint32_t hi = g;
      uint64_t used = 0;
      while (hi < ng && hi - g < 5 && fok(hi) && (hi == g || !is_bound[hi])) {
        const int b = gbit(hi);
        if ((used >> b) & 1) break;
        if (hi > g && (b >= 4) != (gbit(g) >= 4)) break;  // one kind per group
        used |= 1ull << b;
        ++hi;
      }
      if (hi - g < 2) {
        ++g;
        continue;
      }
      Group gr{};
      gr.lo = g;
      gr.k = hi - g;
      for (uint32_t j = 0; j < gr.k; ++j) {
        const uint32_t q = g + j;
        gr.bits[j] = gbit(q);
        const double *G = p->mats.data() + p->moffs[q];
        for (int e = 0; e < 8; ++e) gr.ginv[8 * j + e] = G[e];  // the gate itself
      }
      fgrp_start[g] = (int)fgroups.size();
      fgroups.push_back(gr);
      g = hi;
    }
  }
  const bool any_runs = !runs.empty() || !fruns.empty() || !groups.empty() || !layers.empty();
  auto distinct_params = [&](uint32_t first, uint32_t count) {
    for (uint32_t a = first; a < first + count; ++a)
      for (uint32_t b = a + 1; b < first + count; ++b)
        if (rslots[a] == rslots[b]) return false;
    return true;
  };
  const size_t rg_b = (rgates.size() * sizeof(DiagRunGate) + 255) / 256 * 256;
  const size_t rs_b = (rslots.size() * sizeof(int) + 255) / 256 * 256;
  const size_t rp_b = any_runs ? (size_t)kMaxDiagSlots * 4096 * sizeof(double) : 0;
  // fused runs: a FuseDesc and a FuseTableJob each, and their phase tables
  std::vector<int> rfuse_id(runs.size(), -1), ffuse_id(flayers.size(), -1);
  std::vector<const Run *> fused_runs;
  for (size_t ri = 0; ri < runs.size(); ++ri)
    if (rfuse[ri] >= 0) {
      rfuse_id[ri] = (int)fused_runs.size();
      fused_runs.push_back(&runs[ri]);
    }
  for (size_t fi = 0; fi < flayers.size(); ++fi)
    if (ffuse[fi] >= 0) {
      ffuse_id[fi] = (int)fused_runs.size();
      fused_runs.push_back(&fruns[ffuse[fi]]);
    }
  const size_t nfuse = fused_runs.size();
  // Deferred finalisation: the reverse passes with derivative slots (runs,
  // layer passes, groups) write their partial rows to their own region of
  // one buffer, and one launch at the end adds each parameter's rows in pass
  // order (instead of a finalisation launch per pass).
  size_t drow_doubles = 0, nrows = 0;
  for (const Run &r : runs) {
    drow_doubles += (size_t)r.nslots * diag_run_blocks(n);
    nrows += r.nslots;
  }
  for (const LPass &ps : lpasses) {
    drow_doubles += (size_t)ps.nslots * layer_pass_blocks(n);
    nrows += ps.nslots;
  }
  for (const Group &gr : groups) {
    drow_doubles += (size_t)gr.k * step_group_blocks(n, (int)gr.k, gr.bits);
    nrows += gr.k;
  }
  const bool defer = nrows > 0 && drow_doubles <= (64ull << 20);  // rows of at most 512 MB
  const size_t dp_b = defer ? (drow_doubles * sizeof(double) + 255) / 256 * 256 : 0;
  const size_t rowmeta_b =
      defer ? (nrows * (sizeof(uint64_t) + sizeof(int)) + (num_params + 1) * sizeof(int) + 1024 +
               255) / 256 * 256
            : 0;
  const size_t fz_b = (nfuse * (sizeof(FuseDesc) + sizeof(FuseTableJob)) + 255) / 256 * 256;
  const size_t tab_stride = ((4096 + (n >> 12)) * itemsize(dt) + 255) / 256 * 256;
  const size_t tab_b = nfuse ? nfuse * tab_stride : 0;

  int dev;
  SQ_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mutex[dev]);
  const size_t grad_b = ((2 * (size_t)num_params + 64) * sizeof(double) + 255) / 256 * 256;
  const size_t edev = expect_ws_dev(nterms);
  const size_t apb = nterms * pauli_apply_term_bytes() + 256;
  Workspace *w;
  SQ_TRY(ws_reserve(dev,
                    edev + grad_b + apb + rg_b + rs_b + rp_b + dblk_b + fz_b + tab_b + dp_b +
                        rowmeta_b + 4096,
                    nterms * 64 + grad_b + apb + rg_b + rs_b + dblk_b + fz_b + rowmeta_b + 4096, &w));
  hipStream_t st = (hipStream_t)stream;
  char *dbase = (char *)w->dev;
  double *partial = (double *)dbase;                      // reduction partials
  double *gslots = (double *)al(dbase + edev);            // 2 per param + energy
  void *aterms = al((char *)gslots + grad_b);             // Pauli apply terms
  double *two = gslots + 2 * num_params;                  // {2, 0}
  double *eslot = two + 2;                                // energy (re, im)
  DiagRunGate *rgates_dev = (DiagRunGate *)al((char *)aterms + apb);
  int *rslots_dev = (int *)al((char *)rgates_dev + rg_b);
  double *rpartial = (double *)al((char *)rslots_dev + rs_b);
  char *dblk_dev = al((char *)rpartial + rp_b);           // derivative blocks
  FuseDesc *fdesc_dev = (FuseDesc *)al(dblk_dev + dblk_b);
  FuseTableJob *fjob_dev = (FuseTableJob *)(fdesc_dev + nfuse);
  char *tab_dev = al((char *)fdesc_dev + fz_b);
  double *dpart_dev = (double *)al(tab_dev + tab_b);       // deferred partial rows
  char *rowmeta_dev = al((char *)dpart_dev + dp_b);
  // staged through the pinned host buffer, after the gradient slots
  char *hs = (char *)al((char *)w->host + apb + grad_b);
  if (any_runs) {
    std::memcpy(hs, rgates.data(), rgates.size() * sizeof(DiagRunGate));
    std::memcpy(hs + rg_b, rslots.data(), rslots.size() * sizeof(int));
    SQ_HIP(hipMemcpyAsync(rgates_dev, hs, rg_b + rs_b, hipMemcpyHostToDevice, st));
  }
  if (dblk_b) {
    char *hb = al(hs + rg_b + rs_b);
    block_image(dptr, dt, dblk_off, hb);
    SQ_HIP(hipMemcpyAsync(dblk_dev, hb, dblk_b, hipMemcpyHostToDevice, st));
    for (uint32_t d = 0; d < num_derivs; ++d)
      if (dblk_off[d] != (size_t)-1) dmat_ptr[d] = dblk_dev + dblk_off[d];
  }
  if (nfuse) {
    char *hf = al(hs + rg_b + rs_b + dblk_b);
    FuseDesc *hd = (FuseDesc *)hf;
    FuseTableJob *hj = (FuseTableJob *)(hd + nfuse);
    for (size_t f = 0; f < nfuse; ++f) {
      const Run &r = *fused_runs[f];
      char *tl = tab_dev + f * tab_stride;
      char *tc = tl + 4096 * itemsize(dt);
      FuseDesc d{};
      d.tl = tl;
      d.tc = tc;
      const std::vector<DiagRunGate> mix = run_mixed(r);
      d.nmix = (int)mix.size();
      for (size_t k = 0; k < mix.size(); ++k) d.mix[k] = mix[k];
      std::memcpy(hd + f, &d, sizeof d);
      FuseTableJob j{};
      j.g = rgates_dev + r.first;
      j.m = (int)r.m;
      j.tl = tl;
      j.tc = tc;
      std::memcpy(hj + f, &j, sizeof j);
    }
    SQ_HIP(hipMemcpyAsync(fdesc_dev, hf, fz_b, hipMemcpyHostToDevice, st));
    SQ_TRY(launch_fuse_tables(fjob_dev, (int)nfuse, dt, n, st));
  }

  // forward pass with checkpoints
  SQ_TRY(launch_init(sv, dt, n, true, st));
  uint32_t cnext = 0;
  for (uint32_t g = 0; g < ng; ++g) {
    while (cnext < num_checkpoints && bound[cnext] == g) {
      SQ_HIP(hipMemcpyAsync(checkpoints[cnext], sv, n * itemsize(dt), hipMemcpyDeviceToDevice, st));
      ++cnext;
    }
    if (fwd_start[g] >= 0) {
      const Run &r = fruns[fwd_start[g]];
      SQ_TRY(launch_adjoint_diag_run(sv, nullptr, dt, n, rgates_dev + r.first, (int)r.m, 0,
                                     nullptr, false, rpartial, nullptr, st));
      g = r.lo + r.m - 1;  // the loop steps past the run
      continue;
    }
    if (flayer_start[g] >= 0) {
      const int fi = flayer_start[g];
      const Layer &L = flayers[fi];
      for (uint32_t i = L.first_pass; i < L.first_pass + L.npass; ++i) {
        if ((int)i == ffuse_low[fi]) continue;  // the fused low pass goes last
        const LPass &ps = fpasses[i];
        SQ_TRY(launch_adjoint_layer(sv, nullptr, dt, n, ps.h0, ps.W, ps.C, ps.count, ps.pos, ps.gm,
                                    nullptr, nullptr, 0, nullptr, false, nullptr, nullptr, st));
      }
      if (ffuse[fi] >= 0) {
        const LPass &ps = fpasses[ffuse_low[fi]];
        SQ_TRY(launch_adjoint_layer(sv, nullptr, dt, n, ps.h0, ps.W, ps.C, ps.count, ps.pos, ps.gm,
                                    nullptr, nullptr, 0, nullptr, false, nullptr, nullptr, st,
                                    fdesc_dev + ffuse_id[fi]));
        const Run &r = fruns[ffuse[fi]];
        g = r.lo + r.m - 1;  // the loop steps past the run
      } else {
        g = L.lo + L.m - 1;
      }
      continue;
    }
    if (fgrp_start[g] >= 0) {
      const Group &gr = fgroups[fgrp_start[g]];
      SQ_TRY(launch_adjoint_step_group(sv, nullptr, dt, n, (int)gr.k, gr.bits, gr.ginv, gr.dgg,
                                       nullptr, false, rpartial, nullptr, st));
      g = gr.lo + gr.k - 1;
      continue;
    }
    SQ_TRY(launch_plan(p->fwd[g], sv, dt, n, st, p->fwd_dev[g]));
  }
  // lambda = H psi ; energy = <psi|lambda>
  fill_pauli_apply_terms((int)nterms, x, z, ny, coeffs, w->host);
  SQ_HIP(hipMemcpyAsync(aterms, w->host, nterms * pauli_apply_term_bytes(),
                        hipMemcpyHostToDevice, st));
  SQ_TRY(launch_apply_pauli_sum(sv, work0, dt, n, (int)nterms, aterms, st));
  {
    double *hg = (double *)al((char *)w->host + apb);
    std::memset(hg, 0, grad_b);
    hg[2 * num_params] = 2.0;
    hg[2 * num_params + 1] = 0.0;
    SQ_HIP(hipMemcpyAsync(gslots, hg, grad_b, hipMemcpyHostToDevice, st));
  }
  SQ_TRY(launch_inner(sv, work0, dt, n, partial, eslot, nullptr, 0.0, st));
  // reverse sweep, on `cur`: psi's buffer until the first checkpoint bound
  // is reached, then that checkpoint's buffer, which holds psi at that point
  // (a switch of buffers instead of a copy back; each checkpoint is reached
  // once, and sv, the checkpoints and the work states are scratch after the
  // call)
  void *cur = sv;
  int cidx = (int)num_checkpoints - 1;
  // after a run or layer whose lowest gate is a checkpoint bound, psi is
  // the checkpoint (the state before that gate)
  auto restore_below = [&](uint32_t lo) -> int {
    if (cidx >= 0 && bound[cidx] == lo) {
      cur = checkpoints[cidx];
      while (cidx >= 0 && bound[cidx] == lo) --cidx;
    }
    return SQ_OK;
  };
  // where the next slot-carrying pass writes its rows, and the rows so far
  size_t dcur = 0;
  std::vector<uint64_t> row_off;
  std::vector<int> row_nb, row_param;
  auto rows_at = [&](uint32_t nslots, int nb, uint32_t first_slot) -> double * {
    if (!defer) return rpartial;
    double *at = dpart_dev + dcur;
    for (uint32_t sl = 0; sl < nslots; ++sl) {
      row_off.push_back(dcur + (size_t)sl * nb);
      row_nb.push_back(nb);
      row_param.push_back(rslots[first_slot + sl]);
    }
    dcur += (size_t)nslots * nb;
    return at;
  };
  double *fin = defer ? nullptr : gslots;  // per-pass finalisation target
  for (int g = (int)ng - 1; g >= 0; --g) {
    if (run_top[g] >= 0 && rfuse[run_top[g]] >= 0) {
      const int ri = run_top[g];
      const Run &r = runs[ri];
      if (r.nslots)  // derivative terms only; the low pass undoes the phases
        SQ_TRY(launch_adjoint_diag_run(
            cur, work0, dt, n, rgates_dev + r.first, (int)r.m, (int)r.nslots,
            rslots_dev + r.first_slot, distinct_params(r.first_slot, r.nslots),
            rows_at(r.nslots, diag_run_blocks(n), r.first_slot), fin, st, false));
      const Layer &L = layers[rfuse[ri]];
      auto pass = [&](uint32_t i, const FuseDesc *fz) {
        const LPass &ps = lpasses[i];
        return launch_adjoint_layer(
            cur, work0, dt, n, ps.h0, ps.W, ps.C, ps.count, ps.pos, ps.gm, ps.dm, ps.slot,
            (int)ps.nslots, rslots_dev + ps.first_slot, distinct_params(ps.first_slot, ps.nslots),
            rows_at(ps.nslots, layer_pass_blocks(n), ps.first_slot), fin, st, fz);
      };
      SQ_TRY(pass((uint32_t)rfuse_low[ri], fdesc_dev + rfuse_id[ri]));
      for (uint32_t i = L.first_pass; i < L.first_pass + L.npass; ++i)
        if ((int)i != rfuse_low[ri]) SQ_TRY(pass(i, nullptr));
      SQ_TRY(restore_below(L.lo));
      g = (int)L.lo;
      continue;
    }
    if (run_top[g] >= 0) {
      const Run &r = runs[run_top[g]];
      SQ_TRY(launch_adjoint_diag_run(cur, work0, dt, n, rgates_dev + r.first, (int)r.m,
                                     (int)r.nslots, rslots_dev + r.first_slot,
                                     distinct_params(r.first_slot, r.nslots),
                                     rows_at(r.nslots, diag_run_blocks(n), r.first_slot), fin,
                                     st));
      SQ_TRY(restore_below(r.lo));
      g = (int)r.lo;  // the loop steps below the run
      continue;
    }
    if (layer_top[g] >= 0) {
      const Layer &L = layers[layer_top[g]];
      for (uint32_t i = L.first_pass; i < L.first_pass + L.npass; ++i) {
        const LPass &ps = lpasses[i];
        SQ_TRY(launch_adjoint_layer(
            cur, work0, dt, n, ps.h0, ps.W, ps.C, ps.count, ps.pos, ps.gm, ps.dm, ps.slot,
            (int)ps.nslots, rslots_dev + ps.first_slot, distinct_params(ps.first_slot, ps.nslots),
            rows_at(ps.nslots, layer_pass_blocks(n), ps.first_slot), fin, st));
      }
      SQ_TRY(restore_below(L.lo));
      g = (int)L.lo;
      continue;
    }
    if (grp_top[g] >= 0) {
      const Group &gr = groups[grp_top[g]];
      SQ_TRY(launch_adjoint_step_group(
          cur, work0, dt, n, (int)gr.k, gr.bits, gr.ginv, gr.dgg, rslots_dev + gr.first_slot,
          distinct_params(gr.first_slot, gr.k),
          rows_at(gr.k, step_group_blocks(n, (int)gr.k, gr.bits), gr.first_slot), fin, st));
      g = (int)gr.lo;
      continue;
    }
    bool restored = false;
    if (cidx >= 0 && bound[cidx] == (uint32_t)g) {
      cur = checkpoints[cidx];
      while (cidx >= 0 && bound[cidx] == (uint32_t)g) --cidx;
      restored = true;
    }
    if (p->nt[g] == 1 && by_gate[g].size() == 1) {
      // one pass: undo G on psi and lambda, derivative inner product between
      const uint32_t d = by_gate[g][0];
      const double *G = p->mats.data() + p->moffs[g];
      double gi[8];
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) {  // G^dagger[r][c] = conj(G[c][r])
          gi[2 * (2 * r + c)] = G[2 * (2 * c + r)];
          gi[2 * (2 * r + c) + 1] = -G[2 * (2 * c + r) + 1];
        }
      const size_t doff = doffs[d];
      const int bit = p->nq - 1 - (int)p->targets[p->toff[g]];
      SQ_TRY(launch_adjoint_step1(cur, work0, dt, n, bit, gi, dmatrices + doff, restored ? 0 : 1,
                                  partial, gslots + 2 * deriv_param[d], two, st));
      continue;
    }
    if (!restored) SQ_TRY(launch_plan(p->inv[g], cur, dt, n, st, p->inv_dev[g]));
    for (uint32_t d : by_gate[g]) {
      SQ_HIP(hipMemcpyAsync(work1, cur, n * itemsize(dt), hipMemcpyDeviceToDevice, st));
      SQ_TRY(launch_plan(dplans[d], work1, dt, n, st, dmat_ptr[d]));
      // gslots[2p] += 2 Re <lambda|work1>
      SQ_TRY(launch_inner(work0, work1, dt, n, partial, gslots + 2 * deriv_param[d], two,
                          1.0, st));
    }
    if (g > 0) SQ_TRY(launch_plan(p->inv[g], work0, dt, n, st, p->inv_dev[g]));
  }
  if (defer && !row_off.empty()) {
    if (dcur > drow_doubles || row_off.size() > nrows)
      return fail(SQ_ERR_INVALID, "internal: deferred rows exceed their plan");
    // rows grouped by parameter, pass order kept inside each group
    const size_t nr = row_off.size();
    std::vector<int> ptr(num_params + 1, 0), order(nr);
    for (size_t i = 0; i < nr; ++i) ++ptr[row_param[i] + 1];
    for (uint32_t q = 0; q < num_params; ++q) ptr[q + 1] += ptr[q];
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (size_t i = 0; i < nr; ++i) order[fill[row_param[i]]++] = (int)i;
    char *hm = al(hs + rg_b + rs_b + dblk_b + fz_b);
    uint64_t *h_off = (uint64_t *)hm;
    int *h_nb = (int *)(h_off + nr);
    int *h_ptr = h_nb + nr;
    for (size_t i = 0; i < nr; ++i) {
      h_off[i] = row_off[order[i]];
      h_nb[i] = row_nb[order[i]];
    }
    std::memcpy(h_ptr, ptr.data(), (num_params + 1) * sizeof(int));
    const size_t mb = nr * (sizeof(uint64_t) + sizeof(int)) + (num_params + 1) * sizeof(int);
    SQ_HIP(hipMemcpyAsync(rowmeta_dev, hm, mb, hipMemcpyHostToDevice, st));
    const uint64_t *d_off = (const uint64_t *)rowmeta_dev;
    const int *d_nb = (const int *)(d_off + nr);
    SQ_TRY(launch_finalize_rows(dpart_dev, d_off, d_nb, d_nb + nr, (int)num_params, 2.0, gslots,
                                st));
  }
  double *hg = (double *)al((char *)w->host + apb);
  SQ_HIP(hipMemcpyAsync(hg, gslots, grad_b, hipMemcpyDeviceToHost, st));
  SQ_HIP(hipStreamSynchronize(st));
  for (uint32_t q = 0; q < num_params; ++q) grads[q] = hg[2 * q];
  *energy = hg[2 * num_params + 2];
  return SQ_OK;
}

// ---- batched QAOA --------------------------------------------------------------
int sq_qaoa_create(void **batch, sq_dtype_t dt, uint32_t ninst, const uint32_t *nq,
                   const uint32_t *nedges, const uint32_t *edges, uint32_t p) {
  if (!batch || !nq || !nedges) return fail(SQ_ERR_INVALID, "null argument");
  *batch = nullptr;
  if (!valid_dtype(dt)) return fail(SQ_ERR_INVALID, "bad dtype");
  return qaoa_create(batch, dt, ninst, nq, nedges, edges, p);
}
int sq_qaoa_run(void *batch, const double *gammas, const double *betas, int use_graph,
                double *out, void *stream) {
  if (!batch || !gammas || !betas || !out) return fail(SQ_ERR_INVALID, "null argument");
  return qaoa_run(batch, gammas, betas, use_graph, out, (hipStream_t)stream);
}
int sq_qaoa_stats(void *batch, uint64_t *launches, double *state_bytes, double *traffic_bytes) {
  if (!batch) return fail(SQ_ERR_INVALID, "null batch");
  return qaoa_stats(batch, launches, state_bytes, traffic_bytes);
}
int sq_qaoa_destroy(void *batch) { return qaoa_destroy(batch); }

// ---- circuits of uneven qubit counts ------------------------------------------
int sq_cbatch_create(void **batch, sq_dtype_t dt, uint32_t ncirc, const uint32_t *nq,
                     const uint32_t *ng, const uint32_t *nt, const uint32_t *targets,
                     const double *matrices) {
  if (!batch || !nq || !ng) return fail(SQ_ERR_INVALID, "null argument");
  *batch = nullptr;
  if (!valid_dtype(dt)) return fail(SQ_ERR_INVALID, "bad dtype");
  uint64_t total = 0;
  for (uint32_t j = 0; j < ncirc; ++j) total += ng[j];
  if (total && (!nt || !targets || !matrices)) return fail(SQ_ERR_INVALID, "null argument");
  return cbatch_create(batch, dt, ncirc, nq, ng, nt, targets, matrices);
}
int sq_cbatch_run(void *batch, int use_graph, void *stream) {
  if (!batch) return fail(SQ_ERR_INVALID, "null batch");
  return cbatch_run(batch, use_graph, (hipStream_t)stream);
}
int sq_cbatch_copy_state(void *batch, uint32_t circuit, void *dst, uint64_t num_elems,
                         void *stream) {
  if (!batch || !dst) return fail(SQ_ERR_INVALID, "null argument");
  return cbatch_copy_state(batch, circuit, dst, num_elems, (hipStream_t)stream);
}
int sq_cbatch_stats(void *batch, uint64_t *launches, double *state_bytes, double *traffic_bytes) {
  if (!batch) return fail(SQ_ERR_INVALID, "null batch");
  return cbatch_stats(batch, launches, state_bytes, traffic_bytes);
}
int sq_cbatch_destroy(void *batch) { return cbatch_destroy(batch); }

}  // extern "C"

// ---- parameter batches ---------------------------------------------------------
// One circuit structure, many parameter sets: nsets states of 2^nq amplitudes
// in one device allocation (state y at y * 2^nq), each evolved by the same
// gates with its own matrices, one launch per gate for the whole batch.
namespace sq {

struct Batch {
  sq_dtype_t dt;
  int nq;
  uint32_t nsets;
  int device;
  std::vector<int> nt;
  std::vector<size_t> toff;
  std::vector<uint32_t> targets;
  std::vector<size_t> moffs;    // offset of a gate's matrix within one set's matrices
  size_t set_doubles = 0;       // doubles of one set's matrices
  struct Gate {
    std::vector<GatePlan> plans;  // one per set (empty: the gate does nothing)
    size_t tab_off = 0;           // per-set kernel arguments
    size_t blk_off = (size_t)-1;  // per-set block matrices (4-5 targets)
    uint64_t gstride = 0;         // elements between two sets' blocks
  };
  std::vector<Gate> gates;
  void *dev = nullptr;
  size_t dev_cap = 0;
  uint64_t launches = 0;
  double bytes = 0.0;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  void *graph_sv = nullptr;
};

static void drop_batch_graph(Batch *b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  b->graph_sv = nullptr;
}

// Plans of one gate share their structure when these agree; the kernel
// arguments then differ in matrix entries only.
static bool same_structure(const GatePlan &a, const GatePlan &b) {
  if (a.kind != b.kind || a.kk != b.kk || a.fixed_mask != b.fixed_mask ||
      a.fixed_val != b.fixed_val)
    return false;
  for (int j = 0; j < a.kk; ++j)
    if (a.tbit[j] != b.tbit[j]) return false;
  if (a.kind == Kind::PERM) {
    if (a.moved != b.moved) return false;
    for (int r = 0; r < (1 << a.kk); ++r)
      if (a.perm[r] != b.perm[r]) return false;
  }
  return true;
}

// Plans of every gate and set, the host image of the argument tables and
// block matrices, and its size.  A gate whose sets' matrices fall into
// different classes (e.g. RotY(0) = identity next to RotY(0.3)) is planned as
// a dense gate on all its targets in every set.
static int plan_batch(Batch *b, const double *matrices, std::vector<char> *img) {
  const uint64_t n = 1ull << b->nq;
  const size_t ng = b->nt.size();
  std::vector<Batch::Gate> gates(ng);
  std::vector<char> tab;
  std::vector<std::pair<size_t, const GatePlan *>> blocks;  // (offset, plan)
  size_t blk_total = 0;
  uint64_t launches = 0;
  double bytes = 0.0;
  for (size_t g = 0; g < ng; ++g) {
    const int k = b->nt[g];
    const uint32_t *t = b->targets.data() + b->toff[g];
    std::vector<GatePlan> &P = gates[g].plans;
    P.resize(b->nsets);
    bool same = true, noop = true;
    for (uint32_t y = 0; y < b->nsets; ++y) {
      const double *m = matrices + y * b->set_doubles + b->moffs[g];
      SQ_TRY(make_plan(m, t, k, b->nq, b->nq, &P[y]));
      same = same && same_structure(P[0], P[y]);
      noop = noop && P[y].kind == Kind::NOOP;
    }
    if (noop) {
      P.clear();
      continue;
    }
    if (!same)
      for (uint32_t y = 0; y < b->nsets; ++y) {
        const double *m = matrices + y * b->set_doubles + b->moffs[g];
        SQ_TRY(make_plan(m, t, k, b->nq, b->nq, &P[y], kPlanDense));
      }
    std::vector<const GatePlan *> pp(b->nsets);
    for (uint32_t y = 0; y < b->nsets; ++y) {
      pp[y] = &P[y];
      bytes += 2.0 * P[y].touched * (double)n * itemsize(b->dt);
    }
    tab.resize((tab.size() + 255) / 256 * 256);
    gates[g].tab_off = tab.size();
    SQ_TRY(batch_gate_table(pp.data(), b->nsets, b->dt, n, &tab));
    if (has_block(P[0])) {
      const size_t bb = block_bytes(P[0].kk, b->dt);
      gates[g].blk_off = blk_total;
      gates[g].gstride = bb / (4 * (b->dt == SQ_COMPLEX64 ? 4 : 8));
      for (uint32_t y = 0; y < b->nsets; ++y) blocks.push_back({blk_total + y * bb, &P[y]});
      blk_total += b->nsets * bb;
    }
    ++launches;
  }
  // image: tables, then blocks (256-byte aligned)
  const size_t tab_bytes = (tab.size() + 255) / 256 * 256;
  img->assign(tab_bytes + blk_total, 0);
  std::memcpy(img->data(), tab.data(), tab.size());
  std::vector<const GatePlan *> bp;
  std::vector<size_t> bo;
  for (auto &e : blocks) {
    bp.push_back(e.second);
    bo.push_back(tab_bytes + e.first);
  }
  block_image(bp, b->dt, bo, img->data());
  for (auto &gt : gates)
    if (gt.blk_off != (size_t)-1) gt.blk_off += tab_bytes;
  b->gates.swap(gates);
  b->launches = launches;
  b->bytes = bytes;
  return SQ_OK;
}

static int upload_batch(Batch *b, const std::vector<char> &img, hipStream_t st) {
  if (img.size() > b->dev_cap) {
    if (b->dev) SQ_HIP(hipFree(b->dev));
    b->dev = nullptr;
    b->dev_cap = 0;
    if (hipMalloc(&b->dev, img.size()) != hipSuccess) {
      (void)hipGetLastError();
      b->dev = nullptr;
      return fail(SQ_ERR_OOM, "batch tables: device allocation failed");
    }
    b->dev_cap = img.size();
  }
  if (!img.empty()) {
    SQ_HIP(hipMemcpyAsync(b->dev, img.data(), img.size(), hipMemcpyHostToDevice, st));
    SQ_HIP(hipStreamSynchronize(st));
  }
  return SQ_OK;
}

static int batch_launches(Batch *b, void *states, hipStream_t st) {
  const uint64_t n = 1ull << b->nq;
  for (auto &gt : b->gates) {
    if (gt.plans.empty()) continue;
    const char *d = (const char *)b->dev;
    SQ_TRY(launch_batch_gate(gt.plans[0], b->dt, n, b->nsets, states, d + gt.tab_off,
                             gt.blk_off == (size_t)-1 ? nullptr : d + gt.blk_off, gt.gstride,
                             st));
  }
  return SQ_OK;
}

static int check_batch_states(const Batch *b, const void *states, uint64_t n) {
  if (!states) return fail(SQ_ERR_INVALID, "null states");
  if (n != (1ull << b->nq))
    return fail(SQ_ERR_INVALID, "batch built for " + std::to_string(b->nq) + " qubits, states have " +
                                    std::to_string(n) + " amplitudes");
  return SQ_OK;
}

}  // namespace sq

extern "C" {

int sq_batch_create(void **batch, sq_dtype_t dt, uint32_t nq, uint32_t nsets, uint32_t ng,
                    const uint32_t *num_targets, const uint32_t *targets,
                    const double *matrices) {
  if (!batch) return fail(SQ_ERR_INVALID, "null output pointer");
  *batch = nullptr;
  if (!valid_dtype(dt)) return fail(SQ_ERR_INVALID, "bad dtype");
  if (nq < 1 || nq > 40) return fail(SQ_ERR_INVALID, "num_qubits must be in [1, 40]");
  if (nsets < 1 || nsets > 65535) return fail(SQ_ERR_INVALID, "num_sets must be in [1, 65535]");
  if (ng && (!num_targets || !matrices)) return fail(SQ_ERR_INVALID, "null gate arrays");
  auto b = std::make_unique<Batch>();
  b->dt = dt;
  b->nq = (int)nq;
  b->nsets = nsets;
  SQ_TRY(current_device(&b->device));
  size_t toff = 0, moff = 0;
  for (uint32_t g = 0; g < ng; ++g) {
    const int k = (int)num_targets[g];
    if (k > kMaxK) return fail(SQ_ERR_UNSUPPORTED, "gate with more than 5 targets");
    if (k && !targets) return fail(SQ_ERR_INVALID, "null targets");
    b->nt.push_back(k);
    b->toff.push_back(toff);
    b->targets.insert(b->targets.end(), targets + toff, targets + toff + k);
    b->moffs.push_back(moff);
    toff += k;
    moff += 2ull << (2 * k);
  }
  b->set_doubles = moff;
  std::vector<char> img;
  SQ_TRY(plan_batch(b.get(), matrices, &img));
  SQ_TRY(upload_batch(b.get(), img, nullptr));
  *batch = b.release();
  return SQ_OK;
}

int sq_batch_set_matrices(void *batch, const double *matrices, void *stream) {
  Batch *b = (Batch *)batch;
  if (!b) return fail(SQ_ERR_INVALID, "null batch");
  if (!matrices && !b->nt.empty()) return fail(SQ_ERR_INVALID, "null matrices");
  std::vector<char> img;
  SQ_TRY(plan_batch(b, matrices, &img));
  drop_batch_graph(b);
  // work already queued on `stream` may still read the old tables: the new
  // image is copied behind it
  return upload_batch(b, img, (hipStream_t)stream);
}

int sq_batch_run(void *batch, void *states, uint64_t n, void *stream) {
  Batch *b = (Batch *)batch;
  if (!b) return fail(SQ_ERR_INVALID, "null batch");
  SQ_TRY(check_batch_states(b, states, n));
  return batch_launches(b, states, (hipStream_t)stream);
}

int sq_batch_run_graph(void *batch, void *states, uint64_t n, void *stream) {
  Batch *b = (Batch *)batch;
  if (!b) return fail(SQ_ERR_INVALID, "null batch");
  SQ_TRY(check_batch_states(b, states, n));
  if (b->graph_sv != states) {
    drop_batch_graph(b);
    hipStream_t cs;
    SQ_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    int rc = SQ_OK;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess)
      rc = fail(SQ_ERR_HIP, "hipStreamBeginCapture failed");
    if (rc == SQ_OK) rc = batch_launches(b, states, cs);
    hipGraph_t gr = nullptr;
    if (hipStreamEndCapture(cs, &gr) != hipSuccess && rc == SQ_OK)
      rc = fail(SQ_ERR_HIP, "hipStreamEndCapture failed");
    (void)hipStreamDestroy(cs);
    if (rc == SQ_OK && hipGraphInstantiate(&b->gexec, gr, nullptr, nullptr, 0) != hipSuccess)
      rc = fail(SQ_ERR_HIP, "hipGraphInstantiate failed");
    if (rc != SQ_OK) {
      if (gr) (void)hipGraphDestroy(gr);
      b->gexec = nullptr;
      return rc;
    }
    b->graph = gr;
    b->graph_sv = states;
  }
  SQ_HIP(hipGraphLaunch(b->gexec, (hipStream_t)stream));
  return SQ_OK;
}

int sq_batch_stats(void *batch, uint64_t *num_launches, double *bytes) {
  Batch *b = (Batch *)batch;
  if (!b) return fail(SQ_ERR_INVALID, "null batch");
  if (num_launches) *num_launches = b->launches;
  if (bytes) *bytes = b->bytes;
  return SQ_OK;
}

int sq_batch_destroy(void *batch) {
  Batch *b = (Batch *)batch;
  if (!b) return SQ_OK;
  drop_batch_graph(b);
  if (b->dev) SQ_HIP(hipFree(b->dev));
  delete b;
  return SQ_OK;
}

int sq_initialize_basis_z_batch(void *states, sq_dtype_t dt, uint64_t n, uint32_t nsets,
                                void *stream) {
  SQ_TRY(check_state(states, dt, n));
  if (nsets < 1) return fail(SQ_ERR_INVALID, "num_sets must be >= 1");
  return launch_init_batch(states, dt, n, nsets, (hipStream_t)stream);
}

int sq_expect_pauli_sum_batch(const void *states, sq_dtype_t dt, uint64_t n, uint32_t nsets,
                              uint32_t nterms, const uint64_t *x, const uint64_t *z,
                              const uint32_t *ny, const double *coeffs, double *out,
                              double *term_values, void *stream) {
  SQ_TRY(check_state(states, dt, n));
  if (nsets < 1 || nsets > 65535) return fail(SQ_ERR_INVALID, "num_sets must be in [1, 65535]");
  if (!out) return fail(SQ_ERR_INVALID, "null output");
  if (nterms && (!x || !z || !ny)) return fail(SQ_ERR_INVALID, "null Pauli masks");
  SQ_TRY(check_terms(nterms, x, z, ny, n));
  if (nterms == 0) {
    for (uint32_t s = 0; s < nsets; ++s) out[s] = 0.0;
    return SQ_OK;
  }
  int dev;
  SQ_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mutex[dev]);
  // terms in launches of up to 64; sets in launches whose partials fit 16 MiB
  const uint32_t tc = std::min<uint32_t>(64, nterms);
  const size_t per_set = pauli_batch_partial_doubles(n, 1, (int)tc);
  const uint32_t sc = (uint32_t)std::max<size_t>(1, std::min<size_t>(nsets, (2u << 20) / per_set));
  const size_t part_b = per_set * sc * sizeof(double) + 256;
  const size_t tb = nterms * pauli_term_bytes() + 256, ab = nterms * sizeof(double) + 256;
  const size_t vb = (size_t)nsets * nterms * sizeof(double) + 256;
  Workspace *w;
  SQ_TRY(ws_reserve(dev, part_b + tb + ab + vb + 1024, tb + ab + vb + 1024, &w));
  hipStream_t st = (hipStream_t)stream;
  char *d = (char *)w->dev;
  double *partial = (double *)d;
  char *terms = al(d + part_b);
  double *alpha = (double *)al(terms + tb);
  double *vals = (double *)al((char *)alpha + ab);
  char *h = (char *)w->host;
  double *halpha = (double *)al(h + tb);
  std::vector<double> ones(nterms, 1.0);
  fill_pauli_terms((int)nterms, x, z, ny, ones.data(), h, halpha);
  SQ_HIP(hipMemcpyAsync(terms, h, nterms * pauli_term_bytes(), hipMemcpyHostToDevice, st));
  SQ_HIP(hipMemcpyAsync(alpha, halpha, nterms * sizeof(double), hipMemcpyHostToDevice, st));
  for (uint32_t s0 = 0; s0 < nsets; s0 += sc) {
    const uint32_t ns = std::min(sc, nsets - s0);
    for (uint32_t t0 = 0; t0 < nterms; t0 += tc) {
      const uint32_t c = std::min(tc, nterms - t0);
      SQ_TRY(launch_pauli_terms_batch((const char *)states + s0 * n * itemsize(dt), dt, n, ns,
                                      (int)c, terms + t0 * pauli_term_bytes(), alpha + t0,
                                      partial, vals + (size_t)s0 * nterms + t0, nterms, st));
    }
  }
  double *hv = (double *)al(h + tb + ab);
  SQ_HIP(hipMemcpyAsync(hv, vals, (size_t)nsets * nterms * sizeof(double), hipMemcpyDeviceToHost, st));
  SQ_HIP(hipStreamSynchronize(st));
  for (uint32_t s = 0; s < nsets; ++s) {
    double acc = 0.0;
    for (uint32_t t = 0; t < nterms; ++t) acc += (coeffs ? coeffs[t] : 1.0) * hv[(size_t)s * nterms + t];
    out[s] = acc;
  }
  if (term_values) std::memcpy(term_values, hv, (size_t)nsets * nterms * sizeof(double));
  return SQ_OK;
}

}  // extern "C"
