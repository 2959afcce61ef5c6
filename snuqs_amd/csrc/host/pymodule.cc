// pymodule.cc -- pybind11 module snuqs_amd._C.
//
// Keeps the layout and names of the reference module braket.snuqs._C
// (src/braket/snuqs/csrc/pymodule.cc:66-210): submodules operation,
// operation.gate_operations, stream, event, functionals, result_types, core,
// core.cuda and the DeviceType enum.  Everything that touches the GPU goes
// through the C ABI of libsnuqs_hip.so (include/snuqs_hip.h); this file holds
// no device code.  DeviceType.CUDA names the HIP device.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>
#include <sys/sysinfo.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <sstream>
#include <stdexcept>

#include "gates.h"
#include "snuqs_hip.h"

namespace py = pybind11;
using namespace snuqs;

namespace {

void check(int rc) {
  if (rc == SQ_OK) return;
  std::string msg = sq_last_error();
  switch (rc) {
    case SQ_ERR_INVALID: throw py::value_error(msg);
    case SQ_ERR_OOM: throw std::bad_alloc();
    case SQ_ERR_UNSUPPORTED: throw std::invalid_argument(msg);
    default: throw std::runtime_error("HIP error: " + msg);
  }
}

// ---- Stream / Event (stream/stream.cc:6-34, event/event.cc:5-26) ----------
class Event;
class Stream {
 public:
  explicit Stream(uintptr_t s, bool owned = false) : s_((void *)s), owned_(owned) {}
  ~Stream() {
    if (owned_ && s_) sq_stream_destroy(s_);
  }
  uintptr_t get() const { return (uintptr_t)s_; }
  void *raw() const { return s_; }
  void synchronize() { check(sq_stream_synchronize(s_)); }
  void wait_event(std::shared_ptr<Event> e);
  static std::shared_ptr<Stream> create() {
    void *s;
    check(sq_stream_create(&s, 0));
    return std::make_shared<Stream>((uintptr_t)s, true);
  }
  static std::shared_ptr<Stream> create_nonblocking() {
    void *s;
    check(sq_stream_create(&s, 1));
    return std::make_shared<Stream>((uintptr_t)s, true);
  }

 private:
  void *s_;
  bool owned_;
};

class Event {
 public:
  explicit Event(uintptr_t e, bool owned = false) : e_((void *)e), owned_(owned) {}
  ~Event() {
    if (owned_ && e_) sq_event_destroy(e_);
  }
  uintptr_t get() const { return (uintptr_t)e_; }
  void *raw() const { return e_; }
  void record(std::shared_ptr<Stream> s) { check(sq_event_record(e_, s ? s->raw() : nullptr)); }
  void synchronize() { check(sq_event_synchronize(e_)); }
  float elapsed_ms(const Event &end) const {
    float ms;
    check(sq_event_elapsed_ms(&ms, e_, end.e_));
    return ms;
  }
  static std::shared_ptr<Event> create() {
    void *e;
    check(sq_event_create(&e));
    return std::make_shared<Event>((uintptr_t)e, true);
  }

 private:
  void *e_;
  bool owned_;
};

void Stream::wait_event(std::shared_ptr<Event> e) { check(sq_stream_wait_event(s_, e->raw())); }

void *sraw(const std::shared_ptr<Stream> &s) { return s ? s->raw() : nullptr; }

enum class DeviceType { CPU = 0, CUDA = 1, STORAGE = 2 };

// ---- buffers (buffer/buffer_cpu.cc, buffer/buffer_cuda.cc) ----------------
struct Buffer {
  DeviceType device;
  bool pinned = false;
  void *ptr = nullptr;
  size_t bytes = 0;
  Buffer(DeviceType d, size_t b, bool pin) : device(d), pinned(pin), bytes(b) {
    if (d == DeviceType::CUDA) {
      check(sq_malloc(&ptr, b));
    } else if (d == DeviceType::CPU) {
      if (pin) {
        check(sq_host_alloc_pinned(&ptr, b));
      } else {
        ptr = std::aligned_alloc(4096, ((b + 4095) / 4096) * 4096);
        if (!ptr) throw std::bad_alloc();
      }
    } else {
      throw std::invalid_argument("the STORAGE tier is not part of this engine");
    }
  }
  ~Buffer() {
    if (!ptr) return;
    if (device == DeviceType::CUDA) sq_free(ptr);
    else if (pinned) sq_host_free_pinned(ptr);
    else std::free(ptr);
  }
};

sq_dtype_t parse_dtype(const std::string &s) {
  if (s == "complex128" || s == "c128" || s == "complex") return SQ_COMPLEX128;
  if (s == "complex64" || s == "c64") return SQ_COMPLEX64;
  throw py::value_error("dtype must be 'complex64' or 'complex128'");
}
const char *dtype_name(sq_dtype_t d) { return d == SQ_COMPLEX64 ? "complex64" : "complex128"; }

// ---- StateVector (result_types/state_vector.h:12-42) ----------------------
class StateVector : public std::enable_shared_from_this<StateVector> {
 public:
  StateVector(DeviceType device, size_t num_qubits, sq_dtype_t dt, bool pinned = false)
      : device_(device), num_qubits_(num_qubits), dtype_(dt), offset_(0) {
    if (num_qubits > 40) throw py::value_error("num_qubits must be <= 40");
    buffer_ = std::make_shared<Buffer>(device, bytes(), pinned);
  }
  StateVector(DeviceType device, size_t num_qubits, sq_dtype_t dt,
              std::shared_ptr<Buffer> buf, size_t offset)
      : device_(device), num_qubits_(num_qubits), dtype_(dt), buffer_(buf), offset_(offset) {}

  size_t itemsize() const { return dtype_ == SQ_COMPLEX64 ? 8 : 16; }
  size_t num_elems() const { return (size_t)1 << num_qubits_; }
  size_t bytes() const { return num_elems() * itemsize(); }
  void *ptr() const { return (char *)buffer_->ptr + offset_; }
  DeviceType device() const { return device_; }
  size_t num_qubits() const { return num_qubits_; }
  sq_dtype_t dtype() const { return dtype_; }

  std::shared_ptr<StateVector> cpu(std::shared_ptr<Stream> stream) {
    if (device_ == DeviceType::CPU) return shared_from_this();
    auto out = std::make_shared<StateVector>(DeviceType::CPU, num_qubits_, dtype_, true);
    check(sq_memcpy_d2h(out->ptr(), ptr(), bytes(), sraw(stream)));
    check(sq_stream_synchronize(sraw(stream)));
    return out;
  }
  std::shared_ptr<StateVector> cuda(std::shared_ptr<Stream> stream) {
    if (device_ == DeviceType::CUDA) return shared_from_this();
    auto out = std::make_shared<StateVector>(DeviceType::CUDA, num_qubits_, dtype_);
    check(sq_memcpy_h2d(out->ptr(), ptr(), bytes(), sraw(stream)));
    check(sq_stream_synchronize(sraw(stream)));
    return out;
  }
  void copy(StateVector &other, std::shared_ptr<Stream> stream) {
    if (other.num_elems() != num_elems() || other.dtype_ != dtype_)
      throw py::value_error("copy: size or dtype mismatch");
    void *s = sraw(stream);
    if (device_ == DeviceType::CUDA && other.device_ == DeviceType::CUDA)
      check(sq_memcpy_d2d(ptr(), other.ptr(), bytes(), s));
    else if (device_ == DeviceType::CUDA)
      check(sq_memcpy_h2d(ptr(), other.ptr(), bytes(), s));
    else if (other.device_ == DeviceType::CUDA)
      check(sq_memcpy_d2h(ptr(), other.ptr(), bytes(), s));
    else
      std::memcpy(ptr(), other.ptr(), bytes());
    check(sq_stream_synchronize(s));
  }
  std::shared_ptr<StateVector> slice(size_t num_sliced_qubits, size_t index) {
    if (num_sliced_qubits > num_qubits_ || ((index + 1) << num_sliced_qubits) > num_elems())
      throw py::value_error("slice out of range");
    return std::make_shared<StateVector>(device_, num_sliced_qubits, dtype_, buffer_,
                                         offset_ + (index << num_sliced_qubits) * itemsize());
  }
  void sync() {}
  std::string repr() const {
    std::stringstream ss;
    ss << "StateVector<" << num_qubits_ << " qubits, device: "
       << (device_ == DeviceType::CPU ? "CPU" : "CUDA") << ", dtype: " << dtype_name(dtype_) << ">";
    return ss.str();
  }
  void require_device(const char *what) const {
    if (device_ != DeviceType::CUDA)
      throw std::invalid_argument(std::string(what) +
                                  " runs on the GPU: use StateVector(DeviceType.CUDA, n)");
  }

 private:
  DeviceType device_;
  size_t num_qubits_;
  sq_dtype_t dtype_;
  std::shared_ptr<Buffer> buffer_;
  size_t offset_;
};

std::vector<uint32_t> to_u32(const std::vector<size_t> &v) {
  return std::vector<uint32_t>(v.begin(), v.end());
}

const double *mat_ptr(const GateOperation &op) {
  return reinterpret_cast<const double *>(op.matrix().data());
}

// ---- functionals (functionals/functionals.cc:15-50) ------------------------
void apply(StateVector &sv, GateOperation &op, size_t num_qubits, std::vector<size_t> targets,
           std::shared_ptr<Stream> stream) {
  sv.require_device("apply");
  if (targets.size() != op.num_matrix_qubits())
    throw py::value_error("apply: " + std::to_string(targets.size()) + " targets for a " +
                          std::to_string(op.num_matrix_qubits()) + "-qubit gate matrix");
  auto t = to_u32(targets);
  py::gil_scoped_release nogil;
  check(sq_apply(sv.ptr(), sv.dtype(), sv.num_elems(), (uint32_t)num_qubits, t.data(),
                 (uint32_t)t.size(), mat_ptr(op), sraw(stream)));
}
void initialize_zero(StateVector &sv, std::shared_ptr<Stream> stream) {
  sv.require_device("initialize_zero");
  check(sq_initialize_zero(sv.ptr(), sv.dtype(), sv.num_elems(), sraw(stream)));
}
void initialize_basis_z(StateVector &sv, std::shared_ptr<Stream> stream) {
  sv.require_device("initialize_basis_z");
  check(sq_initialize_basis_z(sv.ptr(), sv.dtype(), sv.num_elems(), sraw(stream)));
}

// Pauli-sum observable in mask form.
struct PauliSum {
  std::vector<uint64_t> x, z;
  std::vector<uint32_t> ny;
  std::vector<double> c;
};
PauliSum make_pauli_sum(size_t nq, const std::vector<std::string> &paulis,
                        const std::vector<std::vector<size_t>> &qubits,
                        const std::vector<double> &coeffs) {
  if (paulis.size() != qubits.size() || paulis.size() != coeffs.size())
    throw py::value_error("paulis, qubits and coeffs must have equal length");
  PauliSum s;
  for (size_t t = 0; t < paulis.size(); ++t) {
    if (paulis[t].size() != qubits[t].size())
      throw py::value_error("Pauli string and qubit list lengths differ");
    uint64_t x = 0, z = 0;
    uint32_t ny = 0;
    for (size_t j = 0; j < paulis[t].size(); ++j) {
      const size_t q = qubits[t][j];
      if (q >= nq) throw py::value_error("Pauli qubit out of range");
      const uint64_t b = 1ull << (nq - 1 - q);
      if ((x | z) & b) throw py::value_error("repeated qubit in a Pauli string");
      switch (paulis[t][j]) {
        case 'I': break;
        case 'X': x |= b; break;
        case 'Y': x |= b; z |= b; ++ny; break;
        case 'Z': z |= b; break;
        default: throw py::value_error("Pauli letters must be I, X, Y or Z");
      }
    }
    s.x.push_back(x);
    s.z.push_back(z);
    s.ny.push_back(ny);
    s.c.push_back(coeffs[t]);
  }
  return s;
}

int probs_into(StateVector &sv, double *d, bool normalize, std::shared_ptr<Stream> stream) {
  return normalize ? sq_probabilities_normalized(sv.ptr(), sv.dtype(), sv.num_elems(), d, sraw(stream))
                   : sq_probabilities(sv.ptr(), sv.dtype(), sv.num_elems(), d, sraw(stream));
}

py::array_t<double> probabilities(StateVector &sv, bool normalize, std::shared_ptr<Stream> stream) {
  sv.require_device("probabilities");
  const size_t n = sv.num_elems();
  void *d;
  check(sq_malloc(&d, n * sizeof(double)));
  std::unique_ptr<void, int (*)(void *)> g(d, sq_free);
  py::array_t<double> out(n);
  {
    py::gil_scoped_release nogil;
    check(probs_into(sv, (double *)d, normalize, stream));
    check(sq_memcpy_d2h(out.mutable_data(), d, n * sizeof(double), sraw(stream)));
    check(sq_stream_synchronize(sraw(stream)));
  }
  return out;
}

py::array_t<int64_t> sample(StateVector &sv, py::array_t<double, py::array::c_style> uniforms,
                            bool exact, bool normalize, std::shared_ptr<Stream> stream) {
  sv.require_device("sample");
  const size_t n = sv.num_elems();
  void *d;
  check(sq_malloc(&d, n * sizeof(double)));
  std::unique_ptr<void, int (*)(void *)> g(d, sq_free);
  py::array_t<int64_t> out(uniforms.size());
  const double *u = uniforms.data();
  int64_t *o = out.mutable_data();
  const size_t shots = uniforms.size();
  {
    py::gil_scoped_release nogil;
    check(probs_into(sv, (double *)d, normalize, stream));
    check(sq_sample((const double *)d, n, u, shots, exact ? 1 : 0, o, sraw(stream)));
  }
  return out;
}

// ---- Program: a gate list uploaded once (replaces the per-gate loop of
// simulation.py:274-284) ---------------------------------------------------
class Program {
 public:
  Program(sq_dtype_t dt, size_t nq, const std::vector<std::shared_ptr<GateOperation>> &ops)
      : dt_(dt), nq_(nq), ops_(ops) {
    std::vector<uint32_t> nt, tg;
    std::vector<double> mats;
    for (auto &op : ops) {
      auto t = op->get_targets();
      if (t.size() != op->num_matrix_qubits())
        throw py::value_error("gate " + op->formatted_string() + ": targets do not match matrix");
      nt.push_back((uint32_t)t.size());
      for (size_t q : t) tg.push_back((uint32_t)q);
      const double *m = mat_ptr(*op);
      mats.insert(mats.end(), m, m + 2 * op->matrix().size());
    }
    check(sq_program_create(&prog_, dt, (uint32_t)nq, (uint32_t)ops.size(), nt.data(),
                            tg.data(), mats.data()));
  }
  ~Program() {
    if (prog_) sq_program_destroy(prog_);
  }
  // New gate objects with the same targets (new parameters); no device
  // allocation.
  void set_gates(const std::vector<std::shared_ptr<GateOperation>> &ops,
                 std::shared_ptr<Stream> stream) {
    if (ops.size() != ops_.size())
      throw py::value_error("set_gates: " + std::to_string(ops.size()) + " gates given, the program has " +
                            std::to_string(ops_.size()));
    std::vector<double> mats;
    for (size_t g = 0; g < ops.size(); ++g) {
      if (ops[g]->get_targets() != ops_[g]->get_targets() ||
          ops[g]->num_matrix_qubits() != ops_[g]->num_matrix_qubits())
        throw py::value_error("set_gates: gate " + std::to_string(g) + " (" +
                              ops[g]->formatted_string() + ") has other targets than " +
                              ops_[g]->formatted_string());
      const double *m = mat_ptr(*ops[g]);
      mats.insert(mats.end(), m, m + 2 * ops[g]->matrix().size());
    }
    {
      py::gil_scoped_release nogil;
      check(sq_program_set_matrices(prog_, mats.data(), sraw(stream)));
    }
    ops_ = ops;
  }
  void run(StateVector &sv, bool use_graph, std::shared_ptr<Stream> stream) {
    sv.require_device("Program.run");
    if (sv.dtype() != dt_) throw py::value_error("state dtype differs from the program's");
    py::gil_scoped_release nogil;
    if (use_graph)
      check(sq_program_run_graph(prog_, sv.ptr(), sv.num_elems(), sraw(stream)));
    else
      check(sq_program_run(prog_, sv.ptr(), sv.num_elems(), sraw(stream)));
  }
  py::array_t<float> run_timed(StateVector &sv, std::shared_ptr<Stream> stream) {
    sv.require_device("Program.run_timed");
    if (sv.dtype() != dt_) throw py::value_error("state dtype differs from the program's");
    py::array_t<float> ms(ops_.size());
    float *m = ms.mutable_data();
    {
      py::gil_scoped_release nogil;
      check(sq_program_run_timed(prog_, sv.ptr(), sv.num_elems(), sraw(stream), m));
    }
    return ms;
  }
  uint32_t tune(StateVector &sv, int reps, std::shared_ptr<Stream> stream) {
    sv.require_device("Program.tune");
    if (sv.dtype() != dt_) throw py::value_error("state dtype differs from the program's");
    uint32_t tuned = 0;
    {
      py::gil_scoped_release nogil;
      check(sq_program_tune(prog_, sv.ptr(), sv.num_elems(), reps, sraw(stream), &tuned));
    }
    return tuned;
  }
  py::tuple gate_info(size_t g) const {
    int cls, kk;
    double b;
    check(sq_program_gate_info(prog_, (uint32_t)g, &cls, &kk, &b));
    return py::make_tuple(cls, kk, b);
  }
  size_t num_gates() const { return ops_.size(); }
  py::dict stats() const {
    uint64_t l;
    double b;
    check(sq_program_stats(prog_, &l, &b));
    py::dict d;
    d["launches"] = l;
    d["algorithmic_bytes"] = b;
    return d;
  }
  // One streaming pass of the host-memory tier (sq_host_tier_pass): this
  // program (on local_bits.size() qubits) over every slice of `host`.
  void run_host_pass(StateVector &host, const std::vector<uint32_t> &local_bits,
                     const std::vector<std::shared_ptr<StateVector>> &buffers, bool zero_slices,
                     std::shared_ptr<Stream> stream) {
    if (host.device() != DeviceType::CPU)
      throw py::value_error("run_host_pass: the state must be a host StateVector (pinned)");
    if (host.dtype() != dt_) throw py::value_error("run_host_pass: state dtype differs from the program's");
    if (local_bits.size() != nq_)
      throw py::value_error("run_host_pass: the program has " + std::to_string(nq_) +
                            " qubits, " + std::to_string(local_bits.size()) + " local bits given");
    if (buffers.empty()) throw py::value_error("run_host_pass: no device buffers");
    std::vector<void *> bp;
    for (auto &b : buffers) {
      if (!b) throw py::value_error("run_host_pass: buffer is None");
      b->require_device("run_host_pass buffer");
      if (b->dtype() != dt_ || b->num_qubits() != nq_)
        throw py::value_error("run_host_pass: buffers must be device states of the program's "
                              "qubit count and dtype");
      bp.push_back(b->ptr());
    }
    py::gil_scoped_release nogil;
    check(sq_host_tier_pass(prog_, host.ptr(), dt_, host.num_elems(), (uint32_t)local_bits.size(),
                            local_bits.data(), bp.data(), (uint32_t)bp.size(), zero_slices ? 1 : 0,
                            sraw(stream)));
  }
  // Energy and gradient w.r.t. `num_params` parameters; param_map[g] lists
  // (angle index, parameter index) pairs of gate g.
  py::tuple adjoint(StateVector &sv, StateVector &w0, StateVector &w1,
                    std::vector<std::shared_ptr<StateVector>> checkpoints,
                    const std::vector<std::vector<std::pair<size_t, size_t>>> &param_map,
                    size_t num_params, const std::vector<std::string> &paulis,
                    const std::vector<std::vector<size_t>> &qubits,
                    const std::vector<double> &coeffs, std::shared_ptr<Stream> stream) {
    // The C ABI takes one size for every buffer, so each one is checked here
    // against the program before anything is launched.
    auto require_match = [&](const StateVector &s, const std::string &what) {
      s.require_device(what.c_str());
      if (s.dtype() != dt_)
        throw py::value_error("adjoint: " + what + " has dtype " + dtype_name(s.dtype()) +
                              ", the program " + dtype_name(dt_));
      if (s.num_qubits() != nq_)
        throw py::value_error("adjoint: " + what + " has " + std::to_string(s.num_qubits()) +
                              " qubits, the program " + std::to_string(nq_));
    };
    require_match(sv, "state_vector");
    require_match(w0, "work0");
    require_match(w1, "work1");
    for (size_t c = 0; c < checkpoints.size(); ++c) {
      if (!checkpoints[c]) throw py::value_error("adjoint: checkpoint " + std::to_string(c) + " is None");
      require_match(*checkpoints[c], "checkpoint " + std::to_string(c));
    }
    if (param_map.size() != ops_.size()) throw py::value_error("param_map needs one entry per gate");
    PauliSum ps = make_pauli_sum(nq_, paulis, qubits, coeffs);
    std::vector<uint32_t> dg, dp;
    std::vector<double> dm;
    for (size_t g = 0; g < ops_.size(); ++g)
      for (auto &ap : param_map[g]) {
        if (ap.second >= num_params) throw py::value_error("parameter index out of range");
        auto d = ops_[g]->derivative(ap.first);
        dg.push_back((uint32_t)g);
        dp.push_back((uint32_t)ap.second);
        const double *m = reinterpret_cast<const double *>(d.data());
        dm.insert(dm.end(), m, m + 2 * d.size());
      }
    std::vector<void *> cps;
    for (auto &c : checkpoints) {
      c->require_device("checkpoint");
      cps.push_back(c->ptr());
    }
    double energy = 0;
    py::array_t<double> grads(num_params);
    double *gp = grads.mutable_data();
    {
      py::gil_scoped_release nogil;
      check(sq_adjoint_gradient(prog_, (uint32_t)dg.size(), dg.data(), dp.data(), dm.data(),
                                (uint32_t)num_params, sv.ptr(), w0.ptr(), w1.ptr(),
                                cps.empty() ? nullptr : cps.data(), (uint32_t)cps.size(),
                                sv.num_elems(), (uint32_t)ps.x.size(), ps.x.data(), ps.z.data(),
                                ps.ny.data(), ps.c.data(), &energy, gp, sraw(stream)));
    }
    return py::make_tuple(energy, grads);
  }

 private:
  sq_dtype_t dt_;
  size_t nq_;
  std::vector<std::shared_ptr<GateOperation>> ops_;
  void *prog_ = nullptr;
};

// Circuits of uneven qubit counts as one launch sequence (sq_cbatch_*).
class CircuitBatch {
 public:
  CircuitBatch(const std::vector<std::pair<size_t, std::vector<std::shared_ptr<GateOperation>>>> &circuits,
               sq_dtype_t dt)
      : dt_(dt) {
    std::vector<uint32_t> nq, ng, nt, tg;
    std::vector<double> mats;
    for (auto &c : circuits) {
      nq.push_back((uint32_t)c.first);
      ng.push_back((uint32_t)c.second.size());
      for (auto &op : c.second) {
        auto t = op->get_targets();
        if (t.size() != op->num_matrix_qubits())
          throw py::value_error("gate " + op->formatted_string() + ": targets do not match matrix");
        nt.push_back((uint32_t)t.size());
        for (size_t q : t) tg.push_back((uint32_t)q);
        const double *m = mat_ptr(*op);
        mats.insert(mats.end(), m, m + 2 * op->matrix().size());
      }
    }
    nq_ = nq;
    check(sq_cbatch_create(&h_, dt, (uint32_t)nq.size(), nq.data(), ng.data(), nt.data(),
                           tg.data(), mats.data()));
  }
  ~CircuitBatch() {
    if (h_) sq_cbatch_destroy(h_);
  }
  void run(bool use_graph, std::shared_ptr<Stream> stream) {
    py::gil_scoped_release nogil;
    check(sq_cbatch_run(h_, use_graph ? 1 : 0, sraw(stream)));
  }
  void copy_state(uint32_t j, StateVector &sv, std::shared_ptr<Stream> stream) {
    sv.require_device("copy_state");
    if (j >= nq_.size()) throw py::index_error("no circuit " + std::to_string(j));
    if (sv.dtype() != dt_) throw py::value_error("state dtype differs from the batch's");
    if (sv.num_qubits() != nq_[j])
      throw py::value_error("state has " + std::to_string(sv.num_qubits()) + " qubits, circuit " +
                            std::to_string(j) + " has " + std::to_string(nq_[j]));
    py::gil_scoped_release nogil;
    check(sq_cbatch_copy_state(h_, j, sv.ptr(), sv.num_elems(), sraw(stream)));
  }
  py::dict stats() const {
    uint64_t l;
    double sb, tb;
    check(sq_cbatch_stats(h_, &l, &sb, &tb));
    py::dict d;
    d["launches"] = l;
    d["state_bytes"] = sb;
    d["traffic_bytes"] = tb;
    return d;
  }
  std::vector<uint32_t> num_qubits() const { return nq_; }

 private:
  sq_dtype_t dt_;
  std::vector<uint32_t> nq_;
  void *h_ = nullptr;
};

// Batched QAOA MaxCut (sq_qaoa_*).
class QaoaBatch {
 public:
  QaoaBatch(const std::vector<uint32_t> &nq, const std::vector<std::vector<std::pair<uint32_t, uint32_t>>> &graphs,
            uint32_t p, sq_dtype_t dt)
      : n_(nq.size()), p_(p) {
    if (graphs.size() != nq.size()) throw py::value_error("one graph per instance");
    std::vector<uint32_t> ne, ed;
    for (auto &gph : graphs) {
      ne.push_back((uint32_t)gph.size());
      for (auto &e : gph) {
        ed.push_back(e.first);
        ed.push_back(e.second);
      }
    }
    check(sq_qaoa_create(&h_, dt, (uint32_t)nq.size(), nq.data(), ne.data(), ed.data(), p));
  }
  ~QaoaBatch() {
    if (h_) sq_qaoa_destroy(h_);
  }
  py::array_t<double> run(py::array_t<double, py::array::c_style | py::array::forcecast> gammas,
                          py::array_t<double, py::array::c_style | py::array::forcecast> betas,
                          bool use_graph, std::shared_ptr<Stream> stream) {
    if ((size_t)gammas.size() != n_ * p_ || (size_t)betas.size() != n_ * p_)
      throw py::value_error("gammas and betas must have num_instances x p entries");
    py::array_t<double> out({n_, (size_t)(1 + 2 * p_)});
    const double *g = gammas.data(), *b = betas.data();
    double *o = out.mutable_data();
    {
      py::gil_scoped_release nogil;
      check(sq_qaoa_run(h_, g, b, use_graph ? 1 : 0, o, sraw(stream)));
    }
    return out;
  }
  py::dict stats() const {
    uint64_t l;
    double sb, tb;
    check(sq_qaoa_stats(h_, &l, &sb, &tb));
    py::dict d;
    d["launches"] = l;
    d["state_bytes"] = sb;
    d["traffic_bytes"] = tb;
    return d;
  }

 private:
  size_t n_;
  uint32_t p_;
  void *h_ = nullptr;
};

// ---- parameter batches (sq_batch_*): num_sets states of one size in one
// device buffer, evolved by one gate structure with per-state matrices ------
class StateBatch {
 public:
  StateBatch(size_t nsets, size_t nq, sq_dtype_t dt) : nsets_(nsets), nq_(nq), dt_(dt) {
    if (nsets < 1 || nsets > 65535) throw py::value_error("num_sets must be in [1, 65535]");
    if (nq < 1 || nq > 40) throw py::value_error("num_qubits must be in [1, 40]");
    buffer_ = std::make_shared<Buffer>(DeviceType::CUDA, bytes(), false);
  }
  size_t num_sets() const { return nsets_; }
  size_t num_qubits() const { return nq_; }
  sq_dtype_t dtype() const { return dt_; }
  size_t num_elems() const { return (size_t)1 << nq_; }
  size_t bytes() const { return nsets_ * num_elems() * (dt_ == SQ_COMPLEX64 ? 8 : 16); }
  void *ptr() const { return buffer_->ptr; }
  // State y as a StateVector sharing the batch's memory.
  std::shared_ptr<StateVector> state(size_t y) {
    if (y >= nsets_) throw py::index_error("state index out of range");
    return std::make_shared<StateVector>(DeviceType::CUDA, nq_, dt_, buffer_,
                                         y * num_elems() * (dt_ == SQ_COMPLEX64 ? 8 : 16));
  }
  void initialize_basis_z(std::shared_ptr<Stream> stream) {
    check(sq_initialize_basis_z_batch(ptr(), dt_, num_elems(), (uint32_t)nsets_, sraw(stream)));
  }
  py::array cpu(std::shared_ptr<Stream> stream) {
    const std::vector<size_t> shape{nsets_, num_elems()};
    py::array out = dt_ == SQ_COMPLEX64 ? (py::array)py::array_t<std::complex<float>>(shape)
                                        : (py::array)py::array_t<std::complex<double>>(shape);
    void *h = out.mutable_data();
    {
      py::gil_scoped_release nogil;
      check(sq_memcpy_d2h(h, ptr(), bytes(), sraw(stream)));
      check(sq_stream_synchronize(sraw(stream)));
    }
    return out;
  }

 private:
  size_t nsets_, nq_;
  sq_dtype_t dt_;
  std::shared_ptr<Buffer> buffer_;
};

class BatchProgram {
 public:
  BatchProgram(sq_dtype_t dt, size_t nq,
               const std::vector<std::vector<std::shared_ptr<GateOperation>>> &sets)
      : dt_(dt), nq_(nq) {
    if (sets.empty()) throw py::value_error("BatchProgram: no parameter sets");
    ref_ = sets[0];
    std::vector<uint32_t> nt, tg;
    for (auto &op : ref_) {
      auto t = op->get_targets();
      if (t.size() != op->num_matrix_qubits())
        throw py::value_error("gate " + op->formatted_string() + ": targets do not match matrix");
      nt.push_back((uint32_t)t.size());
      for (size_t q : t) tg.push_back((uint32_t)q);
    }
    std::vector<double> mats = matrices(sets);
    check(sq_batch_create(&h_, dt, (uint32_t)nq, (uint32_t)sets.size(), (uint32_t)ref_.size(),
                          nt.data(), tg.data(), mats.data()));
    nsets_ = sets.size();
  }
  ~BatchProgram() {
    if (h_) sq_batch_destroy(h_);
  }
  void set_gates(const std::vector<std::vector<std::shared_ptr<GateOperation>>> &sets,
                 std::shared_ptr<Stream> stream) {
    if (sets.size() != nsets_)
      throw py::value_error("set_gates: " + std::to_string(sets.size()) + " sets given, the batch has " +
                            std::to_string(nsets_));
    std::vector<double> mats = matrices(sets);
    py::gil_scoped_release nogil;
    check(sq_batch_set_matrices(h_, mats.data(), sraw(stream)));
  }
  void run(StateBatch &b, bool use_graph, std::shared_ptr<Stream> stream) {
    if (b.dtype() != dt_) throw py::value_error("state batch dtype differs from the program's");
    if (b.num_qubits() != nq_) throw py::value_error("state batch has another qubit count");
    if (b.num_sets() != nsets_) throw py::value_error("state batch has another number of states");
    py::gil_scoped_release nogil;
    if (use_graph)
      check(sq_batch_run_graph(h_, b.ptr(), b.num_elems(), sraw(stream)));
    else
      check(sq_batch_run(h_, b.ptr(), b.num_elems(), sraw(stream)));
  }
  py::dict stats() const {
    uint64_t l;
    double by;
    check(sq_batch_stats(h_, &l, &by));
    py::dict d;
    d["launches"] = l;
    d["algorithmic_bytes"] = by;
    return d;
  }
  size_t num_gates() const { return ref_.size(); }
  size_t num_sets() const { return nsets_; }

 private:
  // every set's matrices, set-major, checked against set 0's targets
  std::vector<double> matrices(const std::vector<std::vector<std::shared_ptr<GateOperation>>> &sets) const {
    std::vector<double> mats;
    for (size_t y = 0; y < sets.size(); ++y) {
      if (sets[y].size() != ref_.size())
        throw py::value_error("parameter set " + std::to_string(y) + " has " +
                              std::to_string(sets[y].size()) + " gates, set 0 has " +
                              std::to_string(ref_.size()));
      for (size_t g = 0; g < ref_.size(); ++g) {
        const auto &op = sets[y][g];
        if (op->get_targets() != ref_[g]->get_targets() ||
            op->num_matrix_qubits() != ref_[g]->num_matrix_qubits())
          throw py::value_error("parameter set " + std::to_string(y) + ", gate " + std::to_string(g) +
                                " (" + op->formatted_string() + ") has other targets than set 0's " +
                                ref_[g]->formatted_string());
        const double *m = mat_ptr(*op);
        mats.insert(mats.end(), m, m + 2 * op->matrix().size());
      }
    }
    return mats;
  }
  sq_dtype_t dt_;
  size_t nq_, nsets_ = 0;
  std::vector<std::shared_ptr<GateOperation>> ref_;
  void *h_ = nullptr;
};

py::buffer_info gate_buffer(GateOperation &g) {
  return py::buffer_info(g.data(), sizeof(cplx), py::format_descriptor<cplx>::format(), 2,
                         g.shape(), g.stride());
}

}  // namespace

// One C++ subclass per gate type, so isinstance() and the reference's
// constructor signatures (pymodule.cc:20-64) hold.
template <GateOperationType T>
struct Gate0 : GateOperation {
  Gate0(const std::vector<size_t> &t, const std::vector<size_t> &c, long p)
      : GateOperation(T, t, {}, c, p) {}
};
template <GateOperationType T>
struct Gate1 : GateOperation {
  Gate1(const std::vector<size_t> &t, double a, const std::vector<size_t> &c, long p)
      : GateOperation(T, t, {a}, c, p) {}
};
struct GateU : GateOperation {
  GateU(const std::vector<size_t> &t, double a1, double a2, double a3,
        const std::vector<size_t> &c, long p)
      : GateOperation(GateOperationType::U, t, {a1, a2, a3}, c, p) {}
};
struct GateUnitary : GateOperation {
  GateUnitary(const std::vector<size_t> &t, const std::vector<cplx> &m,
              const std::vector<size_t> &c, long p)
      : GateOperation(t, m, c, p) {}
};

template <GateOperationType T>
void bind0(py::module_ &mg, const char *name) {
  py::class_<Gate0<T>, GateOperation, std::shared_ptr<Gate0<T>>>(mg, name, py::buffer_protocol())
      .def_buffer([](Gate0<T> &g) { return gate_buffer(g); })
      .def(py::init<const std::vector<size_t> &, const std::vector<size_t> &, long>(),
           py::arg("targets"), py::kw_only(), py::arg("ctrl_modifiers") = std::vector<size_t>(),
           py::arg("power") = 1);
}
template <GateOperationType T>
void bind1(py::module_ &mg, const char *name) {
  py::class_<Gate1<T>, GateOperation, std::shared_ptr<Gate1<T>>>(mg, name, py::buffer_protocol())
      .def_buffer([](Gate1<T> &g) { return gate_buffer(g); })
      .def(py::init<const std::vector<size_t> &, double, const std::vector<size_t> &, long>(),
           py::arg("targets"), py::arg("angle"), py::kw_only(),
           py::arg("ctrl_modifiers") = std::vector<size_t>(), py::arg("power") = 1);
}

PYBIND11_MODULE(_C, m) {
  m.doc() = "SnuQS MI355X state-vector engine (HIP, gfx950).";

  py::enum_<DeviceType>(m, "DeviceType")
      .value("CPU", DeviceType::CPU)
      .value("CUDA", DeviceType::CUDA)
      .value("STORAGE", DeviceType::STORAGE)
      .export_values();

  // operation
  auto m_op = m.def_submodule("operation");
  py::class_<Operation, std::shared_ptr<Operation>>(m_op, "Operation")
      .def(py::init<const std::vector<size_t> &>())
      .def_property("targets", &Operation::get_targets, &Operation::set_targets);

  py::enum_<GateOperationType> gt(m_op, "GateOperationType");
  for (int t = 0; t <= (int)GateOperationType::Unitary; ++t)
    gt.value(type_name((GateOperationType)t).c_str(), (GateOperationType)t);

  py::class_<GateOperation, Operation, std::shared_ptr<GateOperation>>(m_op, "GateOperation",
                                                                        py::buffer_protocol())
      .def(py::init<GateOperationType, const std::vector<size_t> &, const std::vector<double> &,
                    const std::vector<size_t> &, long>(),
           py::arg("type"), py::arg("targets"), py::arg("angles") = std::vector<double>(),
           py::arg("ctrl_modifiers") = std::vector<size_t>(), py::arg("power") = 1)
      .def_buffer(&gate_buffer)
      .def("sliceable", &GateOperation::sliceable)
      .def("slice", &GateOperation::slice)
      .def("__repr__", &GateOperation::formatted_string)
      .def("__eq__", &GateOperation::operator==)
      .def("__hash__", [](const GateOperation &g) { return std::hash<std::string>()(g.formatted_string()); })
      .def_property_readonly("name", &GateOperation::name)
      .def_property_readonly("type", &GateOperation::type)
      .def_property_readonly("angles", &GateOperation::angles)
      .def_property_readonly("ctrl_modifiers", &GateOperation::ctrl_modifiers)
      .def_property_readonly("power", &GateOperation::power)
      .def_property_readonly("num_params", &GateOperation::num_params)
      .def("matrix", [](GateOperation &g) { return py::array(gate_buffer(g)).attr("copy")(); })
      .def("derivative", [](GateOperation &g, size_t i) {
        auto d = g.derivative(i);
        const size_t n = g.ncols();
        py::array_t<cplx> a({n, n});
        std::memcpy(a.mutable_data(), d.data(), d.size() * sizeof(cplx));
        return a;
      });

  // operation.gate_operations: the 33 gate classes of pymodule.cc:87-120 as
  // subclasses with the reference's constructor signatures, plus Unitary.
  auto m_gates = m_op.def_submodule("gate_operations");
  using G = GateOperationType;
  bind0<G::Identity>(m_gates, "Identity");
  bind0<G::Hadamard>(m_gates, "Hadamard");
  bind0<G::PauliX>(m_gates, "PauliX");
  bind0<G::PauliY>(m_gates, "PauliY");
  bind0<G::PauliZ>(m_gates, "PauliZ");
  bind0<G::CX>(m_gates, "CX");
  bind0<G::CY>(m_gates, "CY");
  bind0<G::CZ>(m_gates, "CZ");
  bind0<G::S>(m_gates, "S");
  bind0<G::Si>(m_gates, "Si");
  bind0<G::T>(m_gates, "T");
  bind0<G::Ti>(m_gates, "Ti");
  bind0<G::V>(m_gates, "V");
  bind0<G::Vi>(m_gates, "Vi");
  bind1<G::PhaseShift>(m_gates, "PhaseShift");
  bind1<G::CPhaseShift>(m_gates, "CPhaseShift");
  bind1<G::CPhaseShift00>(m_gates, "CPhaseShift00");
  bind1<G::CPhaseShift01>(m_gates, "CPhaseShift01");
  bind1<G::CPhaseShift10>(m_gates, "CPhaseShift10");
  bind1<G::RotX>(m_gates, "RotX");
  bind1<G::RotY>(m_gates, "RotY");
  bind1<G::RotZ>(m_gates, "RotZ");
  bind0<G::Swap>(m_gates, "Swap");
  bind0<G::ISwap>(m_gates, "ISwap");
  bind1<G::PSwap>(m_gates, "PSwap");
  bind1<G::XY>(m_gates, "XY");
  bind1<G::XX>(m_gates, "XX");
  bind1<G::YY>(m_gates, "YY");
  bind1<G::ZZ>(m_gates, "ZZ");
  bind0<G::CCNot>(m_gates, "CCNot");
  bind0<G::CSwap>(m_gates, "CSwap");
  py::class_<GateU, GateOperation, std::shared_ptr<GateU>>(m_gates, "U", py::buffer_protocol())
      .def_buffer([](GateU &g) { return gate_buffer(g); })
      .def(py::init<const std::vector<size_t> &, double, double, double,
                    const std::vector<size_t> &, long>(),
           py::arg("targets"), py::arg("angle_1"), py::arg("angle_2"), py::arg("angle_3"),
           py::kw_only(), py::arg("ctrl_modifiers") = std::vector<size_t>(),
           py::arg("power") = 1);
  bind1<G::GPhase>(m_gates, "GPhase");
  py::class_<GateUnitary, GateOperation, std::shared_ptr<GateUnitary>>(m_gates, "Unitary",
                                                                       py::buffer_protocol())
      .def_buffer([](GateUnitary &g) { return gate_buffer(g); })
      .def(py::init([](const std::vector<size_t> &targets,
                       py::array_t<cplx, py::array::c_style | py::array::forcecast> matrix,
                       const std::vector<size_t> &ctrl, long power) {
             if (matrix.ndim() != 2 || matrix.shape(0) != matrix.shape(1))
               throw py::value_error("matrix is not a two-dimensional square matrix");
             std::vector<cplx> m(matrix.data(), matrix.data() + matrix.size());
             return std::make_shared<GateUnitary>(targets, m, ctrl, power);
           }),
           py::arg("targets"), py::arg("matrix"), py::kw_only(),
           py::arg("ctrl_modifiers") = std::vector<size_t>(), py::arg("power") = 1);

  // stream / event
  auto m_stream = m.def_submodule("stream");
  py::class_<Stream, std::shared_ptr<Stream>>(m_stream, "Stream")
      .def(py::init<uintptr_t>())
      .def("get", &Stream::get)
      .def("synchronize", &Stream::synchronize)
      .def("wait_event", &Stream::wait_event)
      .def_static("create", &Stream::create)
      .def_static("create_nonblocking", &Stream::create_nonblocking);
  auto m_event = m.def_submodule("event");
  py::class_<Event, std::shared_ptr<Event>>(m_event, "Event")
      .def(py::init<uintptr_t>())
      .def("get", &Event::get)
      .def("record", &Event::record, py::arg("stream") = nullptr)
      .def("synchronize", &Event::synchronize)
      .def("elapsed_ms", &Event::elapsed_ms)
      .def_static("create", &Event::create);

  // result_types
  auto m_rt = m.def_submodule("result_types");
  py::class_<StateVector, std::shared_ptr<StateVector>>(m_rt, "StateVector", py::buffer_protocol())
      .def_buffer([](StateVector &sv) -> py::buffer_info {
        if (sv.device() != DeviceType::CPU)
          throw std::invalid_argument("buffer protocol needs a host state: call .cpu() first");
        if (sv.dtype() == SQ_COMPLEX128)
          return py::buffer_info(sv.ptr(), 16, py::format_descriptor<std::complex<double>>::format(),
                                 1, {sv.num_elems()}, {(size_t)16});
        return py::buffer_info(sv.ptr(), 8, py::format_descriptor<std::complex<float>>::format(), 1,
                               {sv.num_elems()}, {(size_t)8});
      })
      .def(py::init([](size_t n, std::string dtype) {
             return std::make_shared<StateVector>(DeviceType::CPU, n, parse_dtype(dtype));
           }), py::arg("num_qubits"), py::kw_only(), py::arg("dtype") = "complex128")
      .def(py::init([](size_t n, bool pinned, std::string dtype) {
             return std::make_shared<StateVector>(DeviceType::CPU, n, parse_dtype(dtype), pinned);
           }), py::arg("num_qubits"), py::arg("pinned"), py::kw_only(), py::arg("dtype") = "complex128")
      .def(py::init([](DeviceType d, size_t n, std::string dtype) {
             return std::make_shared<StateVector>(d, n, parse_dtype(dtype));
           }), py::arg("device"), py::arg("num_qubits"), py::kw_only(), py::arg("dtype") = "complex128")
      .def("cpu", &StateVector::cpu, py::kw_only(), py::arg("stream") = nullptr)
      .def("cuda", &StateVector::cuda, py::kw_only(), py::arg("stream") = nullptr)
      .def("hip", &StateVector::cuda, py::kw_only(), py::arg("stream") = nullptr)
      .def("copy", &StateVector::copy, py::arg("other"), py::kw_only(), py::arg("stream") = nullptr)
      .def("slice", &StateVector::slice)
      .def("__repr__", &StateVector::repr)
      .def("num_elems", &StateVector::num_elems)
      .def("sync", &StateVector::sync)
      .def_property_readonly("num_qubits", &StateVector::num_qubits)
      .def_property_readonly("device", &StateVector::device)
      .def_property_readonly("dtype", [](const StateVector &s) { return dtype_name(s.dtype()); })
      .def_property_readonly("data_ptr", [](const StateVector &s) { return (uintptr_t)s.ptr(); });

  // functionals
  auto m_fn = m.def_submodule("functionals");
  m_fn.def("apply", &apply, py::arg("state_vector"), py::arg("op"), py::arg("num_qubits"),
           py::arg("targets"), py::kw_only(), py::arg("stream") = nullptr);
  m_fn.def("initialize_zero", &initialize_zero, py::arg("state_vector"), py::kw_only(),
           py::arg("stream") = nullptr);
  m_fn.def("initialize_basis_z", &initialize_basis_z, py::arg("state_vector"), py::kw_only(),
           py::arg("stream") = nullptr);
  m_fn.def("host_initialize_basis_z", [](StateVector &sv, uint32_t threads) {
    if (sv.device() != DeviceType::CPU) throw py::value_error("host_initialize_basis_z: host states only");
    py::gil_scoped_release nogil;
    check(sq_host_initialize_basis_z(sv.ptr(), sv.dtype(), sv.num_elems(), threads));
  }, py::arg("state_vector"), py::kw_only(), py::arg("threads") = 16,
     "|0...0> in a host state (parallel fill).");
  m_fn.def("probabilities", &probabilities, py::arg("state_vector"), py::kw_only(),
           py::arg("normalize") = false, py::arg("stream") = nullptr,
           "|sv|^2 in float64; normalize=True divides by the FP64 norm^2.");
  m_fn.def("sample", &sample, py::arg("state_vector"), py::arg("uniforms"), py::kw_only(),
           py::arg("exact") = true, py::arg("normalize") = false, py::arg("stream") = nullptr,
           "Indices drawn from probabilities(state_vector, normalize=normalize).");
  m_fn.def("copy_kernel", [](StateVector &dst, StateVector &src, std::shared_ptr<Stream> s) {
    dst.require_device("copy_kernel");
    src.require_device("copy_kernel");
    if (dst.num_elems() != src.num_elems() || dst.dtype() != src.dtype())
      throw py::value_error("copy_kernel: states differ in size or dtype");
    check(sq_copy(dst.ptr(), src.ptr(), dst.bytes(), sraw(s)));
  }, py::arg("dst"), py::arg("src"), py::kw_only(), py::arg("stream") = nullptr);
  m_fn.def("marginal_probabilities", [](StateVector &sv, std::vector<uint32_t> targets,
                                        std::shared_ptr<Stream> s) {
    sv.require_device("marginal_probabilities");
    if (targets.empty() || targets.size() > 20)
      throw py::value_error("marginal_probabilities takes 1..20 targets");
    py::array_t<double> out((size_t)1 << targets.size());
    double *o = out.mutable_data();
    {
      py::gil_scoped_release nogil;
      check(sq_marginal_probabilities(sv.ptr(), sv.dtype(), sv.num_elems(),
                                      (uint32_t)sv.num_qubits(), targets.data(),
                                      (uint32_t)targets.size(), o, sraw(s)));
    }
    return out;
  }, py::arg("state_vector"), py::arg("targets"), py::kw_only(), py::arg("stream") = nullptr,
     "Marginal distribution over `targets` (targets[0] = most significant bit).");
  m_fn.def("norm2", [](StateVector &sv, std::shared_ptr<Stream> s) {
    sv.require_device("norm2");
    double out;
    check(sq_norm2(sv.ptr(), sv.dtype(), sv.num_elems(), &out, sraw(s)));
    return out;
  }, py::arg("state_vector"), py::kw_only(), py::arg("stream") = nullptr);
  m_fn.def("renormalize", [](StateVector &sv, std::shared_ptr<Stream> s) {
    sv.require_device("renormalize");
    double n2;
    check(sq_norm2(sv.ptr(), sv.dtype(), sv.num_elems(), &n2, sraw(s)));
    if (!(n2 > 0.0)) throw py::value_error("cannot renormalize a zero state");
    check(sq_scale(sv.ptr(), sv.dtype(), sv.num_elems(), 1.0 / std::sqrt(n2), 0.0, sraw(s)));
    return n2;
  }, py::arg("state_vector"), py::kw_only(), py::arg("stream") = nullptr,
     "Scale the state to unit norm (FP64 norm); returns the norm^2 before scaling.");
  m_fn.def("scale", [](StateVector &sv, cplx f, std::shared_ptr<Stream> s) {
    sv.require_device("scale");
    check(sq_scale(sv.ptr(), sv.dtype(), sv.num_elems(), f.real(), f.imag(), sraw(s)));
  }, py::arg("state_vector"), py::arg("factor"), py::kw_only(), py::arg("stream") = nullptr);
  m_fn.def("inner", [](StateVector &a, StateVector &b, std::shared_ptr<Stream> s) {
    a.require_device("inner");
    b.require_device("inner");
    if (a.dtype() != b.dtype() || a.num_elems() != b.num_elems())
      throw py::value_error("inner: states differ in size or dtype");
    double out[2];
    check(sq_inner(a.ptr(), b.ptr(), a.dtype(), a.num_elems(), out, sraw(s)));
    return cplx(out[0], out[1]);
  }, py::arg("a"), py::arg("b"), py::kw_only(), py::arg("stream") = nullptr);
  m_fn.def("amplitudes", [](StateVector &sv, std::vector<uint64_t> idx, std::shared_ptr<Stream> s) {
    sv.require_device("amplitudes");
    py::array_t<cplx> out(idx.size());
    check(sq_amplitudes(sv.ptr(), sv.dtype(), sv.num_elems(), idx.data(), idx.size(),
                        reinterpret_cast<double *>(out.mutable_data()), sraw(s)));
    return out;
  }, py::arg("state_vector"), py::arg("indices"), py::kw_only(), py::arg("stream") = nullptr);
  m_fn.def("expectation", [](StateVector &sv, const std::vector<std::string> &paulis,
                             const std::vector<std::vector<size_t>> &qubits,
                             const std::vector<double> &coeffs, std::shared_ptr<Stream> s) {
    sv.require_device("expectation");
    PauliSum ps = make_pauli_sum(sv.num_qubits(), paulis, qubits, coeffs);
    double out;
    py::array_t<double> terms(ps.x.size());
    check(sq_expect_pauli_sum(sv.ptr(), sv.dtype(), sv.num_elems(), (uint32_t)ps.x.size(),
                              ps.x.data(), ps.z.data(), ps.ny.data(), ps.c.data(), &out,
                              terms.mutable_data(), sraw(s)));
    return py::make_tuple(out, terms);
  }, py::arg("state_vector"), py::arg("paulis"), py::arg("qubits"), py::arg("coeffs"),
     py::kw_only(), py::arg("stream") = nullptr);
  m_fn.def("apply_pauli_sum", [](StateVector &in, StateVector &out,
                                 const std::vector<std::string> &paulis,
                                 const std::vector<std::vector<size_t>> &qubits,
                                 const std::vector<double> &coeffs, std::shared_ptr<Stream> s) {
    in.require_device("apply_pauli_sum");
    out.require_device("apply_pauli_sum");
    PauliSum ps = make_pauli_sum(in.num_qubits(), paulis, qubits, coeffs);
    check(sq_apply_pauli_sum(in.ptr(), out.ptr(), in.dtype(), in.num_elems(),
                             (uint32_t)ps.x.size(), ps.x.data(), ps.z.data(), ps.ny.data(),
                             ps.c.data(), sraw(s)));
  }, py::arg("input"), py::arg("output"), py::arg("paulis"), py::arg("qubits"),
     py::arg("coeffs"), py::kw_only(), py::arg("stream") = nullptr);

  py::class_<Program, std::shared_ptr<Program>>(m_fn, "Program")
      .def(py::init([](size_t nq, const std::vector<std::shared_ptr<GateOperation>> &ops,
                       std::string dtype) {
             return std::make_shared<Program>(parse_dtype(dtype), nq, ops);
           }), py::arg("num_qubits"), py::arg("operations"), py::kw_only(),
           py::arg("dtype") = "complex128")
      .def("run", &Program::run, py::arg("state_vector"), py::kw_only(),
           py::arg("use_graph") = false, py::arg("stream") = nullptr,
           "Apply the gates in order; use_graph replays a captured hipGraph of the launches.")
      .def("tune", &Program::tune, py::arg("state_vector"), py::kw_only(), py::arg("reps") = 2,
           py::arg("stream") = nullptr,
           "Pick each 5-target gate's fastest launch shape by timing it on state_vector "
           "(overwritten); returns the number of gates tuned.")
      .def("run_timed", &Program::run_timed, py::arg("state_vector"), py::kw_only(),
           py::arg("stream") = nullptr)
      .def("gate_info", &Program::gate_info)
      .def("__len__", &Program::num_gates)
      .def("stats", &Program::stats)
      .def("set_gates", &Program::set_gates, py::arg("ops"), py::kw_only(),
           py::arg("stream") = nullptr)
      .def("run_host_pass", &Program::run_host_pass, py::arg("host_state"), py::arg("local_bits"),
           py::arg("buffers"), py::kw_only(), py::arg("zero_slices") = false,
           py::arg("stream") = nullptr,
           "Stream a pinned host state through the device buffers, running this program on "
           "every slice whose device bit j is host bit local_bits[j].")
      .def("adjoint", &Program::adjoint, py::arg("state_vector"), py::arg("work0"),
           py::arg("work1"), py::arg("checkpoints"), py::arg("param_map"),
           py::arg("num_params"), py::arg("paulis"), py::arg("qubits"), py::arg("coeffs"),
           py::kw_only(), py::arg("stream") = nullptr);

  py::class_<StateBatch, std::shared_ptr<StateBatch>>(m_rt, "StateBatch")
      .def(py::init([](size_t nsets, size_t n, std::string dtype) {
             return std::make_shared<StateBatch>(nsets, n, parse_dtype(dtype));
           }), py::arg("num_sets"), py::arg("num_qubits"), py::kw_only(),
           py::arg("dtype") = "complex128",
           "num_sets device states of 2^num_qubits amplitudes in one buffer.")
      .def("state", &StateBatch::state, py::arg("index"),
           "State `index` as a StateVector sharing the batch's memory.")
      .def("initialize_basis_z", &StateBatch::initialize_basis_z, py::kw_only(),
           py::arg("stream") = nullptr)
      .def("cpu", &StateBatch::cpu, py::kw_only(), py::arg("stream") = nullptr,
           "Host copy, shape (num_sets, 2^num_qubits).")
      .def_property_readonly("num_sets", &StateBatch::num_sets)
      .def_property_readonly("num_qubits", &StateBatch::num_qubits)
      .def_property_readonly("dtype", [](const StateBatch &b) { return dtype_name(b.dtype()); })
      .def_property_readonly("data_ptr", [](const StateBatch &b) { return (uintptr_t)b.ptr(); });

  py::class_<BatchProgram, std::shared_ptr<BatchProgram>>(m_fn, "BatchProgram")
      .def(py::init([](size_t nq,
                       const std::vector<std::vector<std::shared_ptr<GateOperation>>> &sets,
                       std::string dtype) {
             return std::make_shared<BatchProgram>(parse_dtype(dtype), nq, sets);
           }), py::arg("num_qubits"), py::arg("parameter_sets"), py::kw_only(),
           py::arg("dtype") = "complex128",
           "One gate list per parameter set (same gates and targets, own matrices).")
      .def("run", &BatchProgram::run, py::arg("states"), py::kw_only(),
           py::arg("use_graph") = false, py::arg("stream") = nullptr,
           "Apply every gate to every state, one launch per gate for the whole batch.")
      .def("set_gates", &BatchProgram::set_gates, py::arg("parameter_sets"), py::kw_only(),
           py::arg("stream") = nullptr)
      .def("stats", &BatchProgram::stats)
      .def("__len__", &BatchProgram::num_gates)
      .def_property_readonly("num_sets", &BatchProgram::num_sets);

  m_fn.def("expectation_batch", [](StateBatch &b, const std::vector<std::string> &paulis,
                                   const std::vector<std::vector<size_t>> &qubits,
                                   const std::vector<double> &coeffs, std::shared_ptr<Stream> s) {
    PauliSum ps = make_pauli_sum(b.num_qubits(), paulis, qubits, coeffs);
    py::array_t<double> out(b.num_sets());
    py::array_t<double> terms(std::vector<size_t>{b.num_sets(), ps.x.size()});
    double *o = out.mutable_data(), *t = terms.mutable_data();
    {
      py::gil_scoped_release nogil;
      check(sq_expect_pauli_sum_batch(b.ptr(), b.dtype(), b.num_elems(), (uint32_t)b.num_sets(),
                                      (uint32_t)ps.x.size(), ps.x.data(), ps.z.data(),
                                      ps.ny.data(), ps.c.data(), o, t, sraw(s)));
    }
    return py::make_tuple(out, terms);
  }, py::arg("states"), py::arg("paulis"), py::arg("qubits"), py::arg("coeffs"), py::kw_only(),
     py::arg("stream") = nullptr,
     "Per-state energies sum_t c_t <psi_y|P_t|psi_y> (FP64) and per-term values.");

  py::class_<CircuitBatch, std::shared_ptr<CircuitBatch>>(m_fn, "CircuitBatch")
      .def(py::init([](const std::vector<std::pair<size_t, std::vector<std::shared_ptr<GateOperation>>>> &circuits,
                       std::string dtype) {
             return std::make_shared<CircuitBatch>(circuits, parse_dtype(dtype));
           }), py::arg("circuits"), py::kw_only(), py::arg("dtype") = "complex64",
           "Circuits of uneven qubit counts, [(num_qubits, gates), ...], as one launch sequence.")
      .def("run", &CircuitBatch::run, py::kw_only(), py::arg("use_graph") = true,
           py::arg("stream") = nullptr,
           "Every state <- |0...0>, then every circuit's gates (launch s: gate s of each circuit).")
      .def("copy_state", &CircuitBatch::copy_state, py::arg("circuit"), py::arg("sv"),
           py::kw_only(), py::arg("stream") = nullptr,
           "Copy circuit j's state into a device StateVector of its qubit count and dtype.")
      .def("stats", &CircuitBatch::stats)
      .def_property_readonly("num_qubits", &CircuitBatch::num_qubits);

  py::class_<QaoaBatch, std::shared_ptr<QaoaBatch>>(m_fn, "QaoaBatch")
      .def(py::init([](const std::vector<uint32_t> &nq,
                       const std::vector<std::vector<std::pair<uint32_t, uint32_t>>> &graphs,
                       uint32_t p, std::string dtype) {
             return std::make_shared<QaoaBatch>(nq, graphs, p, parse_dtype(dtype));
           }), py::arg("num_qubits"), py::arg("graphs"), py::arg("p"), py::kw_only(),
           py::arg("dtype") = "complex64")
      .def("run", &QaoaBatch::run, py::arg("gammas"), py::arg("betas"), py::kw_only(),
           py::arg("use_graph") = true, py::arg("stream") = nullptr)
      .def("stats", &QaoaBatch::stats);

  // core (core/runtime.cc:19-103) and core.cuda (core/cuda/runtime.cu:4-25)
  auto m_core = m.def_submodule("core");
  m_core.def("mem_info", []() {
    struct sysinfo si;
    if (sysinfo(&si) != 0) throw std::runtime_error("sysinfo failed");
    return std::make_pair((size_t)si.freeram * si.mem_unit, (size_t)si.totalram * si.mem_unit);
  });
  m_core.def("is_initialized_fs", []() { return false; });
  m_core.def("initialize_fs", [](size_t, size_t, const std::vector<std::string> &) {
    throw std::invalid_argument("the storage-offload tier is not part of this engine");
  });
  auto m_cuda = m_core.def_submodule("cuda");
  m_cuda.def("mem_info", []() {
    size_t f, t;
    check(sq_mem_info(&f, &t));
    return std::make_pair(f, t);
  });
  m_cuda.def("device_count", []() {
    int c;
    check(sq_device_count(&c));
    return c;
  });
  m_cuda.def("get_device", []() {
    int d;
    check(sq_get_device(&d));
    return d;
  });
  m_cuda.def("set_device", [](int d) { check(sq_set_device(d)); });
  m_cuda.def("device_synchronize", []() { check(sq_device_synchronize()); });
  m_core.def("tune_set", [](const std::string &k, int v) { check(sq_tune_set(k.c_str(), v)); });
  m_core.def("tune_get", [](const std::string &k) {
    int v;
    check(sq_tune_get(k.c_str(), &v));
    return v;
  });
  m.attr("__version__") = sq_version();
}
