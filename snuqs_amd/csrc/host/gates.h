// gates.h -- host-side gate objects (no device memory per gate).
//
// Mirrors the reference's Operation / GateOperation / GateOperationType
// (src/braket/snuqs/csrc/operation/operation.h:7-90) and the 33 gate classes
// of operation/gate_operations.h.  Differences, all deliberate:
//   - no cudaMalloc per gate (operation.cc:107-109); matrices stay on the host
//     and reach the device through sq_apply / sq_program_create;
//   - ctrl_modifiers and power are honoured (Braket semantics: the first
//     len(ctrl_modifiers) targets are controls, 1 = control on |1>, 0 = on
//     |0>; power is an integer, negative powers invert) where the reference
//     ignores them (operation.cc:96-110);
//   - GPhase with qubit targets is the phase times identity (the reference
//     leaves all but entry [0] uninitialised, gate_operations.cc:702-710);
//   - a working Unitary gate (the reference's cannot be constructed,
//     gate_operations.py:23-40);
//   - derivative(i): d(matrix)/d(angle_i), for adjoint differentiation.
#pragma once

#include <complex>
#include <memory>
#include <string>
#include <vector>

namespace snuqs {

using cplx = std::complex<double>;

class Operation {
 public:
  explicit Operation(const std::vector<size_t> &targets) : targets_(targets) {}
  virtual ~Operation() = default;
  std::vector<size_t> get_targets() const { return targets_; }
  void set_targets(const std::vector<size_t> &t) { targets_ = t; }
  std::vector<size_t> targets_;
};

// Same order as operation.h:17-52, plus Unitary.
enum class GateOperationType {
  Identity, Hadamard, PauliX, PauliY, PauliZ, CX, CY, CZ, S, Si, T, Ti, V, Vi,
  PhaseShift, CPhaseShift, CPhaseShift00, CPhaseShift01, CPhaseShift10,
  RotX, RotY, RotZ, Swap, ISwap, PSwap, XY, XX, YY, ZZ, CCNot, CSwap, U,
  GPhase, SwapA2A, Unitary,
};

std::string type_name(GateOperationType t);

class GateOperation : public Operation {
 public:
  GateOperation(GateOperationType type, const std::vector<size_t> &targets,
                const std::vector<double> &angles,
                const std::vector<size_t> &ctrl_modifiers, long power);
  // Unitary: explicit matrix (row-major, 2^k x 2^k)
  GateOperation(const std::vector<size_t> &targets,
                const std::vector<cplx> &matrix,
                const std::vector<size_t> &ctrl_modifiers, long power);

  GateOperationType type() const { return type_; }
  std::string name() const { return type_name(type_); }
  std::string formatted_string() const;
  const std::vector<double> &angles() const { return angles_; }
  const std::vector<size_t> &ctrl_modifiers() const { return ctrl_modifiers_; }
  long power() const { return power_; }
  bool operator==(const GateOperation &o) const;

  // full matrix on all targets (controls included), row-major
  const std::vector<cplx> &matrix() const { return matrix_; }
  size_t dim() const { return 2; }
  size_t ncols() const { return (size_t)1 << num_matrix_qubits(); }
  std::vector<size_t> shape() const { return {ncols(), ncols()}; }
  std::vector<size_t> stride() const {
    return {ncols() * sizeof(cplx), sizeof(cplx)};
  }
  void *data() { return matrix_.data(); }
  size_t num_matrix_qubits() const;

  size_t num_params() const;
  // d matrix / d angles_[i] (full matrix, controls included)
  std::vector<cplx> derivative(size_t i) const;

  bool sliceable() const { return false; }
  std::shared_ptr<GateOperation> slice(size_t) const {
    return std::make_shared<GateOperation>(*this);
  }

 private:
  std::vector<cplx> base_matrix(const std::vector<double> &angles) const;
  std::vector<cplx> base_derivative(size_t i) const;
  std::vector<cplx> finish(const std::vector<cplx> &base, bool apply_power) const;

  GateOperationType type_;
  std::vector<double> angles_;
  std::vector<size_t> ctrl_modifiers_;
  long power_ = 1;
  std::vector<cplx> explicit_;  // Unitary base matrix
  std::vector<cplx> matrix_;
};

}  // namespace snuqs
