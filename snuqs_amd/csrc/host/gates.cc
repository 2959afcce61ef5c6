// gates.cc -- closed-form gate matrices and their angle derivatives.
//
// Entries follow the gate definitions of
// src/braket/snuqs/csrc/operation/gate_operations.cc:9-734 (row-major, first
// target = MSB of the local index).
#include "gates.h"

#include <cmath>
#include <sstream>
#include <stdexcept>

namespace snuqs {

namespace {
constexpr double kSqrtHalf = 0.70710678118654752440;  // M_SQRT1_2
const cplx I1(0.0, 1.0);

using M = std::vector<cplx>;

M eye(size_t d) {
  M m(d * d, 0.0);
  for (size_t i = 0; i < d; ++i) m[i * d + i] = 1.0;
  return m;
}
M diag(std::initializer_list<cplx> d) {
  const size_t n = d.size();
  M m(n * n, 0.0);
  size_t i = 0;
  for (cplx v : d) {
    m[i * n + i] = v;
    ++i;
  }
  return m;
}
// permutation matrix with out[r] = in[src[r]] times ph[r]
M monomial(std::initializer_list<int> src, std::initializer_list<cplx> ph) {
  const size_t n = src.size();
  M m(n * n, 0.0);
  auto s = src.begin();
  auto p = ph.begin();
  for (size_t r = 0; r < n; ++r, ++s, ++p) m[r * n + *s] = *p;
  return m;
}
M matmul(const M &a, const M &b, size_t d) {
  M c(d * d, 0.0);
  for (size_t i = 0; i < d; ++i)
    for (size_t k = 0; k < d; ++k) {
      const cplx aik = a[i * d + k];
      if (aik == 0.0) continue;
      for (size_t j = 0; j < d; ++j) c[i * d + j] += aik * b[k * d + j];
    }
  return c;
}
M dagger(const M &a, size_t d) {
  M c(d * d);
  for (size_t i = 0; i < d; ++i)
    for (size_t j = 0; j < d; ++j) c[i * d + j] = std::conj(a[j * d + i]);
  return c;
}
size_t isqrt_dim(size_t n) {
  size_t d = 1;
  while (d * d < n) ++d;
  return d;
}
int base_qubits(GateOperationType t) {
  using G = GateOperationType;
  switch (t) {
    case G::GPhase: return 0;
    case G::CX: case G::CY: case G::CZ: case G::CPhaseShift: case G::CPhaseShift00:
    case G::CPhaseShift01: case G::CPhaseShift10: case G::Swap: case G::ISwap:
    case G::PSwap: case G::XY: case G::XX: case G::YY: case G::ZZ: case G::SwapA2A:
      return 2;
    case G::CCNot: case G::CSwap: return 3;
    case G::Unitary: return -1;
    default: return 1;
  }
}
int angle_count(GateOperationType t) {
  using G = GateOperationType;
  switch (t) {
    case G::PhaseShift: case G::CPhaseShift: case G::CPhaseShift00: case G::CPhaseShift01:
    case G::CPhaseShift10: case G::RotX: case G::RotY: case G::RotZ: case G::PSwap:
    case G::XY: case G::XX: case G::YY: case G::ZZ: case G::GPhase:
      return 1;
    case G::U: return 3;
    default: return 0;
  }
}
}  // namespace

std::string type_name(GateOperationType t) {
  static const char *names[] = {
      "Identity", "Hadamard", "PauliX", "PauliY", "PauliZ", "CX", "CY", "CZ", "S", "Si",
      "T", "Ti", "V", "Vi", "PhaseShift", "CPhaseShift", "CPhaseShift00", "CPhaseShift01",
      "CPhaseShift10", "RotX", "RotY", "RotZ", "Swap", "ISwap", "PSwap", "XY", "XX", "YY",
      "ZZ", "CCNot", "CSwap", "U", "GPhase", "SwapA2A", "Unitary"};
  return names[(int)t];
}

M GateOperation::base_matrix(const std::vector<double> &a) const {
  using G = GateOperationType;
  switch (type_) {
    case G::Identity: return eye(2);
    case G::Hadamard: return {kSqrtHalf, kSqrtHalf, kSqrtHalf, -kSqrtHalf};
    case G::PauliX: return {0.0, 1.0, 1.0, 0.0};
    case G::PauliY: return {0.0, -I1, I1, 0.0};
    case G::PauliZ: return diag({1.0, -1.0});
    case G::CX: return monomial({0, 1, 3, 2}, {1.0, 1.0, 1.0, 1.0});
    case G::CY: return monomial({0, 1, 3, 2}, {1.0, 1.0, -I1, I1});
    case G::CZ: return diag({1.0, 1.0, 1.0, -1.0});
    case G::S: return diag({1.0, I1});
    case G::Si: return diag({1.0, -I1});
    case G::T: return diag({1.0, std::exp(I1 * (M_PI / 4.0))});
    case G::Ti: return diag({1.0, std::exp(-I1 * (M_PI / 4.0))});
    case G::V: return {cplx(0.5, 0.5), cplx(0.5, -0.5), cplx(0.5, -0.5), cplx(0.5, 0.5)};
    case G::Vi: return {cplx(0.5, -0.5), cplx(0.5, 0.5), cplx(0.5, 0.5), cplx(0.5, -0.5)};
    case G::PhaseShift: return diag({1.0, std::exp(I1 * a[0])});
    case G::CPhaseShift: return diag({1.0, 1.0, 1.0, std::exp(I1 * a[0])});
    case G::CPhaseShift00: return diag({std::exp(I1 * a[0]), 1.0, 1.0, 1.0});
    case G::CPhaseShift01: return diag({1.0, std::exp(I1 * a[0]), 1.0, 1.0});
    case G::CPhaseShift10: return diag({1.0, 1.0, std::exp(I1 * a[0]), 1.0});
    case G::RotX: {
      const double c = std::cos(a[0] / 2);
      const cplx s = I1 * std::sin(a[0] / 2);
      return {c, -s, -s, c};
    }
    case G::RotY: {
      const double c = std::cos(a[0] / 2), s = std::sin(a[0] / 2);
      return {c, -s, s, c};
    }
    case G::RotZ: return diag({std::exp(-I1 * a[0] / 2.0), std::exp(I1 * a[0] / 2.0)});
    case G::Swap:
    case G::SwapA2A: return monomial({0, 2, 1, 3}, {1.0, 1.0, 1.0, 1.0});
    case G::ISwap: return monomial({0, 2, 1, 3}, {1.0, I1, I1, 1.0});
    case G::PSwap: {
      const cplx e = std::exp(I1 * a[0]);
      return monomial({0, 2, 1, 3}, {1.0, e, e, 1.0});
    }
    case G::XY: {
      const double c = std::cos(a[0] / 2);
      const cplx s = I1 * std::sin(a[0] / 2);
      return {1.0, 0.0, 0.0, 0.0, 0.0, c, s, 0.0, 0.0, s, c, 0.0, 0.0, 0.0, 0.0, 1.0};
    }
    case G::XX: {
      const double c = std::cos(a[0] / 2);
      const cplx s = I1 * std::sin(a[0] / 2);
      return {c, 0.0, 0.0, -s, 0.0, c, -s, 0.0, 0.0, -s, c, 0.0, -s, 0.0, 0.0, c};
    }
    case G::YY: {
      const double c = std::cos(a[0] / 2);
      const cplx s = I1 * std::sin(a[0] / 2);
      return {c, 0.0, 0.0, s, 0.0, c, -s, 0.0, 0.0, -s, c, 0.0, s, 0.0, 0.0, c};
    }
    case G::ZZ: {
      const cplx p = std::exp(I1 * a[0] / 2.0), m = std::exp(-I1 * a[0] / 2.0);
      return diag({m, p, p, m});
    }
    case G::CCNot: return monomial({0, 1, 2, 3, 4, 5, 7, 6}, {1, 1, 1, 1, 1, 1, 1, 1});
    case G::CSwap: return monomial({0, 1, 2, 3, 4, 6, 5, 7}, {1, 1, 1, 1, 1, 1, 1, 1});
    case G::U: {
      const double c = std::cos(a[0] / 2), s = std::sin(a[0] / 2);
      return {c, -std::exp(I1 * a[2]) * s, std::exp(I1 * a[1]) * s,
              std::exp(I1 * (a[1] + a[2])) * c};
    }
    case G::GPhase: return {std::exp(I1 * a[0])};
    case G::Unitary: return explicit_;
  }
  throw std::invalid_argument("unknown gate type");
}

// d(base matrix)/d(angle i), closed forms.
M GateOperation::base_derivative(size_t i) const {
  using G = GateOperationType;
  const auto &a = angles_;
  const double h = 0.5;
  switch (type_) {
    case G::PhaseShift: return diag({0.0, I1 * std::exp(I1 * a[0])});
    case G::CPhaseShift: return diag({0.0, 0.0, 0.0, I1 * std::exp(I1 * a[0])});
    case G::CPhaseShift00: return diag({I1 * std::exp(I1 * a[0]), 0.0, 0.0, 0.0});
    case G::CPhaseShift01: return diag({0.0, I1 * std::exp(I1 * a[0]), 0.0, 0.0});
    case G::CPhaseShift10: return diag({0.0, 0.0, I1 * std::exp(I1 * a[0]), 0.0});
    case G::RotX: {
      const double dc = -h * std::sin(a[0] / 2);
      const cplx ds = I1 * (h * std::cos(a[0] / 2));
      return {dc, -ds, -ds, dc};
    }
    case G::RotY: {
      const double dc = -h * std::sin(a[0] / 2), ds = h * std::cos(a[0] / 2);
      return {dc, -ds, ds, dc};
    }
    case G::RotZ:
      return diag({-I1 * h * std::exp(-I1 * a[0] / 2.0), I1 * h * std::exp(I1 * a[0] / 2.0)});
    case G::PSwap: {
      const cplx e = I1 * std::exp(I1 * a[0]);
      return monomial({0, 2, 1, 3}, {0.0, e, e, 0.0});
    }
    case G::XY: {
      const double dc = -h * std::sin(a[0] / 2);
      const cplx ds = I1 * (h * std::cos(a[0] / 2));
      return {0.0, 0.0, 0.0, 0.0, 0.0, dc, ds, 0.0, 0.0, ds, dc, 0.0, 0.0, 0.0, 0.0, 0.0};
    }
    case G::XX: {
      const double dc = -h * std::sin(a[0] / 2);
      const cplx ds = I1 * (h * std::cos(a[0] / 2));
      return {dc, 0.0, 0.0, -ds, 0.0, dc, -ds, 0.0, 0.0, -ds, dc, 0.0, -ds, 0.0, 0.0, dc};
    }
    case G::YY: {
      const double dc = -h * std::sin(a[0] / 2);
      const cplx ds = I1 * (h * std::cos(a[0] / 2));
      return {dc, 0.0, 0.0, ds, 0.0, dc, -ds, 0.0, 0.0, -ds, dc, 0.0, ds, 0.0, 0.0, dc};
    }
    case G::ZZ: {
      const cplx p = I1 * h * std::exp(I1 * a[0] / 2.0), m = -I1 * h * std::exp(-I1 * a[0] / 2.0);
      return diag({m, p, p, m});
    }
    case G::GPhase: return {I1 * std::exp(I1 * a[0])};
    case G::U: {
      const double c = std::cos(a[0] / 2), s = std::sin(a[0] / 2);
      const cplx el = std::exp(I1 * a[2]), ep = std::exp(I1 * a[1]),
                 epl = std::exp(I1 * (a[1] + a[2]));
      if (i == 0) return {-h * s, -el * (h * c), ep * (h * c), -epl * (h * s)};
      if (i == 1) return {0.0, 0.0, I1 * ep * s, I1 * epl * c};
      return {0.0, -I1 * el * s, 0.0, I1 * epl * c};
    }
    default:
      throw std::invalid_argument(name() + " has no angle parameters");
  }
}

// Apply power (integer; negative = inverse) and control modifiers to a base
// matrix.  Controls come first in the target list; the controlled operator is
// identity except on the block where every control matches its modifier.
M GateOperation::finish(const M &base, bool apply_power) const {
  const size_t bd = isqrt_dim(base.size());
  M b = base;
  if (apply_power && power_ != 1) {
    M r = eye(bd);
    M step = power_ < 0 ? dagger(base, bd) : base;
    for (long p = 0; p < std::labs(power_); ++p) r = matmul(r, step, bd);
    b = r;
  }
  const size_t nc = ctrl_modifiers_.size();
  if (nc == 0) return b;
  const size_t d = bd << nc;
  // control pattern: bits of the control part of the local index (MSB first)
  size_t pattern = 0;
  for (size_t j = 0; j < nc; ++j) pattern = (pattern << 1) | (ctrl_modifiers_[j] ? 1 : 0);
  M m = apply_power ? eye(d) : M(d * d, 0.0);
  for (size_t r = 0; r < bd; ++r)
    for (size_t c = 0; c < bd; ++c) m[(pattern * bd + r) * d + pattern * bd + c] = b[r * bd + c];
  return m;
}

size_t GateOperation::num_matrix_qubits() const {
  size_t d = isqrt_dim(matrix_.size()), k = 0;
  while ((1ull << k) < d) ++k;
  return k;
}

GateOperation::GateOperation(GateOperationType type, const std::vector<size_t> &targets,
                             const std::vector<double> &angles,
                             const std::vector<size_t> &ctrl_modifiers, long power)
    : Operation(targets), type_(type), angles_(angles), ctrl_modifiers_(ctrl_modifiers),
      power_(power) {
  if (type == GateOperationType::Unitary)
    throw std::invalid_argument("Unitary needs a matrix");
  if ((int)angles.size() != angle_count(type))
    throw std::invalid_argument(name() + " takes " + std::to_string(angle_count(type)) +
                                " angle(s)");
  for (size_t c : ctrl_modifiers)
    if (c > 1) throw std::invalid_argument("ctrl_modifiers entries must be 0 or 1");
  const int bq = base_qubits(type);
  const size_t want = (size_t)bq + ctrl_modifiers.size();
  if (type == GateOperationType::GPhase && targets.size() > want) {
    // GPhase on qubits: phase times identity on those qubits (beyond controls)
    const size_t extra = targets.size() - ctrl_modifiers.size();
    const cplx ph = std::exp(I1 * angles[0]);
    M base = eye((size_t)1 << extra);
    for (auto &z : base) z *= ph;
    explicit_ = base;
    matrix_ = finish(base, true);
    return;
  }
  if (targets.size() != want)
    throw std::invalid_argument(name() + " expects " + std::to_string(want) + " target(s), got " +
                                std::to_string(targets.size()));
  matrix_ = finish(base_matrix(angles_), true);
}

GateOperation::GateOperation(const std::vector<size_t> &targets, const M &matrix,
                             const std::vector<size_t> &ctrl_modifiers, long power)
    : Operation(targets), type_(GateOperationType::Unitary), ctrl_modifiers_(ctrl_modifiers),
      power_(power) {
  const size_t base_k = targets.size() - std::min(targets.size(), ctrl_modifiers.size());
  if (ctrl_modifiers.size() > targets.size())
    throw std::invalid_argument("more ctrl_modifiers than targets");
  const size_t d = (size_t)1 << base_k;
  if (matrix.size() != d * d)
    throw std::invalid_argument("`matrix` operates on space of dimension " +
                                std::to_string(isqrt_dim(matrix.size())) + " instead of " +
                                std::to_string(d));
  // unitarity check as in operation_helpers.check_unitary (atol 1e-8, rtol 1e-5)
  M p = matmul(matrix, dagger(matrix, d), d);
  for (size_t i = 0; i < d; ++i)
    for (size_t j = 0; j < d; ++j) {
      const cplx want = (i == j) ? 1.0 : 0.0;
      if (std::abs(p[i * d + j] - want) > 1e-8 + 1e-5 * std::abs(want))
        throw std::invalid_argument("matrix is not unitary");
    }
  explicit_ = matrix;
  matrix_ = finish(matrix, true);
}

size_t GateOperation::num_params() const { return angles_.size(); }

M GateOperation::derivative(size_t i) const {
  if (i >= angles_.size()) throw std::out_of_range("angle index out of range");
  if (power_ != 1) throw std::invalid_argument("derivative of a powered gate is not supported");
  if (type_ == GateOperationType::GPhase && explicit_.size() > 1) {
    M d = explicit_;
    for (auto &z : d) z *= I1;
    return finish(d, false);
  }
  return finish(base_derivative(i), false);
}

std::string GateOperation::formatted_string() const {
  std::stringstream ss;
  ss << name();
  if (!angles_.empty()) {
    ss << "(";
    for (size_t i = 0; i < angles_.size(); ++i) ss << angles_[i] << (i + 1 < angles_.size() ? ", " : "");
    ss << ")";
  }
  ss << " {";
  for (size_t i = 0; i < targets_.size(); ++i) ss << targets_[i] << (i + 1 < targets_.size() ? ", " : "");
  ss << "}";
  return ss.str();
}

bool GateOperation::operator==(const GateOperation &o) const {
  return type_ == o.type_ && targets_ == o.targets_ && angles_ == o.angles_ &&
         ctrl_modifiers_ == o.ctrl_modifiers_ && power_ == o.power_ &&
         (type_ != GateOperationType::Unitary || explicit_ == o.explicit_);
}

}  // namespace snuqs
