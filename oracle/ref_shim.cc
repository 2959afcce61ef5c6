// ref_shim.cc -- TEST INFRASTRUCTURE ONLY.
//
// extern "C" entry points over the reference's own CPU gate path, compiled
// from the reference sources where they lie (see oracle/Makefile):
//   cpu::applyGate         src/braket/snuqs/csrc/operation/gate_operations_impl_cpu.cc:131-148
//   cpu::initializeZero    src/braket/snuqs/csrc/operation/initialization_impl_cpu.cc:7-13
//   cpu::initializeBasis_Z src/braket/snuqs/csrc/operation/initialization_impl_cpu.cc:15-22
// Used to pin the oracle restatement (bitwise on complex128) and as the
// reference CPU baseline in bench.py.  Never part of the product.
#include <cstddef>
#include <vector>

namespace cpu {
void applyGate(void *_buffer, void *_gate, std::vector<size_t> targets,
               size_t nqubits, size_t nelems);
void initializeZero(void *_buf, size_t nelems);
void initializeBasis_Z(void *_buf, size_t nelems);
}  // namespace cpu

extern "C" {
void ref_apply_gate(void *sv, void *gate, const size_t *targets,
                    size_t ntargets, size_t nqubits, size_t nelems) {
  std::vector<size_t> t(targets, targets + ntargets);
  cpu::applyGate(sv, gate, t, nqubits, nelems);
}
void ref_initialize_zero(void *sv, size_t nelems) {
  cpu::initializeZero(sv, nelems);
}
void ref_initialize_basis_z(void *sv, size_t nelems) {
  cpu::initializeBasis_Z(sv, nelems);
}
}
