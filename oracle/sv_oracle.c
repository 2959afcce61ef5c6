/*
 * sv_oracle.c -- CPU restatement of the SnuQS state-vector algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP engine in
 * snuqs_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product path never calls it.
 *
 * What it restates (reference file:line, paths relative to src/braket/snuqs/):
 *   sqo_apply_c128 / sqo_apply_c64
 *       csrc/operation/gate_operations_impl_cpu.cc:10-148 (cpu::applyGate and
 *       applyGlobalPhase / applyOneQubitGate / applyTwoQubitGate /
 *       applyThreeQubitGate): loop over every index i < nelems, skip i unless all
 *       target bits are 0, gather the 2^k amplitudes at i + (sum of a subset of
 *       target strides), stride(t) = 2^(nqubits-1-t), targets[0] is the MSB of the
 *       gate-local index, out_r = sum_c G[r*2^k+c]*a_c summed left to right, in
 *       place.  Generalised from k<=3 to k<=MAXK with the same index math.
 *   sqo_init_zero / sqo_init_basis_z
 *       csrc/operation/initialization_impl_cpu.cc:7-22.
 *   sqo_probabilities
 *       simulation.py:208-214  (np.abs(state_vector) ** 2: hypot, then square;
 *       numpy's complex abs is its own routine, so results agree with numpy to
 *       a few ulp, not bitwise).
 *   sqo_cumsum / sqo_choice
 *       simulation.py:217-220  (np.random.choice(len(sv), p=probabilities,
 *       size=shots)).  numpy's legacy RandomState.choice with p draws
 *       u = random_sample(shots) and returns searchsorted(cdf, u, side='right')
 *       with cdf = cumsum(p) / cumsum(p)[-1]; cumsum is a sequential float64
 *       accumulation.  (Restated from numpy's published behaviour and pinned by
 *       black-box fixtures under tests/golden/.)
 *   sqo_expect_pauli
 *       <psi| P |psi> for a Pauli string P; the reference has no expectation
 *       result type (result_types.py:149-151), so this is the textbook formula
 *       P|i> = i^{nY} (-1)^{popcount(i & zmask)} |i ^ xmask>, pinned in tests by
 *       applying the X/Y/Z gate matrices of gate_operations.cc with sqo_apply.
 *   sqo_inner
 *       <a|b> = sum_i conj(a_i) b_i, sequential.
 */
#include <complex.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#define MAXK 6

/* Offsets of the 2^k gathered amplitudes relative to the base index, in the
 * reference's gate-local order: bit (k-1-j) of r selects targets[j]. */
static void local_offsets(const size_t *targets, size_t k, size_t nqubits,
                          size_t *off, size_t *mask) {
  size_t m = 0;
  for (size_t r = 0; r < ((size_t)1 << k); ++r) {
    size_t o = 0;
    for (size_t j = 0; j < k; ++j) {
      size_t st = (size_t)1 << (nqubits - targets[j] - 1);
      if ((r >> (k - 1 - j)) & 1) o += st;
    }
    off[r] = o;
  }
  for (size_t j = 0; j < k; ++j) m |= (size_t)1 << (nqubits - targets[j] - 1);
  *mask = m;
}

#define DEFINE_APPLY(NAME, CT)                                                  \
  void NAME(CT *sv, const CT *gate, const size_t *targets, size_t k,           \
            size_t nqubits, size_t nelems) {                                   \
    if (k == 0) {                                                               \
      CT g = gate[0];                                                           \
      _Pragma("omp parallel for") for (size_t i = 0; i < nelems; ++i) sv[i] =   \
          sv[i] * g;                                                            \
      return;                                                                   \
    }                                                                           \
    size_t off[1 << MAXK], mask;                                                \
    size_t dim = (size_t)1 << k;                                                \
    local_offsets(targets, k, nqubits, off, &mask);                             \
    _Pragma("omp parallel for") for (size_t i = 0; i < nelems; ++i) {          \
      if ((i & mask) != 0) continue;                                            \
      CT a[1 << MAXK];                                                          \
      for (size_t c = 0; c < dim; ++c) a[c] = sv[i + off[c]];                   \
      for (size_t r = 0; r < dim; ++r) {                                        \
        CT acc = gate[r * dim + 0] * a[0];                                      \
        for (size_t c = 1; c < dim; ++c) acc = acc + gate[r * dim + c] * a[c];  \
        sv[i + off[r]] = acc;                                                   \
      }                                                                         \
    }                                                                           \
  }

DEFINE_APPLY(sqo_apply_c128, double complex)
DEFINE_APPLY(sqo_apply_c64, float complex)

void sqo_init_zero_c128(double complex *sv, size_t nelems) {
#pragma omp parallel for
  for (size_t i = 0; i < nelems; ++i) sv[i] = 0;
}

void sqo_init_basis_z_c128(double complex *sv, size_t nelems) {
  sqo_init_zero_c128(sv, nelems);
  sv[0] = 1;
}

void sqo_probabilities_c128(const double complex *sv, double *p, size_t n) {
#pragma omp parallel for
  for (size_t i = 0; i < n; ++i) {
    double h = hypot(creal(sv[i]), cimag(sv[i]));
    p[i] = h * h;
  }
}

/* Sequential float64 cumulative sum (numpy's add.accumulate order). */
void sqo_cumsum(const double *p, double *c, size_t n) {
  double acc = 0.0;
  for (size_t i = 0; i < n; ++i) {
    acc = (i == 0) ? p[0] : acc + p[i];
    c[i] = acc;
  }
}

/* searchsorted(cdf, u, side='right') on cdf = cumsum(p) / cumsum(p)[-1]. */
void sqo_choice(const double *p, size_t n, const double *u, size_t shots,
                double *scratch_cdf, int64_t *out) {
  sqo_cumsum(p, scratch_cdf, n);
  double last = scratch_cdf[n - 1];
  for (size_t i = 0; i < n; ++i) scratch_cdf[i] = scratch_cdf[i] / last;
  for (size_t s = 0; s < shots; ++s) {
    size_t lo = 0, hi = n; /* first index with cdf[idx] > u */
    double v = u[s];
    while (lo < hi) {
      size_t mid = lo + (hi - lo) / 2;
      if (scratch_cdf[mid] <= v) lo = mid + 1; else hi = mid;
    }
    out[s] = (int64_t)lo;
  }
}

static int popcount64(uint64_t x) { return __builtin_popcountll(x); }

/* <psi|P|psi>; xmask/zmask are in amplitude-index bit space (qubit q is bit
 * nqubits-1-q), ny = number of Y factors.  Result written as (re, im). */
void sqo_expect_pauli_c128(const double complex *sv, size_t n, uint64_t xmask,
                           uint64_t zmask, int ny, double *out) {
  static const double complex ipow[4] = {1.0, I, -1.0, -I};
  double complex ph = ipow[ny & 3];
  double re = 0.0, im = 0.0; /* FP64 partial sums per thread */
#pragma omp parallel for reduction(+ : re, im)
  for (size_t i = 0; i < n; ++i) {
    double complex t = conj(sv[i ^ xmask]) * sv[i];
    if (popcount64(i & zmask) & 1) t = -t;
    re += creal(t);
    im += cimag(t);
  }
  double complex acc = (re + I * im) * ph;
  out[0] = creal(acc);
  out[1] = cimag(acc);
}

void sqo_inner_c128(const double complex *a, const double complex *b, size_t n,
                    double *out) {
  double complex acc = 0;
  for (size_t i = 0; i < n; ++i) acc += conj(a[i]) * b[i];
  out[0] = creal(acc);
  out[1] = cimag(acc);
}
