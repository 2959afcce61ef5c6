// kernels_reduce.hpp -- launchers of the result kernels (kernels_reduce.hip)
// and the sampler (kernels_sample.hip).
#pragma once

#include "sq_common.hpp"

namespace sq {

constexpr int kRedBlocks = 2048;  // partial sums per reduced value
constexpr int kMaxMarginal = 20;  // targets of a marginal distribution (2^20 outcomes)

int launch_probabilities(const void *sv, sq_dtype_t dt, uint64_t n, double *p,
                         hipStream_t st);
int launch_probabilities_div(const void *sv, sq_dtype_t dt, uint64_t n, double *p,
                             const double *norm2_dev, hipStream_t st);
// out_dev[2^k] = marginal distribution of |sv|^2 over the target bits tbit[]
// (tbit[0] the most significant bit of the outcome); partial holds
// 2^k * nb doubles.
int launch_marginal(const void *sv, sq_dtype_t dt, uint64_t n, int nbits, const int *tbit,
                    int k, double *partial, int nb, double *out_dev, hipStream_t st);
int launch_norm2(const void *sv, sq_dtype_t dt, uint64_t n, double *partial,
                 double *out_dev, hipStream_t st);
// out_dev[0..1] = beta*out_dev + alpha_dev[0..1] * (re, im) of <a|b>
int launch_inner(const void *a, const void *b, sq_dtype_t dt, uint64_t n,
                 double *partial, double *out_dev, const double *alpha_dev,
                 double beta, hipStream_t st);
// Fused reverse step of a 1-target gate on amplitude bit `bit` (ginv = G^dagger,
// dgm = dG, row-major complex128 2x2): psi <- ginv psi (if undo_psi),
// out_dev[0] += alpha_dev[0] * Re <lam| dgm |psi>, lam <- ginv lam.
int launch_adjoint_step1(void *psi, void *lam, sq_dtype_t dt, uint64_t n, int bit,
                         const double *ginv, const double *dgm, int undo_psi, double *partial,
                         double *out_dev, const double *alpha_dev, hipStream_t st);
// Reverse step over a run of m diagonal gates (each on 1 or 2 targets): psi
// and lambda both multiplied by the product of the gates' G^dagger diagonals,
// and for every gate with a slot, out_dev[2 slot_param[slot]] += 2 Re
// sum_i conj(lam_i) psi_i dgg(i) at the run's end point (dgg = diagonal of
// dG G^dagger; diagonal gates commute, so this is each gate's sequential
// derivative term).  gates_dev: m DiagRunGate records; slot_param_dev: the
// parameter of each of the nslots slots.  lam == nullptr (and nslots == 0):
// psi alone is multiplied by the records' `ginv` diagonals (used with the
// gates' own diagonals for a forward run).  apply == false (reverse only):
// the derivative terms alone, both states left unchanged.
constexpr int kMaxDiagRun = 64;
constexpr int kMaxDiagSlots = 32;
struct DiagRunGate {
  int b0, b1;   // address bits of local index bit 1 / bit 0 (b1 < 0: one target, local = bit b0)
  int slot;     // partial row of this gate's derivative, or -1
  int uniform;  // 1: both target bits >= 10 (constant over a 1024-amplitude chunk)
  double ginv[4][2];  // diagonal of G^dagger, by local index
  double dgg[4][2];   // diagonal of dG G^dagger, by local index (slot >= 0)
};
// distinct: the slots' parameters differ (their sums may be finalised in
// parallel); otherwise one block adds them in order.
int launch_adjoint_diag_run(void *psi, void *lam, sq_dtype_t dt, uint64_t n,
                            const DiagRunGate *gates_dev, int m, int nslots,
                            const int *slot_param_dev, bool distinct, double *partial,
                            double *out_dev, hipStream_t st, bool apply = true);
// Reverse step over K = 2..4 one-target gates on distinct amplitude bits
// (bits[j] = bit of gate j, none of the lowest 4): psi and lambda undone by
// every gate's G^dagger (ginv, row-major complex128 2x2 each), and for gate j
// out_dev[2 slot_param[j]] += 2 Re <lam| dG_j G_j^dagger |psi> at the group's
// end point (the gates commute).  One pass; partial needs K x kRedBlocks.
// lam == nullptr: forward group, psi <- (every `ginv` matrix) psi, where the
// caller passes the gates themselves; no reduction.
int launch_adjoint_step_group(void *psi, void *lam, sq_dtype_t dt, uint64_t n, int k,
                              const int *bits, const double *ginv, const double *dgg,
                              const int *slot_param_dev, bool distinct, double *partial,
                              double *out_dev, hipStream_t st);
// Partial rows per slot of the adjoint passes (nslots x this many doubles).
// With out_dev == nullptr the passes leave their rows unfinalised (see
// launch_finalize_rows).
int diag_run_blocks(uint64_t n);
int step_group_blocks(uint64_t n, int k, const int *bits);
// Deferred finalisation of rows: out[2p] += alpha * sum over parameter p's
// rows (in the order given) of the row's nb doubles at partial + row_off.
int launch_finalize_rows(const double *partial, const uint64_t *row_off_dev,
                         const int *row_nb_dev, const int *param_ptr_dev, int nparams,
                         double alpha, double *out, hipStream_t st);
// out[2 slot_param[v]] += alpha * sum_b partial[v * nblocks + b], v < nslots
// (distinct: the slots' parameters differ and are summed in parallel).
int launch_finalize_slots(const double *partial, int nblocks, int nslots,
                          const int *slot_param_dev, bool distinct, double alpha, double *out,
                          hipStream_t st);
// Layer pass over a state of n = 2^nq amplitudes (n >= 4096): count one-target
// gates at distinct local positions pos[j] (0..11) of 4096-amplitude chunks
// whose local bits map to global bits as: 0..C-1 -> 0..C-1, C..C+W-1 ->
// h0..h0+W-1, C+W..11 -> C..11-W, with C = 4 or 6 contiguous low bits
// (W = 0, h0 = 12: local = global, bits 0-11; else 12 <= h0, h0 + W <= nq,
// W <= 12 - C).  gm: each gate's 2x2 (row-major complex128).
// lam == nullptr (forward): psi <- gm_j psi for every j.  Reverse: gm_j =
// G_j^dagger, dm_j = dG_j G_j^dagger; for slot[j] >= 0,
// out_dev[2 slot_param[slot[j]]] += 2 Re <lam| dm_j |psi> at the layer's end
// point, and both states are undone by every gm_j.  partial needs nslots x
// layer_pass_blocks(n) doubles.
constexpr int kLayerMaxBlocks = 8192;
// A diagonal run folded into a low pass (h0 = 12, W = 0): every amplitude i
// is multiplied by the product over the run's gates of their diagonal entry
// at i, as tl[i & 4095] (gates on bits 0-11) x tc[i >> 12] (gates on bits 12
// and above) x the nmix gates with one bit on each side (records as the run's
// DiagRunGate: ginv holds the entries applied).  The tables hold complex
// values of the state's precision.  Reverse: before the layer's gates;
// forward: after them.
constexpr int kFuseMix = 4;
struct FuseDesc {
  const void *tl, *tc;
  int nmix, pad;
  DiagRunGate mix[kFuseMix];
};
// Tables of fused runs: job j fills tl (4096 entries) and tc (n >> 12
// entries) from its m gate records, products in FP64 rounded once.
struct FuseTableJob {
  const DiagRunGate *g;
  int m, pad;
  void *tl, *tc;
};
int launch_fuse_tables(const FuseTableJob *jobs_dev, int njobs, sq_dtype_t dt, uint64_t n,
                       hipStream_t st);
int layer_pass_blocks(uint64_t n);
// fuse_dev: a FuseDesc in device memory (low pass only), or nullptr.
int launch_adjoint_layer(void *psi, void *lam, sq_dtype_t dt, uint64_t n, int h0, int W, int C,
                         int count, const int *pos, const double *gm, const double *dm,
                         const int *slot, int nslots, const int *slot_param_dev, bool distinct,
                         double *partial, double *out_dev, hipStream_t st,
                         const FuseDesc *fuse_dev = nullptr);
int launch_pauli_terms(const void *sv, sq_dtype_t dt, uint64_t n, int nterms,
                       const void *terms_dev, const double *alpha_dev,
                       double *partial, double *out_dev, hipStream_t st);
int launch_apply_pauli_sum(const void *in, void *out, sq_dtype_t dt, uint64_t n,
                           int nterms, const void *terms_dev, hipStream_t st);
int launch_scale(void *sv, sq_dtype_t dt, uint64_t n, double re, double im,
                 hipStream_t st);
int launch_gather(const void *sv, sq_dtype_t dt, const uint64_t *idx_dev,
                  uint64_t count, double *out_dev, hipStream_t st);

size_t pauli_term_bytes();
// Batch of nsets consecutive n-amplitude states: out_dev[z * out_stride + t] =
// alpha[t] * <psi_z|P_t|psi_z>.  partial: pauli_batch_partial_doubles(...).
int launch_pauli_terms_batch(const void *sv, sq_dtype_t dt, uint64_t n, uint32_t nsets,
                             int nterms, const void *terms_dev, const double *alpha_dev,
                             double *partial, double *out_dev, uint64_t out_stride,
                             hipStream_t st);
size_t pauli_batch_partial_doubles(uint64_t n, uint32_t nsets, int nterms);
size_t pauli_apply_term_bytes();
void fill_pauli_terms(int nterms, const uint64_t *x, const uint64_t *z,
                      const uint32_t *ny, const double *coeffs, void *terms_host,
                      double *alpha_host);
void fill_pauli_apply_terms(int nterms, const uint64_t *x, const uint64_t *z,
                            const uint32_t *ny, const double *coeffs,
                            void *terms_host);

// Sampling.  cdf (device, n doubles) receives cumsum(p); exact selects the
// bit-exact sequential-order cumsum.  u_dev / out_dev: shots entries.
int sample_cdf(const double *p, uint64_t n, double *cdf, bool exact,
               void *scratch, size_t scratch_bytes, hipStream_t st);
size_t sample_scratch_bytes(uint64_t n, bool exact);
int launch_search(const double *cdf, uint64_t n, const double *u_dev,
                  uint64_t shots, int64_t *out_dev, hipStream_t st);

}  // namespace sq
