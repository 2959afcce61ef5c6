/*
 * snuqs_hip.h -- C ABI of the MI355X state-vector engine (libsnuqs_hip.so).
 *
 * Plain pointers and sizes only.  Every call returns SQ_OK (0) or an error
 * code; sq_last_error() gives the message of the calling thread's last
 * failure.  Nothing here exits the process.
 *
 * Conventions (the reference's, src/braket/snuqs/csrc/operation/
 * gate_operations_impl_cpu.cc:22-66):
 *   - amplitude index i = sum_q b_q * 2^(num_qubits-1-q): qubit 0 is the MSB;
 *   - a gate on targets [t_0..t_{k-1}] is a row-major 2^k x 2^k complex128
 *     matrix whose local index has t_0 as its MSB;
 *   - complex numbers are interleaved (re, im);
 *   - a state is 2^m contiguous amplitudes of dtype SQ_COMPLEX64 (float re,im)
 *     or SQ_COMPLEX128 (double re,im) in device memory.
 *
 * `stream` arguments are hipStream_t values (NULL = the null stream).  All
 * device work is asynchronous on that stream unless the comment says the call
 * returns a host value, in which case the call synchronises the stream.
 *
 * Each declaration names the reference interface it replaces
 * (paths relative to src/braket/snuqs/csrc/).
 */
#ifndef SNUQS_HIP_H_
#define SNUQS_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SQ_OK = 0,
  SQ_ERR_INVALID = 1,     /* bad argument (targets, sizes, dtype, ...)  */
  SQ_ERR_HIP = 2,         /* HIP runtime error                          */
  SQ_ERR_OOM = 3,         /* device allocation failed                   */
  SQ_ERR_UNSUPPORTED = 4  /* valid request outside what the engine does */
};

typedef enum { SQ_COMPLEX64 = 0, SQ_COMPLEX128 = 1 } sq_dtype_t;

/* Message of the calling thread's last failed call ("" if none). */
const char *sq_last_error(void);
/* Library version string. */
const char *sq_version(void);

/* ---- device runtime: core/cuda/runtime.cu:4-25 (cu::mem_info,
 *      device_count, set_device, get_device, device_synchronize) ---------- */
int sq_mem_info(size_t *free_bytes, size_t *total_bytes);
int sq_device_count(int *count);
int sq_set_device(int device);
int sq_get_device(int *device);
int sq_device_synchronize(void);

/* ---- memory: buffer/buffer_cuda.cc:13-20 (cudaMalloc/cudaFree),
 *      buffer/buffer_cpu.cc:13-28 (pinned host), core/runtime.cc:46-65
 *      (memcpyH2D/D2H/D2D) ---------------------------------------------- */
int sq_malloc(void **ptr, size_t bytes);
int sq_free(void *ptr);
int sq_host_alloc_pinned(void **ptr, size_t bytes);
int sq_host_free_pinned(void *ptr);
int sq_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int sq_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int sq_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);

/* Plain device-to-device copy kernel (16-byte loads/stores); the HBM
 * bandwidth ceiling used to calibrate the gate kernels' roofline. */
int sq_copy(void *dst, const void *src, size_t bytes, void *stream);

/* Launch-shape parameters of the kernels (set by the autotuner,
 * snuqs_amd/autotune.py; process-wide), one value per dtype: keys
 * "mfma_blocks_per_cu_c64" / "_c128" (1-4096, persistent grid of the 4-5 target
 * MFMA kernels, default 256), "mfma5_pipeline_c64" / "_c128" (0/1, prefetch
 * the next tile in the 5-target MFMA kernels, default 1) and
 * "mfma5_order_c64" / "_c128" (0 tiles dealt cyclically over the grid,
 * 1 contiguous tile ranges per workgroup, 2 cyclic over the bit-reversed
 * tile index; default 0) and "mfma5_wide_c64"
 * (0/1, 64-group instead of 32-group complex64 tiles: 512 B instead of 256 B
 * runs per access, half the resident waves; default 0) and
 * "qaoa_mid_scatter" (0/1, QAOA batches: window passes above bit 11 take
 * their chunks in bit-reversed order; applies to batches run or captured
 * afterwards) and
 * "tune_scatter" (0/1, sq_program_tune also times tile order 2; default 1) and
 * "adjoint_layers" (0/1, sq_adjoint_gradient applies layers of one-target
 * gates on distinct qubits in passes of up to 12 bits; default 1) and
 * "adjoint_wide_windows" (0/1, those passes take windows of up to 6 high bits
 * with 512 B runs where that needs no more passes than 8-bit windows with
 * 128 B runs; default 1) and "adjoint_fuse_runs" (0-3, diagonal runs folded
 * into the low pass of the layer next to them: bit 0 forward, bit 1 reverse;
 * default 3) and "mfma_tiles_per_wave" (1-64, tiles per wave of the 4-5 target
 * MFMA kernels in programs without a tuned shape and in parameter and circuit
 * batches (which size their grids by it when built), lowered where a launch
 * would drop below 2048 workgroups; the grid cap above still applies;
 * default 4).  A gate tuned by
 * sq_program_tune uses its own shape instead.  No reference counterpart. */
int sq_tune_set(const char *key, int value);
int sq_tune_get(const char *key, int *value);

/* ---- streams and events: stream/stream.cc:6-34, event/event.cc:5-26 ---- */
int sq_stream_create(void **stream, int nonblocking);
int sq_stream_destroy(void *stream);
int sq_stream_synchronize(void *stream);
int sq_stream_wait_event(void *stream, void *event);
int sq_event_create(void **event);
int sq_event_destroy(void *event);
int sq_event_record(void *event, void *stream);
int sq_event_synchronize(void *event);
int sq_event_elapsed_ms(float *ms, void *start, void *stop);

/* ---- initialisation: functionals/functionals.cc:30-50 ->
 *      operation/initialization_impl_cuda.cu:9-38 ------------------------ */
int sq_initialize_zero(void *sv, sq_dtype_t dtype, uint64_t num_elems,
                       void *stream);
int sq_initialize_basis_z(void *sv, sq_dtype_t dtype, uint64_t num_elems,
                          void *stream);

/* ---- gate application: functionals::apply (functionals/functionals.cc:15-28)
 *      -> cu::applyGate (operation/gate_operations_impl_cuda.cu:159-175).
 *
 * Applies `matrix` (row-major complex128, 2^k x 2^k, k = num_targets <= 5) on
 * `targets` of a `num_qubits`-qubit index space to the state `sv` of
 * `num_elems` amplitudes, in place.  As in the reference, num_qubits may
 * exceed log2(num_elems) (the state is then the first slice of a larger
 * space); every target stride 2^(num_qubits-1-t) must be < num_elems.
 * k = 0 multiplies every amplitude by matrix[0] (GPhase).  The gate class
 * (phase / diagonal / permutation / dense, with control qubits factored out)
 * is chosen from the matrix entries.  Asynchronous on `stream`. */
int sq_apply(void *sv, sq_dtype_t dtype, uint64_t num_elems,
             uint32_t num_qubits, const uint32_t *targets,
             uint32_t num_targets, const double *matrix, void *stream);

/* ---- gate programs: a whole gate list uploaded once, run many times.
 * Replaces the per-gate Python loop of simulation.py:274-284 (each gate
 * matrix is classified and copied to the device once, at creation).
 * gates: num_gates entries; gate g has num_targets[g] targets taken in order
 * from `targets` and a (2^k)^2 complex128 matrix taken in order from
 * `matrices`.  k <= 5 for any gate.  A program may be run from several
 * threads at once except through sq_program_tune and sq_program_run_graph,
 * which update the program (its launch shapes, its captured graph). */
int sq_program_create(void **program, sq_dtype_t dtype, uint32_t num_qubits,
                      uint32_t num_gates, const uint32_t *num_targets,
                      const uint32_t *targets, const double *matrices);
int sq_program_destroy(void *program);
/* New matrices for the same gates (no reference counterpart: the reference
 * builds new GateOperation objects, and their device copies, for every
 * parameter set, operation.cc:96-127).  `matrices` has the layout of
 * sq_program_create; every gate keeps its targets.  Allocates nothing: a
 * gate of 4-5 targets owns fixed device slots from creation.  The device
 * copy is ordered after the work already queued on `stream` and the call
 * returns when it is done; work on other streams that runs this program must
 * have finished.  A captured hipGraph is dropped (recaptured on the next
 * sq_program_run_graph); tuned launch shapes are kept. */
int sq_program_set_matrices(void *program, const double *matrices, void *stream);
/* Number of kernels a run launches and its algorithmic HBM bytes
 * (2 x touched amplitudes x itemsize, summed over the gates). */
int sq_program_stats(void *program, uint64_t *num_launches,
                     double *algorithmic_bytes);
int sq_program_run(void *program, void *sv, uint64_t num_elems, void *stream);
/* sq_program_run through a hipGraph (no reference counterpart; the reference
 * launches one kernel per Python call, simulation.py:274-284): the launch
 * sequence is captured on the first call for a state and replayed on later
 * calls with the same state; another state (or sq_program_tune) recaptures.
 * Asynchronous on `stream` like sq_program_run. */
int sq_program_run_graph(void *program, void *sv, uint64_t num_elems, void *stream);
/* Measurement variant of sq_program_run: records a HIP event before every
 * launch and one after the last on `stream`, waits for completion and writes
 * each gate's elapsed time to ms_per_gate[num_gates] (0 for gates that
 * launch nothing). */
int sq_program_run_timed(void *program, void *sv, uint64_t num_elems,
                         void *stream, float *ms_per_gate);
/* Per-gate autotuner (no reference counterpart): for every gate that runs
 * on the 5-target MFMA kernels, time each candidate launch shape (tile order
 * x grid size x prefetch) inside `reps` whole-program runs on sv and keep each
 * gate's fastest for the gate and its inverse.  Choices are cached process-wide by dtype, state
 * size and target/control bits, so later programs with the same gates reuse
 * them without timing.  Results do not depend on the shape (the same
 * arithmetic runs in every shape).  sv is overwritten: the gates are applied
 * to it many times.  *num_tuned (may be NULL) = gates given a shape. */
int sq_program_tune(void *program, void *sv, uint64_t num_elems, int reps,
                    void *stream, uint32_t *num_tuned);
/* Kernel class of gate g (0 none, 1 phase, 2 diagonal, 3 permutation,
 * 4 dense 1-3 targets, 5 dense 4-5 targets on MFMA, 6 dense 4-5 targets on
 * VALU), its target count after control extraction and its algorithmic HBM
 * bytes (2 x touched amplitudes x itemsize). */
int sq_program_gate_info(void *program, uint32_t gate, int *kernel_class,
                         int *num_targets, double *algorithmic_bytes);

/* ---- parameter batches (new; the reference has no batching: its
 * optimize_operations stub, transpile.py:233-234, is where it would go).
 * One gate structure, num_sets parameter sets: the states are num_sets
 * consecutive blocks of num_elems = 2^num_qubits amplitudes in one device
 * buffer (state y at states + y * num_elems), and every gate of the batch
 * runs as ONE launch over all of them, each state with its own matrix.
 * `matrices` holds set 0's gate matrices in the layout of sq_program_create,
 * then set 1's, and so on; the sets must use the same targets per gate.  A
 * gate whose sets' matrices fall into different classes (e.g. an angle of 0
 * in one set) runs as a dense gate in every set.  1 <= num_sets <= 65535. */
int sq_batch_create(void **batch, sq_dtype_t dtype, uint32_t num_qubits,
                    uint32_t num_sets, uint32_t num_gates,
                    const uint32_t *num_targets, const uint32_t *targets,
                    const double *matrices);
int sq_batch_destroy(void *batch);
/* New matrices for the same gates (layout of sq_batch_create).  Ordered
 * after the work already queued on `stream`; returns when the device copy
 * is done.  Reallocates the batch's argument tables only when a gate's plan
 * grows.  A captured hipGraph is dropped. */
int sq_batch_set_matrices(void *batch, const double *matrices, void *stream);
/* Apply every gate to every state (asynchronous on `stream`). */
int sq_batch_run(void *batch, void *states, uint64_t num_elems, void *stream);
/* sq_batch_run through a hipGraph captured on the first call for `states`
 * and replayed on later calls with the same states. */
int sq_batch_run_graph(void *batch, void *states, uint64_t num_elems,
                       void *stream);
/* Launches per run and algorithmic HBM bytes per run (all states). */
int sq_batch_stats(void *batch, uint64_t *num_launches,
                   double *algorithmic_bytes);
/* |0...0> in each of num_sets consecutive states (one launch). */
int sq_initialize_basis_z_batch(void *states, sq_dtype_t dtype,
                                uint64_t num_elems, uint32_t num_sets,
                                void *stream);
/* out[y] = sum_t coeffs[t] <psi_y|P_t|psi_y> for the num_sets consecutive
 * states (terms as in sq_expect_pauli_sum), FP64; term_values (host,
 * num_sets x num_terms, may be NULL) gets the per-term values.  Launches
 * cover every state at once.  Returns when the values are on the host. */
int sq_expect_pauli_sum_batch(const void *states, sq_dtype_t dtype,
                              uint64_t num_elems, uint32_t num_sets,
                              uint32_t num_terms, const uint64_t *xmask,
                              const uint64_t *zmask, const uint32_t *ny,
                              const double *coeffs, double *out,
                              double *term_values, void *stream);

/* ---- states larger than HBM: the reference's hybrid and CPU-offload tiers
 * (simulation.py:286-351, transpile.py:184-298) on one GPU.  The state is
 * num_elems amplitudes in pinned host memory (sq_host_alloc_pinned).
 *
 * sq_host_tier_pass: one streaming pass.  local_bits[0..num_local-1] are the
 * host index bits that form a device slice (device bit j = host bit
 * local_bits[j]); they must start with 0, 1, ..., k-1 for some k >= 1 (the
 * slice is gathered as 2^(num_local-k) contiguous chunks of 2^k
 * amplitudes).  For every assignment of the other bits, the slice is copied
 * into one of the num_buffers device buffers (each 2^num_local amplitudes),
 * `program` (built for num_local qubits, gates on device bits) runs on it,
 * and it is copied back; copies in, compute and copies out overlap on three
 * streams.  zero_slices != 0 skips every slice but the first (correct when
 * the state is |0...0>: the other slices are zero and stay zero).  Ordered
 * after the work queued on `stream`; returns when the host state is
 * written. */
int sq_host_tier_pass(void *program, void *host_state, sq_dtype_t dtype,
                      uint64_t num_elems, uint32_t num_local,
                      const uint32_t *local_bits, void *const *buffers,
                      uint32_t num_buffers, int zero_slices, void *stream);
/* |0...0> in a host state, written by `threads` host threads. */
int sq_host_initialize_basis_z(void *host_state, sq_dtype_t dtype,
                               uint64_t num_elems, uint32_t threads);

/* ---- results: simulation.py:196-220 (state_vector, probabilities,
 *      retrieve_samples) and the expectation/inner-product reductions that
 *      the reference lacks (result_types.py:149-151 registers StateVector
 *      only).  Reductions accumulate in FP64 for both dtypes. --------------- */

/* probs[i] = |sv[i]|^2 in float64, device -> device (np.abs(sv)**2). */
int sq_probabilities(const void *sv, sq_dtype_t dtype, uint64_t num_elems,
                     double *probs, void *stream);
/* probs[i] = |sv[i]|^2 / sum_j |sv[j]|^2, all in FP64 (device -> device):
 * the distribution of a complex64 state whose FP32 norm has drifted, summing
 * to 1 within FP64 rounding (np.abs(sv)**2 of a normalised state,
 * simulation.py:208-211; numpy's choice needs the sum within sqrt(eps)).
 * Returns when probs is written. */
int sq_probabilities_normalized(const void *sv, sq_dtype_t dtype,
                                uint64_t num_elems, double *probs,
                                void *stream);
/* Marginal distribution over k (1..20) target qubits: out[r] (host, 2^k
 * doubles) = sum of |sv[i]|^2 over the i whose target qubits read r, targets[0]
 * the most significant bit of r (Braket's Probability result type with
 * targets; the reference returns only the state vector,
 * simulator.py:578-580).  FP64, fixed summation order. */
int sq_marginal_probabilities(const void *sv, sq_dtype_t dtype, uint64_t num_elems,
                              uint32_t num_qubits, const uint32_t *targets,
                              uint32_t num_targets, double *out, void *stream);
/* sum_i |sv[i]|^2 in FP64; returns a host value. */
int sq_norm2(const void *sv, sq_dtype_t dtype, uint64_t num_elems,
             double *out, void *stream);
/* sv *= (re + i im) in place (norm renormalisation, global phase). */
int sq_scale(void *sv, sq_dtype_t dtype, uint64_t num_elems, double re,
             double im, void *stream);
/* out[2*j], out[2*j+1] = sv[indices[j]] as float64 (host arrays). */
int sq_amplitudes(const void *sv, sq_dtype_t dtype, uint64_t num_elems,
                  const uint64_t *indices, uint64_t count, double *out,
                  void *stream);
/* <a|b> in FP64, returns (re, im) on the host. */
int sq_inner(const void *a, const void *b, sq_dtype_t dtype,
             uint64_t num_elems, double *out, void *stream);
/* sum_t coeffs[t] <sv| P_t |sv> in FP64; P_t given by xmask[t], zmask[t]
 * (amplitude-index bit masks: X or Y set the x bit, Z or Y the z bit) and the
 * number of Y factors ny[t].  Returns the real part on the host; per-term
 * values go to term_values (host, may be NULL). */
int sq_expect_pauli_sum(const void *sv, sq_dtype_t dtype, uint64_t num_elems,
                        uint32_t num_terms, const uint64_t *xmask,
                        const uint64_t *zmask, const uint32_t *ny,
                        const double *coeffs, double *out,
                        double *term_values, void *stream);
/* sq_expect_pauli_sum without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate memory (code under hipGraph
 * capture, tight serving loops, hosts that forbid hipMalloc).  Same arguments,
 * kernels and result; the partial sums and term tables live in a fixed
 * per-device array of about 1 MiB, held until the process ends, and the host
 * staging lives on the stack.  At most 256 terms per call (more returns
 * SQ_ERR_INVALID).  Every call on one GPU shares that GPU's array, so calls
 * that overlapped on one GPU would race: the library takes a per-device lock
 * around each call and synchronises `stream` before returning, so concurrent
 * callers wait for one another instead. */
int sq_expect_pauli_sum_noalloc(const void *sv, sq_dtype_t dtype,
                                uint64_t num_elems, uint32_t num_terms,
                                const uint64_t *xmask, const uint64_t *zmask,
                                const uint32_t *ny, const double *coeffs,
                                double *out, double *term_values,
                                void *stream);
/* out = sum_t coeffs[t] P_t |in>  (device -> device, out != in). */
int sq_apply_pauli_sum(const void *in, void *out, sq_dtype_t dtype,
                       uint64_t num_elems, uint32_t num_terms,
                       const uint64_t *xmask, const uint64_t *zmask,
                       const uint32_t *ny, const double *coeffs, void *stream);

/* Sampling: simulation.py:217-220 (np.random.choice(N, p=|sv|^2, size)).
 * For each host uniform u[s] in [0,1), out[s] = the first index i with
 * cdf[i] > u[s], where cdf = cumsum(p) / cumsum(p)[N-1], p = probs (device,
 * float64, from sq_probabilities).  exact != 0 reproduces numpy's sequential
 * float64 cumsum bit for bit; exact == 0 uses an FP64 parallel scan.  out is
 * a host array of int64.  Returns when the samples are on the host. */
int sq_sample(const double *probs, uint64_t num_elems, const double *uniforms,
              uint64_t shots, int exact, int64_t *out, void *stream);

/* ---- adjoint differentiation (new; no reference counterpart).
 * For the program U = G_{m-1}...G_0 and observable H = sum_t c_t P_t the
 * call runs the forward pass itself (sv <- U|0..0>), returns the energy
 * <psi|H|psi> in *energy and
 *   grads[p] = sum over derivative entries d with deriv_param[d] == p of
 *              2 Re <psi| H U_{>g} dG_d U_{<g} |0..0>,   g = deriv_gate[d],
 * i.e. d<psi|H|psi>/d theta_p when dG_d = dG_g/d theta_p (a gate with several
 * parameters has one entry per parameter; a parameter shared by several gates
 * accumulates).  dmatrices holds the entries' (2^k)^2 complex128 matrices
 * back to back, k = targets of gate deriv_gate[d].
 * The reverse sweep applies G_g^dagger to the state and to lambda = H psi
 * (one FP64 inner product per entry).  num_checkpoints device buffers of
 * num_elems amplitudes store forward states at evenly spaced gates; the
 * reverse sweep continues from them instead of undoing those segments,
 * bounding the drift of complex64 reverse sweeps.  work0/work1: scratch
 * states.  After the call sv, work0, work1 and the checkpoints hold
 * unspecified states (the reverse sweep moves onto each checkpoint's buffer
 * instead of copying it back).
 * Host arrays: deriv_*, dmatrices, xmask/zmask/ny/coeffs, energy, grads. */
int sq_adjoint_gradient(void *program, uint32_t num_derivs,
                        const uint32_t *deriv_gate, const uint32_t *deriv_param,
                        const double *dmatrices, uint32_t num_params,
                        void *sv, void *work0, void *work1,
                        void *const *checkpoints, uint32_t num_checkpoints,
                        uint64_t num_elems, uint32_t num_terms,
                        const uint64_t *xmask, const uint64_t *zmask,
                        const uint32_t *ny, const double *coeffs,
                        double *energy, double *grads, void *stream);

/* ---- batched QAOA MaxCut (new; no reference counterpart).
 * num_instances circuits of uneven qubit counts num_qubits[j] (2..30) on
 * simple graphs with num_edges[j] edges (qubit pairs, back to back in
 * `edges`; a pair given twice, in either order, is rejected with
 * SQ_ERR_INVALID):
 * H on every qubit, then p layers of ZZ(2 gamma_l) on every edge and
 * RX(2 beta_l) on every qubit.  A run computes for every instance the energy
 * <C> = sum_edges (1 - <Z_a Z_b>) / 2 and its adjoint gradient.  Every launch
 * serves all instances (one pass of each); the launch sequence is captured
 * once into a hipGraph and replayed by later runs (use_graph != 0).
 * out: num_instances x (1 + 2p) host doubles: E, dE/dgamma_0..p-1,
 * dE/dbeta_0..p-1.  gammas/betas: num_instances x p host doubles.
 * complex64 states with FP64 reductions (complex128 also accepted). */
int sq_qaoa_create(void **batch, sq_dtype_t dtype, uint32_t num_instances,
                   const uint32_t *num_qubits, const uint32_t *num_edges,
                   const uint32_t *edges, uint32_t p);
int sq_qaoa_run(void *batch, const double *gammas, const double *betas,
                int use_graph, double *out, void *stream);
/* Kernel launches per run, bytes of all states, HBM bytes one run moves. */
int sq_qaoa_stats(void *batch, uint64_t *launches, double *state_bytes,
                  double *traffic_bytes);
int sq_qaoa_destroy(void *batch);

/* ---- Circuits of uneven qubit counts as one launch sequence (no reference
 * counterpart: the reference runs one circuit per simulation and has no
 * batching; transpile.py:233-234 is where it would go).  For callers with
 * many circuits, each too small to fill the GPU alone (e.g. the circuits of
 * an ensemble, a sweep or a training batch).
 * num_circuits circuits; circuit j has num_qubits[j] (1..30) qubits and
 * num_gates[j] gates; num_targets (1..5 each), targets and matrices hold
 * every gate of circuit 0, then of circuit 1, ..., in the layout of
 * sq_program_create (reference qubit order, targets[0] the most significant
 * bit of a matrix's local index; 2^k x 2^k complex128 row-major).  Every
 * gate runs as a dense matrix on its targets.  The batch owns the states
 * (one device buffer).  sq_cbatch_run sets every state to |0...0> and
 * applies every circuit: step s applies gate s of every circuit that has
 * one, in one launch for the gates that run on the 5-target matrix-core
 * kernels and one for the others (at most 2 max(num_gates) + 2 launches,
 * captured into a hipGraph on the first use_graph call and replayed after).
 * Asynchronous on `stream`.
 * Errors (SQ_ERR_INVALID): no circuits, a qubit count outside 1..30, a gate
 * with 0 or more than 5 targets or more targets than qubits, repeated or
 * out-of-range targets. */
int sq_cbatch_create(void **batch, sq_dtype_t dtype, uint32_t num_circuits,
                     const uint32_t *num_qubits, const uint32_t *num_gates,
                     const uint32_t *num_targets, const uint32_t *targets,
                     const double *matrices);
int sq_cbatch_run(void *batch, int use_graph, void *stream);
/* Device-to-device copy of circuit j's state into dst (a device state of
 * 2^num_qubits[j] elements of the batch's dtype), ordered on `stream`. */
int sq_cbatch_copy_state(void *batch, uint32_t circuit, void *dst,
                         uint64_t num_elems, void *stream);
/* Kernel launches per run, bytes of all states, HBM bytes one run moves. */
int sq_cbatch_stats(void *batch, uint64_t *launches, double *state_bytes,
                    double *traffic_bytes);
int sq_cbatch_destroy(void *batch);

#ifdef __cplusplus
}
#endif
#endif /* SNUQS_HIP_H_ */
