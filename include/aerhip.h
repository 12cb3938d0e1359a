/*
 * aerhip.h -- C ABI of the MI355X statevector engine.
 *
 * The engine replaces the QubitVector method family of Qiskit Aer's
 * statevector simulator (the `statevec_t` template parameter of
 * `Statevector::State<statevec_t>`, src/simulators/statevector/
 * statevector_state.hpp:100-101).  Every entry point below names the
 * reference method it stands in for.  Conventions are the reference's:
 *
 *   - amplitudes are complex, interleaved (re, im); bit q of an amplitude
 *     index is qubit q (little endian);
 *   - matrices are column-major vectorized, mat[i + DIM*j] = M(i, j), and
 *     bit k of a row/column index belongs to qubits[k]
 *     (src/simulators/statevector/transformer.hpp:159,
 *      src/simulators/statevector/indexes.hpp:175-250);
 *   - controlled gates list the control qubits first, the target(s) last
 *     (qubitvector.hpp:1450-1451, 1619-1621);
 *   - in Pauli strings character N-1-i acts on qubits[i]
 *     (qubitvector.hpp:2247).
 *
 * Complex values cross the ABI as pairs of doubles.  All host buffers are
 * borrowed for the duration of the call.  Every function returning int
 * returns 0 on success and a non-zero code on failure; the message is then
 * available from aerhip_last_error() on the calling thread (the reference
 * throws std::runtime_error / std::invalid_argument with the same texts).
 *
 * Gates are queued on the handle and executed lazily: any readout
 * (norm, probabilities, sampling, expectation values, copy-out) or
 * aerhip_qv_synchronize() flushes the queue, which the engine plans into
 * tile passes (several gates per HBM sweep).  A handle is used by one host
 * thread at a time; separate handles are independent (one HIP stream each).
 */
#ifndef AERHIP_H
#define AERHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Precision of the amplitudes (reference: QubitVector<double> / <float>,
 * selected by the "precision" option, src/framework/config.hpp:74). */
#define AERHIP_PRECISION_DOUBLE 0 /* complex128 */
#define AERHIP_PRECISION_SINGLE 1 /* complex64  */

/* Rotation axes (reference: QV::Rotation, qubitvector.hpp:47-55). */
#define AERHIP_ROT_X 0
#define AERHIP_ROT_Y 1
#define AERHIP_ROT_Z 2
#define AERHIP_ROT_XX 3
#define AERHIP_ROT_YY 4
#define AERHIP_ROT_ZZ 5
#define AERHIP_ROT_ZX 6

typedef struct aerhip_qv aerhip_qv; /* opaque statevector handle */

/* ---- library ------------------------------------------------------------ */

/* Thread-local message of the last failed call on this thread. */
const char *aerhip_last_error(void);

/* Number of visible HIP devices (0 when no GPU is present). */
int aerhip_device_count(void);

/* Version string of the engine library. */
const char *aerhip_version(void);

/* No counterpart in the reference (its GPU executor queries the devices
 * itself): free and total bytes of device memory on `device`. */
int aerhip_device_memory(int device, uint64_t *free_out, uint64_t *total_out);

/* ---- lifecycle ---------------------------------------------------------- */

/* QubitVector(num_qubits) + set_num_qubits (qubitvector.hpp:922-932):
 * allocates 2^num_qubits amplitudes on `device`.  The contents are
 * undefined until one of the initialize calls.  NULL on failure. */
aerhip_qv *aerhip_qv_create(uint64_t num_qubits, int precision, int device);

/* Frees the device memory of the handle (QubitVector::~QubitVector). */
void aerhip_qv_destroy(aerhip_qv *qv);

/* QubitVector::num_qubits (qubitvector.hpp:105). */
int aerhip_qv_num_qubits(const aerhip_qv *qv, uint64_t *out);

/* QubitVector::required_memory_mb (qubitvector.hpp:984-992). */
uint64_t aerhip_required_memory_mb(uint64_t num_qubits, int precision);

/* QubitVector::initialize (qubitvector.hpp:1088-1091): |0...0>. */
int aerhip_qv_initialize(aerhip_qv *qv);

/* QubitVector::zero (qubitvector.hpp:903-910). */
int aerhip_qv_zero(aerhip_qv *qv);

/* QubitVector::initialize_from_vector (qubitvector.hpp:1095-1109).
 * `vec` holds `len` complex128 values; len must equal 2^num_qubits. */
int aerhip_qv_initialize_from_vector(aerhip_qv *qv, const double *vec,
                                     uint64_t len);

/* QubitVector::copy_to_vector (qubitvector.hpp:861-864): writes `len`
 * complex128 values (len must equal 2^num_qubits). */
int aerhip_qv_copy_to_vector(aerhip_qv *qv, double *out, uint64_t len);

/* QubitVector::initialize_component (qubitvector.hpp:879-896):
 * psi[k-th group][i] = psi[k-th group][0] * state[i]. */
int aerhip_qv_initialize_component(aerhip_qv *qv, const uint64_t *qubits,
                                   uint64_t num_qubits, const double *state);

/* Executes every queued gate and waits for the device
 * (QubitVectorThrust::synchronize in the GPU reference). */
int aerhip_qv_synchronize(aerhip_qv *qv);

/* QubitVector::set_sample_measure_index_size (qubitvector.hpp:457). */
int aerhip_qv_set_sample_measure_index_size(aerhip_qv *qv, int n);

/* No counterpart in the reference (its GPU state is filled from the host):
 * loads a state that already lives in device memory of the same device,
 * `len` amplitudes of `src_precision`, without a PCIe round trip. */
int aerhip_qv_initialize_from_device(aerhip_qv *qv, const void *src,
                                     uint64_t len, int src_precision);

/* No counterpart in the reference (a test-input generator): fills the state
 * with psi = g/|g|, g i.i.d. complex Gaussian from a counter-based hash of
 * (seed, index), generated on the device. */
int aerhip_qv_initialize_random(aerhip_qv *qv, uint64_t seed);

/* No counterpart in the reference: flushes the queue and returns the device
 * address of the amplitudes (interop with device-side callers). */
int aerhip_qv_device_pointer(aerhip_qv *qv, void **out);

/* No counterpart in the reference (its blocking is fixed at build time):
 * tile size (log2 amplitudes per workgroup) and the number of low qubits
 * every tile holds; values <= 0 keep the current setting. */
int aerhip_qv_set_tile_bits(aerhip_qv *qv, int tile_bits, int low_bits);

/* No counterpart in the reference: cumulative counters of the handle,
 * out4 = {passes, kernel launches of passes, queued gate primitives,
 * register layout changes inside tile passes}. */
int aerhip_qv_stats(aerhip_qv *qv, uint64_t *out4);

/* The engine's fusion, configured by Aer's options fusion_enable,
 * fusion_max_qubit and fusion_threshold (src/transpile/fusion.hpp:757-763,
 * 873).  Inside each tile pass the engine merges runs of diagonal gates
 * (phase gates on one or two qubits, single-qubit diagonals) into one
 * table-driven diagonal op, as Aer's DiagonalFusion (fusion.hpp:520-726)
 * merges them into one diagonal matrix.  enable = 0 turns it off; a fused op
 * holds at most max_qubit qubits (>= 2); fusion applies only to states of
 * more than `threshold` qubits.  Handles start enabled, max_qubit 64,
 * threshold 0; aerhip_run_circuit passes the run configuration's values. */
int aerhip_qv_set_fusion(aerhip_qv *qv, int enable, int max_qubit, int threshold);

/* No counterpart in the reference (its GPU executor launches every gate of
 * every run): a flush whose pass launches have the signature of the previous
 * flush's (the same circuit again, or the same circuit shape with other
 * angles) runs them as a HIP graph, captured the second time and replayed
 * from then on.  Returns how many flushes of this handle were captured and
 * replayed. */
int aerhip_qv_graph_stats(aerhip_qv *qv, uint64_t *captures, uint64_t *replays);

/* Block fusion, the engine's form of Aer's fusion of gates into dense
 * matrices (src/transpile/fusion.hpp:849-1136, fusion_max_qubit): unitary
 * gates whose qubits stay within `width` qubits (2..5; 0 turns it off, the
 * default) are multiplied into one matrix, found by looking ahead past gates
 * on other qubits; blocks on 3 to 5 qubits run as GEMMs on MFMA on the tile
 * in LDS.  With cost_check = 1 a block is fused only when the engine's cost
 * model predicts it is cheaper than its gates (Aer weighs its CPU costs the
 * same way, fusion.hpp:1105-1136); 0 fuses every block of two or more gates.
 * The autotuner (qiskit_aer_amd.autotune) picks the width by timing. */
int aerhip_qv_set_block_fusion(aerhip_qv *qv, int width, int cost_check);

/* No counterpart in the reference: number of gate primitives merged into
 * fused ops on this handle so far. */
int aerhip_qv_fused_count(aerhip_qv *qv, uint64_t *out);

/* No counterpart in the reference: number of register qubits per thread in
 * full tile passes (c128: 3 or 4, c64: 4 or 5; 0 keeps every gate on the LDS
 * tile). */
int aerhip_qv_set_register_bits(aerhip_qv *qv, int reg_bits);

/* No counterpart in the reference (its timers are host wall clocks,
 * circuit_executor.hpp:715-723): brackets every flush's pass kernels with
 * HIP events on the handle's stream and accumulates their device time. */
int aerhip_qv_set_event_timing(aerhip_qv *qv, int enable);

/* Device milliseconds and number of pass launches accumulated while event
 * timing was on. */
int aerhip_qv_pass_timing(aerhip_qv *qv, double *ms_out, uint64_t *launches_out);

/* ---- gates (queued) ----------------------------------------------------- */

/* QubitVector::apply_matrix (qubitvector.hpp:1299-1303 ->
 * transformer.hpp:89-140).  `mat` holds 4^nq complex values. */
int aerhip_apply_matrix(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq,
                        const double *mat);

/* QubitVector::apply_diagonal_matrix (qubitvector.hpp:1343-1347 ->
 * transformer.hpp:234-372).  `diag` holds 2^nq complex values. */
int aerhip_apply_diagonal_matrix(aerhip_qv *qv, const uint64_t *qubits,
                                 uint64_t nq, const double *diag);

/* QubitVector::apply_mcx (qubitvector.hpp:1447-1486). */
int aerhip_apply_mcx(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq);

/* QubitVector::apply_mcy (qubitvector.hpp:1489-1537). */
int aerhip_apply_mcy(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq);

/* QubitVector::apply_mcswap (qubitvector.hpp:1540-1573). */
int aerhip_apply_mcswap(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq);

/* QubitVector::apply_mcphase (qubitvector.hpp:1576-1612). */
int aerhip_apply_mcphase(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq,
                         double phase_re, double phase_im);

/* QubitVector::apply_mcu (qubitvector.hpp:1615-1720); `mat` is 2x2. */
int aerhip_apply_mcu(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq,
                     const double *mat);

/* QubitVector::apply_rotation (qubitvector.hpp:1723-1750). */
int aerhip_apply_rotation(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq,
                          int axis, double theta);

/* QubitVector::apply_pauli (qubitvector.hpp:2393-2433). */
int aerhip_apply_pauli(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq,
                       const char *pauli, double coeff_re, double coeff_im);

/* QubitVector::apply_permutation_matrix (qubitvector.hpp:1350-1434);
 * `pairs` holds 2*npairs indices into the 2^nq local basis. */
int aerhip_apply_permutation_matrix(aerhip_qv *qv, const uint64_t *qubits,
                                    uint64_t nq, const uint64_t *pairs,
                                    uint64_t npairs);

/* QubitVector::apply_multiplexer (qubitvector.hpp:1306-1340); `mat` is the
 * stacked (2^nt x 2^(nt+nc)) column-major matrix. */
int aerhip_apply_multiplexer(aerhip_qv *qv, const uint64_t *controls,
                             uint64_t nc, const uint64_t *targets, uint64_t nt,
                             const double *mat);

/* QubitVector::apply_multi_swaps (qubitvector.hpp:1843-1871). */
int aerhip_apply_multi_swaps(aerhip_qv *qv, const uint64_t *qubits,
                             uint64_t nq);

/* ---- readout (flush the queue; FP64 reductions) ------------------------- */

/* QubitVector::norm() (qubitvector.hpp:1879-1886). */
/* <a|b> = sum_i conj(a_i) b_i, accumulated in FP64 (the states may differ in
 * precision; same qubit count and device).  No counterpart in the reference's
 * QubitVector; it serves the complex64-vs-complex128 fidelity check
 * 1 - |<psi64|psi128>|^2 on device-resident states. */
int aerhip_qv_inner_product(aerhip_qv *a, aerhip_qv *b, double *re, double *im);

int aerhip_norm(aerhip_qv *qv, double *out);

/* QubitVector::norm(qubits, mat) (qubitvector.hpp:1889-2054): |M psi|^2 for
 * the 2^nq x 2^nq column-major matrix `mat` on `qubits`, without applying it
 * (State::apply_kraus, statevector_state.hpp:1136).  nq <= 10. */
int aerhip_norm_matrix(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq, const double *mat,
                       double *out);

/* QubitVector::norm_diagonal(qubits, diag) (qubitvector.hpp:1946-2082):
 * |D psi|^2 for the diagonal `diag` (2^nq entries) on `qubits`.  nq <= 10. */
int aerhip_norm_diagonal(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq, const double *diag,
                         double *out);

/* QubitVector::get_state(pos) (qubitvector.hpp:838-841): amplitude `pos` as
 * (re, im) in out2[0..1]. */
int aerhip_qv_get_state(aerhip_qv *qv, uint64_t pos, double *out2);

/* QubitVector::set_state(pos, val) (qubitvector.hpp:833-836). */
int aerhip_qv_set_state(aerhip_qv *qv, uint64_t pos, double re, double im);

/* QubitVector::probability(outcome) (qubitvector.hpp:2090-2092). */
int aerhip_probability(aerhip_qv *qv, uint64_t outcome, double *out);

/* QubitVector::probabilities(qubits) (qubitvector.hpp:2107-2143); with
 * nq == 0 it is probabilities() (qubitvector.hpp:2095-2104).  `out` holds
 * 2^nq doubles (2^num_qubits when nq == 0); bit i of the outcome index is
 * qubits[i]. */
int aerhip_probabilities(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq,
                         double *out);

/* QubitVector::sample_measure(rnds) (qubitvector.hpp:2149-2228): for each
 * rnd the smallest index k with rnd < sum_{j<=k} |psi_j|^2. */
int aerhip_sample_measure(aerhip_qv *qv, const double *rnds, uint64_t shots,
                          uint64_t *out);

/* QubitVector::expval_pauli(qubits, pauli, initial_phase)
 * (qubitvector.hpp:2300-2347). */
int aerhip_expval_pauli(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq,
                        const char *pauli, double phase_re, double phase_im,
                        double *out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): aerhip_expval_pauli on a statevector the caller holds in
 * device memory (`sv`: 2^num_qubits amplitudes of `precision`, num_qubits
 * 1..48), with the same kernels and summation order, so the value is
 * bit-identical.  It allocates nothing: the workgroup partials go to a fixed
 * device array of 4096 doubles per GPU, held until the process ends.  The
 * kernels run on `stream` (NULL: the null stream) on the current device and
 * write the value to `out`, a device pointer; synchronize the stream before
 * reading it.  Calls that overlap on one GPU share that array and race. */
int aerhip_expval_pauli_noalloc(const void *sv, uint64_t num_qubits, int precision,
                                const uint64_t *qubits, uint64_t nq, const char *pauli,
                                double phase_re, double phase_im, double *out, void *stream);

/* Base::apply_save_expval for one Pauli string (src/simulators/state.hpp:
 * 416-447, evaluated with QubitVector::expval_pauli): queues
 * Re <psi| phase * P |psi> at the current point of the gate sequence, without
 * a flush.  The engine evaluates it inside the tile pass that holds that
 * point, so many saved values share the state sweeps of the gates (the GPU
 * reference's batched_expval_pauli, qubitvector_thrust.hpp:2683-2713, also
 * batches terms).  `slot_out` receives the value's index. */
int aerhip_save_expval_pauli(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq,
                             const char *pauli, double phase_re, double phase_im,
                             uint64_t *slot_out);

/* Flushes and copies every saved value in slot order into `out`
 * (`capacity` entries), writes their number to `count`, and clears them. */
int aerhip_saved_expvals(aerhip_qv *qv, double *out, uint64_t capacity, uint64_t *count);

/* No counterpart in the reference (it has no adjoint method): with the top
 * qubit n-1 as a flag, writes sum_t coeffs[t] P_t psi into the half with the
 * flag set, where psi is the half with the flag clear.  `labels` holds nterms
 * strings of n-1 characters, character n-2-i acting on qubit i. */
int aerhip_qv_pauli_sum_to_upper_half(aerhip_qv *qv, const char *labels, const double *coeffs,
                                      uint64_t nterms);

/* ---- circuit runner (controller slice) ---------------------------------- */

/* Operation kinds of the circuit wire format (a subset of Aer's OpType,
 * src/framework/operations.hpp). */
#define AERHIP_OP_GATE 0        /* named gate from the statevector gate set */
#define AERHIP_OP_MATRIX 1      /* "unitary": params = column-major matrix */
#define AERHIP_OP_DIAGONAL 2    /* "diagonal": params = diagonal entries */
#define AERHIP_OP_MEASURE 3     /* measure qubits -> memory clbits */
#define AERHIP_OP_BARRIER 4     /* no-op */
#define AERHIP_OP_RESET 5       /* reset qubits to |0> */
#define AERHIP_OP_INITIALIZE 6  /* initialize qubits to params (complex) */
#define AERHIP_OP_SAVE_STATEVECTOR 7 /* copy the state out (once) */

typedef struct aerhip_op {
  int32_t type;             /* AERHIP_OP_* */
  uint32_t num_qubits;      /* length of `qubits` */
  const char *name;         /* gate name (AERHIP_OP_GATE), Aer spelling */
  const uint64_t *qubits;   /* operand qubits, Aer order */
  uint32_t num_params;      /* real params (gates) or complex count */
  uint32_t reserved;
  const double *params;     /* gate angles; matrix/diag/initialize: complex
                               pairs (re, im) */
  const char *string_param; /* Pauli label for the "pauli" gate */
  const uint64_t *memory;   /* measure: clbit index per qubit */
} aerhip_op;

typedef struct aerhip_run_config {
  int32_t precision;        /* AERHIP_PRECISION_* */
  int32_t device;
  uint64_t shots;
  uint64_t seed;            /* seed_simulator of this circuit */
  int32_t fusion_enable;    /* 1: the engine fuses diagonal runs (aerhip_qv_set_fusion) */
  int32_t fusion_max_qubit; /* qubits of one fused op; Aer default 5 (fusion.hpp:761) */
  int32_t fusion_threshold; /* fuse above this many qubits; Aer default 14 (fusion.hpp:762) */
  int32_t sample_index_size;/* Aer default 10 (config.hpp:108) */
  double global_phase;      /* circuit global phase, applied at
                               initialisation (statevector_state.hpp:446-450) */
} aerhip_run_config;

/* State::apply_ops (src/simulators/state.hpp:295-340) restricted to the
 * unitary operations of the wire format (AERHIP_OP_GATE, _MATRIX, _DIAGONAL,
 * _BARRIER): dispatches each to the handle's gate methods the way
 * Statevector::State::apply_gate does (statevector_state.hpp:733-915). */
int aerhip_apply_ops(aerhip_qv *qv, const aerhip_op *ops, uint64_t num_ops);

/* Fills `cfg` with the reference defaults (precision double, device 0,
 * shots 1024, seed 0, fusion on, max 5, threshold 14, index size 10,
 * global phase 0). */
void aerhip_run_config_default(aerhip_run_config *cfg);

/* Runs one circuit the way Executor::run_circuit_with_sampling does
 * (src/simulators/circuit_executor.hpp:736-825): applies the operations
 * before the first measurement, then samples `shots` outcomes with
 * mt19937_64(seed) draws (measure_sampler, circuit_executor.hpp:
 * 1031-1161).  Circuits whose measurements are followed by gates run
 * shot by shot.  `memory_out` (optional, `shots` entries) receives the
 * classical memory of each shot as an integer (bit c = clbit c) in the
 * reference's memory order; `statevector_out` (optional, 2^num_qubits
 * complex128) receives the state at the first SAVE_STATEVECTOR op, or
 * before the first measurement when there is none.  `timing_out`
 * (optional, 4 doubles): gate-phase seconds, sampling seconds, number of
 * engine passes, number of kernel launches. */
int aerhip_run_circuit(const aerhip_op *ops, uint64_t num_ops,
                       uint64_t num_qubits, uint64_t num_memory,
                       const aerhip_run_config *cfg, uint64_t *memory_out,
                       double *statevector_out, double *timing_out);

/* ---- batched circuits ----------------------------------------------------- */

/* Runs `num_circuits` independent circuits, each from |0> on its own
 * num_qubits[i] qubits, and returns energies[i] = sum_t coeffs[t] <P_t> over
 * the circuit's Pauli terms.  Replaces running each circuit as a separate
 * experiment (the reference's controller loop over circuits,
 * src/controllers/aer_controller.hpp:309-437, with EstimatorV2's expectation
 * values, qiskit_aer/primitives/estimator_v2.py:112-159).
 *   ops          all circuits' unitary operations, concatenated; circuit i owns
 *                ops[op_offsets[i] .. op_offsets[i+1]);
 *   labels       all Pauli strings, concatenated: term t of circuit i has
 *                num_qubits[i] characters (character n-1-q acts on qubit q);
 *   coeffs       one real weight per term; circuit i owns terms
 *                [term_offsets[i], term_offsets[i+1]);
 *   mode         0: every circuit in one kernel launch (one workgroup walks all
 *                passes of its circuit); 1: circuit by circuit, one launch per
 *                pass (the reference's execution order), for comparison;
 *   kernel_ms    (optional) device time of the execution.
 * Dense matrices must act on at most 5 qubits. */
int aerhip_batch_pauli_expval(const aerhip_op *ops, const uint64_t *op_offsets,
                              const uint32_t *num_qubits, uint64_t num_circuits,
                              const char *labels, const double *coeffs,
                              const uint64_t *term_offsets, int precision, int device,
                              int mode, double *energies, double *kernel_ms);

/* ---- parameter bindings -------------------------------------------------- */

/* Energies E_b = <psi_b|H|psi_b>, psi_b = U(theta_b)|0>, of one parametrized
 * circuit over num_bindings parameter rows.  The reference binds each row
 * into a copy of the circuit (Op bind_parameter, src/framework/operations.hpp:
 * 1347-1374) and runs the bound copies as one batch on the GPU
 * (BatchConverter, src/transpile/batch_converter.hpp:54-157) whose Pauli
 * expectation values it evaluates together (QubitVectorThrust::
 * batched_expval_pauli, qubitvector_thrust.hpp:2683-2713).  Here the rows of
 * one launch group (as many num_qubits statevectors as fit in 70 % of the
 * free device memory) advance through the circuit together: each tile pass is
 * one kernel launch for all of them, each row with its own bound matrices.
 *   ops, param_index, params, labels, coeffs: as aerhip_adjoint_expval_gradient;
 *   energies     num_bindings values;
 *   kernel_ms    optional: device time of the pass kernels. */
int aerhip_bind_pauli_expval(const aerhip_op *ops, uint64_t num_ops, uint64_t num_qubits,
                             const int64_t *param_index, uint64_t num_params, uint64_t num_bindings,
                             const double *params, const char *labels, const double *coeffs,
                             uint64_t nterms, int precision, int device, double *energies,
                             double *kernel_ms);

/* ---- adjoint differentiation -------------------------------------------- */

/* No counterpart in the reference (which differentiates by re-running
 * circuits, e.g. the parameter-shift rule through EstimatorV2,
 * qiskit_aer/primitives/estimator_v2.py:112-159).  For each of `num_bindings`
 * parameter rows, computes E = <psi|H|psi>, psi = U(theta)|0> on `num_qubits`
 * qubits, and dE/dtheta by the adjoint method: one forward sweep, then one
 * reverse sweep that uncomputes each gate on psi and on lambda = H psi and
 * accumulates 2 Re <lambda| dU_k |psi_{k-1}> (no intermediate state stored).
 *   ops          unitary operations (gates, matrices, diagonals, barriers);
 *   param_index  per op, the parameter its first angle is bound to, or -1;
 *                rx ry rz rxx ryy rzz rzx p u1 cp cu1 may be parametrized;
 *   params       num_bindings x num_params, row major;
 *   labels       nterms Pauli strings of num_qubits characters (character
 *                num_qubits-1-i acts on qubit i), coeffs their real weights;
 *   energy_out   num_bindings values, grad_out num_bindings x num_params.
 * The bindings run together, as many (num_qubits+1)-qubit statevectors on
 * `device` as fit in 70 % of its free memory: every pass of the forward and
 * of the reverse sweep is one kernel launch for all of them. */
int aerhip_adjoint_expval_gradient(const aerhip_op *ops, uint64_t num_ops, uint64_t num_qubits,
                                   const int64_t *param_index, uint64_t num_params,
                                   uint64_t num_bindings, const double *params,
                                   const char *labels, const double *coeffs, uint64_t nterms,
                                   int precision, int device, double *energy_out,
                                   double *grad_out);

#ifdef __cplusplus
}
#endif

#endif /* AERHIP_H */
