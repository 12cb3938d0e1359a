// Host side of the MI355X statevector engine: one Engine per statevector
// handle.  It owns the device state, a HIP stream, the queue of pending gate
// primitives and the planner that turns the queue into tile passes.
#pragma once

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"
#include "reg_pass.hpp"

namespace aerhip {

using cd = std::complex<double>;

// A queued gate primitive in global qubit numbering.
struct Prim {
  TileOpKind kind;
  std::vector<uint32_t> q;     // operand qubits (meaning depends on kind)
  std::vector<uint32_t> ctrl;  // control qubits
  uint64_t ctrl_val = ~0ull;   // required value of ctrl[i] is bit i
  uint64_t xmask = 0, zmask = 0;
  double pr = 0, pi = 0;       // phase
  uint32_t off = 0, len = 0;   // coefficient slice in the host pool
  uint32_t npairs = 0;
  uint64_t tmask = 0;          // qubits the primitive mixes (must be tile qubits)
  uint64_t touch = 0;          // every qubit it reads
  bool solo = false;           // runs as its own dense-group pass
  bool unitary = true;         // false for projectors and other non-unitary maps
};

struct PassPlan {
  bool solo;
  uint64_t smask;              // tile qubits
  std::vector<uint32_t> prims; // indices into the queue
};

// A flush's device program before upload: pool offsets index `pool`.
struct Program {
  std::vector<PassDesc> passes;
  std::vector<TileOp> ops;
  std::vector<ROp<double>> rops;
  std::vector<cd> pool;
};

struct EngineStats {
  uint64_t passes = 0;
  uint64_t launches = 0;
  uint64_t gates = 0;
  double pass_ms = 0;          // device time of the pass kernels (event timing on)
  uint64_t timed_launches = 0; // launches covered by pass_ms
  uint64_t gathers = 0;        // register layout changes in tile passes
  uint64_t fused = 0;          // gate primitives merged into fused diagonals
  uint64_t graph_captures = 0; // flushes whose pass launches were captured into a HIP graph
  uint64_t graph_replays = 0;  // flushes that replayed the captured graph
};

class Engine {
 public:
  Engine(uint32_t num_qubits, int precision, int device);
  // A handle that only queues and plans (no device state, never executes):
  // the batch runner turns its queue into a program with emit_program().
  struct PlanOnly {};
  Engine(uint32_t num_qubits, int precision, int device, PlanOnly);
  ~Engine();
  Engine(const Engine &) = delete;
  Engine &operator=(const Engine &) = delete;

  uint32_t num_qubits() const { return n_; }
  int precision() const { return prec_; }
  uint64_t size() const { return size_; }
  void *device_data() { return sv_; }
  hipStream_t stream() const { return stream_; }
  EngineStats &stats() { return stats_; }

  // lifecycle
  void initialize();
  void zero();
  void initialize_from_vector(const cd *vec, uint64_t len);
  void initialize_from_device(const void *src, uint64_t len, int src_precision);
  void initialize_random(uint64_t seed);
  void copy_to_vector(cd *out, uint64_t len);
  void initialize_component(const std::vector<uint32_t> &qubits, const std::vector<cd> &state);
  void synchronize();
  void set_sample_measure_index_size(int n) { sample_index_size_ = n; }
  void set_tile_bits(uint32_t nb);
  void set_low_bits(uint32_t c);
  void set_event_timing(bool on);
  void set_reg_bits(uint32_t r);
  // Engine-side fusion of diagonal runs (Aer's fusion options): on/off, the
  // most qubits of one fused op, and the qubit count above which it applies.
  void set_fusion(bool enable, uint32_t max_qubit, uint32_t threshold);
  // Block fusion of unitary gates into dense matrices on at most `width`
  // qubits (0: off); with cost_check, a block is fused only when the cost
  // model predicts a gain.
  void set_block_fusion(uint32_t width, bool cost_check);

  // gates (queued)
  void apply_matrix(const std::vector<uint32_t> &qubits, const std::vector<cd> &mat);
  void apply_diagonal_matrix(const std::vector<uint32_t> &qubits, const std::vector<cd> &diag);
  void apply_mcx(const std::vector<uint32_t> &qubits);
  void apply_mcy(const std::vector<uint32_t> &qubits);
  void apply_mcswap(const std::vector<uint32_t> &qubits);
  void apply_mcphase(const std::vector<uint32_t> &qubits, cd phase);
  void apply_mcu(const std::vector<uint32_t> &qubits, const std::vector<cd> &mat);
  void apply_rotation(const std::vector<uint32_t> &qubits, int axis, double theta);
  void apply_pauli(const std::vector<uint32_t> &qubits, const std::string &pauli, cd coeff);
  void apply_permutation_matrix(const std::vector<uint32_t> &qubits,
                                const std::vector<std::pair<uint64_t, uint64_t>> &pairs);
  void apply_multiplexer(const std::vector<uint32_t> &controls,
                         const std::vector<uint32_t> &targets, const std::vector<cd> &mat);
  void apply_multi_swaps(const std::vector<uint32_t> &qubits);

  // readout
  double norm();
  // QubitVector::norm(qubits, mat): |M psi|^2 with M on `qubits`, not applied.
  double norm_matrix(const std::vector<uint32_t> &qubits, const std::vector<cd> &mat);
  // QubitVector::get_state / set_state: one amplitude.
  cd get_state(uint64_t pos);
  void set_state(uint64_t pos, cd value);
  double probability(uint64_t outcome);
  std::vector<double> probabilities(const std::vector<uint32_t> &qubits);
  std::vector<uint64_t> sample_measure(const std::vector<double> &rnds);
  double expval_pauli(const std::vector<uint32_t> &qubits, const std::string &pauli, cd phase);
  // <this|other>, accumulated in FP64; the states may differ in precision.
  cd inner_product(Engine &other);

  // Queues Re(<psi| phase * P |psi>) at the current point of the gate
  // sequence (evaluated inside the tile passes); returns its slot.
  uint64_t save_expval_pauli(const std::vector<uint32_t> &qubits, const std::string &pauli,
                             cd phase);
  // Flushes and returns every saved value in slot order, then clears them.
  std::vector<double> take_saved_expvals();
  // With the top qubit as a flag: writes sum_t c_t P_t psi (psi = the half
  // with the top qubit 0) into the half with the top qubit 1.  Labels act on
  // qubits 0..n-2 (character n-2-i on qubit i).
  void pauli_sum_to_upper(const std::vector<std::string> &labels, const std::vector<double> &coeffs);
  // The terms of such a Pauli sum on n qubits (flag excluded), sorted by X mask.
  static std::vector<PauliTerm> upper_terms(uint32_t n, const std::vector<std::string> &labels,
                                            const std::vector<double> &coeffs);

  // Executes the queue.
  void flush();
  // Appends the queue's passes to `prog` and clears the queue.
  void emit_program(Program &prog);
  bool plan_only() const { return plan_only_; }
  uint32_t tile_bits() const { return tile_bits_; }
  uint32_t reg_bits() const { return reg_bits_; }
  uint64_t probes_queued() const { return probes_queued_; }

 private:
  void check_qubits(const std::vector<uint32_t> &qubits) const;
  void discard_queue();
  void push(Prim &&p);
  uint32_t pool_add(const cd *v, size_t n);
  void push_dense(const std::vector<uint32_t> &targets, const std::vector<uint32_t> &ctrl,
                  uint64_t ctrl_val, const cd *mat);
  void push_diag(const std::vector<uint32_t> &qubits, const std::vector<uint32_t> &ctrl,
                 uint64_t ctrl_val, const cd *diag);
  void apply_matrix_1(uint32_t q, const std::vector<cd> &mat);
  void apply_diagonal_1(uint32_t q, const cd *diag);
  std::vector<PassPlan> plan() const;
  std::vector<Prim> fuse_blocks(const std::vector<Prim> &in);
  std::vector<Prim> merge_swap_layers(const std::vector<Prim> &in) const;
  PassDesc qperm_pass(const Prim &p, std::vector<TileOp> &ops) const;
  void ensure_partial(size_t n);
  void *scratch(size_t bytes);
  double fetch_scalar(const double *dptr);
  template <typename T> double norm_t();
  template <typename T> double norm_matrix_t(const MarginalDesc &md, const std::vector<cd> &mat);
  template <typename T> void probabilities_t(const std::vector<uint32_t> &qubits, double *out);
  template <typename T> void sample_t(const std::vector<double> &rnds, uint64_t *out);
  template <typename T> double expval_t(uint64_t xmask, uint64_t zmask, uint32_t xmax, cd ph);

  uint32_t n_;
  int prec_;
  int device_;
  bool plan_only_ = false;
  uint64_t size_;
  void *sv_ = nullptr;
  hipStream_t stream_ = nullptr;
  int sample_index_size_ = 10;
  uint32_t tile_bits_;  // largest tile (11 for c128, 12 for c64: 32 KiB of LDS)
  uint32_t low_bits_;   // qubits every tile holds (one 128-byte run)
  uint32_t reg_bits_;   // register qubits per thread in full tiles (0: LDS-only passes)
  bool fusion_enable_ = true;
  uint32_t fusion_max_qubit_ = 64;
  uint32_t fusion_threshold_ = 0;
  uint32_t block_width_ = 0;     // block fusion width (0: off)
  bool block_cost_check_ = true;

  std::vector<Prim> queue_;
  std::vector<cd> pool_;

  // device program of a flush: [passes | ops | pool]
  void *d_prog_ = nullptr;
  size_t d_prog_cap_ = 0;
  void *h_prog_ = nullptr;  // pinned staging
  size_t h_prog_cap_ = 0;
  hipEvent_t prog_ev_ = nullptr;

  double *d_partial_ = nullptr;
  size_t partial_cap_ = 0;
  double *h_scalar_ = nullptr;  // pinned
  void *d_scratch_ = nullptr;   // descriptors, draws and samples of readouts
  size_t scratch_cap_ = 0;

  uint64_t probes_queued_ = 0;        // probes in the current queue
  std::vector<double> saved_expvals_; // values of flushed probes, slot order
  double *d_probes_ = nullptr;        // per-tile probe partials
  size_t probes_cap_ = 0;

  EngineStats stats_;
  // HIP graph of the last repeated flush (its launch signature and the
  // instantiated graph) and the signature of the previous flush.
  std::vector<uint64_t> graph_sig_, last_sig_;
  hipGraphExec_t graph_exec_ = nullptr;
  bool event_timing_ = false;
  hipEvent_t t0_ev_ = nullptr, t1_ev_ = nullptr;
};

}  // namespace aerhip
