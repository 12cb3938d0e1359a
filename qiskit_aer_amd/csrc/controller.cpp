// Controller slice: runs one circuit on the HIP engine through the C ABI.
//
// It restates what sits above the QubitVector seam in the reference for the
// statevector method:
//   - gate name -> engine call (Statevector::State::apply_gate,
//     src/simulators/statevector/statevector_state.hpp:733-865);
//   - matrix / diagonal ops (statevector_state.hpp:879-915);
//   - measure / reset / initialize (statevector_state.hpp:936-1085);
//   - the sampling driver (Executor::run_circuit_with_sampling and
//     measure_sampler, src/simulators/circuit_executor.hpp:736-825,
//     1031-1161) with mt19937_64 draws (src/framework/rng.hpp:32-101).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <random>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "aerhip.h"
#include "gate_matrices.hpp"
#include "host_parallel.hpp"

// Sets the thread-local message returned by aerhip_last_error (engine.hip).
extern "C" __attribute__((visibility("hidden"))) void aerhip_detail_set_error(const char *msg);
extern "C" __attribute__((visibility("hidden"))) aerhip_qv *aerhip_detail_qv_create_planner(
    uint64_t num_qubits, int precision, int device);
extern "C" __attribute__((visibility("hidden"))) int aerhip_detail_batch_run(
    aerhip_qv *const *qvs, uint64_t count, int device, double *values, double *kernel_ms);
// Many states of one size, one planning handle each (engine.hip, MultiState).
struct aerhip_multi;
extern "C" __attribute__((visibility("hidden"))) aerhip_multi *aerhip_detail_multi_create(
    uint64_t num_qubits, int precision, int device, uint64_t count);
extern "C" __attribute__((visibility("hidden"))) void aerhip_detail_multi_destroy(aerhip_multi *m);
extern "C" __attribute__((visibility("hidden"))) aerhip_qv *aerhip_detail_multi_planner(aerhip_multi *m,
                                                                                        uint64_t i);
extern "C" __attribute__((visibility("hidden"))) int aerhip_detail_multi_initialize(aerhip_multi *m);
extern "C" __attribute__((visibility("hidden"))) int aerhip_detail_multi_flush(aerhip_multi *m, double *values,
                                                                               uint64_t capacity,
                                                                               uint64_t *per_state);
extern "C" __attribute__((visibility("hidden"))) int aerhip_detail_multi_pauli_sum_to_upper(
    aerhip_multi *m, const char *labels, const double *coeffs, uint64_t nterms);
extern "C" __attribute__((visibility("hidden"))) double aerhip_detail_multi_kernel_ms(aerhip_multi *m);

namespace {

using cd = std::complex<double>;
namespace G = aerhip::gates;


enum class Gate {
  id, x, y, mcz, s, sdg, h, t, tdg, mcp, mcsx, mcsxdg, mcr, mcrx, mcry, mcrz,
  mcu2, mcu3, mcu, mcswap, rxx, ryy, rzz, rzx, ecr, pauli
};

// The statevector gate set (statevector_state.hpp:314-384); "x"-family and
// "y"-family names map to mcx / mcy.
const std::unordered_map<std::string, Gate> &gate_set() {
  static const std::unordered_map<std::string, Gate> m = {
      {"delay", Gate::id}, {"id", Gate::id}, {"x", Gate::x}, {"y", Gate::y},
      {"z", Gate::mcz}, {"s", Gate::s}, {"sdg", Gate::sdg}, {"h", Gate::h},
      {"t", Gate::t}, {"tdg", Gate::tdg}, {"p", Gate::mcp}, {"sx", Gate::mcsx},
      {"sxdg", Gate::mcsxdg}, {"r", Gate::mcr}, {"rx", Gate::mcrx}, {"ry", Gate::mcry},
      {"rz", Gate::mcrz}, {"u1", Gate::mcp}, {"u2", Gate::mcu2}, {"u3", Gate::mcu3},
      {"u", Gate::mcu3}, {"U", Gate::mcu3}, {"CX", Gate::x}, {"cx", Gate::x},
      {"cy", Gate::y}, {"cz", Gate::mcz}, {"cp", Gate::mcp}, {"cu1", Gate::mcp},
      {"cu2", Gate::mcu2}, {"cu3", Gate::mcu3}, {"cu", Gate::mcu}, {"swap", Gate::mcswap},
      {"rxx", Gate::rxx}, {"ryy", Gate::ryy}, {"rzz", Gate::rzz}, {"rzx", Gate::rzx},
      {"csx", Gate::mcsx}, {"csxdg", Gate::mcsxdg}, {"ecr", Gate::ecr}, {"crx", Gate::mcrx},
      {"cry", Gate::mcry}, {"crz", Gate::mcrz}, {"ccx", Gate::x}, {"ccz", Gate::mcz},
      {"cswap", Gate::mcswap}, {"mcx", Gate::x}, {"mcy", Gate::y}, {"mcz", Gate::mcz},
      {"mcr", Gate::mcr}, {"mcrx", Gate::mcrx}, {"mcry", Gate::mcry}, {"mcrz", Gate::mcrz},
      {"mcphase", Gate::mcp}, {"mcp", Gate::mcp}, {"mcu1", Gate::mcp}, {"mcu2", Gate::mcu2},
      {"mcu3", Gate::mcu3}, {"mcu", Gate::mcu}, {"mcswap", Gate::mcswap}, {"mcsx", Gate::mcsx},
      {"mcsxdg", Gate::mcsxdg}, {"pauli", Gate::pauli}, {"mcx_gray", Gate::x}};
  return m;
}

void ck(int rc) {
  if (rc != 0) throw std::runtime_error(aerhip_last_error());
}

struct Runner {
  aerhip_qv *qv = nullptr;
  uint64_t n = 0;
  bool owns = true;
  std::mt19937_64 rng;

  ~Runner() {
    if (qv && owns) aerhip_qv_destroy(qv);
  }

  double rand01() { return std::uniform_real_distribution<double>(0, 1)(rng); }

  void mcu(const std::vector<uint64_t> &q, const G::cvec &m) {
    ck(aerhip_apply_mcu(qv, q.data(), q.size(), reinterpret_cast<const double *>(m.data())));
  }
  void diag(const std::vector<uint64_t> &q, const G::cvec &d) {
    ck(aerhip_apply_diagonal_matrix(qv, q.data(), q.size(),
                                    reinterpret_cast<const double *>(d.data())));
  }
  void matrix(const std::vector<uint64_t> &q, const G::cvec &m) {
    ck(aerhip_apply_matrix(qv, q.data(), q.size(), reinterpret_cast<const double *>(m.data())));
  }

  void apply_gate(const aerhip_op &op, const std::vector<uint64_t> &q) {
    const std::string name = op.name ? op.name : "";
    auto it = gate_set().find(name);
    if (it == gate_set().end())
      throw std::invalid_argument("QubitVectorState::invalid gate instruction '" + name + "'.");
    auto P = [&](uint32_t i) -> double {
      if (i >= op.num_params)
        throw std::invalid_argument("gate '" + name + "' is missing parameter " +
                                    std::to_string(i));
      return op.params[i];
    };
    const double isqrt2 = 1. / std::sqrt(2);
    switch (it->second) {
      case Gate::x: ck(aerhip_apply_mcx(qv, q.data(), q.size())); break;
      case Gate::y: ck(aerhip_apply_mcy(qv, q.data(), q.size())); break;
      case Gate::mcz: ck(aerhip_apply_mcphase(qv, q.data(), q.size(), -1.0, 0.0)); break;
      case Gate::mcr: mcu(q, G::r(P(0), P(1))); break;
      case Gate::mcrx: ck(aerhip_apply_rotation(qv, q.data(), q.size(), AERHIP_ROT_X, P(0))); break;
      case Gate::mcry: ck(aerhip_apply_rotation(qv, q.data(), q.size(), AERHIP_ROT_Y, P(0))); break;
      case Gate::mcrz: ck(aerhip_apply_rotation(qv, q.data(), q.size(), AERHIP_ROT_Z, P(0))); break;
      case Gate::rxx: ck(aerhip_apply_rotation(qv, q.data(), q.size(), AERHIP_ROT_XX, P(0))); break;
      case Gate::ryy: ck(aerhip_apply_rotation(qv, q.data(), q.size(), AERHIP_ROT_YY, P(0))); break;
      case Gate::rzz: ck(aerhip_apply_rotation(qv, q.data(), q.size(), AERHIP_ROT_ZZ, P(0))); break;
      case Gate::rzx: ck(aerhip_apply_rotation(qv, q.data(), q.size(), AERHIP_ROT_ZX, P(0))); break;
      case Gate::ecr: matrix(q, G::ecr()); break;
      case Gate::id: break;
      case Gate::h: mcu(q, G::u4(M_PI / 2., 0., M_PI, 0.)); break;
      case Gate::s: diag({q[0]}, {1., cd(0., 1.)}); break;
      case Gate::sdg: diag({q[0]}, {1., cd(0., -1.)}); break;
      case Gate::t: diag({q[0]}, {1., cd(isqrt2, isqrt2)}); break;
      case Gate::tdg: diag({q[0]}, {1., cd(isqrt2, -isqrt2)}); break;
      case Gate::mcswap: ck(aerhip_apply_mcswap(qv, q.data(), q.size())); break;
      case Gate::mcu3: mcu(q, G::u4(P(0), P(1), P(2), 0.)); break;
      case Gate::mcu: mcu(q, G::u4(P(0), P(1), P(2), P(3))); break;
      case Gate::mcu2: mcu(q, G::u4(M_PI / 2., P(0), P(1), 0.)); break;
      case Gate::mcp: {
        const cd ph = std::exp(cd(0, 1) * cd(P(0), 0.));
        ck(aerhip_apply_mcphase(qv, q.data(), q.size(), ph.real(), ph.imag()));
        break;
      }
      case Gate::mcsx: mcu(q, G::sx()); break;
      case Gate::mcsxdg: mcu(q, G::sxdg()); break;
      case Gate::pauli: {
        if (!op.string_param) throw std::invalid_argument("pauli gate needs a label");
        ck(aerhip_apply_pauli(qv, q.data(), q.size(), op.string_param, 1.0, 0.0));
        break;
      }
    }
  }

  // State::apply_matrix (statevector_state.hpp:879-899): exactly diagonal
  // matrices take the diagonal path.
  void apply_matrix_op(const aerhip_op &op, const std::vector<uint64_t> &q) {
    const uint64_t dim = 1ull << q.size();
    if (op.num_params != dim * dim) throw std::invalid_argument("matrix op has the wrong size");
    const cd *m = reinterpret_cast<const cd *>(op.params);
    bool diagonal = true;
    for (uint64_t j = 0; j < dim && diagonal; ++j)
      for (uint64_t i = 0; i < dim; ++i)
        if (i != j && std::abs(m[i + dim * j]) > 0.0) {
          diagonal = false;
          break;
        }
    if (diagonal) {
      G::cvec d(dim);
      for (uint64_t i = 0; i < dim; ++i) d[i] = m[i + dim * i];
      diag(q, d);
    } else {
      ck(aerhip_apply_matrix(qv, q.data(), q.size(), op.params));
    }
  }

  // sample_measure_with_prob + measure_reset_update
  // (statevector_state.hpp:950-1014).
  uint64_t measure_collapse(const std::vector<uint64_t> &q, bool reset) {
    const uint64_t dim = 1ull << q.size();
    std::vector<double> probs(dim);
    ck(aerhip_probabilities(qv, q.data(), q.size(), probs.data()));
    const uint64_t meas = std::discrete_distribution<size_t>(probs.begin(), probs.end())(rng);
    const double p = probs[meas];
    const uint64_t final_state = reset ? 0 : meas;
    // The projector diag(..., 1/sqrt(p) at meas, ...).  Up to 10 qubits it is
    // one diagonal, as in the reference; beyond, it is the product of
    // diagonals on groups of 10 qubits (the first carries 1/sqrt(p), the
    // others 0/1 entries), which gives every amplitude the same value.
    const double scale = 1. / std::sqrt(p);
    for (size_t g0 = 0; g0 < q.size(); g0 += 10) {
      const size_t gk = std::min<size_t>(10, q.size() - g0);
      const std::vector<uint64_t> gq(q.begin() + g0, q.begin() + g0 + gk);
      G::cvec mdiag(1ull << gk, 0.);
      mdiag[(meas >> g0) & ((1ull << gk) - 1)] = g0 == 0 ? cd(scale) : cd(1.);
      diag(gq, mdiag);
    }
    if (final_state != meas) {
      // The reference swaps the basis states meas and final_state of the
      // measured qubits (an X for one qubit, a permutation matrix for more,
      // statevector_state.hpp:985-1011).  After the projection every other
      // basis state of those qubits holds exact zeros, so an X on each qubit
      // where meas and final_state differ moves the same values.
      for (size_t i = 0; i < q.size(); ++i)
        if (((meas ^ final_state) >> i) & 1ull) ck(aerhip_apply_mcx(qv, &q[i], 1));
    }
    return meas;
  }
};

// An operation with storage of its own (for re-bound angles and inverses).
struct OwnedOp {
  int32_t type = AERHIP_OP_GATE;
  std::string name;
  std::vector<uint64_t> qubits;
  std::vector<double> params;
  std::string label;
  aerhip_op view() const {
    aerhip_op o;
    std::memset(&o, 0, sizeof(o));
    o.type = type;
    o.name = name.c_str();
    o.num_qubits = static_cast<uint32_t>(qubits.size());
    o.qubits = qubits.data();
    o.num_params = static_cast<uint32_t>(type == AERHIP_OP_GATE ? params.size() : params.size() / 2);
    o.params = params.data();
    o.string_param = label.empty() ? nullptr : label.c_str();
    return o;
  }
};

OwnedOp own(const aerhip_op &op) {
  OwnedOp o;
  o.type = op.type;
  if (op.name) o.name = op.name;
  o.qubits.assign(op.qubits, op.qubits + op.num_qubits);
  const uint64_t np = op.type == AERHIP_OP_GATE ? op.num_params : 2ull * op.num_params;
  if (np) o.params.assign(op.params, op.params + np);
  if (op.string_param) o.label = op.string_param;
  return o;
}

// The inverse of a unitary operation: rotations with negated angles, the
// adjoint pairs s/sdg, t/tdg, sx/sxdg, self-inverse gates unchanged, matrices
// conjugate-transposed and diagonals conjugated.
OwnedOp inverse(const OwnedOp &op) {
  OwnedOp r = op;
  if (op.type == AERHIP_OP_MATRIX) {
    const uint64_t D = 1ull << op.qubits.size();
    for (uint64_t i = 0; i < D; ++i)
      for (uint64_t j = 0; j < D; ++j) {
        r.params[2 * (i + D * j)] = op.params[2 * (j + D * i)];
        r.params[2 * (i + D * j) + 1] = -op.params[2 * (j + D * i) + 1];
      }
    return r;
  }
  if (op.type == AERHIP_OP_DIAGONAL) {
    for (size_t i = 1; i < r.params.size(); i += 2) r.params[i] = -r.params[i];
    return r;
  }
  if (op.type != AERHIP_OP_GATE) return r;
  static const std::map<std::string, std::string> adjoint_name = {
      {"s", "sdg"}, {"sdg", "s"}, {"t", "tdg"}, {"tdg", "t"}, {"sx", "sxdg"}, {"sxdg", "sx"},
      {"csx", "csxdg"}, {"csxdg", "csx"}, {"mcsx", "mcsxdg"}, {"mcsxdg", "mcsx"}};
  static const std::set<std::string> self_inverse = {
      "id", "delay", "x", "y", "z", "h", "cx", "CX", "cy", "cz", "swap", "ccx", "ccz", "cswap",
      "mcx", "mcy", "mcz", "mcswap", "mcx_gray", "ecr", "pauli"};
  static const std::set<std::string> negate_first = {
      "rx", "ry", "rz", "crx", "cry", "crz", "mcrx", "mcry", "mcrz", "rxx", "ryy", "rzz", "rzx",
      "p", "u1", "cp", "cu1", "mcp", "mcphase", "mcu1", "r", "mcr"};
  const std::string &n = op.name;
  if (adjoint_name.count(n)) {
    r.name = adjoint_name.at(n);
  } else if (self_inverse.count(n)) {
  } else if (negate_first.count(n)) {
    r.params[0] = -op.params[0];
  } else if (n == "u3" || n == "u" || n == "U" || n == "cu3" || n == "mcu3") {
    r.params = {-op.params[0], -op.params[2], -op.params[1]};
  } else if (n == "u2" || n == "cu2" || n == "mcu2") {
    r.name = n == "u2" ? "u3" : (n == "cu2" ? "cu3" : "mcu3");
    r.params = {-M_PI / 2., -op.params[1], -op.params[0]};
  } else if (n == "cu" || n == "mcu") {
    r.params = {-op.params[0], -op.params[2], -op.params[1], -op.params[3]};
  } else {
    throw std::invalid_argument("adjoint: no inverse for gate '" + n + "'");
  }
  return r;
}

// d/dtheta of a parametrized gate as sum_g w_g (-i/2) G_g U (U = exp(-i theta
// G/2) for rotations; p = e^{i theta/2} rz; cp through its projector), so
// that dE/dtheta = sum_g w_g Im <lambda| G_g |psi> after the gate.
struct GenTerm {
  std::vector<uint64_t> qubits;
  std::string label;
  double weight;
};

std::vector<GenTerm> generator(const OwnedOp &op) {
  const std::string &n = op.name;
  const auto &q = op.qubits;
  if (op.type != AERHIP_OP_GATE) throw std::invalid_argument("adjoint: parameter on a non-gate op");
  if (n == "rx") return {{{q[0]}, "X", 1.0}};
  if (n == "ry") return {{{q[0]}, "Y", 1.0}};
  if (n == "rz" || n == "p" || n == "u1") return {{{q[0]}, "Z", 1.0}};
  if (n == "rxx") return {{{q[0], q[1]}, "XX", 1.0}};
  if (n == "ryy") return {{{q[0], q[1]}, "YY", 1.0}};
  if (n == "rzz") return {{{q[0], q[1]}, "ZZ", 1.0}};
  if (n == "rzx") return {{{q[0], q[1]}, "XZ", 1.0}};
  if (n == "cp" || n == "cu1")
    return {{{q[0]}, "Z", 0.5}, {{q[1]}, "Z", 0.5}, {{q[0], q[1]}, "ZZ", -0.5}};
  throw std::invalid_argument("adjoint: gate '" + n + "' cannot carry a differentiated parameter");
}

bool is_measure_tail(const aerhip_op *ops, uint64_t num_ops, uint64_t first) {
  for (uint64_t i = first; i < num_ops; ++i)
    if (ops[i].type != AERHIP_OP_MEASURE && ops[i].type != AERHIP_OP_BARRIER) return false;
  return true;
}

std::vector<uint64_t> op_qubits(const aerhip_op &op, uint64_t n) {
  std::vector<uint64_t> q(op.qubits, op.qubits + op.num_qubits);
  for (uint64_t v : q)
    if (v >= n) throw std::invalid_argument("operation qubit out of range");
  return q;
}

}  // namespace

namespace {

// A MultiState (many states of one size advancing together) owned here.
struct Multi {
  aerhip_multi *m = nullptr;
  uint64_t count = 0;
  ~Multi() { reset(); }
  void reset() {
    if (m) aerhip_detail_multi_destroy(m);
    m = nullptr;
    count = 0;
  }
  void make(uint64_t n, int precision, int device, uint64_t cnt) {
    if (m && count == cnt) return;
    reset();
    m = aerhip_detail_multi_create(n, precision, device, cnt);
    if (!m) throw std::runtime_error(aerhip_last_error());
    count = cnt;
  }
  aerhip_qv *planner(uint64_t i) const { return aerhip_detail_multi_planner(m, i); }
};

// Bindings that run together: as many n-qubit states as fit in 70 % of the
// free device memory (at least one), at most `wanted`.
uint64_t bindings_per_launch(uint64_t n, int precision, int device, uint64_t wanted) {
  uint64_t free_b = 0, total_b = 0;
  ck(aerhip_device_memory(device, &free_b, &total_b));
  const uint64_t bytes = (1ull << n) * (precision == AERHIP_PRECISION_DOUBLE ? 16 : 8);
  const uint64_t fit = std::max<uint64_t>(1, free_b / 10 * 7 / bytes);
  return std::min(fit, wanted);
}

// The circuit bound to row `b` of `params` (Circuit::set_params-style
// binding of the first angle of each parametrized op, the way the
// reference's bind_parameter rewrites op params, operations.hpp:1347-1374).
std::vector<OwnedOp> bind_row(const std::vector<OwnedOp> &base, const int64_t *param_index,
                              uint64_t num_params, const double *params, uint64_t b) {
  std::vector<OwnedOp> bound = base;
  for (size_t i = 0; i < base.size(); ++i)
    if (param_index[i] >= 0) bound[i].params[0] = params[b * num_params + param_index[i]];
  return bound;
}

void queue_ops(aerhip_qv *qv, const std::vector<OwnedOp> &ops) {
  std::vector<aerhip_op> views(ops.size());
  for (size_t i = 0; i < ops.size(); ++i) views[i] = ops[i].view();
  ck(aerhip_apply_ops(qv, views.data(), views.size()));
}

std::vector<OwnedOp> checked_ops(const aerhip_op *ops, uint64_t num_ops, uint64_t n, const int64_t *param_index,
                                 uint64_t num_params) {
  std::vector<OwnedOp> base(num_ops);
  for (uint64_t i = 0; i < num_ops; ++i) {
    base[i] = own(ops[i]);
    for (uint64_t q : base[i].qubits)
      if (q >= n) throw std::invalid_argument("operation qubit out of range");
    if (ops[i].type != AERHIP_OP_GATE && ops[i].type != AERHIP_OP_MATRIX &&
        ops[i].type != AERHIP_OP_DIAGONAL && ops[i].type != AERHIP_OP_BARRIER)
      throw std::invalid_argument("parameter bindings: only unitary operations are supported");
    if (param_index[i] >= static_cast<int64_t>(num_params))
      throw std::invalid_argument("parameter index out of range");
    if (param_index[i] >= 0 && base[i].params.empty())
      throw std::invalid_argument("parametrized op without an angle");
  }
  return base;
}

}  // namespace

extern "C" {

void aerhip_run_config_default(aerhip_run_config *cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->precision = AERHIP_PRECISION_DOUBLE;
  cfg->device = 0;
  cfg->shots = 1024;
  cfg->seed = 0;
  cfg->fusion_enable = 1;
  cfg->fusion_max_qubit = 5;
  cfg->fusion_threshold = 14;
  cfg->sample_index_size = 10;
  cfg->global_phase = 0.0;
}

int aerhip_run_circuit(const aerhip_op *ops, uint64_t num_ops, uint64_t num_qubits,
                       uint64_t num_memory, const aerhip_run_config *cfg_in,
                       uint64_t *memory_out, double *statevector_out, double *timing_out) {
  try {
    if (num_ops && !ops) throw std::invalid_argument("null op list");
    if (num_memory > 64) throw std::invalid_argument("at most 64 classical bits are supported");
    aerhip_run_config cfg;
    if (cfg_in) cfg = *cfg_in; else aerhip_run_config_default(&cfg);
    Runner R;
    R.n = num_qubits;
    R.qv = aerhip_qv_create(num_qubits, cfg.precision, cfg.device);
    if (!R.qv) throw std::runtime_error(aerhip_last_error());
    ck(aerhip_qv_set_sample_measure_index_size(R.qv, cfg.sample_index_size));
    ck(aerhip_qv_set_fusion(R.qv, cfg.fusion_enable, std::max(2, cfg.fusion_max_qubit),
                            std::max(0, cfg.fusion_threshold)));

    uint64_t first_meas = num_ops;
    for (uint64_t i = 0; i < num_ops; ++i)
      if (ops[i].type == AERHIP_OP_MEASURE) {
        first_meas = i;
        break;
      }
    const bool can_sample = is_measure_tail(ops, num_ops, first_meas);

    // Global phase (State::apply_global_phase, statevector_state.hpp:446-450);
    // angles within 1e-10 of zero are treated as none (state.hpp:398-414).
    const bool has_gphase = std::abs(cfg.global_phase) > 1e-10;
    const cd gphase = std::exp(cd(0.0, cfg.global_phase));

    bool saved = false;
    double t_gates = 0, t_sample = 0;
    using clk = std::chrono::steady_clock;

    auto run_ops = [&](uint64_t begin, uint64_t end, std::vector<uint64_t> *creg) {
      for (uint64_t i = begin; i < end; ++i) {
        const aerhip_op &op = ops[i];
        const std::vector<uint64_t> q = op_qubits(op, num_qubits);
        switch (op.type) {
          case AERHIP_OP_GATE: R.apply_gate(op, q); break;
          case AERHIP_OP_MATRIX: R.apply_matrix_op(op, q); break;
          case AERHIP_OP_DIAGONAL: {
            if (op.num_params != (1ull << q.size()))
              throw std::invalid_argument("diagonal op has the wrong size");
            ck(aerhip_apply_diagonal_matrix(R.qv, q.data(), q.size(), op.params));
            break;
          }
          case AERHIP_OP_BARRIER: break;
          case AERHIP_OP_MEASURE: {
            const uint64_t m = R.measure_collapse(q, false);
            if (creg && op.memory)
              for (size_t j = 0; j < q.size(); ++j) {
                const uint64_t b = (m >> j) & 1ull;
                (*creg)[0] = ((*creg)[0] & ~(1ull << op.memory[j])) | (b << op.memory[j]);
              }
            break;
          }
          case AERHIP_OP_RESET: R.measure_collapse(q, true); break;
          case AERHIP_OP_INITIALIZE: {
            const uint64_t dim = 1ull << q.size();
            if (op.num_params != dim) throw std::invalid_argument("initialize has the wrong size");
            std::vector<cd> v(reinterpret_cast<const cd *>(op.params),
                              reinterpret_cast<const cd *>(op.params) + dim);
            if (has_gphase)
              for (auto &x : v) x *= gphase;
            bool full = q.size() == num_qubits;
            for (size_t j = 0; full && j < q.size(); ++j) full = q[j] == j;
            if (full) {
              ck(aerhip_qv_initialize_from_vector(R.qv, reinterpret_cast<const double *>(v.data()),
                                                  dim));
            } else {
              R.measure_collapse(q, true);
              ck(aerhip_qv_initialize_component(R.qv, q.data(), q.size(),
                                                reinterpret_cast<const double *>(v.data())));
            }
            break;
          }
          case AERHIP_OP_SAVE_STATEVECTOR:
            if (statevector_out && !saved) {
              ck(aerhip_qv_copy_to_vector(R.qv, statevector_out, 1ull << num_qubits));
              saved = true;
            }
            break;
          default: throw std::invalid_argument("unknown operation type");
        }
      }
    };
    auto init_state = [&] {
      ck(aerhip_qv_initialize(R.qv));
      if (has_gphase) {
        const uint64_t q0 = 0;
        const cd d[2] = {gphase, gphase};
        ck(aerhip_apply_diagonal_matrix(R.qv, &q0, 1, reinterpret_cast<const double *>(d)));
      }
    };

    if (can_sample) {
      R.rng.seed(cfg.seed);
      const auto t0 = clk::now();
      init_state();
      run_ops(0, first_meas, nullptr);
      ck(aerhip_qv_synchronize(R.qv));
      t_gates = std::chrono::duration<double>(clk::now() - t0).count();
      if (statevector_out && !saved) {
        ck(aerhip_qv_copy_to_vector(R.qv, statevector_out, 1ull << num_qubits));
        saved = true;
      }
      // measure_sampler: measured qubits sorted and unique.
      std::vector<uint64_t> mq;
      for (uint64_t i = first_meas; i < num_ops; ++i)
        if (ops[i].type == AERHIP_OP_MEASURE)
          for (uint32_t j = 0; j < ops[i].num_qubits; ++j) mq.push_back(ops[i].qubits[j]);
      std::sort(mq.begin(), mq.end());
      mq.erase(std::unique(mq.begin(), mq.end()), mq.end());
      if (!mq.empty() && cfg.shots > 0) {
        const auto t1 = clk::now();
        std::vector<double> rnds(cfg.shots);
        for (auto &r : rnds) r = R.rand01();
        std::vector<uint64_t> samples(cfg.shots);
        ck(aerhip_sample_measure(R.qv, rnds.data(), rnds.size(), samples.data()));
        t_sample = std::chrono::duration<double>(clk::now() - t1).count();
        // memory bit -> position of its qubit among the measured qubits
        std::map<uint64_t, uint64_t> memory_map;
        for (uint64_t i = first_meas; i < num_ops; ++i)
          if (ops[i].type == AERHIP_OP_MEASURE && ops[i].memory)
            for (uint32_t j = 0; j < ops[i].num_qubits; ++j) {
              const uint64_t pos = std::lower_bound(mq.begin(), mq.end(), ops[i].qubits[j]) - mq.begin();
              memory_map[ops[i].memory[j]] = pos;
            }
        if (memory_out) {
          for (uint64_t is = 0; is < cfg.shots; ++is) {
            const uint64_t s = samples[cfg.shots - is - 1];  // reference memory order
            uint64_t mem = 0;
            for (const auto &kv : memory_map) {
              const uint64_t qubit = mq[kv.second];
              mem |= ((s >> qubit) & 1ull) << kv.first;
            }
            memory_out[is] = mem;
          }
        }
      } else if (memory_out) {
        for (uint64_t is = 0; is < cfg.shots; ++is) memory_out[is] = 0;
      }
    } else {
      // Shot by shot (State::apply_measure with per-shot seeds,
      // circuit_executor.hpp:906-908).
      const auto t0 = clk::now();
      for (uint64_t s = 0; s < std::max<uint64_t>(cfg.shots, 1); ++s) {
        R.rng.seed(cfg.seed + s);
        init_state();
        std::vector<uint64_t> creg(1, 0);
        run_ops(0, num_ops, &creg);
        if (memory_out && s < cfg.shots) memory_out[s] = creg[0];
      }
      ck(aerhip_qv_synchronize(R.qv));
      t_gates = std::chrono::duration<double>(clk::now() - t0).count();
    }
    if (timing_out) {
      uint64_t st[4] = {0, 0, 0, 0};
      ck(aerhip_qv_stats(R.qv, st));
      timing_out[0] = t_gates;
      timing_out[1] = t_sample;
      timing_out[2] = static_cast<double>(st[0]);
      timing_out[3] = static_cast<double>(st[1]);
    }
    return 0;
  } catch (const std::invalid_argument &e) {
    aerhip_detail_set_error(e.what());
    return 2;
  } catch (const std::exception &e) {
    aerhip_detail_set_error(e.what());
    return 1;
  }
}

// State::apply_ops (src/simulators/state.hpp:295-340) for the unitary
// operations of the wire format: applies gates, matrices and diagonals to an
// existing statevector handle (they are queued on the engine).
int aerhip_apply_ops(aerhip_qv *qv, const aerhip_op *ops, uint64_t num_ops) {
  try {
    if (!qv) throw std::invalid_argument("aerhip: null statevector handle");
    if (num_ops && !ops) throw std::invalid_argument("null op list");
    Runner R;
    R.qv = qv;
    R.owns = false;
    ck(aerhip_qv_num_qubits(qv, &R.n));
    for (uint64_t i = 0; i < num_ops; ++i) {
      const aerhip_op &op = ops[i];
      const std::vector<uint64_t> q = op_qubits(op, R.n);
      switch (op.type) {
        case AERHIP_OP_GATE: R.apply_gate(op, q); break;
        case AERHIP_OP_MATRIX: R.apply_matrix_op(op, q); break;
        case AERHIP_OP_DIAGONAL:
          if (op.num_params != (1ull << q.size()))
            throw std::invalid_argument("diagonal op has the wrong size");
          ck(aerhip_apply_diagonal_matrix(qv, q.data(), q.size(), op.params));
          break;
        case AERHIP_OP_BARRIER: break;
        default:
          throw std::invalid_argument("aerhip_apply_ops: only gate, matrix, diagonal and barrier "
                                      "operations can be applied to a handle");
      }
    }
    return 0;
  } catch (const std::invalid_argument &e) {
    aerhip_detail_set_error(e.what());
    return 2;
  } catch (const std::exception &e) {
    aerhip_detail_set_error(e.what());
    return 1;
  }
}

int aerhip_batch_pauli_expval(const aerhip_op *ops, const uint64_t *op_offsets,
                              const uint32_t *num_qubits, uint64_t num_circuits,
                              const char *labels, const double *coeffs,
                              const uint64_t *term_offsets, int precision, int device,
                              int mode, double *energies, double *kernel_ms) {
  std::vector<aerhip_qv *> handles;
  auto release = [&] {
    for (aerhip_qv *h : handles) aerhip_qv_destroy(h);
    handles.clear();
  };
  try {
    if (!num_circuits) return 0;
    if (!op_offsets || !num_qubits || !term_offsets || !energies)
      throw std::invalid_argument("null argument");
    if (op_offsets[num_circuits] && !ops) throw std::invalid_argument("null op list");
    if (term_offsets[num_circuits] && (!labels || !coeffs)) throw std::invalid_argument("null Pauli sum");
    if (mode != 0 && mode != 1) throw std::invalid_argument("mode must be 0 or 1");
    // Character offset of each circuit's first label.
    std::vector<uint64_t> label_at(num_circuits + 1, 0);
    for (uint64_t i = 0; i < num_circuits; ++i) {
      if (op_offsets[i + 1] < op_offsets[i] || term_offsets[i + 1] < term_offsets[i])
        throw std::invalid_argument("offsets must not decrease");
      label_at[i + 1] = label_at[i] + (term_offsets[i + 1] - term_offsets[i]) * num_qubits[i];
    }
    auto queue_circuit = [&](aerhip_qv *qv, uint64_t i) {
      ck(aerhip_apply_ops(qv, ops + op_offsets[i], op_offsets[i + 1] - op_offsets[i]));
      const uint64_t n = num_qubits[i];
      std::vector<uint64_t> all(n);
      for (uint64_t q = 0; q < n; ++q) all[q] = q;
      for (uint64_t t = term_offsets[i]; t < term_offsets[i + 1]; ++t) {
        const std::string lab(labels + label_at[i] + (t - term_offsets[i]) * n, n);
        uint64_t slot = 0;
        ck(aerhip_save_expval_pauli(qv, all.data(), n, lab.c_str(), 1.0, 0.0, &slot));
      }
    };
    auto energy = [&](uint64_t i, const double *vals) {
      double e = 0;
      for (uint64_t t = term_offsets[i]; t < term_offsets[i + 1]; ++t)
        e += coeffs[t] * vals[t - term_offsets[i]];
      return e;
    };
    if (mode == 0) {
      handles.reserve(num_circuits);
      for (uint64_t i = 0; i < num_circuits; ++i) {
        aerhip_qv *h = aerhip_detail_qv_create_planner(num_qubits[i], precision, device);
        if (!h) throw std::runtime_error(aerhip_last_error());
        handles.push_back(h);
      }
      // Planning handles are independent: queue the circuits on host threads.
      aerhip::parallel_for(num_circuits, [&](uint64_t i) { queue_circuit(handles[i], i); });
      std::vector<double> vals(term_offsets[num_circuits] - term_offsets[0] + 1);
      ck(aerhip_detail_batch_run(handles.data(), num_circuits, device, vals.data(), kernel_ms));
      for (uint64_t i = 0; i < num_circuits; ++i)
        energies[i] = energy(i, vals.data() + (term_offsets[i] - term_offsets[0]));
    } else {
      // One statevector per distinct size, reused circuit after circuit.
      std::map<uint64_t, aerhip_qv *> by_n;
      std::vector<double> vals;
      for (uint64_t i = 0; i < num_circuits; ++i) {
        aerhip_qv *&h = by_n[num_qubits[i]];
        if (!h) {
          h = aerhip_qv_create(num_qubits[i], precision, device);
          if (!h) throw std::runtime_error(aerhip_last_error());
          handles.push_back(h);
        }
        ck(aerhip_qv_initialize(h));
        queue_circuit(h, i);
        vals.assign(term_offsets[i + 1] - term_offsets[i] + 1, 0.0);
        uint64_t count = 0;
        ck(aerhip_saved_expvals(h, vals.data(), vals.size(), &count));
        energies[i] = energy(i, vals.data());
      }
      if (kernel_ms) *kernel_ms = 0;
    }
    release();
    return 0;
  } catch (const std::invalid_argument &e) {
    release();
    aerhip_detail_set_error(e.what());
    return 2;
  } catch (const std::exception &e) {
    release();
    aerhip_detail_set_error(e.what());
    return 1;
  }
}


// Energies <psi(theta_b)| H |psi(theta_b)> of one parametrized circuit over
// many parameter rows, the rows of one launch group advancing through each
// pass together (one kernel launch per pass for all of them).  The reference
// binds the parameters of each row into a copy of the circuit
// (operations.hpp:1347-1374) and on the GPU runs the bound copies as a batch
// (BatchConverter, src/transpile/batch_converter.hpp:54-157) whose Pauli
// expectation values it evaluates together (batched_expval_pauli,
// qubitvector_thrust.hpp:2683-2713).
int aerhip_bind_pauli_expval(const aerhip_op *ops, uint64_t num_ops, uint64_t num_qubits,
                             const int64_t *param_index, uint64_t num_params, uint64_t num_bindings,
                             const double *params, const char *labels, const double *coeffs,
                             uint64_t nterms, int precision, int device, double *energies,
                             double *kernel_ms) {
  try {
    if (num_ops && (!ops || !param_index)) throw std::invalid_argument("null op list");
    if (num_bindings && num_params && !params) throw std::invalid_argument("null params");
    if (nterms && (!labels || !coeffs)) throw std::invalid_argument("null Hamiltonian");
    if (num_bindings && !energies) throw std::invalid_argument("null output");
    const uint64_t n = num_qubits;
    const std::vector<OwnedOp> base = checked_ops(ops, num_ops, n, param_index, num_params);
    std::vector<uint64_t> all(n);
    for (uint64_t q = 0; q < n; ++q) all[q] = q;
    Multi M;
    double ms = 0;
    const uint64_t per = num_bindings ? bindings_per_launch(n, precision, device, num_bindings) : 1;
    for (uint64_t b0 = 0; b0 < num_bindings; b0 += per) {
      const uint64_t cnt = std::min(per, num_bindings - b0);
      M.make(n, precision, device, cnt);
      ck(aerhip_detail_multi_initialize(M.m));
      aerhip::parallel_for(cnt, [&](uint64_t i) {
        aerhip_qv *qv = M.planner(i);
        queue_ops(qv, bind_row(base, param_index, num_params, params, b0 + i));
        for (uint64_t t = 0; t < nterms; ++t) {
          const std::string lab(labels + t * n, n);
          uint64_t slot = 0;
          ck(aerhip_save_expval_pauli(qv, all.data(), n, lab.c_str(), 1.0, 0.0, &slot));
        }
      });
      std::vector<double> vals(cnt * nterms + 1);
      uint64_t pps = 0;
      ck(aerhip_detail_multi_flush(M.m, vals.data(), vals.size(), &pps));
      for (uint64_t i = 0; i < cnt; ++i) {
        double e = 0;
        for (uint64_t t = 0; t < nterms; ++t) e += coeffs[t] * vals[i * pps + t];
        energies[b0 + i] = e;
      }
      ms += aerhip_detail_multi_kernel_ms(M.m);
      if (cnt != per || b0 + per >= num_bindings) M.reset();
    }
    if (kernel_ms) *kernel_ms = ms;
    return 0;
  } catch (const std::invalid_argument &e) {
    aerhip_detail_set_error(e.what());
    return 2;
  } catch (const std::exception &e) {
    aerhip_detail_set_error(e.what());
    return 1;
  }
}

// Energy and gradient of E(theta) = <psi(theta)| H |psi(theta)> by the
// adjoint method (no counterpart in the reference, which differentiates by
// re-running circuits, e.g. parameter shift through EstimatorV2).
//
// The state is held as Phi = |0>psi + |1>lambda on n+1 qubits (qubit n is a
// flag): gates on qubits < n act on psi and lambda at once, and
// <Phi| Y_n (x) G |Phi> = -2 Im <lambda|G|psi>, <Phi| X_n |Phi> = 2 Re <psi|lambda>.
// One binding:
//   1. Phi = |0> psi with psi = U(theta)|0>;  2. lambda = H psi (upper half);
//   3. E = <Phi|X_n|Phi>/2;  4. for k = L..1: dE/dtheta_k from
//      <Phi|Y_n (x) G_k|Phi> (queued, evaluated inside the tile passes), then
//      apply U_k^dagger to both halves (uncomputation, nothing is stored).
// The bindings of one launch group run these steps together: each step's
// passes launch once for all of their states.
int aerhip_adjoint_expval_gradient(const aerhip_op *ops, uint64_t num_ops, uint64_t num_qubits,
                                   const int64_t *param_index, uint64_t num_params,
                                   uint64_t num_bindings, const double *params,
                                   const char *labels, const double *coeffs, uint64_t nterms,
                                   int precision, int device, double *energy_out,
                                   double *grad_out) {
  try {
    if (num_ops && (!ops || !param_index)) throw std::invalid_argument("null op list");
    if (num_bindings && num_params && !params) throw std::invalid_argument("null params");
    if (nterms && (!labels || !coeffs)) throw std::invalid_argument("null Hamiltonian");
    const uint64_t n = num_qubits;
    const std::vector<OwnedOp> base = checked_ops(ops, num_ops, n, param_index, num_params);
    const uint64_t flag = n;
    Multi M;
    const uint64_t per = num_bindings ? bindings_per_launch(n + 1, precision, device, num_bindings) : 1;
    for (uint64_t b0 = 0; b0 < num_bindings; b0 += per) {
      const uint64_t cnt = std::min(per, num_bindings - b0);
      M.make(n + 1, precision, device, cnt);
      ck(aerhip_detail_multi_initialize(M.m));
      std::vector<std::vector<OwnedOp>> bound(cnt);
      aerhip::parallel_for(cnt, [&](uint64_t i) {
        bound[i] = bind_row(base, param_index, num_params, params, b0 + i);
        queue_ops(M.planner(i), bound[i]);
      });
      uint64_t pps = 0;
      ck(aerhip_detail_multi_flush(M.m, nullptr, 0, &pps));
      ck(aerhip_detail_multi_pauli_sum_to_upper(M.m, labels, coeffs, nterms));
      // (slot, parameter, weight) of every gradient probe, per binding
      std::vector<std::vector<std::tuple<uint64_t, int64_t, double>>> probes(cnt);
      aerhip::parallel_for(cnt, [&](uint64_t i) {
        aerhip_qv *qv = M.planner(i);
        uint64_t slot = 0;
        ck(aerhip_save_expval_pauli(qv, &flag, 1, "X", 1.0, 0.0, &slot));
        std::vector<OwnedOp> rev;
        for (uint64_t k = num_ops; k-- > 0;) {
          const OwnedOp &o = bound[i][k];
          if (o.type == AERHIP_OP_BARRIER) continue;
          if (param_index[k] >= 0) {
            queue_ops(qv, rev);
            rev.clear();
            for (const GenTerm &g : generator(o)) {
              std::vector<uint64_t> q = g.qubits;
              q.push_back(flag);
              const std::string lab = "Y" + g.label;  // character 0 acts on the last qubit
              ck(aerhip_save_expval_pauli(qv, q.data(), q.size(), lab.c_str(), 1.0, 0.0, &slot));
              probes[i].emplace_back(slot, param_index[k], g.weight);
            }
          }
          rev.push_back(inverse(o));
        }
        queue_ops(qv, rev);
      });
      std::vector<double> vals(cnt * (probes[0].size() + 1) + 1);
      ck(aerhip_detail_multi_flush(M.m, vals.data(), vals.size(), &pps));
      for (uint64_t i = 0; i < cnt; ++i) {
        const double *v = vals.data() + i * pps;
        const uint64_t b = b0 + i;
        if (energy_out) energy_out[b] = 0.5 * v[0];
        if (grad_out) {
          double *g = grad_out + b * num_params;
          for (uint64_t p = 0; p < num_params; ++p) g[p] = 0.0;
          for (const auto &t : probes[i]) g[std::get<1>(t)] += std::get<2>(t) * (-0.5) * v[std::get<0>(t)];
        }
      }
      if (cnt != per || b0 + per >= num_bindings) M.reset();
    }
    return 0;
  } catch (const std::invalid_argument &e) {
    aerhip_detail_set_error(e.what());
    return 2;
  } catch (const std::exception &e) {
    aerhip_detail_set_error(e.what());
    return 1;
  }
}

}  // extern "C"
