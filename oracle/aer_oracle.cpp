// ORACLE -- test infrastructure only (see qubitvector_oracle.hpp).
//
// CPU restatement of the layers of Qiskit Aer that sit above QubitVector for
// the statevector method, driven through a C API that mirrors the engine's
// (include/aerhip.h) so tests can run the same operations on both:
//   - State gate dispatch (src/simulators/statevector/statevector_state.hpp:
//     733-915) with the VMatrix gate formulas
//     (src/framework/linalg/matrix_utils/vmatrix_defs.hpp:197-381);
//   - the Fusion pass (src/transpile/fusion.hpp:29-1146), unitary method;
//   - the sampling driver (src/simulators/circuit_executor.hpp:736-825,
//     1031-1161) with std::mt19937_64 draws (src/framework/rng.hpp:32-101).
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

#include "../include/aerhip.h"
#include "qubitvector_oracle.hpp"

namespace oracle {

thread_local std::string g_err;

// ---------------------------------------------------------------------------
// VMatrix (vmatrix_defs.hpp:197-381, matrix_defs.hpp:150-212)
// ---------------------------------------------------------------------------
namespace vm {
using cv = cvector_t<double>;
const cd I(0., 1.);
inline cv u4(double t, double p, double l, double g) {
  cv m(4);
  m[0] = std::exp(I * g) * std::cos(0.5 * t);
  m[2] = -std::exp(I * (l + g)) * std::sin(0.5 * t);
  m[1] = std::exp(I * (p + g)) * std::sin(0.5 * t);
  m[3] = std::exp(I * (p + l + g)) * std::cos(0.5 * t);
  return m;
}
inline cv r(double t, double p) {
  cv m(4);
  m[0] = std::cos(0.5 * t);
  m[2] = -I * std::exp(-I * p) * std::sin(0.5 * t);
  m[1] = -I * std::exp(I * p) * std::sin(0.5 * t);
  m[3] = std::cos(0.5 * t);
  return m;
}
inline cv rx(double t) {
  cv m(4);
  m[0] = std::cos(0.5 * t);
  m[2] = -I * std::sin(0.5 * t);
  m[1] = m[2];
  m[3] = m[0];
  return m;
}
inline cv ry(double t) {
  cv m(4);
  m[0] = std::cos(0.5 * t);
  m[2] = -1.0 * std::sin(0.5 * t);
  m[1] = -m[2];
  m[3] = m[0];
  return m;
}
inline cv rz(double t) {
  cv m(4);
  m[0] = std::exp(-I * 0.5 * t);
  m[3] = std::exp(I * 0.5 * t);
  return m;
}
inline cv rxx(double t) {
  const double c = std::cos(0.5 * t), s = std::sin(0.5 * t);
  cv m(16);
  m[0] = c; m[12] = -I * s; m[5] = c; m[9] = -I * s; m[6] = -I * s; m[10] = c; m[3] = -I * s; m[15] = c;
  return m;
}
inline cv ryy(double t) {
  const double c = std::cos(0.5 * t), s = std::sin(0.5 * t);
  cv m(16);
  m[0] = c; m[12] = I * s; m[5] = c; m[9] = -I * s; m[6] = -I * s; m[10] = c; m[3] = I * s; m[15] = c;
  return m;
}
inline cv rzz_diag(double t) {
  const cd ep = std::exp(I * 0.5 * t), em = std::exp(-I * 0.5 * t);
  return cv({em, ep, ep, em});
}
inline cv rzx(double t) {
  const double c = std::cos(0.5 * t), s = std::sin(0.5 * t);
  cv m(16);
  m[0] = c; m[8] = -I * s; m[5] = c; m[13] = I * s; m[2] = -I * s; m[10] = c; m[7] = I * s; m[15] = c;
  return m;
}
inline cv sx() { return cv({{0.5, 0.5}, {0.5, -0.5}, {0.5, -0.5}, {0.5, 0.5}}); }
inline cv sxdg() { return cv({{0.5, -0.5}, {0.5, 0.5}, {0.5, 0.5}, {0.5, -0.5}}); }
inline cv ecr() {
  const double s = 1. / std::sqrt(2.);
  const cd row[4][4] = {{0, s, 0, cd(0, s)}, {s, 0, cd(0, -s), 0}, {0, cd(0, s), 0, s}, {cd(0, -s), 0, s, 0}};
  cv m(16);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) m[i + 4 * j] = row[i][j];
  return m;
}
}  // namespace vm

// ---------------------------------------------------------------------------
// Operations
// ---------------------------------------------------------------------------
enum OpType { GATE = 0, MATRIX = 1, DIAGONAL = 2, MEASURE = 3, BARRIER = 4, RESET = 5,
              INITIALIZE = 6, SAVE_SV = 7, NOP = 100 };

struct Op {
  int type = NOP;
  std::string name;
  reg_t qubits;
  std::vector<double> params;
  cvector_t<double> mat;  // matrix (column-major), diagonal, or initialize vector
  std::string sparam;
  reg_t memory;
};

enum class Gate { id, x, y, mcz, s, sdg, h, t, tdg, mcp, mcsx, mcsxdg, mcr, mcrx, mcry, mcrz,
                  mcu2, mcu3, mcu, mcswap, rxx, ryy, rzz, rzx, ecr, pauli };

// statevector_state.hpp:314-384
const std::unordered_map<std::string, Gate> &gateset() {
  static const std::unordered_map<std::string, Gate> m = {
      {"delay", Gate::id}, {"id", Gate::id}, {"x", Gate::x}, {"y", Gate::y}, {"z", Gate::mcz},
      {"s", Gate::s}, {"sdg", Gate::sdg}, {"h", Gate::h}, {"t", Gate::t}, {"tdg", Gate::tdg},
      {"p", Gate::mcp}, {"sx", Gate::mcsx}, {"sxdg", Gate::mcsxdg}, {"r", Gate::mcr},
      {"rx", Gate::mcrx}, {"ry", Gate::mcry}, {"rz", Gate::mcrz}, {"u1", Gate::mcp},
      {"u2", Gate::mcu2}, {"u3", Gate::mcu3}, {"u", Gate::mcu3}, {"U", Gate::mcu3},
      {"CX", Gate::x}, {"cx", Gate::x}, {"cy", Gate::y}, {"cz", Gate::mcz}, {"cp", Gate::mcp},
      {"cu1", Gate::mcp}, {"cu2", Gate::mcu2}, {"cu3", Gate::mcu3}, {"cu", Gate::mcu},
      {"swap", Gate::mcswap}, {"rxx", Gate::rxx}, {"ryy", Gate::ryy}, {"rzz", Gate::rzz},
      {"rzx", Gate::rzx}, {"csx", Gate::mcsx}, {"csxdg", Gate::mcsxdg}, {"ecr", Gate::ecr},
      {"crx", Gate::mcrx}, {"cry", Gate::mcry}, {"crz", Gate::mcrz}, {"ccx", Gate::x},
      {"ccz", Gate::mcz}, {"cswap", Gate::mcswap}, {"mcx", Gate::x}, {"mcy", Gate::y},
      {"mcz", Gate::mcz}, {"mcr", Gate::mcr}, {"mcrx", Gate::mcrx}, {"mcry", Gate::mcry},
      {"mcrz", Gate::mcrz}, {"mcphase", Gate::mcp}, {"mcp", Gate::mcp}, {"mcu1", Gate::mcp},
      {"mcu2", Gate::mcu2}, {"mcu3", Gate::mcu3}, {"mcu", Gate::mcu}, {"mcswap", Gate::mcswap},
      {"mcsx", Gate::mcsx}, {"mcsxdg", Gate::mcsxdg}, {"pauli", Gate::pauli},
      {"mcx_gray", Gate::x}};
  return m;
}

// Gates of the unitary simulator, i.e. fusable gates (unitary_state.hpp:45-54).
const std::set<std::string> &unitary_gates() {
  static const std::set<std::string> s = {
      "u1", "u2", "u3", "u", "U", "CX", "cx", "cz", "cy", "cp", "cu1", "cu2", "cu3", "swap", "id",
      "p", "x", "y", "z", "h", "s", "sdg", "t", "tdg", "r", "rx", "ry", "rz", "rxx", "ryy",
      "rzz", "rzx", "ccx", "ccz", "cswap", "mcx", "mcy", "mcz", "mcu1", "mcu2", "mcu3",
      "mcswap", "mcr", "mcrx", "mcry", "mcrz", "sx", "sxdg", "csx", "mcsx", "csxdg", "mcsxdg",
      "delay", "pauli", "cu", "mcu", "mcp", "ecr", "mcphase", "crx", "cry", "crz"};
  return s;
}

// State::apply_gate / apply_matrix / apply_diagonal_matrix.
template <typename T>
void apply_op_gate(QubitVector<T> &qv, const Op &op) {
  const reg_t &q = op.qubits;
  if (op.type == MATRIX) {
    // State::apply_matrix (statevector_state.hpp:879-890)
    const size_t D = BIT(q.size());
    bool diagonal = true;
    for (size_t j = 0; j < D && diagonal; ++j)
      for (size_t i = 0; i < D; ++i)
        if (i != j && std::abs(op.mat[i + D * j]) > 0.0) { diagonal = false; break; }
    if (diagonal) {
      cvector_t<double> d(D);
      for (size_t i = 0; i < D; ++i) d[i] = op.mat[i + D * i];
      qv.apply_diagonal_matrix(q, d);
    } else {
      qv.apply_matrix(q, op.mat);
    }
    return;
  }
  if (op.type == DIAGONAL) return qv.apply_diagonal_matrix(q, op.mat);
  auto it = gateset().find(op.name);
  if (it == gateset().end())
    throw std::invalid_argument("QubitVectorState::invalid gate instruction '" + op.name + "'.");
  auto P = [&](size_t i) {
    if (i >= op.params.size()) throw std::invalid_argument("missing gate parameter");
    return op.params[i];
  };
  const double isqrt2 = 1. / std::sqrt(2);
  switch (it->second) {
    case Gate::x: qv.apply_mcx(q); break;
    case Gate::y: qv.apply_mcy(q); break;
    case Gate::mcz: qv.apply_mcphase(q, -1); break;
    case Gate::mcr: qv.apply_mcu(q, vm::r(P(0), P(1))); break;
    case Gate::mcrx: qv.apply_mcu(q, vm::rx(P(0))); break;
    case Gate::mcry: qv.apply_mcu(q, vm::ry(P(0))); break;
    case Gate::mcrz: qv.apply_mcu(q, vm::rz(P(0))); break;
    case Gate::rxx: qv.apply_matrix(q, vm::rxx(P(0))); break;
    case Gate::ryy: qv.apply_matrix(q, vm::ryy(P(0))); break;
    case Gate::rzz: qv.apply_diagonal_matrix(q, vm::rzz_diag(P(0))); break;
    case Gate::rzx: qv.apply_matrix(q, vm::rzx(P(0))); break;
    case Gate::ecr: qv.apply_matrix(q, vm::ecr()); break;
    case Gate::id: break;
    case Gate::h: qv.apply_mcu(q, vm::u4(M_PI / 2., 0., M_PI, 0.)); break;
    case Gate::s: qv.apply_diagonal_matrix({q[0]}, {1., cd(0., 1.)}); break;
    case Gate::sdg: qv.apply_diagonal_matrix({q[0]}, {1., cd(0., -1.)}); break;
    case Gate::t: qv.apply_diagonal_matrix({q[0]}, {1., cd(isqrt2, isqrt2)}); break;
    case Gate::tdg: qv.apply_diagonal_matrix({q[0]}, {1., cd(isqrt2, -isqrt2)}); break;
    case Gate::mcswap: qv.apply_mcswap(q); break;
    case Gate::mcu3: qv.apply_mcu(q, vm::u4(P(0), P(1), P(2), 0.)); break;
    case Gate::mcu: qv.apply_mcu(q, vm::u4(P(0), P(1), P(2), P(3))); break;
    case Gate::mcu2: qv.apply_mcu(q, vm::u4(M_PI / 2., P(0), P(1), 0.)); break;
    case Gate::mcp: qv.apply_mcphase(q, std::exp(cd(0, 1) * cd(P(0), 0.))); break;
    case Gate::mcsx: qv.apply_mcu(q, vm::sx()); break;
    case Gate::mcsxdg: qv.apply_mcu(q, vm::sxdg()); break;
    case Gate::pauli: qv.apply_pauli(q, op.sparam); break;
  }
}

// ---------------------------------------------------------------------------
// Fusion (fusion.hpp), unitary method.
// ---------------------------------------------------------------------------
struct Fusion {
  uint_t max_qubit = 5;
  uint_t threshold = 14;
  double cost_factor = 1.8;
  bool avx2 = __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");

  // FusionMethod::can_ignore (fusion.hpp:118-127)
  static bool can_ignore(const Op &op) {
    if (op.type == BARRIER) return true;
    if (op.type == GATE) return op.name == "id" || op.name == "u0";
    return false;
  }
  // UnitaryFusion::can_apply (fusion.hpp:177-192)
  bool can_apply(const Op &op, uint_t maxq) const {
    switch (op.type) {
      case MATRIX: return op.qubits.size() <= maxq;
      case DIAGONAL: return op.qubits.size() <= maxq;
      case GATE: return op.qubits.size() <= maxq && unitary_gates().count(op.name) > 0;
      default: return false;
    }
  }

  // FusionMethod::generate_operation + UnitaryFusion::generate_operation_internal
  // (fusion.hpp:36-111, 159-174): ascending remap of the qubits, fused matrix
  // obtained by applying the ops to the vectorized identity.
  Op generate(std::vector<Op> ops, bool diagonal) const {
    std::set<uint_t> qs;
    for (auto &op : ops) qs.insert(op.qubits.begin(), op.qubits.end());
    reg_t remap(qs.begin(), qs.end());
    std::map<uint_t, uint_t> o2r;
    for (size_t i = 0; i < remap.size(); ++i) o2r[remap[i]] = i;
    for (auto &op : ops)
      for (auto &q : op.qubits) q = o2r[q];
    const size_t k = remap.size();
    const size_t D = BIT(k);
    QubitVector<double> u(2 * k);
    u.zero();
    for (size_t j = 0; j < D; ++j) u.data()[j + D * j] = 1.0;
    for (const auto &op : ops) {
      if (op.type == BARRIER) continue;
      apply_op_gate(u, op);
    }
    Op out;
    out.name = "fusion";
    out.qubits = remap;
    if (diagonal) {
      out.type = DIAGONAL;
      out.mat.resize(D);
      for (size_t i = 0; i < D; ++i) out.mat[i] = u.data()[i + D * i];
    } else {
      out.type = MATRIX;
      out.mat.assign(u.data(), u.data() + D * D);
    }
    return out;
  }

  void allocate(std::vector<Op> &ops, uint_t idx, const std::vector<uint_t> &idxs, bool diag) const {
    std::vector<Op> f;
    for (uint_t i : idxs) f.push_back(ops[i]);
    ops[idx] = generate(f, diag);
    for (uint_t i : idxs)
      if (i != idx) ops[i].type = NOP;
  }

  // DiagonalFusion (fusion.hpp:494-726)
  static bool is_diagonal_op(const Op &op) {
    if (op.type == DIAGONAL) return true;
    if (op.type == GATE) {
      const std::string &n = op.name;
      if (n == "p" || n == "cp" || n == "u1" || n == "cu1" || n == "mcu1" || n == "rz" || n == "rzz")
        return true;
      if (n == "u3") return op.params[0] == 0. && op.params[1] == 0.;
    }
    return false;
  }
  int next_diagonal_end(const std::vector<Op> &ops, int from, int end, std::set<uint_t> &fq) const {
    if (is_diagonal_op(ops[from])) {
      fq.insert(ops[from].qubits.begin(), ops[from].qubits.end());
      return from;
    }
    if (ops[from].type != GATE) return -1;
    int pos = from;
    for (; pos < end; ++pos)
      if (ops[from].type != GATE || ops[pos].name != "cx") break;
    if (pos == from || pos == end) return -1;
    int cx_end = pos - 1;
    bool found = false;
    for (; pos < end; ++pos) {
      if (is_diagonal_op(ops[pos])) found = true;
      else break;
    }
    if (!found || pos == end) return -1;
    const int u1_end = pos;
    for (; pos < end; ++pos) {
      if (ops[pos].type == GATE && ops[pos].name == ops[cx_end].name &&
          ops[pos].qubits == ops[cx_end].qubits) {
        if (cx_end == from) break;
        --cx_end;
      } else {
        return -1;
      }
    }
    if (pos == end) return -1;
    for (int i = from; i < u1_end; ++i) fq.insert(ops[i].qubits.begin(), ops[i].qubits.end());
    return pos;
  }
  void diagonal_fusion(std::vector<Op> &ops, int start, int end) const {
    const uint_t min_qubit = 3;
    for (int op_idx = start; op_idx < end; ++op_idx) {
      std::set<uint_t> fq;
      int nde = next_diagonal_end(ops, op_idx, end, fq);
      if (nde < 0) continue;
      if (fq.size() > max_qubit) continue;
      int next_start = nde + 1;
      while (true) {
        int e = next_diagonal_end(ops, next_start, end, fq);
        if (e < 0 || fq.size() > max_qubit) break;
        next_start = e + 1;
      }
      if (fq.size() < min_qubit) continue;
      std::vector<uint_t> idxs;
      while (op_idx < next_start) idxs.push_back(op_idx++);
      --op_idx;
      allocate(ops, op_idx, idxs, true);
    }
  }

  // NQubitFusion<N> (fusion.hpp:387-492)
  static bool exclude_escaped(std::vector<uint_t> &fq, const Op &t) {
    bool included = true;
    for (uint_t q : t.qubits) included &= std::find(fq.begin(), fq.end(), q) != fq.end();
    if (included) return false;
    for (uint_t q : t.qubits) {
      auto f = std::find(fq.begin(), fq.end(), q);
      if (f != fq.end()) fq.erase(f);
    }
    return true;
  }
  void nqubit_fusion(std::vector<Op> &ops, int start, int end, size_t N) const {
    for (int op_idx = start; op_idx < end; ++op_idx) {
      if (!can_apply(ops[op_idx], max_qubit) || ops[op_idx].type == NOP) continue;
      if (ops[op_idx].qubits.size() != N) continue;
      std::vector<uint_t> idxs = {(uint_t)op_idx};
      std::vector<uint_t> fq(ops[op_idx].qubits.begin(), ops[op_idx].qubits.end());
      for (int j = op_idx - 1; j >= start; --j) {
        auto &t = ops[j];
        if (t.type == NOP) continue;
        if (!can_apply(t, max_qubit)) break;
        if (!exclude_escaped(fq, t)) idxs.push_back(j);
        else if (fq.empty()) break;
      }
      std::reverse(idxs.begin(), idxs.end());
      fq.assign(ops[op_idx].qubits.begin(), ops[op_idx].qubits.end());
      for (int j = op_idx + 1; j < end; ++j) {
        auto &t = ops[j];
        if (t.type == NOP) continue;
        if (!can_apply(t, max_qubit)) break;
        if (!exclude_escaped(fq, t)) idxs.push_back(j);
        else if (fq.empty()) break;
      }
      if (idxs.size() <= 1) continue;
      allocate(ops, op_idx, idxs, false);
    }
  }

  // CostBasedFusion (fusion.hpp:1002-1136)
  static void add_fq(reg_t &fq, const Op &op) {
    for (uint_t q : op.qubits)
      if (std::find(fq.begin(), fq.end(), q) == fq.end()) fq.push_back(q);
  }
  static bool cb_is_diagonal(const std::vector<Op> &ops, uint_t from, uint_t until) {
    for (uint_t i = from; i <= until; ++i) {
      if ((i + 2) <= until && ops[i].name == "cx" && ops[i + 1].name == "u1" &&
          ops[i + 2].name == "cx" && ops[i].qubits[1] == ops[i + 1].qubits[0] &&
          ops[i + 1].qubits[0] == ops[i + 2].qubits[1] && ops[i].qubits[0] == ops[i + 2].qubits[0]) {
        i += 2;
        continue;
      }
      if (ops[i].name == "u1" || ops[i].name == "cu1" || ops[i].name == "cp") continue;
      return false;
    }
    return true;
  }
  double estimate_cost(const std::vector<Op> &ops, uint_t from, uint_t until) const {
    if (cb_is_diagonal(ops, from, until)) return 1.0;
    reg_t fq;
    for (uint_t i = from; i <= until; ++i) add_fq(fq, ops[i]);
    const size_t n = fq.size();
    if (avx2) {
      switch (n) {
        case 1:
        case 2: return 1.0;
        case 3: return 1.1;
        case 4: return 3;
        default: return std::pow(cost_factor, (double)std::max(n - 2, size_t(1)));
      }
    }
    return std::pow(cost_factor, (double)std::max(n - 1, size_t(1)));
  }
  void cost_based(std::vector<Op> &ops, int start, int end) const {
    if (end <= start) return;
    std::vector<double> costs;
    std::vector<int> to;
    to.push_back(start);
    costs.push_back(can_ignore(ops[start]) ? .0 : cost_factor);
    bool applied = false;
    for (int i = start + 1; i < end; ++i) {
      to.push_back(i);
      costs.push_back(costs[i - start - 1] + (can_ignore(ops[i]) ? .0 : cost_factor));
      for (int nf = 2; nf <= (int)max_qubit; ++nf) {
        reg_t fq;
        add_fq(fq, ops[i]);
        for (int j = i - 1; j >= start; --j) {
          add_fq(fq, ops[j]);
          if ((int)fq.size() > nf) break;
          const double est = estimate_cost(ops, (uint_t)j, i) + (j <= start ? 0.0 : costs[j - 1 - start]);
          if (est <= costs[i - start]) {
            costs[i - start] = est;
            to[i - start] = j;
            applied = true;
          }
        }
      }
    }
    if (!applied) return;
    for (int i = end - 1; i >= start;) {
      const int t = to[i - start];
      if (t != i) {
        std::vector<uint_t> idxs;
        for (int j = t; j <= i; ++j) idxs.push_back(j);
        if (!idxs.empty()) allocate(ops, i, idxs, false);
      }
      i = t - 1;
    }
  }

  // Fusion::optimize_circuit (fusion.hpp:849-966)
  bool run(std::vector<Op> &ops, uint_t num_qubits) const {
    if (num_qubits <= threshold || ops.size() < 2) return false;
    bool applied = false;
    for (int fuser = 0; fuser < 4; ++fuser) {
      int fusion_start = 0;
      const int end = (int)ops.size();
      for (int op_idx = 0; op_idx < end; ++op_idx) {
        if (can_ignore(ops[op_idx])) continue;
        if (!can_apply(ops[op_idx], max_qubit) || op_idx == end - 1) {
          switch (fuser) {
            case 0: diagonal_fusion(ops, fusion_start, op_idx); break;
            case 1: nqubit_fusion(ops, fusion_start, op_idx, 1); break;
            case 2: nqubit_fusion(ops, fusion_start, op_idx, 2); break;
            default: cost_based(ops, fusion_start, op_idx); break;
          }
          fusion_start = op_idx + 1;
        }
      }
      size_t idx = 0;
      for (size_t i = 0; i < ops.size(); ++i)
        if (ops[i].type != NOP) {
          if (i != idx) ops[idx] = ops[i];
          ++idx;
        }
      if (idx != ops.size()) {
        applied = true;
        ops.erase(ops.begin() + idx, ops.end());
      }
    }
    return applied;
  }
};

// ---------------------------------------------------------------------------
// Circuit runner
// ---------------------------------------------------------------------------
std::vector<Op> parse_ops(const aerhip_op *ops, uint64_t n) {
  std::vector<Op> out(n);
  for (uint64_t i = 0; i < n; ++i) {
    const aerhip_op &a = ops[i];
    Op &o = out[i];
    o.type = a.type;
    if (a.name) o.name = a.name;
    o.qubits.assign(a.qubits, a.qubits + a.num_qubits);
    if (a.type == GATE) {
      o.params.assign(a.params, a.params + a.num_params);
    } else if (a.type == MATRIX || a.type == DIAGONAL || a.type == INITIALIZE) {
      o.mat.resize(a.num_params);
      for (uint32_t j = 0; j < a.num_params; ++j) o.mat[j] = cd(a.params[2 * j], a.params[2 * j + 1]);
    }
    if (a.string_param) o.sparam = a.string_param;
    if (a.type == MEASURE && a.memory) o.memory.assign(a.memory, a.memory + a.num_qubits);
  }
  return out;
}

template <typename T>
struct Runner {
  QubitVector<T> qv;
  std::mt19937_64 rng;
  explicit Runner(size_t n) : qv(n) {}

  uint_t measure_collapse(const reg_t &q, bool reset) {
    const auto probs = qv.probabilities(q);
    const uint_t meas = std::discrete_distribution<size_t>(probs.begin(), probs.end())(rng);
    const double p = probs[meas];
    const uint_t final_state = reset ? 0 : meas;
    const size_t dim = BIT(q.size());
    cvector_t<double> mdiag(dim, 0.);
    mdiag[meas] = 1. / std::sqrt(p);
    qv.apply_diagonal_matrix(q, mdiag);
    if (final_state != meas) {
      if (q.size() == 1) {
        qv.apply_mcx(q);
      } else {
        cvector_t<double> perm(dim * dim, 0.);
        perm[final_state * dim + meas] = 1.;
        perm[meas * dim + final_state] = 1.;
        for (size_t j = 0; j < dim; ++j)
          if (j != final_state && j != meas) perm[j * dim + j] = 1.;
        Op op;
        op.type = MATRIX;
        op.qubits = q;
        op.mat = perm;
        apply_op_gate(qv, op);
      }
    }
    return meas;
  }
};

template <typename T>
void run_circuit_t(std::vector<Op> ops, uint64_t nq, const aerhip_run_config &cfg, int threads,
                   uint64_t *memory_out, double *sv_out, double *timing) {
  using clk = std::chrono::steady_clock;
  size_t first_meas = ops.size();
  for (size_t i = 0; i < ops.size(); ++i)
    if (ops[i].type == MEASURE) { first_meas = i; break; }
  bool can_sample = true;
  for (size_t i = first_meas; i < ops.size(); ++i)
    if (ops[i].type != MEASURE && ops[i].type != BARRIER) can_sample = false;

  const auto t_fuse = clk::now();
  if (cfg.fusion_enable && can_sample) {
    Fusion f;
    f.max_qubit = cfg.fusion_max_qubit;
    f.threshold = cfg.fusion_threshold;
    std::vector<Op> head(ops.begin(), ops.begin() + first_meas);
    std::vector<Op> tail(ops.begin() + first_meas, ops.end());
    f.run(head, nq);
    first_meas = head.size();
    head.insert(head.end(), tail.begin(), tail.end());
    ops.swap(head);
  }
  const double fuse_s = std::chrono::duration<double>(clk::now() - t_fuse).count();

  Runner<T> R(nq);
  R.qv.set_omp_threads(threads);
  R.qv.set_sample_measure_index_size(cfg.sample_index_size);
  const bool has_gphase = std::abs(cfg.global_phase) > 1e-10;
  const cd gphase = std::exp(cd(0.0, cfg.global_phase));
  bool saved = false;
  auto init = [&] {
    R.qv.initialize();
    if (has_gphase) R.qv.apply_diagonal_matrix({0}, {gphase, gphase});
  };
  auto run_range = [&](size_t b, size_t e, uint64_t *creg) {
    for (size_t i = b; i < e; ++i) {
      const Op &op = ops[i];
      switch (op.type) {
        case GATE:
        case MATRIX:
        case DIAGONAL: apply_op_gate(R.qv, op); break;
        case BARRIER:
        case NOP: break;
        case MEASURE: {
          const uint_t m = R.measure_collapse(op.qubits, false);
          if (creg)
            for (size_t j = 0; j < op.qubits.size() && j < op.memory.size(); ++j)
              *creg = (*creg & ~BIT(op.memory[j])) | (((m >> j) & 1ull) << op.memory[j]);
          break;
        }
        case RESET: R.measure_collapse(op.qubits, true); break;
        case INITIALIZE: {
          cvector_t<double> v = op.mat;
          if (has_gphase)
            for (auto &x : v) x *= gphase;
          bool full = op.qubits.size() == nq;
          for (size_t j = 0; full && j < op.qubits.size(); ++j) full = op.qubits[j] == j;
          if (full) {
            R.qv.initialize_from_vector(v.data(), v.size());
          } else {
            R.measure_collapse(op.qubits, true);
            R.qv.initialize_component(op.qubits, v);
          }
          break;
        }
        case SAVE_SV:
          if (sv_out && !saved) {
            for (size_t k = 0; k < R.qv.size(); ++k) {
              sv_out[2 * k] = R.qv.data()[k].real();
              sv_out[2 * k + 1] = R.qv.data()[k].imag();
            }
            saved = true;
          }
          break;
        default: throw std::invalid_argument("unknown operation type");
      }
    }
  };
  double t_gates = 0, t_sample = 0;
  if (can_sample) {
    R.rng.seed(cfg.seed);
    init();
    const auto t0 = clk::now();  // gate phase only: allocation and |0> excluded
    run_range(0, first_meas, nullptr);
    t_gates = std::chrono::duration<double>(clk::now() - t0).count();
    if (sv_out && !saved) {
      for (size_t k = 0; k < R.qv.size(); ++k) {
        sv_out[2 * k] = R.qv.data()[k].real();
        sv_out[2 * k + 1] = R.qv.data()[k].imag();
      }
    }
    reg_t mq;
    for (size_t i = first_meas; i < ops.size(); ++i)
      if (ops[i].type == MEASURE) mq.insert(mq.end(), ops[i].qubits.begin(), ops[i].qubits.end());
    std::sort(mq.begin(), mq.end());
    mq.erase(std::unique(mq.begin(), mq.end()), mq.end());
    if (!mq.empty() && cfg.shots > 0) {
      const auto t1 = clk::now();
      std::vector<double> rnds;
      rnds.reserve(cfg.shots);
      for (uint64_t s = 0; s < cfg.shots; ++s)
        rnds.push_back(std::uniform_real_distribution<double>(0, 1)(R.rng));
      const auto samples = R.qv.sample_measure(rnds);
      t_sample = std::chrono::duration<double>(clk::now() - t1).count();
      std::map<uint_t, uint_t> memory_map;
      for (size_t i = first_meas; i < ops.size(); ++i)
        if (ops[i].type == MEASURE)
          for (size_t j = 0; j < ops[i].qubits.size() && j < ops[i].memory.size(); ++j)
            memory_map[ops[i].memory[j]] =
                std::lower_bound(mq.begin(), mq.end(), ops[i].qubits[j]) - mq.begin();
      if (memory_out)
        for (uint64_t is = 0; is < cfg.shots; ++is) {
          const uint_t s = samples[cfg.shots - is - 1];
          uint64_t mem = 0;
          for (const auto &kv : memory_map) mem |= ((s >> mq[kv.second]) & 1ull) << kv.first;
          memory_out[is] = mem;
        }
    } else if (memory_out) {
      for (uint64_t is = 0; is < cfg.shots; ++is) memory_out[is] = 0;
    }
  } else {
    const auto t0 = clk::now();
    for (uint64_t s = 0; s < std::max<uint64_t>(cfg.shots, 1); ++s) {
      R.rng.seed(cfg.seed + s);
      init();
      uint64_t creg = 0;
      run_range(0, ops.size(), &creg);
      if (memory_out && s < cfg.shots) memory_out[s] = creg;
    }
    t_gates = std::chrono::duration<double>(clk::now() - t0).count();
  }
  if (timing) {
    timing[0] = t_gates;
    timing[1] = t_sample;
    timing[2] = static_cast<double>(ops.size());
    timing[3] = fuse_s;
  }
}

// Handle of the method-level API.
struct Handle {
  int prec;
  std::unique_ptr<QubitVector<double>> d;
  std::unique_ptr<QubitVector<float>> f;
};

template <typename F>
int guard(F &&f) {
  try {
    f();
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return 1;
  }
}

template <typename F>
void dispatch(Handle *h, F &&f) {
  if (!h) throw std::invalid_argument("null handle");
  if (h->prec == AERHIP_PRECISION_DOUBLE) f(*h->d);
  else f(*h->f);
}

reg_t rv(const uint64_t *q, uint64_t n) { return reg_t(q, q + n); }
cvector_t<double> cv(const double *p, uint64_t n) {
  cvector_t<double> v(n);
  for (uint64_t i = 0; i < n; ++i) v[i] = cd(p[2 * i], p[2 * i + 1]);
  return v;
}

}  // namespace oracle

using namespace oracle;

extern "C" {

const char *oracle_last_error(void) { return g_err.c_str(); }

void *oracle_qv_create(uint64_t nq, int precision, int threads) {
  Handle *h = nullptr;
  guard([&] {
    h = new Handle{precision, nullptr, nullptr};
    if (precision == AERHIP_PRECISION_DOUBLE) {
      h->d = std::make_unique<QubitVector<double>>(nq);
      h->d->set_omp_threads(threads);
    } else {
      h->f = std::make_unique<QubitVector<float>>(nq);
      h->f->set_omp_threads(threads);
    }
  });
  return h;
}
void oracle_qv_destroy(void *h) { delete static_cast<Handle *>(h); }
int oracle_qv_initialize(void *h) {
  return guard([&] { dispatch(static_cast<Handle *>(h), [](auto &q) { q.initialize(); }); });
}
int oracle_qv_initialize_from_vector(void *h, const double *v, uint64_t len) {
  return guard([&] {
    dispatch(static_cast<Handle *>(h), [&](auto &q) {
      q.initialize_from_vector(reinterpret_cast<const cd *>(v), len);
    });
  });
}
int oracle_qv_copy_to_vector(void *h, double *out, uint64_t len) {
  return guard([&] {
    dispatch(static_cast<Handle *>(h), [&](auto &q) {
      if (len != q.size()) throw std::invalid_argument("wrong length");
      for (uint64_t k = 0; k < len; ++k) {
        out[2 * k] = q.data()[k].real();
        out[2 * k + 1] = q.data()[k].imag();
      }
    });
  });
}
int oracle_qv_set_sample_measure_index_size(void *h, int n) {
  return guard([&] { dispatch(static_cast<Handle *>(h), [&](auto &q) { q.set_sample_measure_index_size(n); }); });
}
int oracle_qv_initialize_component(void *h, const uint64_t *q, uint64_t n, const double *s) {
  return guard([&] { dispatch(static_cast<Handle *>(h), [&](auto &v) { v.initialize_component(rv(q, n), cv(s, BIT(n))); }); });
}
int oracle_apply_matrix(void *h, const uint64_t *q, uint64_t n, const double *m) {
  return guard([&] { dispatch(static_cast<Handle *>(h), [&](auto &v) { v.apply_matrix(rv(q, n), cv(m, BIT(2 * n))); }); });
}
int oracle_apply_diagonal_matrix(void *h, const uint64_t *q, uint64_t n, const double *d) {
  return guard([&] { dispatch(static_cast<Handle *>(h), [&](auto &v) { v.apply_diagonal_matrix(rv(q, n), cv(d, BIT(n))); }); });
}
int oracle_apply_mcx(void *h, const uint64_t *q, uint64_t n) {
  return guard([&] { dispatch(static_cast<Handle *>(h), [&](auto &v) { v.apply_mcx(rv(q, n)); }); });
}
int oracle_apply_mcy(void *h, const uint64_t *q, uint64_t n) {
  return guard([&] { dispatch(static_cast<Handle *>(h), [&](auto &v) { v.apply_mcy(rv(q, n)); }); });
}
int oracle_apply_mcswap(void *h, const uint64_t *q, uint64_t n) {
  return guard([&] { dispatch(static_cast<Handle *>(h), [&](auto &v) { v.apply_mcswap(rv(q, n)); }); });
}
int oracle_apply_mcphase(void *h, const uint64_t *q, uint64_t n, double re, double im) {
  return guard([&] { dispatch(static_cast<Handle *>(h), [&](auto &v) { v.apply_mcphase(rv(q, n), cd(re, im)); }); });
}
int oracle_apply_mcu(void *h, const uint64_t *q, uint64_t n, const double *m) {
  return guard([&] { dispatch(static_cast<Handle *>(h), [&](auto &v) { v.apply_mcu(rv(q, n), cv(m, 4)); }); });
}
int oracle_apply_rotation(void *h, const uint64_t *q, uint64_t n, int axis, double t) {
  return guard([&] {
    dispatch(static_cast<Handle *>(h), [&](auto &v) {
      const reg_t qs = rv(q, n);
      switch (axis) {
        case AERHIP_ROT_X: v.apply_mcu(qs, vm::rx(t)); break;
        case AERHIP_ROT_Y: v.apply_mcu(qs, vm::ry(t)); break;
        case AERHIP_ROT_Z: v.apply_mcu(qs, vm::rz(t)); break;
        case AERHIP_ROT_XX: v.apply_matrix(qs, vm::rxx(t)); break;
        case AERHIP_ROT_YY: v.apply_matrix(qs, vm::ryy(t)); break;
        case AERHIP_ROT_ZZ: v.apply_diagonal_matrix(qs, vm::rzz_diag(t)); break;
        case AERHIP_ROT_ZX: v.apply_matrix(qs, vm::rzx(t)); break;
        default: throw std::invalid_argument("QubitVector::invalid rotation axis.");
      }
    });
  });
}
int oracle_apply_pauli(void *h, const uint64_t *q, uint64_t n, const char *p, double re, double im) {
  return guard([&] { dispatch(static_cast<Handle *>(h), [&](auto &v) { v.apply_pauli(rv(q, n), p, cd(re, im)); }); });
}
int oracle_apply_permutation_matrix(void *h, const uint64_t *q, uint64_t n, const uint64_t *pairs, uint64_t np) {
  return guard([&] {
    std::vector<std::pair<uint_t, uint_t>> pv(np);
    for (uint64_t i = 0; i < np; ++i) pv[i] = {pairs[2 * i], pairs[2 * i + 1]};
    dispatch(static_cast<Handle *>(h), [&](auto &v) { v.apply_permutation_matrix(rv(q, n), pv); });
  });
}
int oracle_apply_multiplexer(void *h, const uint64_t *c, uint64_t nc, const uint64_t *t, uint64_t nt, const double *m) {
  return guard([&] {
    dispatch(static_cast<Handle *>(h), [&](auto &v) {
      v.apply_multiplexer(rv(c, nc), rv(t, nt), cv(m, BIT(nc + nt) << nt));
    });
  });
}
int oracle_apply_multi_swaps(void *h, const uint64_t *q, uint64_t n) {
  return guard([&] { dispatch(static_cast<Handle *>(h), [&](auto &v) { v.apply_multi_swaps(rv(q, n)); }); });
}
int oracle_norm(void *h, double *out) {
  return guard([&] { dispatch(static_cast<Handle *>(h), [&](auto &v) { *out = v.norm(); }); });
}
int oracle_probabilities(void *h, const uint64_t *q, uint64_t n, double *out) {
  return guard([&] {
    dispatch(static_cast<Handle *>(h), [&](auto &v) {
      const auto p = v.probabilities(rv(q, n));
      std::memcpy(out, p.data(), p.size() * sizeof(double));
    });
  });
}
int oracle_sample_measure(void *h, const double *rnds, uint64_t shots, uint64_t *out) {
  return guard([&] {
    dispatch(static_cast<Handle *>(h), [&](auto &v) {
      const auto s = v.sample_measure(std::vector<double>(rnds, rnds + shots));
      std::memcpy(out, s.data(), s.size() * sizeof(uint64_t));
    });
  });
}
int oracle_expval_pauli(void *h, const uint64_t *q, uint64_t n, const char *p, double re, double im, double *out) {
  return guard([&] { dispatch(static_cast<Handle *>(h), [&](auto &v) { *out = v.expval_pauli(rv(q, n), p, cd(re, im)); }); });
}

// Runs a circuit the way the reference CPU statevector path does (fusion with
// the reference's defaults unless disabled in cfg).  threads <= 0: all cores.
int oracle_run_circuit(const aerhip_op *ops, uint64_t num_ops, uint64_t nq, uint64_t num_memory,
                       const aerhip_run_config *cfg_in, int threads, uint64_t *memory_out,
                       double *sv_out, double *timing_out) {
  return guard([&] {
    (void)num_memory;
    aerhip_run_config cfg;
    if (cfg_in) {
      cfg = *cfg_in;
    } else {
      std::memset(&cfg, 0, sizeof(cfg));
      cfg.shots = 1024;
      cfg.fusion_enable = 1;
      cfg.fusion_max_qubit = 5;
      cfg.fusion_threshold = 14;
      cfg.sample_index_size = 10;
    }
    const int th = threads > 0 ? threads : omp_get_max_threads();
    auto parsed = parse_ops(ops, num_ops);
    if (cfg.precision == AERHIP_PRECISION_DOUBLE)
      run_circuit_t<double>(parsed, nq, cfg, th, memory_out, sv_out, timing_out);
    else
      run_circuit_t<float>(parsed, nq, cfg, th, memory_out, sv_out, timing_out);
  });
}

// Aer's fused op list for a circuit (fusion.hpp), for inspection by tests:
// writes the number of ops after fusion and, per op, its qubit count.
int oracle_fuse_count(const aerhip_op *ops, uint64_t num_ops, uint64_t nq, int max_qubit,
                      int threshold, uint64_t *num_out, uint64_t *widths_out, uint64_t cap) {
  return guard([&] {
    auto parsed = parse_ops(ops, num_ops);
    Fusion f;
    f.max_qubit = max_qubit;
    f.threshold = threshold;
    f.run(parsed, nq);
    *num_out = parsed.size();
    for (uint64_t i = 0; i < parsed.size() && i < cap; ++i) widths_out[i] = parsed[i].qubits.size();
  });
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

// Installs (or, with nulls, removes) the reference's compiled transformer for
// complex128 apply_matrix / apply_diagonal_matrix (the CPU baseline).
void oracle_set_transformer(void *matrix_fn, void *diagonal_fn) {
  g_matrix_hook = reinterpret_cast<TransformerHook>(matrix_fn);
  g_diagonal_hook = reinterpret_cast<TransformerHook>(diagonal_fn);
}

}  // extern "C"
