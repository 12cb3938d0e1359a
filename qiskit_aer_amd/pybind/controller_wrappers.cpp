// controller_wrappers-compatible Python module for the MI355X statevector
// engine.
//
// It mirrors the surface the reference's Python front end imports from its
// pybind module (qiskit_aer/backends/wrappers/bindings.cc:21-30):
//   - aer_controller_execute().execute(circuits, noise_model, config)
//     (aer_controller_binding.hpp:84-108), returning the result as a dict;
//   - AerConfig with the options of the statevector method
//     (aer_controller_binding.hpp:110-380);
//   - AerCircuit with the builder methods the compiler calls
//     (aer_circuit_binding.hpp:150-190, src/framework/circuit.hpp:136-283);
//   - aer_initialize_libraries (a no-op here).
// Circuits run on the engine through its C ABI (include/aerhip.h): without
// save_expval instructions through aerhip_run_circuit (the sampling executor),
// with them on a statevector handle whose saved expectation values are
// evaluated inside the tile passes.  Noise models, conditional operations and
// the other simulation methods are outside this engine and are reported as
// failed experiments (the reference reports a failed experiment with
// status=error and keeps going, circuit_executor.hpp:726-731).
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <ctime>
#include <map>
#include <optional>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "aerhip.h"

namespace py = pybind11;
using cd = std::complex<double>;

namespace {

void ck(int rc) {
  if (rc != 0) throw std::runtime_error(aerhip_last_error());
}

// One instruction of an AerCircuit.
struct CircOp {
  int32_t type = AERHIP_OP_GATE;  // wire-format kind; -1: save_expval
  std::string name;
  std::vector<uint64_t> qubits;
  std::vector<double> params;  // gate angles, or complex pairs (matrix, diagonal, initialize)
  std::string string_param;    // Pauli label of the "pauli" gate
  std::vector<uint64_t> memory;
  std::string key;             // result key of a save instruction
  bool variance = false;       // save_expval_var
  std::vector<std::string> paulis;
  std::vector<double> coeffs, sq_coeffs;
  std::string unsupported;     // non-empty: why the engine cannot run this op

  aerhip_op view() const {
    aerhip_op o;
    std::memset(&o, 0, sizeof(o));
    o.type = type;
    o.num_qubits = static_cast<uint32_t>(qubits.size());
    o.name = name.c_str();
    o.qubits = qubits.data();
    o.num_params = static_cast<uint32_t>(type == AERHIP_OP_GATE ? params.size() : params.size() / 2);
    o.params = params.data();
    o.string_param = string_param.empty() ? nullptr : string_param.c_str();
    o.memory = memory.empty() ? nullptr : memory.data();
    return o;
  }
};

// AER::Circuit's builder surface (src/framework/circuit.hpp:136-283).
struct AerCircuit {
  uint64_t circ_id = 0;
  uint64_t shots = 1;
  uint64_t num_qubits = 0;
  uint64_t num_memory = 0;
  uint64_t seed = 0;
  double global_phase_angle = 0.0;
  py::object header = py::none();
  std::vector<CircOp> ops;

  static std::vector<double> complex_pairs(const std::vector<cd> &v) {
    std::vector<double> out(2 * v.size());
    for (size_t i = 0; i < v.size(); ++i) {
      out[2 * i] = v[i].real();
      out[2 * i + 1] = v[i].imag();
    }
    return out;
  }

  static void no_condition(int64_t cond_regidx, const py::object &expr, const std::string &what, CircOp &op) {
    if (cond_regidx >= 0 || !expr.is_none())
      op.unsupported = "conditional " + what + " is not supported by the MI355X statevector engine";
  }

  void gate(const std::string &name, const std::vector<uint64_t> &qubits, const std::vector<cd> &params,
            const std::vector<std::string> &string_params, int64_t cond_regidx, const py::object &expr,
            const std::string &label) {
    (void)label;
    CircOp op;
    op.type = AERHIP_OP_GATE;
    op.name = name;
    op.qubits = qubits;
    for (const cd &p : params) op.params.push_back(p.real());  // gate angles are real
    if (!string_params.empty()) op.string_param = string_params[0];
    no_condition(cond_regidx, expr, "gate '" + name + "'", op);
    ops.push_back(std::move(op));
  }

  void diagonal(const std::vector<uint64_t> &qubits, const std::vector<cd> &vec, int64_t cond_regidx,
                const std::string &label) {
    (void)label;
    CircOp op;
    op.type = AERHIP_OP_DIAGONAL;
    op.name = "diagonal";
    op.qubits = qubits;
    op.params = complex_pairs(vec);
    no_condition(cond_regidx, py::none(), "diagonal", op);
    ops.push_back(std::move(op));
  }

  // cmatrix_t arrives as a 2-D complex array; the wire format is column-major.
  void unitary(const std::vector<uint64_t> &qubits, const py::array_t<cd, py::array::forcecast> &mat,
               int64_t cond_regidx, const py::object &expr, const std::string &label) {
    (void)label;
    if (mat.ndim() != 2 || mat.shape(0) != mat.shape(1))
      throw std::invalid_argument("unitary: matrix must be square");
    const uint64_t dim = static_cast<uint64_t>(mat.shape(0));
    if (dim != (1ull << qubits.size())) throw std::invalid_argument("unitary: matrix size does not match qubits");
    CircOp op;
    op.type = AERHIP_OP_MATRIX;
    op.name = "unitary";
    op.qubits = qubits;
    op.params.resize(2 * dim * dim);
    auto m = mat.unchecked<2>();
    for (uint64_t j = 0; j < dim; ++j)
      for (uint64_t i = 0; i < dim; ++i) {
        const cd v = m(i, j);
        op.params[2 * (i + dim * j)] = v.real();
        op.params[2 * (i + dim * j) + 1] = v.imag();
      }
    no_condition(cond_regidx, expr, "unitary", op);
    ops.push_back(std::move(op));
  }

  void initialize(const std::vector<uint64_t> &qubits, const std::vector<cd> &data) {
    CircOp op;
    op.type = AERHIP_OP_INITIALIZE;
    op.name = "initialize";
    op.qubits = qubits;
    op.params = complex_pairs(data);
    ops.push_back(std::move(op));
  }

  // set_statevector on every qubit: the state becomes `data`.
  void set_statevector(const std::vector<uint64_t> &qubits, const std::vector<cd> &data) {
    initialize(qubits, data);
    ops.back().name = "set_statevector";
    if (qubits.size() != num_qubits)
      ops.back().unsupported = "set_statevector must act on every qubit of the circuit";
  }

  void measure(const std::vector<uint64_t> &qubits, const std::vector<uint64_t> &memory,
               const std::vector<uint64_t> &registers) {
    CircOp op;
    op.type = AERHIP_OP_MEASURE;
    op.name = "measure";
    op.qubits = qubits;
    op.memory = memory;
    (void)registers;  // registers feed conditionals, which this engine does not run
    ops.push_back(std::move(op));
  }

  void reset(const std::vector<uint64_t> &qubits, int64_t cond_regidx) {
    CircOp op;
    op.type = AERHIP_OP_RESET;
    op.name = "reset";
    op.qubits = qubits;
    no_condition(cond_regidx, py::none(), "reset", op);
    ops.push_back(std::move(op));
  }

  void barrier(const std::vector<uint64_t> &qubits) {
    CircOp op;
    op.type = AERHIP_OP_BARRIER;
    op.name = "barrier";
    op.qubits = qubits;
    ops.push_back(std::move(op));
  }

  void save_state(const std::vector<uint64_t> &qubits, const std::string &name,
                  const std::string &snapshot_type, const std::string &label) {
    CircOp op;
    op.type = AERHIP_OP_SAVE_STATEVECTOR;
    op.name = name;
    op.qubits = qubits;
    op.key = label.empty() ? name : label;
    if (name != "save_statevector" && name != "save_state")
      op.unsupported = "save instruction '" + name + "' is not supported by this engine";
    else if (qubits.size() != num_qubits)
      op.unsupported = "save_statevector must act on every qubit of the circuit";
    else if (snapshot_type != "single" && snapshot_type != "s")
      op.unsupported = "save_statevector supports the 'single' save type";
    ops.push_back(std::move(op));
  }

  void save_expval(const std::vector<uint64_t> &qubits, const std::string &name,
                   const std::vector<std::string> &pauli_strings, const std::vector<double> &coeff_reals,
                   const std::vector<double> &coeff_imags, const std::string &snapshot_type,
                   const std::string &label) {
    // make_save_expval (operations.hpp:1186-1207): the pairs are (coefficient,
    // coefficient of the squared term) per Pauli string
    if (pauli_strings.size() != coeff_reals.size() || pauli_strings.size() != coeff_imags.size())
      throw std::invalid_argument("save_expval: one coefficient pair per Pauli string");
    CircOp op;
    op.type = -1;
    op.name = name;
    op.qubits = qubits;
    op.key = label.empty() ? name : label;
    op.variance = name == "save_expval_var";
    op.paulis = pauli_strings;
    op.coeffs = coeff_reals;
    op.sq_coeffs = coeff_imags;
    if (op.paulis.empty()) {
      op.paulis.emplace_back(qubits.size(), 'I');
      op.coeffs.push_back(0.0);
      op.sq_coeffs.push_back(0.0);
    }
    if (snapshot_type != "single" && snapshot_type != "average" && snapshot_type != "s" && snapshot_type != "a")
      op.unsupported = "save_expval supports the 'single' and 'average' save types";
    ops.push_back(std::move(op));
  }

  void unsupported_op(const std::string &what) {
    CircOp op;
    op.type = AERHIP_OP_BARRIER;
    op.name = what;
    op.unsupported = "'" + what + "' is not supported by the MI355X statevector engine";
    ops.push_back(std::move(op));
  }
};

// AER::Config's options for the statevector method
// (src/framework/config.hpp; aer_controller_binding.hpp:110-380).
struct AerConfig {
  uint64_t shots = 1024;
  std::string method = "automatic";
  std::string device = "GPU";
  std::string precision = "double";
  bool enable_truncation = true;
  double zero_threshold = 1e-10;
  double validation_threshold = 1e-8;
  std::optional<uint64_t> max_parallel_threads, max_parallel_experiments, max_parallel_shots, max_memory_mb;
  bool fusion_enable = true;
  bool fusion_verbose = false;
  std::optional<uint64_t> fusion_max_qubit, fusion_threshold;
  bool accept_distributed_results = true;
  bool memory = false;
  std::optional<uint64_t> seed_simulator;
  bool cuStateVec_enable = false;
  std::optional<std::vector<uint64_t>> target_gpus;
  bool blocking_enable = false;
  std::optional<uint64_t> blocking_qubits;
  bool batched_shots_gpu = false;
  uint64_t batched_shots_gpu_max_qubits = 16;
  bool shot_branching_enable = false;
  bool shot_branching_sampling_enable = false;
  uint64_t statevector_parallel_threshold = 14;
  uint64_t statevector_sample_measure_opt = 10;
  double chop_threshold = 1e-8;
  std::string library_dir;
  py::object parameterizations = py::list();
  double global_phase = 0.0;
  uint64_t memory_slots = 0;
  std::optional<uint64_t> n_qubits;
  bool runtime_parameter_bind_enable = false;
};

// One experiment's result entry (ExperimentResult::to_python fields).
py::dict run_experiment(AerCircuit &circ, uint64_t index, const AerConfig &cfg, uint64_t seed) {
  py::dict res;
  py::dict data;
  py::dict meta;
  const auto t0 = std::chrono::steady_clock::now();
  res["shots"] = circ.shots;
  res["seed_simulator"] = seed;
  res["header"] = circ.header;
  meta["method"] = "statevector";
  meta["device"] = "GPU";
  meta["precision"] = cfg.precision;
  meta["num_qubits"] = circ.num_qubits;
  meta["num_clbits"] = circ.num_memory;
  meta["fusion"] = py::dict(py::arg("enabled") = cfg.fusion_enable);
  try {
    if (cfg.method != "statevector" && cfg.method != "automatic")
      throw std::invalid_argument("method '" + cfg.method + "' is not provided by the MI355X statevector engine");
    if (cfg.device != "GPU") throw std::invalid_argument("device '" + cfg.device + "' is not this engine's (GPU)");
    if (cfg.precision != "double" && cfg.precision != "single")
      throw std::invalid_argument("precision must be 'double' or 'single'");
    bool has_expval = false, has_measure = false;
    for (const CircOp &op : circ.ops) {
      if (!op.unsupported.empty()) throw std::invalid_argument(op.unsupported);
      for (uint64_t q : op.qubits)
        if (q >= circ.num_qubits) throw std::invalid_argument("instruction qubit out of range");
      has_expval |= op.type == -1;
      has_measure |= op.type == AERHIP_OP_MEASURE || op.type == AERHIP_OP_RESET ||
                     op.type == AERHIP_OP_INITIALIZE;
    }
    const int prec = cfg.precision == "double" ? AERHIP_PRECISION_DOUBLE : AERHIP_PRECISION_SINGLE;
    const int device = cfg.target_gpus && !cfg.target_gpus->empty() ? static_cast<int>((*cfg.target_gpus)[0]) : 0;
    if (!has_expval) {
      // the sampling executor (aerhip_run_circuit)
      aerhip_run_config rc;
      aerhip_run_config_default(&rc);
      rc.precision = prec;
      rc.device = device;
      rc.shots = circ.shots;
      rc.seed = seed;
      rc.fusion_enable = cfg.fusion_enable ? 1 : 0;
      if (cfg.fusion_max_qubit) rc.fusion_max_qubit = static_cast<int32_t>(*cfg.fusion_max_qubit);
      if (cfg.fusion_threshold) rc.fusion_threshold = static_cast<int32_t>(*cfg.fusion_threshold);
      rc.sample_index_size = static_cast<int32_t>(cfg.statevector_sample_measure_opt);
      rc.global_phase = circ.global_phase_angle;
      std::vector<aerhip_op> views;
      std::string sv_key;
      for (const CircOp &op : circ.ops) {
        if (op.type == AERHIP_OP_SAVE_STATEVECTOR) {
          if (!sv_key.empty()) throw std::invalid_argument("one save_statevector per circuit is supported");
          sv_key = op.key;
        }
        views.push_back(op.view());
      }
      std::vector<uint64_t> mem(std::max<uint64_t>(circ.shots, 1));
      py::array_t<cd> sv;
      cd *svp = nullptr;
      if (!sv_key.empty()) {
        sv = py::array_t<cd>(static_cast<py::ssize_t>(1ull << circ.num_qubits));
        svp = sv.mutable_data();
      }
      double timing[4] = {0, 0, 0, 0};
      {
        py::gil_scoped_release nogil;
        ck(aerhip_run_circuit(views.data(), views.size(), circ.num_qubits, circ.num_memory, &rc, mem.data(),
                              reinterpret_cast<double *>(svp), timing));
      }
      if (circ.num_memory && has_measure) {
        // counts keyed by hex memory (SampleVector::to_hex, sample_vector.hpp:78-100)
        std::map<uint64_t, uint64_t> counts;
        for (uint64_t s = 0; s < circ.shots; ++s) counts[mem[s]]++;
        py::dict cdict;
        auto hex = [](uint64_t v) {
          char buf[24];
          std::snprintf(buf, sizeof(buf), "0x%llx", static_cast<unsigned long long>(v));
          return std::string(buf);
        };
        for (const auto &kv : counts) cdict[py::str(hex(kv.first))] = kv.second;
        data["counts"] = cdict;
        if (cfg.memory) {
          py::list ml;
          for (uint64_t s = 0; s < circ.shots; ++s) ml.append(hex(mem[s]));
          data["memory"] = ml;
        }
      }
      if (!sv_key.empty()) data[py::str(sv_key)] = sv;
      meta["time_taken_gates"] = timing[0];
      meta["time_taken_sampling"] = timing[1];
      meta["passes"] = timing[2];
      meta["launches"] = timing[3];
    } else {
      if (has_measure)
        throw std::invalid_argument("save_expval with measure/reset/initialize in one circuit is not supported");
      // A statevector handle: unitary ops are queued, each save_expval term
      // is evaluated inside the tile pass at its point of the circuit.
      aerhip_qv *qv = aerhip_qv_create(circ.num_qubits, prec, device);
      if (!qv) throw std::runtime_error(aerhip_last_error());
      struct Guard {
        aerhip_qv *h;
        ~Guard() { aerhip_qv_destroy(h); }
      } guard{qv};
      ck(aerhip_qv_set_fusion(qv, cfg.fusion_enable ? 1 : 0,
                              static_cast<int>(cfg.fusion_max_qubit.value_or(5)),
                              static_cast<int>(cfg.fusion_threshold.value_or(14))));
      ck(aerhip_qv_initialize(qv));
      if (circ.global_phase_angle != 0.0) {
        const double gp[4] = {std::cos(circ.global_phase_angle), std::sin(circ.global_phase_angle),
                              std::cos(circ.global_phase_angle), std::sin(circ.global_phase_angle)};
        const uint64_t q0 = 0;
        ck(aerhip_apply_diagonal_matrix(qv, &q0, 1, gp));
      }
      struct Saved {
        const CircOp *op;
        std::vector<uint64_t> slots;
      };
      std::vector<Saved> saved;
      std::vector<std::pair<std::string, py::array_t<cd>>> svs;
      std::vector<aerhip_op> run;
      auto flush_run = [&] {
        if (!run.empty()) ck(aerhip_apply_ops(qv, run.data(), run.size()));
        run.clear();
      };
      for (const CircOp &op : circ.ops) {
        if (op.type == -1) {
          flush_run();
          Saved s{&op, {}};
          for (const std::string &p : op.paulis) {
            uint64_t slot = 0;
            ck(aerhip_save_expval_pauli(qv, op.qubits.data(), op.qubits.size(), p.c_str(), 1.0, 0.0, &slot));
            s.slots.push_back(slot);
          }
          saved.push_back(std::move(s));
        } else if (op.type == AERHIP_OP_SAVE_STATEVECTOR) {
          flush_run();
          py::array_t<cd> sv(static_cast<py::ssize_t>(1ull << circ.num_qubits));
          ck(aerhip_qv_copy_to_vector(qv, reinterpret_cast<double *>(sv.mutable_data()), 1ull << circ.num_qubits));
          svs.emplace_back(op.key, sv);
        } else {
          run.push_back(op.view());
        }
      }
      flush_run();
      uint64_t nslots = 0;
      for (const Saved &s : saved) nslots += s.slots.size();
      std::vector<double> vals(nslots + 1);
      uint64_t count = 0;
      ck(aerhip_saved_expvals(qv, vals.data(), vals.size(), &count));
      // Base::apply_save_expval (state.hpp:416-447)
      for (const Saved &s : saved) {
        double ev = 0, sq = 0;
        for (size_t k = 0; k < s.slots.size(); ++k) {
          ev += s.op->coeffs[k] * vals[s.slots[k]];
          sq += s.op->sq_coeffs[k] * vals[s.slots[k]];
        }
        if (s.op->variance) {
          py::list mv;
          mv.append(ev);
          mv.append(sq - ev * ev);
          data[py::str(s.op->key)] = mv;
        } else {
          data[py::str(s.op->key)] = ev;
        }
      }
      for (auto &kv : svs) data[py::str(kv.first)] = kv.second;
    }
    res["status"] = "DONE";
    res["success"] = true;
  } catch (const std::exception &e) {
    res["status"] = std::string("ERROR: ") + e.what();
    res["success"] = false;
  }
  res["data"] = data;
  res["metadata"] = meta;
  res["time_taken"] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  (void)index;
  return res;
}

// controller_execute (src/controllers/controller_execute.hpp:189-211):
// circuit i takes seed_simulator + 2113 * i (a random seed when unset).
py::dict execute(std::vector<AerCircuit *> &circuits, const py::object &noise_model, const AerConfig &cfg) {
  const auto t0 = std::chrono::steady_clock::now();
  py::dict out;
  py::list results;
  bool ok = true;
  std::string err;
  const bool noisy = !noise_model.is_none() && py::len(noise_model) > 0;
  if (noisy) {
    ok = false;
    err = "noise models are not supported by the MI355X statevector engine";
  }
  if (py::len(cfg.parameterizations) > 0) {
    ok = false;
    err = "parameterizations: bind parameters with qiskit_aer_amd.adjoint (aerhip_bind_pauli_expval)";
  }
  uint64_t seed = 0;
  if (cfg.seed_simulator) {
    seed = *cfg.seed_simulator;
  } else {
    std::random_device rd;
    seed = rd() & 0x7FFFFFFF;
  }
  for (size_t i = 0; i < circuits.size(); ++i) {
    if (!circuits[i]) throw std::invalid_argument("null circuit");
    AerCircuit &c = *circuits[i];
    if (!ok) {
      py::dict r;
      r["shots"] = c.shots;
      r["status"] = "ERROR: " + err;
      r["success"] = false;
      r["header"] = c.header;
      r["data"] = py::dict();
      r["metadata"] = py::dict();
      results.append(r);
      continue;
    }
    const uint64_t s = cfg.seed_simulator ? seed + 2113ull * i : (c.seed ? c.seed : seed + 2113ull * i);
    results.append(run_experiment(c, i, cfg, s));
  }
  bool all = true, any = false;
  for (auto r : results) {
    const bool s = r.cast<py::dict>()["success"].cast<bool>();
    all = all && s;
    any = any || s;
  }
  out["results"] = results;
  out["success"] = all;
  out["status"] = all ? "COMPLETED" : (any ? "PARTIAL COMPLETED" : "ERROR");
  if (!ok) out["status"] = "ERROR: " + err;
  py::dict meta;
  meta["omp_enabled"] = false;
  meta["parallel_experiments"] = 1;
  meta["max_gpu_memory_mb"] = 0;
  meta["device"] = "GPU";
  meta["engine"] = aerhip_version();
  out["metadata"] = meta;
  out["backend_name"] = "aer_simulator_statevector_gpu";
  out["backend_version"] = aerhip_version();
  char date[32];
  const std::time_t now = std::time(nullptr);
  std::strftime(date, sizeof(date), "%Y-%m-%dT%H:%M:%S", std::localtime(&now));
  out["date"] = std::string(date);
  out["time_taken"] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return out;
}

}  // namespace

PYBIND11_MODULE(controller_wrappers, m) {
  m.doc() = "controller_wrappers-compatible module of the MI355X statevector engine";
  m.def("aer_initialize_libraries", [](const std::string &) {});

  struct ControllerExecutor {};
  py::class_<ControllerExecutor> ctrl(m, "aer_controller_execute");
  ctrl.def(py::init<>());
  ctrl.def("__reduce__", [ctrl](const ControllerExecutor &) { return py::make_tuple(ctrl, py::tuple()); });
  ctrl.def("execute", [](ControllerExecutor &, std::vector<AerCircuit *> &circuits, const py::object &noise_model,
                         const AerConfig &config) { return execute(circuits, noise_model, config); });
  ctrl.def("available_devices", [](ControllerExecutor &) {
    py::list l;
    if (aerhip_device_count() > 0) l.append("GPU");
    return l;
  });

  py::class_<AerConfig> cfg(m, "AerConfig");
  cfg.def(py::init<>());
#define AERHIP_FIELD(f) cfg.def_readwrite(#f, &AerConfig::f)
  AERHIP_FIELD(shots);
  AERHIP_FIELD(method);
  AERHIP_FIELD(device);
  AERHIP_FIELD(precision);
  AERHIP_FIELD(enable_truncation);
  AERHIP_FIELD(zero_threshold);
  AERHIP_FIELD(validation_threshold);
  AERHIP_FIELD(max_parallel_threads);
  AERHIP_FIELD(max_parallel_experiments);
  AERHIP_FIELD(max_parallel_shots);
  AERHIP_FIELD(max_memory_mb);
  AERHIP_FIELD(fusion_enable);
  AERHIP_FIELD(fusion_verbose);
  AERHIP_FIELD(fusion_max_qubit);
  AERHIP_FIELD(fusion_threshold);
  AERHIP_FIELD(accept_distributed_results);
  AERHIP_FIELD(memory);
  AERHIP_FIELD(seed_simulator);
  AERHIP_FIELD(cuStateVec_enable);
  AERHIP_FIELD(target_gpus);
  AERHIP_FIELD(blocking_enable);
  AERHIP_FIELD(blocking_qubits);
  AERHIP_FIELD(batched_shots_gpu);
  AERHIP_FIELD(batched_shots_gpu_max_qubits);
  AERHIP_FIELD(shot_branching_enable);
  AERHIP_FIELD(shot_branching_sampling_enable);
  AERHIP_FIELD(statevector_parallel_threshold);
  AERHIP_FIELD(statevector_sample_measure_opt);
  AERHIP_FIELD(chop_threshold);
  AERHIP_FIELD(library_dir);
  AERHIP_FIELD(parameterizations);
  AERHIP_FIELD(global_phase);
  AERHIP_FIELD(memory_slots);
  AERHIP_FIELD(n_qubits);
  AERHIP_FIELD(runtime_parameter_bind_enable);
#undef AERHIP_FIELD
  cfg.def_property_readonly_static("GLOBAL_PHASE_POS", [](py::object) { return -1; });

  // Classical-expression types of the reference's module
  // (aer_circuit_binding.hpp:40-130): the compiler imports them, and builds
  // them only for circuits with classical expressions, whose conditional
  // operations this engine reports as unsupported.  They take any arguments.
  enum class UnaryOp { BitNot, LogicNot };
  py::enum_<UnaryOp>(m, "AerUnaryOp").value("BitNot", UnaryOp::BitNot).value("LogicNot", UnaryOp::LogicNot);
  enum class BinaryOp { BitAnd, BitOr, BitXor, LogicAnd, LogicOr, Equal, NotEqual, Less, LessEqual, Greater,
                        GreaterEqual };
  py::enum_<BinaryOp>(m, "AerBinaryOp")
      .value("BitAnd", BinaryOp::BitAnd).value("BitOr", BinaryOp::BitOr).value("BitXor", BinaryOp::BitXor)
      .value("LogicAnd", BinaryOp::LogicAnd).value("LogicOr", BinaryOp::LogicOr)
      .value("Equal", BinaryOp::Equal).value("NotEqual", BinaryOp::NotEqual).value("Less", BinaryOp::Less)
      .value("LessEqual", BinaryOp::LessEqual).value("Greater", BinaryOp::Greater)
      .value("GreaterEqual", BinaryOp::GreaterEqual);
  struct ExprStub {};
  // one Python type per name, each a subclass of one registered base (a C++
  // type is registered once)
  py::class_<ExprStub>(m, "_AerExprBase", py::dynamic_attr())
      .def(py::init([](py::args, py::kwargs) { return ExprStub{}; }));
  {
    py::object base = m.attr("_AerExprBase");
    py::object type = py::module_::import("builtins").attr("type");
    for (const char *name : {"AerScalarType", "AerUint", "AerBool", "AerExpr", "AerCast", "AerVar", "AerValue",
                             "AerUintValue", "AerBoolValue", "AerUnaryExpr", "AerBinaryExpr"})
      m.attr(name) = type(name, py::make_tuple(base), py::dict());
  }

  py::class_<AerCircuit, std::shared_ptr<AerCircuit>> circ(m, "AerCircuit");
  circ.def(py::init<>());
  circ.def("__repr__", [](const AerCircuit &c) {
    return "AerCircuit(num_qubits=" + std::to_string(c.num_qubits) + ", ops=" + std::to_string(c.ops.size()) + ")";
  });
  circ.def_readwrite("circ_id", &AerCircuit::circ_id);
  circ.def_readwrite("shots", &AerCircuit::shots);
  circ.def_readwrite("num_qubits", &AerCircuit::num_qubits);
  circ.def_readwrite("num_memory", &AerCircuit::num_memory);
  circ.def_readwrite("seed", &AerCircuit::seed);
  circ.def_readwrite("global_phase_angle", &AerCircuit::global_phase_angle);
  circ.def_property_readonly("num_ops", [](const AerCircuit &c) { return c.ops.size(); });
  circ.def("set_header", [](AerCircuit &c, const py::handle &h) { c.header = py::reinterpret_borrow<py::object>(h); });
  circ.def("gate", &AerCircuit::gate, py::arg("name"), py::arg("qubits"), py::arg("params"),
           py::arg("string_params"), py::arg("cond_regidx") = -1, py::arg("expr") = py::none(),
           py::arg("label") = "");
  circ.def("diagonal", &AerCircuit::diagonal, py::arg("qubits"), py::arg("vec"), py::arg("cond_regidx") = -1,
           py::arg("label") = "");
  circ.def("unitary", &AerCircuit::unitary, py::arg("qubits"), py::arg("mat"), py::arg("cond_regidx") = -1,
           py::arg("expr") = py::none(), py::arg("label") = "");
  circ.def("initialize", &AerCircuit::initialize);
  circ.def("set_statevector", &AerCircuit::set_statevector);
  circ.def("measure", &AerCircuit::measure);
  circ.def("reset", &AerCircuit::reset, py::arg("qubits"), py::arg("cond_regidx") = -1);
  circ.def("barrier", &AerCircuit::barrier);
  circ.def("save_state", &AerCircuit::save_state, py::arg("qubits"), py::arg("name"),
           py::arg("snapshot_type"), py::arg("label") = "");
  circ.def("save_expval", &AerCircuit::save_expval, py::arg("qubits"), py::arg("name"),
           py::arg("pauli_strings"), py::arg("coeff_reals"), py::arg("coeff_imags"),
           py::arg("snapshot_type"), py::arg("label") = "");
  // Instructions of other methods or of noise/dynamic circuits: recorded,
  // then reported as a failed experiment at execution.
  for (const char *name : {"bfunc", "roerror", "multiplexer", "kraus", "superop", "save_amplitudes",
                           "set_density_matrix", "set_unitary", "set_superop", "set_matrix_product_state",
                           "set_clifford", "jump", "mark", "set_qerror_loc", "store"}) {
    const std::string n = name;
    circ.def(name, [n](AerCircuit &c, py::args, py::kwargs) { c.unsupported_op(n); });
  }
}
