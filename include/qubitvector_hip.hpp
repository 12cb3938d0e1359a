// qubitvector_hip.hpp -- C++ adaptor with the QubitVector method set over the
// engine's C ABI (aerhip.h).
//
// This is the class a maintainer substitutes for QV::QubitVectorThrust<T> as
// the `statevec_t` of Statevector::State<statevec_t>
// (src/simulators/statevector/statevector_state.hpp:100-101; instantiated in
// Controller::make_circuit_executor, src/controllers/aer_controller.hpp:
// 624-649, and AerState::initialize_qreg_state,
// src/controllers/state_controller.hpp:666-677).  Method names, argument
// meaning and error behaviour (std::runtime_error / std::invalid_argument
// with the engine's message) follow QubitVector
// (src/simulators/statevector/qubitvector.hpp:61-651).
//
// Header only; link against libaerhip.so.
#pragma once

#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "aerhip.h"

namespace aerhip {

template <typename data_t = double>
class QubitVectorHip {
 public:
  using reg_t = std::vector<uint64_t>;
  using cvector_t = std::vector<std::complex<double>>;

  QubitVectorHip() = default;
  explicit QubitVectorHip(size_t num_qubits) { set_num_qubits(num_qubits); }
  ~QubitVectorHip() { release(); }
  QubitVectorHip(const QubitVectorHip &) = delete;
  QubitVectorHip &operator=(const QubitVectorHip &) = delete;
  QubitVectorHip(QubitVectorHip &&o) noexcept : h_(std::exchange(o.h_, nullptr)), n_(o.n_) {}
  QubitVectorHip &operator=(QubitVectorHip &&o) noexcept {
    release();
    h_ = std::exchange(o.h_, nullptr);
    n_ = o.n_;
    return *this;
  }

  static constexpr int precision() {
    return sizeof(data_t) == sizeof(double) ? AERHIP_PRECISION_DOUBLE : AERHIP_PRECISION_SINGLE;
  }
  std::string name() const { return "statevector_hip"; }

  // -- lifecycle (qubitvector.hpp:861-1109) ---------------------------------
  void set_num_qubits(size_t n) {
    release();
    h_ = aerhip_qv_create(n, precision(), device_);
    if (!h_) throw std::runtime_error(aerhip_last_error());
    n_ = n;
  }
  size_t num_qubits() const { return n_; }
  size_t size() const { return size_t(1) << n_; }
  size_t required_memory_mb(uint64_t n) const { return aerhip_required_memory_mb(n, precision()); }
  void set_target_gpus(const reg_t &t) { device_ = t.empty() ? 0 : static_cast<int>(t[0]); }
  void initialize() { ck(aerhip_qv_initialize(h_)); }
  void zero() { ck(aerhip_qv_zero(h_)); }
  template <typename list_t>
  void initialize_from_vector(const list_t &vec) {
    std::vector<std::complex<double>> v(vec.begin(), vec.end());
    ck(aerhip_qv_initialize_from_vector(h_, reinterpret_cast<const double *>(v.data()), v.size()));
  }
  void initialize_component(const reg_t &qubits, const cvector_t &state) {
    ck(aerhip_qv_initialize_component(h_, qubits.data(), qubits.size(),
                                      reinterpret_cast<const double *>(state.data())));
  }
  std::vector<std::complex<double>> copy_to_vector() const {
    std::vector<std::complex<double>> v(size());
    ck(aerhip_qv_copy_to_vector(h_, reinterpret_cast<double *>(v.data()), v.size()));
    return v;
  }
  void synchronize() const { ck(aerhip_qv_synchronize(h_)); }
  void set_sample_measure_index_size(int n) { ck(aerhip_qv_set_sample_measure_index_size(h_, n)); }
  // OpenMP settings of the CPU vector have no meaning on the device.
  void set_omp_threads(int) {}
  void set_omp_threshold(int) {}
  void set_json_chop_threshold(double) {}
  bool enable_batch(bool) const { return false; }
  bool support_global_indexing() const { return false; }

  // -- gates (qubitvector.hpp:1299-1871, 2393-2433) --------------------------
  void apply_matrix(const reg_t &q, const cvector_t &mat) {
    ck(aerhip_apply_matrix(h_, q.data(), q.size(), reinterpret_cast<const double *>(mat.data())));
  }
  void apply_diagonal_matrix(const reg_t &q, const cvector_t &diag) {
    ck(aerhip_apply_diagonal_matrix(h_, q.data(), q.size(),
                                    reinterpret_cast<const double *>(diag.data())));
  }
  void apply_mcx(const reg_t &q) { ck(aerhip_apply_mcx(h_, q.data(), q.size())); }
  void apply_mcy(const reg_t &q) { ck(aerhip_apply_mcy(h_, q.data(), q.size())); }
  void apply_mcswap(const reg_t &q) { ck(aerhip_apply_mcswap(h_, q.data(), q.size())); }
  void apply_mcphase(const reg_t &q, const std::complex<double> phase) {
    ck(aerhip_apply_mcphase(h_, q.data(), q.size(), phase.real(), phase.imag()));
  }
  void apply_mcu(const reg_t &q, const cvector_t &mat) {
    ck(aerhip_apply_mcu(h_, q.data(), q.size(), reinterpret_cast<const double *>(mat.data())));
  }
  // axis: AERHIP_ROT_* in the order of QV::Rotation (qubitvector.hpp:47-55)
  void apply_rotation(const reg_t &q, int axis, double theta) {
    ck(aerhip_apply_rotation(h_, q.data(), q.size(), axis, theta));
  }
  void apply_pauli(const reg_t &q, const std::string &pauli,
                   const std::complex<double> &coeff = 1) {
    ck(aerhip_apply_pauli(h_, q.data(), q.size(), pauli.c_str(), coeff.real(), coeff.imag()));
  }
  void apply_permutation_matrix(const reg_t &q,
                                const std::vector<std::pair<uint64_t, uint64_t>> &pairs) {
    std::vector<uint64_t> flat;
    for (const auto &p : pairs) {
      flat.push_back(p.first);
      flat.push_back(p.second);
    }
    ck(aerhip_apply_permutation_matrix(h_, q.data(), q.size(), flat.data(), pairs.size()));
  }
  void apply_multiplexer(const reg_t &controls, const reg_t &targets, const cvector_t &mat) {
    ck(aerhip_apply_multiplexer(h_, controls.data(), controls.size(), targets.data(),
                                targets.size(), reinterpret_cast<const double *>(mat.data())));
  }
  void apply_multi_swaps(const reg_t &q) { ck(aerhip_apply_multi_swaps(h_, q.data(), q.size())); }

  // -- readout (qubitvector.hpp:1879-2347) ---------------------------------
  double norm() const {
    double v;
    ck(aerhip_norm(h_, &v));
    return v;
  }
  double probability(uint64_t outcome) const {
    double v;
    ck(aerhip_probability(h_, outcome, &v));
    return v;
  }
  std::vector<double> probabilities() const {
    std::vector<double> p(size());
    ck(aerhip_probabilities(h_, nullptr, 0, p.data()));
    return p;
  }
  std::vector<double> probabilities(const reg_t &q) const {
    std::vector<double> p(size_t(1) << q.size());
    ck(aerhip_probabilities(h_, q.data(), q.size(), p.data()));
    return p;
  }
  reg_t sample_measure(const std::vector<double> &rnds) const {
    reg_t out(rnds.size());
    ck(aerhip_sample_measure(h_, rnds.data(), rnds.size(), out.data()));
    return out;
  }
  double expval_pauli(const reg_t &q, const std::string &pauli,
                      const std::complex<double> initial_phase = 1.0) const {
    double v;
    ck(aerhip_expval_pauli(h_, q.data(), q.size(), pauli.c_str(), initial_phase.real(),
                           initial_phase.imag(), &v));
    return v;
  }

 private:
  static void ck(int rc) {
    if (rc == 2) throw std::invalid_argument(aerhip_last_error());
    if (rc != 0) throw std::runtime_error(aerhip_last_error());
  }
  void release() {
    if (h_) aerhip_qv_destroy(h_);
    h_ = nullptr;
  }
  aerhip_qv *h_ = nullptr;
  size_t n_ = 0;
  int device_ = 0;
};

}  // namespace aerhip
