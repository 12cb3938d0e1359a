// Compile-and-link check of the QubitVector drop-in (include/qubitvector_hip.hpp).
//
// MockState<statevec_t> repeats every call that Statevector::State<statevec_t>
// makes on its `qreg_` (src/simulators/statevector/statevector_state.hpp),
// with the argument and result types the reference passes there, and the
// executor-side calls of the batched GPU path.  The reference's own types it
// needs (QV::Rotation, AER::Vector, Operations::Op, RngEngine) are given
// minimal stand-ins here with the same shape.  Nothing is executed: the test
// only builds this file against the adaptor and libaerhip.so.
#include <complex>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "qubitvector_hip.hpp"

namespace AER {
using uint_t = uint64_t;
using int_t = int64_t;
using reg_t = std::vector<uint_t>;
using complex_t = std::complex<double>;
using cvector_t = std::vector<complex_t>;
using rvector_t = std::vector<double>;

// Shape of AER::Vector (src/framework/linalg/vector.hpp): (size, fill) and data().
template <typename T>
class Vector {
 public:
  Vector(size_t n, bool fill = true) : v_(n) { (void)fill; }
  T *data() { return v_.data(); }
  size_t size() const { return v_.size(); }

 private:
  std::vector<T> v_;
};

namespace QV {
enum class Rotation { x, y, z, xx, yy, zz, zx };  // qubitvector.hpp:47-55
}
namespace Operations {
struct Op {
  std::string name;
  reg_t qubits;
  std::vector<std::string> string_params;
};
}  // namespace Operations
struct RngEngine {
  double rand(double a, double b) { return a + (b - a) * 0.5; }
};

template <class statevec_t>
struct MockState {
  statevec_t qreg_;
  reg_t target_gpus_;

  void initialize_qreg(uint_t num_qubits) {                       // :394-402
    qreg_.set_omp_threshold(14);                                  // :419
    qreg_.set_omp_threads(4);                                     // :421
    qreg_.set_num_qubits(num_qubits);                             // :398
    qreg_.initialize();                                           // :399
  }
  void initialize_statevector(statevec_t &&state) {               // :404-414
    if (state.num_qubits() != qreg_.num_qubits()) return;
    qreg_ = std::move(state);
  }
  bool allocate(uint_t num_qubits, uint_t block_bits) {           // :424-440
    qreg_.set_max_matrix_bits(5);                                 // :428
    qreg_.set_max_sampling_shots(1024);                           // :430
    qreg_.set_target_gpus(target_gpus_);                          // :432
    qreg_.cuStateVec_enable(false);                               // :434
    qreg_.chunk_setup(block_bits, num_qubits, 0, 1);              // :436
    return true;
  }
  void apply_global_phase(complex_t g) {                          // :446-450
    qreg_.apply_diagonal_matrix({0}, {g, g});
  }
  size_t required_memory_mb(uint_t n) const { return qreg_.required_memory_mb(n); }  // :456
  void set_config(double chop, int index_size) {                  // :461-475
    qreg_.set_json_chop_threshold(chop);
    qreg_.set_sample_measure_index_size(index_size);
  }
  auto move_to_vector() { return std::move(qreg_.move_to_vector()); }  // :479
  auto copy_to_vector() { return qreg_.copy_to_vector(); }              // :484
  void register_blocking(const Operations::Op &op) {              // :532-536
    qreg_.enter_register_blocking(op.qubits);
    qreg_.leave_register_blocking();
  }
  double expval_pauli(const reg_t &qubits, const std::string &pauli) {  // :593
    return qreg_.expval_pauli(qubits, pauli);
  }
  void save_statevector_dict(double chop) {                       // :626
    auto ket = qreg_.vector_ket(chop);
    std::map<std::string, complex_t> out;
    for (auto const &it : ket) out[it.first] = it.second;
  }
  double save_density_trace() { return qreg_.norm(); }            // :644
  void save_amplitudes(const std::vector<uint_t> &idx) {          // :660-672
    std::vector<complex_t> amps(idx.size());
    rvector_t amps_sq(idx.size());
    for (size_t i = 0; i < idx.size(); ++i) {
      amps[i] = qreg_.get_state(idx[i]);
      amps_sq[i] = qreg_.probability(idx[i]);
    }
  }
  void global_index_checks() {                                    // :735-747
    const bool local = qreg_.support_global_indexing();
    const uint_t ci = qreg_.chunk_index();
    (void)local;
    (void)ci;
  }
  void apply_gate(const Operations::Op &op, double theta) {       // :754-858
    qreg_.apply_mcx(op.qubits);
    qreg_.apply_mcy(op.qubits);
    qreg_.apply_mcphase(op.qubits, -1);
    qreg_.apply_mcu(op.qubits, cvector_t(4));
    qreg_.apply_rotation(op.qubits, QV::Rotation::x, theta);
    qreg_.apply_rotation(op.qubits, QV::Rotation::y, theta);
    qreg_.apply_rotation(op.qubits, QV::Rotation::z, theta);
    qreg_.apply_rotation(op.qubits, QV::Rotation::xx, theta);
    qreg_.apply_rotation(op.qubits, QV::Rotation::yy, theta);
    qreg_.apply_rotation(op.qubits, QV::Rotation::zz, theta);
    qreg_.apply_rotation(op.qubits, QV::Rotation::zx, theta);
    qreg_.apply_matrix(op.qubits, cvector_t(16));
    qreg_.apply_mcswap(op.qubits);
    qreg_.apply_mcphase(op.qubits, std::exp(complex_t(0, 1) * theta));
    qreg_.apply_pauli(op.qubits, op.string_params[0]);
  }
  void apply_multiplexer(const reg_t &c, const reg_t &t, const cvector_t &vmat) {  // :874
    qreg_.apply_multiplexer(c, t, vmat);
  }
  void apply_matrix(const reg_t &qubits, const cvector_t &vmat) {  // :884-897
    qreg_.apply_matrix(qubits, vmat);
  }
  void apply_diagonal_matrix(const reg_t &qubits, const cvector_t &diag) {  // :904-913
    qreg_.apply_diagonal_matrix(qubits, diag);
  }
  rvector_t measure_probs(const reg_t &qubits) const {           // :948
    return qreg_.probabilities(qubits);
  }
  void measure_reset_update(const reg_t &qubits, uint_t meas, double p) {  // :970-1011
    cvector_t mdiag(2, 0.);
    mdiag[meas] = 1. / std::sqrt(p);
    qreg_.apply_diagonal_matrix(qubits, mdiag);
    qreg_.apply_mcx(qubits);
  }
  reg_t sample_measure(uint_t shots, RngEngine &rng) {            // :1017-1029
    std::vector<double> rnds;
    for (uint_t i = 0; i < shots; ++i) rnds.push_back(rng.rand(0, 1));
    return qreg_.sample_measure(rnds);
  }
  void apply_initialize(const reg_t &qubits, const cvector_t &params) {  // :1072-1089
    if (qubits.size() == qreg_.num_qubits()) qreg_.initialize_from_vector(params);
    qreg_.initialize_component(qubits, params);
  }
  double kraus_norm(const reg_t &qubits, const cvector_t &vmat) {  // :1136
    return qreg_.norm(qubits, vmat);
  }
  // Batched GPU executor (statevector_executor.hpp): reported unsupported.
  bool batched() { return qreg_.batched_optimization_supported(); }
  void other_methods(uint_t q) {
    qreg_.zero();
    qreg_.synchronize();
    (void)qreg_.size();
    (void)qreg_.norm(q, cvector_t(4));
    (void)qreg_.norm_diagonal(q, cvector_t(2));
    (void)qreg_.norm_diagonal(reg_t({q}), cvector_t(2));
    (void)qreg_.probabilities();
    std::vector<std::pair<uint_t, uint_t>> pairs = {{0, 1}};
    qreg_.apply_permutation_matrix(reg_t({q}), pairs);
    qreg_.apply_multi_swaps(reg_t({q, q + 1}));
    complex_t v(1, 0);
    std::complex<typename std::conditional<sizeof(qreg_.get_state(0).real()) == 8, double,
                                           float>::type> vv(v.real(), v.imag());
    qreg_.set_state(0, vv);
    (void)qreg_.vector();
    (void)statevec_t::name();
  }
};

}  // namespace AER

using QVd = aerhip::QubitVectorHip<double, AER::Vector<std::complex<double>>>;
using QVf = aerhip::QubitVectorHip<float, AER::Vector<std::complex<float>>>;
template class aerhip::QubitVectorHip<double, AER::Vector<std::complex<double>>>;
template class aerhip::QubitVectorHip<float, AER::Vector<std::complex<float>>>;
template struct AER::MockState<QVd>;
template struct AER::MockState<QVf>;

int main() { return 0; }
