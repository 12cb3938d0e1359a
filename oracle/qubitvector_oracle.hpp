// ORACLE -- test infrastructure only.
//
// CPU restatement of Qiskit Aer's QubitVector<T> (generic, non-AVX path) used
// as the correctness checker of the HIP engine and as the timed CPU baseline.
// Nothing in the product links, imports or calls this code; only tests/,
// __graft_entry__.smoke() and bench.py's CPU baseline do.
//
// Every function cites the reference code it restates (paths relative to the
// reference checkout).  Differences from the reference, all deliberate:
//   - apply_diagonal_matrix_1's [[+-i,0],[0,1]] branches update inds[0] (the
//     reference's generic branch updates inds[1], transformer.hpp:311-337;
//     its AVX2 path, used on every AVX2 host, multiplies inds[0] as here);
//   - sample_measure returns size-1 when rnd exceeds the accumulated total
//     (the reference would return size, qubitvector.hpp:2197-2224);
//   - single-precision probabilities accumulate per-thread partials in
//     double (the reference keeps them in float, qubitvector.hpp:2129).
#pragma once

#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace oracle {

using uint_t = uint64_t;
using int_t = int64_t;
using reg_t = std::vector<uint_t>;
using cd = std::complex<double>;
template <typename T> using cvector_t = std::vector<std::complex<T>>;

inline uint_t BIT(uint_t q) { return 1ull << q; }
inline uint_t MASK(uint_t q) { return q >= 64 ? ~0ull : (1ull << q) - 1ull; }

// index0 (indexes.hpp:175-186): insert zero bits at the sorted positions.
inline uint_t index0(const reg_t &qs_sorted, uint_t k) {
  uint_t ret = k;
  for (uint_t q : qs_sorted) {
    const uint_t low = ret & MASK(q);
    ret >>= q;
    ret <<= q + 1;
    ret |= low;
  }
  return ret;
}

// indexes (indexes.hpp:197-250): the 2^N indices of group k; bit i of the
// position is qubits[i].
inline void indexes(const reg_t &qs, const reg_t &qs_sorted, uint_t k, uint_t *ret) {
  const size_t N = qs.size();
  ret[0] = index0(qs_sorted, k);
  for (size_t i = 0; i < N; ++i) {
    const uint_t n = BIT(i), b = BIT(qs[i]);
    for (uint_t j = 0; j < n; ++j) ret[n + j] = ret[j] | b;
  }
}

// Optional transformer of complex128 vectors: the reference's own compiled
// Transformer / TransformerAVX2 (oracle/_ref), installed by the CPU baseline
// so that the timed gate kernels are the reference's.  Arguments: data,
// number of qubits of the state, qubits, number of qubits, column-major
// matrix (or diagonal) as (re, im) pairs, threads.  Null: the restatement.
using TransformerHook = int (*)(double *, uint64_t, const uint64_t *, uint64_t, const double *, int);
inline TransformerHook g_matrix_hook = nullptr;
inline TransformerHook g_diagonal_hook = nullptr;

template <typename data_t>
class QubitVector {
 public:
  using cT = std::complex<data_t>;

  explicit QubitVector(size_t n) { set_num_qubits(n); }
  ~QubitVector() { std::free(data_); }
  QubitVector(const QubitVector &) = delete;
  QubitVector &operator=(const QubitVector &) = delete;

  size_t num_qubits() const { return num_qubits_; }
  size_t size() const { return data_size_; }
  cT *data() { return data_; }
  const cT *data() const { return data_; }
  void set_omp_threads(int n) { omp_threads_ = n > 0 ? n : 1; }
  void set_omp_threshold(int n) { omp_threshold_ = n; }
  void set_sample_measure_index_size(int n) { sample_index_size_ = n; }

  // set_num_qubits / allocate_mem (qubitvector.hpp:922-967)
  void set_num_qubits(size_t n) {
    std::free(data_);
    data_ = nullptr;
    num_qubits_ = n;
    data_size_ = BIT(n);
    void *p = nullptr;
    if (posix_memalign(&p, 64, sizeof(cT) * data_size_) != 0)
      throw std::runtime_error("Cannot allocate memory by posix_memalign");
    data_ = static_cast<cT *>(p);
  }

  // omp_threads_managed (qubitvector.hpp:486-489)
  int threads() const {
    return (num_qubits_ > omp_threshold_ && omp_threads_ > 1) ? omp_threads_ : 1;
  }

  // zero / initialize / initialize_from_vector (qubitvector.hpp:903-1109)
  void zero() {
    const int_t END = data_size_;
#pragma omp parallel for if (threads() > 1) num_threads(threads())
    for (int_t k = 0; k < END; ++k) data_[k] = 0.0;
  }
  void initialize() {
    zero();
    data_[0] = 1.;
  }
  void initialize_from_vector(const cd *vec, size_t len) {
    if (len != data_size_)
      throw std::runtime_error("QubitVector::initialize input vector is incorrect length (" +
                               std::to_string(data_size_) + "!=" + std::to_string(len) + ")");
    const int_t END = data_size_;
#pragma omp parallel for if (threads() > 1) num_threads(threads())
    for (int_t k = 0; k < END; ++k) data_[k] = cT(vec[k]);
  }

  // Generic group loop (apply_lambda, indexes.hpp:280-333).
  template <typename F>
  void for_groups(const reg_t &qs, F &&f) const {
    reg_t sorted = qs;
    std::sort(sorted.begin(), sorted.end());
    const int_t END = data_size_ >> qs.size();
    const size_t D = BIT(qs.size());
#pragma omp parallel num_threads(threads()) if (threads() > 1)
    {
      std::vector<uint_t> inds(D);
#pragma omp for
      for (int_t k = 0; k < END; ++k) {
        indexes(qs, sorted, k, inds.data());
        f(inds.data());
      }
    }
  }

  cvector_t<data_t> convert(const cvector_t<double> &v) const {
    cvector_t<data_t> r(v.size());
    for (size_t i = 0; i < v.size(); ++i) r[i] = v[i];
    return r;
  }

  // initialize_component (qubitvector.hpp:879-896)
  void initialize_component(const reg_t &qs, const cvector_t<double> &state0) {
    const auto state = convert(state0);
    const size_t D = BIT(qs.size());
    for_groups(qs, [&](const uint_t *inds) {
      const cT cache = data_[inds[0]];
      for (size_t i = 0; i < D; ++i) data_[inds[i]] = cache * state[i];
    });
  }

  // Transformer::apply_matrix (transformer.hpp:89-164)
  void apply_matrix(const reg_t &qs, const cvector_t<double> &mat) {
    if (qs.size() > 20) throw std::runtime_error("Maximum size of apply matrix is a 20-qubit matrix.");
    if (std::is_same<data_t, double>::value && g_matrix_hook) {
      if (g_matrix_hook(reinterpret_cast<double *>(data_), num_qubits_, qs.data(), qs.size(),
                        reinterpret_cast<const double *>(mat.data()), threads()) != 0)
        throw std::runtime_error("reference transformer failed");
      return;
    }
    if (qs.size() == 1) return apply_matrix_1(qs[0], mat);
    const auto m = convert(mat);
    const size_t D = BIT(qs.size());
    for_groups(qs, [&](const uint_t *inds) {
      std::vector<cT> cache(D);
      for (size_t i = 0; i < D; ++i) {
        cache[i] = data_[inds[i]];
        data_[inds[i]] = 0.;
      }
      for (size_t i = 0; i < D; ++i)
        for (size_t j = 0; j < D; ++j) data_[inds[i]] += m[i + D * j] * cache[j];
    });
  }

  // Transformer::apply_matrix_1 (transformer.hpp:166-232)
  void apply_matrix_1(uint_t q, const cvector_t<double> &mat) {
    if (mat[1] == 0.0 && mat[2] == 0.0) return apply_diagonal_matrix_1(q, {mat[0], mat[3]});
    const auto m = convert(mat);
    const reg_t qs = {q};
    if (mat[0] == 0.0 && mat[3] == 0.0) {
      if (mat[1] == 1.0 && mat[2] == 1.0) {
        for_groups(qs, [&](const uint_t *inds) { std::swap(data_[inds[0]], data_[inds[1]]); });
        return;
      }
      if (mat[2] == 0.0) {
        for_groups(qs, [&](const uint_t *inds) {
          data_[inds[1]] = m[1] * data_[inds[0]];
          data_[inds[0]] = 0.0;
        });
        return;
      }
      if (mat[1] == 0.0) {
        for_groups(qs, [&](const uint_t *inds) {
          data_[inds[0]] = m[2] * data_[inds[1]];
          data_[inds[1]] = 0.0;
        });
        return;
      }
      for_groups(qs, [&](const uint_t *inds) {
        const cT cache = data_[inds[0]];
        data_[inds[0]] = m[2] * data_[inds[1]];
        data_[inds[1]] = m[1] * cache;
      });
      return;
    }
    for_groups(qs, [&](const uint_t *inds) {
      const cT cache = data_[inds[0]];
      data_[inds[0]] = m[0] * cache + m[2] * data_[inds[1]];
      data_[inds[1]] = m[1] * cache + m[3] * data_[inds[1]];
    });
  }

  // Transformer::apply_diagonal_matrix (transformer.hpp:234-258)
  void apply_diagonal_matrix(const reg_t &qs, const cvector_t<double> &diag) {
    if (std::is_same<data_t, double>::value && g_diagonal_hook) {
      if (g_diagonal_hook(reinterpret_cast<double *>(data_), num_qubits_, qs.data(), qs.size(),
                          reinterpret_cast<const double *>(diag.data()), threads()) != 0)
        throw std::runtime_error("reference transformer failed");
      return;
    }
    if (qs.size() == 1) return apply_diagonal_matrix_1(qs[0], diag);
    const auto d = convert(diag);
    const size_t N = qs.size();
    for_groups(reg_t{qs[0]}, [&](const uint_t *inds) {
      for (int i = 0; i < 2; ++i) {
        const uint_t k = inds[i];
        uint_t iv = 0;
        for (size_t j = 0; j < N; ++j)
          if (k & BIT(qs[j])) iv += BIT(j);
        if (d[iv] != (data_t)1.0) data_[k] *= d[iv];
      }
    });
  }

  // Transformer::apply_diagonal_matrix_1 (transformer.hpp:260-372)
  void apply_diagonal_matrix_1(uint_t q, const cvector_t<double> &diag) {
    const auto d = convert(diag);
    const reg_t qs = {q};
    if (diag[0] == 1.0) {
      if (diag[1] == 1.0) return;
      for_groups(qs, [&](const uint_t *inds) { data_[inds[1]] *= d[1]; });
      return;
    }
    if (diag[1] == 1.0) {
      for_groups(qs, [&](const uint_t *inds) { data_[inds[0]] *= d[0]; });
      return;
    }
    for_groups(qs, [&](const uint_t *inds) {
      data_[inds[0]] *= d[0];
      data_[inds[1]] *= d[1];
    });
  }

  // apply_mcx (qubitvector.hpp:1447-1486)
  void apply_mcx(const reg_t &qs) {
    const size_t N = qs.size();
    const uint_t p0 = MASK(N - 1), p1 = MASK(N);
    for_groups(qs, [&](const uint_t *inds) { std::swap(data_[inds[p0]], data_[inds[p1]]); });
  }

  // apply_mcy (qubitvector.hpp:1489-1537)
  void apply_mcy(const reg_t &qs) {
    const size_t N = qs.size();
    const uint_t p0 = MASK(N - 1), p1 = MASK(N);
    const cT I(0., 1.);
    for_groups(qs, [&](const uint_t *inds) {
      const cT cache = data_[inds[p0]];
      data_[inds[p0]] = -I * data_[inds[p1]];
      data_[inds[p1]] = I * cache;
    });
  }

  // apply_mcswap (qubitvector.hpp:1540-1573)
  void apply_mcswap(const reg_t &qs) {
    const size_t N = qs.size();
    const uint_t p0 = MASK(N - 1), p1 = p0 + BIT(N - 2);
    for_groups(qs, [&](const uint_t *inds) { std::swap(data_[inds[p0]], data_[inds[p1]]); });
  }

  // apply_mcphase (qubitvector.hpp:1576-1612)
  void apply_mcphase(const reg_t &qs, cd phase) {
    const uint_t last = MASK(qs.size());
    for_groups(qs, [&](const uint_t *inds) { data_[inds[last]] *= phase; });
  }

  // apply_mcu (qubitvector.hpp:1615-1720)
  void apply_mcu(const reg_t &qs, const cvector_t<double> &mat) {
    const size_t N = qs.size();
    const uint_t p0 = MASK(N - 1), p1 = MASK(N);
    if (mat[1] == 0.0 && mat[2] == 0.0) {
      if (mat[0] == 1.0) return apply_mcphase(qs, mat[3]);
      const cvector_t<double> diag = {mat[0], mat[3]};
      if (N == 1) return apply_diagonal_matrix(qs, diag);
      const auto d = convert(diag);
      for_groups(qs, [&](const uint_t *inds) {
        data_[inds[p0]] = d[0] * data_[inds[p0]];
        data_[inds[p1]] = d[1] * data_[inds[p1]];
      });
      return;
    }
    if (N == 1) return apply_matrix(qs, mat);
    const auto m = convert(mat);
    for_groups(qs, [&](const uint_t *inds) {
      const cT cache = data_[inds[p0]];
      data_[inds[p0]] = m[0] * data_[inds[p0]] + m[2] * data_[inds[p1]];
      data_[inds[p1]] = m[1] * cache + m[3] * data_[inds[p1]];
    });
  }

  // apply_permutation_matrix (qubitvector.hpp:1350-1434)
  void apply_permutation_matrix(const reg_t &qs, const std::vector<std::pair<uint_t, uint_t>> &pairs) {
    for_groups(qs, [&](const uint_t *inds) {
      for (const auto &p : pairs) std::swap(data_[inds[p.first]], data_[inds[p.second]]);
    });
  }

  // apply_multiplexer (qubitvector.hpp:1306-1340)
  void apply_multiplexer(const reg_t &ctrl, const reg_t &tgt, const cvector_t<double> &mat) {
    const auto m = convert(mat);
    reg_t qs = tgt;
    qs.insert(qs.end(), ctrl.begin(), ctrl.end());
    const uint_t DIM = BIT(tgt.size() + ctrl.size()), cols = BIT(tgt.size()), blocks = BIT(ctrl.size());
    for_groups(qs, [&](const uint_t *inds) {
      std::vector<cT> cache(DIM);
      for (uint_t i = 0; i < DIM; ++i) {
        cache[i] = data_[inds[i]];
        data_[inds[i]] = 0.;
      }
      for (uint_t b = 0; b < blocks; ++b)
        for (uint_t i = 0; i < cols; ++i)
          for (uint_t j = 0; j < cols; ++j)
            data_[inds[i + b * cols]] += m[i + b * cols + DIM * j] * cache[b * cols + j];
    });
  }

  // apply_multi_swaps (qubitvector.hpp:1843-1871), pairs within each group
  // of ten qubits.
  void apply_multi_swaps(const reg_t &qs) {
    for (size_t i = 0; i < qs.size(); i += 10) {
      const size_t n = std::min<size_t>(10, qs.size() - i);
      for (size_t j = 0; j + 1 < n; j += 2) apply_mcswap({qs[i + j], qs[i + j + 1]});
    }
  }

  // norm (qubitvector.hpp:1879-1886)
  double norm() const {
    double re = 0;
    const int_t END = data_size_;
#pragma omp parallel for reduction(+ : re) if (threads() > 1) num_threads(threads())
    for (int_t k = 0; k < END; ++k) re += std::real(data_[k] * std::conj(data_[k]));
    return re;
  }

  // probability (qubitvector.hpp:2090-2092)
  double probability(uint_t k) const { return std::real(data_[k] * std::conj(data_[k])); }

  // probabilities(qubits) (qubitvector.hpp:2095-2143)
  std::vector<double> probabilities(const reg_t &qs) const {
    const size_t N = qs.size();
    if (N == 0) {
      std::vector<double> p(data_size_);
      for (uint_t k = 0; k < data_size_; ++k) p[k] = probability(k);
      return p;
    }
    const uint_t DIM = BIT(N);
    const int_t END = BIT(num_qubits_ - N);
    reg_t sorted = qs;
    std::sort(sorted.begin(), sorted.end());
    if (N == num_qubits_ && qs == sorted) return probabilities({});
    std::vector<double> probs(DIM, 0.);
#pragma omp parallel if (threads() > 1) num_threads(threads())
    {
      std::vector<double> priv(DIM, 0.);
      std::vector<uint_t> idx(DIM);
#pragma omp for
      for (int_t k = 0; k < END; ++k) {
        indexes(qs, sorted, k, idx.data());
        for (uint_t m = 0; m < DIM; ++m) priv[m] += probability(idx[m]);
      }
#pragma omp critical
      for (uint_t m = 0; m < DIM; ++m) probs[m] += priv[m];
    }
    return probs;
  }

  // sample_measure (qubitvector.hpp:2149-2228)
  std::vector<uint_t> sample_measure(const std::vector<double> &rnds) const {
    const int_t END = data_size_;
    const int_t SHOTS = rnds.size();
    std::vector<uint_t> samples(SHOTS, 0);
    const int INDEX_SIZE = sample_index_size_;
    const int_t INDEX_END = BIT(INDEX_SIZE);
    if (END < INDEX_END) {
#pragma omp parallel for if (threads() > 1) num_threads(threads())
      for (int_t i = 0; i < SHOTS; ++i) {
        const double rnd = rnds[i];
        double p = .0;
        int_t sample;
        for (sample = 0; sample < END - 1; ++sample) {
          p += probability(sample);
          if (rnd < p) break;
        }
        samples[i] = sample;
      }
      return samples;
    }
    std::vector<double> idxs(INDEX_END, 0.0);
    const uint_t loop = END >> INDEX_SIZE;
#pragma omp parallel for if (threads() > 1) num_threads(threads())
    for (int_t i = 0; i < INDEX_END; ++i) {
      const uint_t base = loop * i;
      double total = .0;
      for (uint_t j = 0; j < loop; ++j) total += probability(base | j);
      idxs[i] = total;
    }
#pragma omp parallel for if (threads() > 1) num_threads(threads())
    for (int_t i = 0; i < SHOTS; ++i) {
      const double rnd = rnds[i];
      double p = .0;
      int_t sample = 0;
      size_t j = 0;
      for (; j < idxs.size(); ++j) {
        if (rnd < (p + idxs[j])) break;
        p += idxs[j];
        sample += loop;
      }
      if (j == idxs.size()) {
        samples[i] = END - 1;
        continue;
      }
      for (; sample < END - 1; ++sample) {
        p += probability(sample);
        if (rnd < p) break;
      }
      samples[i] = sample;
    }
    return samples;
  }

  // pauli_masks_and_phase + add_y_phase (qubitvector.hpp:2236-2297)
  static void pauli_masks(const reg_t &qs, const std::string &pauli, uint_t &x, uint_t &z,
                          uint_t &ny, uint_t &xmax) {
    const size_t N = qs.size();
    x = z = ny = xmax = 0;
    if (pauli.size() != N) throw std::invalid_argument("Pauli label length mismatch");
    for (size_t i = 0; i < N; ++i) {
      const uint_t b = BIT(qs[i]);
      switch (pauli[N - 1 - i]) {
        case 'I': break;
        case 'X': x += b; xmax = std::max(xmax, qs[i]); break;
        case 'Z': z += b; break;
        case 'Y': x += b; xmax = std::max(xmax, qs[i]); z += b; ny++; break;
        default: throw std::invalid_argument("Invalid Pauli");
      }
    }
  }
  template <typename T>
  static void add_y_phase(uint_t ny, std::complex<T> &c) {
    switch (ny & 3) {
      case 1: c = std::complex<T>(c.imag(), -c.real()); break;
      case 2: c = std::complex<T>(-c.real(), -c.imag()); break;
      case 3: c = std::complex<T>(-c.imag(), c.real()); break;
      default: break;
    }
  }

  // expval_pauli (qubitvector.hpp:2300-2347)
  double expval_pauli(const reg_t &qs, const std::string &pauli, cd initial_phase = 1.0) const {
    uint_t x, z, ny, xmax;
    pauli_masks(qs, pauli, x, z, ny, xmax);
    if (x + z == 0) return norm();
    cT phase = cT(initial_phase);
    add_y_phase(ny, phase);
    double re = 0;
    if (!x) {
      const int_t END = data_size_;
#pragma omp parallel for reduction(+ : re) if (threads() > 1) num_threads(threads())
      for (int_t i = 0; i < END; ++i) {
        double v = std::real(phase * data_[i] * std::conj(data_[i]));
        if (z && (__builtin_popcountll(i & z) & 1)) v = -v;
        re += v;
      }
      return re;
    }
    const uint_t mask_u = ~MASK(xmax + 1), mask_l = MASK(xmax);
    const int_t END = data_size_ >> 1;
#pragma omp parallel for reduction(+ : re) if (threads() > 1) num_threads(threads())
    for (int_t i = 0; i < END; ++i) {
      const uint_t i0 = ((i << 1) & mask_u) | (i & mask_l);
      const uint_t i1 = i0 ^ x;
      double v0 = std::real(phase * data_[i1] * std::conj(data_[i0]));
      double v1 = std::real(phase * data_[i0] * std::conj(data_[i1]));
      re += (z && (__builtin_popcountll(i0 & z) & 1)) ? -v0 : v0;
      re += (z && (__builtin_popcountll(i1 & z) & 1)) ? -v1 : v1;
    }
    return re;
  }

  // apply_pauli (qubitvector.hpp:2393-2433)
  void apply_pauli(const reg_t &qs, const std::string &pauli, cd coeff = 1.0) {
    uint_t x, z, ny, xmax;
    pauli_masks(qs, pauli, x, z, ny, xmax);
    if (x + z == 0) return;
    cT phase = cT(coeff);
    add_y_phase(ny, phase);
    if (!x) {
      const int_t END = data_size_;
#pragma omp parallel for if (threads() > 1) num_threads(threads())
      for (int_t i = 0; i < END; ++i) {
        if (z && (__builtin_popcountll(i & z) & 1)) data_[i] *= -1;
        data_[i] *= phase;
      }
      return;
    }
    const uint_t mask_u = ~MASK(xmax + 1), mask_l = MASK(xmax);
    const int_t END = data_size_ >> 1;
#pragma omp parallel for if (threads() > 1) num_threads(threads())
    for (int_t i = 0; i < END; ++i) {
      uint_t idx[2];
      idx[0] = ((i << 1) & mask_u) | (i & mask_l);
      idx[1] = idx[0] ^ x;
      std::swap(data_[idx[0]], data_[idx[1]]);
      for (int j = 0; j < 2; ++j) {
        if (z && (__builtin_popcountll(idx[j] & z) & 1)) data_[idx[j]] *= -1;
        data_[idx[j]] *= phase;
      }
    }
  }

 private:
  size_t num_qubits_ = 0;
  size_t data_size_ = 0;
  cT *data_ = nullptr;
  int omp_threads_ = 1;
  size_t omp_threshold_ = 14;
  int sample_index_size_ = 10;
};

}  // namespace oracle
