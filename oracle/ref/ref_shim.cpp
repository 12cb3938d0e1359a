// ORACLE -- test infrastructure only.
//
// C entry points over the parts of the reference that compile on their own
// (src/simulators/statevector/indexes.hpp, transformer.hpp,
// transformer_avx2.hpp, qv_avx2.cpp; src/framework/linalg/matrix_utils/
// vmatrix_defs.hpp; src/framework/rng.hpp).  Built from the reference
// checkout by oracle/ref/Makefile into oracle/_ref/; used only by tests to pin
// the restated oracle against the reference's own code.
#include <complex>
#include <cstdint>
#include <cstring>
#include <string>

#include "framework/avx2_detect.hpp"
#include "framework/linalg/matrix_utils/vmatrix_defs.hpp"
#include "framework/rng.hpp"
#include "simulators/statevector/indexes.hpp"
#include "simulators/statevector/qv_avx2.hpp"
#include "simulators/statevector/transformer.hpp"
#include "simulators/statevector/transformer_avx2.hpp"

using cd = std::complex<double>;

extern "C" {

// Transformer<double>::apply_matrix (generic) or TransformerAVX2 (avx != 0).
int ref_apply_matrix(double *data, uint64_t nq_state, const uint64_t *q, uint64_t nq,
                     const double *mat, int avx) {
  try {
    AER::reg_t qs(q, q + nq);
    AER::QV::cvector_t<double> m(1ull << (2 * nq));
    std::memcpy(m.data(), mat, m.size() * sizeof(cd));
    auto *d = reinterpret_cast<cd *>(data);
    if (avx) {
      AER::QV::TransformerAVX2<cd *, double> t;
      t.apply_matrix(d, 1ull << nq_state, 1, qs, m);
    } else {
      AER::QV::Transformer<cd *, double> t;
      t.apply_matrix(d, 1ull << nq_state, 1, qs, m);
    }
    return 0;
  } catch (...) {
    return 1;
  }
}

int ref_apply_diagonal_matrix(double *data, uint64_t nq_state, const uint64_t *q, uint64_t nq,
                              const double *diag, int avx) {
  try {
    AER::reg_t qs(q, q + nq);
    AER::QV::cvector_t<double> m(1ull << nq);
    std::memcpy(m.data(), diag, m.size() * sizeof(cd));
    auto *d = reinterpret_cast<cd *>(data);
    if (avx) {
      AER::QV::TransformerAVX2<cd *, double> t;
      t.apply_diagonal_matrix(d, 1ull << nq_state, 1, qs, m);
    } else {
      AER::QV::Transformer<cd *, double> t;
      t.apply_diagonal_matrix(d, 1ull << nq_state, 1, qs, m);
    }
    return 0;
  } catch (...) {
    return 1;
  }
}

// QubitVector's transformer as the reference selects it at run time
// (set_transformer_method, qubitvector.hpp:491-502: TransformerAVX2 when
// is_avx2_supported(), avx2_detect.hpp:54-88, else Transformer), called with
// the thread count of omp_threads_managed (qubitvector.hpp:486-489).  The
// CPU baseline times the reference's own gate kernels through these.
int ref_avx2_supported(void) { return AER::is_avx2_supported() ? 1 : 0; }

int ref_qv_apply_matrix(double *data, uint64_t nq_state, const uint64_t *q, uint64_t nq,
                        const double *mat, int threads) {
  try {
    AER::reg_t qs(q, q + nq);
    AER::QV::cvector_t<double> m(1ull << (2 * nq));
    std::memcpy(m.data(), mat, m.size() * sizeof(cd));
    auto *d = reinterpret_cast<cd *>(data);
    if (AER::is_avx2_supported()) {
      AER::QV::TransformerAVX2<cd *, double> t;
      t.apply_matrix(d, 1ull << nq_state, threads, qs, m);
    } else {
      AER::QV::Transformer<cd *, double> t;
      t.apply_matrix(d, 1ull << nq_state, threads, qs, m);
    }
    return 0;
  } catch (...) {
    return 1;
  }
}

int ref_qv_apply_diagonal_matrix(double *data, uint64_t nq_state, const uint64_t *q, uint64_t nq,
                                 const double *diag, int threads) {
  try {
    AER::reg_t qs(q, q + nq);
    AER::QV::cvector_t<double> m(1ull << nq);
    std::memcpy(m.data(), diag, m.size() * sizeof(cd));
    auto *d = reinterpret_cast<cd *>(data);
    if (AER::is_avx2_supported()) {
      AER::QV::TransformerAVX2<cd *, double> t;
      t.apply_diagonal_matrix(d, 1ull << nq_state, threads, qs, m);
    } else {
      AER::QV::Transformer<cd *, double> t;
      t.apply_diagonal_matrix(d, 1ull << nq_state, threads, qs, m);
    }
    return 0;
  } catch (...) {
    return 1;
  }
}

// indexes(qubits, sorted, k) (indexes.hpp:197-250).
void ref_indexes(const uint64_t *q, uint64_t nq, uint64_t k, uint64_t *out) {
  AER::reg_t qs(q, q + nq), sorted = qs;
  std::sort(sorted.begin(), sorted.end());
  auto r = AER::QV::indexes(qs, sorted, k);
  for (uint64_t i = 0; i < (1ull << nq); ++i) out[i] = r[i];
}

// VMatrix gate matrices (vmatrix_defs.hpp); writes 2*len doubles.
int ref_vmatrix(const char *name, const double *p, double *out) {
  using V = AER::Linalg::VMatrix;
  const std::string n(name);
  AER::cvector_t m;
  if (n == "u1") m = V::u1(p[0]);
  else if (n == "u2") m = V::u2(p[0], p[1]);
  else if (n == "u3") m = V::u3(p[0], p[1], p[2]);
  else if (n == "u4") m = V::u4(p[0], p[1], p[2], p[3]);
  else if (n == "r") m = V::r(p[0], p[1]);
  else if (n == "rx") m = V::rx(p[0]);
  else if (n == "ry") m = V::ry(p[0]);
  else if (n == "rz") m = V::rz(p[0]);
  else if (n == "rxx") m = V::rxx(p[0]);
  else if (n == "ryy") m = V::ryy(p[0]);
  else if (n == "rzz_diag") m = V::rzz_diag(p[0]);
  else if (n == "rzx") m = V::rzx(p[0]);
  else if (n == "cu") m = V::cu(p[0], p[1], p[2], p[3]);
  else if (n == "cphase") m = V::cphase(p[0]);
  else if (V::allowed_name(n)) m = V::from_name(n);
  else return -1;
  std::memcpy(out, m.data(), m.size() * sizeof(cd));
  return static_cast<int>(m.size());
}

// RngEngine(seed).rand(0, 1) draws (rng.hpp:55-70).
void ref_rng_uniform(uint64_t seed, uint64_t n, double *out) {
  AER::RngEngine rng(seed);
  for (uint64_t i = 0; i < n; ++i) out[i] = rng.rand(0, 1);
}

// RngEngine(seed).rand_int(probs) draws (rng.hpp:94-96).
void ref_rng_discrete(uint64_t seed, const double *probs, uint64_t np, uint64_t n, uint64_t *out) {
  AER::RngEngine rng(seed);
  std::vector<double> p(probs, probs + np);
  for (uint64_t i = 0; i < n; ++i) out[i] = rng.rand_int(p);
}

}  // extern "C"
