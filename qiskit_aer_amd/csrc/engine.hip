// Host implementation of the engine and its C ABI (include/aerhip.h).
#include <algorithm>
#include <array>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

#include "aerhip.h"
#include "engine.hpp"
#include "gate_matrices.hpp"
#include "host_parallel.hpp"
#include "kernels.hpp"
#include "batch.hpp"
#include "reg_pass.hpp"
#include "tile_pass.hpp"

namespace aerhip {

namespace {

constexpr uint32_t kTileBitsC128 = 11;  // 2^11 x 16 B = 32 KiB of LDS
constexpr uint32_t kTileBitsC64 = 12;   // 2^12 x 8 B  = 32 KiB of LDS
constexpr size_t kMaxQueue = 4096;      // flush threshold of the gate queue

inline uint64_t bit(uint32_t q) { return 1ull << q; }

inline uint32_t grid_for(uint64_t work, uint32_t cap = 8192) {
  uint64_t g = (work + kThreads - 1) / kThreads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return static_cast<uint32_t>(g);
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace

Engine::Engine(uint32_t num_qubits, int precision, int device)
    : n_(num_qubits), prec_(precision), device_(device) {
  if (precision != AERHIP_PRECISION_DOUBLE && precision != AERHIP_PRECISION_SINGLE)
    throw std::invalid_argument("aerhip: invalid precision");
  if (num_qubits < 1 || num_qubits > 40)
    throw std::invalid_argument("aerhip: number of qubits must be in [1, 40]");
  size_ = 1ull << n_;
  tile_bits_ = std::min<uint32_t>(n_, prec_ == AERHIP_PRECISION_DOUBLE ? kTileBitsC128 : kTileBitsC64);
  low_bits_ = std::min<uint32_t>(n_, prec_ == AERHIP_PRECISION_DOUBLE ? 3 : 4);
  reg_bits_ = prec_ == AERHIP_PRECISION_DOUBLE ? 3 : 4;
  if (const char *e = std::getenv("AERHIP_REG_BITS")) set_reg_bits(std::atoi(e));
  if (const char *e = std::getenv("AERHIP_LOW_BITS")) set_low_bits(std::atoi(e));
  AERHIP_CHECK(hipSetDevice(device_));
  const size_t bytes = size_ * (prec_ == AERHIP_PRECISION_DOUBLE ? 16 : 8);
  AERHIP_CHECK(hipMalloc(&sv_, bytes));
  AERHIP_CHECK(hipStreamCreate(&stream_));
  AERHIP_CHECK(hipEventCreateWithFlags(&prog_ev_, hipEventDisableTiming));
  AERHIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&h_scalar_), 64 * sizeof(double)));
}

Engine::Engine(uint32_t num_qubits, int precision, int device, PlanOnly)
    : n_(num_qubits), prec_(precision), device_(device), plan_only_(true) {
  if (precision != AERHIP_PRECISION_DOUBLE && precision != AERHIP_PRECISION_SINGLE)
    throw std::invalid_argument("aerhip: invalid precision");
  if (num_qubits < 1 || num_qubits > 40)
    throw std::invalid_argument("aerhip: number of qubits must be in [1, 40]");
  size_ = 1ull << n_;
  tile_bits_ = std::min<uint32_t>(n_, prec_ == AERHIP_PRECISION_DOUBLE ? kTileBitsC128 : kTileBitsC64);
  low_bits_ = std::min<uint32_t>(n_, prec_ == AERHIP_PRECISION_DOUBLE ? 3 : 4);
  reg_bits_ = prec_ == AERHIP_PRECISION_DOUBLE ? 3 : 4;  // 256 threads: the batch kernel's shape
}

Engine::~Engine() {
  if (plan_only_) return;
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (sv_) (void)hipFree(sv_);
  if (d_prog_) (void)hipFree(d_prog_);
  if (d_partial_) (void)hipFree(d_partial_);
  if (d_scratch_) (void)hipFree(d_scratch_);
  if (d_probes_) (void)hipFree(d_probes_);
  if (h_prog_) (void)hipHostFree(h_prog_);
  if (h_scalar_) (void)hipHostFree(h_scalar_);
  if (prog_ev_) (void)hipEventDestroy(prog_ev_);
  if (t0_ev_) (void)hipEventDestroy(t0_ev_);
  if (t1_ev_) (void)hipEventDestroy(t1_ev_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::set_tile_bits(uint32_t nb) {
  const uint32_t cap = prec_ == AERHIP_PRECISION_DOUBLE ? kTileBitsC128 : kTileBitsC64;
  if (nb < 1 || nb > cap) throw std::invalid_argument("aerhip: tile bits out of range");
  flush();
  tile_bits_ = std::min<uint32_t>(nb, n_);
  low_bits_ = std::min(low_bits_, tile_bits_);
  if (tile_bits_ > low_bits_ + 8) low_bits_ = tile_bits_ - 8;
}

void Engine::set_low_bits(uint32_t c) {
  // The tile kernel's offset table holds 2^(tile_bits - low_bits) <= 256 entries.
  if (c < 1 || c > 8 || c + 8 < tile_bits_)
    throw std::invalid_argument("aerhip: low bits out of range");
  flush();
  low_bits_ = std::min<uint32_t>({c, n_, tile_bits_});
}

// ---------------------------------------------------------------------------
// Lifecycle
// ---------------------------------------------------------------------------

void Engine::initialize() {
  discard_queue();
  AERHIP_CHECK(hipSetDevice(device_));
  if (prec_ == AERHIP_PRECISION_DOUBLE)
    k_fill_basis<double><<<grid_for(size_, 16384), kThreads, 0, stream_>>>(
        static_cast<cx<double> *>(sv_), size_, 0);
  else
    k_fill_basis<float><<<grid_for(size_, 16384), kThreads, 0, stream_>>>(
        static_cast<cx<float> *>(sv_), size_, 0);
  AERHIP_CHECK(hipGetLastError());
}

void Engine::zero() {
  discard_queue();
  AERHIP_CHECK(hipSetDevice(device_));
  if (prec_ == AERHIP_PRECISION_DOUBLE)
    k_fill_basis<double><<<grid_for(size_, 16384), kThreads, 0, stream_>>>(
        static_cast<cx<double> *>(sv_), size_, ~0ull);
  else
    k_fill_basis<float><<<grid_for(size_, 16384), kThreads, 0, stream_>>>(
        static_cast<cx<float> *>(sv_), size_, ~0ull);
  AERHIP_CHECK(hipGetLastError());
}

void Engine::initialize_from_vector(const cd *vec, uint64_t len) {
  if (len != size_)
    throw std::runtime_error("QubitVector::initialize input vector is incorrect length (" +
                             std::to_string(size_) + "!=" + std::to_string(len) + ")");
  discard_queue();
  AERHIP_CHECK(hipSetDevice(device_));
  if (prec_ == AERHIP_PRECISION_DOUBLE) {
    AERHIP_CHECK(hipMemcpyAsync(sv_, vec, size_ * sizeof(cd), hipMemcpyHostToDevice, stream_));
  } else {
    // Stage complex128 chunks on the device and narrow them there.
    const uint64_t chunk = std::min<uint64_t>(size_, 1ull << 24);
    void *tmp = scratch(chunk * sizeof(cd));
    for (uint64_t o = 0; o < size_; o += chunk) {
      const uint64_t m = std::min(chunk, size_ - o);
      AERHIP_CHECK(hipMemcpyAsync(tmp, vec + o, m * sizeof(cd), hipMemcpyHostToDevice, stream_));
      k_convert<float, double><<<grid_for(m, 16384), kThreads, 0, stream_>>>(
          static_cast<cx<float> *>(sv_) + o, static_cast<const cx<double> *>(tmp), m);
      AERHIP_CHECK(hipGetLastError());
    }
  }
  AERHIP_CHECK(hipStreamSynchronize(stream_));
}

void Engine::initialize_from_device(const void *src, uint64_t len, int src_precision) {
  if (len != size_)
    throw std::runtime_error("QubitVector::initialize input vector is incorrect length (" +
                             std::to_string(size_) + "!=" + std::to_string(len) + ")");
  discard_queue();
  AERHIP_CHECK(hipSetDevice(device_));
  if (src_precision == prec_) {
    const size_t bytes = size_ * (prec_ == AERHIP_PRECISION_DOUBLE ? 16 : 8);
    AERHIP_CHECK(hipMemcpyAsync(sv_, src, bytes, hipMemcpyDeviceToDevice, stream_));
  } else if (prec_ == AERHIP_PRECISION_SINGLE) {
    k_convert<float, double><<<grid_for(size_, 16384), kThreads, 0, stream_>>>(
        static_cast<cx<float> *>(sv_), static_cast<const cx<double> *>(src), size_);
  } else {
    k_convert<double, float><<<grid_for(size_, 16384), kThreads, 0, stream_>>>(
        static_cast<cx<double> *>(sv_), static_cast<const cx<float> *>(src), size_);
  }
  AERHIP_CHECK(hipGetLastError());
  AERHIP_CHECK(hipStreamSynchronize(stream_));
}

void Engine::initialize_random(uint64_t seed) {
  discard_queue();
  AERHIP_CHECK(hipSetDevice(device_));
  const uint32_t grid = grid_for(size_, 4096);
  ensure_partial(grid + 1);
  if (prec_ == AERHIP_PRECISION_DOUBLE) {
    auto *sv = static_cast<cx<double> *>(sv_);
    k_gaussian<double><<<grid_for(size_, 16384), kThreads, 0, stream_>>>(sv, size_, seed);
    k_sumsq<double><<<grid, kThreads, 0, stream_>>>(sv, size_, d_partial_);
    k_sum_rows<<<1, kThreads, 0, stream_>>>(d_partial_, 1, grid, d_partial_ + grid);
    k_scale<double><<<grid_for(size_, 16384), kThreads, 0, stream_>>>(sv, size_, d_partial_ + grid);
  } else {
    auto *sv = static_cast<cx<float> *>(sv_);
    k_gaussian<float><<<grid_for(size_, 16384), kThreads, 0, stream_>>>(sv, size_, seed);
    k_sumsq<float><<<grid, kThreads, 0, stream_>>>(sv, size_, d_partial_);
    k_sum_rows<<<1, kThreads, 0, stream_>>>(d_partial_, 1, grid, d_partial_ + grid);
    k_scale<float><<<grid_for(size_, 16384), kThreads, 0, stream_>>>(sv, size_, d_partial_ + grid);
  }
  AERHIP_CHECK(hipGetLastError());
  AERHIP_CHECK(hipStreamSynchronize(stream_));
}

void Engine::copy_to_vector(cd *out, uint64_t len) {
  if (len != size_) throw std::runtime_error("aerhip: output vector has the wrong length");
  flush();
  AERHIP_CHECK(hipSetDevice(device_));
  if (prec_ == AERHIP_PRECISION_DOUBLE) {
    AERHIP_CHECK(hipMemcpyAsync(out, sv_, size_ * sizeof(cd), hipMemcpyDeviceToHost, stream_));
  } else {
    const uint64_t chunk = std::min<uint64_t>(size_, 1ull << 24);
    void *tmp = scratch(chunk * sizeof(cd));
    for (uint64_t o = 0; o < size_; o += chunk) {
      const uint64_t m = std::min(chunk, size_ - o);
      k_convert<double, float><<<grid_for(m, 16384), kThreads, 0, stream_>>>(
          static_cast<cx<double> *>(tmp), static_cast<const cx<float> *>(sv_) + o, m);
      AERHIP_CHECK(hipGetLastError());
      AERHIP_CHECK(hipMemcpyAsync(out + o, tmp, m * sizeof(cd), hipMemcpyDeviceToHost, stream_));
    }
  }
  AERHIP_CHECK(hipStreamSynchronize(stream_));
}

void Engine::synchronize() {
  flush();
  AERHIP_CHECK(hipSetDevice(device_));
  AERHIP_CHECK(hipStreamSynchronize(stream_));
}

// ---------------------------------------------------------------------------
// Queue construction (QubitVector semantics)
// ---------------------------------------------------------------------------

void Engine::check_qubits(const std::vector<uint32_t> &qubits) const {
  uint64_t seen = 0;
  for (uint32_t q : qubits) {
    if (q >= n_)
      throw std::invalid_argument("QubitVector: qubit index " + std::to_string(q) +
                                  " > " + std::to_string(n_));
    if (seen & bit(q)) throw std::invalid_argument("QubitVector: duplicate qubit index");
    seen |= bit(q);
  }
}

uint32_t Engine::pool_add(const cd *v, size_t n) {
  const uint32_t off = static_cast<uint32_t>(pool_.size());
  pool_.insert(pool_.end(), v, v + n);
  return off;
}

void Engine::push(Prim &&p) {
  for (uint32_t q : p.q) p.touch |= bit(q);
  for (uint32_t q : p.ctrl) p.touch |= bit(q);
  p.touch |= p.xmask | p.zmask;
  // A non-unitary operator does not commute with an expectation value on
  // other qubits (it rescales them), so nothing is reordered across it.
  if (!p.unitary) p.touch = ~0ull;
  if (p.kind != K_PROBE) stats_.gates++;
  queue_.push_back(std::move(p));
  if (queue_.size() >= kMaxQueue && !plan_only_) flush();
}

namespace {

bool is_unitary(const cd *m, uint64_t D) {
  for (uint64_t i = 0; i < D; ++i)
    for (uint64_t j = 0; j < D; ++j) {
      cd s = 0;
      for (uint64_t r = 0; r < D; ++r) s += std::conj(m[r + D * i]) * m[r + D * j];
      if (std::abs(s - (i == j ? 1.0 : 0.0)) > 1e-10) return false;
    }
  return true;
}

}  // namespace

void Engine::push_dense(const std::vector<uint32_t> &targets, const std::vector<uint32_t> &ctrl,
                        uint64_t ctrl_val, const cd *mat) {
  Prim p;
  const size_t k = targets.size();
  p.kind = k == 1 ? K_DENSE1 : K_DENSE;
  p.q = targets;
  p.ctrl = ctrl;
  p.ctrl_val = ctrl_val;
  p.len = static_cast<uint32_t>(1ull << (2 * k));
  p.off = pool_add(mat, p.len);
  for (uint32_t q : targets) p.tmask |= bit(q);
  p.solo = k > 5;
  p.unitary = k <= 5 && is_unitary(mat, 1ull << k);
  push(std::move(p));
}

void Engine::push_diag(const std::vector<uint32_t> &qubits, const std::vector<uint32_t> &ctrl,
                       uint64_t ctrl_val, const cd *diag) {
  if (qubits.size() > 16)
    throw std::invalid_argument("aerhip: diagonal matrices are limited to 16 qubits");
  Prim p;
  p.kind = K_DIAG;
  p.q = qubits;
  p.ctrl = ctrl;
  p.ctrl_val = ctrl_val;
  p.len = static_cast<uint32_t>(1ull << qubits.size());
  p.off = pool_add(diag, p.len);
  for (uint32_t i = 0; i < p.len; ++i)
    if (std::abs(std::abs(diag[i]) - 1.0) > 1e-12) p.unitary = false;
  push(std::move(p));
}

// Transformer::apply_diagonal_matrix_1 (transformer.hpp:260-372): the
// identity is skipped; every other diagonal is a multiplication (the exact
// +-i and 0 cases of the reference produce the same values).
void Engine::apply_diagonal_1(uint32_t q, const cd *diag) {
  if (diag[0] == 1.0 && diag[1] == 1.0) return;
  push_diag({q}, {}, ~0ull, diag);
}

// Transformer::apply_matrix_1 (transformer.hpp:166-232).
void Engine::apply_matrix_1(uint32_t q, const std::vector<cd> &mat) {
  if (mat[1] == 0.0 && mat[2] == 0.0) {
    const cd d[2] = {mat[0], mat[3]};
    apply_diagonal_1(q, d);
    return;
  }
  if (mat[0] == 0.0 && mat[3] == 0.0 && mat[1] == 1.0 && mat[2] == 1.0) {
    Prim p;
    p.kind = K_MCX;
    p.q = {q};
    p.tmask = bit(q);
    push(std::move(p));
    return;
  }
  // The anti-diagonal and projector forms multiply by exact zeros in the
  // dense form and give the same values.
  push_dense({q}, {}, ~0ull, mat.data());
}

void Engine::apply_matrix(const std::vector<uint32_t> &qubits, const std::vector<cd> &mat) {
  check_qubits(qubits);
  const size_t k = qubits.size();
  if (k == 0) return;
  if (mat.size() != (1ull << (2 * k)))
    throw std::invalid_argument("QubitVector: matrix size does not match the number of qubits");
  if (k > 10)
    throw std::runtime_error("Maximum size of apply matrix is a 10-qubit matrix on the HIP engine.");
  if (k == 1) {
    apply_matrix_1(qubits[0], mat);
    return;
  }
  push_dense(qubits, {}, ~0ull, mat.data());
}

void Engine::apply_diagonal_matrix(const std::vector<uint32_t> &qubits,
                                   const std::vector<cd> &diag) {
  check_qubits(qubits);
  const size_t k = qubits.size();
  if (k == 0) return;
  if (diag.size() != (1ull << k))
    throw std::invalid_argument("QubitVector: diagonal size does not match the number of qubits");
  if (k == 1) {
    apply_diagonal_1(qubits[0], diag.data());
    return;
  }
  push_diag(qubits, {}, ~0ull, diag.data());
}

void Engine::apply_mcx(const std::vector<uint32_t> &qubits) {
  check_qubits(qubits);
  if (qubits.empty()) return;
  Prim p;
  p.kind = K_MCX;
  p.q = {qubits.back()};
  p.ctrl.assign(qubits.begin(), qubits.end() - 1);
  p.tmask = bit(qubits.back());
  push(std::move(p));
}

void Engine::apply_mcy(const std::vector<uint32_t> &qubits) {
  check_qubits(qubits);
  if (qubits.empty()) return;
  Prim p;
  p.kind = K_MCY;
  p.q = {qubits.back()};
  p.ctrl.assign(qubits.begin(), qubits.end() - 1);
  p.tmask = bit(qubits.back());
  push(std::move(p));
}

void Engine::apply_mcswap(const std::vector<uint32_t> &qubits) {
  check_qubits(qubits);
  if (qubits.size() < 2) throw std::invalid_argument("QubitVector: mcswap needs two qubits");
  const size_t N = qubits.size();
  Prim p;
  p.kind = K_MCSWAP;
  p.q = {qubits[N - 2], qubits[N - 1]};
  p.ctrl.assign(qubits.begin(), qubits.end() - 2);
  p.tmask = bit(qubits[N - 2]) | bit(qubits[N - 1]);
  push(std::move(p));
}

void Engine::apply_mcphase(const std::vector<uint32_t> &qubits, cd phase) {
  check_qubits(qubits);
  if (qubits.empty()) return;
  Prim p;
  p.kind = K_MCPHASE;
  p.ctrl = qubits;
  p.pr = phase.real();
  p.pi = phase.imag();
  p.unitary = std::abs(std::abs(phase) - 1.0) <= 1e-12;
  push(std::move(p));
}

// QubitVector::apply_mcu (qubitvector.hpp:1615-1720).
void Engine::apply_mcu(const std::vector<uint32_t> &qubits, const std::vector<cd> &mat) {
  check_qubits(qubits);
  const size_t N = qubits.size();
  if (N == 0) return;
  if (mat.size() != 4) throw std::invalid_argument("QubitVector: mcu needs a 2x2 matrix");
  const std::vector<uint32_t> ctrl(qubits.begin(), qubits.end() - 1);
  const uint32_t tgt = qubits.back();
  if (mat[1] == 0.0 && mat[2] == 0.0) {
    if (mat[0] == 1.0) {
      apply_mcphase(qubits, mat[3]);
      return;
    }
    const cd diag[2] = {mat[0], mat[3]};
    if (N == 1) {
      apply_diagonal_matrix(qubits, {diag[0], diag[1]});
      return;
    }
    push_diag({tgt}, ctrl, ~0ull, diag);
    return;
  }
  if (N == 1) {
    apply_matrix(qubits, mat);
    return;
  }
  push_dense({tgt}, ctrl, ~0ull, mat.data());
}

// QubitVector::apply_rotation (qubitvector.hpp:1723-1750).
void Engine::apply_rotation(const std::vector<uint32_t> &qubits, int axis, double theta) {
  switch (axis) {
    case AERHIP_ROT_X: apply_mcu(qubits, gates::rx(theta)); break;
    case AERHIP_ROT_Y: apply_mcu(qubits, gates::ry(theta)); break;
    case AERHIP_ROT_Z: apply_mcu(qubits, gates::rz(theta)); break;
    case AERHIP_ROT_XX: apply_matrix(qubits, gates::rxx(theta)); break;
    case AERHIP_ROT_YY: apply_matrix(qubits, gates::ryy(theta)); break;
    case AERHIP_ROT_ZZ: apply_diagonal_matrix(qubits, gates::rzz_diag(theta)); break;
    case AERHIP_ROT_ZX: apply_matrix(qubits, gates::rzx(theta)); break;
    default: throw std::invalid_argument("QubitVector::invalid rotation axis.");
  }
}

namespace {

struct PauliMasks {
  uint64_t x = 0, z = 0, num_y = 0;
  uint32_t x_max = 0;
};

// pauli_masks_and_phase (qubitvector.hpp:2236-2272): character N-1-i acts on
// qubits[i].
PauliMasks pauli_masks(const std::vector<uint32_t> &qubits, const std::string &pauli) {
  const size_t N = qubits.size();
  if (pauli.size() != N) throw std::invalid_argument("aerhip: Pauli label length mismatch");
  PauliMasks m;
  for (size_t i = 0; i < N; ++i) {
    const uint64_t b = 1ull << qubits[i];
    switch (pauli[N - 1 - i]) {
      case 'I': break;
      case 'X': m.x += b; m.x_max = std::max(m.x_max, qubits[i]); break;
      case 'Z': m.z += b; break;
      case 'Y': m.x += b; m.x_max = std::max(m.x_max, qubits[i]); m.z += b; m.num_y++; break;
      default:
        throw std::invalid_argument(std::string("Invalid Pauli \"") + pauli[N - 1 - i] + "\".");
    }
  }
  return m;
}

// add_y_phase (qubitvector.hpp:2275-2297): multiply by (-i)^num_y.
cd add_y_phase(uint64_t num_y, cd c) {
  switch (num_y & 3) {
    case 1: return cd(c.imag(), -c.real());
    case 2: return cd(-c.real(), -c.imag());
    case 3: return cd(-c.imag(), c.real());
    default: return c;
  }
}

}  // namespace

void Engine::apply_pauli(const std::vector<uint32_t> &qubits, const std::string &pauli, cd coeff) {
  check_qubits(qubits);
  const PauliMasks m = pauli_masks(qubits, pauli);
  if (m.x + m.z == 0) return;
  cd ph = coeff;
  if (prec_ == AERHIP_PRECISION_SINGLE)  // the phase is formed in the state precision
    ph = cd(static_cast<float>(coeff.real()), static_cast<float>(coeff.imag()));
  ph = add_y_phase(m.num_y, ph);
  const uint64_t lowmask = (1ull << low_bits_) - 1ull;
  if (static_cast<uint32_t>(__builtin_popcountll(m.x | lowmask)) > tile_bits_) {
    // X/Y on more qubits than a tile holds: the sign-and-phase part as a
    // Z-only string, then one X per qubit.  The reference takes the sign from
    // the index an amplitude moves to (qubitvector.hpp:2421-2430); applied
    // before the move it is taken from the source index i ^ x, which differs
    // by parity(x & z) = num_y mod 2, so the phase carries (-1)^num_y.
    Prim d;
    d.kind = K_PAULI;
    d.zmask = m.z;
    if (m.num_y & 1) ph = -ph;
    d.pr = ph.real();
    d.pi = ph.imag();
    d.unitary = std::abs(std::abs(ph) - 1.0) <= 1e-12;
    push(std::move(d));
    for (uint32_t q = 0; q < n_; ++q)
      if (m.x & bit(q)) apply_mcx({q});
    return;
  }
  Prim p;
  p.kind = K_PAULI;
  p.xmask = m.x;
  p.zmask = m.z;
  p.tmask = m.x;
  p.pr = ph.real();
  p.pi = ph.imag();
  p.unitary = std::abs(std::abs(ph) - 1.0) <= 1e-12;
  // pivot qubit: the highest X/Y qubit (x_max of the reference)
  p.q = {m.x_max};
  if (m.x == 0) p.q.clear();
  push(std::move(p));
}

void Engine::apply_permutation_matrix(const std::vector<uint32_t> &qubits,
                                      const std::vector<std::pair<uint64_t, uint64_t>> &pairs) {
  check_qubits(qubits);
  const size_t k = qubits.size();
  if (k == 0 || pairs.empty()) return;
  if (k > 5) {
    // Build the permutation as a dense matrix and let the dense path apply it.
    const uint64_t D = 1ull << k;
    std::vector<uint64_t> perm(D);
    for (uint64_t i = 0; i < D; ++i) perm[i] = i;
    for (const auto &pr : pairs) std::swap(perm[pr.first], perm[pr.second]);
    std::vector<cd> mat(D * D, 0.0);
    // new[i] = old[perm[i]]  ->  M(i, perm[i]) = 1
    for (uint64_t i = 0; i < D; ++i) mat[i + D * perm[i]] = 1.0;
    apply_matrix(qubits, mat);
    return;
  }
  std::vector<cd> enc(pairs.size());
  for (size_t i = 0; i < pairs.size(); ++i) {
    if (pairs[i].first >= (1ull << k) || pairs[i].second >= (1ull << k))
      throw std::invalid_argument("aerhip: permutation index out of range");
    enc[i] = cd(static_cast<double>(pairs[i].first), static_cast<double>(pairs[i].second));
  }
  Prim p;
  p.kind = K_PERM;
  p.q = qubits;
  for (uint32_t q : qubits) p.tmask |= bit(q);
  p.npairs = static_cast<uint32_t>(pairs.size());
  p.len = p.npairs;
  p.off = pool_add(enc.data(), enc.size());
  push(std::move(p));
}

// QubitVector::apply_multiplexer (qubitvector.hpp:1306-1340): for control value
// b the targets get M_b(i, j) = mat[(b*cols + i) + DIM*j].
void Engine::apply_multiplexer(const std::vector<uint32_t> &controls,
                               const std::vector<uint32_t> &targets, const std::vector<cd> &mat) {
  std::vector<uint32_t> all = targets;
  all.insert(all.end(), controls.begin(), controls.end());
  check_qubits(all);
  const uint64_t cols = 1ull << targets.size();
  const uint64_t blocks = 1ull << controls.size();
  const uint64_t DIM = cols * blocks;
  if (mat.size() != DIM * cols) throw std::invalid_argument("aerhip: multiplexer matrix size");
  if (targets.size() > 10) throw std::invalid_argument("aerhip: multiplexer targets > 10");
  std::vector<cd> mb(cols * cols);
  for (uint64_t b = 0; b < blocks; ++b) {
    for (uint64_t i = 0; i < cols; ++i)
      for (uint64_t j = 0; j < cols; ++j) mb[i + cols * j] = mat[(b * cols + i) + DIM * j];
    push_dense(targets, controls, b, mb.data());
  }
}

// QubitVector::apply_multi_swaps (qubitvector.hpp:1843-1871): in groups of ten
// qubits, swaps qubit pairs (0,1), (2,3), ... of the group.
void Engine::apply_multi_swaps(const std::vector<uint32_t> &qubits) {
  check_qubits(qubits);
  for (size_t i = 0; i < qubits.size(); i += 10) {
    const size_t n = std::min<size_t>(10, qubits.size() - i);
    for (size_t j = 0; j + 1 < n; j += 2) apply_mcswap({qubits[i + j], qubits[i + j + 1]});
  }
}

// QubitVector::initialize_component (qubitvector.hpp:879-896): new[i] =
// state[i] * old[0] within every group, i.e. the matrix whose first column is
// `state`.
void Engine::initialize_component(const std::vector<uint32_t> &qubits,
                                  const std::vector<cd> &state) {
  check_qubits(qubits);
  if (qubits.empty()) return;
  if (qubits.size() > 40) throw std::invalid_argument("aerhip: initialize_component size");
  const uint64_t D = 1ull << qubits.size();
  if (state.size() != D) throw std::invalid_argument("aerhip: initialize_component size");
  if (qubits.size() <= 5) {
    // Up to 5 qubits: queued as the dense matrix whose first column is
    // `state` (its other entries are zeros, so the products are the same).
    std::vector<cd> mat(D * D, 0.0);
    for (uint64_t i = 0; i < D; ++i) mat[i] = state[i];
    if (qubits.size() == 1)
      push_dense(qubits, {}, ~0ull, mat.data());
    else
      apply_matrix(qubits, mat);
    return;
  }
  flush();
  AERHIP_CHECK(hipSetDevice(device_));
  MarginalDesc md;
  std::memset(&md, 0, sizeof(md));
  std::vector<uint32_t> sorted = qubits;
  std::sort(sorted.begin(), sorted.end());
  md.nq = static_cast<uint32_t>(qubits.size());
  md.inner = size_ >> qubits.size();
  for (uint32_t i = 0; i < md.nq; ++i) {
    md.sorted[i] = static_cast<uint8_t>(sorted[i]);
    md.order[i] = static_cast<uint8_t>(qubits[i]);
  }
  const size_t amp = prec_ == AERHIP_PRECISION_DOUBLE ? 16 : 8;
  const size_t desc_off = 0, state_off = 256, cache_off = align_up(state_off + D * sizeof(cd), 256);
  char *sc = static_cast<char *>(scratch(cache_off + md.inner * amp));
  AERHIP_CHECK(hipMemcpyAsync(sc + desc_off, &md, sizeof(md), hipMemcpyHostToDevice, stream_));
  AERHIP_CHECK(hipMemcpyAsync(sc + state_off, state.data(), D * sizeof(cd), hipMemcpyHostToDevice,
                              stream_));
  const auto *dmd = reinterpret_cast<const MarginalDesc *>(sc + desc_off);
  const auto *dst = reinterpret_cast<const cx<double> *>(sc + state_off);
  if (prec_ == AERHIP_PRECISION_DOUBLE) {
    auto *sv = static_cast<cx<double> *>(sv_);
    auto *cache = reinterpret_cast<cx<double> *>(sc + cache_off);
    k_gather_group0<double><<<grid_for(md.inner, 16384), kThreads, 0, stream_>>>(sv, dmd, cache);
    k_init_component<double><<<grid_for(size_, 16384), kThreads, 0, stream_>>>(sv, dmd, size_, cache, dst);
  } else {
    auto *sv = static_cast<cx<float> *>(sv_);
    auto *cache = reinterpret_cast<cx<float> *>(sc + cache_off);
    k_gather_group0<float><<<grid_for(md.inner, 16384), kThreads, 0, stream_>>>(sv, dmd, cache);
    k_init_component<float><<<grid_for(size_, 16384), kThreads, 0, stream_>>>(sv, dmd, size_, cache, dst);
  }
  AERHIP_CHECK(hipGetLastError());
  AERHIP_CHECK(hipStreamSynchronize(stream_));
}

// ---------------------------------------------------------------------------
// Planner: greedy tile passes.  Primitives stay in program order except that a
// primitive may run in an earlier pass than a primitive it shares no qubit
// with.  A pass holds the low qubits plus the mixed qubits of its primitives,
// at most tile_bits_ qubits; primitives that do not fit wait for a later pass,
// and so does every later primitive that touches their qubits.
// ---------------------------------------------------------------------------

// Register program of a full tile pass (k_reg_pass).  Each TileOp is
// classified: dense 2x2 / 4x4, X and Y and probes need their mixed tile bits
// to be register bits; diagonals and phases run in any layout; everything
// else runs on the LDS tile.  Ops keep program order except that an op may
// move ahead of earlier ops it does not conflict with (disjoint tile bits, or
// both diagonal, or both read-only).  The register set of each segment is the
// R-subset of the tile bits needed by the next pending gates that lets the
// most gates run before the next layout change.
namespace {

struct RItem {
  TileOp op;
  int cls;        // 0: register gate, 1: any layout, 2: LDS only
  int cat;        // 0: general, 1: diagonal, 2: read-only
  uint32_t need;  // tile bits that must be register bits
  uint32_t touch; // tile bits read or written
};

constexpr uint32_t kProbeSentinel = 1u << 31;

RItem classify_item(const TileOp &op, uint32_t R) {
  RItem it;
  it.op = op;
  it.cls = 2;
  it.cat = 0;
  it.need = 0;
  it.touch = op.cl;
  uint32_t pos = 0;
  for (uint32_t j = 0; j < op.k && j < 16; ++j)
    if (op.pos[j] != kNotLocal) pos |= 1u << op.pos[j];
  switch (op.kind) {
    case K_DENSE1:
    case K_MCX:
    case K_MCY:
      it.cls = 0;
      it.need = 1u << op.t;
      it.touch |= it.need;
      break;
    case K_DENSE:
      it.touch |= pos;
      if (op.k == 2 && R >= 2) {
        it.cls = 0;
        it.need = pos;
      }
      break;
    case K_DIAG:
      it.cls = 1;
      it.cat = 1;
      it.touch |= pos;
      break;
    case K_MCPHASE:
      it.cls = 1;
      it.cat = 1;
      break;
    case K_PHASES:
      it.cls = 1;
      it.cat = 1;
      it.touch |= op.xl;  // tile bits of the members
      break;
    case K_PROBE:
      it.cat = 2;
      it.touch = op.xl | op.zl;
      if (static_cast<uint32_t>(__builtin_popcount(op.xl)) <= R) {
        it.cls = 0;
        it.need = op.xl;
      }
      break;
    case K_PAULI:
      it.touch |= op.xl | op.zl;
      break;
    default:
      it.touch |= pos;
      break;
  }
  // A probe never moves across a non-unitary op (a projector or a scale
  // rescales every expectation value): probes carry the sentinel bit 31 (no
  // tile bit), non-unitary ops touch every bit.
  if (op.kind == K_PROBE) it.touch |= kProbeSentinel;
  if (op.nonunitary) it.touch = ~0u;
  return it;
}

// Scans `rem` in order; an item is taken if it conflicts with no earlier item
// left behind and `ok(item)` holds.  Returns the taken indices.
template <typename F>
std::vector<size_t> scan_items(const std::vector<RItem> &rem, F ok) {
  uint32_t blocked[3] = {0, 0, 0};
  std::vector<size_t> taken;
  for (size_t i = 0; i < rem.size(); ++i) {
    const RItem &it = rem[i];
    uint32_t conflict = blocked[0];
    for (int c = 1; c < 3; ++c)
      if (c != it.cat) conflict |= blocked[c];
    if (!(it.touch & conflict) && ok(it)) {
      taken.push_back(i);
    } else {
      blocked[it.cat] |= it.touch;
    }
  }
  return taken;
}

RLayout make_layout(uint32_t rs, uint32_t B, const uint8_t *spos) {
  RLayout L;
  std::memset(&L, 0, sizeof(L));
  int r = 0, u = 0;
  for (uint32_t b = 0; b < B; ++b) {
    if (rs & (1u << b)) {
      set_byte(L.rp, r, b);
      set_byte(L.rq, r, spos[b]);
      ++r;
    } else {
      set_byte(L.tp, u, b);
      set_byte(L.tq, u, spos[b]);
      ++u;
    }
  }
  return L;
}

}  // namespace

// Engine-side diagonal fusion (the counterpart of Aer's DiagonalFusion,
// src/transpile/fusion.hpp:520-726, for this engine's tile passes).  Within
// one register pass, phase gates on one or two qubits (mcphase, i.e. p / cp /
// cz / cu1 / z / s / t) and uncontrolled single-qubit diagonals (rz, u1 ...)
// are gathered into groups whose product is one K_PHASES op:
//   - a gate on two qubits {a, b} joins the open group whose pivot is a or b
//     (its other qubit becomes a member), or opens a group pivoted on the
//     higher qubit (the target of a controlled-phase ladder);
//   - a single-qubit phase or diagonal joins the one group without a pivot.
// A member contributes u where its qubit is 0 and v where it is 1; repeated
// members multiply.  An op that is not fused closes (emits) every open group
// that shares a qubit with it, just before it: members are moved later only
// past ops on other qubits, which commute with them.  Probes close every
// group.  A group that holds a single op is emitted as that op.  `max_qubits`
// bounds the qubits of one group (pivot and members).  Member lists go to
// `pool` as (qubit, u, v) triples.
std::vector<TileOp> fuse_phase_runs(const std::vector<TileOp> &in, const uint8_t *spos, uint32_t ns,
                                    std::vector<cd> &pool, uint32_t max_qubits,
                                    uint64_t *fused_ops) {
  struct Group {
    int pivot = -1;
    uint64_t qmask = 0;
    std::vector<uint32_t> order;  // member qubits in first-seen order
    std::map<uint32_t, std::pair<cd, cd>> members;
    std::vector<TileOp> originals;
    bool nonunitary = false;
  };
  int loc[64];
  for (int &v : loc) v = -1;
  for (uint32_t b = 0; b < ns; ++b) loc[spos[b]] = static_cast<int>(b);
  auto local_qubits = [&](uint32_t lmask) {
    uint64_t g = 0;
    for (uint32_t b = 0; b < ns; ++b)
      if (lmask & (1u << b)) g |= 1ull << spos[b];
    return g;
  };
  std::vector<Group> open;
  std::vector<TileOp> out;
  out.reserve(in.size());
  auto emit = [&](Group &g) {
    if (g.originals.size() == 1) {
      out.push_back(g.originals[0]);
      return;
    }
    TileOp op;
    std::memset(&op, 0, sizeof(op));
    op.kind = K_PHASES;
    if (g.pivot >= 0) {
      const int lb = loc[g.pivot];
      if (lb >= 0) {
        op.cl = op.cvl = 1u << lb;
      } else {
        op.cg = op.cvg = 1ull << g.pivot;
      }
    }
    op.off = static_cast<uint32_t>(pool.size());
    op.npairs = static_cast<uint32_t>(g.order.size());
    for (uint32_t q : g.order) {
      const auto &uv = g.members[q];
      pool.push_back(cd(static_cast<double>(q), 0.0));
      pool.push_back(uv.first);
      pool.push_back(uv.second);
      if (loc[q] >= 0) op.xl |= 1u << loc[q];
    }
    op.nonunitary = g.nonunitary ? 1u : 0u;
    out.push_back(op);
    if (fused_ops) *fused_ops += g.originals.size();
  };
  auto close_if = [&](uint64_t touch) {
    std::vector<Group> keep;
    for (Group &g : open) {
      if (g.qmask & touch)
        emit(g);
      else
        keep.push_back(std::move(g));
    }
    open.swap(keep);
  };
  auto add = [&](int pivot, uint32_t q, cd u, cd v, const TileOp &orig) {
    const uint64_t need = (pivot >= 0 ? 1ull << pivot : 0ull) | (1ull << q);
    Group *g = nullptr;
    for (Group &c : open)
      if (c.pivot == pivot) g = &c;
    if (g && static_cast<uint32_t>(__builtin_popcountll(g->qmask | need)) > max_qubits) {
      close_if(g->qmask);
      g = nullptr;
    }
    if (!g) {
      open.emplace_back();
      g = &open.back();
      g->pivot = pivot;
    }
    g->qmask |= need;
    auto it = g->members.find(q);
    if (it == g->members.end()) {
      g->members[q] = {u, v};
      g->order.push_back(q);
    } else {
      it->second.first *= u;
      it->second.second *= v;
    }
    g->originals.push_back(orig);
    g->nonunitary = g->nonunitary || orig.nonunitary;
  };
  for (const TileOp &op : in) {
    if (op.kind == K_MCPHASE && op.cvl == op.cl && op.cvg == op.cg) {
      const uint64_t qs = local_qubits(op.cl) | op.cg;
      const int nq = __builtin_popcountll(qs);
      const cd ph(op.pr, op.pi);
      if (nq == 1) {
        add(-1, static_cast<uint32_t>(__builtin_ctzll(qs)), cd(1.0, 0.0), ph, op);
        continue;
      }
      if (nq == 2) {
        const uint32_t a = static_cast<uint32_t>(__builtin_ctzll(qs));
        const uint32_t b = 63u - static_cast<uint32_t>(__builtin_clzll(qs));
        bool has_a = false, has_b = false;
        for (const Group &g : open) {
          has_a = has_a || g.pivot == static_cast<int>(a);
          has_b = has_b || g.pivot == static_cast<int>(b);
        }
        if (has_b || !has_a)
          add(static_cast<int>(b), a, cd(1.0, 0.0), ph, op);
        else
          add(static_cast<int>(a), b, cd(1.0, 0.0), ph, op);
        continue;
      }
    }
    if (op.kind == K_DIAG && op.k == 1 && op.cl == 0 && op.cg == 0) {
      const uint32_t q = op.pos[0] != kNotLocal ? spos[op.pos[0]] : op.gpos[0];
      add(-1, q, pool[op.off], pool[op.off + 1], op);
      continue;
    }
    if (op.kind == K_PROBE) {
      close_if(~0ull);
      out.push_back(op);
      continue;
    }
    uint64_t touch = local_qubits(op.cl | op.xl | op.zl) | op.cg | op.zg;
    for (uint32_t j = 0; j < op.k && j < 16; ++j) touch |= 1ull << op.gpos[j];
    if (op.kind == K_DENSE1 || op.kind == K_MCX || op.kind == K_MCY || op.kind == K_PAULI)
      touch |= local_qubits(1u << op.t);
    close_if(touch);
    out.push_back(op);
  }
  close_if(~0ull);
  return out;
}

// Walks a register program the way k_reg_pass does and checks its shape and
// every layout, so that a malformed program fails on the host and never
// reaches the device.
// Moves a register op's pool and TileOp references by the offsets of its
// program inside a concatenated one.
void rebase_rop(ROp<double> &r, uint32_t pool_base, uint32_t op_base) {
  const uint32_t kk = r.kind & ~F_ALL;
  if (kk == Q_U2 || kk == Q_DIAGF) r.extra += pool_base;
  else if (kk == Q_DIAG || kk == Q_LDS) r.extra += op_base;
  else if ((kk == Q_LOAD || kk == Q_FILL) && r.a) r.b += pool_base;
}

void validate_register_program(const ROp<double> *p, size_t n, uint32_t B, uint32_t R,
                               const uint8_t *spos, size_t ntops) {
  auto fail = [](const char *what) { throw std::logic_error(std::string("aerhip: register program: ") + what); };
  auto check_layout = [&](const RLayout &L) {
    uint32_t seen = 0;
    for (uint32_t u = 0; u < B - R; ++u) {
      const uint32_t b = byte_at(L.tp, static_cast<int>(u));
      if (b >= B || byte_at(L.tq, static_cast<int>(u)) != spos[b]) fail("thread bit");
      seen |= 1u << b;
    }
    for (uint32_t r = 0; r < R; ++r) {
      const uint32_t b = byte_at(L.rp, static_cast<int>(r));
      if (b >= B || byte_at(L.rq, static_cast<int>(r)) != spos[b]) fail("register bit");
      seen |= 1u << b;
    }
    if (seen != (1u << B) - 1u) fail("layout is not a permutation");
  };
  size_t o = 0;
  for (;;) {
    if (o >= n) fail("missing phase");
    const ROp<double> &e = p[o++];
    if (e.kind != Q_LOAD && e.kind != Q_FILL) fail("phase must start with LOAD or FILL");
    uint32_t regmask = 0;  // tile bits of the register set
    for (uint32_t r = 0; r < R; ++r) regmask |= 1u << byte_at(e.u.lay.rp, static_cast<int>(r));
    auto check_map = [&](const ROp<double> &r) {
      if (!(r.kind & F_MAP)) return;
      uint32_t span = 0;
      for (uint32_t q = 0; q < R; ++q) {
        if (r.map_bm[q] & ~regmask) fail("relabelling outside the register bits");
        span |= r.map_bm[q];
      }
      if (r.map_t0 & ~regmask) fail("relabelling translation outside the register bits");
      (void)span;
    };
    if ((o == 1) != (e.kind == Q_LOAD)) fail("LOAD must come first, and only there");
    check_layout(e.u.lay);
    if (e.extra & 1u) fail("odd phase length");
    if (e.a > kDiagSlots) fail("fused-diagonal slots");
    if (o + e.extra >= n) fail("phase overruns the program");
    for (uint32_t i = 0; i < e.extra; ++i) {
      const ROp<double> &r = p[o++];
      check_map(r);
      switch (r.kind & ~F_ALL) {
        case Q_RELAYOUT:
          check_layout(r.u.lay);
          regmask = 0;
          for (uint32_t q = 0; q < R; ++q) regmask |= 1u << byte_at(r.u.lay.rp, static_cast<int>(q));
          break;
        case Q_U1M: case Q_U1MX:
          if (r.a == 0 || r.a >= (1u << R) || r.extra >= (1u << ((1u << R) / 2))) fail("relabelled 2x2 gate");
          break;
        case Q_U1: case Q_U1X: case Q_U1C: case Q_X: case Q_Y:
          if (r.a >= R) fail("register bit of a gate");
          break;
        case Q_U2:
          if (r.a >= R || r.b >= R || r.a == r.b) fail("register bits of a 2-qubit gate");
          break;
        case Q_D1: if (r.a >= B) fail("diagonal bit"); break;
        case Q_DIAG: if (r.extra >= ntops) fail("diagonal index"); break;
        case Q_PROBE: if (r.u.pr.xr >= (1u << R)) fail("probe mask"); break;
        case Q_PHASE: case Q_NOP: break;
        case Q_DIAGF:
          if ((r.a & 0xFFu) > 8 || (r.a & 0xFF00u) || (r.a >> 16) > e.a) fail("fused diagonal tables");
          break;
        default: fail("op kind inside a register phase");
      }
    }
    const ROp<double> &l = p[o++];
    check_map(l);
    if ((l.kind & ~F_MAP) == Q_STORE) {
      check_layout(l.u.lay);
      if (o != n) fail("ops after STORE");
      return;
    }
    if ((l.kind & ~F_MAP) != Q_SPILL) fail("phase must end with STORE or SPILL");
    while (o < n && p[o].kind == Q_LDS) {
      if (p[o].extra >= ntops) fail("LDS op index");
      ++o;
    }
  }
}

// Builds the register program of one pass.  ROps go to `rops` (matrices in
// double; converted for complex64 at upload), TileOps run on the LDS tile go to
// `tops`.  `lowb`: tile bits that must be lane bits for a full 128-byte run per
// HBM access.
void register_program(const std::vector<TileOp> &in, uint32_t B, uint32_t R, uint32_t lowb,
                      const uint8_t *spos, std::vector<cd> &pool,
                      std::vector<ROp<double>> &rops, std::vector<TileOp> &tops,
                      uint64_t *relayouts) {
  std::vector<RItem> rem;
  rem.reserve(in.size());
  for (const TileOp &op : in) rem.push_back(classify_item(op, R));
  const uint32_t lowmask = (1u << lowb) - 1u;
  auto fits = [](const RItem &it, uint32_t rs) { return it.cls == 1 || (it.cls == 0 && (it.need & ~rs) == 0); };
  auto score = [&](uint32_t rs) {
    size_t n = 0;
    for (size_t i : scan_items(rem, [&](const RItem &it) { return fits(it, rs); }))
      if (rem[i].cls == 0) ++n;
    return n;
  };
  // Fills rs up to R bits: bits that keep the low bits on lanes first, highest first.
  auto fill = [&](uint32_t rs) {
    for (int b = static_cast<int>(B) - 1; b >= 0 && static_cast<uint32_t>(__builtin_popcount(rs)) < R; --b)
      if (!(lowmask & (1u << b))) rs |= 1u << b;
    for (int b = static_cast<int>(B) - 1; b >= 0 && static_cast<uint32_t>(__builtin_popcount(rs)) < R; --b)
      rs |= 1u << b;
    return rs;
  };
  // Best register set for the pending gates; `compat` restricts to sets that
  // leave the low bits on lanes.  Returns {set, score}.
  auto best_set = [&](bool compat) {
    uint32_t cand_bits = 0;
    size_t seen = 0;
    for (size_t i : scan_items(rem, [&](const RItem &it) { return it.cls != 2; })) {
      if (rem[i].cls != 0) continue;
      const uint32_t nb = cand_bits | rem[i].need;
      if (static_cast<uint32_t>(__builtin_popcount(nb)) > R + 4) break;
      cand_bits = nb;
      if (++seen >= 24) break;
    }
    if (compat) cand_bits &= ~lowmask;
    uint32_t best = fill(0);
    size_t best_score = 0;
    int best_low = 99;
    // subsets of cand_bits with at most R bits, completed by fill()
    std::vector<uint32_t> bits;
    for (uint32_t b = 0; b < B; ++b)
      if (cand_bits & (1u << b)) bits.push_back(b);
    const uint32_t nbits = static_cast<uint32_t>(bits.size());
    for (uint32_t m = 1; m < (1u << nbits); ++m) {
      if (static_cast<uint32_t>(__builtin_popcount(m)) > R) continue;
      if (static_cast<uint32_t>(__builtin_popcount(m)) < std::min<uint32_t>(R, nbits)) continue;
      uint32_t rs = 0;
      for (uint32_t i = 0; i < nbits; ++i)
        if (m & (1u << i)) rs |= 1u << bits[i];
      rs = fill(rs);
      if (compat && (rs & lowmask)) continue;
      const size_t sc = score(rs);
      const int low = __builtin_popcount(rs & lowmask);
      if (sc > best_score || (sc == best_score && sc > 0 && low < best_low)) {
        best = rs;
        best_score = sc;
        best_low = low;
      }
    }
    return std::make_pair(best, best_score);
  };
  // A register phase is ENTER (LOAD/FILL, whose `extra` counts the register
  // ops that follow), an even number of register ops (Q_NOP pads), then
  // LEAVE (STORE/SPILL).
  size_t enter_at = 0;
  // Register relabelling (see RegIdx in reg_pass.hpp): physical register j
  // holds the logical register index lin(j) ^ cinv, lin(j) = XOR of colinv[s]
  // over the set bits s of j; logical register bit r is tile bit rp[r] of the
  // current layout.  X and CX between register qubits only change the map.
  uint32_t colinv[8], cinv = 0, cur_rs = 0;
  bool map_dirty = false;
  auto map_reset = [&] {
    for (uint32_t q = 0; q < 8; ++q) colinv[q] = 1u << q;
    cinv = 0;
    map_dirty = false;
  };
  map_reset();
  auto tile_mask = [&](uint32_t logical) {  // logical register mask -> tile bits
    uint32_t t = 0;
    uint32_t r = 0;
    for (uint32_t b = 0; b < B; ++b)
      if (cur_rs & (1u << b)) {
        if (logical & (1u << r)) t |= 1u << b;
        ++r;
      }
    return t;
  };
  auto lin = [&](uint32_t x) {
    uint32_t y = 0;
    for (uint32_t q = 0; q < R; ++q)
      if (x & (1u << q)) y ^= colinv[q];
    return y;
  };
  // Tile bits of register j under the relabelling in force (the device's
  // RegIdx::at(j)).
  auto reg_tile_bits = [&](uint32_t j) { return tile_mask(cinv ^ lin(j)); };
  auto phys_of = [&](uint32_t y) {  // the physical mask x with lin(x) == y
    for (uint32_t x = 0; x < (1u << R); ++x)
      if (lin(x) == y) return x;
    throw std::logic_error("aerhip: register relabelling is not invertible");
  };
  auto row_of = [&](uint32_t b) {  // physical bits whose logical image has bit b
    uint32_t f = 0;
    for (uint32_t q = 0; q < R; ++q)
      if (colinv[q] & (1u << b)) f |= 1u << q;
    return f;
  };
  auto identity_on = [&](uint32_t b) {
    return phys_of(1u << b) == (1u << b) && row_of(b) == (1u << b) && !((cinv >> b) & 1u);
  };
  auto stamp = [&](ROp<double> &r) {
    if (!map_dirty) return;
    r.kind |= F_MAP;
    for (uint32_t q = 0; q < R; ++q) r.map_bm[q] = static_cast<uint16_t>(tile_mask(colinv[q]));
    r.map_t0 = static_cast<uint16_t>(tile_mask(cinv));
    map_dirty = false;
  };
  // Fused diagonals of the phase that ends: slot k for the k-th (up to
  // kDiagSlots), lane record 8k + c for its table c and staging distance
  // 32k + e for entry e of its lane tables, after an entry equal to 1 (the
  // table of unused lanes); distances are counted back from the first record,
  // so that moving the pool moves nothing else.
  auto assign_diag_slots = [&] {
    const uint32_t NR = 1u << R;
    std::vector<ROp<double> *> df;
    for (size_t i = enter_at + 1; i < rops.size() && df.size() < kDiagSlots; ++i)
      if ((rops[i].kind & ~F_ALL) == Q_DIAGF) df.push_back(&rops[i]);
    if (df.empty()) return;
    pool.push_back(cd(1.0, 0.0));
    const size_t first = pool.size();
    if (first + 96 > 0xFFFFFFFFull) throw std::length_error("aerhip: pool offset");
    DiagLane lanes[64];
    for (uint32_t l = 0; l < 64; ++l) {
      lanes[l].back = 1;
      lanes[l].fq = 0xFFFFFFFFu;
      const uint32_t k = l / 8, c = l % 8;
      if (k >= df.size() || c >= (df[k]->a & 0xFFu)) continue;
      lanes[l].back = static_cast<uint32_t>(first - (df[k]->extra + NR + 32 + 16 * c));
      lanes[l].fq = 0;
      for (uint32_t w = 0; w < 4; ++w) lanes[l].fq |= static_cast<uint32_t>(df[k]->u.fq[4 * c + w]) << (8 * w);
    }
    uint32_t stage[32 * kDiagSlots];
    for (uint32_t i = 0; i < 32 * kDiagSlots; ++i) {
      const uint32_t k = i / 32, e = i % 32;
      stage[i] = k < df.size() ? static_cast<uint32_t>(first - (df[k]->extra + NR + e)) : 1u;
    }
    cd w[32 + 64];
    std::memcpy(w, lanes, sizeof(lanes));
    std::memcpy(w + 32, stage, sizeof(stage));
    pool.insert(pool.end(), w, w + 96);
    for (size_t k = 0; k < df.size(); ++k) df[k]->a = (df[k]->a & 0xFFu) | static_cast<uint32_t>(k + 1) << 16;
    ROp<double> &e = rops[enter_at];
    e.a = static_cast<uint32_t>(df.size());
    e.b = static_cast<uint32_t>(first);
  };
  auto emit_raw = [&](uint32_t kind, uint32_t rs) {
    if (kind == Q_STORE || kind == Q_SPILL) {
      uint32_t cnt = static_cast<uint32_t>(rops.size() - enter_at - 1);
      if (cnt & 1u) {
        ROp<double> nop;
        std::memset(&nop, 0, sizeof(nop));
        nop.kind = Q_NOP;
        rops.push_back(nop);
        ++cnt;
      }
      rops[enter_at].extra = cnt;
      assign_diag_slots();
    }
    if (kind == Q_LOAD || kind == Q_FILL) enter_at = rops.size();
    ROp<double> r;
    std::memset(&r, 0, sizeof(r));
    r.kind = kind;
    r.u.lay = make_layout(rs, B, spos);
    if (kind == Q_STORE || kind == Q_SPILL || kind == Q_RELAYOUT) stamp(r);
    rops.push_back(r);
    if (kind == Q_LOAD || kind == Q_FILL || kind == Q_RELAYOUT) {
      map_reset();
      cur_rs = rs;
    }
  };
  auto regidx = [](uint32_t rs, uint32_t b) {
    return static_cast<uint32_t>(__builtin_popcount(rs & ((1u << b) - 1u)));
  };
  // Logical register bits of the register tile bits in `tmask`.
  auto reg_bits = [&](uint32_t tmask) {
    uint32_t r = 0;
    for (uint32_t b = 0; b < B; ++b)
      if ((tmask & cur_rs & (1u << b))) r |= 1u << regidx(cur_rs, b);
    return r;
  };
  // X gates controlled only by qubits fixed by the tile, waiting on logical
  // register bits (pend) to fold into the next 2x2 gate on that bit (Q_U1X,
  // Q_U1MX);
  // anything else that touches the bit, and every layout change, first
  // emits them as Q_X.
  uint32_t pend = 0;
  uint64_t pcg[8] = {}, pcvg[8] = {};
  auto materialize = [&](uint32_t bits) {
    bits &= pend;
    if (!bits) return;
    bool fresh = true;
    for (uint32_t b = 0; b < R; ++b)
      if ((bits >> b) & 1u) fresh = fresh && identity_on(b);
    if (!fresh) {
      emit_raw(Q_RELAYOUT, cur_rs);  // Q_X pairs the layout's own register bits
      if (relayouts) ++*relayouts;
    }
    for (uint32_t b = 0; b < R; ++b) {
      if (!((bits >> b) & 1u)) continue;
      ROp<double> x;
      std::memset(&x, 0, sizeof(x));
      x.kind = Q_X;
      x.a = b;
      x.cg = pcg[b];
      x.cvg = pcvg[b];
      stamp(x);
      rops.push_back(x);
    }
    pend &= ~bits;
  };
  auto emit = [&](uint32_t kind, uint32_t rs) {
    if (kind == Q_RELAYOUT || kind == Q_STORE || kind == Q_SPILL) materialize(pend);
    emit_raw(kind, rs);
  };
  // Ops that need the registers in the layout's own order (pair structure on
  // their register bits) start from a fresh layout when relabelled.
  auto need_identity = [&](uint32_t rs, std::initializer_list<uint32_t> bits) {
    bool ok = true;
    for (uint32_t b : bits) ok = ok && identity_on(b);
    if (ok) return;
    emit(Q_RELAYOUT, rs);
    if (relayouts) ++*relayouts;
  };
  auto emit_gate = [&](const TileOp &op, uint32_t rs) {
    ROp<double> r;
    std::memset(&r, 0, sizeof(r));
    r.cl = op.cl;
    r.cvl = op.cvl;
    r.cg = op.cg;
    r.cvg = op.cvg;
    switch (op.kind) {
      case K_DENSE1: {
        const uint32_t b = regidx(rs, op.t);
        for (int e = 0; e < 4; ++e) {
          r.u.m[2 * e] = pool[op.off + e].real();
          r.u.m[2 * e + 1] = pool[op.off + e].imag();
        }
        materialize(reg_bits(op.cl));
        bool fold = false;
        if (!op.cl && !op.cg && ((pend >> b) & 1u)) {
          // fold the waiting X: U X where its tile predicate holds
          fold = true;
          r.cg = pcg[b];
          r.cvg = pcvg[b];
          pend &= ~(1u << b);
        }
        if (op.cl || op.cg) {
          materialize(1u << b);
          need_identity(rs, {b});
          r.kind = Q_U1C;
          r.a = b;
        } else if (identity_on(b)) {
          r.kind = fold ? Q_U1X : Q_U1;
          r.a = b;
        } else {
          // relabelled qubit: pairs {j, j ^ m}, roles from the map
          const uint32_t m = phys_of(1u << b), f = row_of(b), cb = (cinv >> b) & 1u;
          const uint32_t hb = 31u - static_cast<uint32_t>(__builtin_clz(m));
          uint32_t roles = 0, k = 0;
          for (uint32_t j = 0; j < (1u << R); ++j) {
            if (j & (1u << hb)) continue;
            if ((static_cast<uint32_t>(__builtin_popcount(j & f)) & 1u) ^ cb) roles |= 1u << k;
            ++k;
          }
          r.kind = fold ? Q_U1MX : Q_U1M;
          r.a = m;
          r.extra = roles;
        }
        break;
      }
      case K_MCX:
      case K_MCY: {
        const uint32_t t = regidx(rs, op.t);
        if (op.kind == K_MCX && op.cl == 0 && op.cg != 0) {
          // controls all fixed by the tile: wait for the next 2x2 on t
          materialize(1u << t);
          pend |= 1u << t;
          pcg[t] = op.cg;
          pcvg[t] = op.cvg;
          return;
        }
        materialize((1u << t) | reg_bits(op.cl));
        if (op.kind == K_MCX && op.cg == 0 && (op.cl & ~rs) == 0 && __builtin_popcount(op.cl) <= 1) {
          // X with at most one register control: relabel the registers
          if (op.cl == 0) {
            cinv ^= 1u << t;
          } else {
            const uint32_t c = regidx(rs, static_cast<uint32_t>(__builtin_ctz(op.cl)));
            const bool v = (op.cvl & op.cl) != 0;
            auto cxm = [&](uint32_t y) { return y ^ (((y >> c) & 1u) << t); };
            for (uint32_t q = 0; q < R; ++q) colinv[q] = cxm(colinv[q]);
            cinv = cxm(cinv) ^ (v ? 0u : (1u << t));
          }
          map_dirty = true;
          return;
        }
        need_identity(rs, {t});
        r.kind = op.kind == K_MCX ? Q_X : Q_Y;
        r.a = t;
        break;
      }
      case K_DENSE: {
        const uint32_t b0 = regidx(rs, op.pos[0]), b1 = regidx(rs, op.pos[1]);
        materialize((1u << b0) | (1u << b1) | reg_bits(op.cl));
        need_identity(rs, {b0, b1});
        r.kind = Q_U2;
        r.a = b0;
        r.b = b1;
        r.extra = op.off;
        break;
      }
      case K_DIAG:
        materialize(pend);
        if (op.k == 1 && op.pos[0] != kNotLocal) {
          r.kind = Q_D1;
          r.a = op.pos[0];
          r.b = 0;  // registers whose target bit is 1 (a register target)
          for (uint32_t j = 0; j < (1u << R); ++j)
            if ((reg_tile_bits(j) >> op.pos[0]) & 1u) r.b |= 1u << j;
          for (int e = 0; e < 2; ++e) {
            r.u.m[2 * e] = pool[op.off + e].real();
            r.u.m[2 * e + 1] = pool[op.off + e].imag();
          }
        } else {
          r.kind = Q_DIAG;
          r.extra = static_cast<uint32_t>(tops.size());
          tops.push_back(op);
        }
        break;
      case K_MCPHASE:
        materialize(reg_bits(op.cl));
        r.kind = Q_PHASE;
        r.b = 0;  // registers that meet the register part of the predicate
        for (uint32_t j = 0; j < (1u << R); ++j)
          if ((reg_tile_bits(j) & op.cl & cur_rs) == (op.cvl & op.cl & cur_rs)) r.b |= 1u << j;
        r.u.m[0] = op.pr;
        r.u.m[1] = op.pi;
        break;
      case K_PHASES: {
        // Tables of the fused diagonal for the layout in force: register
        // factors R_j, lane factors over thread bits 0-3 and 4-7, and
        // 16-entry tables over groups of four fixed qubits.
        materialize(reg_bits(op.cl | op.xl));
        r.kind = Q_DIAGF;
        r.b = 0;
        for (uint32_t j = 0; j < (1u << R); ++j)
          if ((reg_tile_bits(j) & op.cl & cur_rs) == (op.cvl & op.cl & cur_rs)) r.b |= 1u << j;
        int loc[64];
        for (int &v : loc) v = -1;
        for (uint32_t b = 0; b < B; ++b) loc[spos[b]] = static_cast<int>(b);
        uint32_t tp[16], nt = 0;  // tile bit of thread bit u
        for (uint32_t b = 0; b < B; ++b)
          if (!(cur_rs & (1u << b))) tp[nt++] = b;
        const uint32_t NR = 1u << R;
        std::vector<cd> blk(NR + 32, cd(1.0, 0.0));
        std::vector<std::pair<uint32_t, std::pair<cd, cd>>> fixed;
        for (uint32_t m = 0; m < op.npairs; ++m) {
          const uint32_t q = static_cast<uint32_t>(pool[op.off + 3 * m].real());
          const cd u = pool[op.off + 3 * m + 1], v = pool[op.off + 3 * m + 2];
          const int lb = loc[q];
          if (lb < 0) {
            fixed.push_back({q, {u, v}});
          } else if (cur_rs & (1u << lb)) {
            for (uint32_t j = 0; j < NR; ++j) blk[j] *= ((reg_tile_bits(j) >> lb) & 1u) ? v : u;
          } else {
            uint32_t tu = 0;
            while (tp[tu] != static_cast<uint32_t>(lb)) ++tu;
            cd *lt = blk.data() + NR + (tu < 4 ? 0 : 16);
            const uint32_t sh = tu & 3u;
            for (uint32_t i = 0; i < 16; ++i) lt[i] *= ((i >> sh) & 1u) ? v : u;
          }
        }
        const uint32_t nchunk = static_cast<uint32_t>((fixed.size() + 3) / 4);
        if (nchunk > 8) throw std::logic_error("aerhip: fused diagonal with too many fixed qubits");
        std::memset(r.u.fq, 0xFF, sizeof(r.u.fq));
        for (uint32_t c = 0; c < nchunk; ++c) {
          cd tab[16];
          for (cd &t : tab) t = cd(1.0, 0.0);
          for (uint32_t w = 0; w < 4 && 4 * c + w < fixed.size(); ++w) {
            const auto &fm = fixed[4 * c + w];
            r.u.fq[4 * c + w] = static_cast<uint8_t>(fm.first);
            for (uint32_t i = 0; i < 16; ++i) tab[i] *= ((i >> w) & 1u) ? fm.second.second : fm.second.first;
          }
          blk.insert(blk.end(), tab, tab + 16);
        }
        r.a = nchunk;
        r.extra = static_cast<uint32_t>(pool.size());
        pool.insert(pool.end(), blk.begin(), blk.end());
        break;
      }
      case K_PROBE: {
        materialize(pend);
        r.kind = Q_PROBE;
        r.extra = op.npairs;
        uint32_t xl = 0;
        for (uint32_t b = 0; b < B; ++b)
          if (op.xl & (1u << b)) xl |= 1u << regidx(rs, b);
        r.u.pr.xr = phys_of(xl);
        r.u.pr.zl = op.zl;
        r.u.pr.zg = op.zg;
        r.u.pr.pr = op.pr;
        r.u.pr.pi = op.pi;
        break;
      }
      default:
        throw std::logic_error("aerhip: op cannot run in registers");
    }
    stamp(r);
    rops.push_back(r);
  };
  auto emit_lds = [&](const TileOp &op) {
    ROp<double> r;
    std::memset(&r, 0, sizeof(r));
    r.kind = Q_LDS;
    r.extra = static_cast<uint32_t>(tops.size());
    tops.push_back(op);
    rops.push_back(r);
  };
  auto take = [&](const std::vector<size_t> &idx, uint32_t rs, bool lds_mode) {
    for (size_t i : idx) {
      if (lds_mode) emit_lds(rem[i].op);
      else emit_gate(rem[i].op, rs);
    }
    std::vector<RItem> keep;
    keep.reserve(rem.size() - idx.size());
    size_t k = 0;
    for (size_t i = 0; i < rem.size(); ++i) {
      if (k < idx.size() && idx[k] == i) {
        ++k;
        continue;
      }
      keep.push_back(rem[i]);
    }
    rem.swap(keep);
  };

  // Load in the best set that keeps the low bits on lanes.
  uint32_t rs = best_set(true).first;
  emit(Q_LOAD, rs);
  bool lds_mode = false;
  while (!rem.empty()) {
    if (!lds_mode) {
      const auto idx = scan_items(rem, [&](const RItem &it) { return fits(it, rs); });
      if (!idx.empty()) {
        take(idx, rs, false);
        continue;
      }
      const auto b = best_set(false);
      if (b.second > 0) {
        rs = b.first;
        emit(Q_RELAYOUT, rs);
        if (relayouts) ++*relayouts;
        continue;
      }
      emit(Q_SPILL, rs);
      lds_mode = true;
    }
    // LDS mode: run every ready op that cannot run in registers (and ready
    // diagonals; fused diagonals run only in registers), then return to
    // registers if register gates are pending.
    const auto idx = scan_items(rem, [&](const RItem &it) {
      return it.cls == 2 || (it.cls == 1 && it.op.kind != K_PHASES);
    });
    if (!idx.empty()) {
      take(idx, rs, true);
      continue;
    }
    rs = best_set(false).first;
    emit(Q_FILL, rs);
    if (relayouts) ++*relayouts;
    lds_mode = false;
  }
  if (lds_mode || (rs & lowmask)) {
    rs = fill(0);
    emit(lds_mode ? Q_FILL : Q_RELAYOUT, rs);
    if (relayouts) ++*relayouts;
  }
  emit(Q_STORE, rs);
}

std::vector<PassPlan> Engine::plan() const {
  std::vector<PassPlan> out;
  const uint32_t nb = tile_bits_;
  const uint64_t lowmask = (1ull << low_bits_) - 1ull;
  size_t seg_begin = 0;
  for (const Prim &p : queue_)
    if (!p.solo && static_cast<uint32_t>(__builtin_popcountll(p.tmask | lowmask)) > nb)
      throw std::logic_error("aerhip: a primitive mixes more qubits than a tile holds");
  while (seg_begin < queue_.size()) {
    size_t seg_end = seg_begin;
    while (seg_end < queue_.size() && !queue_[seg_end].solo) ++seg_end;
    std::vector<uint32_t> remaining;
    for (size_t i = seg_begin; i < seg_end; ++i) remaining.push_back(static_cast<uint32_t>(i));
    while (!remaining.empty()) {
      // One candidate pass: the tile starts as `s0`, then takes the
      // primitives in program order while their mixed qubits fit.  With
      // `take` null it only counts, and stops once every qubit the remaining
      // primitives touch is blocked.
      uint64_t active = 0;
      for (uint32_t idx : remaining) active |= queue_[idx].touch;
      auto scan = [&](uint64_t s0, PassPlan *take, std::vector<uint32_t> *rest) {
        uint64_t s = s0;
        uint64_t blocked = 0;
        bool nonunitary_blocked = false;  // a deferred op rescales amplitudes
        size_t count = 0;
        for (size_t r = 0; r < remaining.size(); ++r) {
          const uint32_t idx = remaining[r];
          const Prim &p = queue_[idx];
          bool fits = !(p.touch & blocked) && !(p.kind == K_PROBE && nonunitary_blocked);
          if (fits) {
            const uint64_t ns = s | p.tmask;
            fits = static_cast<uint32_t>(__builtin_popcountll(ns)) <= nb;
            if (fits) s = ns;
          }
          if (fits) {
            ++count;
            if (take) take->prims.push_back(idx);
            continue;
          }
          if (rest) rest->push_back(idx);
          blocked |= p.touch;
          nonunitary_blocked |= !p.unitary;
          if (!take && (blocked & active) == active) break;
        }
        if (take) take->smask = s;
        return count;
      };
      // Candidates: the tile seeded by nothing but the low qubits, and by
      // each window of tile_bits - low_bits consecutive qubits.  A window
      // lets the pass run a whole light cone of a nearest-neighbour circuit
      // (the 30-qubit U3+CX brickwork takes 29 passes instead of 66); the
      // pass keeps the candidate that takes the most primitives, the first
      // on ties.
      uint64_t seed = lowmask;
      size_t best = scan(lowmask, nullptr, nullptr);
      const uint32_t w = nb - low_bits_;
      if (n_ > nb && w > 0)
        for (uint32_t a = 0; a + w <= n_; ++a) {
          const uint64_t s0 = lowmask | (((1ull << w) - 1ull) << a);
          const size_t c = scan(s0, nullptr, nullptr);
          if (c > best) {
            best = c;
            seed = s0;
          }
        }
      PassPlan pp;
      pp.solo = false;
      std::vector<uint32_t> rest;
      scan(seed, &pp, &rest);
      uint64_t s = pp.smask;
      // Complete the tile with the lowest free qubits (longer contiguous runs).
      for (uint32_t q = 0; q < n_ && static_cast<uint32_t>(__builtin_popcountll(s)) < nb; ++q)
        s |= bit(q);
      pp.smask = s;
      out.push_back(std::move(pp));
      remaining.swap(rest);
    }
    if (seg_end < queue_.size()) {
      PassPlan pp;
      pp.solo = true;
      pp.smask = queue_[seg_end].tmask;
      pp.prims.push_back(static_cast<uint32_t>(seg_end));
      out.push_back(std::move(pp));
      ++seg_end;
    }
    seg_begin = seg_end;
  }
  return out;
}

// Turns the queue into passes: PassDescs, TileOps (LDS passes and the LDS
// primitives of register passes), ROps (register passes) and the coefficient
// pool, appended to `prog` (pool offsets are relative to prog.pool).  Consumes
// the queue.
// ---------------------------------------------------------------------------
// Engine-side block fusion: runs of unitary gates whose qubits (targets and
// controls) stay within `width` qubits are multiplied into one dense matrix
// (FP64, on the host), which the tile passes apply as a GEMM on MFMA (k = 3..5
// qubits) or as a register 4x4 / 2x2 gate.  This is the aggregation of Aer's
// fusion (src/transpile/fusion.hpp:849-966; the fused matrix is the product
// of the members' matrices, as UnitaryFusion builds it, :159-174), with a
// cost model of this engine instead of Aer's CPU one (:1105-1136): a block is
// fused only when its dense form costs less than its members.
// ---------------------------------------------------------------------------
namespace {

// Qubits a primitive acts on (targets and controls) and whether block
// fusion can take it.
bool block_fusable(const Prim &p, uint32_t width) {
  if (!p.unitary || p.solo) return false;
  switch (p.kind) {
    case K_DENSE1: case K_DENSE: case K_DIAG: case K_MCX: case K_MCY: case K_MCSWAP: case K_MCPHASE:
      break;
    default:
      return false;
  }
  return static_cast<uint32_t>(p.q.size() + p.ctrl.size()) <= width;
}

uint64_t prim_qubits(const Prim &p) {
  uint64_t m = 0;
  for (uint32_t q : p.q) m |= 1ull << q;
  for (uint32_t q : p.ctrl) m |= 1ull << q;
  return m;
}

// Applies primitive p to the 2^w-amplitude vector v over the block qubits
// `bq` (bit i of an index = qubit bq[i]).
void apply_small(const Prim &p, const std::vector<uint32_t> &bq, const std::vector<cd> &pool, cd *v) {
  const uint32_t w = static_cast<uint32_t>(bq.size());
  auto loc = [&](uint32_t q) {
    for (uint32_t i = 0; i < w; ++i)
      if (bq[i] == q) return i;
    throw std::logic_error("aerhip: block fusion: qubit outside the block");
  };
  uint32_t cm = 0, cv = 0;  // control mask / required value (local bits)
  for (size_t i = 0; i < p.ctrl.size(); ++i) {
    const uint32_t b = 1u << loc(p.ctrl[i]);
    cm |= b;
    if ((p.ctrl_val >> i) & 1ull) cv |= b;
  }
  const uint32_t S = 1u << w;
  switch (p.kind) {
    case K_MCPHASE: {
      const cd ph(p.pr, p.pi);
      for (uint32_t x = 0; x < S; ++x)
        if ((x & cm) == cm) v[x] *= ph;
      return;
    }
    case K_MCX: case K_MCY: {
      const uint32_t t = 1u << loc(p.q[0]);
      for (uint32_t x = 0; x < S; ++x) {
        if ((x & t) || (x & cm) != cv) continue;
        const cd a0 = v[x], a1 = v[x | t];
        if (p.kind == K_MCX) {
          v[x] = a1;
          v[x | t] = a0;
        } else {
          v[x] = cd(0, -1) * a1;
          v[x | t] = cd(0, 1) * a0;
        }
      }
      return;
    }
    case K_MCSWAP: {
      const uint32_t a = 1u << loc(p.q[0]), b = 1u << loc(p.q[1]);
      for (uint32_t x = 0; x < S; ++x)
        if ((x & a) && !(x & b) && (x & cm) == cv) std::swap(v[x], v[(x & ~a) | b]);
      return;
    }
    case K_DIAG: {
      const uint32_t k = static_cast<uint32_t>(p.q.size());
      std::vector<uint32_t> tb(k);
      for (uint32_t j = 0; j < k; ++j) tb[j] = loc(p.q[j]);
      for (uint32_t x = 0; x < S; ++x) {
        if ((x & cm) != cv) continue;
        uint32_t i = 0;
        for (uint32_t j = 0; j < k; ++j) i |= ((x >> tb[j]) & 1u) << j;
        v[x] *= pool[p.off + i];
      }
      return;
    }
    default: {  // K_DENSE1 / K_DENSE: column-major matrix, bit j of its index = q[j]
      const uint32_t k = static_cast<uint32_t>(p.q.size()), D = 1u << k;
      uint32_t tm = 0;
      std::vector<uint32_t> tb(k);
      for (uint32_t j = 0; j < k; ++j) {
        tb[j] = loc(p.q[j]);
        tm |= 1u << tb[j];
      }
      std::vector<cd> in(D);
      for (uint32_t x = 0; x < S; ++x) {
        if ((x & tm) || (x & cm) != cv) continue;
        auto at = [&](uint32_t i) {
          uint32_t y = x;
          for (uint32_t j = 0; j < k; ++j) y |= ((i >> j) & 1u) << tb[j];
          return y;
        };
        for (uint32_t i = 0; i < D; ++i) in[i] = v[at(i)];
        for (uint32_t r = 0; r < D; ++r) {
          cd acc = 0;
          for (uint32_t c = 0; c < D; ++c) acc += pool[p.off + r + D * c] * in[c];
          v[at(r)] = acc;
        }
      }
      return;
    }
  }
}

// Relative cost per amplitude of running a primitive (or a fused block) in a
// tile pass: arithmetic in units of one FP64 FMA (complex64 FMAs on MFMA cost
// half), plus the fixed cost of an op and of moving the tile to LDS for a
// dense matrix on 3 or more qubits.  Measured op times on 28 qubits (c128 /
// c64 us): U3 60 / 42, 4x4 167 / 108, CZ 77 / 51, layout change 95 / 62.
double block_cost(uint32_t kind, uint32_t k, bool single) {
  const double op = 4.0;                      // dispatch and register bookkeeping
  switch (kind) {
    case K_MCX: case K_MCY: case K_MCSWAP: return 1.0;
    case K_MCPHASE: case K_DIAG: return op + 4.0;
    // A 2x2 gate is priced at half its complex128 FMA count for both
    // precisions: pricing it at 8 leaves runs of two 2x2 gates and a CX on a
    // pair to be fused into 4x4 gates, which makes the planner fit fewer
    // gates per tile pass (the 30-qubit brickwork: 77 passes and 728 ms at
    // width 2 against 51 passes and 500 ms when only larger runs are fused).
    case K_DENSE1: return op + 4.0;
    default: break;
  }
  const double fma = 4.0 * static_cast<double>(1u << k);  // complex MACs x 4 FMAs per amplitude
  if (k <= 2) return op + fma;                            // register 4x4 on the vector ALU
  return op + 12.0 + (single ? fma / 2.0 : fma);          // LDS round trip + MFMA GEMM
}

}  // namespace

// A run of SWAPs without controls on disjoint qubit pairs (a layer, such as
// the final reversal of a QFT) whose qubits outside the low ones are more than
// one tile pass can hold becomes one K_QPERM primitive: a single pass that
// permutes the index bits, where the SWAPs one by one would take a tile pass
// per few pairs.  At most 16 pairs per layer (the TileOp's pair list).
std::vector<Prim> Engine::merge_swap_layers(const std::vector<Prim> &in) const {
  const uint64_t lowmask = (1ull << low_bits_) - 1ull;
  // the low qubits and their partners must fit the permutation tile
  if (tile_bits_ < 2 * low_bits_ + 1) return in;
  std::vector<Prim> out;
  out.reserve(in.size());
  size_t i = 0;
  while (i < in.size()) {
    size_t j = i;
    uint64_t used = 0;
    while (j < in.size() && j - i < 16 && in[j].kind == K_MCSWAP && in[j].ctrl.empty() &&
           !(in[j].tmask & used)) {
      used |= in[j].tmask;
      ++j;
    }
    if (j - i >= 2 &&
        static_cast<uint32_t>(__builtin_popcountll(used & ~lowmask)) > tile_bits_ - low_bits_) {
      Prim p;
      p.kind = K_QPERM;
      p.solo = true;
      p.tmask = p.touch = used;
      for (size_t k = i; k < j; ++k) {
        p.q.push_back(in[k].q[0]);
        p.q.push_back(in[k].q[1]);
      }
      out.push_back(std::move(p));
      i = j;
    } else {
      out.push_back(in[i]);
      ++i;
    }
  }
  return out;
}

// The pass of a K_QPERM primitive: tile qubits S closed under the pair
// exchange, holding the low qubits and their partners (so that both tiles of
// a pair are read and written in runs of 2^low_bits amplitudes), then the
// lowest other qubits with their partners, up to tile_bits - 1 qubits (two
// tiles share the workgroup's LDS).
PassDesc Engine::qperm_pass(const Prim &p, std::vector<TileOp> &ops) const {
  uint32_t partner[64];
  for (uint32_t q = 0; q < 64; ++q) partner[q] = q;
  for (size_t k = 0; k + 1 < p.q.size(); k += 2) {
    partner[p.q[k]] = p.q[k + 1];
    partner[p.q[k + 1]] = p.q[k];
  }
  const uint32_t sb = tile_bits_ - 1;
  uint64_t S = 0;
  for (uint32_t q = 0; q < low_bits_; ++q) S |= bit(q) | bit(partner[q]);
  for (uint32_t q = low_bits_; q < n_; ++q) {
    const uint64_t orbit = bit(q) | bit(partner[q]);
    if (!(S & orbit) && static_cast<uint32_t>(__builtin_popcountll(S | orbit)) <= sb) S |= orbit;
  }
  if (static_cast<uint32_t>(__builtin_popcountll(S)) > sb) throw std::logic_error("aerhip: permutation tile");
  PassDesc pd;
  std::memset(&pd, 0, sizeof(pd));
  int loc[64];
  for (int &v : loc) v = -1;
  uint32_t ns = 0, nc = 0;
  for (uint32_t q = 0; q < n_; ++q) {
    if (S & bit(q)) {
      loc[q] = static_cast<int>(ns);
      pd.spos[ns++] = static_cast<uint8_t>(q);
    } else {
      pd.comp[nc++] = static_cast<uint8_t>(q);
    }
  }
  pd.kind = 4;
  pd.nb = ns;
  pd.ncomp = nc;
  while (pd.c < ns && pd.spos[pd.c] == pd.c) ++pd.c;
  pd.op_begin = static_cast<uint32_t>(ops.size());
  pd.nops = 1;
  TileOp op;
  std::memset(&op, 0, sizeof(op));
  op.kind = K_QPERM;
  for (uint32_t i = 0; i < ns; ++i)
    op.cg |= static_cast<uint64_t>(loc[partner[pd.spos[i]]]) << (4 * i);
  // LDS order: local index k at k ^ g(k), g(k) bit j = bit P_j of k where the
  // lane bit j of the reads lands (P_j = image of local bit j) if that is
  // above the bank slot bits (16-byte slots of c128: 4 bits, 8-byte of c64:
  // 5), so that the permuted reads, like the writes, meet distinct banks.
  const uint32_t slot_bits = prec_ == AERHIP_PRECISION_DOUBLE ? 4 : 5;
  op.xl = op.zl = 0xFFFFFFFFu;
  for (uint32_t j = 0; j < slot_bits && j < ns; ++j) {
    const uint32_t pj = static_cast<uint32_t>(loc[partner[pd.spos[j]]]);
    if (pj < slot_bits) continue;
    uint32_t &w = j < 4 ? op.xl : op.zl;
    const uint32_t sh = 8 * (j & 3);
    w = (w & ~(0xFFu << sh)) | (pj << sh);
  }
  for (size_t k = 0; k + 1 < p.q.size(); k += 2) {
    if (S & bit(p.q[k])) continue;  // exchanged inside the tile
    op.pos[op.k] = static_cast<uint8_t>(p.q[k]);
    op.gpos[op.k] = static_cast<uint8_t>(p.q[k + 1]);
    ++op.k;
  }
  ops.push_back(op);
  return pd;
}

std::vector<Prim> Engine::fuse_blocks(const std::vector<Prim> &in) {
  const uint32_t width = block_width_;
  const bool single = prec_ == AERHIP_PRECISION_SINGLE;
  const size_t window = 128;
  std::vector<char> used(in.size(), 0);
  std::vector<Prim> out;
  out.reserve(in.size());
  for (size_t i = 0; i < in.size(); ++i) {
    if (used[i]) continue;
    const Prim &p0 = in[i];
    if (!block_fusable(p0, width)) {
      out.push_back(p0);
      continue;
    }
    // Members: p0 and later fusable primitives on the block's qubits (at most
    // `width`) that commute with every primitive they are moved past (those
    // share no qubit with them).
    uint64_t Q = prim_qubits(p0), blocked = 0;
    std::vector<size_t> members = {i};
    for (size_t j = i + 1; j < in.size() && j < i + window; ++j) {
      if (used[j]) continue;
      const Prim &p = in[j];
      const uint64_t t = p.touch;
      if (!block_fusable(p, width)) {
        blocked |= t;
      } else {
        const uint64_t qp = prim_qubits(p);
        if (!(qp & blocked) && static_cast<uint32_t>(__builtin_popcountll(Q | qp)) <= width) {
          members.push_back(j);
          Q |= qp;
        } else {
          blocked |= qp;
        }
      }
      if ((blocked & Q) == Q) break;
    }
    if (members.size() < 2) {
      out.push_back(p0);
      continue;
    }
    std::vector<uint32_t> bq;
    for (uint32_t q = 0; q < 64; ++q)
      if (Q & (1ull << q)) bq.push_back(q);
    const uint32_t w = static_cast<uint32_t>(bq.size());
    double member_cost = 0;
    for (size_t m : members) member_cost += block_cost(in[m].kind, static_cast<uint32_t>(in[m].q.size()), single);
    if (block_cost_check_ && block_cost(K_DENSE, w, single) >= member_cost) {
      out.push_back(p0);
      continue;
    }
    // the block's matrix: column c is the image of basis state c
    const uint32_t D = 1u << w;
    std::vector<cd> M(static_cast<size_t>(D) * D);
    for (uint32_t c = 0; c < D; ++c) {
      cd *col = M.data() + static_cast<size_t>(D) * c;
      col[c] = 1.0;
      for (size_t m : members) apply_small(in[m], bq, pool_, col);
    }
    for (size_t m : members) used[m] = 1;
    Prim f;
    f.kind = w == 1 ? K_DENSE1 : K_DENSE;
    f.q = bq;
    f.len = D * D;
    f.off = pool_add(M.data(), f.len);
    f.tmask = Q;
    f.touch = Q;
    f.unitary = true;
    out.push_back(std::move(f));
    stats_.fused += members.size();
  }
  return out;
}

void Engine::set_block_fusion(uint32_t width, bool cost_check) {
  if (width == 1 || width > 5) throw std::invalid_argument("aerhip: block fusion width must be 0 or 2..5");
  flush();
  block_width_ = width;
  block_cost_check_ = cost_check;
}

void Engine::emit_program(Program &prog) {
  if (block_width_ >= 2 && n_ > fusion_threshold_) queue_ = fuse_blocks(queue_);
  queue_ = merge_swap_layers(queue_);
  const std::vector<PassPlan> plans = plan();
  std::vector<PassDesc> &passes = prog.passes;
  std::vector<TileOp> &ops = prog.ops;
  std::vector<ROp<double>> &rops = prog.rops;
  const uint32_t pool_base = static_cast<uint32_t>(prog.pool.size());
  prog.pool.insert(prog.pool.end(), pool_.begin(), pool_.end());
  passes.reserve(passes.size() + plans.size());
  for (const PassPlan &pp : plans) {
    if (pp.solo && queue_[pp.prims[0]].kind == K_QPERM) {
      passes.push_back(qperm_pass(queue_[pp.prims[0]], ops));
      continue;
    }
    PassDesc pd;
    std::memset(&pd, 0, sizeof(pd));
    int loc[64];
    for (int &v : loc) v = -1;
    uint32_t ns = 0, nc = 0;
    for (uint32_t q = 0; q < n_; ++q) {
      if (pp.smask & bit(q)) {
        loc[q] = static_cast<int>(ns);
        pd.spos[ns++] = static_cast<uint8_t>(q);
      } else {
        pd.comp[nc++] = static_cast<uint8_t>(q);
      }
    }
    pd.kind = pp.solo ? 1u : 0u;
    if (pp.solo && queue_[pp.prims[0]].kind == K_PROBE) {
      // probe sweep: one TileOp with the masks over all qubits
      const Prim &p = queue_[pp.prims[0]];
      TileOp op;
      std::memset(&op, 0, sizeof(op));
      op.kind = K_PROBE;
      op.cg = p.xmask;
      op.zg = p.zmask;
      op.t = p.q.empty() ? 0u : p.q[0];
      op.npairs = p.npairs;
      op.pr = p.pr;
      op.pi = p.pi;
      PassDesc sd;
      std::memset(&sd, 0, sizeof(sd));
      sd.kind = 3;
      sd.readonly = 1;
      sd.nb = tile_bits_;
      sd.ncomp = n_ - tile_bits_;
      sd.op_begin = static_cast<uint32_t>(ops.size());
      sd.nops = 1;
      ops.push_back(op);
      passes.push_back(sd);
      continue;
    }
    pd.nb = ns;
    pd.readonly = 1;
    for (uint32_t idx : pp.prims)
      if (queue_[idx].kind != K_PROBE) pd.readonly = 0;
    uint32_t c = 0;
    while (c < ns && pd.spos[c] == c) ++c;
    pd.c = c;
    pd.ncomp = nc;
    pd.op_begin = static_cast<uint32_t>(ops.size());
    std::vector<TileOp> pops;
    for (uint32_t idx : pp.prims) {
      const Prim &p = queue_[idx];
      TileOp op;
      std::memset(&op, 0, sizeof(op));
      op.kind = p.kind;
      op.off = p.off + pool_base;
      op.npairs = p.npairs;
      op.nonunitary = p.unitary ? 0u : 1u;
      op.pr = p.pr;
      op.pi = p.pi;
      for (size_t i = 0; i < p.ctrl.size(); ++i) {
        const uint32_t q = p.ctrl[i];
        const uint64_t v = (p.ctrl_val >> i) & 1ull;
        if (loc[q] >= 0) {
          op.cl |= 1u << loc[q];
          op.cvl |= static_cast<uint32_t>(v) << loc[q];
        } else {
          op.cg |= bit(q);
          op.cvg |= v << q;
        }
      }
      op.k = static_cast<uint32_t>(p.q.size());
      for (size_t j = 0; j < p.q.size(); ++j) {
        const uint32_t q = p.q[j];
        op.gpos[j] = static_cast<uint8_t>(q);
        op.pos[j] = loc[q] >= 0 ? static_cast<uint8_t>(loc[q]) : kNotLocal;
      }
      switch (p.kind) {
        case K_DENSE1:
        case K_MCX:
        case K_MCY:
          op.t = op.pos[0];
          break;
        case K_PAULI:
        case K_PROBE: {
          for (uint32_t q = 0; q < n_; ++q) {
            if (p.xmask & bit(q)) op.xl |= 1u << loc[q];
            if (p.zmask & bit(q)) {
              if (loc[q] >= 0) op.zl |= 1u << loc[q];
              else op.zg |= bit(q);
            }
          }
          op.t = p.q.empty() ? 0 : op.pos[0];
          op.k = 0;
          break;
        }
        default:
          break;
      }
      pops.push_back(op);
    }
    const uint32_t full_bits =
        prec_ == AERHIP_PRECISION_DOUBLE ? kTileBitsC128 : kTileBitsC64;
    if (!pp.solo && reg_bits_ > 0 && ns == full_bits) {
      if (fusion_enable_ && n_ > fusion_threshold_)
        pops = fuse_phase_runs(pops, pd.spos, ns, prog.pool, fusion_max_qubit_, &stats_.fused);
      pd.kind = 2;  // register-layout tile pass: op_begin/nops index the ROps
      pd.op_begin = static_cast<uint32_t>(rops.size());
      register_program(pops, ns, reg_bits_, low_bits_, pd.spos, prog.pool, rops, ops, &stats_.gathers);
      pd.nops = static_cast<uint32_t>(rops.size()) - pd.op_begin;
      if (pd.readonly) rops.back().extra = 1;  // the final STORE writes nothing
      validate_register_program(rops.data() + pd.op_begin, pd.nops, ns, reg_bits_, pd.spos, ops.size());
    } else {
      pd.nops = static_cast<uint32_t>(pops.size());
      ops.insert(ops.end(), pops.begin(), pops.end());
    }
    passes.push_back(pd);
  }

  queue_.clear();
  pool_.clear();
}

namespace {
// Register passes of at most this many ops (LOAD, STORE and a few gates)
// launch the wide lean variant.
constexpr uint32_t kWidePassOps = 8;

// The device program of a flush as the pass kernels see it, and the states
// they run on: `count` states of 2^n amplitudes, state b at sv + b *
// state_stride amplitudes, taking pass descriptors pd + b * pd_stride and
// probe partials probes + b * probes_per_state * (tiles of the pass).
struct DeviceProgram {
  const PassDesc *passes;
  const TileOp *ops;
  const void *rops;
  const cx<double> *pool;
  double *probes;
  uint32_t probes_per_state;
};

template <typename T, int B, int R>
void launch_reg_pass(bool lean, bool wide, bool diag, uint32_t grid, hipStream_t stream, cx<T> *sv,
                     const PassDesc *pd, const ROp<T> *rops, const TileOp *tops, const cx<double> *pool,
                     double *probes, uint64_t state_stride, uint32_t pd_stride, uint32_t tile_log,
                     uint32_t pps) {
  // the staged lane tables of fused diagonals, only for passes that have them
  const size_t dyn = diag ? kDiagLdsBytes : 0;
  if constexpr (sizeof(T) == 8 && R == 3) {
    if (lean && wide) {
      k_reg_pass<T, B, R, false, true><<<grid, 1 << (B - R), dyn, stream>>>(
          sv, pd, rops, tops, pool, probes, state_stride, pd_stride, tile_log, pps);
      return;
    }
  }
  if (lean)
    k_reg_pass<T, B, R, false><<<grid, 1 << (B - R), dyn, stream>>>(sv, pd, rops, tops, pool, probes,
                                                                    state_stride, pd_stride, tile_log, pps);
  else
    k_reg_pass<T, B, R, true><<<grid, 1 << (B - R), dyn, stream>>>(sv, pd, rops, tops, pool, probes,
                                                                   state_stride, pd_stride, tile_log, pps);
}

// Launches pass index i (host descriptor `pd`, the same pass of every state)
// over `count` states; `lean` / `wide` pick the register-pass variant, `diag`
// gives it the LDS of staged fused-diagonal lane tables.
void launch_pass(int prec, uint32_t reg_bits, uint32_t n, uint32_t tile_bits, const PassDesc &pd,
                 uint32_t i, bool lean, bool wide, bool diag, const DeviceProgram &dp, void *sv,
                 uint64_t count, uint32_t pd_stride, hipStream_t stream) {
  const uint64_t state_stride = 1ull << n;
  const PassDesc *dpd = dp.passes + i;
  const uint32_t pps = dp.probes_per_state;
  auto grid_of = [&](uint32_t tile_log) {
    const uint64_t g = count << tile_log;
    if (g > 0xFFFFFFFFull) throw std::invalid_argument("aerhip: too many states for one launch");
    return static_cast<uint32_t>(g);
  };
  if (pd.kind == 2) {
    const uint32_t tl = pd.ncomp;
    if (prec == AERHIP_PRECISION_DOUBLE) {
      auto *s = static_cast<cx<double> *>(sv);
      auto *ro = static_cast<const ROp<double> *>(dp.rops);
      if (reg_bits == 3)
        launch_reg_pass<double, kTileBitsC128, 3>(lean, wide, diag, grid_of(tl), stream, s, dpd, ro, dp.ops, dp.pool,
                                                  dp.probes, state_stride, pd_stride, tl, pps);
      else
        launch_reg_pass<double, kTileBitsC128, 4>(lean, wide, diag, grid_of(tl), stream, s, dpd, ro, dp.ops, dp.pool,
                                                  dp.probes, state_stride, pd_stride, tl, pps);
    } else {
      auto *s = static_cast<cx<float> *>(sv);
      auto *ro = static_cast<const ROp<float> *>(dp.rops);
      if (reg_bits == 4)
        launch_reg_pass<float, kTileBitsC64, 4>(lean, wide, diag, grid_of(tl), stream, s, dpd, ro, dp.ops, dp.pool,
                                                dp.probes, state_stride, pd_stride, tl, pps);
      else
        launch_reg_pass<float, kTileBitsC64, 5>(lean, wide, diag, grid_of(tl), stream, s, dpd, ro, dp.ops, dp.pool,
                                                dp.probes, state_stride, pd_stride, tl, pps);
    }
  } else if (pd.kind == 3) {
    // probe sweep over the whole state, one workgroup per tile-sized slice
    const uint32_t tl = n - tile_bits;
    if (prec == AERHIP_PRECISION_DOUBLE)
      k_probe_sweep<double><<<grid_of(tl), kThreads, 0, stream>>>(
          static_cast<const cx<double> *>(sv), state_stride, dpd, dp.ops, dp.probes, state_stride, pd_stride,
          tl, pps);
    else
      k_probe_sweep<float><<<grid_of(tl), kThreads, 0, stream>>>(
          static_cast<const cx<float> *>(sv), state_stride, dpd, dp.ops, dp.probes, state_stride, pd_stride,
          tl, pps);
  } else if (pd.kind == 4) {
    // qubit permutation (a layer of SWAPs): four tiles per workgroup, each
    // exchanged with its partner tile or skipped
    const uint32_t per = std::min<uint32_t>(2, pd.ncomp), tl = pd.ncomp - per;
    if (prec == AERHIP_PRECISION_DOUBLE)
      k_qperm<double, kTileBitsC128 - 1><<<grid_of(tl), kThreads, 0, stream>>>(
          static_cast<cx<double> *>(sv), dpd, dp.ops, state_stride, pd_stride, tl, per);
    else
      k_qperm<float, kTileBitsC64 - 1><<<grid_of(tl), kThreads, 0, stream>>>(
          static_cast<cx<float> *>(sv), dpd, dp.ops, state_stride, pd_stride, tl, per);
  } else if (pd.kind == 0) {
    // LDS-resident tile pass (tiles smaller than the register layouts need,
    // or register layouts switched off).
    const uint32_t tl = pd.ncomp;
    if (prec == AERHIP_PRECISION_DOUBLE)
      k_tile_lds<double, kTileBitsC128><<<grid_of(tl), kThreads, 0, stream>>>(
          static_cast<cx<double> *>(sv), dpd, dp.ops, dp.pool, dp.probes, state_stride, pd_stride, tl, pps);
    else
      k_tile_lds<float, kTileBitsC64><<<grid_of(tl), kThreads, 0, stream>>>(
          static_cast<cx<float> *>(sv), dpd, dp.ops, dp.pool, dp.probes, state_stride, pd_stride, tl, pps);
  } else {
    if (count != 1) throw std::invalid_argument("aerhip: dense matrices on 6 or more qubits run on one state at a time");
    // One workgroup per group up to 2^16 workgroups; beyond that each
    // workgroup loops over groups (the grid stays far below the launch
    // limit of 2^32 work-items per dimension).
    const uint64_t ngroups = 1ull << (n - pd.nb);
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>(ngroups, 1ull << 16));
    if (prec == AERHIP_PRECISION_DOUBLE)
      k_dense_group<double><<<grid, kThreads, 0, stream>>>(static_cast<cx<double> *>(sv), dpd, dp.ops,
                                                           dp.pool, ngroups);
    else
      k_dense_group<float><<<grid, kThreads, 0, stream>>>(static_cast<cx<float> *>(sv), dpd, dp.ops,
                                                          dp.pool, ngroups);
  }
  AERHIP_CHECK(hipGetLastError());
}

// AERHIP_GRAPHS=0 launches every flush's passes one by one (no HIP graphs).
bool graphs_enabled() {
  static const bool on = [] {
    const char *e = std::getenv("AERHIP_GRAPHS");
    return !(e && e[0] == '0');
  }();
  return on;
}

// The register-pass variant of a pass: lean if every op is a lean kind, wide
// if it also has at most kWidePassOps ops.
void pass_variant(const PassDesc &pd, const std::vector<ROp<double>> &rops, bool &lean, bool &wide,
                  bool &diag) {
  lean = true;
  wide = pd.nops <= kWidePassOps;
  diag = false;
  if (pd.kind != 2) return;
  for (uint32_t k = 0; k < pd.nops; ++k) {
    const ROp<double> &r = rops[pd.op_begin + k];
    lean = lean && lean_kind(r.kind);
    const uint32_t kk = r.kind & ~F_ALL;
    diag = diag || ((kk == Q_LOAD || kk == Q_FILL) && r.a != 0);
  }
}
}  // namespace

void Engine::flush() {
  if (queue_.empty()) return;
  if (plan_only_) throw std::logic_error("aerhip: a planning handle cannot execute");
  AERHIP_CHECK(hipSetDevice(device_));
  const uint64_t nprobes = probes_queued_;
  Program prog;
  emit_program(prog);
  const std::vector<PassDesc> &passes = prog.passes;
  const std::vector<TileOp> &ops = prog.ops;
  const std::vector<ROp<double>> &rops = prog.rops;
  const std::vector<cd> &pool = prog.pool;

  // Upload [passes | ops | rops | pool] with one copy from pinned staging.
  const size_t pass_bytes = passes.size() * sizeof(PassDesc);
  const size_t ops_off = align_up(pass_bytes, 256);
  const size_t ops_bytes = ops.size() * sizeof(TileOp);
  const size_t rops_off = align_up(ops_off + ops_bytes, 256);
  const size_t rops_bytes = rops.size() * sizeof(ROp<double>);
  const size_t pool_off = align_up(rops_off + rops_bytes, 256);
  const size_t total = pool_off + pool.size() * sizeof(cd) + 16;
  AERHIP_CHECK(hipEventSynchronize(prog_ev_));  // previous upload finished with staging
  if (total > h_prog_cap_) {
    if (h_prog_) AERHIP_CHECK(hipHostFree(h_prog_));
    h_prog_cap_ = align_up(total * 2, 1 << 16);
    AERHIP_CHECK(hipHostMalloc(&h_prog_, h_prog_cap_));
  }
  if (total > d_prog_cap_) {
    AERHIP_CHECK(hipStreamSynchronize(stream_));
    if (d_prog_) AERHIP_CHECK(hipFree(d_prog_));
    d_prog_cap_ = align_up(total * 2, 1 << 16);
    AERHIP_CHECK(hipMalloc(&d_prog_, d_prog_cap_));
  }
  char *h = static_cast<char *>(h_prog_);
  std::memcpy(h, passes.data(), pass_bytes);
  std::memcpy(h + ops_off, ops.data(), ops_bytes);
  if (prec_ == AERHIP_PRECISION_DOUBLE) {
    std::memcpy(h + rops_off, rops.data(), rops_bytes);
  } else {
    // complex64 programs carry their matrices in float
    rops_to_float(rops.data(), rops.size(), reinterpret_cast<ROp<float> *>(h + rops_off));
  }
  if (!pool.empty()) std::memcpy(h + pool_off, pool.data(), pool.size() * sizeof(cd));
  AERHIP_CHECK(hipMemcpyAsync(d_prog_, h_prog_, total, hipMemcpyHostToDevice, stream_));
  AERHIP_CHECK(hipEventRecord(prog_ev_, stream_));

  char *d = static_cast<char *>(d_prog_);
  const PassDesc *dpass = reinterpret_cast<const PassDesc *>(d);
  const TileOp *dops = reinterpret_cast<const TileOp *>(d + ops_off);
  const void *drops = d + rops_off;
  const cx<double> *dpool = reinterpret_cast<const cx<double> *>(d + pool_off);
  // Probe partials: one double per (probe, tile); every tile pass has
  // 2^(n - tile_bits) tiles.
  double *dprobes = nullptr;
  const uint64_t ntiles = 1ull << (n_ - tile_bits_);
  if (nprobes) {
    const size_t need = (nprobes * ntiles + nprobes) * sizeof(double);
    if (need > probes_cap_) {
      AERHIP_CHECK(hipStreamSynchronize(stream_));
      if (d_probes_) AERHIP_CHECK(hipFree(d_probes_));
      probes_cap_ = align_up(need, 1 << 20);
      AERHIP_CHECK(hipMalloc(reinterpret_cast<void **>(&d_probes_), probes_cap_));
    }
    dprobes = d_probes_;
  }
  if (event_timing_) AERHIP_CHECK(hipEventRecord(t0_ev_, stream_));
  const DeviceProgram dp{dpass, dops, drops, dpool, dprobes, 0};
  // The launch sequence is a function of this signature: the kernel arguments
  // (device addresses, offsets into the uploaded program) and each pass's
  // kernel and grid.  A flush with the signature of the previous one (the same
  // circuit again, or the same shape with other angles: the new program was
  // uploaded to the same buffer above) replays the sequence as a HIP graph:
  // captured on the second occurrence, replayed from the third on.
  std::vector<uint64_t> sig = {reinterpret_cast<uint64_t>(d_prog_), reinterpret_cast<uint64_t>(sv_),
                               reinterpret_cast<uint64_t>(dprobes), static_cast<uint64_t>(prec_),
                               reg_bits_, n_, tile_bits_, passes.size(), ops_off, rops_off, pool_off};
  std::vector<std::array<bool, 3>> variant(passes.size());
  for (size_t i = 0; i < passes.size(); ++i) {
    bool lean = false, wide = false, diag = false;
    pass_variant(passes[i], rops, lean, wide, diag);
    variant[i] = {lean, wide, diag};
    const PassDesc &pd = passes[i];
    sig.push_back((uint64_t(pd.kind) << 48) | (uint64_t(pd.ncomp) << 32) | (uint64_t(pd.nb) << 16) |
                  (uint64_t(diag) << 2) | (uint64_t(lean) << 1) | uint64_t(wide));
  }
  auto launch_all = [&] {
    for (size_t i = 0; i < passes.size(); ++i)
      launch_pass(prec_, reg_bits_, n_, tile_bits_, passes[i], static_cast<uint32_t>(i), variant[i][0],
                  variant[i][1], variant[i][2], dp, sv_, 1, 0, stream_);
  };
  if (graphs_enabled() && graph_exec_ && sig == graph_sig_) {
    AERHIP_CHECK(hipGraphLaunch(graph_exec_, stream_));
    stats_.graph_replays++;
  } else if (graphs_enabled() && sig == last_sig_) {
    if (graph_exec_) AERHIP_CHECK(hipGraphExecDestroy(graph_exec_));
    graph_exec_ = nullptr;
    hipGraph_t g = nullptr;
    AERHIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
      launch_all();
    } catch (...) {
      (void)hipStreamEndCapture(stream_, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    AERHIP_CHECK(hipStreamEndCapture(stream_, &g));
    const hipError_t ie = hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    AERHIP_CHECK(ie);
    graph_sig_ = sig;
    AERHIP_CHECK(hipGraphLaunch(graph_exec_, stream_));
    stats_.graph_captures++;
  } else {
    launch_all();
  }
  last_sig_ = std::move(sig);
  stats_.launches += passes.size();
  stats_.passes += passes.size();
  if (nprobes) {
    // Sum each probe's tile partials in tile order and fetch the values.
    double *sums = d_probes_ + nprobes * ntiles;
    k_sum_rows<<<grid_for(nprobes, 1u << 30), kThreads, 0, stream_>>>(
        d_probes_, nprobes, static_cast<uint32_t>(ntiles), sums);
    AERHIP_CHECK(hipGetLastError());
    const size_t base = saved_expvals_.size();
    saved_expvals_.resize(base + nprobes);
    AERHIP_CHECK(hipMemcpyAsync(saved_expvals_.data() + base, sums, nprobes * sizeof(double),
                                hipMemcpyDeviceToHost, stream_));
    AERHIP_CHECK(hipStreamSynchronize(stream_));
    probes_queued_ = 0;
  }
  if (event_timing_) {
    AERHIP_CHECK(hipEventRecord(t1_ev_, stream_));
    AERHIP_CHECK(hipEventSynchronize(t1_ev_));
    float ms = 0;
    AERHIP_CHECK(hipEventElapsedTime(&ms, t0_ev_, t1_ev_));
    stats_.pass_ms += ms;
    stats_.timed_launches += passes.size();
  }
}

void Engine::set_fusion(bool enable, uint32_t max_qubit, uint32_t threshold) {
  if (max_qubit < 2) throw std::invalid_argument("aerhip: fusion_max_qubit must be at least 2");
  flush();
  fusion_enable_ = enable;
  fusion_max_qubit_ = max_qubit;
  fusion_threshold_ = threshold;
}

void Engine::set_reg_bits(uint32_t r) {
  const bool ok = r == 0 || (prec_ == AERHIP_PRECISION_DOUBLE ? (r == 3 || r == 4) : (r == 4 || r == 5));
  if (!ok) throw std::invalid_argument("aerhip: register bits must be 0, 3 or 4 (c128), 0, 4 or 5 (c64)");
  flush();
  reg_bits_ = r;
}

void Engine::set_event_timing(bool on) {
  flush();
  if (on && !t0_ev_) {
    AERHIP_CHECK(hipEventCreate(&t0_ev_));
    AERHIP_CHECK(hipEventCreate(&t1_ev_));
  }
  event_timing_ = on;
}

// ---------------------------------------------------------------------------
// Readout
// ---------------------------------------------------------------------------

void Engine::ensure_partial(size_t n) {
  if (n <= partial_cap_) return;
  AERHIP_CHECK(hipStreamSynchronize(stream_));
  if (d_partial_) AERHIP_CHECK(hipFree(d_partial_));
  partial_cap_ = std::max<size_t>(n, 1 << 16);
  AERHIP_CHECK(hipMalloc(reinterpret_cast<void **>(&d_partial_), partial_cap_ * sizeof(double)));
}

void *Engine::scratch(size_t bytes) {
  if (bytes > scratch_cap_) {
    AERHIP_CHECK(hipStreamSynchronize(stream_));
    if (d_scratch_) AERHIP_CHECK(hipFree(d_scratch_));
    scratch_cap_ = std::max<size_t>(align_up(bytes, 1 << 16), 1 << 20);
    AERHIP_CHECK(hipMalloc(&d_scratch_, scratch_cap_));
  }
  return d_scratch_;
}

double Engine::fetch_scalar(const double *dptr) {
  AERHIP_CHECK(hipMemcpyAsync(h_scalar_, dptr, sizeof(double), hipMemcpyDeviceToHost, stream_));
  AERHIP_CHECK(hipStreamSynchronize(stream_));
  return h_scalar_[0];
}

template <typename T>
double Engine::norm_t() {
  const uint32_t grid = grid_for(size_, 4096);
  ensure_partial(grid + 1);
  k_sumsq<T><<<grid, kThreads, 0, stream_>>>(static_cast<const cx<T> *>(sv_), size_, d_partial_);
  k_sum_rows<<<1, kThreads, 0, stream_>>>(d_partial_, 1, grid, d_partial_ + grid);
  AERHIP_CHECK(hipGetLastError());
  return fetch_scalar(d_partial_ + grid);
}

double Engine::norm() {
  flush();
  AERHIP_CHECK(hipSetDevice(device_));
  return prec_ == AERHIP_PRECISION_DOUBLE ? norm_t<double>() : norm_t<float>();
}

template <typename T>
double Engine::norm_matrix_t(const MarginalDesc &md, const std::vector<cd> &mat) {
  const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>(md.inner, 4096));
  const size_t mat_off = 256;
  char *sc = static_cast<char *>(scratch(mat_off + mat.size() * sizeof(cd)));
  AERHIP_CHECK(hipMemcpyAsync(sc, &md, sizeof(md), hipMemcpyHostToDevice, stream_));
  AERHIP_CHECK(hipMemcpyAsync(sc + mat_off, mat.data(), mat.size() * sizeof(cd),
                              hipMemcpyHostToDevice, stream_));
  ensure_partial(grid + 1);
  k_norm_group<T><<<grid, kThreads, 0, stream_>>>(static_cast<const cx<T> *>(sv_),
                                                  reinterpret_cast<const MarginalDesc *>(sc),
                                                  reinterpret_cast<const cx<double> *>(sc + mat_off),
                                                  d_partial_);
  k_sum_rows<<<1, kThreads, 0, stream_>>>(d_partial_, 1, grid, d_partial_ + grid);
  AERHIP_CHECK(hipGetLastError());
  return fetch_scalar(d_partial_ + grid);
}

// QubitVector::norm(qubits, mat) (qubitvector.hpp:1889-2054) and, with a
// diagonal written as a matrix, norm_diagonal (qubitvector.hpp:1946-2082).
double Engine::norm_matrix(const std::vector<uint32_t> &qubits, const std::vector<cd> &mat) {
  check_qubits(qubits);
  const size_t k = qubits.size();
  if (k == 0) return norm();
  if (k > 10) throw std::invalid_argument("aerhip: norm of a matrix on more than 10 qubits");
  const uint64_t D = 1ull << k;
  if (mat.size() != D * D)
    throw std::invalid_argument("QubitVector: matrix size does not match the number of qubits");
  flush();
  AERHIP_CHECK(hipSetDevice(device_));
  MarginalDesc md;
  std::memset(&md, 0, sizeof(md));
  std::vector<uint32_t> sorted = qubits;
  std::sort(sorted.begin(), sorted.end());
  md.nq = static_cast<uint32_t>(k);
  md.inner = size_ >> k;
  for (uint32_t i = 0; i < md.nq; ++i) {
    md.sorted[i] = static_cast<uint8_t>(sorted[i]);
    md.order[i] = static_cast<uint8_t>(qubits[i]);
  }
  return prec_ == AERHIP_PRECISION_DOUBLE ? norm_matrix_t<double>(md, mat)
                                          : norm_matrix_t<float>(md, mat);
}

cd Engine::get_state(uint64_t pos) {
  if (pos >= size_) throw std::invalid_argument("aerhip: amplitude index out of range");
  flush();
  AERHIP_CHECK(hipSetDevice(device_));
  if (prec_ == AERHIP_PRECISION_DOUBLE) {
    AERHIP_CHECK(hipMemcpyAsync(h_scalar_, static_cast<cx<double> *>(sv_) + pos, 16,
                                hipMemcpyDeviceToHost, stream_));
    AERHIP_CHECK(hipStreamSynchronize(stream_));
    return cd(h_scalar_[0], h_scalar_[1]);
  }
  AERHIP_CHECK(hipMemcpyAsync(h_scalar_, static_cast<cx<float> *>(sv_) + pos, 8,
                              hipMemcpyDeviceToHost, stream_));
  AERHIP_CHECK(hipStreamSynchronize(stream_));
  const float *f = reinterpret_cast<const float *>(h_scalar_);
  return cd(f[0], f[1]);
}

void Engine::set_state(uint64_t pos, cd value) {
  if (pos >= size_) throw std::invalid_argument("aerhip: amplitude index out of range");
  flush();
  AERHIP_CHECK(hipSetDevice(device_));
  AERHIP_CHECK(hipStreamSynchronize(stream_));  // the staging word is free
  if (prec_ == AERHIP_PRECISION_DOUBLE) {
    h_scalar_[0] = value.real();
    h_scalar_[1] = value.imag();
    AERHIP_CHECK(hipMemcpyAsync(static_cast<cx<double> *>(sv_) + pos, h_scalar_, 16,
                                hipMemcpyHostToDevice, stream_));
  } else {
    float *f = reinterpret_cast<float *>(h_scalar_);
    f[0] = static_cast<float>(value.real());
    f[1] = static_cast<float>(value.imag());
    AERHIP_CHECK(hipMemcpyAsync(static_cast<cx<float> *>(sv_) + pos, h_scalar_, 8,
                                hipMemcpyHostToDevice, stream_));
  }
  AERHIP_CHECK(hipStreamSynchronize(stream_));
}

cd Engine::inner_product(Engine &other) {
  if (other.n_ != n_) throw std::invalid_argument("aerhip: inner product of states of different size");
  if (other.device_ != device_) throw std::invalid_argument("aerhip: inner product across devices");
  flush();
  other.flush();
  AERHIP_CHECK(hipStreamSynchronize(other.stream_));
  AERHIP_CHECK(hipSetDevice(device_));
  const uint32_t grid = grid_for(size_, 4096);
  ensure_partial(2 * grid + 2);
  const bool a64 = prec_ == AERHIP_PRECISION_DOUBLE, b64 = other.prec_ == AERHIP_PRECISION_DOUBLE;
  if (a64 && b64)
    k_inner<double, double><<<grid, kThreads, 0, stream_>>>(static_cast<const cx<double> *>(sv_),
        static_cast<const cx<double> *>(other.sv_), size_, d_partial_);
  else if (a64)
    k_inner<double, float><<<grid, kThreads, 0, stream_>>>(static_cast<const cx<double> *>(sv_),
        static_cast<const cx<float> *>(other.sv_), size_, d_partial_);
  else if (b64)
    k_inner<float, double><<<grid, kThreads, 0, stream_>>>(static_cast<const cx<float> *>(sv_),
        static_cast<const cx<double> *>(other.sv_), size_, d_partial_);
  else
    k_inner<float, float><<<grid, kThreads, 0, stream_>>>(static_cast<const cx<float> *>(sv_),
        static_cast<const cx<float> *>(other.sv_), size_, d_partial_);
  k_sum_rows<<<1, kThreads, 0, stream_>>>(d_partial_, 2, grid, d_partial_ + 2 * grid);
  AERHIP_CHECK(hipGetLastError());
  const double re = fetch_scalar(d_partial_ + 2 * grid);
  const double im = fetch_scalar(d_partial_ + 2 * grid + 1);
  return cd(re, im);
}

double Engine::probability(uint64_t outcome) {
  if (outcome >= size_) throw std::invalid_argument("aerhip: outcome out of range");
  flush();
  AERHIP_CHECK(hipSetDevice(device_));
  double x, y;
  if (prec_ == AERHIP_PRECISION_DOUBLE) {
    cx<double> a;
    AERHIP_CHECK(hipMemcpyAsync(&a, static_cast<cx<double> *>(sv_) + outcome, sizeof(a),
                                hipMemcpyDeviceToHost, stream_));
    AERHIP_CHECK(hipStreamSynchronize(stream_));
    x = a.x;
    y = a.y;
  } else {
    cx<float> a;
    AERHIP_CHECK(hipMemcpyAsync(&a, static_cast<cx<float> *>(sv_) + outcome, sizeof(a),
                                hipMemcpyDeviceToHost, stream_));
    AERHIP_CHECK(hipStreamSynchronize(stream_));
    x = a.x;
    y = a.y;
  }
  return x * x + y * y;
}

template <typename T>
void Engine::probabilities_t(const std::vector<uint32_t> &qubits, double *out) {
  const cx<T> *sv = static_cast<const cx<T> *>(sv_);
  const uint32_t nq = static_cast<uint32_t>(qubits.size());
  std::vector<uint32_t> sorted = qubits;
  std::sort(sorted.begin(), sorted.end());
  const uint64_t outcomes = 1ull << nq;
  double *dout = nullptr;
  if (nq == 0 || (nq == n_ && sorted == qubits)) {
    ensure_partial(size_);
    k_probs_all<T><<<grid_for(size_, 16384), kThreads, 0, stream_>>>(sv, size_, d_partial_);
    AERHIP_CHECK(hipGetLastError());
    AERHIP_CHECK(hipMemcpyAsync(out, d_partial_, size_ * sizeof(double), hipMemcpyDeviceToHost,
                                stream_));
    AERHIP_CHECK(hipStreamSynchronize(stream_));
    return;
  }
  MarginalDesc md;
  std::memset(&md, 0, sizeof(md));
  md.nq = nq;
  md.inner = 1ull << (n_ - nq);
  for (uint32_t i = 0; i < nq; ++i) {
    md.sorted[i] = static_cast<uint8_t>(sorted[i]);
    md.order[i] = static_cast<uint8_t>(qubits[i]);
  }
  MarginalDesc *dmd = static_cast<MarginalDesc *>(scratch(sizeof(MarginalDesc)));
  if (md.inner <= 256) {
    AERHIP_CHECK(hipMemcpy(dmd, &md, sizeof(md), hipMemcpyHostToDevice));
    ensure_partial(outcomes);
    dout = d_partial_;
    k_marginal_small<T><<<grid_for(outcomes, 1u << 30), kThreads, 0, stream_>>>(sv, dmd, outcomes,
                                                                                 dout);
  } else {
    uint64_t nchunk = std::max<uint64_t>(1, 8192 / outcomes);
    nchunk = std::min<uint64_t>(nchunk, std::max<uint64_t>(1, md.inner / 1024));
    md.nchunk = static_cast<uint32_t>(nchunk);
    AERHIP_CHECK(hipMemcpy(dmd, &md, sizeof(md), hipMemcpyHostToDevice));
    ensure_partial(outcomes * nchunk + outcomes);
    // outcomes is a power of two: grid.y * grid.z == outcomes exactly.
    dim3 grid(static_cast<uint32_t>(nchunk), 1, 1);
    grid.y = static_cast<uint32_t>(std::min<uint64_t>(outcomes, 32768));
    grid.z = static_cast<uint32_t>(outcomes / grid.y);
    k_marginal_chunks<T><<<grid, kThreads, 0, stream_>>>(sv, dmd, d_partial_);
    dout = d_partial_ + outcomes * nchunk;
    k_sum_rows<<<grid_for(outcomes, 1u << 30), kThreads, 0, stream_>>>(
        d_partial_, outcomes, static_cast<uint32_t>(nchunk), dout);
  }
  AERHIP_CHECK(hipGetLastError());
  AERHIP_CHECK(hipMemcpyAsync(out, dout, outcomes * sizeof(double), hipMemcpyDeviceToHost, stream_));
  AERHIP_CHECK(hipStreamSynchronize(stream_));
}

std::vector<double> Engine::probabilities(const std::vector<uint32_t> &qubits) {
  check_qubits(qubits);
  if (qubits.size() > 40) throw std::invalid_argument("aerhip: too many qubits");
  flush();
  AERHIP_CHECK(hipSetDevice(device_));
  std::vector<double> out(qubits.empty() ? size_ : (1ull << qubits.size()));
  if (prec_ == AERHIP_PRECISION_DOUBLE)
    probabilities_t<double>(qubits, out.data());
  else
    probabilities_t<float>(qubits, out.data());
  return out;
}

template <typename T>
void Engine::sample_t(const std::vector<double> &rnds, uint64_t *out) {
  const cx<T> *sv = static_cast<const cx<T> *>(sv_);
  const uint64_t shots = rnds.size();
  char *sc = static_cast<char *>(scratch(2 * shots * sizeof(double)));
  double *drnd = reinterpret_cast<double *>(sc);
  uint64_t *dout = reinterpret_cast<uint64_t *>(sc + shots * sizeof(double));
  AERHIP_CHECK(hipMemcpyAsync(drnd, rnds.data(), shots * sizeof(double), hipMemcpyHostToDevice,
                              stream_));
  const uint32_t isz = static_cast<uint32_t>(sample_index_size_);
  if (size_ < (1ull << isz)) {
    k_sample_linear<T><<<grid_for(shots, 1u << 30), kThreads, 0, stream_>>>(sv, size_, drnd, shots,
                                                                             dout);
  } else {
    const uint32_t blockbits = n_ - isz;               // log2 of amplitudes per index block
    const uint32_t sbits = std::min<uint32_t>(blockbits, 10);
    const uint64_t nruns = size_ >> sbits;
    const uint64_t nblocks = 1ull << isz;
    ensure_partial(nruns + nblocks);
    double *runs = d_partial_;
    double *blocks = d_partial_ + nruns;
    k_run_sums<T><<<static_cast<uint32_t>((nruns + 3) / 4), kThreads, 0, stream_>>>(sv, sbits, nruns,
                                                                                   runs);
    k_sum_rows<<<grid_for(nblocks, 1u << 30), kThreads, 0, stream_>>>(
        runs, nblocks, static_cast<uint32_t>(nruns / nblocks), blocks);
    k_sample<T><<<grid_for(shots, 1u << 30), kThreads, 0, stream_>>>(sv, size_, isz, sbits, blocks,
                                                                      runs, drnd, shots, dout);
  }
  AERHIP_CHECK(hipGetLastError());
  AERHIP_CHECK(hipMemcpyAsync(out, dout, shots * sizeof(uint64_t), hipMemcpyDeviceToHost, stream_));
  AERHIP_CHECK(hipStreamSynchronize(stream_));
}

std::vector<uint64_t> Engine::sample_measure(const std::vector<double> &rnds) {
  flush();
  AERHIP_CHECK(hipSetDevice(device_));
  std::vector<uint64_t> out(rnds.size());
  if (rnds.empty()) return out;
  if (sample_index_size_ < 1 || static_cast<uint32_t>(sample_index_size_) > n_ + 30)
    throw std::invalid_argument("aerhip: sample_measure index size out of range");
  if (prec_ == AERHIP_PRECISION_DOUBLE)
    sample_t<double>(rnds, out.data());
  else
    sample_t<float>(rnds, out.data());
  return out;
}

template <typename T>
double Engine::expval_t(uint64_t xmask, uint64_t zmask, uint32_t xmax, cd ph) {
  const uint32_t grid = grid_for(xmask ? size_ / 2 : size_, 4096);
  ensure_partial(grid + 1);
  k_expval_pauli<T><<<grid, kThreads, 0, stream_>>>(static_cast<const cx<T> *>(sv_), size_, xmask,
                                                     zmask, xmax, ph.real(), ph.imag(), d_partial_);
  k_sum_rows<<<1, kThreads, 0, stream_>>>(d_partial_, 1, grid, d_partial_ + grid);
  AERHIP_CHECK(hipGetLastError());
  return fetch_scalar(d_partial_ + grid);
}

// Pending gates are dropped when the state is overwritten, but saved
// expectation values are evaluated first.
void Engine::discard_queue() {
  if (probes_queued_) flush();
  queue_.clear();
  pool_.clear();
}

uint64_t Engine::save_expval_pauli(const std::vector<uint32_t> &qubits, const std::string &pauli,
                                   cd phase) {
  check_qubits(qubits);
  const PauliMasks m = pauli_masks(qubits, pauli);
  const uint64_t slot = saved_expvals_.size() + probes_queued_;
  Prim p;
  p.kind = K_PROBE;
  const cd ph = add_y_phase(m.num_y, phase);
  p.pr = ph.real();
  p.pi = ph.imag();
  p.xmask = m.x;
  p.zmask = m.z;
  p.tmask = m.x;
  if (m.x) p.q = {m.x_max};
  // X/Y on more qubits than a tile holds: a sweep of its own over the state.
  const uint64_t lowmask = (1ull << low_bits_) - 1ull;
  p.solo = static_cast<uint32_t>(__builtin_popcountll(m.x | lowmask)) > tile_bits_;
  p.npairs = static_cast<uint32_t>(probes_queued_++);
  push(std::move(p));
  return slot;
}

std::vector<double> Engine::take_saved_expvals() {
  flush();
  std::vector<double> out;
  out.swap(saved_expvals_);
  return out;
}

std::vector<PauliTerm> Engine::upper_terms(uint32_t n, const std::vector<std::string> &labels,
                                           const std::vector<double> &coeffs) {
  if (n < 2) throw std::invalid_argument("aerhip: pauli_sum_to_upper needs two or more qubits");
  if (labels.size() != coeffs.size()) throw std::invalid_argument("aerhip: labels/coeffs size");
  std::vector<uint32_t> lower(n - 1);
  for (uint32_t q = 0; q + 1 < n; ++q) lower[q] = q;
  std::vector<PauliTerm> terms;
  for (size_t t = 0; t < labels.size(); ++t) {
    const PauliMasks m = pauli_masks(lower, labels[t]);
    const cd c = add_y_phase(m.num_y, cd(coeffs[t], 0.0));
    terms.push_back(PauliTerm{m.x, m.z, c.real(), c.imag()});
  }
  std::stable_sort(terms.begin(), terms.end(),
                   [](const PauliTerm &a, const PauliTerm &b) { return a.x < b.x; });
  return terms;
}

void Engine::pauli_sum_to_upper(const std::vector<std::string> &labels,
                                const std::vector<double> &coeffs) {
  const std::vector<PauliTerm> terms = upper_terms(n_, labels, coeffs);
  flush();
  AERHIP_CHECK(hipSetDevice(device_));
  PauliTerm *dt = static_cast<PauliTerm *>(scratch(std::max<size_t>(1, terms.size()) * sizeof(PauliTerm)));
  AERHIP_CHECK(hipMemcpy(dt, terms.data(), terms.size() * sizeof(PauliTerm), hipMemcpyHostToDevice));
  const uint64_t half = size_ >> 1;
  if (prec_ == AERHIP_PRECISION_DOUBLE)
    k_pauli_sum_upper<double><<<grid_for(half, 16384), kThreads, 0, stream_>>>(
        static_cast<cx<double> *>(sv_), half, dt, static_cast<uint32_t>(terms.size()));
  else
    k_pauli_sum_upper<float><<<grid_for(half, 16384), kThreads, 0, stream_>>>(
        static_cast<cx<float> *>(sv_), half, dt, static_cast<uint32_t>(terms.size()));
  AERHIP_CHECK(hipGetLastError());
  AERHIP_CHECK(hipStreamSynchronize(stream_));
}

double Engine::expval_pauli(const std::vector<uint32_t> &qubits, const std::string &pauli, cd phase) {
  check_qubits(qubits);
  const PauliMasks m = pauli_masks(qubits, pauli);
  if (m.x + m.z == 0) return norm();
  flush();
  AERHIP_CHECK(hipSetDevice(device_));
  const cd ph = add_y_phase(m.num_y, phase);
  return prec_ == AERHIP_PRECISION_DOUBLE ? expval_t<double>(m.x, m.z, m.x_max, ph)
                                          : expval_t<float>(m.x, m.z, m.x_max, ph);
}

}  // namespace aerhip

// ===========================================================================
// C ABI
// ===========================================================================

namespace aerhip {

// Runs the queues of `count` planning engines (same precision) as one batch:
// every circuit starts from |0>; the saved expectation values of circuit i
// are written to values[offset_i ..], in circuit order and slot order.
// `kernel_ms` receives the device time of the batch kernel.
void run_batch(Engine *const *engines, uint64_t count, int device, double *values,
               double *kernel_ms) {
  if (!count) return;
  if (!engines || !values) throw std::invalid_argument("aerhip: null argument");
  const int prec = engines[0]->precision();
  for (uint64_t i = 0; i < count; ++i)
    if (!engines[i]->plan_only() || engines[i]->precision() != prec)
      throw std::invalid_argument("aerhip: batch needs planning handles of one precision");
  // Plan every circuit on the host threads (the handles are independent),
  // then concatenate the programs in circuit order.
  std::vector<Program> parts(count);
  std::vector<uint32_t> nprobes(count);
  parallel_for(count, [&](uint64_t i) {
    nprobes[i] = static_cast<uint32_t>(engines[i]->probes_queued());
    engines[i]->emit_program(parts[i]);
  });
  // Offsets of each circuit's passes / ops / register ops / pool entries in
  // the concatenated program.
  std::vector<uint64_t> pass_at(count + 1, 0), op_at(count + 1, 0), rop_at(count + 1, 0),
      pool_at(count + 1, 0);
  std::vector<BatchCircuit> circ(count);
  std::vector<double> cost(count);
  uint64_t sv_off = 0, probe_off = 0;
  for (uint64_t i = 0; i < count; ++i) {
    const Program &p = parts[i];
    for (const PassDesc &d : p.passes)
      if (d.kind == 1)
        throw std::invalid_argument("aerhip: batched circuits take dense matrices on at most 5 qubits");
    pass_at[i + 1] = pass_at[i] + p.passes.size();
    op_at[i + 1] = op_at[i] + p.ops.size();
    rop_at[i + 1] = rop_at[i] + p.rops.size();
    pool_at[i + 1] = pool_at[i] + p.pool.size();
    BatchCircuit &c = circ[i];
    std::memset(&c, 0, sizeof(c));
    c.sv_off = sv_off;
    c.num_qubits = engines[i]->num_qubits();
    c.pass_begin = static_cast<uint32_t>(pass_at[i]);
    c.pass_count = static_cast<uint32_t>(p.passes.size());
    c.probe_base = static_cast<uint32_t>(probe_off);
    c.num_probes = nprobes[i];
    cost[i] = static_cast<double>(c.pass_count + 1) * static_cast<double>(1ull << c.num_qubits);
    sv_off += align_up(1ull << c.num_qubits, 256);
    probe_off += c.num_probes;
  }
  if (op_at[count] > UINT32_MAX || rop_at[count] > UINT32_MAX || pool_at[count] > UINT32_MAX)
    throw std::invalid_argument("aerhip: batch program too large");
  // Largest circuits first: workgroups start in index order.
  std::vector<uint64_t> order(count);
  for (uint64_t i = 0; i < count; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return cost[a] > cost[b]; });
  std::vector<BatchCircuit> sorted(count);
  for (uint64_t i = 0; i < count; ++i) sorted[i] = circ[order[i]];

  const size_t amp = prec == AERHIP_PRECISION_DOUBLE ? 16 : 8;
  const size_t circ_bytes = count * sizeof(BatchCircuit);
  const size_t pass_off = align_up(circ_bytes, 256);
  const size_t ops_off = align_up(pass_off + pass_at[count] * sizeof(PassDesc), 256);
  const size_t rops_off = align_up(ops_off + op_at[count] * sizeof(TileOp), 256);
  const size_t pool_off = align_up(rops_off + rop_at[count] * sizeof(ROp<double>), 256);
  const size_t total = pool_off + pool_at[count] * sizeof(cd) + 16;
  // The upload image, written in place: each circuit's part is copied to its
  // offsets with its indices rebased (pass op ranges, TileOp pool offsets,
  // the pool offsets and TileOp indices carried by register ops).
  std::unique_ptr<char[]> h(new char[total]);
  std::memcpy(h.get(), sorted.data(), circ_bytes);
  auto *hp = reinterpret_cast<PassDesc *>(h.get() + pass_off);
  auto *ho = reinterpret_cast<TileOp *>(h.get() + ops_off);
  auto *hpool = reinterpret_cast<cd *>(h.get() + pool_off);
  parallel_for(count, [&](uint64_t i) {
    Program &p = parts[i];
    const uint32_t ob = static_cast<uint32_t>(op_at[i]), rb = static_cast<uint32_t>(rop_at[i]),
                   pb = static_cast<uint32_t>(pool_at[i]);
    for (size_t k = 0; k < p.passes.size(); ++k) {
      PassDesc d = p.passes[k];
      d.op_begin += d.kind == 2 ? rb : ob;
      hp[pass_at[i] + k] = d;
    }
    for (size_t k = 0; k < p.ops.size(); ++k) {
      TileOp t = p.ops[k];
      t.off += pb;
      ho[op_at[i] + k] = t;
    }
    for (ROp<double> &r : p.rops) rebase_rop(r, pb, ob);
    if (prec == AERHIP_PRECISION_DOUBLE)
      std::memcpy(h.get() + rops_off + rop_at[i] * sizeof(ROp<double>), p.rops.data(),
                  p.rops.size() * sizeof(ROp<double>));
    else
      rops_to_float(p.rops.data(), p.rops.size(),
                    reinterpret_cast<ROp<float> *>(h.get() + rops_off) + rop_at[i]);
    if (!p.pool.empty()) std::memcpy(hpool + pool_at[i], p.pool.data(), p.pool.size() * sizeof(cd));
    std::vector<PassDesc>().swap(p.passes);  // free the part early
    std::vector<TileOp>().swap(p.ops);
    std::vector<ROp<double>>().swap(p.rops);
    std::vector<cd>().swap(p.pool);
  });

  AERHIP_CHECK(hipSetDevice(device));
  void *d_states = nullptr, *d_prog = nullptr;
  double *d_probes = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto release = [&] {
    if (st) (void)hipStreamSynchronize(st);
    if (d_states) (void)hipFree(d_states);
    if (d_prog) (void)hipFree(d_prog);
    if (d_probes) (void)hipFree(d_probes);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
  };
  try {
    AERHIP_CHECK(hipStreamCreate(&st));
    AERHIP_CHECK(hipEventCreate(&e0));
    AERHIP_CHECK(hipEventCreate(&e1));
    AERHIP_CHECK(hipMalloc(&d_states, sv_off * amp));
    AERHIP_CHECK(hipMalloc(&d_prog, total));
    AERHIP_CHECK(hipMalloc(reinterpret_cast<void **>(&d_probes), (probe_off + 1) * sizeof(double)));
    AERHIP_CHECK(hipMemcpyAsync(d_prog, h.get(), total, hipMemcpyHostToDevice, st));
    char *d = static_cast<char *>(d_prog);
    const auto *dc = reinterpret_cast<const BatchCircuit *>(d);
    const auto *dp = reinterpret_cast<const PassDesc *>(d + pass_off);
    const auto *dops = reinterpret_cast<const TileOp *>(d + ops_off);
    const auto *dpool = reinterpret_cast<const cx<double> *>(d + pool_off);
    const uint32_t grid = static_cast<uint32_t>(count);
    AERHIP_CHECK(hipEventRecord(e0, st));
    if (prec == AERHIP_PRECISION_DOUBLE)
      k_batch<double, kTileBitsC128, 3><<<grid, kThreads, 0, st>>>(
          static_cast<cx<double> *>(d_states), dc, dp,
          reinterpret_cast<const ROp<double> *>(d + rops_off), dops, dpool, d_probes);
    else
      k_batch<float, kTileBitsC64, 4><<<grid, kThreads, 0, st>>>(
          static_cast<cx<float> *>(d_states), dc, dp,
          reinterpret_cast<const ROp<float> *>(d + rops_off), dops, dpool, d_probes);
    AERHIP_CHECK(hipGetLastError());
    AERHIP_CHECK(hipEventRecord(e1, st));
    AERHIP_CHECK(hipMemcpyAsync(values, d_probes, probe_off * sizeof(double), hipMemcpyDeviceToHost, st));
    AERHIP_CHECK(hipStreamSynchronize(st));
    float ms = 0;
    AERHIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    if (kernel_ms) *kernel_ms = ms;
  } catch (...) {
    release();
    throw;
  }
  release();
}

// ---------------------------------------------------------------------------
// Many states of one size advancing together.  State b runs the program its
// planning engine b emits, and each pass index launches once for all states
// (launch_pass with count > 1: workgroup w takes tile w mod 2^ncomp of state
// w >> ncomp).  The parameter bindings of one circuit run this way: the
// reference binds parameters per binding and, on the GPU, batches the bound
// experiments (BatchConverter, src/transpile/batch_converter.hpp:54-157;
// batched_expval_pauli, qubitvector_thrust.hpp:2683-2713).
// ---------------------------------------------------------------------------
class MultiState {
 public:
  MultiState(uint32_t n, int prec, int device, std::vector<Engine *> planners)
      : n_(n), prec_(prec), device_(device), planners_(std::move(planners)) {
    if (planners_.empty()) throw std::invalid_argument("aerhip: no states");
    for (Engine *e : planners_)
      if (!e->plan_only() || e->num_qubits() != n || e->precision() != prec)
        throw std::invalid_argument("aerhip: states need planning handles of one size and precision");
    AERHIP_CHECK(hipSetDevice(device_));
    AERHIP_CHECK(hipStreamCreate(&st_));
    AERHIP_CHECK(hipEventCreate(&e0_));
    AERHIP_CHECK(hipEventCreate(&e1_));
    AERHIP_CHECK(hipMalloc(&states_, planners_.size() * (1ull << n_) * amp_bytes()));
  }
  ~MultiState() {
    (void)hipSetDevice(device_);
    if (st_) (void)hipStreamSynchronize(st_);
    if (states_) (void)hipFree(states_);
    if (d_prog_) (void)hipFree(d_prog_);
    if (d_probes_) (void)hipFree(d_probes_);
    if (d_terms_) (void)hipFree(d_terms_);
    if (e0_) (void)hipEventDestroy(e0_);
    if (e1_) (void)hipEventDestroy(e1_);
    if (st_) (void)hipStreamDestroy(st_);
  }
  MultiState(const MultiState &) = delete;
  MultiState &operator=(const MultiState &) = delete;

  uint64_t count() const { return planners_.size(); }
  double kernel_ms() const { return kernel_ms_; }

  // |0...0> in every state.
  void initialize() {
    AERHIP_CHECK(hipSetDevice(device_));
    const uint64_t size = 1ull << n_;
    for (uint64_t b = 0; b < count(); ++b) {
      if (prec_ == AERHIP_PRECISION_DOUBLE)
        k_fill_basis<double><<<grid_for(size, 16384), kThreads, 0, st_>>>(state<double>(b), size, 0);
      else
        k_fill_basis<float><<<grid_for(size, 16384), kThreads, 0, st_>>>(state<float>(b), size, 0);
    }
    AERHIP_CHECK(hipGetLastError());
  }

  // With the top qubit as a flag, sum_t c_t P_t psi into the flagged half of
  // every state (Engine::pauli_sum_to_upper on each).
  void pauli_sum_to_upper(const std::vector<PauliTerm> &terms) {
    AERHIP_CHECK(hipSetDevice(device_));
    AERHIP_CHECK(hipStreamSynchronize(st_));
    if (d_terms_) AERHIP_CHECK(hipFree(d_terms_));
    d_terms_ = nullptr;
    AERHIP_CHECK(hipMalloc(&d_terms_, std::max<size_t>(1, terms.size()) * sizeof(PauliTerm)));
    AERHIP_CHECK(hipMemcpy(d_terms_, terms.data(), terms.size() * sizeof(PauliTerm), hipMemcpyHostToDevice));
    const uint64_t half = 1ull << (n_ - 1);
    const auto *dt = static_cast<const PauliTerm *>(d_terms_);
    const uint32_t nt = static_cast<uint32_t>(terms.size());
    for (uint64_t b = 0; b < count(); ++b) {
      if (prec_ == AERHIP_PRECISION_DOUBLE)
        k_pauli_sum_upper<double><<<grid_for(half, 16384), kThreads, 0, st_>>>(state<double>(b), half, dt, nt);
      else
        k_pauli_sum_upper<float><<<grid_for(half, 16384), kThreads, 0, st_>>>(state<float>(b), half, dt, nt);
    }
    AERHIP_CHECK(hipGetLastError());
  }

  // Runs what every planning engine has queued and returns the saved values,
  // state-major (per_state values of state 0, then of state 1, ...).  States
  // whose programs have the same pass shapes run together; any state whose
  // shape differs from state 0's runs on its own.
  std::vector<double> flush(uint64_t *per_state) {
    const uint64_t cnt = count();
    std::vector<Program> parts(cnt);
    std::vector<uint64_t> np(cnt);
    parallel_for(cnt, [&](uint64_t i) {
      np[i] = planners_[i]->probes_queued();
      planners_[i]->emit_program(parts[i]);
    });
    const uint64_t pps = np[0];
    for (uint64_t i = 1; i < cnt; ++i)
      if (np[i] != pps) throw std::invalid_argument("aerhip: states save different numbers of values");
    *per_state = pps;
    std::vector<double> values(cnt * pps, 0.0);
    if (pps > 0xFFFFFFFFull) throw std::invalid_argument("aerhip: too many saved values");
    auto same_shape = [&](const Program &a, const Program &b) {
      if (a.passes.size() != b.passes.size()) return false;
      for (size_t k = 0; k < a.passes.size(); ++k)
        if (a.passes[k].kind != b.passes[k].kind || a.passes[k].ncomp != b.passes[k].ncomp ||
            a.passes[k].kind == 1)
          return false;
      return true;
    };
    bool together = true;
    for (uint64_t i = 1; i < cnt && together; ++i) together = same_shape(parts[0], parts[i]);
    AERHIP_CHECK(hipSetDevice(device_));
    AERHIP_CHECK(hipEventRecord(e0_, st_));
    if (together) {
      run(parts, 0, cnt, pps, values.data());
    } else {
      for (uint64_t i = 0; i < cnt; ++i) run(parts, i, 1, pps, values.data() + i * pps);
    }
    AERHIP_CHECK(hipEventRecord(e1_, st_));
    AERHIP_CHECK(hipStreamSynchronize(st_));
    float ms = 0;
    AERHIP_CHECK(hipEventElapsedTime(&ms, e0_, e1_));
    kernel_ms_ += ms;
    return values;
  }

 private:
  size_t amp_bytes() const { return prec_ == AERHIP_PRECISION_DOUBLE ? 16 : 8; }
  template <typename T>
  cx<T> *state(uint64_t b) {
    return static_cast<cx<T> *>(states_) + b * (1ull << n_);
  }

  // Runs the programs of states [first, first + cnt) (same pass shapes) and
  // writes their saved values to values (state-major).
  void run(const std::vector<Program> &parts, uint64_t first, uint64_t cnt, uint64_t pps, double *values) {
    const size_t P = parts[first].passes.size();
    // concatenated program, state-major, indices rebased per state
    std::vector<PassDesc> passes;
    std::vector<TileOp> ops;
    std::vector<ROp<double>> rops;
    std::vector<cd> pool;
    std::vector<char> lean(P, 1), wide(P, 1), diag(P, 0);
    for (uint64_t b = first; b < first + cnt; ++b) {
      const Program &p = parts[b];
      const uint32_t ob = static_cast<uint32_t>(ops.size()), rb = static_cast<uint32_t>(rops.size()),
                     pb = static_cast<uint32_t>(pool.size());
      for (size_t k = 0; k < P; ++k) {
        bool l = false, w = false, dg = false;
        pass_variant(p.passes[k], p.rops, l, w, dg);
        lean[k] = lean[k] && l;
        wide[k] = wide[k] && w;
        diag[k] = diag[k] || dg;
        PassDesc d = p.passes[k];
        d.op_begin += d.kind == 2 ? rb : ob;
        passes.push_back(d);
      }
      for (TileOp t : p.ops) {
        t.off += pb;
        ops.push_back(t);
      }
      for (ROp<double> r : p.rops) {
        rebase_rop(r, pb, ob);
        rops.push_back(r);
      }
      pool.insert(pool.end(), p.pool.begin(), p.pool.end());
      if (pool.size() > 0xFFFFFFFFull || ops.size() > 0xFFFFFFFFull || rops.size() > 0xFFFFFFFFull)
        throw std::invalid_argument("aerhip: batched program too large");
    }
    const size_t pass_bytes = passes.size() * sizeof(PassDesc);
    const size_t ops_off = align_up(pass_bytes, 256);
    const size_t rops_off = align_up(ops_off + ops.size() * sizeof(TileOp), 256);
    const size_t pool_off = align_up(rops_off + rops.size() * sizeof(ROp<double>), 256);
    const size_t total = pool_off + pool.size() * sizeof(cd) + 16;
    std::vector<char> h(total);
    std::memcpy(h.data(), passes.data(), pass_bytes);
    std::memcpy(h.data() + ops_off, ops.data(), ops.size() * sizeof(TileOp));
    if (prec_ == AERHIP_PRECISION_DOUBLE)
      std::memcpy(h.data() + rops_off, rops.data(), rops.size() * sizeof(ROp<double>));
    else
      rops_to_float(rops.data(), rops.size(), reinterpret_cast<ROp<float> *>(h.data() + rops_off));
    if (!pool.empty()) std::memcpy(h.data() + pool_off, pool.data(), pool.size() * sizeof(cd));
    if (total > prog_cap_) {
      AERHIP_CHECK(hipStreamSynchronize(st_));
      if (d_prog_) AERHIP_CHECK(hipFree(d_prog_));
      d_prog_ = nullptr;
      prog_cap_ = align_up(total, 1 << 16);
      AERHIP_CHECK(hipMalloc(&d_prog_, prog_cap_));
    }
    // pageable source: the copy is complete when hipMemcpyAsync returns
    AERHIP_CHECK(hipMemcpyAsync(d_prog_, h.data(), total, hipMemcpyHostToDevice, st_));
    AERHIP_CHECK(hipStreamSynchronize(st_));
    const uint32_t tb = planners_[first]->tile_bits();
    const uint64_t ntiles = 1ull << (n_ - tb);
    if (pps) {
      const size_t need = (cnt * pps * ntiles + cnt * pps) * sizeof(double);
      if (need > probes_cap_) {
        if (d_probes_) AERHIP_CHECK(hipFree(d_probes_));
        d_probes_ = nullptr;
        probes_cap_ = align_up(need, 1 << 20);
        AERHIP_CHECK(hipMalloc(reinterpret_cast<void **>(&d_probes_), probes_cap_));
      }
    }
    char *d = static_cast<char *>(d_prog_);
    const DeviceProgram dp{reinterpret_cast<const PassDesc *>(d), reinterpret_cast<const TileOp *>(d + ops_off),
                           d + rops_off, reinterpret_cast<const cx<double> *>(d + pool_off),
                           pps ? d_probes_ : nullptr, static_cast<uint32_t>(pps)};
    void *sv = static_cast<char *>(states_) + first * (1ull << n_) * amp_bytes();
    for (size_t k = 0; k < P; ++k)
      launch_pass(prec_, planners_[first]->reg_bits(), n_, tb, parts[first].passes[k],
                  static_cast<uint32_t>(k), lean[k] != 0, wide[k] != 0, diag[k] != 0, dp, sv, cnt,
                  static_cast<uint32_t>(P), st_);
    if (pps) {
      double *sums = d_probes_ + cnt * pps * ntiles;
      k_sum_rows<<<grid_for(cnt * pps, 1u << 30), kThreads, 0, st_>>>(d_probes_, cnt * pps,
                                                                      static_cast<uint32_t>(ntiles), sums);
      AERHIP_CHECK(hipGetLastError());
      AERHIP_CHECK(hipMemcpyAsync(values, sums, cnt * pps * sizeof(double), hipMemcpyDeviceToHost, st_));
    }
    AERHIP_CHECK(hipStreamSynchronize(st_));
  }

  uint32_t n_;
  int prec_;
  int device_;
  std::vector<Engine *> planners_;
  void *states_ = nullptr;
  hipStream_t st_ = nullptr;
  hipEvent_t e0_ = nullptr, e1_ = nullptr;
  void *d_prog_ = nullptr;
  size_t prog_cap_ = 0;
  double *d_probes_ = nullptr;
  size_t probes_cap_ = 0;
  void *d_terms_ = nullptr;
  double kernel_ms_ = 0;
};

}  // namespace aerhip

using aerhip::cd;
using aerhip::Engine;

struct aerhip_qv {
  std::unique_ptr<Engine> eng;
};

namespace {

thread_local std::string g_last_error;

template <typename F>
int guard(F &&f) {
  try {
    f();
    return 0;
  } catch (const std::invalid_argument &e) {
    g_last_error = e.what();
    return 2;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return 1;
  } catch (...) {
    g_last_error = "aerhip: unknown error";
    return 1;
  }
}

Engine &E(aerhip_qv *qv) {
  if (!qv || !qv->eng) throw std::invalid_argument("aerhip: null statevector handle");
  return *qv->eng;
}

std::vector<uint32_t> qvec(const uint64_t *q, uint64_t n) {
  if (n && !q) throw std::invalid_argument("aerhip: null qubit list");
  std::vector<uint32_t> v(n);
  for (uint64_t i = 0; i < n; ++i) {
    if (q[i] >= 64) throw std::invalid_argument("aerhip: qubit index out of range");
    v[i] = static_cast<uint32_t>(q[i]);
  }
  return v;
}

std::vector<cd> cvec(const double *p, uint64_t n) {
  if (n && !p) throw std::invalid_argument("aerhip: null complex buffer");
  std::vector<cd> v(n);
  for (uint64_t i = 0; i < n; ++i) v[i] = cd(p[2 * i], p[2 * i + 1]);
  return v;
}

}  // namespace

extern "C" {

const char *aerhip_last_error(void) { return g_last_error.c_str(); }

__attribute__((visibility("hidden"))) void aerhip_detail_set_error(const char *msg) {
  g_last_error = msg ? msg : "";
}

int aerhip_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int aerhip_device_memory(int device, uint64_t *free_out, uint64_t *total_out) {
  return guard([&] {
    if (!free_out || !total_out) throw std::invalid_argument("aerhip: null output");
    int cur = 0;
    AERHIP_CHECK(hipGetDevice(&cur));
    AERHIP_CHECK(hipSetDevice(device));
    size_t f = 0, t = 0;
    const hipError_t e = hipMemGetInfo(&f, &t);
    (void)hipSetDevice(cur);
    AERHIP_CHECK(e);
    *free_out = f;
    *total_out = t;
  });
}

const char *aerhip_version(void) { return "aerhip 0.1.0 (gfx950)"; }

// Internal to the library: planning handles and the batch runner used by
// aerhip_batch_pauli_expval (controller.cpp).
__attribute__((visibility("hidden"))) aerhip_qv *aerhip_detail_qv_create_planner(
    uint64_t num_qubits, int precision, int device) {
  aerhip_qv *h = nullptr;
  const int rc = guard([&] {
    auto eng = std::make_unique<Engine>(static_cast<uint32_t>(num_qubits), precision, device,
                                        Engine::PlanOnly{});
    h = new aerhip_qv{std::move(eng)};
  });
  return rc == 0 ? h : nullptr;
}

// Many states of one size (aerhip::MultiState) with one planning handle per
// state, used by the parameter-binding runners (controller.cpp).
struct aerhip_multi {
  std::vector<aerhip_qv> planners;
  std::unique_ptr<aerhip::MultiState> m;
};

__attribute__((visibility("hidden"))) aerhip_multi *aerhip_detail_multi_create(uint64_t num_qubits,
                                                                                int precision, int device,
                                                                                uint64_t count) {
  aerhip_multi *h = nullptr;
  const int rc = guard([&] {
    if (!count) throw std::invalid_argument("aerhip: no states");
    auto m = std::make_unique<aerhip_multi>();
    m->planners.resize(count);
    std::vector<Engine *> engines(count);
    for (uint64_t i = 0; i < count; ++i) {
      m->planners[i].eng = std::make_unique<Engine>(static_cast<uint32_t>(num_qubits), precision, device,
                                                    Engine::PlanOnly{});
      engines[i] = m->planners[i].eng.get();
    }
    m->m = std::make_unique<aerhip::MultiState>(static_cast<uint32_t>(num_qubits), precision, device,
                                                std::move(engines));
    h = m.release();
  });
  return rc == 0 ? h : nullptr;
}

__attribute__((visibility("hidden"))) void aerhip_detail_multi_destroy(aerhip_multi *m) { delete m; }

__attribute__((visibility("hidden"))) aerhip_qv *aerhip_detail_multi_planner(aerhip_multi *m, uint64_t i) {
  if (!m || i >= m->planners.size()) return nullptr;
  return &m->planners[i];
}

__attribute__((visibility("hidden"))) int aerhip_detail_multi_initialize(aerhip_multi *m) {
  return guard([&] {
    if (!m) throw std::invalid_argument("aerhip: null argument");
    m->m->initialize();
  });
}

__attribute__((visibility("hidden"))) int aerhip_detail_multi_flush(aerhip_multi *m, double *values,
                                                                    uint64_t capacity, uint64_t *per_state) {
  return guard([&] {
    if (!m || !per_state) throw std::invalid_argument("aerhip: null argument");
    const std::vector<double> v = m->m->flush(per_state);
    if (v.size() > capacity) throw std::invalid_argument("aerhip: saved values exceed the capacity");
    if (!v.empty()) std::memcpy(values, v.data(), v.size() * sizeof(double));
  });
}

__attribute__((visibility("hidden"))) int aerhip_detail_multi_pauli_sum_to_upper(aerhip_multi *m,
                                                                                  const char *labels,
                                                                                  const double *coeffs,
                                                                                  uint64_t nterms) {
  return guard([&] {
    if (!m || (nterms && (!labels || !coeffs))) throw std::invalid_argument("aerhip: null argument");
    const uint32_t n = m->planners[0].eng->num_qubits();
    std::vector<std::string> labs(nterms);
    for (uint64_t t = 0; t < nterms; ++t) labs[t].assign(labels + t * (n - 1), n - 1);
    m->m->pauli_sum_to_upper(Engine::upper_terms(n, labs, std::vector<double>(coeffs, coeffs + nterms)));
  });
}

__attribute__((visibility("hidden"))) double aerhip_detail_multi_kernel_ms(aerhip_multi *m) {
  return m ? m->m->kernel_ms() : 0.0;
}

// Runs the queues of `count` planning handles (same precision) as one batch
// (aerhip::run_batch).
__attribute__((visibility("hidden"))) int aerhip_detail_batch_run(aerhip_qv *const *qvs,
                                                                   uint64_t count, int device,
                                                                   double *values,
                                                                   double *kernel_ms) {
  return guard([&] {
    if (!count) return;
    if (!qvs) throw std::invalid_argument("aerhip: null argument");
    std::vector<Engine *> engines(count);
    for (uint64_t i = 0; i < count; ++i) engines[i] = &E(qvs[i]);
    aerhip::run_batch(engines.data(), count, device, values, kernel_ms);
  });
}

aerhip_qv *aerhip_qv_create(uint64_t num_qubits, int precision, int device) {
  aerhip_qv *h = nullptr;
  const int rc = guard([&] {
    auto eng = std::make_unique<Engine>(static_cast<uint32_t>(num_qubits), precision, device);
    h = new aerhip_qv{std::move(eng)};
  });
  return rc == 0 ? h : nullptr;
}

void aerhip_qv_destroy(aerhip_qv *qv) { delete qv; }

int aerhip_qv_num_qubits(const aerhip_qv *qv, uint64_t *out) {
  return guard([&] {
    if (!qv || !qv->eng || !out) throw std::invalid_argument("aerhip: null argument");
    *out = qv->eng->num_qubits();
  });
}

uint64_t aerhip_required_memory_mb(uint64_t num_qubits, int precision) {
  // QubitVector::required_memory_mb (qubitvector.hpp:984-992)
  const int64_t unit = precision == AERHIP_PRECISION_SINGLE ? 3 : 4;
  const int64_t shift = std::max<int64_t>(0, static_cast<int64_t>(num_qubits) + unit - 20);
  if (shift >= 63) return UINT64_MAX;
  return 1ull << shift;
}

int aerhip_qv_initialize(aerhip_qv *qv) { return guard([&] { E(qv).initialize(); }); }
int aerhip_qv_zero(aerhip_qv *qv) { return guard([&] { E(qv).zero(); }); }

int aerhip_qv_initialize_from_vector(aerhip_qv *qv, const double *vec, uint64_t len) {
  return guard([&] {
    if (!vec) throw std::invalid_argument("aerhip: null vector");
    E(qv).initialize_from_vector(reinterpret_cast<const cd *>(vec), len);
  });
}

int aerhip_qv_initialize_from_device(aerhip_qv *qv, const void *src, uint64_t len,
                                     int src_precision) {
  return guard([&] {
    if (!src) throw std::invalid_argument("aerhip: null device pointer");
    E(qv).initialize_from_device(src, len, src_precision);
  });
}

int aerhip_qv_initialize_random(aerhip_qv *qv, uint64_t seed) {
  return guard([&] { E(qv).initialize_random(seed); });
}

int aerhip_qv_device_pointer(aerhip_qv *qv, void **out) {
  return guard([&] {
    if (!out) throw std::invalid_argument("aerhip: null argument");
    E(qv).synchronize();
    *out = E(qv).device_data();
  });
}

int aerhip_qv_copy_to_vector(aerhip_qv *qv, double *out, uint64_t len) {
  return guard([&] {
    if (!out) throw std::invalid_argument("aerhip: null output");
    E(qv).copy_to_vector(reinterpret_cast<cd *>(out), len);
  });
}

int aerhip_qv_initialize_component(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq,
                                   const double *state) {
  return guard([&] { E(qv).initialize_component(qvec(qubits, nq), cvec(state, 1ull << nq)); });
}

int aerhip_qv_synchronize(aerhip_qv *qv) { return guard([&] { E(qv).synchronize(); }); }

int aerhip_qv_set_sample_measure_index_size(aerhip_qv *qv, int n) {
  return guard([&] { E(qv).set_sample_measure_index_size(n); });
}

int aerhip_qv_set_tile_bits(aerhip_qv *qv, int tile_bits, int low_bits) {
  return guard([&] {
    if (tile_bits > 0) E(qv).set_tile_bits(static_cast<uint32_t>(tile_bits));
    if (low_bits > 0) E(qv).set_low_bits(static_cast<uint32_t>(low_bits));
  });
}

int aerhip_qv_stats(aerhip_qv *qv, uint64_t *out3) {  // 4 values
  return guard([&] {
    if (!out3) throw std::invalid_argument("aerhip: null argument");
    out3[0] = E(qv).stats().passes;
    out3[1] = E(qv).stats().launches;
    out3[2] = E(qv).stats().gates;
    out3[3] = E(qv).stats().gathers;
  });
}

int aerhip_qv_set_block_fusion(aerhip_qv *qv, int width, int cost_check) {
  return guard([&] {
    if (width < 0) throw std::invalid_argument("aerhip: block fusion width must be 0 or 2..5");
    E(qv).set_block_fusion(static_cast<uint32_t>(width), cost_check != 0);
  });
}

int aerhip_qv_graph_stats(aerhip_qv *qv, uint64_t *captures, uint64_t *replays) {
  return guard([&] {
    if (!captures || !replays) throw std::invalid_argument("aerhip: null argument");
    *captures = E(qv).stats().graph_captures;
    *replays = E(qv).stats().graph_replays;
  });
}

int aerhip_save_expval_pauli(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq,
                             const char *pauli, double phase_re, double phase_im,
                             uint64_t *slot_out) {
  return guard([&] {
    if (!pauli) throw std::invalid_argument("aerhip: null Pauli label");
    const uint64_t slot = E(qv).save_expval_pauli(qvec(qubits, nq), pauli, cd(phase_re, phase_im));
    if (slot_out) *slot_out = slot;
  });
}

int aerhip_saved_expvals(aerhip_qv *qv, double *out, uint64_t capacity, uint64_t *count) {
  return guard([&] {
    const auto v = E(qv).take_saved_expvals();
    if (count) *count = v.size();
    if (v.size() > capacity) throw std::invalid_argument("aerhip: output buffer too small");
    if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(double));
  });
}

int aerhip_qv_pauli_sum_to_upper_half(aerhip_qv *qv, const char *labels, const double *coeffs,
                                      uint64_t nterms) {
  return guard([&] {
    const uint64_t w = E(qv).num_qubits() - 1;
    std::vector<std::string> labs(nterms);
    for (uint64_t t = 0; t < nterms; ++t) labs[t].assign(labels + t * w, w);
    E(qv).pauli_sum_to_upper(labs, std::vector<double>(coeffs, coeffs + nterms));
  });
}

int aerhip_qv_set_fusion(aerhip_qv *qv, int enable, int max_qubit, int threshold) {
  return guard([&] {
    if (threshold < 0) throw std::invalid_argument("aerhip: fusion_threshold must be >= 0");
    E(qv).set_fusion(enable != 0, static_cast<uint32_t>(max_qubit), static_cast<uint32_t>(threshold));
  });
}

int aerhip_qv_fused_count(aerhip_qv *qv, uint64_t *out) {
  return guard([&] {
    if (!out) throw std::invalid_argument("aerhip: null output");
    *out = E(qv).stats().fused;
  });
}

int aerhip_qv_set_register_bits(aerhip_qv *qv, int reg_bits) {
  return guard([&] { E(qv).set_reg_bits(static_cast<uint32_t>(reg_bits)); });
}

int aerhip_qv_set_event_timing(aerhip_qv *qv, int enable) {
  return guard([&] { E(qv).set_event_timing(enable != 0); });
}

int aerhip_qv_pass_timing(aerhip_qv *qv, double *ms_out, uint64_t *launches_out) {
  return guard([&] {
    if (!ms_out || !launches_out) throw std::invalid_argument("aerhip: null argument");
    *ms_out = E(qv).stats().pass_ms;
    *launches_out = E(qv).stats().timed_launches;
  });
}

int aerhip_apply_matrix(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq, const double *mat) {
  return guard([&] {
    if (nq > 20) throw std::runtime_error("Maximum size of apply matrix is a 20-qubit matrix.");
    E(qv).apply_matrix(qvec(qubits, nq), cvec(mat, 1ull << (2 * nq)));
  });
}

int aerhip_apply_diagonal_matrix(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq,
                                 const double *diag) {
  return guard([&] {
    if (nq > 40) throw std::invalid_argument("aerhip: too many qubits");
    E(qv).apply_diagonal_matrix(qvec(qubits, nq), cvec(diag, 1ull << nq));
  });
}

int aerhip_apply_mcx(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq) {
  return guard([&] { E(qv).apply_mcx(qvec(qubits, nq)); });
}
int aerhip_apply_mcy(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq) {
  return guard([&] { E(qv).apply_mcy(qvec(qubits, nq)); });
}
int aerhip_apply_mcswap(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq) {
  return guard([&] { E(qv).apply_mcswap(qvec(qubits, nq)); });
}
int aerhip_apply_mcphase(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq, double re,
                         double im) {
  return guard([&] { E(qv).apply_mcphase(qvec(qubits, nq), cd(re, im)); });
}
int aerhip_apply_mcu(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq, const double *mat) {
  return guard([&] { E(qv).apply_mcu(qvec(qubits, nq), cvec(mat, 4)); });
}
int aerhip_apply_rotation(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq, int axis,
                          double theta) {
  return guard([&] { E(qv).apply_rotation(qvec(qubits, nq), axis, theta); });
}
int aerhip_apply_pauli(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq, const char *pauli,
                       double re, double im) {
  return guard([&] {
    if (!pauli) throw std::invalid_argument("aerhip: null Pauli label");
    E(qv).apply_pauli(qvec(qubits, nq), pauli, cd(re, im));
  });
}
int aerhip_apply_permutation_matrix(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq,
                                    const uint64_t *pairs, uint64_t npairs) {
  return guard([&] {
    std::vector<std::pair<uint64_t, uint64_t>> pv(npairs);
    for (uint64_t i = 0; i < npairs; ++i) pv[i] = {pairs[2 * i], pairs[2 * i + 1]};
    E(qv).apply_permutation_matrix(qvec(qubits, nq), pv);
  });
}
int aerhip_apply_multiplexer(aerhip_qv *qv, const uint64_t *controls, uint64_t nc,
                             const uint64_t *targets, uint64_t nt, const double *mat) {
  return guard([&] {
    if (nc + nt > 30) throw std::invalid_argument("aerhip: multiplexer too large");
    const uint64_t dim = 1ull << (nc + nt);
    E(qv).apply_multiplexer(qvec(controls, nc), qvec(targets, nt), cvec(mat, dim << nt));
  });
}
int aerhip_apply_multi_swaps(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq) {
  return guard([&] { E(qv).apply_multi_swaps(qvec(qubits, nq)); });
}

int aerhip_qv_inner_product(aerhip_qv *a, aerhip_qv *b, double *re, double *im) {
  return guard([&] {
    if (!a || !b || !re || !im) throw std::invalid_argument("aerhip: null argument");
    const cd v = E(a).inner_product(E(b));
    *re = v.real();
    *im = v.imag();
  });
}

int aerhip_norm_matrix(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq, const double *mat,
                       double *out) {
  return guard([&] {
    if (!out) throw std::invalid_argument("aerhip: null output");
    if (nq > 10) throw std::invalid_argument("aerhip: norm of a matrix on more than 10 qubits");
    *out = E(qv).norm_matrix(qvec(qubits, nq), cvec(mat, 1ull << (2 * nq)));
  });
}

int aerhip_norm_diagonal(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq, const double *diag,
                         double *out) {
  return guard([&] {
    if (!out) throw std::invalid_argument("aerhip: null output");
    if (nq > 10) throw std::invalid_argument("aerhip: norm of a diagonal on more than 10 qubits");
    const std::vector<cd> d = cvec(diag, 1ull << nq);
    const uint64_t D = 1ull << nq;
    std::vector<cd> m(D * D, 0.0);
    for (uint64_t i = 0; i < D; ++i) m[i + D * i] = d[i];
    *out = E(qv).norm_matrix(qvec(qubits, nq), m);
  });
}

int aerhip_qv_get_state(aerhip_qv *qv, uint64_t pos, double *out2) {
  return guard([&] {
    if (!out2) throw std::invalid_argument("aerhip: null output");
    const cd v = E(qv).get_state(pos);
    out2[0] = v.real();
    out2[1] = v.imag();
  });
}

int aerhip_qv_set_state(aerhip_qv *qv, uint64_t pos, double re, double im) {
  return guard([&] { E(qv).set_state(pos, cd(re, im)); });
}

int aerhip_norm(aerhip_qv *qv, double *out) {
  return guard([&] {
    if (!out) throw std::invalid_argument("aerhip: null output");
    *out = E(qv).norm();
  });
}

int aerhip_probability(aerhip_qv *qv, uint64_t outcome, double *out) {
  return guard([&] {
    if (!out) throw std::invalid_argument("aerhip: null output");
    *out = E(qv).probability(outcome);
  });
}

int aerhip_probabilities(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq, double *out) {
  return guard([&] {
    if (!out) throw std::invalid_argument("aerhip: null output");
    const auto v = E(qv).probabilities(qvec(qubits, nq));
    std::memcpy(out, v.data(), v.size() * sizeof(double));
  });
}

int aerhip_sample_measure(aerhip_qv *qv, const double *rnds, uint64_t shots, uint64_t *out) {
  return guard([&] {
    if (shots && (!rnds || !out)) throw std::invalid_argument("aerhip: null buffer");
    const auto v = E(qv).sample_measure(std::vector<double>(rnds, rnds + shots));
    std::memcpy(out, v.data(), v.size() * sizeof(uint64_t));
  });
}

int aerhip_expval_pauli(aerhip_qv *qv, const uint64_t *qubits, uint64_t nq, const char *pauli,
                        double re, double im, double *out) {
  return guard([&] {
    if (!pauli || !out) throw std::invalid_argument("aerhip: null argument");
    *out = E(qv).expval_pauli(qvec(qubits, nq), pauli, cd(re, im));
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Pauli expectation value that allocates nothing.  Its workgroup partials and
// nothing else live in `scratch`, a file-scope device array: one copy per GPU,
// held until the process ends.  Host code never passes `scratch` itself (its
// host-side name is not a device address): the device address comes from
// hipGetSymbolAddress, once per device, with that device current.
// ---------------------------------------------------------------------------
namespace {
constexpr uint32_t kNoAllocPartials = 4096;  // workgroup partials: the grid cap of expval_t
constexpr int kNoAllocMaxDevices = 64;
}  // namespace

static __device__ __align__(256) double scratch[kNoAllocPartials];

namespace {

std::mutex g_scratch_mu;
double *g_scratch_ptr[kNoAllocMaxDevices] = {};

// The device address of `scratch` on the current device (looked up on the
// first call on that device).
double *scratch_on_current_device() {
  int dev = 0;
  AERHIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) throw std::runtime_error("aerhip: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("aerhip: scratch symbol address: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

int aerhip_expval_pauli_noalloc(const void *sv, uint64_t num_qubits, int precision,
                                const uint64_t *qubits, uint64_t nq, const char *pauli,
                                double phase_re, double phase_im, double *out, void *stream) {
  using namespace aerhip;
  return guard([&] {
    if (!sv || !out || !pauli || (nq && !qubits)) throw std::invalid_argument("aerhip: null argument");
    if (precision != AERHIP_PRECISION_DOUBLE && precision != AERHIP_PRECISION_SINGLE)
      throw std::invalid_argument("aerhip: unknown precision");
    if (num_qubits < 1 || num_qubits > 48)
      throw std::invalid_argument("aerhip: expval_pauli_noalloc takes 1 to 48 qubits");
    if (std::strlen(pauli) != nq) throw std::invalid_argument("aerhip: Pauli label length mismatch");
    // pauli_masks_and_phase (qubitvector.hpp:2236-2272): character N-1-i acts on qubits[i]
    uint64_t xm = 0, zm = 0, ny = 0;
    uint32_t xmax = 0;
    for (uint64_t i = 0; i < nq; ++i) {
      if (qubits[i] >= num_qubits) throw std::invalid_argument("aerhip: qubit index out of range");
      const uint64_t b = 1ull << qubits[i];
      switch (pauli[nq - 1 - i]) {
        case 'I': break;
        case 'X': xm |= b; xmax = std::max(xmax, static_cast<uint32_t>(qubits[i])); break;
        case 'Z': zm |= b; break;
        case 'Y':
          xm |= b; zm |= b; ++ny;
          xmax = std::max(xmax, static_cast<uint32_t>(qubits[i]));
          break;
        default: throw std::invalid_argument("aerhip: invalid Pauli character");
      }
    }
    const cd ph = add_y_phase(ny, cd(phase_re, phase_im));
    const uint64_t size = 1ull << num_qubits;
    const uint32_t grid = grid_for(xm ? size / 2 : size, kNoAllocPartials);
    if (grid > kNoAllocPartials) throw std::logic_error("aerhip: partials exceed the scratch array");
    double *part = scratch_on_current_device();
    hipStream_t st = static_cast<hipStream_t>(stream);
    if ((xm | zm) == 0) {
      // an all-identity label gives the norm (expval_pauli returns norm() there)
      if (precision == AERHIP_PRECISION_DOUBLE)
        k_sumsq<double><<<grid, kThreads, 0, st>>>(static_cast<const cx<double> *>(sv), size, part);
      else
        k_sumsq<float><<<grid, kThreads, 0, st>>>(static_cast<const cx<float> *>(sv), size, part);
    } else if (precision == AERHIP_PRECISION_DOUBLE)
      k_expval_pauli<double><<<grid, kThreads, 0, st>>>(static_cast<const cx<double> *>(sv), size, xm, zm,
                                                        xmax, ph.real(), ph.imag(), part);
    else
      k_expval_pauli<float><<<grid, kThreads, 0, st>>>(static_cast<const cx<float> *>(sv), size, xm, zm,
                                                       xmax, ph.real(), ph.imag(), part);
    k_sum_rows<<<1, kThreads, 0, st>>>(part, 1, grid, out);
    AERHIP_CHECK(hipGetLastError());
  });
}

}  // extern "C"
