// Shared device/host definitions of the MI355X statevector engine.
//
// The amplitude layout is the reference's: interleaved complex values, bit q
// of the index is qubit q (src/simulators/statevector/qubitvector.hpp:61-651).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace aerhip {

constexpr int kThreads = 256;  // workgroup size of every sweep kernel (4 wave64)

// Complex amplitude stored as an aligned pair so that a c128 amplitude is one
// 16-byte load and a c64 amplitude one 8-byte load.
template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cx<T> mk(T a, T b) {
  cx<T> r;
  r.x = a;
  r.y = b;
  return r;
}

template <typename T>
__host__ __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <typename T>
__host__ __device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) {
  return mk<T>(a.x + b.x, a.y + b.y);
}

// a*b + c, accumulated in the order the reference's transformer uses
// (data[i] += mat * cache, transformer.hpp:159).
template <typename T>
__host__ __device__ __forceinline__ cx<T> cfma(cx<T> a, cx<T> b, cx<T> c) {
  return mk<T>(c.x + (a.x * b.x - a.y * b.y), c.y + (a.x * b.y + a.y * b.x));
}

template <typename T>
__host__ __device__ __forceinline__ cx<T> cneg(cx<T> a) {
  return mk<T>(-a.x, -a.y);
}

template <typename T>
__host__ __device__ __forceinline__ cx<T> to_t(cx<double> a) {
  return mk<T>(static_cast<T>(a.x), static_cast<T>(a.y));
}

[[noreturn]] inline void hip_fail(hipError_t e, const char *what, const char *file,
                                  int line) {
  throw std::runtime_error(std::string("HIP error in ") + what + " (" + file + ":" +
                           std::to_string(line) + "): " + hipGetErrorString(e));
}

#define AERHIP_CHECK(expr)                                      \
  do {                                                          \
    hipError_t aerhip_e_ = (expr);                              \
    if (aerhip_e_ != hipSuccess)                                \
      ::aerhip::hip_fail(aerhip_e_, #expr, __FILE__, __LINE__); \
  } while (0)

// ---------------------------------------------------------------------------
// Device program of one flush.  A flush is a list of passes; each pass is one
// kernel launch.  A tile pass loads 2^nb amplitudes per workgroup (the tile
// qubits S, with every other qubit fixed by the workgroup index), applies a
// list of TileOps to the tile in LDS and writes it back: one HBM read and one
// HBM write of the state for the whole list.
// ---------------------------------------------------------------------------

enum TileOpKind : uint32_t {
  K_DENSE1 = 0,   // 2x2 matrix on local bit t (optionally controlled)
  K_DENSE = 1,    // 2^k x 2^k matrix, k <= 5, local bits pos[0..k-1]
  K_DIAG = 2,     // diagonal on k qubits, local (pos) or fixed by tile (gpos)
  K_MCX = 3,      // swap |t=0> <-> |t=1> where the controls hold
  K_MCY = 4,      // Y on t where the controls hold
  K_MCSWAP = 5,   // swap |a=1,b=0> <-> |a=0,b=1> where the controls hold
  K_MCPHASE = 6,  // multiply by phase where every control bit is 1
  K_PAULI = 7,    // Pauli string: flip xl, sign by parity(z), times phase
  K_PERM = 8,     // swap pairs of basis states of k local qubits
  K_PROBE = 9,    // read-only: tile share of Re<psi|phase P|psi> into probe slot npairs
  K_PHASES = 10,  // fused diagonal of register passes: where the controls hold, the product
                  // over npairs members of (bit ? v : u); pool[off + 3m .. +2] = (qubit, u, v)
  K_QPERM = 11,   // a layer of SWAPs on disjoint qubit pairs (an involution of the index
                  // bits) as one pass: k pairs (pos[j], gpos[j]) among the qubits outside
                  // the tile; cg: local bit i goes to local bit (cg >> 4i) & 15; bytes of
                  // xl (j = 0..3) and zl (j = 4): the local bit XORed into LDS slot bit j
                  // (0xFF: none)
};

constexpr uint8_t kNotLocal = 0xFF;

struct TileOp {
  uint32_t kind;
  uint32_t k;    // number of entries of pos/gpos
  uint32_t cl;   // control mask over local bits
  uint32_t cvl;  // required value of the local control bits
  uint64_t cg;   // control mask over global bits (fixed per tile)
  uint64_t cvg;  // required value of the global control bits
  uint32_t xl;   // Pauli: local X mask
  uint32_t zl;   // Pauli: local Z mask
  uint32_t t;    // target / pivot local bit
  uint32_t off;  // offset into the coefficient pool (complex<double> units)
  uint64_t zg;   // Pauli: Z mask over global bits
  uint32_t npairs;
  uint32_t nonunitary;  // 1: the op rescales amplitudes (projector, |phase| != 1)
  double pr, pi;  // phase
  uint8_t pos[16];   // local bit of entry j (kNotLocal when fixed by the tile)
  uint8_t gpos[16];  // global qubit of entry j
};

struct PassDesc {
  uint32_t kind;      // 0 = tile pass, 1 = dense group pass, 2 = register pass,
                      // 3 = probe sweep, 4 = qubit-permutation pass (K_QPERM)
  uint32_t nb;        // tile qubits
  uint32_t c;         // S contains qubits 0..c-1
  uint32_t op_begin;  // first TileOp of the pass
  uint32_t nops;
  uint32_t ncomp;     // number of qubits outside the tile
  uint32_t readonly;  // 1: the pass only reads the state (probes); nothing is written back
  uint32_t pad;
  uint8_t spos[16];   // tile qubits, ascending
  uint8_t comp[64];   // qubits outside the tile, ascending
};

}  // namespace aerhip
