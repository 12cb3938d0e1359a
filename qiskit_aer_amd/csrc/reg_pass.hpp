// Register-layout tile pass (gfx950).
//
// A workgroup of NT = 2^(B-R) threads owns a tile of 2^B amplitudes (the
// tile qubits of the pass; every other qubit is fixed by the workgroup
// index).  Each thread keeps 2^R amplitudes in VGPRs for the whole pass.  A
// *layout* says which R tile bits are register bits (register index j) and
// which B-R tile bits are thread bits (thread index t = wave * 64 + lane):
//
//     local index  l(t, j) = tl(t) | rl[j]
//     HBM offset   base + gt(t) + gr[j]
//
// tl/gt are per-thread VGPRs, rl/gr are uniform (SGPRs).  A gate whose mixed
// qubits are register bits runs in registers with no LDS traffic and no
// barrier; controls and diagonals may sit on any tile bit or on a bit fixed by
// the tile.  The host plans the register sets (Q_RELAYOUT moves the tile
// through LDS into a new layout) and loads/stores the tile straight from and
// to HBM in layouts whose lanes are tile bits 0..5, so that every wave access
// is a run of consecutive amplitudes.  Gates that do not run in registers are
// applied to the tile in LDS (Q_SPILL, TileOp primitives, Q_FILL).
#pragma once

#include <cstddef>
#include <cstring>
#include <utility>

#include "tile_pass.hpp"

namespace aerhip {

enum RegPassOpKind : uint32_t {
  Q_LOAD = 40,      // HBM -> registers in layout u.lay; extra: register ops that follow;
                    // a: fused-diagonal slots of the phase, b: pool offset of their
                    // DiagLane records
  Q_STORE = 41,     // registers -> HBM in layout u.lay (extra = 1: read-only pass, no write)
  Q_RELAYOUT = 42,  // registers -> LDS -> registers in layout u.lay
  Q_SPILL = 43,     // registers -> LDS (the tile is then in LDS)
  Q_FILL = 44,      // LDS -> registers in layout u.lay; extra, a, b: as Q_LOAD
  Q_LDS = 45,       // TileOp primitive extra on the LDS tile
  Q_U1 = 46,        // 2x2 matrix u.m on register bit a
  Q_U1C = 47,       // 2x2 matrix on register bit a where (l & cl) == cvl
  Q_X = 48,         // X on register bit a where (l & cl) == cvl
  Q_Y = 49,         // Y on register bit a where (l & cl) == cvl
  Q_U2 = 50,        // 4x4 matrix pool[extra..] on register bits a (matrix bit 0), b (bit 1)
  Q_D1 = 51,        // diag(u.m[0..1], u.m[2..3]) on local bit a where (l & cl) == cvl; bit j
                    // of b: tile bit a of register j is 1
  Q_PHASE = 52,     // multiply by u.m[0..1] where (l & cl) == cvl; bit j of b: register j
                    // meets the register part of the predicate
  Q_DIAG = 53,      // diagonal of TileOp extra (local and fixed bits)
  Q_PROBE = 54,     // Pauli probe whose X/Y bits are register bits (u.pr)
  Q_NOP = 55,       // keeps the amplitudes (pads a register phase to an even length)
  Q_U1M = 57,       // 2x2 matrix on the register pairs {j, j ^ a}; bit k of extra set:
                    // the k-th pair's lower register holds the |1> amplitude
  Q_U1X = 58,       // Q_U1 with an X on the same bit folded in where (base & cg) == cvg
  Q_U1MX = 59,      // Q_U1M with an X on the same logical bit folded in likewise
  Q_DIAGF = 60,     // fused diagonal: where (l & cl) == cvl (registers: bit j of b), the
                    // product over its member qubits of (bit ? v : u), from host tables at
                    // pool[extra..]: NR register factors, 16 + 16 lane factors (thread bits
                    // 0-3, 4-7), then a & 0xFF = number of 16-entry tables over fixed
                    // qubits u.fq[4c .. 4c+3] (0xFF: unused); a >> 16 = slot + 1 of the
                    // phase's lane tables staged in LDS (0: none, read the tables)
};

// Fused diagonals of a register phase: the factor over the qubits fixed by
// the tile is uniform per tile, so it is computed once per phase for up to
// kDiagSlots ops, while the tile loads are in flight, instead of as a chain of
// dependent table loads inside each op.  Lane l of a wave reads table l % 8 of
// slot l / 8, and the eight lanes of a slot multiply their entries together.
// The slots' lane tables (16 entries over thread bits 0-3, with that factor
// folded in, and 16 over thread bits 4-7) are then staged in LDS, so that an
// op reads its lane factor from LDS instead of through the vector caches.
// In the pool, after an entry equal to 1: one 8-byte record per lane (the
// distance back from the first record to the table, the entry equal to 1 for
// an unused table or slot, and the table's four qubits, 0xFF: unused), then
// for staging thread i = 32k + e the distance back from the first record to
// entry e of slot k's lane tables (the entry equal to 1 past the last slot).
constexpr uint32_t kDiagSlots = 8;
constexpr uint32_t kDiagLdsBytes = kDiagSlots * 32 * sizeof(cx<double>);
struct DiagLane {
  uint32_t back;
  uint32_t fq;
};
static_assert(sizeof(DiagLane) * 64 == 32 * sizeof(cx<double>), "64 records in 32 pool entries");
// Flag on kind: the op carries a new register relabelling (map_bm, map_t0).
constexpr uint32_t F_MAP = 0x80000000u;
// Q_U1X / Q_U1MX: where the tile predicate (base & cg) == cvg holds, an X on
// the gate's qubit comes first (the matrix is applied as U X: its two columns
// swapped).  An X controlled by qubits fixed by the tile folds into the next
// 2x2 gate on its target this way.
constexpr uint32_t F_ALL = F_MAP;

// Kinds the lean register-pass variant runs.
constexpr bool lean_kind(uint32_t kind) {
  const uint32_t k = kind & ~F_ALL;
  return k == Q_LOAD || k == Q_STORE || k == Q_RELAYOUT || k == Q_U1 || k == Q_U1X || k == Q_U1M ||
         k == Q_U1MX || k == Q_X || k == Q_D1 || k == Q_PHASE || k == Q_NOP || k == Q_DIAGF;
}

// Bit positions packed four to a word (byte i of word w is entry 4w+i), so
// that the device reads them with scalar loads.
struct RLayout {
  uint32_t tp[2];  // tile bit of thread bit u
  uint32_t rp[2];  // tile bit of register bit r
  uint32_t tq[2];  // qubit of thread bit u
  uint32_t rq[2];  // qubit of register bit r
};

__host__ __device__ __forceinline__ uint32_t byte_at(const uint32_t *w, int i) {
  return (w[i >> 2] >> (8 * (i & 3))) & 0xFFu;
}
__host__ __device__ __forceinline__ void set_byte(uint32_t *w, int i, uint32_t v) {
  w[i >> 2] = (w[i >> 2] & ~(0xFFu << (8 * (i & 3)))) | ((v & 0xFFu) << (8 * (i & 3)));
}

struct RProbe {
  uint32_t xr;  // X/Y mask in register-index space
  uint32_t zl;  // Z mask over tile bits
  uint64_t zg;  // Z mask over qubits fixed by the tile
  double pr, pi;
};

// 128-byte op record: a 64-byte header and a 64-byte payload, each read with
// one scalar load (both issued together, so an op costs one load latency).
template <typename T>
struct alignas(128) ROp {
  uint32_t kind;
  uint32_t a, b;
  uint32_t extra;    // probe slot, TileOp index or pool offset
  uint32_t cl, cvl;  // local control mask and value
  uint64_t cg, cvg;  // controls on qubits fixed by the tile
  uint16_t map_bm[5];  // with F_MAP in kind: the relabelling in force from this op on
  uint16_t map_t0;
  uint32_t pad[3];
  union {
    T m[8];          // 2x2 complex matrix (column major), diagonal, phase
    float mr[16];    // complex64 Q_U1 / Q_U1X / Q_U1M: the matrix as a real 4x4, rows over
                     // (re a0', im a0', re a1', im a1'), columns over (re a0, im a0, re a1, im a1)
    RLayout lay;
    RProbe pr;
    uint8_t fq[32];  // Q_DIAGF: fixed qubits of the tables, four per table
  } u;
};
static_assert(sizeof(ROp<double>) == 128 && sizeof(ROp<float>) == 128, "ROp is one 128-byte record");

// complex64 copy of a register program: matrices rounded to float; the 2x2
// gates (Q_U1, Q_U1X, Q_U1M) carry their matrix as the real 4x4 that the MFMA
// path multiplies by.
inline void rops_to_float(const ROp<double> *in, size_t n, ROp<float> *out) {
  for (size_t i = 0; i < n; ++i) {
    std::memcpy(&out[i], &in[i], sizeof(ROp<float>));
    const uint32_t k = in[i].kind & ~F_ALL;
    if (k == Q_U1 || k == Q_U1X || k == Q_U1M || k == Q_U1MX) {
      const double *m = in[i].u.m;  // m0 = u00, m1 = u10, m2 = u01, m3 = u11 (re, im)
      const double u00r = m[0], u00i = m[1], u10r = m[2], u10i = m[3];
      const double u01r = m[4], u01i = m[5], u11r = m[6], u11i = m[7];
      const double rows[16] = {u00r, -u00i, u01r, -u01i,   // re a0'
                               u00i, u00r, u01i, u01r,     // im a0'
                               u10r, -u10i, u11r, -u11i,   // re a1'
                               u10i, u10r, u11i, u11r};    // im a1'
      for (int e = 0; e < 16; ++e) out[i].u.mr[e] = static_cast<float>(rows[e]);
    } else if (k == Q_U1C || k == Q_D1 || k == Q_PHASE) {
      for (int e = 0; e < 8; ++e) out[i].u.m[e] = static_cast<float>(in[i].u.m[e]);
    }
  }
}

// Fused multiply-add in the amplitude precision.
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// m0*a0 + m1*a1 with one multiply and three fused multiply-adds per component.
template <typename T>
__device__ __forceinline__ cx<T> mv2(cx<T> m0, cx<T> a0, cx<T> m1, cx<T> a1) {
  T x = m1.x * a1.x;
  x = fma_t(-m1.y, a1.y, x);
  x = fma_t(m0.x, a0.x, x);
  x = fma_t(-m0.y, a0.y, x);
  T y = m1.x * a1.y;
  y = fma_t(m1.y, a1.x, y);
  y = fma_t(m0.x, a0.y, y);
  y = fma_t(m0.y, a0.x, y);
  return mk<T>(x, y);
}

// m0*a0 + m1*a1 + m2*a2 + m3*a3: one multiply and seven fused multiply-adds
// per component.
template <typename T>
__device__ __forceinline__ cx<T> mv4(cx<T> m0, cx<T> a0, cx<T> m1, cx<T> a1, cx<T> m2, cx<T> a2, cx<T> m3,
                                     cx<T> a3) {
  T x = m3.x * a3.x;
  x = fma_t(-m3.y, a3.y, x);
  x = fma_t(m2.x, a2.x, x);
  x = fma_t(-m2.y, a2.y, x);
  x = fma_t(m1.x, a1.x, x);
  x = fma_t(-m1.y, a1.y, x);
  x = fma_t(m0.x, a0.x, x);
  x = fma_t(-m0.y, a0.y, x);
  T y = m3.x * a3.y;
  y = fma_t(m3.y, a3.x, y);
  y = fma_t(m2.x, a2.y, y);
  y = fma_t(m2.y, a2.x, y);
  y = fma_t(m1.x, a1.y, y);
  y = fma_t(m1.y, a1.x, y);
  y = fma_t(m0.x, a0.y, y);
  y = fma_t(m0.y, a0.x, y);
  return mk<T>(x, y);
}

template <typename T>
__device__ __forceinline__ cx<T> csel(bool c, cx<T> a, cx<T> b) {
  return mk<T>(c ? a.x : b.x, c ? a.y : b.y);
}

// Uniform layout state: register j holds the amplitude whose tile index is
// at(j) = t0 ^ XOR of bm[s] over the set bits s of j (folded at compile time
// since j always is a constant).  In a fresh layout bm[s] is the tile bit of
// register bit s and t0 = 0; X and CX gates between register qubits are
// applied by changing bm/t0 (an affine relabelling of the registers) instead
// of moving amplitudes.  R + 1 words instead of 2^R keep the scalar register
// file free.
template <int NR>
struct RegIdx {
  static constexpr int R = NR == 8 ? 3 : NR == 16 ? 4 : 5;
  uint32_t bm[R];  // tile-bit mask contributed by register bit s
  uint32_t t0;     // tile bits set for every register (a pending X)
  __device__ __forceinline__ uint32_t at(int j) const {
    uint32_t v = t0;
#pragma unroll
    for (int r = 0; r < R; ++r)
      if ((j >> r) & 1) v ^= bm[r];
    return v;
  }
};
// HBM offset gr[j] of register j.
template <int NR>
struct RegOff {
  uint64_t gr[NR];
};

template <int R>
__device__ __forceinline__ RegIdx<1 << R> reg_index(const RLayout &lay) {
  RegIdx<1 << R> r;
#pragma unroll
  for (int b = 0; b < R; ++b) r.bm[b] = 1u << byte_at(lay.rp, b);
  r.t0 = 0;
  return r;
}

// HBM offsets of the registers under a relabelling: the tile bits of at(j)
// mapped to their qubits (only register tile bits occur in at(j)).
template <int R>
__device__ __forceinline__ RegOff<1 << R> reg_offset_mapped(const RLayout &lay,
                                                            const RegIdx<1 << R> &ri) {
  RegOff<1 << R> o;
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) {
    const uint32_t l = ri.at(j);
    uint64_t g = 0;
#pragma unroll
    for (int b = 0; b < R; ++b)
      if ((l >> byte_at(lay.rp, b)) & 1u) g |= 1ull << byte_at(lay.rq, b);
    o.gr[j] = g;
  }
  return o;
}

template <int R>
__device__ __forceinline__ RegOff<1 << R> reg_offset(const RLayout &lay) {
  RegOff<1 << R> r;
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) {
    uint64_t g = 0;
#pragma unroll
    for (int b = 0; b < R; ++b)
      if (j & (1 << b)) g |= 1ull << byte_at(lay.rq, b);
    r.gr[j] = g;
  }
  return r;
}

template <int TB>
__device__ __forceinline__ uint32_t thread_index(const RLayout &lay, uint32_t tid) {
  uint32_t l = 0;
#pragma unroll
  for (int u = 0; u < TB; ++u) l |= ((tid >> u) & 1u) << byte_at(lay.tp, u);
  return l;
}

template <int TB>
__device__ __forceinline__ uint64_t thread_offset(const RLayout &lay, uint32_t tid) {
  uint64_t g = 0;
#pragma unroll
  for (int u = 0; u < TB; ++u) g |= static_cast<uint64_t>((tid >> u) & 1u) << byte_at(lay.tq, u);
  return g;
}

// ---- register gates, one code path per register bit ----------------------
//
// A gate on register bit b pairs registers j and j | 2^b (b is a template
// parameter, so every index is a compile-time constant); a small dispatch
// picks the instantiation.  Every path writes all of its outputs, so the op
// loop can alternate between two register banks.

template <typename T, int NR, int b>
__device__ __forceinline__ Amps<T, NR> q_u1(const Amps<T, NR> &a, cx<T> m0, cx<T> m1, cx<T> m2,
                                            cx<T> m3) {
  Amps<T, NR> o;
#pragma unroll
  for (int p = 0; p < NR / 2; ++p) {
    const int j = pair_lo<b>(p), j1 = j | (1 << b);
    o.v[j] = mv2(m0, a.v[j], m2, a.v[j1]);
    o.v[j1] = mv2(m1, a.v[j], m3, a.v[j1]);
  }
  return o;
}

// 2x2 gate on the register pairs {j, j ^ m} (m: a relabelled qubit spans
// several register bits).  The k-th pair (k-th register j with the top bit of
// m clear) is |0>,|1> = (j, j^m) if bit k of roles is 0, else (j^m, j).
template <int m>
constexpr int top_bit() {
  return m >= 16 ? 4 : m >= 8 ? 3 : m >= 4 ? 2 : m >= 2 ? 1 : 0;
}

// Index of the pair whose lower register is j (j has bit hb clear): j with
// bit hb removed.
template <int hb>
constexpr int k_of(int j) {
  return ((j >> (hb + 1)) << hb) | (j & ((1 << hb) - 1));
}

template <typename T, int NR, int m>
__device__ __forceinline__ Amps<T, NR> q_u1m(const Amps<T, NR> &a, uint32_t roles, cx<T> m0,
                                             cx<T> m1, cx<T> m2, cx<T> m3) {
  constexpr int hb = top_bit<m>();
  Amps<T, NR> o;
  int k = 0;
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    if (j & (1 << hb)) continue;
    const int jj = j ^ m;
    if constexpr (sizeof(T) == 8) {
      // a uniform branch on the pair's role: no per-pair selects
      if ((roles >> k_of<hb>(j)) & 1u) {
        asm volatile("");  // keeps the branch (no if-conversion into selects)
        o.v[j] = mv2(m3, a.v[j], m1, a.v[jj]);
        o.v[jj] = mv2(m2, a.v[j], m0, a.v[jj]);
      } else {
        o.v[j] = mv2(m0, a.v[j], m2, a.v[jj]);
        o.v[jj] = mv2(m1, a.v[j], m3, a.v[jj]);
      }
    } else {
      const bool sw = (roles >> k) & 1u;
      const cx<T> c00 = sw ? m3 : m0, c01 = sw ? m1 : m2;
      const cx<T> c10 = sw ? m2 : m1, c11 = sw ? m0 : m3;
      o.v[j] = mv2(c00, a.v[j], c01, a.v[jj]);
      o.v[jj] = mv2(c10, a.v[j], c11, a.v[jj]);
    }
    ++k;
  }
  return o;
}

template <typename T, int NR, int M = 1>
__device__ __forceinline__ Amps<T, NR> d_u1m(uint32_t m, uint32_t roles, const Amps<T, NR> &a,
                                             cx<T> m0, cx<T> m1, cx<T> m2, cx<T> m3) {
  if constexpr (M == NR - 1) {
    return q_u1m<T, NR, M>(a, roles, m0, m1, m2, m3);
  } else {
    if (m == M) return q_u1m<T, NR, M>(a, roles, m0, m1, m2, m3);
    return d_u1m<T, NR, M + 1>(m, roles, a, m0, m1, m2, m3);
  }
}

// complex64 2x2 gate with MFMA: v_mfma_f32_4x4x1f32 runs 16 independent 4x4
// blocks (lanes 4k..4k+3 form block k).  With A = row (lane % 4) of the real
// 4x4 matrix and B = one real component of the lane's own pair, four
// accumulating issues leave the pair's four output components in the lane's
// accumulator: twice the FP32 rate of FMAs on the vector ALU.
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int NR, int b>
__device__ __forceinline__ Amps<float, NR> q_u1_mfma(const Amps<float, NR> &a, float A0, float A1,
                                                     float A2, float A3) {
  // Issue k of every pair before issue k+1 of any: the accumulation chains
  // interleave, so no issue waits on the previous one's result.
  f32x4 acc[NR / 2];
#pragma unroll
  for (int p = 0; p < NR / 2; ++p) {
    const int j = pair_lo<b>(p);
    acc[p] = __builtin_amdgcn_mfma_f32_4x4x1f32(A0, a.v[j].x, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
  }
#pragma unroll
  for (int p = 0; p < NR / 2; ++p) {
    const int j = pair_lo<b>(p);
    acc[p] = __builtin_amdgcn_mfma_f32_4x4x1f32(A1, a.v[j].y, acc[p], 0, 0, 0);
  }
#pragma unroll
  for (int p = 0; p < NR / 2; ++p) {
    const int j1 = pair_lo<b>(p) | (1 << b);
    acc[p] = __builtin_amdgcn_mfma_f32_4x4x1f32(A2, a.v[j1].x, acc[p], 0, 0, 0);
  }
#pragma unroll
  for (int p = 0; p < NR / 2; ++p) {
    const int j1 = pair_lo<b>(p) | (1 << b);
    acc[p] = __builtin_amdgcn_mfma_f32_4x4x1f32(A3, a.v[j1].y, acc[p], 0, 0, 0);
  }
  Amps<float, NR> o;
#pragma unroll
  for (int p = 0; p < NR / 2; ++p) {
    const int j = pair_lo<b>(p), j1 = j | (1 << b);
    o.v[j] = mk<float>(acc[p][0], acc[p][1]);
    o.v[j1] = mk<float>(acc[p][2], acc[p][3]);
  }
  return o;
}

// complex64 2x2 gate on the relabelled pairs {j, j ^ m} (Q_U1M) with MFMA.
// A pair whose lower register holds |1> (role bit set) takes the matrix
// P R P (P swaps the pair's two amplitudes): its A operand is row (lane ^ 2)
// of R with the column halves swapped, `As`; the others take `An`.
template <int NR, int m>
__device__ __forceinline__ Amps<float, NR> q_u1m_mfma(const Amps<float, NR> &a, uint32_t roles,
                                                      f32x4 An, f32x4 As) {
  constexpr int hb = top_bit<m>();
  constexpr int NP = NR / 2;
  constexpr int G = NP < 4 ? NP : 4;  // pairs whose accumulation chains interleave
  // lower register of pair p: the p-th index with bit hb clear
  auto lo = [](int p) { return ((p >> hb) << (hb + 1)) | (p & ((1 << hb) - 1)); };
  Amps<float, NR> o;
#pragma unroll
  for (int g = 0; g < NP; g += G) {
    f32x4 acc[G];
    f32x4 A[G];
#pragma unroll
    for (int p = 0; p < G; ++p) A[p] = ((roles >> (g + p)) & 1u) ? As : An;
#pragma unroll
    for (int p = 0; p < G; ++p)
      acc[p] = __builtin_amdgcn_mfma_f32_4x4x1f32(A[p].x, a.v[lo(g + p)].x, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
#pragma unroll
    for (int p = 0; p < G; ++p)
      acc[p] = __builtin_amdgcn_mfma_f32_4x4x1f32(A[p].y, a.v[lo(g + p)].y, acc[p], 0, 0, 0);
#pragma unroll
    for (int p = 0; p < G; ++p)
      acc[p] = __builtin_amdgcn_mfma_f32_4x4x1f32(A[p].z, a.v[lo(g + p) ^ m].x, acc[p], 0, 0, 0);
#pragma unroll
    for (int p = 0; p < G; ++p)
      acc[p] = __builtin_amdgcn_mfma_f32_4x4x1f32(A[p].w, a.v[lo(g + p) ^ m].y, acc[p], 0, 0, 0);
#pragma unroll
    for (int p = 0; p < G; ++p) {
      o.v[lo(g + p)] = mk<float>(acc[p][0], acc[p][1]);
      o.v[lo(g + p) ^ m] = mk<float>(acc[p][2], acc[p][3]);
    }
  }
  return o;
}

template <int NR, int M = 1>
__device__ __forceinline__ Amps<float, NR> d_u1m_mfma(uint32_t m, uint32_t roles, const Amps<float, NR> &a,
                                                      f32x4 An, f32x4 As) {
  if constexpr (M == NR - 1) {
    return q_u1m_mfma<NR, M>(a, roles, An, As);
  } else {
    if (m == M) return q_u1m_mfma<NR, M>(a, roles, An, As);
    return d_u1m_mfma<NR, M + 1>(m, roles, a, An, As);
  }
}

// Row (lane ^ 2) of a quad's A operands, read across the quad (DPP
// quad_perm [2,3,0,1]), with its column halves swapped: the A operand of
// P R P for the row this lane computes.
__device__ __forceinline__ f32x4 mfma_row_swapped(f32x4 An) {
  auto perm = [](float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
  };
  return f32x4{perm(An.z), perm(An.w), perm(An.x), perm(An.y)};
}

template <typename T, int NR, int b>
__device__ __forceinline__ Amps<T, NR> q_u1c(const Amps<T, NR> &a, uint32_t tlc,
                                             const RegIdx<NR> &ri, uint32_t cl, uint32_t cvl,
                                             cx<T> m0, cx<T> m1, cx<T> m2, cx<T> m3) {
  Amps<T, NR> o;
#pragma unroll
  for (int p = 0; p < NR / 2; ++p) {
    const int j = pair_lo<b>(p), j1 = j | (1 << b);
    const bool on = (tlc | (ri.at(j) & cl)) == cvl;
    o.v[j] = csel(on, mv2(m0, a.v[j], m2, a.v[j1]), a.v[j]);
    o.v[j1] = csel(on, mv2(m1, a.v[j], m3, a.v[j1]), a.v[j1]);
  }
  return o;
}

template <typename T, int NR, int b, bool isy>
__device__ __forceinline__ Amps<T, NR> q_xy(const Amps<T, NR> &a, uint32_t tlc,
                                            const RegIdx<NR> &ri, uint32_t cl, uint32_t cvl) {
  Amps<T, NR> o;
#pragma unroll
  for (int p = 0; p < NR / 2; ++p) {
    const int j = pair_lo<b>(p), j1 = j | (1 << b);
    const bool on = (tlc | (ri.at(j) & cl)) == cvl;
    const cx<T> a0 = a.v[j], a1 = a.v[j1];
    cx<T> n0 = a1, n1 = a0;
    if constexpr (isy) {
      n0 = mk<T>(a1.y, -a1.x);
      n1 = mk<T>(-a0.y, a0.x);
    }
    o.v[j] = csel(on, n0, a0);
    o.v[j1] = csel(on, n1, a1);
  }
  return o;
}

// 4x4 matrix on register bits b0 (matrix bit 0) and b1 (matrix bit 1).
template <typename T, int NR, int b0, int b1>
__device__ __forceinline__ Amps<T, NR> q_u2(const Amps<T, NR> &a, uint32_t tlc,
                                            const RegIdx<NR> &ri, uint32_t cl, uint32_t cvl,
                                            const cx<double> *__restrict__ mp) {
  constexpr int lo = b0 < b1 ? b0 : b1, hi = b0 < b1 ? b1 : b0;
  Amps<T, NR> o;
  if (cl == 0) {
    // no local control: the whole tile or nothing (a uniform branch), and
    // no per-output selects
    if (cvl != 0) return pass_through(a);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const cx<T> c0 = to_t<T>(mp[r]), c1 = to_t<T>(mp[r + 4]);
      const cx<T> c2 = to_t<T>(mp[r + 8]), c3 = to_t<T>(mp[r + 12]);
#pragma unroll
      for (int p = 0; p < NR / 4; ++p) {
        const int j = pair_lo<hi>(pair_lo<lo>(p));
        const int i1 = j | (1 << b0), i2 = j | (1 << b1), i3 = j | (1 << b0) | (1 << b1);
        const int dst = r == 0 ? j : r == 1 ? i1 : r == 2 ? i2 : i3;
        o.v[dst] = mv4(c0, a.v[j], c1, a.v[i1], c2, a.v[i2], c3, a.v[i3]);
      }
    }
    return o;
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const cx<T> c0 = to_t<T>(mp[r]), c1 = to_t<T>(mp[r + 4]);
    const cx<T> c2 = to_t<T>(mp[r + 8]), c3 = to_t<T>(mp[r + 12]);
#pragma unroll
    for (int p = 0; p < NR / 4; ++p) {
      const int j = pair_lo<hi>(pair_lo<lo>(p));
      const int i1 = j | (1 << b0), i2 = j | (1 << b1), i3 = j | (1 << b0) | (1 << b1);
      const int dst = r == 0 ? j : r == 1 ? i1 : r == 2 ? i2 : i3;
      const bool on = (tlc | (ri.at(j) & cl)) == cvl;
      o.v[dst] = csel(on, mv4(c0, a.v[j], c1, a.v[i1], c2, a.v[i2], c3, a.v[i3]), a.v[dst]);
    }
  }
  return o;
}

// Dispatchers: b < R; the last instantiation is the default so that every
// path computes the outputs.
template <int NR>
constexpr int reg_bits_of() {
  return NR == 8 ? 3 : NR == 16 ? 4 : 5;
}

template <typename T, int NR, int B0 = 0>
__device__ __forceinline__ Amps<T, NR> d_u1(uint32_t b, const Amps<T, NR> &a, cx<T> m0, cx<T> m1,
                                            cx<T> m2, cx<T> m3) {
  if constexpr (B0 == reg_bits_of<NR>() - 1) {
    return q_u1<T, NR, B0>(a, m0, m1, m2, m3);
  } else {
    if (b == B0) return q_u1<T, NR, B0>(a, m0, m1, m2, m3);
    return d_u1<T, NR, B0 + 1>(b, a, m0, m1, m2, m3);
  }
}

template <int NR, int B0 = 0>
__device__ __forceinline__ Amps<float, NR> d_u1_mfma(uint32_t b, const Amps<float, NR> &a,
                                                     float A0, float A1, float A2, float A3) {
  if constexpr (B0 == reg_bits_of<NR>() - 1) {
    return q_u1_mfma<NR, B0>(a, A0, A1, A2, A3);
  } else {
    if (b == B0) return q_u1_mfma<NR, B0>(a, A0, A1, A2, A3);
    return d_u1_mfma<NR, B0 + 1>(b, a, A0, A1, A2, A3);
  }
}

template <typename T, int NR, int B0 = 0>
__device__ __forceinline__ Amps<T, NR> d_u1c(uint32_t b, const Amps<T, NR> &a, uint32_t tlc,
                                             const RegIdx<NR> &ri, uint32_t cl, uint32_t cvl,
                                             cx<T> m0, cx<T> m1, cx<T> m2, cx<T> m3) {
  if constexpr (B0 == reg_bits_of<NR>() - 1) {
    return q_u1c<T, NR, B0>(a, tlc, ri, cl, cvl, m0, m1, m2, m3);
  } else {
    if (b == B0) return q_u1c<T, NR, B0>(a, tlc, ri, cl, cvl, m0, m1, m2, m3);
    return d_u1c<T, NR, B0 + 1>(b, a, tlc, ri, cl, cvl, m0, m1, m2, m3);
  }
}

template <typename T, int NR, int B0 = 0>
__device__ __forceinline__ Amps<T, NR> d_xy(uint32_t b, const Amps<T, NR> &a, uint32_t tlc,
                                            const RegIdx<NR> &ri, uint32_t cl, uint32_t cvl,
                                            bool isy) {
  if constexpr (B0 == reg_bits_of<NR>() - 1) {
    if (isy) return q_xy<T, NR, B0, true>(a, tlc, ri, cl, cvl);
    return q_xy<T, NR, B0, false>(a, tlc, ri, cl, cvl);
  } else {
    if (b == B0) {
      if (isy) return q_xy<T, NR, B0, true>(a, tlc, ri, cl, cvl);
      return q_xy<T, NR, B0, false>(a, tlc, ri, cl, cvl);
    }
    return d_xy<T, NR, B0 + 1>(b, a, tlc, ri, cl, cvl, isy);
  }
}

// 4x4 on (b0, b1), enumerated as P = b0 * R + b1 over distinct pairs.
template <typename T, int NR, int P = 1>
__device__ __forceinline__ Amps<T, NR> d_u2(uint32_t p, const Amps<T, NR> &a, uint32_t tlc,
                                            const RegIdx<NR> &ri, uint32_t cl, uint32_t cvl,
                                            const cx<double> *mp) {
  constexpr int R = reg_bits_of<NR>();
  constexpr int b0 = P / R, b1 = P % R;
  constexpr int last = (R - 1) * R + (R - 2);
  if constexpr (P >= last) {
    return q_u2<T, NR, R - 1, R - 2>(a, tlc, ri, cl, cvl, mp);
  } else if constexpr (b0 == b1) {
    return d_u2<T, NR, P + 1>(p, a, tlc, ri, cl, cvl, mp);
  } else {
    if (p == static_cast<uint32_t>(P)) return q_u2<T, NR, b0, b1>(a, tlc, ri, cl, cvl, mp);
    return d_u2<T, NR, P + 1>(p, a, tlc, ri, cl, cvl, mp);
  }
}

// sum_j sign_j * Re(phase * a[j ^ xr] * conj(a[j])), sign_j the parity of
// (l_j & zl) xor sgb: the thread's share of Re<psi| phase P |psi>.
template <typename T, int NR, int xr>
__device__ __forceinline__ double q_probe(const Amps<T, NR> &a, uint32_t tl,
                                          const RegIdx<NR> &ri, uint32_t zl, uint32_t sgb,
                                          double pr, double pi) {
  const uint32_t st = (__popc(tl & zl) & 1u) ^ sgb;
  double acc = 0;
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    const double x0 = a.v[j].x, y0 = a.v[j].y;
    const double x1 = a.v[j ^ xr].x, y1 = a.v[j ^ xr].y;
    const double px = pr * x1 - pi * y1, py = pr * y1 + pi * x1;
    double v = px * x0 + py * y0;
    if (st ^ (__popc(ri.at(j) & zl) & 1u)) v = -v;
    acc += v;
  }
  return acc;
}

template <typename T, int NR, int X>
__device__ __forceinline__ double q_probe_dispatch(uint32_t xr, const Amps<T, NR> &a, uint32_t tl,
                                                   const RegIdx<NR> &ri, uint32_t zl,
                                                   uint32_t sgb, double pr, double pi) {
  if constexpr (X < NR) {
    if (xr == X) return q_probe<T, NR, X>(a, tl, ri, zl, sgb, pr, pi);
    return q_probe_dispatch<T, NR, X + 1>(xr, a, tl, ri, zl, sgb, pr, pi);
  } else {
    return 0.0;
  }
}

// Layout state of a thread: tile index of its thread part and of each register.
template <int NR>
struct LayoutState {
  uint32_t tl;
  RegIdx<NR> ri;
};

template <int R, int TB>
__device__ __forceinline__ LayoutState<1 << R> make_state(const RLayout &lay, uint32_t tid) {
  LayoutState<1 << R> s;
  s.tl = thread_index<TB>(lay, tid);
  s.ri = reg_index<R>(lay);
  return s;
}

// The relabelling carried by an op with F_MAP, read as three dwords (the
// 16-bit fields sit at a 4-byte aligned offset, so scalar loads can fetch them).
template <int R, typename T>
__device__ __forceinline__ void load_map(const ROp<T> &op, RegIdx<1 << R> &ri) {
  static_assert(offsetof(ROp<T>, map_bm) % 4 == 0 && offsetof(ROp<T>, map_t0) == offsetof(ROp<T>, map_bm) + 10,
                "map fields are three aligned dwords");
  uint32_t w[3];
  __builtin_memcpy(w, op.map_bm, sizeof(w));
#pragma unroll
  for (int s = 0; s < R; ++s) ri.bm[s] = (w[s >> 1] >> (16 * (s & 1))) & 0xFFFFu;
  ri.t0 = w[2] >> 16;
}

// The amplitudes unchanged, as the result of an op that leaves them alone.
// The empty asm pins the copy into the other register bank to this path:
// without it the compiler hoists the copy into the shared dispatch blocks,
// where every op would pay for it.
template <typename T, int NR>
__device__ __forceinline__ Amps<T, NR> pass_through(const Amps<T, NR> &a) {
  Amps<T, NR> o = a;
#pragma unroll
  for (int j = 0; j < NR; ++j) asm volatile("" : "+v"(o.v[j].x), "+v"(o.v[j].y));
  return o;
}

// Amplitude loads and stores of the HBM sweep, marked non-temporal: each
// amplitude is read and written once per pass, so caching it gains nothing.
template <typename T>
__device__ __forceinline__ cx<T> ld_stream(const cx<T> *p) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2 *>(p));
  return mk<T>(v.x, v.y);
}
template <typename T>
__device__ __forceinline__ void st_stream(cx<T> *p, cx<T> a) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  __builtin_nontemporal_store(v2{a.x, a.y}, reinterpret_cast<v2 *>(p));
}

// A wave-uniform 64-bit value moved to scalar registers.
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// The value itself, hidden from the optimizer (no code is emitted).
__device__ __forceinline__ uint32_t opaque_u32(uint32_t v) {
  asm volatile("" : "+s"(v));
  return v;
}

// The head of an op, read one op ahead: its scalar fields; in the lean
// variant (PF) also the controls and the first words of the payload (for
// complex128 the 2x2 matrix, for complex64 the diagonal or phase), as scalar
// loads, so that they are in SGPRs when the op starts (the full variant has
// no SGPRs to spare for them); for complex64, the lane's row of the real 4x4
// matrix that the MFMA paths multiply by (row lane % 4, a per-lane vector
// load).
template <typename T, bool PF>
struct OpHead {
  uint32_t kind, a, b, extra;
  uint32_t cl, cvl;
  uint64_t cg, cvg;
  f32x4 arow;
  double m[8];
  float mf[4];
};

template <bool PF, typename T>
__device__ __forceinline__ OpHead<T, PF> load_head(const ROp<T> &op) {
  OpHead<T, PF> h;
  h.kind = op.kind;
  h.a = op.a;
  h.b = op.b;
  h.extra = op.extra;
  h.cl = h.cvl = 0;
  h.cg = h.cvg = 0;
  h.arow = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int e = 0; e < 8; ++e) h.m[e] = 0.0;
#pragma unroll
  for (int e = 0; e < 4; ++e) h.mf[e] = 0.f;
  if constexpr (PF) {
    h.cl = op.cl;
    h.cvl = op.cvl;
    h.cg = op.cg;
    h.cvg = op.cvg;
  }
  if constexpr (sizeof(T) == 4) {
    h.arow = *reinterpret_cast<const f32x4 *>(op.u.mr + 4 * (threadIdx.x & 3u));
    if constexpr (PF) {
#pragma unroll
      for (int e = 0; e < 4; ++e) h.mf[e] = op.u.mr[e];
    }
  } else if constexpr (PF) {
#pragma unroll
    for (int e = 0; e < 8; ++e) h.m[e] = op.u.m[e];
  }
  return h;
}

// The op's controls and its first payload words, from the head where it
// holds them.
template <typename T, bool PF>
__device__ __forceinline__ uint32_t op_cl(const ROp<T> &op, const OpHead<T, PF> &hd) {
  return PF ? hd.cl : op.cl;
}
template <typename T, bool PF>
__device__ __forceinline__ uint32_t op_cvl(const ROp<T> &op, const OpHead<T, PF> &hd) {
  return PF ? hd.cvl : op.cvl;
}
template <typename T, bool PF>
__device__ __forceinline__ bool op_tile_ok(const ROp<T> &op, const OpHead<T, PF> &hd, uint64_t base) {
  return PF ? (base & hd.cg) == hd.cvg : (base & op.cg) == op.cvg;
}
template <typename T, bool PF>
__device__ __forceinline__ cx<T> op_m(const ROp<T> &op, const OpHead<T, PF> &hd, int e) {
  if constexpr (PF) {
    if constexpr (sizeof(T) == 4)
      return mk<T>(hd.mf[2 * e], hd.mf[2 * e + 1]);
    else
      return mk<T>(hd.m[2 * e], hd.m[2 * e + 1]);
  }
  return mk<T>(op.u.m[2 * e], op.u.m[2 * e + 1]);
}

// A register op that rewrites every amplitude (a layout change or a gate):
// returns the new amplitudes and never passes the input through unchanged
// (except probes and Q_NOP), so that the op loop can alternate between two
// register banks without copies.  Controls fixed by the tile are folded into
// the local predicate (an unmet one turns it into "never"; Q_U1 carries no
// controls).
template <typename T>
__device__ __forceinline__ cx<T> mat_at(const T *m, int e) {
  return mk<T>(m[2 * e], m[2 * e + 1]);
}

// The 2x2 matrix of a 2x2 gate op (column major: u00, u10, u01, u11), its
// columns swapped when an X is folded in on this tile (fx).  complex64 ops
// carry the real 4x4 form, whose rows 0 and 2 hold the real parts and rows 1
// and 3 the imaginary parts.
template <typename T, bool PF>
__device__ __forceinline__ void u1_matrix(const ROp<T> &op, const OpHead<T, PF> &hd, bool fx, cx<T> &m0,
                                          cx<T> &m1, cx<T> &m2, cx<T> &m3) {
  cx<T> c0, c1, c2, c3;
  if constexpr (sizeof(T) == 4) {
    const float *r = op.u.mr;
    c0 = mk<T>(r[0], r[4]);
    c1 = mk<T>(r[8], r[12]);
    c2 = mk<T>(r[2], r[6]);
    c3 = mk<T>(r[10], r[14]);
  } else {
    const double *m = PF ? hd.m : reinterpret_cast<const double *>(op.u.m);
    c0 = mat_at<double>(m, 0);
    c1 = mat_at<double>(m, 1);
    c2 = mat_at<double>(m, 2);
    c3 = mat_at<double>(m, 3);
  }
  m0 = fx ? c2 : c0;
  m1 = fx ? c3 : c1;
  m2 = fx ? c0 : c2;
  m3 = fx ? c1 : c3;
}

template <typename T, int B, int R, bool FULL>
__device__ __forceinline__ Amps<T, 1 << R> rw_op(const ROp<T> &op, const OpHead<T, !FULL> &hd,
                                                 const Amps<T, 1 << R> &a,
                                                 LayoutState<1 << R> &st, cx<T> *lds,
                                                 double *red, uint64_t base, uint32_t tid,
                                                 const TileOp *__restrict__ tops,
                                                 const cx<double> *__restrict__ pool,
                                                 const ProbeSink &sink, const cx<double> *dlt) {
  constexpr int NR = 1 << R;
  constexpr int TB = B - R;
  const uint32_t kind = hd.kind & ~F_ALL, ra = hd.a, rb = hd.b, extra = hd.extra;
  if (hd.kind & F_MAP) {
    load_map<R>(op, st.ri);
  }
  // Q_U1, the most frequent kind, is tested first, on a copy of the kind
  // that the compiler cannot fold into the comparison tree of the others.
  if (opaque_u32(kind) == Q_U1) {
    if constexpr (sizeof(T) == 4) {
      // A operand of issue k: row (lane % 4), column k of the real 4x4 matrix
      return d_u1_mfma<NR>(ra, a, hd.arow.x, hd.arow.y, hd.arow.z, hd.arow.w);
    } else {
      cx<T> m0, m1, m2, m3;
      u1_matrix(op, hd, false, m0, m1, m2, m3);
      return d_u1<T, NR>(ra, a, m0, m1, m2, m3);
    }
  }
  if (kind == Q_U1X) {
    const bool fx = op_tile_ok(op, hd, base);
    if constexpr (sizeof(T) == 4) {
      // A operand of issue k: row (lane % 4), column k of the real 4x4
      // matrix; U X swaps the columns that act on a0 and a1 (k and k ^ 2).
      const float A0 = fx ? hd.arow.z : hd.arow.x, A1 = fx ? hd.arow.w : hd.arow.y;
      const float A2 = fx ? hd.arow.x : hd.arow.z, A3 = fx ? hd.arow.y : hd.arow.w;
      return d_u1_mfma<NR>(ra, a, A0, A1, A2, A3);
    } else {
      // a uniform branch on the tile predicate: the folded X swaps the
      // matrix columns (no selects)
      cx<T> m0, m1, m2, m3;
      u1_matrix(op, hd, false, m0, m1, m2, m3);
      if (fx) {
        asm volatile("");  // keeps the branch (no if-conversion into selects)
        return d_u1<T, NR>(ra, a, m2, m3, m0, m1);
      }
      return d_u1<T, NR>(ra, a, m0, m1, m2, m3);
    }
  }
  if (kind == Q_U1M) {  // relabelled pairs
    if constexpr (sizeof(T) == 4 && !FULL) {
      return d_u1m_mfma<NR>(ra, extra, a, hd.arow, mfma_row_swapped(hd.arow));
    } else {
      cx<T> m0, m1, m2, m3;
      u1_matrix(op, hd, false, m0, m1, m2, m3);
      return d_u1m<T, NR>(ra, extra, a, m0, m1, m2, m3);
    }
  }
  if (kind == Q_U1MX) {  // the same, with an X folded in where its tile predicate holds
    const bool fx = op_tile_ok(op, hd, base);
    if constexpr (sizeof(T) == 4 && !FULL) {
      const f32x4 An = fx ? f32x4{hd.arow.z, hd.arow.w, hd.arow.x, hd.arow.y} : hd.arow;
      return d_u1m_mfma<NR>(ra, extra, a, An, mfma_row_swapped(An));
    } else {
      cx<T> m0, m1, m2, m3;
      u1_matrix(op, hd, fx, m0, m1, m2, m3);
      return d_u1m<T, NR>(ra, extra, a, m0, m1, m2, m3);
    }
  }
  if (kind == Q_PROBE || kind == Q_NOP) {
    if (FULL && kind == Q_PROBE) {
      const RProbe &pr = op.u.pr;
      const uint32_t sgb = __popcll(base & pr.zg) & 1u;
      double acc = q_probe_dispatch<T, NR, 0>(pr.xr, a, st.tl, st.ri, pr.zl, sgb, pr.pr, pr.pi);
      acc = block_sum_nt<(1 << TB)>(acc, red);
      if (tid == 0) sink.put(extra, acc);
    }
    return pass_through(a);
  }
  if (kind == Q_RELAYOUT) {
    // complex64: the tile passes through LDS in a swizzled order, slot x at
    // x ^ ((x >> 4) & 15), so that consecutive lanes reach distinct banks also
    // when the register bits are the low tile bits; complex128 in the full
    // variant (the passes with 4x4 gates, which change layout most): slot x at
    // x ^ ((x >> 3) & 7).  Each swizzle is linear and the thread and register
    // parts of x are disjoint, so it splits into a per-lane and a per-register
    // term.
    auto swz = [](uint32_t x) {
      if (sizeof(T) == 4) return x ^ ((x >> 4) & 15u);
      return FULL ? x ^ ((x >> 3) & 7u) : x;
    };
    __syncthreads();
    const uint32_t tw = swz(st.tl);
#pragma unroll
    for (int j = 0; j < NR; ++j) lds[tw ^ swz(st.ri.at(j))] = a.v[j];
    __syncthreads();
    st = make_state<R, TB>(op.u.lay, tid);
    const uint32_t tr = swz(st.tl);
    Amps<T, NR> o;
#pragma unroll
    for (int j = 0; j < NR; ++j) o.v[j] = lds[tr ^ swz(st.ri.at(j))];
    return o;
  }
  const bool tile_ok = op_tile_ok(op, hd, base);
  const uint32_t cl = tile_ok ? op_cl(op, hd) : 0u;
  const uint32_t cvl = tile_ok ? op_cvl(op, hd) : 1u;
  const uint32_t tlc = st.tl & cl;
  if (FULL && kind == Q_U1C) {
    const cx<T> m0 = mat_at<T>(op.u.m, 0), m1 = mat_at<T>(op.u.m, 1), m2 = mat_at<T>(op.u.m, 2), m3 = mat_at<T>(op.u.m, 3);
    return d_u1c<T, NR>(ra, a, tlc, st.ri, cl, cvl, m0, m1, m2, m3);
  }
  if (kind == Q_X || (FULL && kind == Q_Y)) {
    const bool isy = FULL && kind == Q_Y;
    if (cl == 0) {
      // No local control: the gate acts on the whole tile or not at all, so
      // the registers are only moved (no per-lane selects).
      if (cvl != 0) return pass_through(a);
      return d_xy<T, NR>(ra, a, 0u, st.ri, 0u, 0u, isy);
    }
    return d_xy<T, NR>(ra, a, tlc, st.ri, cl, cvl, isy);
  }
  if (FULL && kind == Q_U2) return d_u2<T, NR>(ra * reg_bits_of<NR>() + rb, a, tlc, st.ri, cl, cvl, pool + extra);
  const cx<T> one = mk<T>(1, 0);
  Amps<T, NR> o;
  if (kind == Q_DIAGF) {
    // Fused diagonal (host tables for this layout): the factor of register j
    // on this lane is G * L * R_j, G over the qubits fixed by the tile
    // (uniform), L over the thread bits (per lane), R_j over the register
    // bits (uniform per register).
    if (cvl & ~cl) return pass_through(a);  // a tile-level control is not met
    uint32_t rbits = 0;
#pragma unroll
    for (int s = 0; s < R; ++s) rbits |= st.ri.bm[s];
    const uint32_t clt = cl & ~rbits;
    const bool lane_ok = (st.tl & clt) == (cvl & clt);
    const cx<double> *blk = pool + extra;
    // the lane indices are derived here from an opaque copy of the thread
    // index: hoisted out of the op loop they would hold registers for the
    // whole pass
    uint32_t t = tid;
    asm volatile("" : "+v"(t));
    cx<double> g;
    if (const uint32_t slot = ra >> 16) {
      // staged in LDS at the phase start, the fixed-qubit factor included
      const cx<double> *row = dlt + 32u * (slot - 1u);
      g = cmul(row[t & 15u], row[16u + ((t >> 4) & 15u)]);
    } else {
      g = cmul(blk[NR + (t & 15u)], blk[NR + 16 + ((t >> 4) & 15u)]);
      const cx<double> *tab = blk + NR + 32;
      for (uint32_t c = 0; c < (ra & 0xFFu); ++c, tab += 16) {
        uint32_t idx = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const uint32_t q = op.u.fq[4 * c + i];
          if (q < 64) idx |= static_cast<uint32_t>((base >> q) & 1ull) << i;
        }
        g = cmul(g, tab[idx]);
      }
    }
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      if ((rb >> j) & 1u) {
        asm volatile("");  // keeps the branch (no if-conversion into selects)
        const cx<T> f = to_t<T>(cmul(g, blk[j]));
        o.v[j] = cmul(a.v[j], csel(lane_ok, f, one));
      } else {
        o.v[j] = a.v[j];
      }
    }
    return o;
  }
  // Diagonal kinds.  The predicate (l & cl) == cvl, l = tl | at(j), splits
  // into a lane part (thread tile bits) and a register part (the register
  // tile bits rbits), which is uniform per register j.
  if (kind == Q_D1 || kind == Q_PHASE) {
    if (cvl & ~cl) return pass_through(a);  // a tile-level control is not met
    uint32_t rbits = 0;
#pragma unroll
    for (int s = 0; s < R; ++s) rbits |= st.ri.bm[s];
    const uint32_t clr = cl & rbits, clt = cl & ~rbits;
    const bool lane_ok = (st.tl & clt) == (cvl & clt);
    if (kind == Q_D1) {
      const cx<T> d0 = op_m(op, hd, 0), d1 = op_m(op, hd, 1);
      const uint32_t t = ra;
      if (cl == 0) {
        if ((rbits >> t) & 1u) {
          // target on a register bit: the factor of each register is uniform
#pragma unroll
          for (int j = 0; j < NR; ++j) {
            if constexpr (sizeof(T) == 8) {
              // one test of the register's bit in b and two 64-bit scalar selects
              cx<T> d;
              asm("s_bitcmp1_b32 %[v], %[j]\n\t"
                  "s_cselect_b64 %[dx], %[d1x], %[d0x]\n\t"
                  "s_cselect_b64 %[dy], %[d1y], %[d0y]"
                  : [dx] "=&s"(d.x), [dy] "=&s"(d.y)
                  : [v] "s"(rb), [j] "n"(j), [d0x] "s"(d0.x), [d0y] "s"(d0.y), [d1x] "s"(d1.x),
                    [d1y] "s"(d1.y)
                  : "scc");
              o.v[j] = cmul(a.v[j], d);
            } else {
              o.v[j] = cmul(a.v[j], ((rb >> j) & 1u) ? d1 : d0);
            }
          }
        } else {
          const cx<T> dl = csel(((st.tl >> t) & 1u) != 0, d1, d0);
#pragma unroll
          for (int j = 0; j < NR; ++j) o.v[j] = cmul(a.v[j], dl);
        }
        return o;
      }
#pragma unroll
      for (int j = 0; j < NR; ++j) {
        const uint32_t l = st.tl | st.ri.at(j);
        const cx<T> d = lane_ok && (st.ri.at(j) & clr) == (cvl & clr) ? csel(((l >> t) & 1u) != 0, d1, d0) : one;
        o.v[j] = cmul(a.v[j], d);
      }
      return o;
    }
    // Q_PHASE
    const cx<T> ph = op_m(op, hd, 0);
    if constexpr (sizeof(T) == 8) {
      // The lane part of the predicate selects the factor once; the register
      // part, uniform per register, comes as a mask (bit j of b): a branch
      // (registers it excludes are only copied).
      const cx<T> f = csel(lane_ok, ph, one);
#pragma unroll
      for (int j = 0; j < NR; ++j) {
        if ((rb >> j) & 1u) {
          asm volatile("");  // keeps the branch (no if-conversion into selects)
          o.v[j] = cmul(a.v[j], f);
        } else {
          o.v[j] = a.v[j];
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < NR; ++j) o.v[j] = cmul(a.v[j], csel((tlc | (st.ri.at(j) & cl)) == cvl, ph, one));
    }
    return o;
  }
  if constexpr (!FULL) return pass_through(a);  // not reached: lean programs hold no other kinds
  // Q_DIAG (the last kind, so that every path computes new amplitudes)
  const TileOp &top = tops[extra];
  const uint32_t k = top.k;
  uint32_t ivb = 0;
  for (uint32_t d = 0; d < k; ++d)
    if (top.pos[d] == kNotLocal) ivb |= static_cast<uint32_t>((base >> top.gpos[d]) & 1ull) << d;
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    const uint32_t l = st.tl | st.ri.at(j);
    uint32_t iv = ivb;
    for (uint32_t d = 0; d < k; ++d)
      if (top.pos[d] != kNotLocal) iv |= ((l >> top.pos[d]) & 1u) << d;
    o.v[j] = cmul(a.v[j], csel((l & cl) == cvl, to_t<T>(pool[top.off + iv]), one));
  }
  return o;
}

// The product of the slot's table entries at the tile's fixed-qubit bits,
// in every lane of the slot's eight.
__device__ __forceinline__ cx<double> slot_factor(uint2 rec, const cx<double> *dlane, uint64_t base) {
  uint32_t idx = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t q = (rec.y >> (8 * i)) & 0xFFu;
    if (q < 64) idx |= static_cast<uint32_t>((base >> q) & 1ull) << i;
  }
  cx<double> g = *(dlane - static_cast<int64_t>(rec.x) + idx);
#pragma unroll
  for (int m = 1; m < 8; m <<= 1) g = cmul(g, mk<double>(__shfl_xor(g.x, m), __shfl_xor(g.y, m)));
  return g;
}

// Stages the lane tables of the phase's fused-diagonal slots in LDS:
// dlt[32k + e] = entry e of slot k's lane tables, entries 0-15 times the
// slot's fixed-qubit factor (which lanes 8k..8k+7 of every wave hold); the
// caller waits for the workgroup.  Staging thread i reads the entry at distance
// sb[i / NT] back from the records (read before the tile loads).
template <int NT>
__device__ __forceinline__ void stage_diag_tables(uint2 drec, const cx<double> *dlane, uint32_t nslot,
                                                  uint64_t base, uint32_t tid, const uint32_t (&sb)[256 / NT],
                                                  cx<double> *dlt) {
  const cx<double> g = slot_factor(drec, dlane, base);
#pragma unroll
  for (int j = 0; j < 256 / NT; ++j) {
    const uint32_t i = static_cast<uint32_t>(j) * NT + tid, k = i >> 5;
    const cx<double> gk = mk<double>(__shfl(g.x, 8 * static_cast<int>(k)), __shfl(g.y, 8 * static_cast<int>(k)));
    if (i < 32u * nslot) {
      cx<double> v = *(dlane - static_cast<int64_t>(sb[j]));
      if ((i & 31u) < 16u) v = cmul(v, gk);
      dlt[i] = v;
    }
  }
}

// One tile of a register-layout pass: ops[pd->op_begin .. + pd->nops) are
// ROps; TileOp primitives (Q_LDS, Q_DIAG) index `tops`.  The op list has the
// form
//     LOAD  { register ops }  ( SPILL { LDS ops } FILL { register ops } )*  STORE
// where each register phase holds an even number of ops (Q_NOP pads).
template <typename T, int B, int R, int MATCAP, bool FULL>
__device__ __forceinline__ void reg_tile(cx<T> *__restrict__ sv, const PassDesc *__restrict__ pd,
                                         const ROp<T> *__restrict__ rops,
                                         const TileOp *__restrict__ tops,
                                         const cx<double> *__restrict__ pool,
                                         const ProbeSink &sink, uint64_t tile_id, cx<T> *lds,
                                         cx<T> *mat, uint32_t *doff, double *red, cx<double> *dlt) {
  constexpr int TB = B - R;
  constexpr int NT = 1 << TB;
  constexpr int NR = 1 << R;
  constexpr uint32_t N = 1u << B;
  const uint32_t tid = threadIdx.x;
  uint64_t base = tile_id;
#pragma unroll
  for (int i = 0; i < B; ++i) {
    const uint32_t q = pd->spos[i];
    base = ((base >> q) << (q + 1)) | (base & ((1ull << q) - 1ull));
  }
  base = uniform64(base);  // keeps the tile predicates of the ops on the scalar unit
  const ROp<T> *opl = rops + pd->op_begin;
  uint32_t o = 0;
  for (;;) {
    // Enter a register phase: LOAD from HBM or FILL from LDS; `extra` register
    // ops follow (an even number), then STORE or SPILL.
    const ROp<T> &e = opl[o++];
    if (e.kind == Q_FILL) __syncthreads();
    LayoutState<NR> st = make_state<R, TB>(e.u.lay, tid);
    // The lane records are read before the tile, so that their wait does not
    // wait for the tile loads; the table entries after, overlapping them.
    const uint32_t nslot = e.a;
    const cx<double> *dlane = pool + e.b;
    uint2 drec = make_uint2(0u, 0u);
    uint32_t sb[256 / NT] = {};
    if (nslot) {
      drec = reinterpret_cast<const uint2 *>(dlane)[tid & 63u];
      const uint32_t *stage_back = reinterpret_cast<const uint32_t *>(dlane + 32);
#pragma unroll
      for (int j = 0; j < 256 / NT; ++j) sb[j] = stage_back[j * NT + tid];
    }
    Amps<T, NR> a;
    if (e.kind == Q_LOAD) {
      const RegOff<NR> ro = reg_offset<R>(e.u.lay);
      const cx<T> *p = sv + base + thread_offset<TB>(e.u.lay, tid);
#pragma unroll
      for (int j = 0; j < NR; ++j) a.v[j] = ld_stream(p + ro.gr[j]);
    } else {
#pragma unroll
      for (int j = 0; j < NR; ++j) a.v[j] = lds[st.tl | st.ri.at(j)];
    }
    if (nslot) {
      // A workgroup that loops over tiles may still read the previous tile's
      // tables: wait before overwriting them (a FILL phase waited already).
      if (e.kind == Q_LOAD) __syncthreads();
      stage_diag_tables<NT>(drec, dlane, nslot, base, tid, sb, dlt);
      __syncthreads();
    }
    // Register ops alternate between banks a and b.  The head of the op
    // after the current one is loaded while the current op runs.
    const uint32_t cnt = e.extra;
    const ROp<T> *opp = opl + o;
    OpHead<T, !FULL> h0 = load_head<!FULL>(opp[0]);
    for (uint32_t i = 0; i < cnt; i += 2) {
      const OpHead<T, !FULL> h1 = load_head<!FULL>(opp[1]);
      const Amps<T, NR> b =
          rw_op<T, B, R, FULL>(opp[0], h0, a, st, lds, red, base, tid, tops, pool, sink, dlt);
      h0 = load_head<!FULL>(opp[2]);
      a = rw_op<T, B, R, FULL>(opp[1], h1, b, st, lds, red, base, tid, tops, pool, sink, dlt);
      opp += 2;
    }
    o += cnt;
    const ROp<T> &l = opl[o++];
    if (l.kind & F_MAP) {
      load_map<R>(l, st.ri);
    }
    if (!FULL || (l.kind & ~F_MAP) == Q_STORE) {
      if (l.extra) return;  // read-only pass
      const RegOff<NR> ro = reg_offset_mapped<R>(l.u.lay, st.ri);
      cx<T> *p = sv + base + thread_offset<TB>(l.u.lay, tid);
#pragma unroll
      for (int j = 0; j < NR; ++j) st_stream(p + ro.gr[j], a.v[j]);
      return;
    }
    // SPILL: the tile goes to LDS in natural order for the LDS primitives.
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NR; ++j) lds[st.tl | st.ri.at(j)] = a.v[j];
    while (opl[o].kind == Q_LDS) {
      const TileOp &top = tops[opl[o].extra];
      if ((base & top.cg) == top.cvg) {
        __syncthreads();
        lds_apply<T, NT, B, MATCAP>(top, lds, mat, doff, red, N, base, pool, sink);
      }
      ++o;
    }
  }
}

// Dynamic LDS of the pass kernels (the fused-diagonal lane tables).
extern __shared__ __align__(16) unsigned char aerhip_dyn_lds[];

// Waves per SIMD that the tile's LDS allows (160 KiB per CU, four SIMDs;
// 1 KiB per workgroup for the small arrays beside the tile); register
// allocation is held to that occupancy.  WIDE: the lean variant, which uses
// no small arrays, at five 32 KiB tiles per CU (its complex128 registers fit
// the 96 VGPRs of five waves); launched for short passes, whose time is the
// HBM sweep.
template <typename T, int B, int R, bool WIDE = false>
constexpr int reg_pass_waves() {
  const int blocks = (160 * 1024) / ((static_cast<int>(sizeof(cx<T>)) << B) + (WIDE ? 0 : 1024));
  const int waves = blocks * ((1 << (B - R)) / 64) / 4;
  return waves < 1 ? 1 : waves > 8 ? 8 : waves;
}

// One workgroup per tile.  FULL = false: a lean variant for programs whose
// ops are all LOAD, STORE, RELAYOUT, U1, U1M, X, D1, PHASE and NOP (a single
// register phase); the smaller op loop keeps the frequent gates faster.
template <typename T, int B, int R, bool FULL, bool WIDE = false>
__global__ __launch_bounds__(1 << (B - R))
__attribute__((amdgpu_waves_per_eu(reg_pass_waves<T, B, R, WIDE>()))) void k_reg_pass(
    cx<T> *__restrict__ sv, const PassDesc *__restrict__ pd, const ROp<T> *__restrict__ rops,
    const TileOp *__restrict__ tops, const cx<double> *__restrict__ pool,
    double *__restrict__ probes, uint64_t state_stride, uint32_t pd_stride, uint32_t tile_log,
    uint32_t probes_per_state) {
  static_assert(!WIDE || (!FULL && sizeof(T) == 8), "the wide variant is the lean complex128 one");
  __shared__ cx<T> lds[1 << B];
  __shared__ cx<T> mat[16];
  __shared__ uint32_t doff[32];
  __shared__ double red[(1 << (B - R)) / 64];
  const StateTile st = state_tile(tile_log);
  const ProbeSink sink{probes + st.b * probes_per_state * st.ntiles, st.t, st.ntiles, false};
  // kDiagLdsBytes of dynamic LDS when the pass has fused-diagonal slots
  reg_tile<T, B, R, 16, FULL>(sv + st.b * state_stride, pd + st.b * pd_stride, rops, tops, pool,
                              sink, st.t, lds, mat, doff, red, reinterpret_cast<cx<double> *>(aerhip_dyn_lds));
}

// A layer of SWAPs on disjoint qubit pairs permutes the index bits by an
// involution pi: amplitude i moves to pi(i).  The tile qubits S are closed
// under pi, so tile (S, b) maps onto tile (S, pi(b)) with its local bits
// permuted; one workgroup exchanges such a pair of tiles (or permutes a tile
// with pi(b) = b in place), and the other tile's workgroup has nothing to do,
// so the pass is in place without races.  Both tiles go through LDS: 2^SB
// amplitudes each.
template <typename T, int SB, int NT>
struct QPerm {
  static constexpr int NE = (1 << SB) >= NT ? (1 << SB) / NT : 1;  // elements per thread and tile
  uint32_t tid, ns, npairs, N;
  uint64_t cg;
  uint32_t xl, zl;
  uint32_t sp[SB], px[16], py[16];
  uint64_t et;     // expand() of the thread's part of the local index
  uint32_t wt, rt;  // its LDS slots for the writes and the permuted reads

  // The descriptors, read once into scalar registers (compile-time loop
  // bounds keep them there).
  __device__ __forceinline__ void init(const PassDesc *__restrict__ pdp, const TileOp *__restrict__ opp) {
    tid = threadIdx.x;
    ns = pdp->nb;
    npairs = opp->k;
    cg = opp->cg;
    xl = opp->xl;
    zl = opp->zl;
#pragma unroll
    for (int i = 0; i < SB; ++i) sp[i] = pdp->spos[i];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      px[j] = opp->pos[j];
      py[j] = opp->gpos[j];
    }
    N = 1u << ns;
    const uint32_t tk = tid & (N - 1u);
    et = expand(tk);
    wt = swz(tk);
    rt = swz(lperm(tk));
  }
  // The maps below are linear in the local index: the thread's part once,
  // the part of each step (k = tid + k0, k0 a multiple of NT) per step.
  __device__ __forceinline__ uint64_t expand(uint32_t k) const {
    uint64_t e = 0;
#pragma unroll
    for (int i = 0; i < SB; ++i)
      if (static_cast<uint32_t>(i) < ns) e |= static_cast<uint64_t>((k >> i) & 1u) << sp[i];
    return e;
  }
  __device__ __forceinline__ uint32_t lperm(uint32_t k) const {
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < SB; ++i)
      if (static_cast<uint32_t>(i) < ns) m |= ((k >> i) & 1u) << ((cg >> (4 * i)) & 15u);
    return m;
  }
  // LDS slot of local index k: k ^ g(k) (bank-conflict-free permuted reads)
  __device__ __forceinline__ uint32_t swz(uint32_t k) const {
    uint32_t g = 0;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const uint32_t p = ((j < 4 ? xl : zl) >> (8 * (j & 3))) & 0xFFu;
      if (p < 32) g |= ((k >> p) & 1u) << j;
    }
    return k ^ g;
  }
  // Base of tile t (t with zeros inserted at the tile qubits) and of its
  // partner tile (the pairs outside the tile exchanged); both uniform.
  __device__ __forceinline__ void bases(uint64_t t, uint64_t &b, uint64_t &pb) const {
    b = t;
#pragma unroll
    for (int i = 0; i < SB; ++i)
      if (static_cast<uint32_t>(i) < ns) {
        const uint32_t q = sp[i];
        b = ((b >> q) << (q + 1)) | (b & ((1ull << q) - 1ull));
      }
    pb = b;
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if (static_cast<uint32_t>(j) < npairs && (((b >> px[j]) ^ (b >> py[j])) & 1ull))
        pb ^= (1ull << px[j]) | (1ull << py[j]);
    b = uniform64(b);
    pb = uniform64(pb);
  }
  // This thread's amplitudes of tiles b and pb into registers.
  __device__ __forceinline__ void load(const cx<T> *sv, uint64_t b, uint64_t pb, cx<T> (&va)[NE],
                                       cx<T> (&vb)[NE]) const {
#pragma unroll
    for (int s = 0; s < NE; ++s) {
      const uint32_t k0 = static_cast<uint32_t>(s) * NT;
      if (k0 < N && tid < N) {
        const uint64_t e = et | expand(k0);
        va[s] = ld_stream(sv + b + e);
        if (pb != b) vb[s] = ld_stream(sv + pb + e);
      }
    }
  }
  __device__ __forceinline__ void to_lds(cx<T> *lds, bool pair, const cx<T> (&va)[NE], const cx<T> (&vb)[NE]) const {
#pragma unroll
    for (int s = 0; s < NE; ++s) {
      const uint32_t k0 = static_cast<uint32_t>(s) * NT;
      if (k0 < N && tid < N) {
        const uint32_t w = wt ^ swz(k0);
        lds[w] = va[s];
        if (pair) lds[N + w] = vb[s];
      }
    }
  }
  // Output tile pb at local k = tile b at the permuted local index, and the
  // other way round.
  __device__ __forceinline__ void store(cx<T> *sv, uint64_t b, uint64_t pb, const cx<T> *lds) const {
#pragma unroll
    for (int s = 0; s < NE; ++s) {
      const uint32_t k0 = static_cast<uint32_t>(s) * NT;
      if (k0 < N && tid < N) {
        const uint64_t e = et | expand(k0);
        const uint32_t r = rt ^ swz(lperm(k0));
        st_stream(sv + pb + e, lds[r]);
        if (pb != b) st_stream(sv + b + e, lds[N + r]);
      }
    }
  }
};

// One tile (the batched-circuit kernel's tile loop): the workgroup of tile t
// exchanges it with its partner, or skips it when the partner's base is the
// smaller one (uniform).
template <typename T, int SB, int NT>
__device__ __forceinline__ void qperm_tile(cx<T> *__restrict__ sv, const PassDesc *__restrict__ pdp,
                                           const TileOp *__restrict__ opp, uint64_t t, cx<T> *lds) {
  QPerm<T, SB, NT> q;
  q.init(pdp, opp);
  uint64_t b, pb;
  q.bases(t, b, pb);
  if (pb < b) return;
  cx<T> va[QPerm<T, SB, NT>::NE], vb[QPerm<T, SB, NT>::NE];
  q.load(sv, b, pb, va, vb);
  q.to_lds(lds, pb != b, va, vb);
  __syncthreads();
  q.store(sv, b, pb, lds);
  __syncthreads();  // the LDS tiles are free again
}

// Up to four consecutive tiles per workgroup (about half of all tiles are the
// larger base of their pair and skipped), software-pipelined: the loads of
// the next exchanged tile pair are issued before the stores of the current
// one.
template <typename T, int SB>
__global__ __launch_bounds__(kThreads) void k_qperm(cx<T> *__restrict__ sv, const PassDesc *__restrict__ pd,
                                                    const TileOp *__restrict__ ops, uint64_t state_stride,
                                                    uint32_t pd_stride, uint32_t tile_log,
                                                    uint32_t per_wg_log) {
  using Q = QPerm<T, SB, kThreads>;
  __shared__ cx<T> lds[2 << SB];
  const StateTile st = state_tile(tile_log);
  const PassDesc *d = pd + st.b * pd_stride;
  cx<T> *s = sv + st.b * state_stride;
  Q q;
  q.init(d, ops + d->op_begin);
  const uint32_t count = 1u << per_wg_log;  // at most 4
  uint64_t b[4], pb[4];
  bool act[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    b[i] = pb[i] = 0;
    act[i] = false;
    if (static_cast<uint32_t>(i) < count) {
      q.bases((st.t << per_wg_log) + static_cast<uint32_t>(i), b[i], pb[i]);
      act[i] = pb[i] >= b[i];
    }
  }
  cx<T> va[Q::NE], vb[Q::NE];
  if (act[0]) q.load(s, b[0], pb[0], va, vb);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (act[i]) {
      q.to_lds(lds, pb[i] != b[i], va, vb);
      __syncthreads();
    }
    if (i + 1 < 4 && act[i + 1]) q.load(s, b[i + 1], pb[i + 1], va, vb);  // the registers are free again
    if (act[i]) {
      q.store(s, b[i], pb[i], lds);
      __syncthreads();  // the LDS tiles are free again
    }
  }
}

}  // namespace aerhip
