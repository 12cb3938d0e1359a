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

#include "tile_pass.hpp"

namespace aerhip {

enum RegPassOpKind : uint32_t {
  Q_LOAD = 40,      // HBM -> registers in layout u.lay; extra: register ops that follow
  Q_STORE = 41,     // registers -> HBM in layout u.lay (the current layout)
  Q_RELAYOUT = 42,  // registers -> LDS -> registers in layout u.lay
  Q_SPILL = 43,     // registers -> LDS (the tile is then in LDS)
  Q_FILL = 44,      // LDS -> registers in layout u.lay; extra: register ops that follow
  Q_LDS = 45,       // TileOp primitive extra on the LDS tile
  Q_U1 = 46,        // 2x2 matrix u.m on register bit a
  Q_U1C = 47,       // 2x2 matrix on register bit a where (l & cl) == cvl
  Q_X = 48,         // X on register bit a where (l & cl) == cvl
  Q_Y = 49,         // Y on register bit a where (l & cl) == cvl
  Q_U2 = 50,        // 4x4 matrix pool[extra..] on register bits a (bit 0), b (bit 1)
  Q_D1 = 51,        // diag(u.m[0..1], u.m[2..3]) on local bit a where (l & cl) == cvl
  Q_PHASE = 52,     // multiply by u.m[0..1] where (l & cl) == cvl
  Q_DIAG = 53,      // diagonal of TileOp extra (local and fixed bits)
  Q_PROBE = 54,     // Pauli probe whose X/Y bits are register bits (u.pr)
  Q_NOP = 55,       // keeps the amplitudes (pads a register phase to an even length)
};

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

template <typename T>
struct alignas(128) ROp {
  uint32_t kind;
  uint32_t a, b;
  uint32_t cl, cvl;  // local control mask and value
  uint32_t extra;    // probe slot, TileOp index or pool offset
  uint64_t cg, cvg;  // controls on qubits fixed by the tile
  union {
    T m[8];
    RLayout lay;
    RProbe pr;
  } u;
};
static_assert(sizeof(ROp<double>) == 128 && sizeof(ROp<float>) == 128, "ROp is one 128-byte record");

// m0*a0 + m1*a1 with one multiply and three fused multiply-adds per component.
template <typename T>
__device__ __forceinline__ cx<T> mv2(cx<T> m0, cx<T> a0, cx<T> m1, cx<T> a1) {
  T x = m1.x * a1.x;
  x = __builtin_fma(-m1.y, a1.y, x);
  x = __builtin_fma(m0.x, a0.x, x);
  x = __builtin_fma(-m0.y, a0.y, x);
  T y = m1.x * a1.y;
  y = __builtin_fma(m1.y, a1.x, y);
  y = __builtin_fma(m0.x, a0.y, y);
  y = __builtin_fma(m0.y, a0.x, y);
  return mk<T>(x, y);
}

template <typename T>
__device__ __forceinline__ cx<T> csel(bool c, cx<T> a, cx<T> b) {
  return mk<T>(c ? a.x : b.x, c ? a.y : b.y);
}

// Uniform layout state: rl[j] (tile index) and gr[j] (HBM offset) of register j.
template <int NR>
struct RegIdx {
  uint32_t rl[NR];
};
template <int NR>
struct RegOff {
  uint64_t gr[NR];
};

template <int R>
__device__ __forceinline__ RegIdx<1 << R> reg_index(const RLayout &lay) {
  RegIdx<1 << R> r;
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) {
    uint32_t l = 0;
#pragma unroll
    for (int b = 0; b < R; ++b)
      if (j & (1 << b)) l |= 1u << byte_at(lay.rp, b);
    r.rl[j] = l;
  }
  return r;
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

// ---- register gates, one instantiation per register bit -----------------

template <typename T, int NR, int b>
__device__ __forceinline__ Amps<T, NR> q_u1(Amps<T, NR> a, cx<T> m0, cx<T> m1, cx<T> m2,
                                            cx<T> m3) {
#pragma unroll
  for (int p = 0; p < NR / 2; ++p) {
    const int j = pair_lo<b>(p), j1 = j | (1 << b);
    const cx<T> a0 = a.v[j], a1 = a.v[j1];
    a.v[j] = mv2(m0, a0, m2, a1);
    a.v[j1] = mv2(m1, a0, m3, a1);
  }
  return a;
}

template <typename T, int NR, int b>
__device__ __forceinline__ Amps<T, NR> q_u1c(Amps<T, NR> a, uint32_t tlc, const RegIdx<NR> &ri,
                                             uint32_t cl, uint32_t cvl, cx<T> m0, cx<T> m1,
                                             cx<T> m2, cx<T> m3) {
#pragma unroll
  for (int p = 0; p < NR / 2; ++p) {
    const int j = pair_lo<b>(p), j1 = j | (1 << b);
    const bool on = (tlc | (ri.rl[j] & cl)) == cvl;
    const cx<T> a0 = a.v[j], a1 = a.v[j1];
    a.v[j] = csel(on, mv2(m0, a0, m2, a1), a0);
    a.v[j1] = csel(on, mv2(m1, a0, m3, a1), a1);
  }
  return a;
}

template <typename T, int NR, int b>
__device__ __forceinline__ Amps<T, NR> q_xy(Amps<T, NR> a, uint32_t tlc, const RegIdx<NR> &ri,
                                            uint32_t cl, uint32_t cvl, bool isy) {
#pragma unroll
  for (int p = 0; p < NR / 2; ++p) {
    const int j = pair_lo<b>(p), j1 = j | (1 << b);
    const bool on = (tlc | (ri.rl[j] & cl)) == cvl;
    const cx<T> a0 = a.v[j], a1 = a.v[j1];
    const cx<T> n0 = isy ? mk<T>(a1.y, -a1.x) : a1;
    const cx<T> n1 = isy ? mk<T>(-a0.y, a0.x) : a0;
    a.v[j] = csel(on, n0, a0);
    a.v[j1] = csel(on, n1, a1);
  }
  return a;
}

template <typename T, int NR, int b0, int b1>
__device__ __forceinline__ Amps<T, NR> q_u2(Amps<T, NR> a, uint32_t tlc, const RegIdx<NR> &ri,
                                            uint32_t cl, uint32_t cvl,
                                            const cx<double> *__restrict__ mp) {
  constexpr int lo = b0 < b1 ? b0 : b1, hi = b0 < b1 ? b1 : b0;
  cx<T> m[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) m[e] = to_t<T>(mp[e]);
#pragma unroll
  for (int p = 0; p < NR / 4; ++p) {
    const int j = pair_lo<hi>(pair_lo<lo>(p));
    const int i1 = j | (1 << b0), i2 = j | (1 << b1), i3 = j | (1 << b0) | (1 << b1);
    const bool on = (tlc | (ri.rl[j] & cl)) == cvl;
    const cx<T> x0 = a.v[j], x1 = a.v[i1], x2 = a.v[i2], x3 = a.v[i3];
    cx<T> y[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const cx<T> s01 = mv2(m[r], x0, m[r + 4], x1);
      const cx<T> s23 = mv2(m[r + 8], x2, m[r + 12], x3);
      y[r] = mk<T>(s01.x + s23.x, s01.y + s23.y);
    }
    a.v[j] = csel(on, y[0], x0);
    a.v[i1] = csel(on, y[1], x1);
    a.v[i2] = csel(on, y[2], x2);
    a.v[i3] = csel(on, y[3], x3);
  }
  return a;
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
    if (st ^ (__popc(ri.rl[j] & zl) & 1u)) v = -v;
    acc += v;
  }
  return acc;
}

template <typename T, int NR>
__device__ __forceinline__ Amps<T, NR> q_u1_dispatch(uint32_t b, Amps<T, NR> a, cx<T> m0,
                                                     cx<T> m1, cx<T> m2, cx<T> m3) {
  // b < R; the top register bit is the default so that no path returns `a`.
  if (b == 0) return q_u1<T, NR, 0>(a, m0, m1, m2, m3);
  if (b == 1) return q_u1<T, NR, 1>(a, m0, m1, m2, m3);
  if constexpr (NR == 8) return q_u1<T, NR, 2>(a, m0, m1, m2, m3);
  if constexpr (NR >= 16) if (b == 2) return q_u1<T, NR, 2>(a, m0, m1, m2, m3);
  if constexpr (NR == 16) return q_u1<T, NR, 3>(a, m0, m1, m2, m3);
  if constexpr (NR == 32) {
    if (b == 3) return q_u1<T, NR, 3>(a, m0, m1, m2, m3);
    return q_u1<T, NR, 4>(a, m0, m1, m2, m3);
  }
}

template <typename T, int NR>
__device__ __forceinline__ Amps<T, NR> q_u1c_dispatch(uint32_t b, Amps<T, NR> a, uint32_t tlc,
                                                      const RegIdx<NR> &ri, uint32_t cl,
                                                      uint32_t cvl, cx<T> m0, cx<T> m1,
                                                      cx<T> m2, cx<T> m3) {
  if (b == 0) return q_u1c<T, NR, 0>(a, tlc, ri, cl, cvl, m0, m1, m2, m3);
  if (b == 1) return q_u1c<T, NR, 1>(a, tlc, ri, cl, cvl, m0, m1, m2, m3);
  if constexpr (NR == 8) return q_u1c<T, NR, 2>(a, tlc, ri, cl, cvl, m0, m1, m2, m3);
  if constexpr (NR >= 16) if (b == 2) return q_u1c<T, NR, 2>(a, tlc, ri, cl, cvl, m0, m1, m2, m3);
  if constexpr (NR == 16) return q_u1c<T, NR, 3>(a, tlc, ri, cl, cvl, m0, m1, m2, m3);
  if constexpr (NR == 32) {
    if (b == 3) return q_u1c<T, NR, 3>(a, tlc, ri, cl, cvl, m0, m1, m2, m3);
    return q_u1c<T, NR, 4>(a, tlc, ri, cl, cvl, m0, m1, m2, m3);
  }
}

template <typename T, int NR>
__device__ __forceinline__ Amps<T, NR> q_xy_dispatch(uint32_t b, Amps<T, NR> a, uint32_t tlc,
                                                     const RegIdx<NR> &ri, uint32_t cl,
                                                     uint32_t cvl, bool isy) {
  if (b == 0) return q_xy<T, NR, 0>(a, tlc, ri, cl, cvl, isy);
  if (b == 1) return q_xy<T, NR, 1>(a, tlc, ri, cl, cvl, isy);
  if constexpr (NR == 8) return q_xy<T, NR, 2>(a, tlc, ri, cl, cvl, isy);
  if constexpr (NR >= 16) if (b == 2) return q_xy<T, NR, 2>(a, tlc, ri, cl, cvl, isy);
  if constexpr (NR == 16) return q_xy<T, NR, 3>(a, tlc, ri, cl, cvl, isy);
  if constexpr (NR == 32) {
    if (b == 3) return q_xy<T, NR, 3>(a, tlc, ri, cl, cvl, isy);
    return q_xy<T, NR, 4>(a, tlc, ri, cl, cvl, isy);
  }
}

// Pair (b0, b1) of distinct register bits, enumerated as P = b0 * R + b1; the
// last valid pair is the default so that no path returns `a`.
template <typename T, int NR, int R, int P>
__device__ __forceinline__ Amps<T, NR> q_u2_pick(uint32_t p, Amps<T, NR> a, uint32_t tlc,
                                                 const RegIdx<NR> &ri, uint32_t cl, uint32_t cvl,
                                                 const cx<double> *mp) {
  constexpr int b0 = P / R, b1 = P % R;
  constexpr int last = (R - 1) * R + (R - 2);
  if constexpr (P >= last) {
    return q_u2<T, NR, R - 1, R - 2>(a, tlc, ri, cl, cvl, mp);
  } else if constexpr (b0 == b1) {
    return q_u2_pick<T, NR, R, P + 1>(p, a, tlc, ri, cl, cvl, mp);
  } else {
    if (p == P) return q_u2<T, NR, b0, b1>(a, tlc, ri, cl, cvl, mp);
    return q_u2_pick<T, NR, R, P + 1>(p, a, tlc, ri, cl, cvl, mp);
  }
}

template <typename T, int NR>
__device__ __forceinline__ Amps<T, NR> q_u2_dispatch(uint32_t b0, uint32_t b1, Amps<T, NR> a,
                                                     uint32_t tlc, const RegIdx<NR> &ri,
                                                     uint32_t cl, uint32_t cvl,
                                                     const cx<double> *mp) {
  constexpr int R = NR == 8 ? 3 : NR == 16 ? 4 : 5;
  return q_u2_pick<T, NR, R, 1>(b0 * R + b1, a, tlc, ri, cl, cvl, mp);
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

// A register op that rewrites every amplitude (a layout change or a gate):
// returns the new amplitudes and never passes the input through unchanged,
// so that the op loop can alternate between two register banks without
// copies.  Controls fixed by the tile are folded into the local predicate
// (an unmet one turns it into "never"); a 2x2 matrix becomes the identity.
template <typename T, int B, int R>
__device__ __forceinline__ Amps<T, 1 << R> rw_op(const ROp<T> &op, const Amps<T, 1 << R> &a,
                                                 LayoutState<1 << R> &st, cx<T> *lds,
                                                 double *red, uint64_t base, uint32_t tid,
                                                 const TileOp *__restrict__ tops,
                                                 const cx<double> *__restrict__ pool,
                                                 double *__restrict__ probes) {
  constexpr int NR = 1 << R;
  constexpr int TB = B - R;
  const uint32_t kind = op.kind;
  if (kind == Q_PROBE || kind == Q_NOP) {
    if (kind == Q_PROBE) {
      const RProbe &pr = op.u.pr;
      const uint32_t sgb = __popcll(base & pr.zg) & 1u;
      double acc = q_probe_dispatch<T, NR, 0>(pr.xr, a, st.tl, st.ri, pr.zl, sgb, pr.pr, pr.pi);
      acc = block_sum_nt<(1 << TB)>(acc, red);
      if (tid == 0) probes[static_cast<uint64_t>(op.extra) * gridDim.x + blockIdx.x] = acc;
    }
    return a;
  }
  if (kind == Q_RELAYOUT) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NR; ++j) lds[st.tl | st.ri.rl[j]] = a.v[j];
    __syncthreads();
    st = make_state<R, TB>(op.u.lay, tid);
    Amps<T, NR> o;
#pragma unroll
    for (int j = 0; j < NR; ++j) o.v[j] = lds[st.tl | st.ri.rl[j]];
    return o;
  }
  // Controls fixed by the tile fold into the local predicate: an unmet one
  // makes it "never" (Q_U1 carries no controls; the host emits Q_U1C then).
  const bool tile_ok = (base & op.cg) == op.cvg;
  const uint32_t cl = tile_ok ? op.cl : 0u;
  const uint32_t cvl = tile_ok ? op.cvl : 1u;
  const uint32_t tlc = st.tl & cl;
  if (kind == Q_U1) {
    const cx<T> m0 = mk<T>(op.u.m[0], op.u.m[1]), m1 = mk<T>(op.u.m[2], op.u.m[3]);
    const cx<T> m2 = mk<T>(op.u.m[4], op.u.m[5]), m3 = mk<T>(op.u.m[6], op.u.m[7]);
    return q_u1_dispatch<T, NR>(op.a, a, m0, m1, m2, m3);
  }
  if (kind == Q_U1C) {
    const cx<T> m0 = mk<T>(op.u.m[0], op.u.m[1]), m1 = mk<T>(op.u.m[2], op.u.m[3]);
    const cx<T> m2 = mk<T>(op.u.m[4], op.u.m[5]), m3 = mk<T>(op.u.m[6], op.u.m[7]);
    return q_u1c_dispatch<T, NR>(op.a, a, tlc, st.ri, cl, cvl, m0, m1, m2, m3);
  }
  if (kind == Q_X || kind == Q_Y) return q_xy_dispatch<T, NR>(op.a, a, tlc, st.ri, cl, cvl, kind == Q_Y);
  if (kind == Q_U2) return q_u2_dispatch<T, NR>(op.a, op.b, a, tlc, st.ri, cl, cvl, pool + op.extra);
  const cx<T> one = mk<T>(1, 0);
  Amps<T, NR> o;
  if (kind == Q_D1) {
    const cx<T> d0 = mk<T>(op.u.m[0], op.u.m[1]), d1 = mk<T>(op.u.m[2], op.u.m[3]);
    const uint32_t t = op.a;
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      const uint32_t l = st.tl | st.ri.rl[j];
      const cx<T> d = ((l & cl) == cvl) ? csel(((l >> t) & 1u) != 0, d1, d0) : one;
      o.v[j] = cmul(a.v[j], d);
    }
    return o;
  }
  if (kind == Q_PHASE) {
    const cx<T> ph = mk<T>(op.u.m[0], op.u.m[1]);
#pragma unroll
    for (int j = 0; j < NR; ++j)
      o.v[j] = cmul(a.v[j], csel((tlc | (st.ri.rl[j] & cl)) == cvl, ph, one));
    return o;
  }
  // Q_DIAG (the last kind, so that every path computes new amplitudes)
  const TileOp &top = tops[op.extra];
  const uint32_t k = top.k;
  uint32_t ivb = 0;
  for (uint32_t d = 0; d < k; ++d)
    if (top.pos[d] == kNotLocal) ivb |= static_cast<uint32_t>((base >> top.gpos[d]) & 1ull) << d;
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    const uint32_t l = st.tl | st.ri.rl[j];
    uint32_t iv = ivb;
    for (uint32_t d = 0; d < k; ++d)
      if (top.pos[d] != kNotLocal) iv |= ((l >> top.pos[d]) & 1u) << d;
    o.v[j] = cmul(a.v[j], csel((l & cl) == cvl, to_t<T>(pool[top.off + iv]), one));
  }
  return o;
}

// One tile pass: ops[pd->op_begin .. + pd->nops) are ROps; TileOp primitives
// (Q_LDS, Q_DIAG) index `tops`.  Probe partials go to probes[slot * ntiles +
// tile].  The op list has the form
//     LOAD  { register ops }  ( SPILL { LDS ops } FILL { register ops } )*  STORE
// where each register phase holds an even number of ops (Q_NOP pads).
template <typename T, int B, int R>
__global__ __launch_bounds__(1 << (B - R)) void k_reg_pass(
    cx<T> *__restrict__ sv, const PassDesc *__restrict__ pd, const ROp<T> *__restrict__ rops,
    const TileOp *__restrict__ tops, const cx<double> *__restrict__ pool,
    double *__restrict__ probes) {
  constexpr int TB = B - R;
  constexpr int NT = 1 << TB;
  constexpr int NR = 1 << R;
  constexpr uint32_t N = 1u << B;
  __shared__ cx<T> lds[N];
  __shared__ cx<T> mat[16];
  __shared__ uint32_t doff[32];
  __shared__ double red[NT / 64];

  const uint32_t tid = threadIdx.x;
  // HBM offset of the tile: the workgroup index spread over the qubits
  // outside the tile (zeros inserted at the tile qubits, ascending).
  uint64_t base = blockIdx.x;
#pragma unroll
  for (int i = 0; i < B; ++i) {
    const uint32_t q = pd->spos[i];
    base = ((base >> q) << (q + 1)) | (base & ((1ull << q) - 1ull));
  }
  const ROp<T> *opl = rops + pd->op_begin;
  uint32_t o = 0;
  for (;;) {
    // Enter a register phase: LOAD from HBM or FILL from LDS; `extra` register
    // ops follow (an even number), then STORE or SPILL.
    const ROp<T> &e = opl[o++];
    if (e.kind == Q_FILL) __syncthreads();
    LayoutState<NR> st = make_state<R, TB>(e.u.lay, tid);
    Amps<T, NR> a;
    if (e.kind == Q_LOAD) {
      const RegOff<NR> ro = reg_offset<R>(e.u.lay);
      const cx<T> *p = sv + base + thread_offset<TB>(e.u.lay, tid);
#pragma unroll
      for (int j = 0; j < NR; ++j) a.v[j] = p[ro.gr[j]];
    } else {
#pragma unroll
      for (int j = 0; j < NR; ++j) a.v[j] = lds[st.tl | st.ri.rl[j]];
    }
    // Register ops alternate between banks a and b.
    const uint32_t cnt = e.extra;
    for (uint32_t i = 0; i < cnt; i += 2) {
      const Amps<T, NR> b = rw_op<T, B, R>(opl[o], a, st, lds, red, base, tid, tops, pool, probes);
      a = rw_op<T, B, R>(opl[o + 1], b, st, lds, red, base, tid, tops, pool, probes);
      o += 2;
    }
    const ROp<T> &l = opl[o++];
    if (l.kind == Q_STORE) {
      const RegOff<NR> ro = reg_offset<R>(l.u.lay);
      cx<T> *p = sv + base + thread_offset<TB>(l.u.lay, tid);
#pragma unroll
      for (int j = 0; j < NR; ++j) p[ro.gr[j]] = a.v[j];
      break;
    }
    // SPILL: the tile goes to LDS in natural order for the LDS primitives.
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NR; ++j) lds[st.tl | st.ri.rl[j]] = a.v[j];
    while (opl[o].kind == Q_LDS) {
      __syncthreads();
      lds_apply<T, NT, B, 16>(tops[opl[o].extra], lds, mat, doff, red, N, base, pool, probes,
                              gridDim.x);
      ++o;
    }
  }
}

}  // namespace aerhip
