// Register-blocked tile pass (gfx950).
//
// A workgroup of NT = 2^(B-R) threads owns one tile of 2^B amplitudes.  The
// tile is staged through LDS once per pass; in between, every thread keeps
// 2^R amplitudes in registers.  Which R local qubits are the register qubits
// is the current *layout*: a gate whose mixed qubits are register qubits runs
// entirely in registers (no LDS traffic, no barrier); controls and diagonals
// may sit on any qubit (thread bits, register bits, or bits fixed by the tile).
// The host inserts GATHER (LDS -> registers with a new layout) and SCATTER
// (registers -> LDS) instructions where the next gates need other register
// qubits, and falls back to LDS-resident primitives (the K_* kinds of
// kernels.hpp) for gates that do not fit in registers.
#pragma once

#include "kernels.hpp"

namespace aerhip {

// Register-resident instruction kinds (continue TileOpKind).
enum RegOpKind : uint32_t {
  K_GATHER = 16,    // pos[0..R-1]: register qubits (local bits), pos[R..]: thread bits
  K_SCATTER = 17,   // registers -> LDS with the current layout
  K_RDENSE1 = 18,   // 2x2 on register bit t
  K_RDENSE2 = 19,   // 4x4 on register bits pos[0] (matrix bit 0), pos[1] (matrix bit 1)
  K_RMCX = 20,      // X on register bit t where the controls hold
  K_RMCY = 21,      // Y on register bit t
  K_RDIAG = 22,     // diagonal on k qubits (local pos / fixed gpos)
  K_RMCPHASE = 23,  // phase where (l & cl) == cvl
};

// Register-resident amplitudes and their local tile indices, passed by value
// so that every access uses a compile-time index and the arrays stay in VGPRs.
template <typename T, int NR>
struct Amps {
  cx<T> v[NR];
};
template <int NR>
struct Locs {
  uint32_t v[NR];
};

// Index of the p-th register pair member with bit b clear.
template <int b>
__device__ __forceinline__ constexpr int pair_lo(int p) {
  return ((p >> b) << (b + 1)) | (p & ((1 << b) - 1));
}

// LDS address of tile element l.  The identity: an XOR swizzle of the slot
// within a 256-byte bank row measured slower on both the LDS and the register
// layouts (the address arithmetic costs more than the conflicts it removes).
template <typename T>
__device__ __forceinline__ uint32_t swz(uint32_t l) {
  return l;
}

template <typename T, int NR, int b>
__device__ __forceinline__ Amps<T, NR> r_dense1(Amps<T, NR> a, const Locs<NR> &lr, uint32_t cl,
                                                uint32_t cvl, cx<T> m0, cx<T> m1, cx<T> m2,
                                                cx<T> m3) {
#pragma unroll
  for (int p = 0; p < NR / 2; ++p) {
    const int j = pair_lo<b>(p), j1 = j | (1 << b);
    if ((lr.v[j] & cl) == cvl) {
      const cx<T> a0 = a.v[j], a1 = a.v[j1];
      a.v[j] = cfma(m2, a1, cmul(m0, a0));
      a.v[j1] = cfma(m3, a1, cmul(m1, a0));
    }
  }
  return a;
}

template <typename T, int NR, int b>
__device__ __forceinline__ Amps<T, NR> r_mcx(Amps<T, NR> a, const Locs<NR> &lr, uint32_t cl,
                                             uint32_t cvl, bool isy) {
#pragma unroll
  for (int p = 0; p < NR / 2; ++p) {
    const int j = pair_lo<b>(p), j1 = j | (1 << b);
    if ((lr.v[j] & cl) == cvl) {
      const cx<T> a0 = a.v[j], a1 = a.v[j1];
      a.v[j] = isy ? mk<T>(a1.y, -a1.x) : a1;
      a.v[j1] = isy ? mk<T>(-a0.y, a0.x) : a0;
    }
  }
  return a;
}

// 4x4 on register bits b0 (matrix bit 0) and b1 (matrix bit 1), b0 != b1.
template <typename T, int NR, int b0, int b1>
__device__ __forceinline__ Amps<T, NR> r_dense2(Amps<T, NR> a, const Locs<NR> &lr, uint32_t cl,
                                                uint32_t cvl, const cx<double> *__restrict__ mp) {
  constexpr int lo = b0 < b1 ? b0 : b1, hi = b0 < b1 ? b1 : b0;
  cx<T> m[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) m[e] = to_t<T>(mp[e]);
#pragma unroll
  for (int p = 0; p < NR / 4; ++p) {
    const int j = pair_lo<hi>(pair_lo<lo>(p));
    if ((lr.v[j] & cl) != cvl) continue;
    const int i1 = j | (1 << b0), i2 = j | (1 << b1), i3 = j | (1 << b0) | (1 << b1);
    const cx<T> x0 = a.v[j], x1 = a.v[i1], x2 = a.v[i2], x3 = a.v[i3];
    cx<T> y[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      cx<T> acc = cmul(m[r], x0);
      acc = cfma(m[r + 4], x1, acc);
      acc = cfma(m[r + 8], x2, acc);
      acc = cfma(m[r + 12], x3, acc);
      y[r] = acc;
    }
    a.v[j] = y[0];
    a.v[i1] = y[1];
    a.v[i2] = y[2];
    a.v[i3] = y[3];
  }
  return a;
}

template <typename T, int NR>
__device__ __forceinline__ Amps<T, NR> r_dense1_dispatch(uint32_t b, Amps<T, NR> a,
                                                         const Locs<NR> &lr, uint32_t cl,
                                                         uint32_t cvl, cx<T> m0, cx<T> m1,
                                                         cx<T> m2, cx<T> m3) {
  switch (b) {
    case 0: return r_dense1<T, NR, 0>(a, lr, cl, cvl, m0, m1, m2, m3);
    case 1: return r_dense1<T, NR, 1>(a, lr, cl, cvl, m0, m1, m2, m3);
    case 2: return r_dense1<T, NR, 2>(a, lr, cl, cvl, m0, m1, m2, m3);
    case 3: if constexpr (NR > 8) return r_dense1<T, NR, 3>(a, lr, cl, cvl, m0, m1, m2, m3); break;
    case 4: if constexpr (NR > 16) return r_dense1<T, NR, 4>(a, lr, cl, cvl, m0, m1, m2, m3); break;
    default: break;
  }
  return a;
}

template <typename T, int NR>
__device__ __forceinline__ Amps<T, NR> r_mcx_dispatch(uint32_t b, Amps<T, NR> a,
                                                      const Locs<NR> &lr, uint32_t cl,
                                                      uint32_t cvl, bool isy) {
  switch (b) {
    case 0: return r_mcx<T, NR, 0>(a, lr, cl, cvl, isy);
    case 1: return r_mcx<T, NR, 1>(a, lr, cl, cvl, isy);
    case 2: return r_mcx<T, NR, 2>(a, lr, cl, cvl, isy);
    case 3: if constexpr (NR > 8) return r_mcx<T, NR, 3>(a, lr, cl, cvl, isy); break;
    case 4: if constexpr (NR > 16) return r_mcx<T, NR, 4>(a, lr, cl, cvl, isy); break;
    default: break;
  }
  return a;
}

template <typename T, int NR, int b0>
__device__ __forceinline__ Amps<T, NR> r_dense2_second(uint32_t b1, Amps<T, NR> a,
                                                       const Locs<NR> &lr, uint32_t cl,
                                                       uint32_t cvl,
                                                       const cx<double> *__restrict__ mp) {
  switch (b1) {
    case 0: if constexpr (b0 != 0) return r_dense2<T, NR, b0, 0>(a, lr, cl, cvl, mp); break;
    case 1: if constexpr (b0 != 1) return r_dense2<T, NR, b0, 1>(a, lr, cl, cvl, mp); break;
    case 2: if constexpr (b0 != 2) return r_dense2<T, NR, b0, 2>(a, lr, cl, cvl, mp); break;
    case 3: if constexpr (NR > 8 && b0 != 3) return r_dense2<T, NR, b0, 3>(a, lr, cl, cvl, mp); break;
    case 4: if constexpr (NR > 16 && b0 != 4) return r_dense2<T, NR, b0, 4>(a, lr, cl, cvl, mp); break;
    default: break;
  }
  return a;
}

template <typename T, int NR>
__device__ __forceinline__ Amps<T, NR> r_dense2_dispatch(uint32_t b0, uint32_t b1,
                                                         Amps<T, NR> a, const Locs<NR> &lr,
                                                         uint32_t cl, uint32_t cvl,
                                                         const cx<double> *__restrict__ mp) {
  switch (b0) {
    case 0: return r_dense2_second<T, NR, 0>(b1, a, lr, cl, cvl, mp);
    case 1: return r_dense2_second<T, NR, 1>(b1, a, lr, cl, cvl, mp);
    case 2: return r_dense2_second<T, NR, 2>(b1, a, lr, cl, cvl, mp);
    case 3: if constexpr (NR > 8) return r_dense2_second<T, NR, 3>(b1, a, lr, cl, cvl, mp); break;
    case 4: if constexpr (NR > 16) return r_dense2_second<T, NR, 4>(b1, a, lr, cl, cvl, mp); break;
    default: break;
  }
  return a;
}

// Sum over the workgroup of NT threads (fixed order); valid in thread 0.
template <int NT>
__device__ __forceinline__ double block_sum_nt(double v, double *red) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < NT / 64; ++i) s += red[i];
  __syncthreads();
  return s;
}

// LDS-resident primitive (the tile is in `tile`); every thread takes part.
// K_PROBE writes the tile's share of a Pauli expectation value to
// probes[slot * ntiles + tile].
// MATCAP: complex entries of `mat`; larger matrices are read from the pool.
template <typename T, int NT, int BMAX, int MATCAP = 1024>
__device__ __forceinline__ void lds_apply(const TileOp &op, cx<T> *tile, cx<T> *mat,
                                          uint32_t *doff, double *red, uint32_t N, uint64_t base,
                                          const cx<double> *__restrict__ pool,
                                          double *__restrict__ probes, uint64_t ntiles) {
  constexpr int RMAX = (1 << BMAX) / NT;
  const uint32_t tid = threadIdx.x;
  const uint32_t cl = op.cl, cvl = op.cvl;
  switch (op.kind) {
    case K_PROBE: {
      // Re(phase * conj(psi_l) * (P psi)_l) summed over the tile, in the
      // pairing of QubitVector::expval_pauli (qubitvector.hpp:2300-2347).
      const double pr = op.pr, pi = op.pi;
      const uint32_t sgb = __popcll(base & op.zg) & 1u;
      const uint32_t xl = op.xl, zl = op.zl;
      double acc = 0;
      if (xl == 0) {
        for (uint32_t l = tid; l < N; l += NT) {
          const double x = tile[swz<T>(l)].x, y = tile[swz<T>(l)].y;
          double v = pr * (x * x + y * y);
          if ((__popc(l & zl) & 1u) ^ sgb) v = -v;
          acc += v;
        }
      } else {
        const uint32_t t = op.t;
        for (uint32_t p = tid; p < (N >> 1); p += NT) {
          const uint32_t l0 = insert_zero(p, t), l1 = l0 ^ xl;
          const double x0 = tile[swz<T>(l0)].x, y0 = tile[swz<T>(l0)].y, x1 = tile[swz<T>(l1)].x, y1 = tile[swz<T>(l1)].y;
          const double a1x = pr * x1 - pi * y1, a1y = pr * y1 + pi * x1;
          const double a0x = pr * x0 - pi * y0, a0y = pr * y0 + pi * x0;
          double v0 = a1x * x0 + a1y * y0;
          double v1 = a0x * x1 + a0y * y1;
          if ((__popc(l0 & zl) & 1u) ^ sgb) v0 = -v0;
          if ((__popc(l1 & zl) & 1u) ^ sgb) v1 = -v1;
          acc += v0 + v1;
        }
      }
      acc = block_sum_nt<NT>(acc, red);
      if (tid == 0) probes[(uint64_t)op.npairs * ntiles + blockIdx.x] = acc;
      break;
    }
    case K_DENSE1: {
      const cx<T> m0 = to_t<T>(pool[op.off + 0]), m1 = to_t<T>(pool[op.off + 1]);
      const cx<T> m2 = to_t<T>(pool[op.off + 2]), m3 = to_t<T>(pool[op.off + 3]);
      const uint32_t t = op.t;
      for (uint32_t p = tid; p < (N >> 1); p += NT) {
        const uint32_t l0 = insert_zero(p, t), l1 = l0 | (1u << t);
        if ((l0 & cl) != cvl) continue;
        const cx<T> a0 = tile[swz<T>(l0)], a1 = tile[swz<T>(l1)];
        tile[swz<T>(l0)] = cfma(m2, a1, cmul(m0, a0));
        tile[swz<T>(l1)] = cfma(m3, a1, cmul(m1, a0));
      }
      break;
    }
    case K_DENSE: {
      const uint32_t k = op.k, D = 1u << k;
      uint32_t tmask = 0;
      for (uint32_t b = 0; b < k; ++b) tmask |= 1u << op.pos[b];
      const bool staged = D * D <= static_cast<uint32_t>(MATCAP);
      if (staged)
        for (uint32_t e = tid; e < D * D; e += NT) mat[e] = to_t<T>(pool[op.off + e]);
      if (tid < D) {
        uint32_t d = 0;
        for (uint32_t b = 0; b < k; ++b) d |= ((tid >> b) & 1u) << op.pos[b];
        doff[tid] = d;
      }
      __syncthreads();
      cx<T> out[RMAX];
#pragma unroll
      for (int r = 0; r < RMAX; ++r) {
        const uint32_t l = tid + r * NT;
        if (l < N) {
          if ((l & cl) != cvl) {
            out[r] = tile[swz<T>(l)];
          } else {
            uint32_t i = 0;
            for (uint32_t b = 0; b < k; ++b) i |= ((l >> op.pos[b]) & 1u) << b;
            const uint32_t lb = l & ~tmask;
            cx<T> acc = mk<T>(0, 0);
            if (staged)
              for (uint32_t j = 0; j < D; ++j) acc = cfma(mat[i + D * j], tile[swz<T>(lb | doff[j])], acc);
            else
              for (uint32_t j = 0; j < D; ++j)
                acc = cfma(to_t<T>(pool[op.off + i + D * j]), tile[swz<T>(lb | doff[j])], acc);
            out[r] = acc;
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < RMAX; ++r) {
        const uint32_t l = tid + r * NT;
        if (l < N) tile[swz<T>(l)] = out[r];
      }
      break;
    }
    case K_DIAG: {
      const uint32_t k = op.k;
      uint32_t ivb = 0;
      for (uint32_t j = 0; j < k; ++j)
        if (op.pos[j] == kNotLocal) ivb |= (uint32_t)((base >> op.gpos[j]) & 1ull) << j;
      for (uint32_t l = tid; l < N; l += NT) {
        if ((l & cl) != cvl) continue;
        uint32_t iv = ivb;
        for (uint32_t j = 0; j < k; ++j)
          if (op.pos[j] != kNotLocal) iv |= ((l >> op.pos[j]) & 1u) << j;
        tile[swz<T>(l)] = cmul(tile[swz<T>(l)], to_t<T>(pool[op.off + iv]));
      }
      break;
    }
    case K_MCX:
    case K_MCY: {
      const uint32_t t = op.t;
      const bool isy = op.kind == K_MCY;
      for (uint32_t p = tid; p < (N >> 1); p += NT) {
        const uint32_t l0 = insert_zero(p, t), l1 = l0 | (1u << t);
        if ((l0 & cl) != cvl) continue;
        const cx<T> a0 = tile[swz<T>(l0)], a1 = tile[swz<T>(l1)];
        if (isy) {
          tile[swz<T>(l0)] = mk<T>(a1.y, -a1.x);
          tile[swz<T>(l1)] = mk<T>(-a0.y, a0.x);
        } else {
          tile[swz<T>(l0)] = a1;
          tile[swz<T>(l1)] = a0;
        }
      }
      break;
    }
    case K_MCSWAP: {
      const uint32_t pa = op.pos[0], pb = op.pos[1];
      const uint32_t lo = pa < pb ? pa : pb, hi = pa < pb ? pb : pa;
      for (uint32_t p = tid; p < (N >> 2); p += NT) {
        const uint32_t l = insert_zero(insert_zero(p, lo), hi);
        if ((l & cl) != cvl) continue;
        const uint32_t la = l | (1u << pa), lb = l | (1u << pb);
        const cx<T> va = tile[swz<T>(la)];
        tile[swz<T>(la)] = tile[swz<T>(lb)];
        tile[swz<T>(lb)] = va;
      }
      break;
    }
    case K_MCPHASE: {
      const cx<T> ph = mk<T>(static_cast<T>(op.pr), static_cast<T>(op.pi));
      for (uint32_t l = tid; l < N; l += NT)
        if ((l & cl) == cvl) tile[swz<T>(l)] = cmul(tile[swz<T>(l)], ph);
      break;
    }
    case K_PAULI: {
      const cx<T> ph = mk<T>(static_cast<T>(op.pr), static_cast<T>(op.pi));
      const uint32_t sgb = __popcll(base & op.zg) & 1u;
      const uint32_t xl = op.xl, zl = op.zl;
      if (xl == 0) {
        for (uint32_t l = tid; l < N; l += NT) {
          cx<T> v = tile[swz<T>(l)];
          if ((__popc(l & zl) & 1u) ^ sgb) v = cneg(v);
          tile[swz<T>(l)] = cmul(v, ph);
        }
      } else {
        const uint32_t t = op.t;
        for (uint32_t p = tid; p < (N >> 1); p += NT) {
          const uint32_t l0 = insert_zero(p, t), l1 = l0 ^ xl;
          cx<T> v0 = tile[swz<T>(l1)], v1 = tile[swz<T>(l0)];
          if ((__popc(l0 & zl) & 1u) ^ sgb) v0 = cneg(v0);
          if ((__popc(l1 & zl) & 1u) ^ sgb) v1 = cneg(v1);
          tile[swz<T>(l0)] = cmul(v0, ph);
          tile[swz<T>(l1)] = cmul(v1, ph);
        }
      }
      break;
    }
    case K_PERM: {
      const uint32_t k = op.k, D = 1u << k;
      uint32_t tmask = 0;
      for (uint32_t b = 0; b < k; ++b) tmask |= 1u << op.pos[b];
      if (tid < D) {
        uint32_t d = 0;
        for (uint32_t b = 0; b < k; ++b) d |= ((tid >> b) & 1u) << op.pos[b];
        doff[tid] = d;
      }
      __syncthreads();
      for (uint32_t l = tid; l < N; l += NT) {
        if (l & tmask) continue;
        if ((l & cl) != cvl) continue;
        for (uint32_t q = 0; q < op.npairs; ++q) {
          const cx<double> pr = pool[op.off + q];
          const uint32_t ia = l | doff[(uint32_t)pr.x], ib = l | doff[(uint32_t)pr.y];
          const cx<T> va = tile[swz<T>(ia)];
          tile[swz<T>(ia)] = tile[swz<T>(ib)];
          tile[swz<T>(ib)] = va;
        }
      }
      break;
    }
    default:
      break;
  }
}

template <typename T, int B, int R>
__global__ __launch_bounds__(1 << (B - R)) void k_tile_reg(
    cx<T> *__restrict__ sv, const PassDesc *__restrict__ pd, const TileOp *__restrict__ ops,
    const cx<double> *__restrict__ pool, double *__restrict__ probes) {
  constexpr int NT = 1 << (B - R);
  constexpr int NR = 1 << R;
  __shared__ cx<T> tile[1 << B];
  __shared__ uint64_t hioff[256];
  __shared__ cx<T> mat[1024];
  __shared__ uint32_t doff[32];
  __shared__ double red[NT / 64];

  const uint32_t tid = threadIdx.x;
  const uint32_t nb = pd->nb, c = pd->c;
  const uint32_t N = 1u << nb;
  const uint32_t ncomp = pd->ncomp;
  uint64_t base = 0;
  const uint64_t tile_id = blockIdx.x;
  for (uint32_t i = 0; i < ncomp; ++i) base |= ((tile_id >> i) & 1ull) << pd->comp[i];
  const uint32_t nh = nb - c;
  for (uint32_t h = tid; h < (1u << nh); h += NT) {
    uint64_t o = 0;
    for (uint32_t j = 0; j < nh; ++j) o |= (uint64_t)((h >> j) & 1u) << pd->spos[c + j];
    hioff[h] = o;
  }
  __syncthreads();
  const uint32_t cmask = (1u << c) - 1u;
  for (uint32_t l = tid; l < N; l += NT) tile[swz<T>(l)] = sv[base + (l & cmask) + hioff[l >> c]];

  Amps<T, NR> a;
  Locs<NR> lr;
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    a.v[j] = mk<T>(0, 0);
    lr.v[j] = 0;
  }
  const uint32_t nops = pd->nops;
  const TileOp *opl = ops + pd->op_begin;
  for (uint32_t o = 0; o < nops; ++o) {
    const TileOp &op = opl[o];
    const uint32_t kind = op.kind;
    if (kind == K_GATHER) {
      __syncthreads();
      uint32_t lb = 0;
      for (uint32_t i = 0; i < nb - R; ++i) lb |= ((tid >> i) & 1u) << op.pos[R + i];
#pragma unroll
      for (int j = 0; j < NR; ++j) {
        uint32_t l = lb;
#pragma unroll
        for (int b = 0; b < R; ++b)
          if (j & (1 << b)) l |= 1u << op.pos[b];
        lr.v[j] = l;
        a.v[j] = tile[swz<T>(l)];
      }
      continue;
    }
    if (kind == K_SCATTER) {
#pragma unroll
      for (int j = 0; j < NR; ++j) tile[swz<T>(lr.v[j])] = a.v[j];
      __syncthreads();
      continue;
    }
    if ((base & op.cg) != op.cvg) continue;  // uniform over the workgroup
    if (kind >= K_RDENSE1) {
      const uint32_t cl = op.cl, cvl = op.cvl;
      switch (kind) {
        case K_RDENSE1:
          a = r_dense1_dispatch<T, NR>(op.t, a, lr, cl, cvl, to_t<T>(pool[op.off]),
                                  to_t<T>(pool[op.off + 1]), to_t<T>(pool[op.off + 2]),
                                  to_t<T>(pool[op.off + 3]));
          break;
        case K_RDENSE2:
          a = r_dense2_dispatch<T, NR>(op.pos[0], op.pos[1], a, lr, cl, cvl, pool + op.off);
          break;
        case K_RMCX:
        case K_RMCY:
          a = r_mcx_dispatch<T, NR>(op.t, a, lr, cl, cvl, kind == K_RMCY);
          break;
        case K_RDIAG: {
          const uint32_t k = op.k;
          uint32_t ivb = 0;
          for (uint32_t d = 0; d < k; ++d)
            if (op.pos[d] == kNotLocal) ivb |= (uint32_t)((base >> op.gpos[d]) & 1ull) << d;
#pragma unroll
          for (int j = 0; j < NR; ++j) {
            const uint32_t l = lr.v[j];
            if ((l & cl) != cvl) continue;
            uint32_t iv = ivb;
            for (uint32_t d = 0; d < k; ++d)
              if (op.pos[d] != kNotLocal) iv |= ((l >> op.pos[d]) & 1u) << d;
            a.v[j] = cmul(a.v[j], to_t<T>(pool[op.off + iv]));
          }
          break;
        }
        case K_RMCPHASE: {
          const cx<T> ph = mk<T>(static_cast<T>(op.pr), static_cast<T>(op.pi));
#pragma unroll
          for (int j = 0; j < NR; ++j)
            if ((lr.v[j] & cl) == cvl) a.v[j] = cmul(a.v[j], ph);
          break;
        }
        default:
          break;
      }
      continue;
    }
    __syncthreads();
    lds_apply<T, NT, B>(op, tile, mat, doff, red, N, base, pool, probes, gridDim.x);
  }
  __syncthreads();
  for (uint32_t l = tid; l < N; l += NT) sv[base + (l & cmask) + hioff[l >> c]] = tile[swz<T>(l)];
}

}  // namespace aerhip
