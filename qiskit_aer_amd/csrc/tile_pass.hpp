// LDS tile pass (gfx950).
//
// A workgroup owns one tile of 2^nb amplitudes: the tile qubits S of the
// pass (S holds the lowest qubits, so every HBM access is a run of
// consecutive amplitudes); every other qubit is fixed by the tile index.  The
// tile is loaded into LDS, the pass's TileOp primitives are applied one after
// the other with a barrier in between, and the tile is written back: one HBM
// read and one HBM write for the whole list.  lds_apply() is also the
// fallback of the register-layout pass (reg_pass.hpp) for primitives that do
// not run in registers.
#pragma once

#include "kernels.hpp"

namespace aerhip {

// Tile of this workgroup.  Workgroups are dealt to the 8 XCDs round-robin
// (workgroup w runs on XCD w % 8); giving each XCD a contiguous range of
// tiles keeps the concurrent accesses of one XCD (and its L2) on neighbouring
// memory: +6-16 % sweep bandwidth on this part.  A bijection when the grid is
// a multiple of 8 (the identity otherwise).
__device__ __forceinline__ uint64_t xcd_tile(uint32_t w, uint32_t grid) {
  if (grid & 7u) return w;
  return static_cast<uint64_t>(w & 7u) * (grid >> 3) + (w >> 3);
}


// Register-resident amplitudes, passed by value so that every access uses a
// compile-time index and the array stays in VGPRs.
template <typename T, int NR>
struct Amps {
  cx<T> v[NR];
};

// Index of the p-th register pair member with bit b clear.
template <int b>
__device__ __forceinline__ constexpr int pair_lo(int p) {
  return ((p >> b) << (b + 1)) | (p & ((1 << b) - 1));
}

// Where a tile's share of a saved expectation value goes: p[slot * ntiles +
// tile] (one workgroup per tile; the host sums the row in tile order), or,
// with one workgroup walking every tile of a state, p[slot] += share.
struct ProbeSink {
  double *p;
  uint64_t tile, ntiles;
  bool accumulate;
  __device__ __forceinline__ void put(uint32_t slot, double v) const {
    if (accumulate)
      p[slot] += v;
    else
      p[static_cast<uint64_t>(slot) * ntiles + tile] = v;
  }
};

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

// Dense matrix M on k = 3..5 tile qubits without local controls, applied to
// the LDS tile as a GEMM on MFMA: with D = 2^k and C = N / D columns (the
// other tile bits), Out = R V, R the 2D x 2D real form of M and V the 2D x C
// real form of the tile (rows re / im of each amplitude of a column).  R V is
// cut into 16 x 16 output blocks, each a chain of 2D / 4 MFMA issues with
// K = 4 (v_mfma_f32_16x16x4f32 for complex64, v_mfma_f64_16x16x4f64 for
// complex128): lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16].
// The real K index 4 kk + l / 16 is (column 2 kk + (l / 16 >> 1) of M,
// part l / 16 & 1), and the real rows are ordered so that each lane's four
// outputs are the (re, im) of two amplitudes of its column: complex64 blocks
// leave D[4 (l / 16) + r][l % 16] in lane l, complex128 blocks
// D[l / 16 + 4 r][l % 16].  `doff[c]` (c < D) holds the tile offset of row c.
// Returns false (nothing done) when the tile is too small for 16-wide blocks.
template <typename T, int NT, int BMAX>
__device__ __forceinline__ bool lds_gemm(const TileOp &op, cx<T> *tile, uint32_t N,
                                         const cx<double> *__restrict__ pool, const uint32_t *doff) {
  constexpr uint32_t NW = NT / 64;
  constexpr int RES = ((1 << BMAX) / 128 / NW) > 0 ? (1 << BMAX) / 128 / NW : 1;
  const uint32_t k = op.k, D = 1u << k, Dr = 2u * D;
  const uint32_t C = N >> k;
  if (k < 3 || k > 5 || C < 16 || op.cl != 0) return false;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t kq = lane >> 4, l16 = lane & 15u;
  uint32_t tmask = 0;
  for (uint32_t b = 0; b < k; ++b) tmask |= 1u << op.pos[b];
  const uint32_t row_tiles = Dr / 16u, ntiles = row_tiles * (C / 16u), ksteps = Dr / 4u;
  // this lane's row of A (complex row within the block and part)
  uint32_t arow, apart;
  if constexpr (sizeof(T) == 4) {
    arow = l16 >> 1;
    apart = l16 & 1u;
  } else {
    arow = 2u * (l16 & 3u) + ((l16 >> 2) >> 1);
    apart = (l16 >> 2) & 1u;
  }
  cx<T> res[RES][2];
  uint32_t coffs[RES];
  uint32_t s = 0;
  for (uint32_t t = wave; t < ntiles; t += NW, ++s) {
    const uint32_t ti = t % row_tiles, tj = t / row_tiles;
    // tile offset of column j = 16 tj + l16: its bits deposited on the
    // non-target tile bits, ascending
    uint32_t j = 16u * tj + l16, coff = 0;
    for (uint32_t b = 0; j && (1u << b) < N; ++b)
      if (!(tmask & (1u << b))) {
        coff |= (j & 1u) << b;
        j >>= 1;
      }
    const uint32_t c = 8u * ti + arow;
    typedef T v4 __attribute__((ext_vector_type(4)));
    v4 acc = {T(0), T(0), T(0), T(0)};
    for (uint32_t kk = 0; kk < ksteps; ++kk) {
      const uint32_t cc = 2u * kk + (kq >> 1), q = kq & 1u;
      const cx<double> m = pool[op.off + c + D * cc];
      const double av = apart ? (q ? m.x : m.y) : (q ? -m.y : m.x);
      const cx<T> v = tile[doff[cc] | coff];
      const T bv = q ? v.y : v.x;
      if constexpr (sizeof(T) == 4)
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(static_cast<float>(av), bv, acc, 0, 0, 0);
      else
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
    }
    if (s < static_cast<uint32_t>(RES)) {
      res[s][0] = mk<T>(acc[0], acc[1]);
      res[s][1] = mk<T>(acc[2], acc[3]);
      coffs[s] = coff | (ti << 16);
    }
  }
  __syncthreads();  // every read of the tile is done before the writes
  s = 0;
  for (uint32_t t = wave; t < ntiles && s < static_cast<uint32_t>(RES); t += NW, ++s) {
    const uint32_t coff = coffs[s] & 0xFFFFu, ti = coffs[s] >> 16;
    const uint32_t c0 = 8u * ti + 2u * kq;  // the lane's two output rows
    tile[doff[c0] | coff] = res[s][0];
    tile[doff[c0 + 1] | coff] = res[s][1];
  }
  return true;
}

// LDS-resident primitive (the tile is in `tile`); every thread takes part.
// K_PROBE hands the tile's share of a Pauli expectation value to `sink`.
// MATCAP: complex entries of `mat`; larger matrices are read from the pool.
template <typename T, int NT, int BMAX, int MATCAP = 1024>
__device__ __forceinline__ void lds_apply(const TileOp &op, cx<T> *tile, cx<T> *mat,
                                          uint32_t *doff, double *red, uint32_t N, uint64_t base,
                                          const cx<double> *__restrict__ pool,
                                          const ProbeSink &sink) {
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
          const double x = tile[(l)].x, y = tile[(l)].y;
          double v = pr * (x * x + y * y);
          if ((__popc(l & zl) & 1u) ^ sgb) v = -v;
          acc += v;
        }
      } else {
        const uint32_t t = op.t;
        for (uint32_t p = tid; p < (N >> 1); p += NT) {
          const uint32_t l0 = insert_zero(p, t), l1 = l0 ^ xl;
          const double x0 = tile[(l0)].x, y0 = tile[(l0)].y, x1 = tile[(l1)].x, y1 = tile[(l1)].y;
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
      if (tid == 0) sink.put(op.npairs, acc);
      break;
    }
    case K_DENSE1: {
      const cx<T> m0 = to_t<T>(pool[op.off + 0]), m1 = to_t<T>(pool[op.off + 1]);
      const cx<T> m2 = to_t<T>(pool[op.off + 2]), m3 = to_t<T>(pool[op.off + 3]);
      const uint32_t t = op.t;
      for (uint32_t p = tid; p < (N >> 1); p += NT) {
        const uint32_t l0 = insert_zero(p, t), l1 = l0 | (1u << t);
        if ((l0 & cl) != cvl) continue;
        const cx<T> a0 = tile[(l0)], a1 = tile[(l1)];
        tile[(l0)] = cfma(m2, a1, cmul(m0, a0));
        tile[(l1)] = cfma(m3, a1, cmul(m1, a0));
      }
      break;
    }
    case K_DENSE: {
      const uint32_t k = op.k, D = 1u << k;
      uint32_t tmask = 0;
      for (uint32_t b = 0; b < k; ++b) tmask |= 1u << op.pos[b];
      if (k >= 3 && k <= 5 && cl == 0 && (N >> k) >= 16) {
        // matrices on 3 to 5 tile qubits: a GEMM on MFMA
        if (tid < D) {
          uint32_t d = 0;
          for (uint32_t b = 0; b < k; ++b) d |= ((tid >> b) & 1u) << op.pos[b];
          doff[tid] = d;
        }
        __syncthreads();
        lds_gemm<T, NT, BMAX>(op, tile, N, pool, doff);
        break;
      }
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
            out[r] = tile[(l)];
          } else {
            uint32_t i = 0;
            for (uint32_t b = 0; b < k; ++b) i |= ((l >> op.pos[b]) & 1u) << b;
            const uint32_t lb = l & ~tmask;
            cx<T> acc = mk<T>(0, 0);
            if (staged)
              for (uint32_t j = 0; j < D; ++j) acc = cfma(mat[i + D * j], tile[(lb | doff[j])], acc);
            else
              for (uint32_t j = 0; j < D; ++j)
                acc = cfma(to_t<T>(pool[op.off + i + D * j]), tile[(lb | doff[j])], acc);
            out[r] = acc;
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < RMAX; ++r) {
        const uint32_t l = tid + r * NT;
        if (l < N) tile[(l)] = out[r];
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
        tile[(l)] = cmul(tile[(l)], to_t<T>(pool[op.off + iv]));
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
        const cx<T> a0 = tile[(l0)], a1 = tile[(l1)];
        if (isy) {
          tile[(l0)] = mk<T>(a1.y, -a1.x);
          tile[(l1)] = mk<T>(-a0.y, a0.x);
        } else {
          tile[(l0)] = a1;
          tile[(l1)] = a0;
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
        const cx<T> va = tile[(la)];
        tile[(la)] = tile[(lb)];
        tile[(lb)] = va;
      }
      break;
    }
    case K_MCPHASE: {
      const cx<T> ph = mk<T>(static_cast<T>(op.pr), static_cast<T>(op.pi));
      for (uint32_t l = tid; l < N; l += NT)
        if ((l & cl) == cvl) tile[(l)] = cmul(tile[(l)], ph);
      break;
    }
    case K_PAULI: {
      const cx<T> ph = mk<T>(static_cast<T>(op.pr), static_cast<T>(op.pi));
      const uint32_t sgb = __popcll(base & op.zg) & 1u;
      const uint32_t xl = op.xl, zl = op.zl;
      if (xl == 0) {
        for (uint32_t l = tid; l < N; l += NT) {
          cx<T> v = tile[(l)];
          if ((__popc(l & zl) & 1u) ^ sgb) v = cneg(v);
          tile[(l)] = cmul(v, ph);
        }
      } else {
        const uint32_t t = op.t;
        for (uint32_t p = tid; p < (N >> 1); p += NT) {
          const uint32_t l0 = insert_zero(p, t), l1 = l0 ^ xl;
          cx<T> v0 = tile[(l1)], v1 = tile[(l0)];
          if ((__popc(l0 & zl) & 1u) ^ sgb) v0 = cneg(v0);
          if ((__popc(l1 & zl) & 1u) ^ sgb) v1 = cneg(v1);
          tile[(l0)] = cmul(v0, ph);
          tile[(l1)] = cmul(v1, ph);
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
          const cx<T> va = tile[(ia)];
          tile[(ia)] = tile[(ib)];
          tile[(ib)] = va;
        }
      }
      break;
    }
    default:
      break;
  }
}

// HBM offset of tile `tile_id`: its bits spread over the qubits outside the
// tile (zeros inserted at the tile qubits, ascending).
__device__ __forceinline__ uint64_t tile_base(const PassDesc *pd, uint64_t tile_id, uint32_t nb) {
  uint64_t base = tile_id;
  for (uint32_t i = 0; i < nb; ++i) {
    const uint32_t q = pd->spos[i];
    base = ((base >> q) << (q + 1)) | (base & ((1ull << q) - 1ull));
  }
  return base;
}

// One tile of an LDS pass, by a workgroup of NT threads.  `hioff` holds
// 2^(nb - c) entries (c: the run of lowest qubits in the tile).
template <typename T, int NT, int BMAX, int MATCAP>
__device__ __forceinline__ void lds_tile(cx<T> *__restrict__ sv, const PassDesc *__restrict__ pd,
                                         const TileOp *__restrict__ ops,
                                         const cx<double> *__restrict__ pool,
                                         const ProbeSink &sink, uint64_t tile_id, cx<T> *tile,
                                         uint64_t *hioff, cx<T> *mat, uint32_t *doff,
                                         double *red) {
  const uint32_t tid = threadIdx.x;
  const uint32_t nb = pd->nb, c = pd->c;
  const uint32_t N = 1u << nb;
  const uint64_t base = tile_base(pd, tile_id, nb);
  const uint32_t nh = nb - c;
  __syncthreads();  // the previous tile of this workgroup is done with LDS
  for (uint32_t h = tid; h < (1u << nh); h += NT) {
    uint64_t o = 0;
    for (uint32_t j = 0; j < nh; ++j) o |= static_cast<uint64_t>((h >> j) & 1u) << pd->spos[c + j];
    hioff[h] = o;
  }
  __syncthreads();
  const uint32_t cmask = (1u << c) - 1u;
  for (uint32_t l = tid; l < N; l += NT) tile[l] = sv[base + (l & cmask) + hioff[l >> c]];
  const uint32_t nops = pd->nops;
  const TileOp *opl = ops + pd->op_begin;
  for (uint32_t o = 0; o < nops; ++o) {
    const TileOp &op = opl[o];
    if ((base & op.cg) != op.cvg) continue;  // uniform over the workgroup
    __syncthreads();
    lds_apply<T, NT, BMAX, MATCAP>(op, tile, mat, doff, red, N, base, pool, sink);
  }
  if (pd->readonly) return;
  __syncthreads();
  for (uint32_t l = tid; l < N; l += NT) sv[base + (l & cmask) + hioff[l >> c]] = tile[l];
}

// A saved Pauli expectation value whose X/Y qubits do not fit a tile: block
// `blk` of `nblk` equal slices of k_expval_pauli's iteration space (pairs
// idx0 = insert-0 at the pivot, idx1 = idx0 ^ X; or single amplitudes when X
// is empty).  op: K_PROBE with cg = X mask, zg = Z mask (all qubits), t = the
// pivot.  Returns the slice's share in thread 0.
template <typename T, int NT>
__device__ __forceinline__ double probe_sweep_block(const cx<T> *__restrict__ sv, uint64_t size,
                                                    const TileOp &op, uint64_t blk,
                                                    uint64_t nblk, double *red) {
  const uint64_t xmask = op.cg, zmask = op.zg;
  const uint32_t xmax = op.t;
  const double pr = op.pr, pi = op.pi;
  const uint64_t work = xmask ? size >> 1 : size;
  const uint64_t chunk = (work + nblk - 1) / nblk;
  const uint64_t lo = blk * chunk;
  const uint64_t hi = lo + chunk < work ? lo + chunk : work;
  double acc = 0;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += NT) {
    if (xmask == 0) {
      const double x = sv[i].x, y = sv[i].y;
      double v = pr * (x * x + y * y);
      if (__popcll(i & zmask) & 1) v = -v;
      acc += v;
    } else {
      const uint64_t i0 = insert_zero64(i, xmax), i1 = i0 ^ xmask;
      const double x0 = sv[i0].x, y0 = sv[i0].y, x1 = sv[i1].x, y1 = sv[i1].y;
      const double a1x = pr * x1 - pi * y1, a1y = pr * y1 + pi * x1;
      const double a0x = pr * x0 - pi * y0, a0y = pr * y0 + pi * x0;
      double v0 = a1x * x0 + a1y * y0;
      double v1 = a0x * x1 + a0y * y1;
      if (__popcll(i0 & zmask) & 1) v0 = -v0;
      if (__popcll(i1 & zmask) & 1) v1 = -v1;
      acc += v0 + v1;
    }
  }
  return block_sum_nt<NT>(acc, red);
}

// Workgroup -> (state, tile) of a pass over `count` states of one size that
// run the same pass index: workgroup w (after the XCD-aware order) takes tile
// w mod 2^tile_log of state w >> tile_log.  State b's amplitudes start at
// b * state_stride, its pass descriptor at pd + b * pd_stride, its probe
// partials at probes + b * probes_per_state * 2^tile_log.  A single state is
// count 1 (strides unused).
struct StateTile {
  uint64_t b, t, ntiles;
};
__device__ __forceinline__ StateTile state_tile(uint32_t tile_log) {
  const uint64_t w = xcd_tile(blockIdx.x, gridDim.x);
  return StateTile{w >> tile_log, w & ((1ull << tile_log) - 1ull), 1ull << tile_log};
}

// One workgroup per slice (2^tile_log slices per state):
// probes[(b * probes_per_state + slot) * 2^tile_log + slice].  The probe of
// state b is ops[pd_b->op_begin], pd_b = pd + b * pd_stride.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_probe_sweep(const cx<T> *__restrict__ sv,
                                                          uint64_t size,
                                                          const PassDesc *__restrict__ pd,
                                                          const TileOp *__restrict__ ops,
                                                          double *__restrict__ probes,
                                                          uint64_t state_stride, uint32_t pd_stride,
                                                          uint32_t tile_log, uint32_t probes_per_state) {
  __shared__ double red[kThreads / 64];
  const StateTile st = state_tile(tile_log);
  const TileOp &o = ops[pd[st.b * pd_stride].op_begin];
  const double v = probe_sweep_block<T, kThreads>(sv + st.b * state_stride, size, o, st.t, st.ntiles, red);
  if (threadIdx.x == 0)
    probes[(st.b * probes_per_state + o.npairs) * st.ntiles + st.t] = v;
}

// One workgroup per tile.
template <typename T, int BMAX>
__global__ __launch_bounds__(kThreads) void k_tile_lds(cx<T> *__restrict__ sv,
                                                       const PassDesc *__restrict__ pd,
                                                       const TileOp *__restrict__ ops,
                                                       const cx<double> *__restrict__ pool,
                                                       double *__restrict__ probes,
                                                       uint64_t state_stride, uint32_t pd_stride,
                                                       uint32_t tile_log, uint32_t probes_per_state) {
  __shared__ cx<T> tile[1 << BMAX];
  __shared__ uint64_t hioff[256];  // 2^(nb - c) <= 256: the planner keeps nb - c <= 8
  __shared__ cx<T> mat[1024];
  __shared__ uint32_t doff[32];
  __shared__ double red[kThreads / 64];
  const StateTile st = state_tile(tile_log);
  const ProbeSink sink{probes + st.b * probes_per_state * st.ntiles, st.t, st.ntiles, false};
  lds_tile<T, kThreads, BMAX, 1024>(sv + st.b * state_stride, pd + st.b * pd_stride, ops, pool,
                                    sink, st.t, tile, hioff, mat, doff, red);
}

}  // namespace aerhip
