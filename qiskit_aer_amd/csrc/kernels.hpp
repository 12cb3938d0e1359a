// Device kernels of the MI355X statevector engine (gfx950, wave64).
//
// Gate kernels work on tiles: a workgroup owns 2^nb amplitudes whose indices
// differ only in the tile qubits S (which always contain the lowest qubits so
// that global loads are coalesced 128-byte runs) and applies a whole list of
// gates to them in LDS.  Readout kernels accumulate in FP64 for both
// precisions and reduce in a fixed order, so results are reproducible run to
// run.
#pragma once

#include "common.hpp"

namespace aerhip {

__device__ __forceinline__ uint32_t insert_zero(uint32_t p, uint32_t t) {
  const uint32_t low = (1u << t) - 1u;
  return ((p & ~low) << 1) | (p & low);
}

__device__ __forceinline__ uint64_t insert_zero64(uint64_t p, uint32_t t) {
  const uint64_t low = (1ull << t) - 1ull;
  return ((p & ~low) << 1) | (p & low);
}

// ---------------------------------------------------------------------------
// Tile pass.
// BMAX is the largest tile (2^BMAX amplitudes, 32 KiB of LDS for both
// precisions); RMAX amplitudes per thread.
// ---------------------------------------------------------------------------
template <typename T, int BMAX>
__global__ __launch_bounds__(kThreads) void k_tile_pass(
    cx<T> *__restrict__ sv, const PassDesc *__restrict__ pd,
    const TileOp *__restrict__ ops, const cx<double> *__restrict__ pool) {
  constexpr int RMAX = (1 << BMAX) / kThreads;
  __shared__ cx<T> tile[1 << BMAX];
  __shared__ uint64_t hioff[256];
  __shared__ cx<T> mat[1024];
  __shared__ uint32_t doff[32];

  const uint32_t tid = threadIdx.x;
  const uint32_t nb = pd->nb;
  const uint32_t c = pd->c;
  const uint32_t N = 1u << nb;
  const uint32_t ncomp = pd->ncomp;

  uint64_t base = 0;
  const uint64_t tile_id = blockIdx.x;
  for (uint32_t i = 0; i < ncomp; ++i)
    base |= ((tile_id >> i) & 1ull) << pd->comp[i];

  const uint32_t nh = nb - c;
  for (uint32_t h = tid; h < (1u << nh); h += kThreads) {
    uint64_t o = 0;
    for (uint32_t j = 0; j < nh; ++j)
      o |= (uint64_t)((h >> j) & 1u) << pd->spos[c + j];
    hioff[h] = o;
  }
  __syncthreads();
  const uint32_t cmask = (1u << c) - 1u;
  for (uint32_t l = tid; l < N; l += kThreads)
    tile[l] = sv[base + (l & cmask) + hioff[l >> c]];

  const uint32_t nops = pd->nops;
  const TileOp *opl = ops + pd->op_begin;
  for (uint32_t o = 0; o < nops; ++o) {
    const TileOp &op = opl[o];
    if ((base & op.cg) != op.cvg) continue;  // uniform over the workgroup
    __syncthreads();
    const uint32_t cl = op.cl, cvl = op.cvl;
    switch (op.kind) {
      case K_DENSE1: {
        const cx<T> m0 = to_t<T>(pool[op.off + 0]);
        const cx<T> m1 = to_t<T>(pool[op.off + 1]);
        const cx<T> m2 = to_t<T>(pool[op.off + 2]);
        const cx<T> m3 = to_t<T>(pool[op.off + 3]);
        const uint32_t t = op.t;
        for (uint32_t p = tid; p < (N >> 1); p += kThreads) {
          const uint32_t l0 = insert_zero(p, t);
          const uint32_t l1 = l0 | (1u << t);
          if ((l0 & cl) != cvl) continue;
          const cx<T> a0 = tile[l0];
          const cx<T> a1 = tile[l1];
          tile[l0] = cfma(m2, a1, cmul(m0, a0));
          tile[l1] = cfma(m3, a1, cmul(m1, a0));
        }
        break;
      }
      case K_DENSE: {
        const uint32_t k = op.k;
        const uint32_t D = 1u << k;
        uint32_t tmask = 0;
        for (uint32_t b = 0; b < k; ++b) tmask |= 1u << op.pos[b];
        for (uint32_t e = tid; e < D * D; e += kThreads) mat[e] = to_t<T>(pool[op.off + e]);
        if (tid < D) {
          uint32_t d = 0;
          for (uint32_t b = 0; b < k; ++b) d |= ((tid >> b) & 1u) << op.pos[b];
          doff[tid] = d;
        }
        __syncthreads();
        cx<T> out[RMAX];
#pragma unroll
        for (int r = 0; r < RMAX; ++r) {
          const uint32_t l = tid + r * kThreads;
          if (l < N) {
            if ((l & cl) != cvl) {
              out[r] = tile[l];
            } else {
              uint32_t i = 0;
              for (uint32_t b = 0; b < k; ++b) i |= ((l >> op.pos[b]) & 1u) << b;
              const uint32_t lb = l & ~tmask;
              cx<T> acc = mk<T>(0, 0);
              for (uint32_t j = 0; j < D; ++j) acc = cfma(mat[i + D * j], tile[lb | doff[j]], acc);
              out[r] = acc;
            }
          }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RMAX; ++r) {
          const uint32_t l = tid + r * kThreads;
          if (l < N) tile[l] = out[r];
        }
        break;
      }
      case K_DIAG: {
        const uint32_t k = op.k;
        uint32_t ivb = 0;
        for (uint32_t j = 0; j < k; ++j)
          if (op.pos[j] == kNotLocal) ivb |= (uint32_t)((base >> op.gpos[j]) & 1ull) << j;
        for (uint32_t l = tid; l < N; l += kThreads) {
          if ((l & cl) != cvl) continue;
          uint32_t iv = ivb;
          for (uint32_t j = 0; j < k; ++j)
            if (op.pos[j] != kNotLocal) iv |= ((l >> op.pos[j]) & 1u) << j;
          tile[l] = cmul(tile[l], to_t<T>(pool[op.off + iv]));
        }
        break;
      }
      case K_MCX:
      case K_MCY: {
        const uint32_t t = op.t;
        const bool isy = op.kind == K_MCY;
        for (uint32_t p = tid; p < (N >> 1); p += kThreads) {
          const uint32_t l0 = insert_zero(p, t);
          const uint32_t l1 = l0 | (1u << t);
          if ((l0 & cl) != cvl) continue;
          const cx<T> a0 = tile[l0];
          const cx<T> a1 = tile[l1];
          if (isy) {
            tile[l0] = mk<T>(a1.y, -a1.x);  // -i * a1
            tile[l1] = mk<T>(-a0.y, a0.x);  //  i * a0
          } else {
            tile[l0] = a1;
            tile[l1] = a0;
          }
        }
        break;
      }
      case K_MCSWAP: {
        const uint32_t pa = op.pos[0], pb = op.pos[1];
        const uint32_t lo = pa < pb ? pa : pb, hi = pa < pb ? pb : pa;
        for (uint32_t p = tid; p < (N >> 2); p += kThreads) {
          const uint32_t l = insert_zero(insert_zero(p, lo), hi);
          if ((l & cl) != cvl) continue;
          const uint32_t la = l | (1u << pa), lb = l | (1u << pb);
          const cx<T> va = tile[la];
          tile[la] = tile[lb];
          tile[lb] = va;
        }
        break;
      }
      case K_MCPHASE: {
        const cx<T> ph = mk<T>(static_cast<T>(op.pr), static_cast<T>(op.pi));
        for (uint32_t l = tid; l < N; l += kThreads)
          if ((l & cl) == cvl) tile[l] = cmul(tile[l], ph);
        break;
      }
      case K_PAULI: {
        const cx<T> ph = mk<T>(static_cast<T>(op.pr), static_cast<T>(op.pi));
        const uint32_t sgb = __popcll(base & op.zg) & 1u;
        const uint32_t xl = op.xl, zl = op.zl;
        if (xl == 0) {
          for (uint32_t l = tid; l < N; l += kThreads) {
            cx<T> v = tile[l];
            if ((__popc(l & zl) & 1u) ^ sgb) v = cneg(v);
            tile[l] = cmul(v, ph);
          }
        } else {
          const uint32_t t = op.t;
          for (uint32_t p = tid; p < (N >> 1); p += kThreads) {
            const uint32_t l0 = insert_zero(p, t);
            const uint32_t l1 = l0 ^ xl;
            cx<T> v0 = tile[l1];
            cx<T> v1 = tile[l0];
            if ((__popc(l0 & zl) & 1u) ^ sgb) v0 = cneg(v0);
            if ((__popc(l1 & zl) & 1u) ^ sgb) v1 = cneg(v1);
            tile[l0] = cmul(v0, ph);
            tile[l1] = cmul(v1, ph);
          }
        }
        break;
      }
      case K_PERM: {
        const uint32_t k = op.k;
        const uint32_t D = 1u << k;
        uint32_t tmask = 0;
        for (uint32_t b = 0; b < k; ++b) tmask |= 1u << op.pos[b];
        if (tid < D) {
          uint32_t d = 0;
          for (uint32_t b = 0; b < k; ++b) d |= ((tid >> b) & 1u) << op.pos[b];
          doff[tid] = d;
        }
        __syncthreads();
        // Groups: tile indices with every permuted bit clear.
        for (uint32_t l = tid; l < N; l += kThreads) {
          if (l & tmask) continue;
          if ((l & cl) != cvl) continue;
          for (uint32_t q = 0; q < op.npairs; ++q) {
            const cx<double> pr = pool[op.off + q];
            const uint32_t ia = l | doff[(uint32_t)pr.x];
            const uint32_t ib = l | doff[(uint32_t)pr.y];
            const cx<T> va = tile[ia];
            tile[ia] = tile[ib];
            tile[ib] = va;
          }
        }
        break;
      }
      default:
        break;
    }
  }
  __syncthreads();
  for (uint32_t l = tid; l < N; l += kThreads)
    sv[base + (l & cmask) + hioff[l >> c]] = tile[l];
}

// ---------------------------------------------------------------------------
// Dense matrix on 6..10 qubits: one workgroup per group of 2^k amplitudes.
// The matrix (column-major) is read from global memory; consecutive threads
// read consecutive rows of one column, so the reads coalesce.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void k_dense_group(
    cx<T> *__restrict__ sv, const PassDesc *__restrict__ pd,
    const TileOp *__restrict__ ops, const cx<double> *__restrict__ pool) {
  __shared__ cx<T> g[1024];
  const TileOp &op = ops[pd->op_begin];
  const uint32_t k = op.k;
  const uint32_t D = 1u << k;
  const uint32_t tid = threadIdx.x;
  // Group base: insert zeros at the ascending target positions
  // (index0, indexes.hpp:183-193).
  uint64_t base = blockIdx.x;
  for (uint32_t j = 0; j < pd->nb; ++j) base = insert_zero64(base, pd->spos[j]);
  for (uint32_t j = tid; j < D; j += kThreads) {
    uint64_t d = 0;
    for (uint32_t b = 0; b < k; ++b) d |= (uint64_t)((j >> b) & 1u) << op.gpos[b];
    g[j] = sv[base + d];
  }
  __syncthreads();
  cx<T> out[4];
#pragma unroll
  for (uint32_t r = 0; r < 4; ++r) {
    const uint32_t i = tid + r * kThreads;
    if (i < D) {
      cx<T> acc = mk<T>(0, 0);
      for (uint32_t j = 0; j < D; ++j)
        acc = cfma(to_t<T>(pool[op.off + i + (uint64_t)D * j]), g[j], acc);
      out[r] = acc;
    }
  }
#pragma unroll
  for (uint32_t r = 0; r < 4; ++r) {
    const uint32_t i = tid + r * kThreads;
    if (i < D) {
      uint64_t d = 0;
      for (uint32_t b = 0; b < k; ++b) d |= (uint64_t)((i >> b) & 1u) << op.gpos[b];
      sv[base + d] = out[r];
    }
  }
}

// ---------------------------------------------------------------------------
// Initialisation and conversion.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_fill_basis(cx<T> *__restrict__ sv, uint64_t n, uint64_t one_at) {
  for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kThreads)
    sv[i] = mk<T>(i == one_at ? T(1) : T(0), T(0));
}

template <typename TO, typename TI>
__global__ void k_convert(cx<TO> *__restrict__ out, const cx<TI> *__restrict__ in, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kThreads)
    out[i] = mk<TO>(static_cast<TO>(in[i].x), static_cast<TO>(in[i].y));
}

// ---------------------------------------------------------------------------
// FP64 reductions.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[kThreads / 64];
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kThreads / 64; ++i) s += red[i];
  __syncthreads();
  return s;  // valid in thread 0
}

template <typename T>
__device__ __forceinline__ double prob(const cx<T> &a) {
  const double x = a.x, y = a.y;
  return x * x + y * y;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_sumsq(const cx<T> *__restrict__ sv, uint64_t n,
                                                   double *__restrict__ partial) {
  double acc = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kThreads)
    acc += prob(sv[i]);
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// Sums `cols` consecutive partials of each of `rows` rows (one thread per row,
// sequential order).
__global__ void k_sum_rows(const double *__restrict__ partial, uint64_t rows, uint32_t cols,
                           double *__restrict__ out) {
  const uint64_t r = blockIdx.x * (uint64_t)kThreads + threadIdx.x;
  if (r >= rows) return;
  double s = 0;
  for (uint32_t c = 0; c < cols; ++c) s += partial[r * cols + c];
  out[r] = s;
}

template <typename T>
__global__ void k_probs_all(const cx<T> *__restrict__ sv, uint64_t n, double *__restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kThreads)
    out[i] = prob(sv[i]);
}

struct MarginalDesc {
  uint32_t nq;         // measured qubits
  uint32_t nchunk;     // partial sums per outcome
  uint64_t inner;      // 2^(n - nq) indices per outcome
  uint8_t sorted[64];  // measured qubits ascending
  uint8_t order[64];   // measured qubits in outcome-bit order
};

// Block (c, m) sums |psi|^2 over inner indices k in chunk c for outcome m.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_marginal_chunks(const cx<T> *__restrict__ sv,
                                                              const MarginalDesc *__restrict__ md,
                                                              double *__restrict__ partial) {
  const uint64_t m = blockIdx.y + (uint64_t)gridDim.y * blockIdx.z;
  const uint32_t nq = md->nq;
  uint64_t mbits = 0;
  for (uint32_t i = 0; i < nq; ++i) mbits |= ((m >> i) & 1ull) << md->order[i];
  const uint64_t inner = md->inner;
  const uint64_t chunk = (inner + md->nchunk - 1) / md->nchunk;
  const uint64_t k0 = blockIdx.x * chunk;
  const uint64_t k1 = min(inner, k0 + chunk);
  double acc = 0;
  for (uint64_t k = k0 + threadIdx.x; k < k1; k += kThreads) {
    uint64_t idx = k;
    for (uint32_t j = 0; j < nq; ++j) idx = insert_zero64(idx, md->sorted[j]);
    acc += prob(sv[idx | mbits]);
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[m * md->nchunk + blockIdx.x] = acc;
}

// One thread per outcome for marginals with few inner indices.
template <typename T>
__global__ void k_marginal_small(const cx<T> *__restrict__ sv, const MarginalDesc *__restrict__ md,
                                 uint64_t outcomes, double *__restrict__ out) {
  const uint64_t m = blockIdx.x * (uint64_t)kThreads + threadIdx.x;
  if (m >= outcomes) return;
  const uint32_t nq = md->nq;
  uint64_t mbits = 0;
  for (uint32_t i = 0; i < nq; ++i) mbits |= ((m >> i) & 1ull) << md->order[i];
  double acc = 0;
  for (uint64_t k = 0; k < md->inner; ++k) {
    uint64_t idx = k;
    for (uint32_t j = 0; j < nq; ++j) idx = insert_zero64(idx, md->sorted[j]);
    acc += prob(sv[idx | mbits]);
  }
  out[m] = acc;
}

// Sum of |psi|^2 over consecutive runs of 2^sbits amplitudes, one wave per run.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_run_sums(const cx<T> *__restrict__ sv,
                                                       uint32_t sbits, uint64_t nruns,
                                                       double *__restrict__ out) {
  const uint64_t run = blockIdx.x * (uint64_t)(kThreads / 64) + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  if (run >= nruns) return;
  const uint64_t len = 1ull << sbits;
  const uint64_t b = run << sbits;
  double acc = 0;
  for (uint64_t j = lane; j < len; j += 64) acc += prob(sv[b + j]);
  acc = wave_sum(acc);
  if (lane == 0) out[run] = acc;
}

// Sampling by a three-level cumulative search in the order of the reference's
// QubitVector::sample_measure (qubitvector.hpp:2149-2228): index blocks of
// 2^(n-isz) amplitudes (the reference's idxs), sub-runs of 2^sbits
// amplitudes inside a block, then amplitudes; comparisons are rnd < cumsum.
template <typename T>
__global__ void k_sample(const cx<T> *__restrict__ sv, uint64_t size, uint32_t isz,
                         uint32_t sbits, const double *__restrict__ blocks,
                         const double *__restrict__ runs, const double *__restrict__ rnds,
                         uint64_t shots, uint64_t *__restrict__ out) {
  const uint64_t s = blockIdx.x * (uint64_t)kThreads + threadIdx.x;
  if (s >= shots) return;
  const double rnd = rnds[s];
  const uint64_t nblocks = 1ull << isz;
  const uint64_t loop = size >> isz;
  double p = 0;
  uint64_t sample = 0;
  uint64_t j = 0;
  for (; j < nblocks; ++j) {
    if (rnd < p + blocks[j]) break;
    p += blocks[j];
    sample += loop;
  }
  if (j == nblocks) {  // rnd beyond the accumulated total: last amplitude
    out[s] = size - 1;
    return;
  }
  const uint64_t rpb = loop >> sbits;  // runs per block
  const uint64_t r0 = sample >> sbits;
  for (uint64_t r = 0; r + 1 < rpb; ++r) {
    const double v = runs[r0 + r];
    if (rnd < p + v) break;
    p += v;
    sample += 1ull << sbits;
  }
  for (; sample < size - 1; ++sample) {
    p += prob(sv[sample]);
    if (rnd < p) break;
  }
  out[s] = sample;
}

// Linear scan for small states (END < 2^isz in the reference).
template <typename T>
__global__ void k_sample_linear(const cx<T> *__restrict__ sv, uint64_t size,
                                const double *__restrict__ rnds, uint64_t shots,
                                uint64_t *__restrict__ out) {
  const uint64_t s = blockIdx.x * (uint64_t)kThreads + threadIdx.x;
  if (s >= shots) return;
  const double rnd = rnds[s];
  double p = 0;
  uint64_t sample = 0;
  for (; sample < size - 1; ++sample) {
    p += prob(sv[sample]);
    if (rnd < p) break;
  }
  out[s] = sample;
}

// <P> for one Pauli string (qubitvector.hpp:2300-2347), FP64 accumulation.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_expval_pauli(
    const cx<T> *__restrict__ sv, uint64_t size, uint64_t xmask, uint64_t zmask,
    uint32_t xmax, double pr, double pi, double *__restrict__ partial) {
  double acc = 0;
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  if (xmask == 0) {
    for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < size; i += stride) {
      const double x = sv[i].x, y = sv[i].y;
      const double ax = pr * x - pi * y, ay = pr * y + pi * x;  // phase * psi
      double v = ax * x + ay * y;                               // Re(. * conj psi)
      if (__popcll(i & zmask) & 1) v = -v;
      acc += v;
    }
  } else {
    const uint64_t half = size >> 1;
    for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < half; i += stride) {
      const uint64_t i0 = insert_zero64(i, xmax);
      const uint64_t i1 = i0 ^ xmask;
      const double x0 = sv[i0].x, y0 = sv[i0].y, x1 = sv[i1].x, y1 = sv[i1].y;
      // Re(phase * psi1 * conj psi0) and Re(phase * psi0 * conj psi1)
      const double a1x = pr * x1 - pi * y1, a1y = pr * y1 + pi * x1;
      const double a0x = pr * x0 - pi * y0, a0y = pr * y0 + pi * x0;
      double v0 = a1x * x0 + a1y * y0;
      double v1 = a0x * x1 + a0y * y1;
      if (__popcll(i0 & zmask) & 1) v0 = -v0;
      if (__popcll(i1 & zmask) & 1) v1 = -v1;
      acc += v0 + v1;
    }
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

}  // namespace aerhip
