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
// Dense matrix on 6..10 qubits: a workgroup takes groups of 2^k amplitudes
// (group g and every gridDim.x-th after it, 64-bit group index, so the grid
// stays within the launch limits at any state size).  Controls on qubits
// outside the group (op.cg / op.cvg) select the groups the matrix applies to
// (the multiplexer pushes one controlled matrix per control value).  The
// matrix (column-major) is read from global memory; consecutive threads read
// consecutive rows of one column, so the reads coalesce.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void k_dense_group(
    cx<T> *__restrict__ sv, const PassDesc *__restrict__ pd,
    const TileOp *__restrict__ ops, const cx<double> *__restrict__ pool, uint64_t ngroups) {
  __shared__ cx<T> g[1024];
  const TileOp &op = ops[pd->op_begin];
  const uint32_t k = op.k;
  const uint32_t D = 1u << k;
  const uint32_t tid = threadIdx.x;
  const uint64_t cg = op.cg, cvg = op.cvg;
  for (uint64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    // Group base: insert zeros at the ascending target positions
    // (index0, indexes.hpp:183-193).
    uint64_t base = grp;
    for (uint32_t j = 0; j < pd->nb; ++j) base = insert_zero64(base, pd->spos[j]);
    if ((base & cg) != cvg) continue;  // uniform over the workgroup
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
    __syncthreads();  // g is refilled for the next group
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

// splitmix64 finaliser: a counter-based 64-bit hash.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// psi_i = (g1 + i g2)/sqrt(2) with g1, g2 standard normal (Box-Muller on two
// 53-bit uniforms from the hash of (seed, i)); normalised afterwards.
template <typename T>
__global__ void k_gaussian(cx<T> *__restrict__ sv, uint64_t n, uint64_t seed) {
  const uint64_t key = mix64(seed + 0x9e3779b97f4a7c15ull);
  for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kThreads) {
    const uint64_t a = mix64(key ^ (2 * i)), b = mix64(key ^ (2 * i + 1));
    const double u1 = ((a >> 11) + 1) * (1.0 / 9007199254740993.0);  // (0, 1]
    const double u2 = (b >> 11) * (1.0 / 9007199254740992.0);        // [0, 1)
    const double r = sqrt(-2.0 * log(u1)) * 0.70710678118654752;
    double s, c;
    sincospi(2.0 * u2, &s, &c);
    sv[i] = mk<T>(static_cast<T>(r * c), static_cast<T>(r * s));
  }
}

template <typename T>
__global__ void k_scale(cx<T> *__restrict__ sv, uint64_t n, const double *__restrict__ sumsq) {
  const double f = 1.0 / sqrt(*sumsq);
  for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kThreads)
    sv[i] = mk<T>(static_cast<T>(sv[i].x * f), static_cast<T>(sv[i].y * f));
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

// <a|b> = sum_i conj(a_i) b_i in FP64: real parts to partial[block], imaginary
// parts to partial[gridDim.x + block].
template <typename TA, typename TB>
__global__ __launch_bounds__(kThreads) void k_inner(const cx<TA> *__restrict__ a,
                                                  const cx<TB> *__restrict__ b, uint64_t n,
                                                  double *__restrict__ partial) {
  double re = 0, im = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kThreads) {
    const double ax = a[i].x, ay = a[i].y, bx = b[i].x, by = b[i].y;
    re += ax * bx + ay * by;
    im += ax * by - ay * bx;
  }
  re = block_sum(re);
  im = block_sum(im);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = re;
    partial[gridDim.x + blockIdx.x] = im;
  }
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

// QubitVector::initialize_component (qubitvector.hpp:879-896) on k > 5
// qubits, in two sweeps so that no group reads an amplitude another thread has
// already overwritten: first cache[g] = psi[index0(g)] for every group g (the
// amplitude whose measured bits are all zero), then psi[idx] = cache[g(idx)] *
// state[i(idx)] with i(idx) the bits of idx on md->order (bit j = qubits[j]).
template <typename T>
__global__ void k_gather_group0(const cx<T> *__restrict__ sv, const MarginalDesc *__restrict__ md,
                                cx<T> *__restrict__ cache) {
  const uint32_t nq = md->nq;
  for (uint64_t g = blockIdx.x * (uint64_t)kThreads + threadIdx.x; g < md->inner;
       g += (uint64_t)gridDim.x * kThreads) {
    uint64_t idx = g;
    for (uint32_t j = 0; j < nq; ++j) idx = insert_zero64(idx, md->sorted[j]);
    cache[g] = sv[idx];
  }
}

template <typename T>
__global__ void k_init_component(cx<T> *__restrict__ sv, const MarginalDesc *__restrict__ md,
                                 uint64_t size, const cx<T> *__restrict__ cache,
                                 const cx<double> *__restrict__ state) {
  const uint32_t nq = md->nq;
  for (uint64_t idx = blockIdx.x * (uint64_t)kThreads + threadIdx.x; idx < size;
       idx += (uint64_t)gridDim.x * kThreads) {
    uint64_t i = 0;
    for (uint32_t j = 0; j < nq; ++j) i |= ((idx >> md->order[j]) & 1ull) << j;
    // group index: idx with the measured bits removed (highest first)
    uint64_t g = idx;
    for (int j = static_cast<int>(nq) - 1; j >= 0; --j) {
      const uint32_t q = md->sorted[j];
      g = ((g >> (q + 1)) << q) | (g & ((1ull << q) - 1ull));
    }
    sv[idx] = cmul(cache[g], to_t<T>(state[i]));
  }
}

// QubitVector::norm(qubits, mat) (qubitvector.hpp:1889-2054): sum over the
// groups of |M psi_group|^2, M (D x D, column-major, D = 2^k <= 1024) from the
// pool, products in the state precision, accumulated in FP64.  Workgroup w
// takes groups w, w + gridDim.x, ... and writes its partial to partial[w].
template <typename T>
__global__ __launch_bounds__(kThreads) void k_norm_group(const cx<T> *__restrict__ sv,
                                                         const MarginalDesc *__restrict__ md,
                                                         const cx<double> *__restrict__ mat,
                                                         double *__restrict__ partial) {
  __shared__ cx<T> g[1024];
  __shared__ double red[kThreads / 64];
  const uint32_t k = md->nq;
  const uint32_t D = 1u << k;
  const uint32_t tid = threadIdx.x;
  double acc = 0;
  for (uint64_t grp = blockIdx.x; grp < md->inner; grp += gridDim.x) {
    uint64_t base = grp;
    for (uint32_t j = 0; j < k; ++j) base = insert_zero64(base, md->sorted[j]);
    for (uint32_t j = tid; j < D; j += kThreads) {
      uint64_t d = 0;
      for (uint32_t b = 0; b < k; ++b) d |= (uint64_t)((j >> b) & 1u) << md->order[b];
      g[j] = sv[base + d];
    }
    __syncthreads();
    for (uint32_t i = tid; i < D; i += kThreads) {
      cx<T> v = mk<T>(0, 0);
      for (uint32_t j = 0; j < D; ++j) v = cfma(to_t<T>(mat[i + (uint64_t)D * j]), g[j], v);
      acc += static_cast<double>(v.x) * v.x + static_cast<double>(v.y) * v.y;
    }
    __syncthreads();
  }
  acc = wave_sum(acc);
  if ((tid & 63) == 0) red[tid >> 6] = acc;
  __syncthreads();
  if (tid == 0) {
    double s = 0;
    for (int w = 0; w < kThreads / 64; ++w) s += red[w];
    partial[blockIdx.x] = s;
  }
}

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

struct PauliTerm {
  uint64_t x, z;  // flip and sign masks
  double cr, ci;  // coefficient times the Pauli phase (-i)^#Y
};

// out[i] = sum_t c_t (-1)^popcount(i & z_t) in[i ^ x_t] for i < half, where
// in = sv[0, half) and out = sv[half, 2 half): the action of a Pauli sum,
// (P psi)_i = phase (-1)^popcount(i & z) psi_{i ^ x}, as in
// QubitVector::apply_pauli (qubitvector.hpp:2393-2433).  Terms arrive sorted
// by x so each distinct flip is loaded once.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_pauli_sum_upper(cx<T> *__restrict__ sv,
                                                              uint64_t half,
                                                              const PauliTerm *__restrict__ terms,
                                                              uint32_t nterms) {
  for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < half;
       i += (uint64_t)gridDim.x * kThreads) {
    double ar = 0, ai = 0;
    uint64_t xprev = ~0ull;
    double vr = 0, vi = 0;
    for (uint32_t t = 0; t < nterms; ++t) {
      const PauliTerm pt = terms[t];
      if (pt.x != xprev) {
        const cx<T> v = sv[i ^ pt.x];
        vr = v.x;
        vi = v.y;
        xprev = pt.x;
      }
      const double s = (__popcll(i & pt.z) & 1) ? -1.0 : 1.0;
      ar += s * (pt.cr * vr - pt.ci * vi);
      ai += s * (pt.cr * vi + pt.ci * vr);
    }
    sv[half + i] = mk<T>(static_cast<T>(ar), static_cast<T>(ai));
  }
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
