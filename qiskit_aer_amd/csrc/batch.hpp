// Batched circuits in one launch (gfx950).
//
// Many independent statevectors of up to ~20 qubits (different sizes) run in
// a single kernel launch: one workgroup per circuit walks every pass and every
// tile of its own circuit (the same tile programs as the one-state kernels),
// so no workgroup ever waits for another and the only synchronisation is the
// workgroup barrier between passes.  Saved expectation values accumulate per
// circuit in tile order.  The host orders circuits by decreasing cost so that
// the largest ones start first.
#pragma once

#include "reg_pass.hpp"

namespace aerhip {

struct BatchCircuit {
  uint64_t sv_off;      // first amplitude of the circuit's state in the batch buffer
  uint32_t num_qubits;
  uint32_t pass_begin;  // passes[pass_begin .. + pass_count)
  uint32_t pass_count;
  uint32_t probe_base;  // saved values go to probes[probe_base .. + num_probes)
  uint32_t num_probes;
  uint32_t pad;
};

template <typename T, int B, int R>
__global__ __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(reg_pass_waves<T, B, R>()))) void k_batch(cx<T> *__restrict__ states,
                                                    const BatchCircuit *__restrict__ circ,
                                                    const PassDesc *__restrict__ passes,
                                                    const ROp<T> *__restrict__ rops,
                                                    const TileOp *__restrict__ tops,
                                                    const cx<double> *__restrict__ pool,
                                                    double *__restrict__ probes) {
  static_assert((1 << (B - R)) == kThreads, "register passes and LDS passes share the workgroup");
  __shared__ cx<T> lds[1 << B];
  // the LDS-pass offset table and the register passes' staged tables never
  // live at the same time: one array holds either
  __shared__ __align__(16) unsigned char aux_lds[kDiagLdsBytes > 256 * sizeof(uint64_t) ? kDiagLdsBytes
                                                                                         : 256 * sizeof(uint64_t)];
  uint64_t *hioff = reinterpret_cast<uint64_t *>(aux_lds);
  cx<double> *dlt = reinterpret_cast<cx<double> *>(aux_lds);
  __shared__ cx<T> mat[16];
  __shared__ uint32_t doff[32];
  __shared__ double red[kThreads / 64];
  const BatchCircuit c = circ[blockIdx.x];
  cx<T> *sv = states + c.sv_off;
  const uint64_t size = 1ull << c.num_qubits;
  for (uint64_t i = threadIdx.x; i < size; i += kThreads) sv[i] = mk<T>(i == 0 ? T(1) : T(0), T(0));
  for (uint32_t i = threadIdx.x; i < c.num_probes; i += kThreads) probes[c.probe_base + i] = 0.0;
  __syncthreads();
  const ProbeSink sink{probes + c.probe_base, 0, 1, true};
  for (uint32_t p = 0; p < c.pass_count; ++p) {
    const PassDesc *pd = passes + c.pass_begin + p;
    const uint64_t ntiles = 1ull << pd->ncomp;
    if (pd->kind == 2) {
      for (uint64_t t = 0; t < ntiles; ++t)
        reg_tile<T, B, R, 16, true>(sv, pd, rops, tops, pool, sink, t, lds, mat, doff, red, dlt);
    } else if (pd->kind == 4) {
      for (uint64_t t = 0; t < ntiles; ++t) qperm_tile<T, B - 1, kThreads>(sv, pd, tops + pd->op_begin, t, lds);
    } else if (pd->kind == 3) {
      const TileOp &op = tops[pd->op_begin];
      for (uint64_t t = 0; t < ntiles; ++t) {
        const double v = probe_sweep_block<T, kThreads>(sv, size, op, t, ntiles, red);
        if (threadIdx.x == 0) sink.put(op.npairs, v);
      }
    } else {
      for (uint64_t t = 0; t < ntiles; ++t)
        lds_tile<T, kThreads, B, 16>(sv, pd, tops, pool, sink, t, lds, hioff, mat, doff, red);
    }
    __syncthreads();  // the pass's HBM writes are visible to the whole workgroup
  }
}

}  // namespace aerhip
