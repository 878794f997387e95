// mcv_kernel.h -- skeleton of a fused instrument kernel.
//
// The reference's generated kernel (acc/run_script.py:320-334) gives each
// thread a contiguous block of neutron indices [tid*npt, min((tid+1)*npt, N))
// drawn from its own xoroshiro128+ stream, with one global atomic per neutron
// on a single counter.  Here one lane owns one stream: its state is loaded
// once into registers, the per-neutron chain body (emitted by the fuser) runs
// in a loop, and the processed count is reduced across the wavefront before a
// single atomic per wave.
#pragma once
#include "mcv_core.h"
#include "mcv_components.h"
#include "mcv_samples.h"
#include "mcv_launch.h"

namespace mcv {

__device__ __forceinline__ void dump_neutron(const LaunchArgs& A, int64_t idx, const Neutron& n, int alive) {
  const int64_t k = idx - A.dump_first;
  if (A.dump != nullptr && k >= 0 && k < A.dump_count) {
    double* d = A.dump + 10 * k;
    d[0] = n.x; d[1] = n.y; d[2] = n.z; d[3] = n.vx; d[4] = n.vy; d[5] = n.vz;
    d[6] = n.s1; d[7] = n.s2; d[8] = n.t; d[9] = n.p;
    A.dump_alive[k] = alive;
  }
}

__device__ __forceinline__ void count_processed(const LaunchArgs& A, int64_t mine) {
  // wavefront reduction, one atomic per wave
  unsigned long long v = (unsigned long long)mine;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(A.counter, v);
}

}  // namespace mcv

// Per-stream loop around a chain body `int BODY(Rng&, Neutron&, const LaunchArgs&)`
// returning the dump flag: 1/0 whether the neutron reached the end of the
// chain alive, or, for multiple-scattering chains, the number of outputs.
#define MCV_DEFINE_FUSED_KERNEL(NAME, BODY)                                                   \
  extern "C" __global__ void __launch_bounds__(256) NAME(const mcv::LaunchArgs A) {           \
    const int64_t tid = A.tid_base + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;          \
    int64_t mine = 0;                                                                         \
    if (tid < A.nthreads) {                                                                   \
      mcv::Rng rng;                                                                           \
      rng.s0 = A.rng[2 * tid];                                                                \
      rng.s1 = A.rng[2 * tid + 1];                                                            \
      const int64_t start = tid * A.npt;                                                      \
      const int64_t end = start + A.npt < A.N ? start + A.npt : A.N;                          \
      mcv::Neutron n;                                                                         \
      n.s1 = 0;                                                                               \
      n.s2 = 0;                                                                               \
      for (int64_t idx = start; idx < end; ++idx) {                                           \
        const int alive = (int)BODY(rng, n, A);                                               \
        mcv::dump_neutron(A, idx, n, alive);                                                  \
      }                                                                                       \
      mine = end > start ? end - start : 0;                                                   \
    }                                                                                         \
    mcv::count_processed(A, mine);                                                            \
  }
