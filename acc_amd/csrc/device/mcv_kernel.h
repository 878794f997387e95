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

// ------------------------------------------------------------------ compacted chains
// A chain whose random draws all happen in its first components (phase A) and
// whose remaining components (phase B) are deterministic is run in two
// phases per workgroup.  Every lane produces neutrons from its own stream in
// stream order through phase A; survivors are appended to an LDS queue (slots
// from a wavefront ballot + prefix count, one LDS atomic per wave); whenever
// the queue holds a full workgroup of neutrons, every lane pops one and runs
// phase B.  Phase B draws nothing, so running it later keeps each stream's
// draw sequence identical to the reference's neutron-by-neutron loop, and its
// expensive part (e.g. the segmented guide) runs with all lanes busy instead
// of one lane in four.
#define MCV_QCAP 512   // >= (block size - 1) + block size
#define MCV_QBLOCK 256

namespace mcv {

struct QueueLDS {
  double x[MCV_QCAP], y[MCV_QCAP], z[MCV_QCAP], vx[MCV_QCAP], vy[MCV_QCAP], vz[MCV_QCAP], t[MCV_QCAP],
      p[MCV_QCAP];
  int64_t idx[MCV_QCAP];
  int count;
  int active;
};

__device__ __forceinline__ void q_store(QueueLDS& q, int s, const Neutron& n, int64_t idx) {
  q.x[s] = n.x; q.y[s] = n.y; q.z[s] = n.z; q.vx[s] = n.vx; q.vy[s] = n.vy; q.vz[s] = n.vz;
  q.t[s] = n.t; q.p[s] = n.p; q.idx[s] = idx;
}

__device__ __forceinline__ int64_t q_load(const QueueLDS& q, int s, Neutron& n) {
  n.x = q.x[s]; n.y = q.y[s]; n.z = q.z[s]; n.vx = q.vx[s]; n.vy = q.vy[s]; n.vz = q.vz[s];
  n.t = q.t[s]; n.p = q.p[s]; n.s1 = 0; n.s2 = 0;
  return q.idx[s];
}

}  // namespace mcv

#define MCV_DEFINE_COMPACTED_KERNEL(NAME, PHASE_A, PHASE_B)                                    \
  extern "C" __global__ void __launch_bounds__(MCV_QBLOCK) NAME(const mcv::LaunchArgs A) {    \
    __shared__ mcv::QueueLDS q;                                                                \
    const int lt = threadIdx.x;                                                                \
    const int64_t tid = A.tid_base + (int64_t)blockIdx.x * blockDim.x + lt;                    \
    if (lt == 0) q.count = 0;                                                                  \
    mcv::Rng rng;                                                                              \
    int64_t idx = 0, end = 0;                                                                  \
    if (tid < A.nthreads) {                                                                    \
      rng.s0 = A.rng[2 * tid];                                                                 \
      rng.s1 = A.rng[2 * tid + 1];                                                             \
      idx = tid * A.npt;                                                                       \
      end = idx + A.npt < A.N ? idx + A.npt : A.N;                                             \
      if (end < idx) end = idx;                                                                \
    }                                                                                          \
    const int64_t mine = end - idx;                                                            \
    const unsigned lane = threadIdx.x & 63u;                                                   \
    __syncthreads();                                                                           \
    while (true) {                                                                             \
      /* phase A: at most one neutron per lane per round */                                    \
      bool has = false;                                                                        \
      mcv::Neutron n;                                                                          \
      n.s1 = 0;                                                                                \
      n.s2 = 0;                                                                                \
      const int64_t my = idx;                                                                  \
      if (idx < end) {                                                                         \
        has = PHASE_A(rng, n, A) != 0;                                                         \
        if (!has) mcv::dump_neutron(A, my, n, 0);                                              \
        ++idx;                                                                                 \
      }                                                                                        \
      const unsigned long long m = __ballot(has);                                              \
      int base = 0;                                                                            \
      if (lane == 0 && m) base = atomicAdd(&q.count, (int)__popcll(m));                        \
      base = __shfl(base, 0, 64);                                                              \
      if (has) mcv::q_store(q, base + (int)__popcll(m & ((1ull << lane) - 1ull)), n, my);      \
      const int more = __syncthreads_or(idx < end);                                            \
      const int cnt = q.count;                                                                 \
      if (cnt >= MCV_QBLOCK || (!more && cnt > 0)) {                                           \
        const int take = cnt < MCV_QBLOCK ? cnt : MCV_QBLOCK;                                  \
        if (lt < take) {                                                                       \
          mcv::Neutron b;                                                                      \
          const int64_t bidx = mcv::q_load(q, cnt - take + lt, b);                             \
          const int alive = (int)PHASE_B(b, A);                                                \
          mcv::dump_neutron(A, bidx, b, alive);                                                \
        }                                                                                      \
        __syncthreads();                                                                       \
        if (lt == 0) q.count = cnt - take;                                                     \
        __syncthreads();                                                                       \
        if (!more && cnt - take == 0) break;                                                   \
      } else if (!more) {                                                                      \
        break;                                                                                 \
      }                                                                                        \
    }                                                                                          \
    mcv::count_processed(A, mine);                                                             \
  }

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
