// mcv_kernel.h -- skeleton of the fused instrument kernels.
//
// The reference's generated kernel (acc/run_script.py:320-334) gives each
// thread a contiguous block of neutron indices [tid*npt, min((tid+1)*npt, N))
// drawn from its own xoroshiro128+ stream, with one global atomic per neutron
// on a single counter.  Here one lane owns one stream: its state is loaded
// once into registers, the chain (emitted by the fuser as a struct with static
// device functions) runs per neutron, and the processed count is reduced
// across the wavefront before one atomic per wave.
//
// Every code object carries two kernels for its chain:
//   NAME        one instrument, one launch over its streams;
//   NAME_batch  a persistent kernel for a scan of many instances of the same
//               chain (own parameters, histograms and neutron counts each):
//               workgroups pull (instance, block of 256 streams) items from
//               one atomic counter until all are done, so uneven instances
//               share the GPU instead of queueing one launch after another.
#pragma once
#include "mcv_core.h"
#include "mcv_components.h"
#include "mcv_samples.h"
#include "mcv_launch.h"

#define MCV_BLOCK 256
#define MCV_QCAP 512  // >= (MCV_BLOCK - 1) + MCV_BLOCK

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

__device__ __forceinline__ void count_processed(unsigned long long* counter, int64_t mine) {
  // wavefront reduction, one atomic per wave
  unsigned long long v = (unsigned long long)mine;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(counter, v);
}

// Streams of one lane: [tid*npt, min((tid+1)*npt, N)), or empty.
__device__ __forceinline__ void stream_range(const LaunchArgs& A, int64_t tid, Rng& rng, int64_t& idx,
                                             int64_t& end) {
  idx = end = 0;
  if (tid < A.nthreads) {
    rng.s0 = A.rng[2 * tid];
    rng.s1 = A.rng[2 * tid + 1];
    idx = tid * A.npt;
    end = idx + A.npt < A.N ? idx + A.npt : A.N;
    if (end < idx) end = idx;
  }
}

// ------------------------------------------------------------------ plain chains
// Chain::body(Rng&, Neutron&, const LaunchArgs&) returns the dump flag: 1/0
// whether the neutron reached the end alive, or, for multiple-scattering
// chains, the number of outputs.
template <class Chain>
__device__ __forceinline__ void run_lane(const LaunchArgs& A, int64_t tid) {
  Rng rng;
  int64_t idx, end;
  stream_range(A, tid, rng, idx, end);
  Neutron n;
  n.s1 = 0;
  n.s2 = 0;
  for (int64_t i = idx; i < end; ++i) {
    const int alive = (int)Chain::body(rng, n, A);
    dump_neutron(A, i, n, alive);
  }
  count_processed(A.counter, end - idx);
}

// ------------------------------------------------------------------ compacted chains
// A chain whose random draws all happen in its first components (phase A) and
// whose remaining components (phase B) are deterministic runs in two phases per
// workgroup.  Every lane produces neutrons from its own stream, in stream
// order, through phase A; survivors are appended to an LDS queue (slots from a
// wavefront ballot + prefix count, one LDS atomic per wave); whenever the queue
// holds a full workgroup of neutrons, every lane pops one and runs phase B.
// Phase B draws nothing, so running it later keeps each stream's draw sequence
// identical to the reference's neutron-by-neutron loop, and its expensive part
// (e.g. the segmented guide) runs with all lanes busy instead of one in four.
struct QueueLDS {
  double x[MCV_QCAP], y[MCV_QCAP], z[MCV_QCAP], vx[MCV_QCAP], vy[MCV_QCAP], vz[MCV_QCAP], t[MCV_QCAP],
      p[MCV_QCAP];
  int64_t idx[MCV_QCAP];
  int count;
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

// All MCV_BLOCK threads of the workgroup must call this together (barriers).
template <class Chain>
__device__ __forceinline__ void run_block_compacted(const LaunchArgs& A, int64_t tid, QueueLDS& q) {
  const int lt = threadIdx.x;
  if (lt == 0) q.count = 0;
  Rng rng;
  int64_t idx, end;
  stream_range(A, tid, rng, idx, end);
  const int64_t mine = end - idx;
  const unsigned lane = threadIdx.x & 63u;
  __syncthreads();
  while (true) {
    // phase A: at most one neutron per lane per round
    bool has = false;
    Neutron n;
    n.s1 = 0;
    n.s2 = 0;
    const int64_t my = idx;
    if (idx < end) {
      has = Chain::phase_a(rng, n, A) != 0;
      if (!has) dump_neutron(A, my, n, 0);
      ++idx;
    }
    const unsigned long long m = __ballot(has);
    int base = 0;
    if (lane == 0 && m) base = atomicAdd(&q.count, (int)__popcll(m));
    base = __shfl(base, 0, 64);
    if (has) q_store(q, base + (int)__popcll(m & ((1ull << lane) - 1ull)), n, my);
    const int more = __syncthreads_or(idx < end);
    const int cnt = q.count;
    if (cnt >= MCV_BLOCK || (!more && cnt > 0)) {
      const int take = cnt < MCV_BLOCK ? cnt : MCV_BLOCK;
      if (lt < take) {
        Neutron b;
        const int64_t bidx = q_load(q, cnt - take + lt, b);
        const int alive = (int)Chain::phase_b(b, A);
        dump_neutron(A, bidx, b, alive);
      }
      __syncthreads();
      if (lt == 0) q.count = cnt - take;
      __syncthreads();
      if (!more && cnt - take == 0) break;
    } else if (!more) {
      break;
    }
  }
  count_processed(A.counter, mine);
  __syncthreads();  // the queue is reused by the next block of streams
}

// ------------------------------------------------------------------ batched scans
// Next (instance, stream block) for the workgroup; false when all are taken.
__device__ __forceinline__ bool next_item(const BatchArgs& B, LaunchArgs& A, int64_t& tid) {
  __shared__ int64_t s_item;
  if (threadIdx.x == 0) s_item = (int64_t)atomicAdd(B.work, 1ull);
  __syncthreads();
  const int64_t item = s_item;
  __syncthreads();
  if (item >= B.item0[B.ninst]) return false;
  int lo = 0, hi = B.ninst - 1;  // largest instance with item0 <= item
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (B.item0[mid] <= item) lo = mid;
    else hi = mid - 1;
  }
  A = B.base;
  A.P = B.base.P + lo * B.pstride;
  A.H = B.base.H + lo * B.hstride;
  A.N = B.inst_n[lo];
  A.nthreads = B.inst_nthreads[lo];
  A.npt = B.inst_npt[lo];
  A.counter = B.inst_processed + lo;
  A.dump = nullptr;
  A.dump_count = 0;
  tid = (item - B.item0[lo]) * MCV_BLOCK + threadIdx.x;
  return true;
}

}  // namespace mcv

#define MCV_DEFINE_KERNELS_PLAIN(NAME, CHAIN)                                                  \
  extern "C" __global__ void __launch_bounds__(MCV_BLOCK) NAME(const mcv::LaunchArgs A) {     \
    const int64_t tid = A.tid_base + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;          \
    mcv::run_lane<CHAIN>(A, tid);                                                             \
  }                                                                                           \
  extern "C" __global__ void __launch_bounds__(MCV_BLOCK) NAME##_batch(const mcv::BatchArgs B) { \
    mcv::LaunchArgs A;                                                                        \
    int64_t tid;                                                                              \
    while (mcv::next_item(B, A, tid)) mcv::run_lane<CHAIN>(A, tid);                           \
  }

#define MCV_DEFINE_KERNELS_COMPACTED(NAME, CHAIN)                                              \
  extern "C" __global__ void __launch_bounds__(MCV_BLOCK) NAME(const mcv::LaunchArgs A) {     \
    __shared__ mcv::QueueLDS q;                                                               \
    const int64_t tid = A.tid_base + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;          \
    mcv::run_block_compacted<CHAIN>(A, tid, q);                                               \
  }                                                                                           \
  extern "C" __global__ void __launch_bounds__(MCV_BLOCK) NAME##_batch(const mcv::BatchArgs B) { \
    __shared__ mcv::QueueLDS q;                                                               \
    mcv::LaunchArgs A;                                                                        \
    int64_t tid;                                                                              \
    while (mcv::next_item(B, A, tid)) mcv::run_block_compacted<CHAIN>(A, tid, q);             \
  }
