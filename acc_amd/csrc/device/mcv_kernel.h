// mcv_kernel.h -- skeleton of the fused instrument kernels.
//
// The reference's generated kernel (acc/run_script.py:182-196) gives each
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
//               workgroups pull (instance, block of MCV_BLOCK streams) items from
//               one atomic counter until all are done, so uneven instances
//               share the GPU instead of queueing one launch after another.
#pragma once
#include "mcv_core.h"
#include "mcv_components.h"
#include "mcv_samples.h"
#include "mcv_launch.h"

// Workgroup size of the chain kernels, fixed per code object (the fuser passes
// -DMCV_BLOCK=...); the runtime reads it back from `mcv_block_size`.
#ifndef MCV_BLOCK
#define MCV_BLOCK 256
#endif
#if MCV_BLOCK % 64 != 0 || MCV_BLOCK < 64 || MCV_BLOCK > 1024
#error "MCV_BLOCK must be a multiple of 64 in [64, 1024]"
#endif
// The LDS bin cache is compiled into the plain kernels only (the fuser
// defines MCV_LDS_CACHE 1 for them); elsewhere its code would cost registers.
#ifndef MCV_LDS_CACHE
#define MCV_LDS_CACHE 0
#endif
#define MCV_QCAP (2 * MCV_BLOCK)  // >= (MCV_BLOCK - 1) + MCV_BLOCK queued neutrons

// Random streams per lane (1 except in stream-carrying compacted kernels): a
// workgroup serves MCV_BLOCK * MCV_SPL consecutive streams, lane l owning
// streams l, l + MCV_BLOCK, ...  The runtime reads it from `mcv_streams_per_lane`.
#ifndef MCV_SPL
#define MCV_SPL 1
#endif

// Streams per work item of the stream-carrying batch kernels: a workgroup
// refills its stream slots from its item (an LDS counter), so an item of
// several workgroups' worth of streams ends in one drained queue, not several.
// (64 instances, 1.6e8 neutrons, items in longest-first order: 2 workgroups'
// worth 48.1 ms, 4: 40.8, 8: 37.4, 16: 36.0, 32: 36.3, 64: 43.7.)
#ifndef MCV_RNG_ITEM_STREAMS
#define MCV_RNG_ITEM_STREAMS (16 * MCV_BLOCK * MCV_SPL)
#endif

extern "C" __device__ int mcv_block_size = MCV_BLOCK;
extern "C" __device__ int mcv_streams_per_lane = MCV_SPL;

// Optional occupancy target (waves per SIMD) the register allocator must meet;
// above the natural occupancy of a chain it trades spills for latency hiding.
#ifdef MCV_WAVES_PER_EU
#define MCV_KERNEL_ATTR __launch_bounds__(MCV_BLOCK) __attribute__((amdgpu_waves_per_eu(MCV_WAVES_PER_EU)))
#else
#define MCV_KERNEL_ATTR __launch_bounds__(MCV_BLOCK)
#endif

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

// ------------------------------------------------------------------ monitor histograms
// Small monitor histograms live in LDS for the lifetime of a block of streams
// (LDS FP64 atomics), and are added to the global histograms with one
// device-scope atomic per non-empty bin when the block is done.  Larger ones
// are replicated in global memory when that fits a budget: adds to one bin
// serialise at the memory-side atomic unit, and a direct beam puts thousands
// of adds on each of a few thousand bins, so the workgroups spread theirs
// over the replicas (the runtime folds them afterwards); the rest go straight
// to the histogram.
extern __shared__ double mcv_lds_hist[];

// Global destination of monitor comp's (3, nbins) block for this workgroup:
// its replica when replicated, else the histogram itself.
__device__ __forceinline__ double* hist_block(const LaunchArgs& A, int comp, int64_t nbins) {
  if (A.roff[comp] >= 0) {
    const int64_t r = (int64_t)(blockIdx.x & (unsigned)A.rmask[comp]);
    return A.R + A.roff[comp] + r * 3 * nbins;
  }
  return A.H + A.hoff[comp];
}

// ------------------------------------------------------------------ LDS bin cache
// A direct beam puts thousands of events on each of a few thousand bins, and
// every event of a large monitor would otherwise be three device-scope atomics
// performed past the XCD's L2 (one 32-byte memory write each).  A workgroup
// keeps the bins its events hit in an open-addressed LDS table: key =
// (ckbase[comp] << MCV_CACHE_TILE_SHIFT) + bin, up to kCacheProbes linear
// probes, a slot claimed with one LDS compare-and-swap; an event whose probed
// slots all hold other bins goes to the global block.  At its end the
// workgroup sorts its claimed slots by tile (a counting sort in LDS) and
// writes them to its dump with plain stores; the runtime's merge kernel then
// adds each tile of the dumps into the histograms, one workgroup per tile, so
// no bin takes more than one global write per launch from the cache.
// Layout (doubles from cache_off): keys [cache_slots / 2] (int32), N, sum p
// and sum p^2 [cache_slots each], then the tile counters [cache_ntiles + 1]
// (int32).
constexpr int32_t kCacheEmpty = -1;
// c1's PSD (about 4,500 beam bins of 62,500, events spread flatly over them)
// through a 4,096-slot table, 39,000 neutrons per workgroup, simulated: 18 %
// of events miss with 2 probes, 4.3 % with 4, 0.3 % with 8.
constexpr int kCacheProbes = 8;

__device__ __forceinline__ bool cache_add(const LaunchArgs& A, int comp, int64_t b, double p) {
  double* base = mcv_lds_hist + A.cache_off;
  int32_t* keys = reinterpret_cast<int32_t*>(base);
  double* val = base + (A.cache_slots >> 1);
  const int32_t key = (int32_t)(((uint32_t)A.ckbase[comp] << MCV_CACHE_TILE_SHIFT) + (uint32_t)b);
  const uint32_t mask = (uint32_t)A.cache_slots - 1u;
  uint32_t s = ((uint32_t)key * 2654435761u) >> 7;
  for (int probe = 0; probe < kCacheProbes; ++probe, s += 1u) {
    const int j = (int)(s & mask);
    int32_t k = keys[j];
    if (k == kCacheEmpty) {
      k = atomicCAS(keys + j, kCacheEmpty, key);
      if (k == kCacheEmpty) k = key;
    }
    if (k == key) {
      atomicAdd(val + j, 1.0);
      atomicAdd(val + A.cache_slots + j, p);
      atomicAdd(val + 2 * A.cache_slots + j, p * p);
      return true;
    }
  }
  return false;
}

__device__ __forceinline__ void hist_count(const LaunchArgs& A, int comp, int64_t nbins, int64_t b, double p) {
  const int lo = A.loff[comp];
  if (lo >= 0) {
    double* L = mcv_lds_hist + lo;
    atomicAdd(L + b, 1.0);
    atomicAdd(L + nbins + b, p);
    atomicAdd(L + 2 * nbins + b, p * p);
  } else if (MCV_LDS_CACHE && ((A.cmask >> comp) & 1u) && cache_add(A, comp, b, p)) {
    return;
  } else {
    hist_add3(hist_block(A, comp, nbins), nbins, b, p);
  }
}

// All threads of the workgroup call these (barriers when LDS histograms or the cache are in use).
__device__ __forceinline__ void lds_hist_zero(const LaunchArgs& A) {
  const bool cache = MCV_LDS_CACHE && A.cache_slots > 0;
  if (A.lds_nseg == 0 && !cache) return;
  for (int j = threadIdx.x; j < A.lds_total; j += blockDim.x) mcv_lds_hist[j] = 0;
  if (cache) {
    double* base = mcv_lds_hist + A.cache_off;
    int32_t* keys = reinterpret_cast<int32_t*>(base);
    double* val = base + (A.cache_slots >> 1);
    for (int j = threadIdx.x; j < A.cache_slots; j += blockDim.x) keys[j] = kCacheEmpty;
    for (int j = threadIdx.x; j < 3 * A.cache_slots; j += blockDim.x) val[j] = 0;
  }
  __syncthreads();
}

// The workgroup's claimed slots, ordered by tile, to its dump (after a barrier).
__device__ __forceinline__ void cache_dump(const LaunchArgs& A) {
  const int S = A.cache_slots, T = A.cache_ntiles;
  double* base = mcv_lds_hist + A.cache_off;
  const int32_t* keys = reinterpret_cast<const int32_t*>(base);
  const double* val = base + (S >> 1);
  int32_t* cnt = reinterpret_cast<int32_t*>(base + (S >> 1) + 3 * S);
  for (int t = threadIdx.x; t < T; t += blockDim.x) cnt[t] = 0;
  __syncthreads();
  for (int j = threadIdx.x; j < S; j += blockDim.x) {
    const int32_t key = keys[j];
    if (key != kCacheEmpty) atomicAdd(cnt + (key >> MCV_CACHE_TILE_SHIFT), 1);
  }
  __syncthreads();
  int32_t* off = A.cdoff + (int64_t)blockIdx.x * (T + 1);
  if (threadIdx.x < 64) {  // exclusive scan of the tile counts by wave 0, 64 tiles a step
    const int lane = (int)threadIdx.x;
    int carry = 0;
    for (int t0 = 0; t0 < T; t0 += 64) {
      const int t = t0 + lane;
      const int v = t < T ? cnt[t] : 0;
      int incl = v;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(incl, o, 64);
        if (lane >= o) incl += u;
      }
      if (t < T) {
        cnt[t] = carry + incl - v;
        off[t] = carry + incl - v;
      }
      carry += __shfl(incl, 63, 64);
    }
    if (lane == 0) off[T] = carry;
  }
  __syncthreads();
  double* dump = A.cdump + (int64_t)blockIdx.x * ((S >> 1) + 3 * S);
  int32_t* dkey = reinterpret_cast<int32_t*>(dump);
  double* dval = dump + (S >> 1);
  for (int j = threadIdx.x; j < S; j += blockDim.x) {
    const int32_t key = keys[j];
    if (key == kCacheEmpty) continue;
    const int pos = atomicAdd(cnt + (key >> MCV_CACHE_TILE_SHIFT), 1);
    dkey[pos] = key;
    dval[pos] = val[j];
    dval[S + pos] = val[S + j];
    dval[2 * S + pos] = val[2 * S + j];
  }
}

__device__ __forceinline__ void lds_hist_flush(const LaunchArgs& A) {
  const bool cache = MCV_LDS_CACHE && A.cache_slots > 0;
  if (A.lds_nseg == 0 && !cache) return;
  __syncthreads();
  if (cache) cache_dump(A);
  // constant indices only: the batch kernel keeps its LaunchArgs copy out of
  // scratch memory only if no field is indexed at run time
#pragma unroll
  for (int k = 0; k < MCV_LDS_SEG; ++k) {
    if (k < A.lds_nseg) {
      const double* L = mcv_lds_hist + A.lds_loff[k];
      double* G = A.H + A.lds_hoff[k];
      for (int j = threadIdx.x; j < A.lds_len[k]; j += blockDim.x) {
        const double v = L[j];
        if (v != 0) unsafeAtomicAdd(G + j, v);
      }
    }
  }
  __syncthreads();
}

__device__ __forceinline__ void count_processed(unsigned long long* counter, int64_t mine) {
  // wavefront reduction, one atomic per wave
  unsigned long long v = (unsigned long long)mine;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(counter, v);
}

// Neutrons of stream tid: [tid*npt, min((tid+1)*npt, N)), or this launch's
// slice of them, or empty for a stream outside this launch's range (a run
// split into launches by stream ranges, or by slices of every stream with the
// states carried between launches, computes every stream exactly as one
// launch over all of them does).  Returns whether the stream is in range.
__device__ __forceinline__ bool stream_range(const LaunchArgs& A, int64_t tid, Rng& rng, int64_t& idx,
                                             int64_t& end) {
  idx = end = 0;
  if (tid < A.nthreads && tid < A.tid_end) {
    const int64_t r = tid - A.rng_first;
    rng.s0 = A.rng[2 * r];
    rng.s1 = A.rng[2 * r + 1];
    const int64_t first = tid * A.npt;
    int64_t e = first + A.npt;
    idx = first + A.slice0;
    if (A.slice_len > 0 && idx + A.slice_len < e) e = idx + A.slice_len;
    end = e < A.N ? e : A.N;
    if (end < idx) end = idx;
    return true;
  }
  return false;
}

// ------------------------------------------------------------------ plain chains
// Chain::body(Rng&, Neutron&, const LaunchArgs&, int64_t nidx) returns the dump flag: 1/0
// whether the neutron reached the end alive, or, for multiple-scattering
// chains, the number of outputs.
template <class Chain>
__device__ __forceinline__ void run_lane(const LaunchArgs& A, int64_t tid) {
  lds_hist_zero(A);
  Rng rng;
  int64_t idx, end;
  stream_range(A, tid, rng, idx, end);
  Neutron n;
  n.s1 = 0;
  n.s2 = 0;
  for (int64_t i = idx; i < end; ++i) {
    const int alive = (int)Chain::body(rng, n, A, i);
    dump_neutron(A, i, n, alive);
  }
  count_processed(A.counter, end - idx);
  lds_hist_flush(A);
}

// The plain kernel: workgroup g runs the blocks of MCV_BLOCK streams g, g +
// gridDim.x, ... of the launch's range, so that with a grid of the resident
// workgroups their LDS histograms and bin caches serve many streams before
// one flush (with one block per workgroup it is run_lane).
template <class Chain>
__device__ __forceinline__ void run_lanes(const LaunchArgs& A) {
  lds_hist_zero(A);
  const int64_t last = A.tid_end < A.nthreads ? A.tid_end : A.nthreads;
  int64_t mine = 0;
  Neutron n;
  n.s1 = 0;
  n.s2 = 0;
  for (int64_t first = A.tid_base + (int64_t)blockIdx.x * MCV_BLOCK; first < last;
       first += (int64_t)gridDim.x * MCV_BLOCK) {
    Rng rng;
    int64_t idx, end;
    const bool live = stream_range(A, first + threadIdx.x, rng, idx, end);
    for (int64_t i = idx; i < end; ++i) {
      const int alive = (int)Chain::body(rng, n, A, i);
      dump_neutron(A, i, n, alive);
    }
    mine += end - idx;
    if (A.rng_out != nullptr && live) {  // the next slice continues this stream
      const int64_t r = first + threadIdx.x - A.rng_first;
      A.rng_out[2 * r] = rng.s0;
      A.rng_out[2 * r + 1] = rng.s1;
    }
  }
  count_processed(A.counter, mine);
  lds_hist_flush(A);
}

// ------------------------------------------------------------------ buffered mode
// acc/components/ComponentBase.py:79-107: the chain runs over the rows of a
// neutron array in place, with the partition of the array's length.
__device__ __forceinline__ void load_row(const double* __restrict__ r, Neutron& n) {
  n.x = r[0]; n.y = r[1]; n.z = r[2]; n.vx = r[3]; n.vy = r[4]; n.vz = r[5];
  n.s1 = r[6]; n.s2 = r[7]; n.t = r[8]; n.p = r[9];
}

__device__ __forceinline__ void store_row(double* __restrict__ r, const Neutron& n) {
  r[0] = n.x; r[1] = n.y; r[2] = n.z; r[3] = n.vx; r[4] = n.vy; r[5] = n.vz;
  r[6] = n.s1; r[7] = n.s2; r[8] = n.t; r[9] = n.p;
}

template <class Chain>
__device__ __forceinline__ void run_lane_process(const LaunchArgs& A, int64_t tid) {
  lds_hist_zero(A);
  Rng rng;
  int64_t idx, end;
  stream_range(A, tid, rng, idx, end);
  for (int64_t i = idx; i < end; ++i) {
    Neutron n;
    load_row(A.io + 10 * i, n);
    Chain::body(rng, n, A, i);
    store_row(A.io + 10 * i, n);
  }
  count_processed(A.counter, end - idx);
  lds_hist_flush(A);
}

// Chain::scatter(rng, n, A, out) of a lone multiple-scattering sample: up to
// A.io_m outputs, the unused rows are written as zeros.
template <class Chain>
__device__ __forceinline__ void run_lane_process_ms(const LaunchArgs& A, int64_t tid) {
  Rng rng;
  int64_t idx, end;
  stream_range(A, tid, rng, idx, end);
  for (int64_t i = idx; i < end; ++i) {
    Neutron n;
    load_row(A.io + 10 * i, n);
    Chain::scatter(rng, n, A, A.io_out + 10 * (int64_t)A.io_m * i);
  }
  count_processed(A.counter, end - idx);
}

// ------------------------------------------------------------------ compacted chains
// A chain whose components after its first culling point (the first
// component after the source that drops neutrons in numbers, e.g. a guide
// entrance) are expensive runs in two phases per workgroup.  Every lane
// produces neutrons from its own stream, in stream order, through phase A;
// survivors are appended to an LDS queue (slots from a wavefront ballot +
// prefix count, one LDS atomic per wave), and phase B runs over the queue with
// all lanes of whole wavefronts busy instead of one in four.
//
// Chain::phase_a returns r > 0 to queue the neutron (2: phase B is to skip
// the rest of the component phase A ended in, which did not let the neutron
// in); r <= 0 when it is done in phase A, with -r its dump flag (0 for a
// dropped neutron; a multiple-scattering chain whose neutron misses the
// sample finishes its one output in phase A and returns -1).
// Chain::phase_b(..., skip) returns the dump flag; the queue keeps `skip` in
// the sign of the neutron index or the top bit of the owner slot.
//
// Control decisions need the workgroup-wide queue length and whether any lane
// still has work; each wave publishes its two counts in a per-round slot
// (double-buffered by round parity, waves are never more than one round
// apart), so one barrier per round makes every lane see the same values.
// Queued neutrons keep their spins only where a source can set them
// (MCV_SPINS, e.g. stored neutrons); every other source leaves them 0, and
// no component changes them, so the queue need not hold them.
#ifndef MCV_SPINS
#define MCV_SPINS 0
#endif

template <int CAP>
struct QueueCore {
  double x[CAP], y[CAP], z[CAP], vx[CAP], vy[CAP], vz[CAP], t[CAP], p[CAP];
#if MCV_SPINS
  double s1[CAP], s2[CAP];
#endif
  int64_t idx[CAP];
  int count;  // slot allocator
  int wave_new[2][MCV_BLOCK / 64], wave_more[2][MCV_BLOCK / 64];
};

// deterministic phase B: up to (MCV_BLOCK - 1) + MCV_BLOCK queued neutrons
struct QueueLDS : QueueCore<MCV_QCAP> {};

static_assert(MCV_BLOCK * MCV_SPL <= 32768, "stream slots must fit the owner field below its skip bit");
constexpr uint16_t kOwnerSkip = 0x8000;

// drawing phase B: the states and next neutron indices of all the
// workgroup's streams live here.  A queued neutron names its stream slot;
// that stream is paused while the neutron waits, so its slot holds the
// neutron's stream state and its index is the slot's next index - 1.
struct QueueRngLDS {
  double x[MCV_QCAP], y[MCV_QCAP], z[MCV_QCAP], vx[MCV_QCAP], vy[MCV_QCAP], vz[MCV_QCAP], t[MCV_QCAP],
      p[MCV_QCAP];
#if MCV_SPINS
  double s1[MCV_QCAP], s2[MCV_QCAP];
#endif
  // narrow fields keep a 128-lane, 2-slot workgroup's queue under a sixth of the
  // CU's LDS, so LDS does not cap residency below the register limit
  uint16_t owner[MCV_QCAP];  // stream slot: lane * MCV_SPL + stream of the lane
  uint64_t st0[MCV_BLOCK * MCV_SPL], st1[MCV_BLOCK * MCV_SPL];
  int64_t next[MCV_BLOCK * MCV_SPL], end[MCV_BLOCK * MCV_SPL];
  uint8_t returned[MCV_BLOCK * MCV_SPL];
  int count;  // slot allocator
  int wave_new[2][MCV_BLOCK / 64], wave_more[2][MCV_BLOCK / 64];
};

template <class Q>
__device__ __forceinline__ void q_store_state(Q& q, int s, const Neutron& n) {
  q.x[s] = n.x; q.y[s] = n.y; q.z[s] = n.z; q.vx[s] = n.vx; q.vy[s] = n.vy; q.vz[s] = n.vz;
  q.t[s] = n.t; q.p[s] = n.p;
#if MCV_SPINS
  q.s1[s] = n.s1; q.s2[s] = n.s2;
#endif
}

template <class Q>
__device__ __forceinline__ void q_load_state(const Q& q, int s, Neutron& n) {
  n.x = q.x[s]; n.y = q.y[s]; n.z = q.z[s]; n.vx = q.vx[s]; n.vy = q.vy[s]; n.vz = q.vz[s];
  n.t = q.t[s]; n.p = q.p[s];
#if MCV_SPINS
  n.s1 = q.s1[s]; n.s2 = q.s2[s];
#else
  n.s1 = 0; n.s2 = 0;
#endif
}

template <class Q>
__device__ __forceinline__ void q_store(Q& q, int s, const Neutron& n, int64_t idx) {
  q_store_state(q, s, n);
  q.idx[s] = idx;
}

template <class Q>
__device__ __forceinline__ int64_t q_load(const Q& q, int s, Neutron& n) {
  q_load_state(q, s, n);
  return q.idx[s];
}

// Append `has` lanes' neutrons; returns the slot of this lane (or -1).
template <class Q>
__device__ __forceinline__ int q_append(Q& q, bool has) {
  const unsigned lane = threadIdx.x & 63u;
  const unsigned long long m = __ballot(has);
  int base = 0;
  if (lane == 0 && m) base = atomicAdd(&q.count, (int)__popcll(m));
  base = __shfl(base, 0, 64);
  return has ? base + (int)__popcll(m & ((1ull << lane) - 1ull)) : -1;
}

// Barrier; every lane gets the workgroup totals of `has` (neutrons appended
// this round) and whether any lane has `more`.
template <class Q>
__device__ __forceinline__ void round_totals(Q& q, int& par, bool has, bool more, int& n_new,
                                             bool& any_more) {
  const unsigned long long mh = __ballot(has), mm = __ballot(more);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    q.wave_new[par][w] = (int)__popcll(mh);
    q.wave_more[par][w] = mm != 0ull;
  }
  __syncthreads();
  n_new = 0;
  int mo = 0;
#pragma unroll
  for (int k = 0; k < MCV_BLOCK / 64; ++k) {
    n_new += q.wave_new[par][k];
    mo |= q.wave_more[par][k];
  }
  any_more = mo != 0;
  par ^= 1;
}

// Phase B draws nothing: each stream runs ahead through phase A while its
// earlier neutrons wait in the queue, and the draw order per stream is the
// reference's.  All MCV_BLOCK threads must call this together (barriers).
template <class Chain>
__device__ __forceinline__ void run_block_compacted(const LaunchArgs& A, int64_t tid, QueueLDS& q) {
  const int lt = threadIdx.x;
  lds_hist_zero(A);
  if (lt == 0) q.count = 0;
  Rng rng;
  int64_t idx, end;
  stream_range(A, tid, rng, idx, end);
  const int64_t mine = end - idx;
  int par = 0, cnt = 0;
  __syncthreads();
  while (true) {
    bool has = false;
    Neutron n;
    n.s1 = 0;
    n.s2 = 0;
    const int64_t my = idx;
    int r = 0;
    if (idx < end) {
      r = Chain::phase_a(rng, n, A, my);
      has = r > 0;
      if (!has) dump_neutron(A, my, n, -r);
      ++idx;
    }
    const int s = q_append(q, has);
    if (has) q_store(q, s, n, r > 1 ? ~my : my);
    int n_new;
    bool more;
    round_totals(q, par, has, idx < end, n_new, more);
    cnt += n_new;
    if (cnt >= MCV_BLOCK || (!more && cnt > 0)) {
      const int take = cnt < MCV_BLOCK ? cnt : MCV_BLOCK;
      if (lt < take) {
        Neutron b;
        const int64_t bidx = q_load(q, cnt - take + lt, b);
        const bool skip = bidx < 0;
        const int alive = (int)Chain::phase_b(b, A, skip);
        dump_neutron(A, skip ? ~bidx : bidx, b, alive);
      }
      cnt -= take;
      __syncthreads();
      if (lt == 0) q.count = cnt;
      __syncthreads();
    } else if (!more) {
      break;
    }
  }
  count_processed(A.counter, mine);
  lds_hist_flush(A);
  __syncthreads();  // the queue is reused by the next block of streams
}

// Stream slots of the stream-carrying kernels (QueueRngLDS and the queues
// derived from it): slot k of lane lt starts with stream tid + k * MCV_BLOCK;
// `done` marks slots whose stream has no neutron left.
template <class Q>
__device__ __forceinline__ void slots_init(const LaunchArgs& A, Q& q, int64_t tid, unsigned& done, int64_t& mine) {
  const int lt = threadIdx.x;
#pragma unroll
  for (int k = 0; k < MCV_SPL; ++k) {
    Rng r;
    int64_t idx, end;
    stream_range(A, tid + (int64_t)k * MCV_BLOCK, r, idx, end);
    const int o = lt * MCV_SPL + k;
    q.st0[o] = r.s0;
    q.st1[o] = r.s1;
    q.next[o] = idx;
    q.end[o] = end;
    q.returned[o] = 0;
    mine += end - idx;
    if (idx >= end) done |= 1u << k;
  }
}

// An exhausted, unpaused slot of this lane takes the pool's next stream (one
// atomic per wave; the pool counter may live in global memory or LDS); a lane
// that draws an empty stream or one at or past `pool_end` marks the pool dry
// (streams past the last non-empty one are empty).
template <class Q>
__device__ __forceinline__ void slots_refill(const LaunchArgs& A, Q& q, unsigned long long* pool, int64_t pool_base,
                                             int64_t pool_end, unsigned busy, unsigned& done, bool& dry,
                                             int64_t& mine) {
  constexpr unsigned ALL = (1u << MCV_SPL) - 1u;
  const unsigned lane = threadIdx.x & 63u;
  const unsigned freeslots = done & ~busy & ALL;
  const bool want = freeslots != 0u;
  const unsigned long long m = __ballot(want);
  if (!m) return;
  const int first = __builtin_ctzll(m);
  unsigned long long base = 0;
  if ((int)lane == first) base = atomicAdd(pool, (unsigned long long)__popcll(m));
  base = __shfl(base, first, 64);
  if (!want) return;
  const int64_t sid = pool_base + (int64_t)(base + __popcll(m & ((1ull << lane) - 1ull)));
  Rng r;
  int64_t idx = 0, end = 0;
  if (sid < pool_end) stream_range(A, sid, r, idx, end);
  if (idx < end) {
    const int k = __builtin_ctz(freeslots);
    const int o = (int)threadIdx.x * MCV_SPL + k;
    q.st0[o] = r.s0;
    q.st1[o] = r.s1;
    q.next[o] = idx;
    q.end[o] = end;
    mine += end - idx;
    done &= ~(1u << k);
  } else {
    dry = true;
  }
}

// Phase B draws too: a surviving neutron carries its stream's state into the
// queue and that stream pauses until phase B hands the state back, so every
// stream still draws for its neutrons one after another in the reference's
// order.  Each lane owns MCV_SPL stream slots and produces from any that is
// not paused, so lanes keep producing while their earlier neutrons wait.
// Phase B runs over a full workgroup of queued neutrons, or over everything
// queued when no lane can produce.  `tid` is the lane's first stream (slot k
// starts with stream tid + k * MCV_BLOCK).
//
// With a stream pool (`pool` non-null: a launch-wide counter, streams from
// `pool_base` on), a slot whose stream is exhausted takes the next stream of
// the pool, so a persistent grid keeps its queue filling until the pool is
// dry.  Without it, the streams of a workgroup run out together and the last
// phase-B batches of every workgroup run partly empty (the stream pauses make
// the queue fill ever slower towards the end).  Which lane runs a stream
// changes nothing in its neutrons.
template <class Chain>
__device__ __forceinline__ void run_block_compacted_rng(const LaunchArgs& A, int64_t tid, QueueRngLDS& q,
                                                        unsigned long long* pool = nullptr, int64_t pool_base = 0,
                                                        int64_t pool_end = INT64_MAX) {
  const int lt = threadIdx.x;
  constexpr unsigned ALL = (1u << MCV_SPL) - 1u;
  lds_hist_zero(A);
  if (lt == 0) q.count = 0;
  unsigned busy = 0, done = 0;  // per stream slot of this lane: paused / no neutrons left
  bool dry = pool == nullptr;   // the pool has no stream left for this lane
  int64_t mine = 0;
  slots_init(A, q, tid, done, mine);
  int par = 0, cnt = 0, rr = 0;
  __syncthreads();
  while (true) {
    if (!dry) slots_refill(A, q, pool, pool_base, pool_end, busy, done, dry, mine);
    // a stream that is neither paused nor exhausted, round robin from rr
    const unsigned avail = ~(busy | done) & ALL;
    bool has = false;
    Neutron n;
    n.s1 = 0;
    n.s2 = 0;
    Rng w;
    int64_t my = 0;
    int sel = -1, ra = 0;
    if (avail) {
      const unsigned rot = ((avail >> rr) | (avail << (MCV_SPL - rr))) & ALL;
      sel = (rr + __builtin_ctz(rot)) % MCV_SPL;
      const int o = lt * MCV_SPL + sel;
      w.s0 = q.st0[o];
      w.s1 = q.st1[o];
      my = q.next[o];
      ra = Chain::phase_a(w, n, A, my);
      has = ra > 0;
      if (!has) dump_neutron(A, my, n, -ra);
      q.st0[o] = w.s0;
      q.st1[o] = w.s1;
      q.next[o] = my + 1;
      if (my + 1 >= q.end[o]) done |= 1u << sel;
      if (has) busy |= 1u << sel;
      rr = (sel + 1) % MCV_SPL;
    }
    const int s = q_append(q, has);
    if (has) {
      q_store_state(q, s, n);
      q.owner[s] = (uint16_t)(lt * MCV_SPL + sel) | (ra > 1 ? kOwnerSkip : 0);
    }
    int n_new;
    bool can_a;
    round_totals(q, par, has, (~(busy | done) & ALL) != 0 || (!dry && (done & ~busy & ALL) != 0), n_new, can_a);
    cnt += n_new;  // < 2 * MCV_BLOCK: a full workgroup is taken as soon as one is queued
    if (cnt >= MCV_BLOCK || (!can_a && cnt > 0)) {
      const int take = cnt >= MCV_BLOCK ? MCV_BLOCK : cnt;
      if (lt < take) {
        Neutron b;
        const int slot = cnt - take + lt;
        q_load_state(q, slot, b);
        const int ow = q.owner[slot];
        const int o = ow & (kOwnerSkip - 1);
        Rng r;
        r.s0 = q.st0[o];
        r.s1 = q.st1[o];
        const int alive = (int)Chain::phase_b(r, b, A, (ow & kOwnerSkip) != 0);
        dump_neutron(A, q.next[o] - 1, b, alive);
        q.st0[o] = r.s0;  // the stream continues from here
        q.st1[o] = r.s1;
        q.returned[o] = 1;
      }
      cnt -= take;
      __syncthreads();
#pragma unroll
      for (int k = 0; k < MCV_SPL; ++k) {
        const int o = lt * MCV_SPL + k;
        if (q.returned[o]) {
          q.returned[o] = 0;
          busy &= ~(1u << k);
        }
      }
      if (lt == 0) q.count = cnt;
      __syncthreads();
    } else if (!can_a) {
      break;  // nothing queued and no stream has neutrons left
    }
  }
  count_processed(A.counter, mine);
  lds_hist_flush(A);
  __syncthreads();
}

// ------------------------------------------------------------------ batched scans
// Next (instance, stream block) for the workgroup; false when all are taken.
__device__ __forceinline__ bool next_item(const BatchArgs& B, LaunchArgs& A, int64_t& tid,
                                          int64_t item_streams = MCV_BLOCK * MCV_SPL) {
  __shared__ int64_t s_item;
  if (threadIdx.x == 0) s_item = (int64_t)atomicAdd(B.work, 1ull);
  __syncthreads();
  const int64_t item = s_item;
  __syncthreads();
  if (item >= B.item0[B.ninst]) return false;
  int lo = 0, hi = B.ninst - 1;  // last position with item0 <= item
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (B.item0[mid] <= item) lo = mid;
    else hi = mid - 1;
  }
  const int64_t k = B.inst_index[lo];
  A = B.base;
  A.P = B.base.P + k * B.pstride;
  A.H = B.base.H + k * B.hstride;
  A.N = B.inst_n[lo];
  A.nthreads = B.inst_nthreads[lo];
  A.npt = B.inst_npt[lo];
  A.counter = B.inst_processed + k;
  A.dump = nullptr;
  A.dump_count = 0;
  tid = (item - B.item0[lo]) * item_streams + threadIdx.x;
  return true;
}

}  // namespace mcv

#define MCV_DEFINE_KERNELS_PLAIN(NAME, CHAIN)                                                  \
  extern "C" __device__ int mcv_lds_cache = MCV_LDS_CACHE;                                    \
  extern "C" __device__ int mcv_slices = 1;                                                   \
  extern "C" __global__ void MCV_KERNEL_ATTR NAME(const mcv::LaunchArgs A) {                  \
    mcv::run_lanes<CHAIN>(A);                                                                 \
  }                                                                                           \
  extern "C" __global__ void MCV_KERNEL_ATTR NAME##_batch(const mcv::BatchArgs B) { \
    mcv::LaunchArgs A;                                                                        \
    int64_t tid;                                                                              \
    while (mcv::next_item(B, A, tid)) mcv::run_lane<CHAIN>(A, tid);                           \
  }

#define MCV_DEFINE_KERNELS_COMPACTED(NAME, CHAIN)                                              \
  extern "C" __global__ void MCV_KERNEL_ATTR NAME(const mcv::LaunchArgs A) {     \
    __shared__ mcv::QueueLDS q;                                                               \
    const int64_t tid = A.tid_base + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;          \
    mcv::run_block_compacted<CHAIN>(A, tid, q);                                               \
  }                                                                                           \
  extern "C" __global__ void MCV_KERNEL_ATTR NAME##_batch(const mcv::BatchArgs B) { \
    __shared__ mcv::QueueLDS q;                                                               \
    mcv::LaunchArgs A;                                                                        \
    int64_t tid;                                                                              \
    while (mcv::next_item(B, A, tid)) mcv::run_block_compacted<CHAIN>(A, tid, q);             \
  }

#define MCV_DEFINE_KERNELS_COMPACTED_RNG(NAME, CHAIN)                                          \
  extern "C" __device__ int mcv_stream_pool = 1;                                              \
  extern "C" __device__ int mcv_item_streams = MCV_RNG_ITEM_STREAMS;                          \
  extern "C" __global__ void MCV_KERNEL_ATTR NAME(const mcv::LaunchArgs A) {     \
    __shared__ mcv::QueueRngLDS q;                                                            \
    const int64_t tid = A.tid_base + (int64_t)blockIdx.x * (blockDim.x * MCV_SPL) + threadIdx.x; \
    mcv::run_block_compacted_rng<CHAIN>(A, tid, q, A.pool,                                    \
                                        A.tid_base + (int64_t)gridDim.x * (MCV_BLOCK * MCV_SPL)); \
  }                                                                                           \
  extern "C" __global__ void MCV_KERNEL_ATTR NAME##_batch(const mcv::BatchArgs B) { \
    __shared__ mcv::QueueRngLDS q;                                                            \
    __shared__ unsigned long long item_pool;                                                  \
    mcv::LaunchArgs A;                                                                        \
    int64_t tid;                                                                              \
    while (mcv::next_item(B, A, tid, MCV_RNG_ITEM_STREAMS)) {                                  \
      const int64_t first = tid - threadIdx.x;                                                \
      if (threadIdx.x == 0) item_pool = 0;                                                    \
      mcv::run_block_compacted_rng<CHAIN>(A, tid, q, &item_pool, first + MCV_BLOCK * MCV_SPL, \
                                          first + MCV_RNG_ITEM_STREAMS);                      \
    }                                                                                         \
  }

#define MCV_DEFINE_PROCESS_KERNEL(NAME, CHAIN)                                                 \
  extern "C" __global__ void MCV_KERNEL_ATTR NAME(const mcv::LaunchArgs A) {                  \
    const int64_t tid = A.tid_base + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;          \
    mcv::run_lane_process<CHAIN>(A, tid);                                                     \
  }

#define MCV_DEFINE_PROCESS_MS_KERNEL(NAME, CHAIN)                                              \
  extern "C" __global__ void MCV_KERNEL_ATTR NAME(const mcv::LaunchArgs A) {                  \
    const int64_t tid = A.tid_base + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;          \
    mcv::run_lane_process_ms<CHAIN>(A, tid);                                                  \
  }
