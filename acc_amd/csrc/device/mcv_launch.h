// mcv_launch.h -- argument block of a fused instrument kernel, shared by the
// host runtime (mcv_runtime.hip) and the generated chain kernels.
#pragma once
#include <stdint.h>

#define MCV_MAX_COMP 32
#define MCV_LDS_SEG 8  // monitors whose histograms a workgroup keeps in LDS
// Bins of the LDS bin cache's key space per merge tile (a power of two).
#define MCV_CACHE_TILE_SHIFT 12
#define MCV_CACHE_TILE (1 << MCV_CACHE_TILE_SHIFT)
#define MCV_CACHE_MAX_TILES 1024

namespace mcv {

// Passed by value as the single kernel argument; every field is uniform, so
// the kernel keeps it in SGPRs.  Offsets index the packed parameter blob P
// (per-component parameters, data tables and the (R, o) frame changes) and the
// histogram buffer H.
struct LaunchArgs {
  const double* P;
  double* H;
  const uint64_t* rng;               // stream states (2 words each) of streams rng_first, rng_first + 1, ...
  unsigned long long* counter;       // processed neutrons
  unsigned long long* pool;          // streams taken from the pool (persistent stream-carrying kernels)
  double* dump;                      // (dump_count, 10) final states, or null
  int32_t* dump_alive;               // (dump_count)
  int64_t dump_first, dump_count;
  int64_t N, nthreads, npt;          // reference partition
  int64_t tid_base;                  // first stream of this launch
  int64_t tid_end;                   // streams of this launch: [tid_base, min(tid_end, nthreads))
  int64_t rng_first;                 // stream whose state is rng[0..1]
  // Slice of every stream this launch runs: neutrons [tid*npt + slice0,
  // tid*npt + slice0 + slice_len) of stream tid (clipped to the stream and to
  // N; slice_len = 0: to the stream's end).  Only plain kernels (mcv_slices =
  // 1) take a slice_len, and then write each stream's state back to
  // rng_out[2 * (tid - rng_first)] for the next slice.
  int64_t slice0, slice_len;
  uint64_t* rng_out;
  int32_t gate;                      // MCV_FLAG_GATE_DEAD
  int32_t ncomp;
  int64_t off[MCV_MAX_COMP];         // parameters of component i
  int64_t doff[MCV_MAX_COMP];        // data table of component i
  int64_t hoff[MCV_MAX_COMP];        // (3, nbins) histogram of monitor i
  int64_t xoff[MCV_MAX_COMP];        // frame change i -> i+1: R[9], o[3]
  // Monitor histograms privatised in LDS (dynamic shared memory of lds_total
  // doubles): loff[i] = LDS offset of monitor i's (3, nbins) block or -1;
  // segment k copies LDS [lds_loff[k], +lds_len[k]) onto H + lds_hoff[k].
  int32_t loff[MCV_MAX_COMP];
  int32_t lds_nseg, lds_total;
  int32_t lds_loff[MCV_LDS_SEG], lds_len[MCV_LDS_SEG];
  int64_t lds_hoff[MCV_LDS_SEG];
  // Monitor histograms replicated in global memory: monitor i adds into
  // replica (workgroup & rmask[i]) of its block at R + roff[i] (roff[i] = -1:
  // straight into H); the runtime folds the replicas into H after the launch.
  double* R;
  int64_t roff[MCV_MAX_COMP];
  int32_t rmask[MCV_MAX_COMP];
  // LDS bin cache: monitors whose bit is set in cmask add through an
  // open-addressed table of cache_slots (a power of two) bins in dynamic LDS,
  // starting cache_off doubles in (after the privatised histograms).  Bin b of
  // monitor i has key (ckbase[i] << MCV_CACHE_TILE_SHIFT) + b: the monitors'
  // bins laid end to end in tiles of MCV_CACHE_TILE, cache_ntiles tiles in
  // all.  When the workgroup ends it writes its claimed bins, ordered by tile,
  // to its dump at cdump + blockIdx.x * (cache_slots / 2 + 3 * cache_slots)
  // doubles (keys [cache_slots] int32, then N, sum p, sum p^2 [cache_slots]
  // each) with tile t's entries at [cdoff[g][t], cdoff[g][t + 1]) of
  // cdoff + blockIdx.x * (cache_ntiles + 1); the runtime's merge kernel adds
  // the dumps into H.  cache_slots = 0: no cache.
  int32_t cache_slots, cache_off;
  uint32_t cmask;
  int32_t cache_ntiles;
  int32_t ckbase[MCV_MAX_COMP];
  double* cdump;
  int32_t* cdoff;
  // buffered mode: neutron i of the run is row i of io (N, 10), updated in
  // place; a multiple-scattering sample writes its io_m outputs per neutron to
  // io_out rows [i*io_m, (i+1)*io_m)
  double* io;
  double* io_out;
  int32_t io_m;
};

// Persistent kernel of a batched scan: instances of one chain structure with
// their own parameter blocks, histograms and neutron counts; all instances
// share the stream states (same seed, as separate runs would).
struct BatchArgs {
  LaunchArgs base;           // offsets, rng, gate; P/H of instance 0
  int64_t pstride, hstride;  // per-instance strides of the parameter blob and histograms
  // The instances in work order (the longest work items first): position j
  // is instance inst_index[j]; item0, inst_n, inst_nthreads and inst_npt are
  // listed in that order.
  const int64_t* item0;      // (ninst + 1) first work item of each position (prefix sums)
  const int64_t* inst_n;     // (ninst) neutrons
  const int64_t* inst_nthreads;
  const int64_t* inst_npt;
  const int64_t* inst_index;  // (ninst)
  unsigned long long* inst_processed;  // (ninst)
  unsigned long long* work;  // work-item counter (zeroed before the launch)
  int32_t ninst;
};

// Second argument of an adjoint kernel (mcv_adjoint.h).
struct AdjointArgs {
  const double* G;   // upstream gradient, laid out like the histograms (3 rows per monitor)
  double* grad;      // (nscore) result, zeroed before the launch
  double* dense;     // (ndense) gradient of a sample's S(Q,E) grid values, or null
};

}  // namespace mcv
