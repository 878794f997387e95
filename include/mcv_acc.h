/*
 * mcv_acc.h -- C ABI of the MI355X neutron ray-tracing engine (libmcvacc.so).
 *
 * The reference (mcvine.acc) has no native boundary: its Python component
 * classes hand `propagate_params` tuples to Numba, which JIT-compiles one fused
 * CUDA kernel per instrument (acc/run_script.py:37-139, templates :165-289) and launches it from
 * SourceBase.call_process_no_buffer (acc/components/sources/SourceBase.py:36-52).
 * This header is the boundary that replaces that launch: the Python mirror of
 * the component API packs each component into an mcv_comp_desc (same parameter
 * order as the reference's propagate_params), the fuser compiles the chain into
 * one gfx950 code object, and these entry points own device memory, RNG
 * streams, launches and histogram read-back.
 *
 * Conventions: 0 = success, negative = error (message via mcv_last_error(),
 * thread-local).  All host arrays are caller-owned, C-contiguous float64/int64.
 * Device buffers are owned by the mcv_instr handle.  A handle is not
 * thread-safe; calls block until the device work has finished.
 */
#ifndef MCV_ACC_H
#define MCV_ACC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCV_ABI_VERSION 2

/* ---- error classes (mirrors the exception types the reference raises) ---- */
#define MCV_OK 0
#define MCV_ERR_ARG -1     /* bad argument / parameter (RuntimeError, AssertionError) */
#define MCV_ERR_HIP -2     /* HIP runtime failure (no device, launch failure) */
#define MCV_ERR_MODULE -3  /* code object missing or kernel symbol absent */
#define MCV_ERR_COUNT -4   /* processed-neutron count mismatch (run_script.py:218-222) */
#define MCV_ERR_NOMEM -5   /* a launch's lane scratch exceeds the budget (MCV_SCRATCH_BUDGET_MB) */

/* ---- component type tags ------------------------------------------------
 * params/iparams/data layouts (float64 unless stated):
 *
 * MCV_ARM                no parameters (acc/components/optics/arm.py:14-43)
 * MCV_SOURCE_SIMPLE      params[13] = square, width, height, radius, wl_distr,
 *                        Lambda0, dLambda, E0, dE, xw, yh, dist, pmul
 *                        (acc/components/sources/source_simple.py:104-108)
 * MCV_GUIDE              params[10] = ww, hh, hw1, hh1, l, R0, Qc, alpha, m, W
 *                        (acc/components/optics/guide.py:45-50)
 * MCV_GUIDE_TAPERING     params[9] = l_seg, R0, Qcx, Qcy, alphax, alphay, mx, my, W
 *                        iparams[1] = nseg; data[10*nseg] = rows w1, w2, h1, h2,
 *                        ww, hh, whalf, hhalf, lwhalf, lhhalf (row-major, 10 x nseg)
 *                        (acc/components/optics/guide_tapering.py:73-75,117-139)
 * MCV_SAMPLE_HSS         single scatterer (homogeneous_single_scatterer.py:26-102)
 * MCV_SAMPLE_HMS         multiple scatterer (homogeneous_multiple_scatterer.py:22-342)
 *     iparams[8] = n_shape_nodes, kernel_type, max_scattered_neutrons,
 *                  max_ms_loops, max_ms_loops_path1, kernel_data_offset,
 *                  kernel_n0, kernel_n1
 *     params[16] = mu (absorption coefficient at 2200 m/s, 1/m),
 *                  sigma (scattering coefficient, 1/m),
 *                  minimum_neutron_event_probability, kernel params[13]
 *     data = CSG nodes (MCV_CSG_NODE_SIZE doubles each, root = node 0),
 *            then kernel tables starting at kernel_data_offset.
 * MCV_MON_PSD            params[4] = xmin, xmax, ymin, ymax; iparams[2] = nx, ny
 *                        (acc/components/monitors/psd_monitor.py:53-56)
 * MCV_MON_WAVELENGTH     params[6] = xmin, xmax, ymin, ymax, Lmin, Lmax;
 *                        iparams[1] = nchan (wavelength_monitor.py:39-42)
 * MCV_MON_IQE            params[9] = Ei, Qmin, Qmax, Emin, Emax, max_angle_in_plane,
 *                        min_angle_in_plane, max_angle_out_of_plane,
 *                        min_angle_out_of_plane; iparams[2] = nQ, nE
 *                        (acc/components/monitors/iqe_monitor.py:43-49)
 * MCV_DET_TUBES          cylindrical array of vertical tubes (no reference
 *                        counterpart): params[8] = radius, height, phi_min,
 *                        phi_max (deg), tof_min, tof_max (s), tube_diameter,
 *                        unused; iparams[3] = ntubes, npixels, ntof
 *
 * Monitors own a (3, n_bins) float64 histogram: rows N, sum p, sum p^2.
 */
#define MCV_ARM 0
#define MCV_SOURCE_SIMPLE 1
#define MCV_GUIDE 2
#define MCV_GUIDE_TAPERING 3
#define MCV_SAMPLE_HSS 4
#define MCV_SAMPLE_HMS 5
#define MCV_SLIT 6      /* params[6] = xmin, xmax, ymin, ymax, radius^2, cut (acc/components/optics/slit.py:56-58) */
#define MCV_BEAMSTOP 7  /* params[6] as MCV_SLIT (acc/components/optics/beamstop.py:60-62) */
#define MCV_MON_PSD 10
#define MCV_MON_WAVELENGTH 11
#define MCV_MON_IQE 12
#define MCV_DET_TUBES 13
#define MCV_MON_PSD_4PI 14 /* params[5] = phimin, phimax, thetamin, thetamax, radius; iparams = nphi, ntheta
                              (acc/components/monitors/psd_monitor_4pi.py:42-45) */
#define MCV_MON_POSDIV 15  /* params[5] = xmin, xmax, ymin, ymax, maxdiv; iparams = npos, ndiv
                              (acc/components/monitors/posdiv_monitor.py:54-57) */
/* engine-internal sources */
#define MCV_SRC_IQE_NORMALIZER 20 /* params[3] = Ei, Emin, Emax: the I(Q,E) monitor's
                                     Monte-Carlo normalizer neutrons
                                     (acc/components/monitors/iqe_monitor.py:146-174) */
#define MCV_SRC_STORAGE 21        /* iparams[1] = N; data = (N, 10) stored neutrons: neutron i is
                                     row i % N (acc/components/sources/neutronfromstorage.py:70-82) */

/* scattering kernels of samples (acc/kernels/__init__.py:8-22,36-44) */
#define MCV_KERNEL_ISOTROPIC 1  /* acc/kernels/isotropic.py:5-21 */
#define MCV_KERNEL_SQE_GRID 2   /* S(Q,E) on a regular grid (no reference) */
#define MCV_KERNEL_PHONON_DOS 3 /* 3-peak phonon DOS kernel (no reference) */

/* CSG nodes: {type, a, b, c, left, right} (acc/geometry/arrow_intersect.py:18-123) */
#define MCV_CSG_NODE_SIZE 6
#define MCV_CSG_SPHERE 1       /* a = radius */
#define MCV_CSG_CYLINDER 2     /* a = radius, b = height; axis = local z */
#define MCV_CSG_BLOCK 3        /* a = width (x), b = height (y), c = thickness (z) */
#define MCV_CSG_UNION 10
#define MCV_CSG_DIFFERENCE 11
#define MCV_CSG_INTERSECTION 12

/* run flags */
#define MCV_FLAG_GATE_DEAD 1u  /* drop a neutron once p <= 0 (mcvine semantics);
                                  clear = acc fused-kernel semantics, where an
                                  absorbed neutron keeps flowing downstream */

typedef struct mcv_comp_desc {
  int32_t type;
  int32_t n_params;
  const double* params;
  int32_t n_iparams;
  const int64_t* iparams;
  int64_t n_data;
  const double* data;
  int64_t n_bins; /* monitors: bins of the (3, n_bins) histogram; else 0 */
} mcv_comp_desc;

typedef struct mcv_run_stats {
  uint64_t processed;     /* neutron counter (run_script.py:195) */
  int64_t nblocks;        /* reference partition (SourceBase.py:40-43) */
  int64_t nthreads;       /* streams */
  int64_t npt;            /* neutrons per stream */
  double ms_rng_init;     /* device time of the stream-seeding kernel */
  double ms_kernel;       /* device time of the fused propagate kernel(s) */
  double ms_total;        /* host wall time of the call */
  int64_t nchunks;        /* launches the run was split into (buffer_size, scratch budget) */
  int64_t scratch_per_lane; /* private segment bytes per lane of the launched kernel */
} mcv_run_stats;

typedef struct mcv_instr mcv_instr;

/* Version of this ABI (MCV_ABI_VERSION). */
int mcv_abi_version(void);

/* Message of the last failing call on this host thread. */
const char* mcv_last_error(void);

/* Number of visible HIP devices (0 when none). */
int mcv_device_count(void);

/* Build an instrument handle from a component chain.  Replaces the reference's
 * InstrumentWrapper (acc/run_script.py:198-208): offsets[(n-1)*3] and
 * rotmats[(n-1)*9] are the frame changes between consecutive components
 * (calcTransformations, run_script.py:141-162).  code_object/kernel_name name
 * the fused kernel that the fuser compiled for exactly this chain. */
mcv_instr* mcv_instr_create(const mcv_comp_desc* comps, int32_t ncomp,
                            const double* offsets, const double* rotmats,
                            const char* code_object, const char* kernel_name);

/* Replace the float parameters of component `comp` (same layout as at create
 * time) without recompiling: the reference bakes them in at JIT time
 * (homogeneous_single_scatterer.py:36-40); here they are device data. */
int mcv_instr_set_params(mcv_instr* h, int32_t comp, const double* params,
                         int32_t n_params);

/* Run ncount neutrons.  Replaces the generated run() and
 * SourceBase.call_process_no_buffer (run_script.py:210-223, SourceBase.py:36-52):
 * the reference partition (ntotalthreads, threads_per_block -> nblocks, npt)
 * and numba's xoroshiro128+ seeding are reproduced, so stream tid runs neutrons
 * [tid*npt, min((tid+1)*npt, ncount)).  Histograms accumulate across calls
 * until mcv_monitor_clear.  ntotalthreads/tpb <= 0 select 1e6 / 512
 * (acc/config.py:1-2).  flags: MCV_FLAG_*.
 * buffer_size (run_script.run's keyword, which the reference accepts and
 * ignores) > 0 splits the run into launches of about buffer_size neutrons:
 * for a chain compiled as a plain kernel, launch c runs neutrons [c*k,
 * (c+1)*k) of every stream (k = buffer_size / streams, at least 1) and keeps
 * each stream's state for the next launch; otherwise launch c runs whole
 * streams [c*S, (c+1)*S) of the same partition (S = buffer_size / npt, at
 * least 1), seeded as in one launch.  Either way every neutron is computed
 * exactly as in a single launch.  The runtime also splits a launch whose
 * lanes' scratch would exceed the budget. */
int mcv_instr_run(mcv_instr* h, uint64_t ncount, uint64_t seed,
                  int64_t ntotalthreads, int32_t tpb, uint32_t flags,
                  uint64_t buffer_size, mcv_run_stats* stats);

/* mcv_instr_run without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No reference counterpart.  Same arguments, partition,
 * streams and results as mcv_instr_run; the stream states and launch
 * counters live in a static per-device buffer of the library (16 MiB + 64 B,
 * held until the process ends) instead of the handle's allocations, so a
 * launch may hold at most 2^20 streams (the default partition holds
 * 1,000,448): a larger one is refused with MCV_ERR_ARG unless buffer_size
 * splits it.  Counters are cleared by the call itself.  Every call on one GPU
 * shares that buffer: calls that overlap on one GPU (from several host
 * threads) race. */
int mcv_instr_run_noalloc(mcv_instr* h, uint64_t ncount, uint64_t seed,
                          int64_t ntotalthreads, int32_t tpb, uint32_t flags,
                          uint64_t buffer_size, mcv_run_stats* stats);

/* Copy the (3, n_bins) histogram of monitor component `comp` into out3. */
int mcv_monitor_read(mcv_instr* h, int32_t comp, double* out3);

/* Zero every monitor histogram of the handle. */
int mcv_monitor_clear(mcv_instr* h);

/* Parity support: on the next mcv_instr_run, record the final neutron state
 * (10 doubles, after the last component) of neutron indices [first, first+count)
 * into the handle; mcv_instr_read_dump copies them out (count x 10) together
 * with a per-neutron flag (1 = reached the end of the chain alive). */
int mcv_instr_set_dump(mcv_instr* h, int64_t first, int64_t count);
int mcv_instr_read_dump(mcv_instr* h, double* states, int32_t* alive);

/* Buffered mode (acc/components/ComponentBase.py:79-107 `process`, launched as
 * acc/components/StochasticComponentBase.py:21-54 does): `code_object` /
 * `kernel_name` name the chain's buffered-mode kernel (the fuser's process
 * mode for the same chain).  neutrons (n x 10 doubles, x y z vx vy vz s1 s2 t p)
 * are run through the chain in place, stream tid taking rows
 * [tid*npt, min((tid+1)*npt, n)) of the partition of n, streams seeded from
 * `seed`; monitors add to the handle's histograms (not cleared).  For a lone
 * multiple-scattering sample (m = its max_scattered_neutrons > 0) the rows
 * are inputs only and out_ms (n*m x 10) receives neutron i's outputs in rows
 * [i*m, i*m + num) with zeros after them.  Rows are not filtered: the
 * reference's p > 0 selection is the caller's. */
int mcv_instr_process(mcv_instr* h, const char* code_object, const char* kernel_name, double* neutrons, int64_t n,
                      uint64_t seed, int64_t ntotalthreads, int32_t tpb, uint32_t flags, double* out_ms,
                      int32_t m, mcv_run_stats* stats);

/* Vector-Jacobian product of the monitor histograms.  No reference
 * counterpart (the reference only simulates forward).  `code_object` /
 * `kernel_name` name the chain's adjoint kernel (generated by the fuser's
 * adjoint mode for the same chain); it re-runs the same ncount neutrons with
 * the same seed and partition as mcv_instr_run and returns, in grad[nscore],
 *   sum_i sum_monitors (G[1][bin] p_i + 2 G[2][bin] p_i^2) d log p_i / d theta
 * for the chain's weight-carrying parameters theta in score-slot order, with
 * g_hist laid out like the histograms (hist_total doubles: per monitor rows
 * N, sum p, sum p^2; the N row is ignored).  grad_dense (ndense = the node
 * count of the chain's S(Q,E)-grid sample, or 0) receives the gradient with
 * respect to that grid's values. */
int mcv_instr_adjoint(mcv_instr* h, const char* code_object, const char* kernel_name, uint64_t ncount,
                      uint64_t seed, int64_t ntotalthreads, int32_t tpb, uint32_t flags, const double* g_hist,
                      int32_t nscore, double* grad, int64_t ndense, double* grad_dense, mcv_run_stats* stats);

/* Seed streams [first, first + n) exactly as numba's create_xoroshiro128p_states
 * does (SplitMix64(seed) into both words, stream i = stream i-1 jumped 2^64) and
 * copy the (n, 2) uint64 states to host memory.  Used to pin the device seeding. */
int mcv_rng_states(uint64_t seed, int64_t first, int64_t n, uint64_t* out);

/* Scratch guard arithmetic (no device work): how many workgroups of `block`
 * lanes needing scratch_per_lane bytes each may be resident together within
 * budget_bytes, at most resident_groups; 0 when not even one fits.  Every
 * launch applies it with the kernel's private segment size and the device's
 * budget (MCV_SCRATCH_BUDGET_MB, default an eighth of the free device memory):
 * persistent grids are narrowed, mcv_instr_run splits its launch, and the other
 * entry points return MCV_ERR_NOMEM.  No reference counterpart. */
int64_t mcv_scratch_fit(int64_t scratch_per_lane, int32_t block, int64_t resident_groups, int64_t budget_bytes);

/* Release the handle and its device memory. */
void mcv_instr_destroy(mcv_instr* h);

/* ---- batched scans ----------------------------------------------------------
 * No reference counterpart: the reference runs one instrument per
 * run_script.run call (acc/run_script.py:10-35), so a parameter scan is a loop
 * of separate launches.  A batch holds `ninst` instances of one chain structure
 * (comps = ninst x ncomp descriptors, instance-major; offsets/rotmats per
 * instance) with their own parameters and histograms; tables and histogram
 * sizes must match across instances. */
typedef struct mcv_batch mcv_batch;

mcv_batch* mcv_batch_create(const mcv_comp_desc* comps, int32_t ninst, int32_t ncomp,
                            const double* offsets, const double* rotmats,
                            const char* code_object, const char* kernel_name);

/* All instances in one persistent launch (kernel `<kernel_name>_batch`):
 * workgroups take (instance, block of 256 streams) items from one atomic
 * counter.  Instance i runs ncounts[i] neutrons with the same partition and
 * streams as a separate mcv_instr_run with the same seed would.  grid <= 0
 * selects 1024 workgroups; use_graph != 0 captures seeding + launch in a
 * hipGraph and replays it. */
int mcv_batch_run(mcv_batch* b, const uint64_t* ncounts, uint64_t seed, int64_t ntotalthreads,
                  int32_t tpb, uint32_t flags, int32_t grid, int32_t use_graph, mcv_run_stats* stats);

/* The same instances as one launch each, in order (the loop of separate
 * runs a scan would otherwise be): the comparison point of mcv_batch_run. */
int mcv_batch_run_sequential(mcv_batch* b, const uint64_t* ncounts, uint64_t seed,
                             int64_t ntotalthreads, int32_t tpb, uint32_t flags, mcv_run_stats* stats);

int mcv_batch_monitor_read(mcv_batch* b, int32_t inst, int32_t comp, double* out3);

/* Every instance's histograms in one copy: out holds n = ninst x H doubles,
 * H = 3 x (sum of the monitors' bins), instance-major, each instance's
 * monitors in chain order as mcv_batch_monitor_read returns them (one
 * device-to-host copy instead of one per instance and monitor). */
int mcv_batch_histograms_read(mcv_batch* b, double* out, int64_t n);

/* Every instance's histograms delivered to the caller's arrays: block j is
 * the len[j] doubles from offset src[j] of the layout mcv_batch_histograms_read
 * returns, written to dst[j] (accumulate 0) or added to it (accumulate 1).
 * One device-to-host copy into the batch's pinned staging buffer (allocated on
 * first use, freed with the batch), in 8 MB pieces, while up to 8 host
 * threads write each block out as soon as its piece is in: what a scan's
 * caller needs to fill each instance's monitor
 * arrays (the `out` arrays of acc/components/monitors), without a temporary copy.
 * Blocks must not overlap in dst. */
int mcv_batch_histograms_deliver(mcv_batch* b, int64_t nblocks, const int64_t* src, const int64_t* len,
                                 double* const* dst, int32_t accumulate);
int mcv_batch_clear(mcv_batch* b);
void mcv_batch_destroy(mcv_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* MCV_ACC_H */
