"""Launch-shape autotuner for the fused kernels.

The reference fixes its launch shape by configuration (acc/config.py:1-2:
ntotalthreads 1e6, 512 threads per block) and derives neutrons per thread from
it (acc/components/sources/SourceBase.py:36-52).  Here the shape is measured:

 * block: workgroup size the chain is compiled for (multiples of 64).  Each
   stream's work does not depend on it, so results are identical for every
   block; it changes occupancy and the width of the compaction queue.
 * waves_per_eu: occupancy target the register allocator must meet (None =
   the compiler's choice).  Chains whose natural allocation leaves 2 waves per
   SIMD can run faster at 3-4 with some spills, or slower: it is measured.
 * streams_per_lane: random streams each lane serves in compacted kernels
   whose later stages draw (each stream pauses while its neutron waits in the
   queue; more streams keep lanes producing, at the price of LDS).  Results
   do not depend on it.
 * grid: number of persistent workgroups of a batched scan (0 = default).
 * ntotalthreads (only when `neutrons_per_thread=True`): number of random
   streams, i.e. neutrons per thread.  This changes which stream draws which
   neutron, so histograms change within their statistics and the reference's
   default partition is no longer reproduced; off by default.

Each candidate is timed with the kernels' own HIP events (minimum over `reps`
runs after one warmup).  Choices are cached in a JSON file next to the kernel
cache, keyed by the chain's source (without its block size) and the workload.
"""
import hashlib
import json
import os
import threading

from . import config

DEFAULT_BLOCKS = (128, 256, 512)
DEFAULT_WAVES = (None, 3, 4)
DEFAULT_SPL = (1, 2, 4)
DEFAULT_GRIDS = (0,)
DEFAULT_NTOTALTHREADS = (250000, 500000, 1000000, 2000000)
_lock = threading.Lock()


def _cache_path():
    from .fuser import CACHE_DIR
    return os.path.join(CACHE_DIR, "autotune.json")


def _load():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key, value):
    with _lock:
        d = _load()
        d[key] = value
        path = _cache_path()
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp%d_%d" % (os.getpid(), threading.get_ident())
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, path)


def _key(kind, instrument, workload):
    from .fuser import generate_source
    src = generate_source(instrument, block=64, waves_per_eu=0)[0]  # text without the launch shape
    h = hashlib.sha1(src.encode())
    h.update(json.dumps(workload, sort_keys=True).encode())
    return "%s:%s" % (kind, h.hexdigest()[:16])


def _best(trials):
    return min(trials, key=lambda t: t["ms"])


def _spls(instrument, spls):
    """Streams-per-lane candidates: only stream-carrying compacted chains have the knob."""
    from .fuser import generate_source
    src = generate_source(instrument, block=64, waves_per_eu=0)[0]
    carried = "MCV_DEFINE_KERNELS_COMPACTED_RNG" in src
    return tuple(spls) if carried else (None,)


def tune_instrument(instrument, ncount, seed=0, blocks=DEFAULT_BLOCKS, waves=DEFAULT_WAVES, spls=DEFAULT_SPL,
                    neutrons_per_thread=False, ntotalthreads=DEFAULT_NTOTALTHREADS, reps=2, use_cache=True):
    """Best {block, waves_per_eu, streams_per_lane, ntotalthreads, ms} for
    `ncount` neutrons of `instrument` (the histograms it leaves behind are
    those of the last trial)."""
    from .run_script import get_handle
    from ._abi import MCV_FLAG_GATE_DEAD
    nts = tuple(ntotalthreads) if neutrons_per_thread else (config.ntotalthreads,)
    spls = _spls(instrument, spls)
    workload = {"ncount": int(ncount), "blocks": list(blocks), "waves": list(waves), "spls": list(spls),
                "nts": list(nts)}
    key = _key("instrument", instrument, workload)
    if use_cache:
        hit = _load().get(key)
        if hit:
            return hit
    trials = []
    for block in blocks:
        for w in waves:
            for spl in spls:
                h = get_handle(instrument, block=block, waves_per_eu=w, streams_per_lane=spl)
                for nt in nts:
                    h.clear()
                    h.run(int(ncount), seed, nt, config.threads_per_block, MCV_FLAG_GATE_DEAD)
                    ms = min(h.run(int(ncount), seed, nt, config.threads_per_block, MCV_FLAG_GATE_DEAD)["ms_kernel"]
                             for _ in range(reps))
                    trials.append({"block": int(block), "waves_per_eu": w, "streams_per_lane": spl,
                                   "ntotalthreads": int(nt), "ms": float(ms)})
    best = dict(_best(trials), trials=trials)
    _store(key, best)
    return best


def tune_scan(instruments, ncounts, seed=0, blocks=DEFAULT_BLOCKS, waves=DEFAULT_WAVES, spls=DEFAULT_SPL,
              grids=DEFAULT_GRIDS, neutrons_per_thread=False, ntotalthreads=DEFAULT_NTOTALTHREADS, reps=2,
              use_cache=True):
    """Best {block, waves_per_eu, streams_per_lane, grid, ntotalthreads, ms} for a batched scan."""
    from .batch import BatchedScan
    nts = tuple(ntotalthreads) if neutrons_per_thread else (config.ntotalthreads,)
    spls = _spls(instruments[0], spls)
    workload = {"ncounts": [int(n) for n in ncounts], "blocks": list(blocks), "waves": list(waves),
                "spls": list(spls), "grids": list(grids), "nts": list(nts), "ninst": len(instruments)}
    key = _key("scan", instruments[0], workload)
    if use_cache:
        hit = _load().get(key)
        if hit:
            return hit
    trials = []
    for block in blocks:
        for w in waves:
            for spl in spls:
                scan = BatchedScan(instruments, block=block, waves_per_eu=w, streams_per_lane=spl)
                try:
                    for grid in grids:
                        for nt in nts:
                            scan.run(ncounts, seed=seed, ntotalthreads=nt, grid=grid)
                            ms = min(scan.run(ncounts, seed=seed, ntotalthreads=nt, grid=grid)["ms_kernel"]
                                     for _ in range(reps))
                            trials.append({"block": int(block), "waves_per_eu": w, "streams_per_lane": spl,
                                           "grid": int(grid), "ntotalthreads": int(nt), "ms": float(ms)})
                finally:
                    scan.close()
    best = dict(_best(trials), trials=trials)
    _store(key, best)
    return best


def tuned_scan(instruments, ncounts, **kw):
    """(BatchedScan built for the tuned launch shape, keyword arguments for its run())."""
    from .batch import BatchedScan
    best = tune_scan(instruments, ncounts, **kw)
    scan = BatchedScan(instruments, block=best["block"], waves_per_eu=best["waves_per_eu"],
                       streams_per_lane=best["streams_per_lane"])
    return scan, {"grid": best["grid"], "ntotalthreads": best["ntotalthreads"]}


def tuned_handle(instrument, ncount, **kw):
    """(device handle built for the tuned launch shape, the tuning record)."""
    from .run_script import get_handle
    best = tune_instrument(instrument, ncount, **kw)
    return get_handle(instrument, block=best["block"], waves_per_eu=best["waves_per_eu"],
                      streams_per_lane=best["streams_per_lane"]), best
