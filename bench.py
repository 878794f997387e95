#!/usr/bin/env python
"""Benchmark: neutrons/s of the fused propagate path on one MI355X (FP64).

A step is one run of the workload instrument over `ncount` synthetic neutrons
(reference partition, numba-seeded xoroshiro128+ streams), with the
instrument's parameters and tables already resident in HBM: device-side stream
seeding + the fused kernel + the processed-count check.  Default workload:
the source -> guide -> sample -> monitor chain (Source_simple -> 10 m guide ->
100-segment elliptic tapered guide -> isotropic sphere -> I(Q,E) + PSD
monitors), 1e8 neutrons per step.  The line also carries the other configured
chains side by side (`configs`: c1 at 1e7 neutrons in launches of 1e6, c2 with
and without dead-neutron gating, c3 with the 400 x 128 x 500 tube array), the
64-instance batched scan and the adjoint/forward cost.

Prints one JSON line.  `roofline` prices the fused kernel against the FP64
vector peak with the algorithmic work counted by the oracle's counting build
(mean FP64 ops per neutron, transcendentals weighted 20); `cpu_baseline` times
the CPU oracle on this host's cores over a bounded sample of the same workload.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "neutrons/s per instrument at 1 MI355X (FP64); adjoint/forward cost ratio"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "c2": ("guide transport: Source_simple (r 5 cm, E 5-100 meV) -> 10 m straight guide 5x5 cm m=2 -> "
           "10 m 100-segment elliptic guide 6->3 cm -> wavelength (200 bins) + PSD 100x100", int(1e8)),
    "c2s": ("source -> guide -> sample -> monitor: Source_simple (r 5 cm, E 5-100 meV) -> 10 m straight guide "
            "5x5 cm m=2 -> 10 m 100-segment elliptic guide 6->3 cm -> isotropic Al sphere (r 1 cm, 5 b / 2 b) -> "
            "I(Q,E) 120x100 + PSD 250x250", int(1e8)),
    "c1": ("Source_simple (E 60+-10 meV) -> isotropic sphere (r 1 cm, 5 b / 2 b) -> I(Q,E) 120x100 + "
           "PSD 250x250", int(1e7)),
    "c3": ("Source_simple (E 60+-10 meV) -> hollow cylinder (r 1.0/1.5 cm, h 5 cm), S(Q,E) grid 200x200, "
           "multiple scattering -> 400 tubes x 128 pixels x 500 TOF bins", int(1e8)),
}


def oracle_flops_per_neutron(instr, n=1000000, seed=0, gate=True):
    """Mean FP64 ops per neutron over the first n neutrons (counting oracle)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle
    from acc_amd.instrument import calcTransformations
    path = os.path.join(ROOT, "oracle", "liboracle_flops.so")
    L = ctypes.CDLL(path)
    saved_lib, saved_path = pyoracle._lib, pyoracle.LIB
    try:
        pyoracle._lib, pyoracle.LIB = None, path
        pyoracle.lib()
        L.orc_flops_read.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
        L.orc_flops_reset()
        descs = [c.mcv_desc() for c in instr.components]
        off, rot = calcTransformations(instr)
        _, processed, _, _ = pyoracle.run(descs, off, rot, n, seed, gate=gate, nthreads=1)
        c = np.zeros(3, dtype=np.uint64)
        L.orc_flops_read(c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    finally:
        pyoracle._lib, pyoracle.LIB = saved_lib, saved_path
    return {"arith": float(c[0]) / processed, "trans": float(c[1]) / processed,
            "draws": float(c[2]) / processed, "F_alg": (float(c[0]) + 20.0 * float(c[1])) / processed,
            "neutrons": int(processed)}


def cpu_baseline(instr, ncount, seconds=12.0, ntotalthreads=0, tpb=0, cores=None, gate=True):
    """Oracle on this host's cores (or `cores` threads) over a bounded leading
    part of the workload (whole streams of the reference partition)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from acc_amd.instrument import calcTransformations
    if cores is None:
        cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16") or 16))
    descs = [c.mcv_desc() for c in instr.components]
    off, rot = calcTransformations(instr)
    nblocks, nthreads, npt = pyoracle.partition(ncount, ntotalthreads, tpb)
    # probe, then size the sample to ~`seconds`
    probe = max(1, min(nthreads, 2000000 // max(npt, 1)))
    t = time.perf_counter()
    _, done, _, _ = pyoracle.run(descs, off, rot, ncount, 0, ntotalthreads, tpb, gate, (0, probe), cores)
    dt = time.perf_counter() - t
    rate = done / dt
    streams = min(nthreads, max(probe, int(rate * seconds / max(npt, 1))))
    t = time.perf_counter()
    _, done, _, _ = pyoracle.run(descs, off, rot, ncount, 0, ntotalthreads, tpb, gate, (0, streams), cores)
    dt = time.perf_counter() - t
    return {"value": done / dt, "unit": "neutrons/s", "cores": cores, "kind": "port",
            "sample": "first %d of %d streams of the same partition (%d neutrons, seed 0), %.1f s" % (
                streams, nthreads, done, dt)}


def batch_scan(ninst=64, reps=3):
    """The 64-instance scan (Ei 20-200 meV log-uniform, sample rotation
    0-180 deg, counts log-uniform 1e5-1e7, default_rng(0)) two ways, each
    like for like:
     * on the device: one persistent hipGraph launch of the batch kernel vs
       the same instances launched back to back (stream seeding + kernel per
       instance, no host work between), both timed by HIP events and by wall
       clock around the launches;
     * end to end through the Python API, histograms delivered into every
       instance's monitor `out` arrays: BatchedScan.run + store() vs a loop of
       run_instrument calls (the one-run-per-instrument API).
    Every run checks every instance's processed count."""
    from acc_amd import configs
    from acc_amd.batch import BatchedScan
    from acc_amd.run_script import run_instrument
    from acc_amd.autotune import tuned_scan
    ins, n = configs.scan(ninst)
    total = int(n.sum())
    scan, kw = tuned_scan(ins, n, use_cache=False)
    tuned = {"block": scan.block, "waves_per_eu": scan.waves_per_eu, "streams_per_lane": scan.streams_per_lane}

    def checked(st):
        assert st["processed"] == total, st  # (the runtime also checks every instance's count)
        return st

    try:
        checked(scan.run(n, seed=0, **kw))
        checked(scan.run_sequential(n, seed=0))
        b = [checked(scan.run(n, seed=0, **kw))["ms_kernel"] for _ in range(reps)]
        t = time.perf_counter()
        for _ in range(reps):
            checked(scan.run(n, seed=0, **kw))
        b_wall = (time.perf_counter() - t) / reps * 1e3
        s = [checked(scan.run_sequential(n, seed=0))["ms_kernel"] for _ in range(reps)]
        t = time.perf_counter()
        for _ in range(reps):
            checked(scan.run_sequential(n, seed=0))
        s_wall = (time.perf_counter() - t) / reps * 1e3
        scan.store()
        t = time.perf_counter()
        for _ in range(reps):
            checked(scan.run(n, seed=0, **kw))
            scan.store()
        b_e2e = (time.perf_counter() - t) / reps * 1e3
    finally:
        scan.close()
    from acc_amd import config
    singles = [configs.scan_instance(float(e), float(r), "one%d" % k)
               for k, (e, r) in enumerate(zip(*configs.scan_plan(ninst)[:2]))]
    saved = config.kernel_block, config.kernel_waves_per_eu
    saved_spl = config.kernel_streams_per_lane
    config.kernel_block, config.kernel_waves_per_eu = tuned["block"], tuned["waves_per_eu"]
    config.kernel_streams_per_lane = tuned["streams_per_lane"]
    try:
        for instr, c in zip(singles, n):  # builds the handles (module loads) outside the timing
            run_instrument(instr, int(min(c, 1000)), seed=0, accumulate=False)
        t = time.perf_counter()
        for instr, c in zip(singles, n):
            assert run_instrument(instr, int(c), seed=0, accumulate=False)["processed"] == c
        api_wall = (time.perf_counter() - t) * 1e3
    finally:
        config.kernel_block, config.kernel_waves_per_eu = saved
        config.kernel_streams_per_lane = saved_spl
    bk = sum(b) / len(b)
    return {"instances": ninst, "neutrons": total, "ms_batched_graph_kernel": bk, "ms_batched_graph_wall": b_wall,
            "ms_back_to_back_kernels": sum(s) / len(s), "ms_back_to_back_wall": s_wall,
            "speedup_vs_back_to_back": s_wall / b_wall,
            "ms_batched_with_histograms_wall": b_e2e, "ms_run_instrument_loop_wall": api_wall,
            "speedup_end_to_end": api_wall / b_e2e,
            "neutrons_per_s_batched": total / (b_wall * 1e-3), "launch_shape": tuned,
            "note": "same tuned kernel everywhere; back-to-back = one launch per instance, no host work between; "
                    "end to end = histograms of every instance copied into its monitors' out arrays"}


def adjoint_cost(ncount=int(1e7), reps=3):
    """Adjoint (vector-Jacobian product) vs forward kernel time on the
    differentiable-fit chain: c5 (source at 60 meV -> straight + tapered guide
    -> sphere with the 3-peak phonon-DOS kernel -> I(Q,E) + PSD), 14
    parameters (9 peak parameters, 2 cross-sections, shared m/W/alpha),
    1e7 neutrons, the forward kernel at its autotuned launch shape."""
    import numpy as np
    from acc_amd import configs
    from acc_amd.adjoint import DifferentiableRun
    from acc_amd.autotune import tune_instrument
    best = tune_instrument(configs.c5(), ncount, use_cache=False)
    run = DifferentiableRun(configs.c5(), configs.C5_PARAMS, ncount=ncount, seed=0, block=best["block"],
                            waves_per_eu=best["waves_per_eu"], streams_per_lane=best["streams_per_lane"])
    theta = run.theta0()
    g = np.random.default_rng(0).uniform(0.5, 1.5, sum(run.sizes))
    run.forward(theta)
    run.vjp(g, theta)
    fwd, adj = [], []
    for _ in range(reps):
        run.forward(theta)
        fwd.append(run.last_forward_stats["ms_kernel"])
        run.vjp(g, theta)
        adj.append(run.last_adjoint_stats["ms_kernel"])
    f, b = min(fwd), min(adj)
    return {"workload": "c5 differentiable fit chain, 14 parameters", "ncount": ncount,
            "ms_forward_kernel": f, "ms_adjoint_kernel": b, "adjoint_forward_ratio": b / f,
            "forward_launch_shape": {"block": best["block"], "waves_per_eu": best["waves_per_eu"],
                                     "streams_per_lane": best["streams_per_lane"]},
            "check": "adjoint vs CRN central differences (h=1e-6 rel) of forward kernel and CPU oracle"}


def copy_kernel(nbytes=1 << 31, reps=3):
    """A plain device copy of known size: rocprofv3's FETCH_SIZE / WRITE_SIZE
    for it give the factor between those counters and bytes moved."""
    import torch
    x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_kernel_bytes_read": nbytes, "copy_kernel_bytes_written": nbytes, "reps": reps}))
    return 0


# (name, chain, neutrons per run, dead-neutron gating, neutrons per launch or 0).  c1 is
# configured in buffers of 1e6 neutrons; the reference's runner ignores buffer_size
# (acc/run_script.py:316-328) and runs it as one launch, which "c1-one-launch" times.
CONFIG_ROWS = (("c1", "c1", int(1e7), True, int(1e6)), ("c1-one-launch", "c1", int(1e7), True, 0),
               ("c2", "c2", int(1e8), True, 0),
               ("c2-ungated", "c2", int(1e8), False, 0), ("c2s", "c2s", int(1e8), True, 0),
               ("c2s-ungated", "c2s", int(1e8), False, 0), ("c3", "c3", int(1e8), True, 0))


def config_rows(names=None, steps=3, cpu_seconds=4.0, cpu_one_core=False, falg_neutrons=1000000):
    """The configured chains side by side: fused-kernel time (HIP events, summed
    over a run's launches) and wall time per run at the reference partition,
    neutrons/s, F_alg and the FP64 fraction, and the CPU oracle over a bounded
    sample (all host cores; one core too with cpu_one_core).  c2 and c2s run
    at their autotuned shape, c1's one launch at the better of 512- and
    1,024-lane workgroups, c1 in launches and c3 at the default one; c1 runs 1e7
    neutrons in launches of 1e6 as configured and "c1-one-launch" in one
    launch; "c2-ungated" keeps absorbed neutrons flowing (the reference's fused
    kernel exactly).  Rows of the same chain share its F_alg and CPU figures."""
    from acc_amd import configs
    from acc_amd._abi import MCV_FLAG_GATE_DEAD
    from acc_amd.autotune import tuned_handle
    from acc_amd.run_script import get_handle
    rows, measured = {}, {}  # measured: (chain, gate) -> F_alg and CPU figures of an earlier row
    for name, chain, ncount, gate, buf in CONFIG_ROWS:
        if names is not None and name not in names:
            continue
        print("[configs] %s: %d neutrons%s" % (name, ncount, " in launches of %d" % buf if buf else ""),
              file=sys.stderr, flush=True)
        instr = configs.BUILDERS[chain]()
        flags = MCV_FLAG_GATE_DEAD if gate else 0
        if chain in ("c2", "c2s"):
            h, best = tuned_handle(instr, ncount, use_cache=False)
            shape = "%s/%s/%s" % (best["block"], best["waves_per_eu"], best["streams_per_lane"])
        elif name == "c1-one-launch":
            # one launch gains from 1,024-lane workgroups (16 waves a CU, a few
            # registers spilled); launches of 1e6 lose from them
            h, best = tuned_handle(instr, ncount, blocks=(512, 1024), waves=(None,), use_cache=False)
            shape = "%s/%s" % (best["block"], best["waves_per_eu"])
        else:
            h, shape = get_handle(instr), "default"
        h.clear()
        h.run(ncount, 0, 0, 0, flags, buffer_size=buf)
        kms, walls, launches = [], [], 1
        for k in range(steps):
            h.clear()
            t = time.perf_counter()
            st = h.run(ncount, k, 0, 0, flags, buffer_size=buf)
            walls.append((time.perf_counter() - t) * 1e3)
            assert st["processed"] == ncount, st
            kms.append(st["ms_kernel"])
            launches = st["nchunks"]
        if (chain, gate) not in measured:
            m = {"fl": oracle_flops_per_neutron(configs.BUILDERS[chain](), n=falg_neutrons, gate=gate),
                 "cpu": cpu_baseline(configs.BUILDERS[chain](), ncount, seconds=cpu_seconds, gate=gate)}
            if cpu_one_core:
                m["cpu1"] = cpu_baseline(configs.BUILDERS[chain](), ncount, seconds=cpu_seconds, cores=1, gate=gate)
            measured[(chain, gate)] = m
        m = measured[(chain, gate)]
        fl = m["fl"]
        k_avg, w_avg = sum(kms) / steps, sum(walls) / steps
        row = {"ncount": ncount, "gated": gate, "launches": launches, "launch_shape": shape,
               "ms_kernel": k_avg, "ms_wall": w_avg,
               "neutrons_per_s_kernel": ncount / (k_avg * 1e-3), "neutrons_per_s_wall": ncount / (w_avg * 1e-3),
               "falg_ops_per_neutron": fl["F_alg"], "draws_per_neutron": fl["draws"],
               "fp64_frac": fl["F_alg"] * ncount / (k_avg * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
               "cpu_all_cores": m["cpu"]}
        if cpu_one_core:
            row["cpu_one_core"] = m["cpu1"]
        rows[name] = row
    return rows


def config_table(steps=3, cpu_seconds=5.0, falg_neutrons=100000):
    """--configs: the rows above with one-core CPU rates as well."""
    return {"configs": config_rows(steps=steps, cpu_seconds=cpu_seconds, cpu_one_core=True,
                                   falg_neutrons=falg_neutrons), "cpu_model": cpu_model()}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def traffic_from_profiles(workload, shape):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE for
    this workload's kernel at this launch shape (the shapes' spills differ), or
    None when that shape was not measured."""
    p = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        shapes = json.load(f).get("shapes", {})
    key = "%s/%s" % (shape["block"], shape["waves_per_eu"])
    if shape.get("streams_per_lane"):
        key += "/%s" % shape["streams_per_lane"]
    entry = shapes.get(key)
    return entry["hbm_bytes_per_launch"] if entry else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2s", choices=sorted(WORKLOADS))
    ap.add_argument("--ncount", type=float, default=None)
    ap.add_argument("--buffer-size", type=float, default=0,
                    help="neutrons per launch (whole streams; 0: one launch), as run_script.run's buffer_size")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--falg-neutrons", type=int, default=1000000)
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-instance batched scan")
    ap.add_argument("--no-tune", action="store_true", help="default launch shape instead of the autotuned one")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward cost measurement")
    ap.add_argument("--no-configs", action="store_true", help="skip the other configured chains (c1, c2, c3)")
    ap.add_argument("--shape", default=None, help="fixed launch shape BLOCK/WAVES[/STREAMS_PER_LANE] (e.g. 256/3) instead of tuning")
    ap.add_argument("--copy-kernel", action="store_true",
                    help="only run a plain 2 GiB device copy 3 times (HBM counter calibration) and exit")
    ap.add_argument("--configs", action="store_true",
                    help="only print the c1 / c2 / c2s / c3 table (GPU and CPU oracle rates) and exit")
    a = ap.parse_args()
    if a.copy_kernel:
        return copy_kernel()
    if a.configs:
        print(json.dumps(config_table()))
        return 0
    if a.gpus != 1:
        print("bench.py measures one GPU: the propagate path runs on a single MI355X here "
              "(sharding by stream ranges over several GPUs is not built); use --gpus 1", file=sys.stderr)
        return 2

    from acc_amd import configs
    from acc_amd._abi import MCV_FLAG_GATE_DEAD
    from acc_amd.autotune import tuned_handle
    from acc_amd.run_script import get_handle

    desc, default_n = WORKLOADS[a.workload]
    ncount = int(a.ncount or default_n)
    instr = configs.BUILDERS[a.workload]()
    if a.shape:
        parts = a.shape.split("/") + ["None", "None"]
        blk = int(parts[0])
        wv = None if parts[1] in ("None", "", "0") else int(parts[1])
        spl = None if parts[2] in ("None", "", "0") else int(parts[2])
        h = get_handle(instr, block=blk, waves_per_eu=wv, streams_per_lane=spl)
        shape = {"block": blk, "waves_per_eu": wv, "streams_per_lane": spl, "tuned": False}
    elif a.no_tune:
        h, shape = get_handle(instr), {"block": None, "waves_per_eu": None, "tuned": False}
    else:
        h, best = tuned_handle(instr, ncount, use_cache=False)
        shape = {"block": best["block"], "waves_per_eu": best["waves_per_eu"], "tuned": True,
                 "streams_per_lane": best["streams_per_lane"],
                 "trials_ms": {"%d/%s/%s" % (t["block"], t["waves_per_eu"], t["streams_per_lane"]): round(t["ms"], 2)
                               for t in best["trials"]}}
    seed = 0
    buf = int(a.buffer_size)
    for _ in range(a.warmup):
        h.clear()
        h.run(ncount, seed, 0, 0, MCV_FLAG_GATE_DEAD, buffer_size=buf)
    kernel_ms, rng_ms = [], []
    h.clear()
    t0 = time.perf_counter()
    for k in range(a.steps):
        st = h.run(ncount, seed + k, 0, 0, MCV_FLAG_GATE_DEAD, buffer_size=buf)
        kernel_ms.append(st["ms_kernel"])
        rng_ms.append(st["ms_rng_init"])
    elapsed = time.perf_counter() - t0
    value = ncount * a.steps / elapsed

    fl = oracle_flops_per_neutron(configs.BUILDERS[a.workload](), n=a.falg_neutrons)
    k_avg = sum(kernel_ms) / len(kernel_ms)
    achieved = fl["F_alg"] * ncount / (k_avg * 1e-3) / 1e12
    # "mfma" is the contract's name for the compute roof: the kernel is bound by FP64
    # vector arithmetic (no dense contraction, so no matrix instructions), and
    # MI355X's FP64 peak, 78.6 TFLOP/s, is the same for vector and matrix units
    roofline = {"bound": "mfma", "compute_unit": "fp64 VALU (MFMA unused)",
                "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic_from_profiles(a.workload, shape),
                "kernel": "mcv_fused (%s)" % a.workload, "kernel_ms_avg": k_avg,
                "falg_ops_per_neutron": fl["F_alg"], "falg_detail": fl,
                "rng_seed_kernel_ms_avg": sum(rng_ms) / len(rng_ms)}
    cpu = None if a.no_cpu_baseline else cpu_baseline(configs.BUILDERS[a.workload](), ncount)
    line = {
        "metric": METRIC, "value": value, "unit": "neutrons/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": elapsed / a.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{a.workload}: {desc}", "ncount_per_step": ncount, "buffer_size": buf or None,
                   "partition": "reference default (ntotalthreads 1e6, 512 threads/block)",
                   "semantics": "dead neutrons (p<=0) dropped", "launch_shape": shape},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    if not a.no_configs:
        line["configs"] = config_rows([r[0] for r in CONFIG_ROWS if r[0] != a.workload])
        line["cpu_model"] = cpu_model()
    if not a.no_batch:
        line["batch_scan"] = batch_scan()
    if not a.no_adjoint:
        line["adjoint"] = adjoint_cost()
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
