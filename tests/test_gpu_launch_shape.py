"""Launch shape: the workgroup size and the streams per lane a chain is
compiled for change nothing in the results (each stream's neutrons are
computed the same way), and the autotuner returns a usable choice."""
import numpy as np
import pytest

from acc_amd import autotune, configs, prebuilt
from acc_amd._abi import MCV_FLAG_GATE_DEAD
from acc_amd.run_script import get_handle

pytestmark = pytest.mark.gpu


def _run(instr, block, N, ntot, spl=None):
    h = get_handle(instr, block=block, streams_per_lane=spl)
    h.set_dump(0, N)
    h.clear()
    st = h.run(N, 3, ntot, 512, MCV_FLAG_GATE_DEAD)
    assert st["processed"] == N
    states, alive = h.read_dump()
    hists = {i: h.read_monitor(i) for i in range(len(instr.components)) if h.nbins[i] > 0}
    return states, alive, hists


@pytest.mark.parametrize("name,block,spl", [("c2", 128, None), ("c2", 512, None), ("c2s", 128, 1), ("c2s", 512, 2),
                                            ("c2s", 256, 4)])
def test_block_size_invariance(name, block, spl):
    print(f"\n[launch] {name} block {block} streams/lane {spl}", flush=True)
    instr = prebuilt.instruments()[name]()
    N, ntot = 300000, 20000
    s0, a0, h0 = _run(instr, 256, N, ntot, spl=1 if spl else None)
    s1, a1, h1 = _run(instr, block, N, ntot, spl=spl)
    np.testing.assert_array_equal(a1, a0)
    np.testing.assert_array_equal(s1, s0)
    for i in h0:
        np.testing.assert_array_equal(h1[i][0], h0[i][0])
        np.testing.assert_allclose(h1[i][1:], h0[i][1:], rtol=1e-12, atol=1e-300)


def test_autotune_scan_small():
    print("\n[launch] autotune scan", flush=True)
    Ei, rot = [30.0, 90.0, 150.0], [0.0, 60.0, 120.0]
    ins = [configs.scan_instance(e, r, "t%d" % k) for k, (e, r) in enumerate(zip(Ei, rot))]
    n = [20000, 150000, 3000]
    kw = dict(blocks=(128, 256), waves=(None, 3), spls=(1, 4), grids=(0, 256))
    best = autotune.tune_scan(ins, n, use_cache=False, **kw)
    assert best["block"] in (128, 256) and best["grid"] in (0, 256) and best["waves_per_eu"] in (None, 3)
    assert best["streams_per_lane"] in (1, 4)
    assert len(best["trials"]) == 16 and all(t["ms"] > 0 for t in best["trials"])
    scan, kw = autotune.tuned_scan(ins, n, **kw)
    try:
        st = scan.run(n, seed=1, **kw)
        assert st["processed"] == sum(n)
    finally:
        scan.close()


@pytest.mark.parametrize("name", ["c3-small", "hms-ref"])
def test_ms_compaction_invariance(name):
    """Compacting a multiple-scattering chain at its sample (misses finish in
    phase A, interacting neutrons queued with their streams) leaves every
    neutron's outputs as the one-lane-per-stream kernel computes them."""
    from acc_amd._abi import Handle
    from acc_amd.fuser import fuse
    from acc_amd.instrument import calcTransformations
    print(f"\n[launch] {name} compacted vs plain", flush=True)
    instr = prebuilt.instruments()[name]()
    N, ntot = 60000, 4096
    off, rot = calcTransformations(instr)
    res = []
    for compact in (False, True):
        p, k = fuse(instr, compact=compact)
        h = Handle([c.mcv_desc() for c in instr.components], off, rot, p, k)
        try:
            h.set_dump(0, N)
            h.clear()
            assert h.run(N, 5, ntot, 512, MCV_FLAG_GATE_DEAD)["processed"] == N
            s, a = h.read_dump()
            hists = {i: h.read_monitor(i) for i in range(len(instr.components)) if h.nbins[i] > 0}
            res.append((s, a, hists))
        finally:
            h.close()
    (s0, a0, h0), (s1, a1, h1) = res
    np.testing.assert_array_equal(a1, a0)
    np.testing.assert_array_equal(s1, s0)
    assert (a0 > 1).any()  # some neutrons did scatter into several outputs
    for i in h0:
        np.testing.assert_array_equal(h1[i][0], h0[i][0])
        np.testing.assert_allclose(h1[i][1:], h0[i][1:], rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("name", ["c1", "c2s"])
def test_replicated_histograms(name, monkeypatch):
    """Monitors whose adds are spread over replicas (and folded after the
    launch) give the histograms of unreplicated adds, on every run of a
    handle (the fold leaves the replicas zeroed for the next one).  The LDS
    bin cache is off: monitors behind it are not replicated."""
    from acc_amd._abi import Handle
    from acc_amd.fuser import fuse
    from acc_amd.instrument import calcTransformations
    print(f"\n[launch] {name} replicated histograms", flush=True)
    instr = prebuilt.instruments()[name]()
    off, rot = calcTransformations(instr)
    p, k = fuse(instr)
    N, ntot = 400000, 20000
    res = {}
    monkeypatch.setenv("MCV_LDS_CACHE_SLOTS", "0")
    for reps in ("1", "64"):
        monkeypatch.setenv("MCV_HIST_REPLICAS", reps)
        h = Handle([c.mcv_desc() for c in instr.components], off, rot, p, k)
        try:
            runs = []
            for seed in (1, 2):
                h.clear()
                assert h.run(N, seed, ntot, 512, MCV_FLAG_GATE_DEAD)["processed"] == N
                runs.append({i: h.read_monitor(i) for i in range(len(instr.components)) if h.nbins[i] > 0})
            res[reps] = runs
        finally:
            h.close()
    for a, b in zip(res["1"], res["64"]):
        for i in a:
            assert a[i][0].sum() > 0
            np.testing.assert_array_equal(b[i][0], a[i][0])
            np.testing.assert_allclose(b[i][1:], a[i][1:], rtol=1e-12, atol=1e-300)
