"""The 64-instance scan as configured (configs.scan(64): Ei log-uniform on
20-200 meV, sample rotation uniform on 0-180 deg, counts log-uniform on
1e5-1e7, numpy default_rng(0)) in one persistent launch:

 * with the counts scaled down 400 times (their spread kept), every
   instance's histograms equal the CPU oracle's for that instance, count and
   seed (counts exact, sums to 1e-9);
 * at the configured counts (1.6e8 neutrons), every instance's histograms
   equal its own one-by-one run's, and every instance processes exactly its
   count."""
import numpy as np
import pytest

import pyoracle
from acc_amd import configs, run_script
from acc_amd.batch import BatchedScan
from acc_amd.instrument import calcTransformations

pytestmark = pytest.mark.gpu

SEED = 3


@pytest.fixture(scope="module")
def scan64():
    print("\n[scan64] building 64 instances", flush=True)
    instruments, counts = configs.scan(64)
    scan = BatchedScan(instruments)
    yield scan, counts
    scan.close()


def _monitors(instr):
    return [i for i, c in enumerate(instr.components) if c.mcv_desc().get("nbins", 0) > 0]


def test_scan64_scaled_counts_match_oracle(scan64):
    scan, counts = scan64
    small = np.maximum(counts // 400, 1)
    ntot = 4096
    print(f"[scan64] scaled counts {small.min()}..{small.max()} (total {small.sum()})", flush=True)
    st = scan.run(small, seed=SEED, ntotalthreads=ntot)
    assert st["processed"] == small.sum()
    for k, instr in enumerate(scan.instruments):
        descs = [c.mcv_desc() for c in instr.components]
        off, rot = calcTransformations(instr)
        hists, processed, _, _ = pyoracle.run(descs, off, rot, int(small[k]), SEED, ntotalthreads=ntot, tpb=512,
                                              gate=True)
        assert processed == small[k]
        for i in _monitors(instr):
            g = scan.monitor(k, i)
            np.testing.assert_array_equal(g[0], hists[i][0], err_msg="instance %d comp %d counts" % (k, i))
            np.testing.assert_allclose(g[1:], hists[i][1:], rtol=1e-9, atol=1e-300,
                                       err_msg="instance %d comp %d sums" % (k, i))
    seen = sum(scan.monitor(k, i)[0].sum() for k in range(scan.ninst) for i in _monitors(scan.instruments[k]))
    assert seen > 0


def test_scan64_full_counts_match_single_runs(scan64):
    scan, counts = scan64
    print(f"[scan64] full counts {counts.min()}..{counts.max()} (total {counts.sum()})", flush=True)
    st = scan.run(counts, seed=SEED)
    assert st["processed"] == counts.sum()
    for k, (e, r) in enumerate(zip(*configs.scan_plan(64)[:2])):
        single = configs.scan_instance(float(e), float(r), "single%d" % k)
        run_script.run_instrument(single, int(counts[k]), seed=SEED, accumulate=False)
        for i in _monitors(single):
            exp = single.components[i].out.reshape(3, -1)
            got = scan.monitor(k, i)
            np.testing.assert_array_equal(got[0], exp[0], err_msg="instance %d comp %d counts" % (k, i))
            np.testing.assert_allclose(got[1:], exp[1:], rtol=1e-9, atol=1e-300,
                                       err_msg="instance %d comp %d sums" % (k, i))
        print(f"[scan64] instance {k}: {counts[k]} neutrons, equal", flush=True)
