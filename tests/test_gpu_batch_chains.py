"""Batched scans of each kernel kind: every instance of a batch equals its own
one-by-one run (counts exactly, weight sums up to the order of additions).

The 64-instance scan and test_gpu_batch.py cover the stream-carrying
compacted chain of the scan; here the batch kernel of a plain chain (whose
large monitors go through the LDS bin cache), of a compacted chain that draws
nothing after its culling point, and of a multiple-scattering chain, each
with parameters and neutron counts differing per instance."""
import numpy as np
import pytest

from acc_amd import prebuilt, run_script
from acc_amd.batch import BatchedScan

pytestmark = pytest.mark.gpu

SEED = 11


# (the ungated compacted case queues neutrons the guide's entrance absorbs
# with the flag that makes phase B skip the rest of the guide)
CASES = [("plain", prebuilt.batch_plain, [400000, 0, 123457], True),
         ("compacted", prebuilt.batch_compacted, [300000, 77777, 5], True),
         ("compacted-ungated", prebuilt.batch_compacted, [200000, 3, 99999], False),
         ("multiple-scattering", prebuilt.batch_ms, [200000, 1, 64001], True)]


@pytest.mark.parametrize("name,make,counts,gate", CASES)
def test_batch_instances_equal_single_runs(name, make, counts, gate):
    print(f"\n[batch-chains] {name} counts={counts}", flush=True)
    scan = BatchedScan([make(k) for k in range(len(counts))])
    try:
        st = scan.run(counts, seed=SEED, use_graph=False, gate_dead_neutrons=gate)
        assert st["processed"] == sum(counts)
        for k, n in enumerate(counts):
            single = make(k)
            if n:
                run_script.run_instrument(single, n, seed=SEED, accumulate=False, gate_dead_neutrons=gate)
            h = run_script.get_handle(single)
            for i, nb in enumerate(scan.handle.nbins):
                if nb <= 0:
                    continue
                got = scan.monitor(k, i)
                if not n:
                    assert not got.any(), "instance %d (no neutrons) has counts" % k
                    continue
                exp = h.read_monitor(i)
                np.testing.assert_array_equal(got[0], exp[0], err_msg="%s instance %d comp %d counts" % (name, k, i))
                np.testing.assert_allclose(got[1:], exp[1:], rtol=1e-9, atol=1e-300,
                                           err_msg="%s instance %d comp %d sums" % (name, k, i))
                assert n < 100 or exp[0].sum() > 0, "%s instance %d: monitor %d saw nothing" % (name, k, i)
    finally:
        scan.close()
