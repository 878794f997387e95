"""Batched scans: one persistent launch over many instances of one chain gives
each instance the histograms of its own one-by-one run.

Instance k with ncounts[k] neutrons uses the partition and streams of a
separate run with the same seed, so per-bin neutron counts must be equal and
weight sums equal up to the order of atomic additions.  The one-by-one runs
here go through `run_script.run_instrument` (the reference's single-run path,
acc/run_script.py:10-35), which is checked against the oracle by the parity
tests."""
import numpy as np
import pytest

from acc_amd import configs, run_script
from acc_amd.batch import BatchedScan

pytestmark = pytest.mark.gpu

SEED = 5
# a few instances with uneven counts, one of them empty; the last has 3 neutrons
# per stream (2.5e6 neutrons over the default 1e6 streams), so the batch hands
# its work items out first (longest items first)
PLAN = [(25.0, 0.0, 30000), (60.0, 45.0, 250000), (120.0, 90.0, 0), (180.0, 170.0, 1300), (40.0, 10.0, 70001),
        (90.0, 120.0, 2500001)]


def _instances():
    return [configs.scan_instance(e, r, "scan%d" % k) for k, (e, r, _) in enumerate(PLAN)]


def _expected():
    out = []
    for k, (e, r, n) in enumerate(PLAN):
        instr = configs.scan_instance(e, r, "single%d" % k)
        if n:
            run_script.run_instrument(instr, n, seed=SEED, accumulate=False)
            h = run_script.get_handle(instr)
            out.append({i: h.read_monitor(i) for i, c in enumerate(instr.components) if h.nbins[i] > 0})
        else:
            out.append(None)
    return out


def _check(scan, expected):
    for k, exp in enumerate(expected):
        for i, nb in enumerate(scan.handle.nbins):
            if nb <= 0:
                continue
            got = scan.monitor(k, i)
            if exp is None:
                assert not got.any(), "instance %d (no neutrons) has counts" % k
                continue
            np.testing.assert_array_equal(got[0], exp[i][0], err_msg="instance %d comp %d counts" % (k, i))
            np.testing.assert_allclose(got[1:], exp[i][1:], rtol=1e-9, atol=1e-300,
                                       err_msg="instance %d comp %d sums" % (k, i))


@pytest.fixture(scope="module")
def scan_and_expected():
    print("batch: building", flush=True)
    scan = BatchedScan(_instances())
    exp = _expected()
    yield scan, exp
    scan.close()


def test_batch_matches_single_runs(scan_and_expected):
    scan, exp = scan_and_expected
    n = [p[2] for p in PLAN]
    st = scan.run(n, seed=SEED, use_graph=False)
    assert st["processed"] == sum(n)
    _check(scan, exp)


def test_batch_graph_replays(scan_and_expected):
    scan, exp = scan_and_expected
    n = [p[2] for p in PLAN]
    for rep in range(3):  # capture, then two replays
        st = scan.run(n, seed=SEED, use_graph=True, grid=96)
        assert st["processed"] == sum(n)
        _check(scan, exp)


def test_batch_histograms_in_one_copy(scan_and_expected):
    """mcv_batch_histograms_read returns the blocks mcv_batch_monitor_read
    does, and store() fills every instance's monitor out arrays with them."""
    scan, exp = scan_and_expected
    n = [p[2] for p in PLAN]
    scan.run(n, seed=SEED)
    allh = scan.handle.read_histograms()
    nb = scan.handle.nbins
    assert allh.shape == (len(PLAN), 3 * sum(nb))
    for k in range(len(PLAN)):
        off = 0
        for i, b in enumerate(nb):
            if b > 0:
                np.testing.assert_array_equal(allh[k, off:off + 3 * b].reshape(3, -1), scan.monitor(k, i))
            off += 3 * b
    scan.store()
    for k, instr in enumerate(scan.instruments):
        for i, c in enumerate(instr.components):
            if nb[i] > 0:
                np.testing.assert_array_equal(c.out.reshape(3, -1), scan.monitor(k, i))
    # adding: out + h + h is 2h exactly; an out array the runtime cannot write
    # (Fortran order here) goes through a temporary
    c0 = [c for c in scan.instruments[0].components if nb[scan.instruments[0].components.index(c)] > 0][0]
    c0.out = np.asfortranarray(np.zeros(c0.out.shape))
    scan.store(accumulate=True)
    scan.store(accumulate=True)
    for k, instr in enumerate(scan.instruments):
        for i, c in enumerate(instr.components):
            if nb[i] > 0:
                want = scan.monitor(k, i) * (2.0 if c is c0 else 3.0)
                np.testing.assert_array_equal(np.ascontiguousarray(c.out).reshape(3, -1), want)


def test_batch_deliver_rejects_bad_blocks(scan_and_expected):
    """A block reaching past the batch's histograms, or a destination the
    runtime cannot write, is refused before anything is copied."""
    from acc_amd._abi import McvError
    scan, _ = scan_and_expected
    h = scan.handle
    n = h.ninst * 3 * sum(h.nbins)
    dst = np.zeros(16)
    with pytest.raises(McvError, match="outside"):
        h.deliver_histograms([(n - 8, dst)])
    with pytest.raises(McvError, match="outside"):
        h.deliver_histograms([(-1, dst)])
    with pytest.raises(McvError, match="C-contiguous"):
        h.deliver_histograms([(0, np.zeros((4, 8))[:, ::2])])
    assert not dst.any()
    h.deliver_histograms([(8, dst)])  # in range: the 16 doubles from offset 8
    np.testing.assert_array_equal(dst, h.read_histograms().reshape(-1)[8:24])


def test_batch_sequential_matches(scan_and_expected):
    scan, exp = scan_and_expected
    n = [p[2] for p in PLAN]
    st = scan.run_sequential(n, seed=SEED)
    assert st["processed"] == sum(n)
    _check(scan, exp)


def test_batch_new_counts_recapture(scan_and_expected):
    """Changing the counts after a capture re-captures; an all-zero scan runs nothing."""
    scan, _ = scan_and_expected
    n = [1000, 0, 0, 0, 5] + [0] * (len(PLAN) - 5)
    st = scan.run(n, seed=SEED, use_graph=True)
    assert st["processed"] == sum(n)
    st = scan.run([0] * len(PLAN), seed=SEED, use_graph=True)
    assert st["processed"] == 0
    for k in range(len(PLAN)):
        assert not scan.monitor(k, "psd").any()


def test_batch_rejects_different_structure():
    from acc_amd.prebuilt import instruments
    with pytest.raises(ValueError):
        BatchedScan([configs.scan_instance(60.0, 0.0), instruments()["c2"]()])
