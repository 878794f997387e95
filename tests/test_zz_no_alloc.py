"""mcv_instr_run_noalloc: the run entry point that allocates no device
memory (stream states and counters in the library's static per-device
buffer) gives the results of mcv_instr_run, call after call, and refuses a
launch larger than its buffer."""
import numpy as np
import pytest

from acc_amd import prebuilt
from acc_amd._abi import Handle, MCV_ERR_ARG, MCV_FLAG_GATE_DEAD, McvError
from acc_amd.fuser import fuse
from acc_amd.instrument import calcTransformations

pytestmark = pytest.mark.gpu


def _handle(instr):
    path, name = fuse(instr)
    off, rot = calcTransformations(instr)
    return Handle([c.mcv_desc() for c in instr.components], off, rot, path, name)


def _run(h, instr, fn, N, ntot, seed, buffer_size=0):
    h.set_dump(0, min(N, 200000))
    h.clear()
    st = fn(N, seed, ntot, 512, MCV_FLAG_GATE_DEAD, buffer_size=buffer_size)
    assert st["processed"] == N
    states, alive = h.read_dump()
    hists = {i: h.read_monitor(i) for i in range(len(instr.components)) if h.nbins[i] > 0}
    return states, alive, hists


def _same(a, b):
    s0, a0, h0 = a
    s1, a1, h1 = b
    np.testing.assert_array_equal(a1, a0)
    np.testing.assert_array_equal(s1, s0)
    for i in h0:
        assert h0[i][0].sum() > 0
        np.testing.assert_array_equal(h1[i][0], h0[i][0])
        np.testing.assert_allclose(h1[i][1:], h0[i][1:], rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("name", ["c2", "c2s", "c1", "c3-small"])
def test_noalloc_matches_run_twice(name):
    print(f"\n[noalloc] {name}", flush=True)
    instr = prebuilt.instruments()[name]()
    ref, nal = _handle(instr), _handle(instr)
    try:
        for N, ntot, seed, buf in ((600000, 100000, 2, 0), (90001, 20000, 9, 0), (200000, 20000, 4, 33000)):
            print(f"[noalloc] {name} N={N} ntot={ntot} seed={seed} buffer_size={buf}", flush=True)
            a = _run(ref, instr, ref.run, N, ntot, seed, buf)
            b = _run(nal, instr, nal.run_noalloc, N, ntot, seed, buf)
            _same(a, b)
    finally:
        ref.close()
        nal.close()


def _free_device_memory():
    """hipMemGetInfo's free bytes, from the HIP runtime the engine library uses."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_noalloc_allocates_nothing_and_refuses_oversized_launches():
    instr = prebuilt.instruments()["c2"]()
    warm = _handle(instr)
    warm.run(100000, 1, 0, 0, MCV_FLAG_GATE_DEAD)  # the runtime's own first-launch set-up
    h = _handle(instr)
    try:
        free0 = _free_device_memory()
        # the default partition: 1,000,448 streams (16 MB of states that mcv_instr_run would allocate)
        st = h.run_noalloc(int(2e7), 5, 0, 0, MCV_FLAG_GATE_DEAD)
        free1 = _free_device_memory()
        print(f"\n[noalloc] free memory {free0} -> {free1}", flush=True)
        assert st["processed"] == int(2e7) and st["nthreads"] == 1000448
        assert free0 - free1 < 8 << 20
        with pytest.raises(McvError) as e:
            h.run_noalloc(int(4e6), 5, 2 ** 21, 512, MCV_FLAG_GATE_DEAD)
        assert e.value.code == MCV_ERR_ARG and "scratch buffer" in str(e.value)
        st = h.run_noalloc(int(4e6), 5, 2 ** 21, 512, MCV_FLAG_GATE_DEAD, buffer_size=2 ** 20)
        assert st["processed"] == int(4e6) and st["nchunks"] == 4
    finally:
        warm.close()
        h.close()
