"""Runs split into launches: `buffer_size` (run_script.run's keyword) and the
scratch guard split a run into launches over ranges of whole streams of the
one reference partition, each stream seeded as in a single launch (plain
kernels: over slices of every stream, each launch continuing the states the
previous one left), so every neutron's final state and every histogram count
must equal the single launch's (weight sums up to the order of atomic
additions).

Also: seeding a stream range directly, the scratch guard's narrowing and
refusal, and a handle used for buffered processing, the adjoint and
buffered processing again (the adjoint's module load must leave the buffered
state alone)."""
import os

import numpy as np
import pytest

from acc_amd import configs, prebuilt
from acc_amd._abi import MCV_FLAG_GATE_DEAD, McvError, MCV_ERR_ARG, MCV_ERR_NOMEM, rng_states
from acc_amd.run_script import get_handle

pytestmark = pytest.mark.gpu


def _run(h, instr, N, ntot, seed, buffer_size, flags=MCV_FLAG_GATE_DEAD):
    h.set_dump(0, N)
    h.clear()
    st = h.run(N, seed, ntot, 512, flags, buffer_size=buffer_size)
    assert st["processed"] == N
    states, alive = h.read_dump()
    hists = {i: h.read_monitor(i) for i in range(len(instr.components)) if h.nbins[i] > 0}
    return st, states, alive, hists


def _same(a, b):
    st0, s0, a0, h0 = a
    st1, s1, a1, h1 = b
    np.testing.assert_array_equal(a1, a0)
    np.testing.assert_array_equal(s1, s0)
    for i in h0:
        np.testing.assert_array_equal(h1[i][0], h0[i][0])
        np.testing.assert_allclose(h1[i][1:], h0[i][1:], rtol=1e-12, atol=1e-300)


# plain (c1: slices of every stream; the last case's slice is one neutron;
# stored neutrons, whose row is the neutron's index), compacted (c2),
# stream-carrying pooled (c2s; stored neutrons with spins), multiple
# scattering (c3-small)
PLAIN = ("c1", "storage-beam-psd")


@pytest.mark.parametrize("name,N,ntot,buf", [("c1", 400000, 20000, 57000), ("c1", 399999, 20000, 1000),
                                             ("c2", 300000, 20000, 61000), ("c2s", 300000, 20000, 45000),
                                             ("c3-small", 60000, 4096, 9000), ("c2", 300000, 20000, 7),
                                             ("storage-beam-psd", 100003, 4096, 9000),
                                             ("storage-guide-sample", 60000, 4096, 7000)])
def test_chunked_run_equals_one_launch(name, N, ntot, buf):
    print(f"\n[chunks] {name} N={N} ntot={ntot} buffer_size={buf}", flush=True)
    instr = prebuilt.instruments()[name]()
    h = get_handle(instr)
    one = _run(h, instr, N, ntot, 4, 0)
    many = _run(h, instr, N, ntot, 4, buf)
    npt = one[0]["npt"]
    streams = -(-N // npt)
    assert one[0]["nchunks"] == 1
    if name in PLAIN:  # slices of max(1, buf // streams) neutrons of every stream
        per = max(1, buf // streams)
        assert many[0]["nchunks"] == -(-npt // per), many[0]
    else:  # ranges of max(1, buf // npt) whole streams
        per = max(1, buf // npt)
        assert many[0]["nchunks"] == -(-streams // per), many[0]
    _same(one, many)


def test_stream_range_seeding():
    """Streams [first, first + n) seeded directly equal the same streams of a
    seeding from stream 0 (the host jumps to `first`, the device goes on)."""
    full = rng_states(3, 5000)
    for first, n in ((0, 5000), (1, 17), (777, 1000), (4095, 905), (4999, 1)):
        np.testing.assert_array_equal(rng_states(3, n, first=first), full[first:first + n], err_msg=str(first))


@pytest.mark.parametrize("compact", [True, False])
def test_scratch_guard_splits_and_refuses(compact):
    """With a small scratch budget the multiple-scattering chain (the
    reference's 100-neutron lists) runs with the same results: the compacted
    kernel on a narrower persistent grid, the one-lane-per-stream kernel in
    more launches; with no budget the run is refused before any launch."""
    from acc_amd._abi import Handle
    from acc_amd.fuser import fuse
    from acc_amd.instrument import calcTransformations
    from codeobj_tools import private_segment_sizes
    instr = prebuilt.instruments()["hms-ref"]()
    path, kname = fuse(instr, compact=compact)
    off, rot = calcTransformations(instr)
    h = Handle([c.mcv_desc() for c in instr.components], off, rot, path, kname)
    N, ntot = 30000, 2048
    saved = os.environ.get("MCV_SCRATCH_BUDGET_MB")
    try:
        os.environ.pop("MCV_SCRATCH_BUDGET_MB", None)
        one = _run(h, instr, N, ntot, 0, 0)
        per_lane = one[0]["scratch_per_lane"]
        print(f"\n[chunks] hms-ref compact={compact} scratch {per_lane} B per lane", flush=True)
        assert per_lane == private_segment_sizes(path)[kname]  # what HIP reports = the code object's metadata
        # room for a few workgroups at once
        os.environ["MCV_SCRATCH_BUDGET_MB"] = str(max(1, (4 * h_block(path) * per_lane) >> 20))
        few = _run(h, instr, N, ntot, 0, 0)
        if not compact:
            assert few[0]["nchunks"] > 1, few[0]
        _same(one, few)
        os.environ["MCV_SCRATCH_BUDGET_MB"] = "0"
        with pytest.raises(McvError) as e:
            h.run(N, 0, ntot, 512, MCV_FLAG_GATE_DEAD)
        assert e.value.code == MCV_ERR_NOMEM
    finally:
        if saved is None:
            os.environ.pop("MCV_SCRATCH_BUDGET_MB", None)
        else:
            os.environ["MCV_SCRATCH_BUDGET_MB"] = saved
    again = _run(h, instr, N, ntot, 0, 0)
    _same(one, again)
    h.close()


def h_block(path):
    """Workgroup size a fused code object was built for (its source's MCV_BLOCK)."""
    import re
    with open(path[:-len(".hsaco")] + ".hip") as f:
        return int(re.search(r"#define MCV_BLOCK (\d+)", f.read()).group(1))


def test_process_adjoint_process_one_handle():
    """Buffered processing, the adjoint and buffered processing again on one
    handle: the second buffered run equals the first; an adjoint call with the
    wrong score count is refused."""
    from acc_amd.buffered import _handle
    from acc_amd.fuser_adj import fuse_adjoint
    instr = configs.c5()
    h, path, name, m = _handle(instr)
    n = 20000
    arr = np.zeros((n, 10))
    first, _ = h.process(path, name, arr, 3, m)
    h.clear()
    apath, aname, layout = fuse_adjoint(instr)
    nscore = sum(len(dp) for _, _, dp in layout)
    g = np.ones(h.hist_total)
    grad, _, st = h.adjoint(apath, aname, 50000, 0, g, nscore)
    assert st["processed"] == 50000 and np.isfinite(grad).all() and np.any(grad != 0)
    with pytest.raises(McvError) as e:
        h.adjoint(apath, aname, 50000, 0, g, nscore - 1)
    assert e.value.code == MCV_ERR_ARG
    second, _ = h.process(path, name, arr, 3, m)
    np.testing.assert_array_equal(second, first)
    assert (first[:, 9] > 0).sum() > 0
