"""The LDS bin cache of plain kernels: a workgroup's events on a large
monitor accumulate in an open-addressed LDS table, the table is written out
sorted by tile of bins at the workgroup's end and a merge kernel adds the
tiles into the histograms.  Whatever the table size (none, a tiny table where
most events miss and go straight to global memory, the default), and whether
the run is one launch or several, the counts must be identical and the weight
sums equal up to the order of additions."""
import numpy as np
import pytest

from acc_amd import prebuilt
from acc_amd._abi import MCV_FLAG_GATE_DEAD

pytestmark = pytest.mark.gpu


def _hists(name, slots, N, ntot, seed, buffer_size, monkeypatch):
    from acc_amd._abi import Handle
    from acc_amd.fuser import fuse
    from acc_amd.instrument import calcTransformations
    monkeypatch.setenv("MCV_LDS_CACHE_SLOTS", slots)
    instr = prebuilt.instruments()[name]()
    off, rot = calcTransformations(instr)
    p, k = fuse(instr)
    h = Handle([c.mcv_desc() for c in instr.components], off, rot, p, k)
    try:
        out = []
        for rep in range(2):  # the second run of a handle reuses its dumps
            h.clear()
            st = h.run(N, seed, ntot, 512, MCV_FLAG_GATE_DEAD, buffer_size=buffer_size)
            assert st["processed"] == N
            out.append({i: h.read_monitor(i) for i in range(len(instr.components)) if h.nbins[i] > 0})
        for i in out[0]:
            np.testing.assert_array_equal(out[1][i][0], out[0][i][0])
        return out[0]
    finally:
        h.close()


# c1: I(Q,E) 120 x 100 (3 tiles of 4096 bins) and PSD 250 x 250 (16 tiles)
@pytest.mark.parametrize("name,N,ntot,buf", [("c1", 2000000, 100000, 0), ("c1", 2000000, 100000, 300000),
                                             ("c1", 50000, 512, 0), ("iqe-normalizer", 400000, 8192, 0)])
def test_cache_sizes_agree(name, N, ntot, buf, monkeypatch):
    print(f"\n[cache] {name} N={N} ntot={ntot} buffer={buf}", flush=True)
    ref = _hists(name, "0", N, ntot, 11, buf, monkeypatch)
    for slots in ("64", "4096"):
        got = _hists(name, slots, N, ntot, 11, buf, monkeypatch)
        for i in ref:
            assert ref[i][0].sum() > 0
            np.testing.assert_array_equal(got[i][0], ref[i][0], err_msg=f"{name} monitor {i} slots {slots}")
            np.testing.assert_allclose(got[i][1:], ref[i][1:], rtol=1e-11, atol=1e-300)
