"""The oracle's conditioning probe: nudging every transcendental by one ulp
changes states only at rounding level for almost every neutron, and leaves the
unnudged oracle bit-identical when switched off."""
import numpy as np

import pyoracle
from acc_amd import prebuilt
from acc_amd.instrument import calcTransformations


def test_nudge_is_ulp_level_and_reversible():
    instr = prebuilt.instruments()["c1"]()
    descs = [c.mcv_desc() for c in instr.components]
    off, rot = calcTransformations(instr)
    N = 50000
    _, _, o, oa = pyoracle.run(descs, off, rot, N, 0, ntotalthreads=4096, tpb=512, dump=(0, N))
    try:
        pyoracle.set_trans_nudge(1)
        _, _, on, ona = pyoracle.run(descs, off, rot, N, 0, ntotalthreads=4096, tpb=512, dump=(0, N))
    finally:
        pyoracle.set_trans_nudge(0)
    _, _, o2, _ = pyoracle.run(descs, off, rot, N, 0, ntotalthreads=4096, tpb=512, dump=(0, N))
    assert np.array_equal(o, o2)
    assert np.array_equal(oa, ona)
    rel = np.abs(on - o) / np.maximum(np.abs(o), 1e-3)
    assert np.quantile(rel.max(axis=1), 0.99) < 1e-13
    assert (rel > 0).any()
