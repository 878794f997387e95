"""Histograms agree with the CPU engine within Monte-Carlo error.

The device runs with one seed, the oracle (the CPU engine here) with another,
so the two histograms are statistically independent estimates of the same
intensity.  Over the bins with at least 20 events on both sides,
chi^2 = sum (I1 - I2)^2 / (s1^2 + s2^2) with s^2 = sum p^2 per bin must give
chi^2/dof <= 1.1 (expectation 1, standard deviation sqrt(2/dof)).
"""
import numpy as np
import pytest

import pyoracle
from acc_amd import prebuilt
from acc_amd._abi import MCV_FLAG_GATE_DEAD
from acc_amd.instrument import calcTransformations
from acc_amd.run_script import get_handle

pytestmark = pytest.mark.gpu


def chi2_per_dof(a, b, min_count=20):
    sel = (a[0] >= min_count) & (b[0] >= min_count)
    d = a[1][sel] - b[1][sel]
    v = a[2][sel] + b[2][sel]
    return float((d * d / v).sum() / sel.sum()), int(sel.sum())


@pytest.mark.parametrize("name,N", [("c1", 4000000), ("c2", 4000000), ("c2s", 4000000), ("c3-small", 2000000),
                                    ("c5s", 2000000)])
def test_chi2_against_cpu_engine(name, N):
    print(f"\n[stats] {name} N={N}", flush=True)
    instr = prebuilt.instruments()[name]()
    h = get_handle(instr)
    h.clear()
    h.run(N, 11, 0, 0, MCV_FLAG_GATE_DEAD)
    descs = [c.mcv_desc() for c in instr.components]
    off, rot = calcTransformations(instr)
    hists, processed, _, _ = pyoracle.run(descs, off, rot, N, 12, gate=True)
    assert processed == N
    for i, c in enumerate(instr.components):
        if hists[i] is None:
            continue
        g = h.read_monitor(i)
        x, dof = chi2_per_dof(g, hists[i])
        print(f"[stats] {name}/{c.name}: chi2/dof = {x:.3f} over {dof} bins", flush=True)
        assert dof >= 10
        assert x <= 1.1
