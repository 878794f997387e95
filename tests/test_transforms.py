"""Frame changes between components: acc_amd.instrument.calcTransformations and
the oracle's applyTransformation against
tests/test_coord_transform_among_comps.py:25-57 of the reference."""
import numpy as np

import pyoracle as orc
from acc_amd.instrument import Instrument, calcTransformations
from acc_amd.components.optics.arm import Arm


def build(kats):
    instr = Instrument()
    comps = []
    for k, spec in enumerate(kats["coord_transforms"]["components"]):
        c = Arm(f"comp{k + 1}")
        rel = comps[spec["relativeTo"]] if spec["relativeTo"] is not None else None
        instr.append(c, position=spec["position"], orientation=spec["orientation"], relativeTo=rel)
        comps.append(c)
    return instr


def test_transform_kats(kats):
    instr = build(kats)
    offsets, rotmats = calcTransformations(instr)
    assert len(offsets) == len(rotmats) == len(instr.components) - 1
    for i, r, v, r_want, v_want in kats["coord_transforms"]["cases"]:
        n = np.zeros(10)
        n[0:3] = r
        n[3:6] = v
        out = orc.transform(n, rotmats[i], offsets[i])
        np.testing.assert_allclose(out[0:3], r_want, atol=1e-12)
        np.testing.assert_allclose(out[3:6], v_want, atol=1e-12)


def test_transform_round_trip():
    # chain of arbitrary orientations: going through every frame and back to
    # the global frame is the identity
    rng = np.random.default_rng(0)
    instr = Instrument()
    for k in range(5):
        instr.append(Arm(f"a{k}"), position=rng.normal(size=3), orientation=rng.uniform(-180, 180, size=3))
    offsets, rotmats = calcTransformations(instr)
    for R in rotmats:
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)
        assert np.isclose(np.linalg.det(R), 1.0)
    g = instr.geometer
    p = rng.normal(size=3)
    local = g.rotation(instr.components[0]).T @ (p - g.position(instr.components[0]))
    for i in range(4):
        local = rotmats[i] @ (local - offsets[i])
    expect = g.rotation(instr.components[4]).T @ (p - g.position(instr.components[4]))
    np.testing.assert_allclose(local, expect, atol=1e-12)
