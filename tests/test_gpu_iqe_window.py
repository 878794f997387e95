"""The I(Q,E) monitor's angular window on the GPU against the oracle.

The device decides the window (acc/components/monitors/iqe_monitor.py:80-113:
in-plane angle atan2(vx, vz), out-of-plane angle 90 -+ atan(r / vy), strict
bounds in degrees) from pseudo-angles and cross-product signs, and falls back
to the reference's atan2 / atan for directions within 1e-12 of a bound.  Here
directions are placed on purpose at 1e-13 to 1e-6 rad from every bound of
several windows (both fallback and fast decisions near the edges), along the
+-180 degree seam, in the out-of-plane plane (vy = 0), straight up (vx = vz =
0) and with vx = +-0 behind the monitor, plus directions spread over the whole
sphere; the monitor's count histogram must equal the oracle's exactly."""
import numpy as np
import pytest

import pyoracle
from acc_amd import prebuilt
from acc_amd.buffered import process_chain
from acc_amd.components.monitors.iqe_monitor import IQE_monitor
from acc_amd.instrument import calcTransformations

pytestmark = pytest.mark.gpu

SPEED = 3000.0  # m/s: E = 60 - 47 meV inside the energy window, Q below 12 / AA

WINDOWS = [
    (120, 0, 30, -30),        # the reference's defaults
    (180, -180, 90, -90),     # everything
    (200, -45, 60, 5),        # a bound beyond the range, an out-of-plane band
    (170, 10, 90, 0),
    (-10, -170, 45, 20),
    (0.5, -0.5, 1e-3, 0),     # narrow
]


def _dirs(phis, els):
    """Unit directions (vx, vy, vz) x SPEED for in-plane angles phi and elevations el (radians)."""
    phis, els = np.broadcast_arrays(np.asarray(phis, float), np.asarray(els, float))
    c = np.cos(els)
    return np.stack([SPEED * c * np.sin(phis), SPEED * np.sin(els), SPEED * c * np.cos(phis)], axis=1)


def _neutrons(window, seed):
    max_in, min_in, max_out, min_out = window
    rng = np.random.default_rng(seed)
    offs = np.array([1e-13, 3e-12, 1e-11, 1e-9, 1e-6])
    offs = np.concatenate([-offs, offs])
    v = []
    # spread over the sphere
    u = rng.uniform(-1, 1, 20000)
    v.append(_dirs(rng.uniform(-np.pi, np.pi, 20000), np.arcsin(u)))
    # near each in-plane bound, at elevations inside and outside the band
    for b in (min_in, max_in):
        if -180 < b < 180:
            phi = np.deg2rad(b) + offs
            for el in np.deg2rad([0.0, 0.5 * (min_out + max_out) if max_out > min_out else 0.0, 10.0, -10.0]):
                v.append(_dirs(phi, el))
    # near each out-of-plane bound (elevation el and -el: the reference's angle is |el|)
    for b in (min_out, max_out):
        if 0 <= b <= 90:
            el = np.deg2rad(b) + offs
            el = el[(el >= 0) & (el <= np.pi / 2)]
            for phi in np.deg2rad([0.5 * (min_in + max_in), 1.0, -1.0, 90.0, 179.0]):
                v.append(_dirs(phi, el))
                v.append(_dirs(phi, -el))
    # the +-180 seam, vy = 0, straight up / down, vx = +-0 behind and in front
    seam = np.pi - np.array([0.0, 1e-15, 1e-13, 1e-9])
    v.append(_dirs(seam, 0.0))
    v.append(_dirs(-seam, 0.0))
    v.append(_dirs(rng.uniform(-np.pi, np.pi, 200), 0.0))
    v.append(np.array([[0.0, SPEED, 0.0], [0.0, -SPEED, 0.0], [0.0, 0.0, -SPEED], [-0.0, 0.0, -SPEED],
                       [0.0, 0.0, SPEED], [-0.0, 0.0, SPEED], [0.0, 1.0, -SPEED], [-0.0, -1.0, -SPEED]]))
    vel = np.concatenate(v)
    arr = np.zeros((len(vel), 10))
    arr[:, 3:6] = vel
    arr[:, 9] = rng.uniform(0.5, 1.5, len(vel))
    return arr


@pytest.mark.parametrize("window", WINDOWS, ids=lambda w: "in%s..%s_out%s..%s" % (w[1], w[0], w[3], w[2]))
def test_iqe_window_matches_oracle(window):
    max_in, min_in, max_out, min_out = window
    mon = IQE_monitor('iqe', Ei=60., Qmin=0., Qmax=12., nQ=24, Emin=-50., Emax=50., nE=20,
                      max_angle_in_plane=max_in, min_angle_in_plane=min_in,
                      max_angle_out_of_plane=max_out, min_angle_out_of_plane=min_out)
    instr = prebuilt.single(mon)
    arr = _neutrons(window, seed=11)
    print("\n[iqe-window] %s: %d neutrons" % (window, len(arr)), flush=True)
    mon.out[...] = 0
    process_chain(instr, arr, seed=1)
    descs = [c.mcv_desc() for c in instr.components]
    off, rot = calcTransformations(instr)
    _, hists, processed = pyoracle.process(descs, off, rot, arr, 1, gate=True)
    assert processed == len(arr)
    got = mon.out.reshape(3, -1)
    assert got[0].sum() > 0
    np.testing.assert_array_equal(got[0], hists[0][0])
    np.testing.assert_allclose(got[1], hists[0][1], rtol=1e-12, atol=0)
