"""Instrument container and geometer.

Stands in for what the reference takes from mcvine-core: `mcvine.instrument()`
and `Instrument.append(component, position, orientation, relativeTo)` (used by
every instrument script, e.g. tests/acc_sgm_instrument.py:8-33), and
`neutron_coordinates_transformer.relativePositionOrientation`, which
acc/run_script.py:141-162 (calcTransformations) calls for each consecutive
pair of components.

Orientation convention: Euler angles in degrees, McStas style (mcni follows
McStas).  The component's axes are the global axes rotated by
M = Rx(rx) Ry(ry) Rz(rz) (active rotations), so coordinates transform with
M^T; with `relativeTo`, position and rotation compose with the reference
component's.  The single-axis cases are pinned by
tests/test_coord_transform_among_comps.py:25-57; the order of composition of
several non-zero angles follows McStas' rot_set_rotation and is not pinned by
any reference test.
"""
import math

import numpy as np


def _exact_trig(deg):
    """cos/sin of an angle in degrees, exact at multiples of 90 degrees."""
    r = math.radians(deg)
    q = deg / 90.0
    if q == int(q):
        k = int(q) % 4
        return [(1.0, 0.0), (0.0, 1.0), (-1.0, 0.0), (0.0, -1.0)][k]
    return math.cos(r), math.sin(r)


def orientation_matrix(angles):
    """Active rotation matrix M = Rx Ry Rz for Euler angles (degrees)."""
    rx, ry, rz = angles
    m = np.eye(3)
    for axis, a in (("x", rx), ("y", ry), ("z", rz)):
        if a == 0:
            continue
        c, s = _exact_trig(a)
        if axis == "x":
            r = np.array([[1, 0, 0], [0, c, -s], [0, s, c]], dtype=float)
        elif axis == "y":
            r = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]], dtype=float)
        else:
            r = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]], dtype=float)
        m = m @ r
    return m


class Geometer:
    """Absolute position and rotation of each registered component."""

    def __init__(self):
        self._pos = {}
        self._rot = {}

    def register(self, comp, position, orientation, relativeTo=None):
        pos = np.asarray(position, dtype=float)
        rot = orientation_matrix(orientation)
        if relativeTo is not None:
            ppos, prot = self._pos[id(relativeTo)], self._rot[id(relativeTo)]
            pos = ppos + prot @ pos
            rot = prot @ rot
        self._pos[id(comp)] = pos
        self._rot[id(comp)] = rot

    def position(self, comp):
        return self._pos[id(comp)]

    def rotation(self, comp):
        return self._rot[id(comp)]


def relativePositionOrientation(pos1, rot1, pos2, rot2):
    """(offset, rotmat) taking coordinates in frame 1 to frame 2:
    r2 = rotmat @ (r1 - offset), v2 = rotmat @ v1 (acc/vec3.py:71-75)."""
    offset = rot1.T @ (pos2 - pos1)
    rotmat = rot2.T @ rot1
    return offset, rotmat


class Instrument:
    """Ordered component chain with a geometer (mcni.Instrument's role)."""

    def __init__(self, name="instrument"):
        self.name = name
        self.components = []
        self.geometer = Geometer()

    def append(self, component, position=(0, 0, 0), orientation=(0, 0, 0), relativeTo=None):
        self.geometer.register(component, position, orientation, relativeTo)
        self.components.append(component)
        return component

    def __repr__(self):
        return "Instrument(%s: %s)" % (self.name, ", ".join(getattr(c, "name", "?") for c in self.components))


def instrument(name="instrument"):
    """`mcvine.instrument()` equivalent."""
    return Instrument(name)


def calcTransformations(instrument):
    """offsets (n-1, 3) and rotmats (n-1, 3, 3) between consecutive components
    (acc/run_script.py:141-162)."""
    comps = instrument.components
    g = instrument.geometer
    offsets, rotmats = [], []
    for i in range(len(comps) - 1):
        o, r = relativePositionOrientation(g.position(comps[i]), g.rotation(comps[i]),
                                           g.position(comps[i + 1]), g.rotation(comps[i + 1]))
        offsets.append(o)
        rotmats.append(r)
    return np.array(offsets, dtype=float).reshape(-1, 3), np.array(rotmats, dtype=float).reshape(-1, 3, 3)
