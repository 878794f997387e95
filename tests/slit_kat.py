"""The reference's exhaustive slit/beamstop known-answer set
(tests/components/optics/test_slit_beamstop.py:91-140), restated: neutrons
arriving at every integer (x, y) in [-9, 9]^2 from 2 m upstream with integer
offsets in [-3, 3]^2, weights cycling through 0.2..1.0; the expected output is
each passing neutron at its arrival point with time = its flight time."""
from collections import deque
from math import sqrt

import numpy as np


def cases():
    """(name, component kwargs, is Beamstop, passes(x, y, p)) of the reference's parametrizations."""
    out = []
    for invert in (False, True):
        for cut in np.arange(0.15, 1, 0.2):
            def passes(x, y, p, cut=cut, invert=invert):
                return (abs(x) <= 3.5 and abs(y) <= 2.5) != invert and p >= cut
            out.append(("width_height_cut%.2f_%s" % (cut, invert), dict(width=7, height=5, cut=cut), invert, passes))
        for radius in np.arange(1.5, 5, 1):
            def passes(x, y, p, radius=radius, invert=invert):
                return (x ** 2 + y ** 2 <= radius ** 2) != invert
            out.append(("radius%.1f_%s" % (radius, invert), dict(radius=radius), invert, passes))
        for xmin in (-7.5, 2.5):
            for xmax in (3.5, 9.5):
                for ymin in (-7.5, 2.5):
                    for ymax in (3.5, 9.5):
                        def passes(x, y, p, b=(xmin, xmax, ymin, ymax), invert=invert):
                            return (b[0] <= x <= b[1] and b[2] <= y <= b[3]) != invert and p >= 0.65
                        out.append(("box%g_%g_%g_%g_%s" % (xmin, xmax, ymin, ymax, invert),
                                    dict(xmin=xmin, xmax=xmax, ymin=ymin, ymax=ymax, cut=0.65), invert, passes))
    return out


def neutrons(passes):
    """(incoming (N, 10) rows, expected outgoing rows of the passing ones)."""
    ws = deque(np.arange(0.2, 1.1, 0.2))
    start, expected = [], []
    for xa in range(-9, 10):
        for ya in range(-9, 10):
            for dx in range(-3, 4):
                for dy in range(-3, 4):
                    distance = sqrt(2 + dx ** 2 + dy ** 2)
                    v = [-dx / distance, -dy / distance, 2 / distance]
                    w = ws[0]
                    ws.rotate()
                    start.append([xa + dx, ya + dy, -2] + v + [0, 0, 0, w])
                    if passes(xa, ya, w):
                        expected.append([xa, ya, 0] + v + [0, 0, distance, w])
    return np.array(start, dtype=float), np.array(expected, dtype=float)
