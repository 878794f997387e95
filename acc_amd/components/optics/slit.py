"""Slit at z = 0, rectangular or circular (acc/components/optics/slit.py:17-108).

A neutron that never reaches z = 0 passes unchanged; one below the weight cut
or outside the opening is absorbed (p = -1); the others move to z = 0.
Device routine: mcv::slit.
"""
import numpy as np

from ..._abi import MCV_SLIT
from ..ComponentBase import ComponentBase, DeviceFunction

category = 'optics'


def aperture_params(kind, xmin, xmax, ymin, ymax, radius, cut, width, height):
    """The reference's argument handling for Slit and Beamstop (slit.py:45-60)."""
    if width > 0:
        xmax = width / 2
        xmin = -xmax
    if height > 0:
        ymax = height / 2
        ymin = -ymax
    if not (xmin or xmax or ymin or ymax or radius):
        raise ValueError("a %s must have some extent" % kind)
    return (np.array([xmin, xmax, ymin, ymax, radius * radius, cut], dtype=float),)


class Slit(ComponentBase):

    mcv_type = MCV_SLIT

    def __init__(self, name, xmin=0, xmax=0, ymin=0, ymax=0, radius=0, cut=0, width=0, height=0, **kwargs):
        """xmin..ymax (m): rectangular opening; radius (m): circular opening
        centred on x = y = 0 (used when non-zero); cut: smallest weight that
        passes; width/height (m): centred rectangle."""
        self.name = name
        self.propagate_params = aperture_params("slit", xmin, xmax, ymin, ymax, radius, cut, width, height)

    propagate = DeviceFunction("mcv::slit({n}, P + A.off[{i}]);")

    def _mcv_desc(self):
        return dict(type=self.mcv_type, params=self.propagate_params[0])
