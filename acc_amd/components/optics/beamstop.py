"""Beamstop at z = 0, rectangular or circular (acc/components/optics/beamstop.py:17-112).

The complement of the slit: neutrons inside the stop (or below the weight
cut) are absorbed (p = -1); the others move to z = 0; a neutron that never
reaches z = 0 passes unchanged.  Device routine: mcv::beamstop.
"""
from ..._abi import MCV_BEAMSTOP
from ..ComponentBase import ComponentBase, DeviceFunction
from .slit import aperture_params

category = 'optics'


class Beamstop(ComponentBase):

    mcv_type = MCV_BEAMSTOP

    def __init__(self, name, xmin=0, xmax=0, ymin=0, ymax=0, radius=0, cut=0, width=0, height=0, **kwargs):
        """Arguments as Slit's: the region given is the one that stops neutrons."""
        self.name = name
        self.propagate_params = aperture_params("beamstop", xmin, xmax, ymin, ymax, radius, cut, width, height)

    propagate = DeviceFunction("mcv::beamstop({n}, P + A.off[{i}]);")

    def _mcv_desc(self):
        return dict(type=self.mcv_type, params=self.propagate_params[0])
