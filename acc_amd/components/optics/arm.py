"""Arm: a component that does nothing to the neutron (acc/components/optics/arm.py:14-43)."""
import numpy as np

from ..._abi import MCV_ARM
from ..ComponentBase import ComponentBase, DeviceFunction

category = 'optics'


class Arm(ComponentBase):

    mcv_type = MCV_ARM

    def __init__(self, name, **kwargs):
        self.name = name
        self.propagate_params = ()

    propagate = DeviceFunction("/* Arm */")

    def _mcv_desc(self):
        return dict(type=self.mcv_type, params=np.zeros(0))
