"""Straight / linearly tapered guide with four supermirrors
(acc/components/optics/guide.py:21-162).

The guide is centred on the z axis with its entrance at z=0.  Device routine:
mcv::guide (analytic bounce loop, McStas supermirror reflectivity
acc/components/optics/_guide_utils.py:6-23).
"""
import numpy as np

from ..._abi import MCV_GUIDE
from ..ComponentBase import ComponentBase, DeviceFunction

category = 'optics'

max_bounces = 100000


class Guide(ComponentBase):

    mcv_type = MCV_GUIDE

    def __init__(
            self, name,
            w1, h1, w2, h2, l,
            R0=0.99, Qc=0.0219, alpha=6.07, m=2, W=0.003, **kwargs):
        """
        name (str): the name of this component
        w1, h1 (m): width and height at the guide entry
        w2, h2 (m): width and height at the guide exit
        l (m): length of guide
        R0: low-angle reflectivity
        Qc: critical scattering vector
        alpha: slope of reflectivity
        m: m-value of material (0 is complete absorption)
        W: width of supermirror cutoff
        """
        self.name = name
        ww = .5 * (w2 - w1)
        hh = .5 * (h2 - h1)
        hw1 = 0.5 * w1
        hh1 = 0.5 * h1
        self.propagate_params = (
            float(ww), float(hh), float(hw1), float(hh1), float(l),
            float(R0), float(Qc), float(alpha), float(m), float(W),
        )

    propagate = DeviceFunction(
        "mcv::guide({n}, P + A.off[{i}]);",
        adjoint="mcv::guide_t({n}, P + A.off[{i}], mcv::GuideScore({s}));",
        dparams=lambda comp: [("R0", 5), ("Qc", 6), ("alpha", 7), ("m", 8), ("W", 9)],
        split=("mcv::guide_enter({n}, P + A.off[{i}])", "mcv::guide_bounces({n}, P + A.off[{i}]);",
               "mcv::guide_bounces_t({n}, P + A.off[{i}], mcv::GuideScore({s}));"))

    def _mcv_desc(self):
        return dict(type=self.mcv_type, params=np.array(self.propagate_params, dtype=float))
