"""Source_simple (acc/components/sources/source_simple.py:45-161).

A disk (radius != 0) or rectangle at z=0 aimed at a rectangular target at
distance `dist`, with a flat energy or wavelength distribution.  Parameter
validation and `propagate_params` are the reference's (:87-108); the device
routine is mcv::source_simple.
"""
import math
import warnings

import numpy as np

from ..._abi import MCV_SOURCE_SIMPLE
from ..ComponentBase import DeviceFunction
from .SourceBase import SourceBase

category = 'sources'


class Source_simple(SourceBase):

    mcv_type = MCV_SOURCE_SIMPLE

    def __init__(
            self, name,
            radius=0.05, height=0, width=0, dist=10.0,
            xw=0.1, yh=0.1,
            E0=60, dE=10, Lambda0=0, dLambda=0,
            flux=1, gauss=0, N=1, **kwargs
    ):
        self.name = name
        if (radius != 0 and height == 0 and width == 0):
            square = False
            srcArea = math.pi * radius * radius
        elif (radius == 0 and height != 0 and width != 0):
            square = True
            srcArea = width * height
        else:
            msg = f"Source_simple: confused! Both radius ({radius}) and width({width})/height({height}) are specified"
            raise RuntimeError(msg)
        pmul = flux * 1e4 * srcArea / N
        if (srcArea <= 0 or dist < 0 or xw < 0 or yh < 0):
            raise RuntimeError("bad source geometry")
        if (Lambda0 == 0 and dLambda == 0 and (E0 <= 0 or dE < 0 or E0 - dE <= 0)):
            raise RuntimeError("bad energy distribution spec")
        if (E0 == 0 and dE == 0 and (Lambda0 <= 0 or dLambda < 0 or Lambda0 - dLambda <= 0)):
            raise RuntimeError("bad wavelength distribution spec")
        if gauss:
            # accepted and ignored, as in the reference (source_simple.py:82-83, 104-108)
            warnings.warn("Source_simple: gauss is ignored; the distribution is flat")
        wl_distr = Lambda0 != 0
        self.propagate_params = (
            square, width, height, radius,
            wl_distr, Lambda0, dLambda, E0, dE,
            xw, yh, dist, pmul
        )

    propagate = DeviceFunction("mcv::source_simple(rng, {n}, P + A.off[{i}]);", stochastic=True)

    def _mcv_desc(self):
        return dict(type=self.mcv_type, params=np.array([float(v) for v in self.propagate_params]))
