"""PSD_monitor_4Pi: intensity on a sphere around the origin vs (phi, theta)
(acc/components/monitors/psd_monitor_4pi.py:17-79).

Destructive as the reference's: the neutron is moved onto the sphere (its
first crossing ahead), phi = atan2(x, z), theta = asin(y / radius), with the
reference's bin formulas (theta's rounds half up).  A neutron whose line
misses the sphere is not counted (the reference would bin NaN coordinates).
Device routine: mcv::psd4pi_bin.
"""
import math

import numpy as np

from ..._abi import MCV_MON_PSD_4PI
from ..ComponentBase import DeviceFunction
from .MonitorBase import MonitorBase as base

category = 'monitors'


class PSD_monitor_4Pi(base):

    mcv_type = MCV_MON_PSD_4PI

    def __init__(self, name, radius=1., nphi=90, ntheta=90, filename="psd_monitor_4pi.h5"):
        self.name = name
        self.radius = radius
        self.filename = filename
        thetamin, thetamax = -math.pi / 2, math.pi / 2
        phimin, phimax = -math.pi, math.pi
        dphi = (phimax - phimin) / nphi
        self.phi_centers = np.arange(phimin + dphi / 2, phimax, dphi)
        dtheta = (thetamax - thetamin) / ntheta
        self.theta_centers = np.arange(-thetamax + dtheta / 2, thetamax, dtheta)
        self.out = np.zeros((3, nphi, ntheta))
        self.out_N = self.out[0]
        self.out_p = self.out[1]
        self.out_p2 = self.out[2]
        self.propagate_params = (np.array([phimin, phimax, thetamin, thetamax]), radius, nphi, ntheta, self.out)

    def getHistogram(self, scale_factor=1.):
        from ...histogram import Histogram
        axes = [('phi', self.phi_centers, 'radian'), ('theta', self.theta_centers, 'radian')]
        return Histogram('I_thetaphi', axes, data=self.out_p * scale_factor,
                         errors=self.out_p2 * scale_factor * scale_factor)

    propagate = DeviceFunction(
        "{{ int64_t nb; const int64_t b = mcv::psd4pi_bin({n}, P + A.off[{i}], nb);"
        " if (b >= 0) mcv::hist_count(A, {i}, nb, b, {n}.p); }}",
        monitor=True, binning="mcv::psd4pi_bin({n}, P + A.off[{i}], nb)")

    def _mcv_desc(self):
        limits, radius, nphi, ntheta, _ = self.propagate_params
        return dict(type=self.mcv_type, params=np.concatenate([limits, [radius]]),
                    iparams=np.array([nphi, ntheta], dtype=np.int64), nbins=nphi * ntheta)
