"""Cylindrical array of vertical detector tubes (no counterpart in the reference).

Tubes stand vertically (along local y) on a cylinder of `radius` around the
component origin; tube j covers the azimuth phi = atan2(x, z) in
[phi_min + j dphi, phi_min + (j+1) dphi) (degrees), each split into `npixels`
along y over `height`, and each pixel histograms time of flight in `ntof`
bins over [tof_min, tof_max) seconds.  Like the reference monitors
(acc/components/monitors/psd_monitor.py:66-83) it does not change the neutron
and accumulates N, sum p, sum p^2 -- here with FP64 atomics straight into HBM,
since the (3, ntubes*npixels*ntof) histogram is larger than any cache.
Device routine: mcv::tube_bin.
"""
import numpy as np

from ..._abi import MCV_DET_TUBES
from ..ComponentBase import DeviceFunction
from .MonitorBase import MonitorBase as base

category = 'monitors'


class TubeDetectorArray(base):

    mcv_type = MCV_DET_TUBES

    def __init__(self, name, radius=3.0, height=1.28, phi_min=5.0, phi_max=135.0,
                 tof_min=2.5e-3, tof_max=6.0e-3, ntubes=400, npixels=128, ntof=500,
                 tube_diameter=0.0254, filename="detector.h5"):
        self.name = name
        self.filename = filename
        assert radius > 0 and height > 0 and phi_min < phi_max and tof_min < tof_max
        self.ntubes, self.npixels, self.ntof = int(ntubes), int(npixels), int(ntof)
        self.out = np.zeros((3, self.ntubes, self.npixels, self.ntof))
        self.out_N, self.out_p, self.out_p2 = self.out[0], self.out[1], self.out[2]
        dphi = (phi_max - phi_min) / ntubes
        self.phi_centers = phi_min + dphi * (np.arange(ntubes) + 0.5)
        dy = height / npixels
        self.y_centers = -height / 2 + dy * (np.arange(npixels) + 0.5)
        dt = (tof_max - tof_min) / ntof
        self.tof_centers = tof_min + dt * (np.arange(ntof) + 0.5)
        self.propagate_params = (
            np.array([radius, height, phi_min, phi_max, tof_min, tof_max, tube_diameter, 0.0]),
            self.ntubes, self.npixels, self.ntof, self.out)

    def getHistogram(self, scale_factor=1.):
        from ...histogram import Histogram
        axes = [('tube_phi', self.phi_centers, 'deg'), ('pixel_y', self.y_centers, 'm'),
                ('tof', self.tof_centers, 's')]
        return Histogram('I(tube,pixel,tof)', axes, data=self.out_p * scale_factor,
                         errors=self.out_p2 * scale_factor * scale_factor)

    propagate = DeviceFunction(
        "{{ int64_t nb; const int64_t b = mcv::tube_bin({n}, P + A.off[{i}], nb);"
        " if (b >= 0) mcv::hist_count(A, {i}, nb, b, {n}.p); }}",
        monitor=True, binning="mcv::tube_bin({n}, P + A.off[{i}], nb)")

    def _mcv_desc(self):
        params, nt, npx, ntof, _ = self.propagate_params
        return dict(type=self.mcv_type, params=params, iparams=np.array([nt, npx, ntof], dtype=np.int64),
                    nbins=nt * npx * ntof)
