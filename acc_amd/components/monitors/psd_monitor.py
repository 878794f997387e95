"""PSD_monitor (acc/components/monitors/psd_monitor.py:16-83).

Propagates the neutron to z=0 without writing it back, bins (x, y) strictly
inside the window, adds N, p, p^2.  Device routine: mcv::psd_bin.
"""
import math

import numpy as np

from ..._abi import MCV_MON_PSD
from ..ComponentBase import DeviceFunction
from .MonitorBase import MonitorBase as base

category = 'monitors'
RAD2DEG = 180. / math.pi


class PSD_monitor(base):

    mcv_type = MCV_MON_PSD

    def __init__(
            self, name,
            xmin=0., xmax=0., ymin=0., ymax=0.,
            xwidth=0., yheight=0.,
            nx=100, ny=100,
            filename="psd.h5",
            **kwargs
    ):
        self.name = name
        self.filename = filename
        if xwidth > 0:
            xmax = xwidth / 2
            xmin = -xmax
        if yheight > 0:
            ymax = yheight / 2
            ymin = -ymax
        assert xmin < xmax
        assert ymin < ymax
        dx = (xmax - xmin) / nx
        dy = (ymax - ymin) / ny
        self.x_centers = np.arange(xmin + dx / 2, xmax, dx)
        self.y_centers = np.arange(ymin + dy / 2, ymax, dy)
        shape = 3, nx, ny
        self.out = np.zeros(shape)
        self.out_N = self.out[0]
        self.out_p = self.out[1]
        self.out_p2 = self.out[2]
        self.propagate_params = (
            np.array([xmin, xmax, ymin, ymax]),
            nx, ny, self.out
        )

    def getHistogram(self, scale_factor=1.):
        from ...histogram import Histogram
        axes = [('x', self.x_centers, 'm'), ('y', self.y_centers, 'm')]
        return Histogram(
            'Iyx', axes,
            data=self.out_p * scale_factor,
            errors=self.out_p2 * scale_factor * scale_factor)

    propagate = DeviceFunction(
        "{{ int64_t nb; const int64_t b = mcv::psd_bin({n}, P + A.off[{i}], nb);"
        " if (b >= 0) mcv::hist_count(A, {i}, nb, b, {n}.p); }}",
        monitor=True, binning="mcv::psd_bin({n}, P + A.off[{i}], nb)")

    def _mcv_desc(self):
        limits, nx, ny, _ = self.propagate_params
        return dict(type=self.mcv_type, params=limits, iparams=np.array([nx, ny], dtype=np.int64),
                    nbins=nx * ny)
