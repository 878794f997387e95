"""PosDiv_monitor: intensity vs x and horizontal divergence at z = 0
(acc/components/monitors/posdiv_monitor.py:16-89).

Non-destructive: propagates to z = 0 without writing back, keeps neutrons
strictly inside the window and the divergence range, bins
(x, atan(vx/vz) in degrees).  Device routine: mcv::posdiv_bin.
"""
import numpy as np

from ..._abi import MCV_MON_POSDIV
from ..ComponentBase import DeviceFunction
from .MonitorBase import MonitorBase as base

category = 'monitors'


class PosDiv_monitor(base):

    mcv_type = MCV_MON_POSDIV

    def __init__(self, name, xmin=0., xmax=0., ymin=0., ymax=0., xwidth=0., yheight=0., maxdiv=2.,
                 npos=20., ndiv=20., filename="posdiv.h5"):
        """xwidth/yheight (m) or xmin..ymax: window; maxdiv (deg): divergence
        range +-maxdiv; npos, ndiv: bins."""
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
        npos, ndiv = int(npos), int(ndiv)
        dx = (xmax - xmin) / npos
        self.x_centers = np.arange(xmin + dx / 2, xmax, dx)
        ddiv = 2 * maxdiv / ndiv
        self.div_centers = np.arange(-maxdiv + ddiv / 2, maxdiv, ddiv)
        self.out = np.zeros((3, npos, ndiv))
        self.out_N = self.out[0]
        self.out_p = self.out[1]
        self.out_p2 = self.out[2]
        self.propagate_params = (np.array([xmin, xmax, ymin, ymax, maxdiv]), npos, ndiv, self.out)

    def getHistogram(self, scale_factor=1.):
        from ...histogram import Histogram
        axes = [('x', self.x_centers, 'm'), ('div', self.div_centers, 'deg')]
        return Histogram('Idiv_x', axes, data=self.out_p * scale_factor,
                         errors=self.out_p2 * scale_factor * scale_factor)

    propagate = DeviceFunction(
        "{{ int64_t nb; const int64_t b = mcv::posdiv_bin({n}, P + A.off[{i}], nb);"
        " if (b >= 0) mcv::hist_count(A, {i}, nb, b, {n}.p); }}",
        monitor=True, binning="mcv::posdiv_bin({n}, P + A.off[{i}], nb)")

    def _mcv_desc(self):
        limits, npos, ndiv, _ = self.propagate_params
        return dict(type=self.mcv_type, params=limits, iparams=np.array([npos, ndiv], dtype=np.int64),
                    nbins=npos * ndiv)
