"""Wavelength_monitor (acc/components/monitors/wavelength_monitor.py:15-74).

Propagates to z=0 without writing back, lambda = 2 pi / (v V2K), adds N, p,
p^2.  Device routine: mcv::wavelength_bin.
"""
import numpy as np

from ..._abi import MCV_MON_WAVELENGTH
from ..ComponentBase import DeviceFunction
from .MonitorBase import MonitorBase as base

category = 'monitors'


class Wavelength_monitor(base):

    mcv_type = MCV_MON_WAVELENGTH

    def __init__(
            self, name,
            xmin=0., xmax=0., ymin=0., ymax=0.,
            xwidth=0., yheight=0.,
            Lmin=0., Lmax=10., nchan=200,
            filename="IL.h5",
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
        dL = (Lmax - Lmin) / nchan
        self.L_centers = np.arange(Lmin + dL / 2, Lmax, dL)
        self.out = np.zeros((3, nchan))
        self.out_N = self.out[0]
        self.out_p = self.out[1]
        self.out_p2 = self.out[2]
        self.propagate_params = (
            np.array([xmin, xmax, ymin, ymax, Lmin, Lmax]),
            nchan, self.out
        )

    def getHistogram(self, scale_factor=1.):
        from ...histogram import Histogram
        axes = [('wavelength', self.L_centers, 'angstrom')]
        return Histogram(
            'I(wavelength)', axes,
            data=self.out_p * scale_factor,
            errors=self.out_p2 * scale_factor * scale_factor)

    propagate = DeviceFunction(
        "{{ int64_t nb; const int64_t b = mcv::wavelength_bin({n}, P + A.off[{i}], nb);"
        " if (b >= 0) mcv::hist_count(A, {i}, nb, b, {n}.p); }}",
        monitor=True, binning="mcv::wavelength_bin({n}, P + A.off[{i}], nb)")

    def _mcv_desc(self):
        limits, nchan, _ = self.propagate_params
        return dict(type=self.mcv_type, params=limits, iparams=np.array([nchan], dtype=np.int64), nbins=nchan)
