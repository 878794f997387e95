"""IQE_monitor (acc/components/monitors/iqe_monitor.py:15-174).

Bins (Q, E) of the scattered neutron inside an angular window; position is not
used.  `getHistogram` divides by a Monte-Carlo normalizer (getNormalization,
:115-174) computed on the device with a fixed seed, where the reference uses
unseeded numpy draws.  Device routine: mcv::iqe_bin.
"""
import numpy as np

from ..._abi import MCV_MON_IQE
from ..ComponentBase import DeviceFunction
from .MonitorBase import MonitorBase as base

category = 'monitors'


class IQE_monitor(base):

    mcv_type = MCV_MON_IQE

    def __init__(
            self, name,
            Ei=60.,
            Qmin=0., Qmax=10., nQ=100,
            Emin=-45., Emax=45., nE=90,
            max_angle_in_plane=120, min_angle_in_plane=0,
            max_angle_out_of_plane=30, min_angle_out_of_plane=-30,
            filename="iqe.h5",
    ):
        self.name = name
        self.filename = filename
        self.Ei = Ei
        self.nQ, self.Qmin, self.Qmax = nQ, Qmin, Qmax
        self.nE, self.Emin, self.Emax = nE, Emin, Emax
        dQ = (Qmax - Qmin) / nQ
        self.Q_centers = np.arange(Qmin + dQ / 2, Qmax, dQ)
        dE = (Emax - Emin) / nE
        self.E_centers = np.arange(Emin + dE / 2, Emax, dE)
        shape = 3, nQ, nE
        self.out = np.zeros(shape)
        self.out_N = self.out[0]
        self.out_p = self.out[1]
        self.out_p2 = self.out[2]
        self.propagate_params = (
            np.array([Ei, Qmin, Qmax, Emin, Emax,
                      max_angle_in_plane, min_angle_in_plane,
                      max_angle_out_of_plane, min_angle_out_of_plane,
                      ]),
            nQ, nE, self.out
        )

    def copy(self):
        (Ei, Qmin, Qmax, Emin, Emax,
         max_angle_in_plane, min_angle_in_plane,
         max_angle_out_of_plane, min_angle_out_of_plane,
         ), nQ, nE, out = self.propagate_params
        return self.__class__(
            self.name,
            Ei=Ei,
            Qmin=Qmin, Qmax=Qmax, nQ=nQ,
            Emin=Emin, Emax=Emax, nE=nE,
            max_angle_in_plane=max_angle_in_plane,
            min_angle_in_plane=min_angle_in_plane,
            max_angle_out_of_plane=max_angle_out_of_plane,
            min_angle_out_of_plane=min_angle_out_of_plane,
            filename=self.filename)

    def getHistogram(self, scale_factor=1.):
        h = self._getHistogram(scale_factor)
        n = getNormalization(self, N=None, epsilon=1e-7)
        return h / n

    def _getHistogram(self, scale_factor=1.):
        from ...histogram import Histogram
        axes = [('Q', self.Q_centers, '1./angstrom'), ('E', self.E_centers, 'meV')]
        return Histogram(
            'IQE', axes,
            data=self.out_p * scale_factor,
            errors=self.out_p2 * scale_factor * scale_factor)

    propagate = DeviceFunction(
        "{{ int64_t nb; const int64_t b = mcv::iqe_bin({n}, P + A.off[{i}], nb);"
        " if (b >= 0) mcv::hist_count(A, {i}, nb, b, {n}.p); }}",
        monitor=True, binning="mcv::iqe_bin({n}, P + A.off[{i}], nb)")

    def _mcv_desc(self):
        params, nQ, nE, _ = self.propagate_params
        return dict(type=self.mcv_type, params=params, iparams=np.array([nQ, nE], dtype=np.int64),
                    nbins=nQ * nE)


def getNormalization(monitor, N=None, epsilon=1e-7, seed=0):
    """Monte-Carlo normalizer of the IQE histogram (iqe_monitor.py:115-144):
    N = nQ*nE*1e4 neutrons with E uniform in [Emin, Emax], isotropic final
    directions, speed e2v(Ei - E) and weight vf/vi, binned by a copy of the
    monitor; bins below epsilon are set to 1."""
    from ...normalization import iqe_normalizer
    if N is None:
        N = monitor.nQ * monitor.nE * 10000
    h = iqe_normalizer(monitor, int(N), seed=seed)
    h.I[h.I < epsilon] = 1
    return h
