"""Scattering kernels of homogeneous samples (acc/kernels/__init__.py).

The reference bakes a kernel's coefficients into JIT constants
(ScatterFuncFactory, :8-145); here each kernel packs into the sample's
parameter block (include/mcv_acc.h, MCV_SAMPLE_HSS/HMS) so the values stay
runtime data -- needed to differentiate with respect to them and to scan them
in batched runs.

 * IsotropicKernel: p *= sigma; uniform direction on 4 pi, |v| kept
   (acc/kernels/isotropic.py:5-21).  sigma, mu are the scattering and
   absorption (at 2200 m/s) coefficients in 1/m; mu is scaled by 2200/|v|
   (acc/kernels/__init__.py:149-161).
 * SQEGridKernel (no reference counterpart): S(Q,E) tabulated on an nQ x nE
   grid of node values at Qmin + i dQ, Emin + j dE (dQ = (Qmax-Qmin)/nQ), with
   the bilinear lookup and index clamping of the reference's only grid kernel
   (acc/kernels/SANS2D_ongrid.py:16-56).  Q and E are drawn uniformly and the
   kinematics closed as in acc/kernels/E_Q.py:37-99; weight
   sigma S(Q,E) (vf/vi) Q dQ dE / (2 ki kf).
 * PhononDOSKernel (no reference counterpart): same sampling, with
   S(Q,E) = Q^2 sum_k A_k exp(-(E - c_k)^2 / (2 w_k^2)) for three peaks.
"""
import numpy as np

from .._abi import MCV_KERNEL_ISOTROPIC, MCV_KERNEL_SQE_GRID, MCV_KERNEL_PHONON_DOS

N_KERNEL_PARAMS = 13


class Kernel:
    kernel_type = None
    absorption_coefficient = 0.0
    scattering_coefficient = 0.0

    def kernel_params(self):
        return np.zeros(N_KERNEL_PARAMS)

    def kernel_data(self):
        return np.zeros(0)

    def kernel_dims(self):
        return 0, 0


class IsotropicKernel(Kernel):
    kernel_type = MCV_KERNEL_ISOTROPIC

    def __init__(self, absorption_coefficient, scattering_coefficient):
        self.absorption_coefficient = float(absorption_coefficient)
        self.scattering_coefficient = float(scattering_coefficient)

    def identify(self, visitor):
        return visitor.onIsotropicKernel(self)


class SQEGridKernel(Kernel):
    kernel_type = MCV_KERNEL_SQE_GRID

    def __init__(self, S, Qmin, Qmax, Emin, Emax, absorption_coefficient, scattering_coefficient):
        self.S = np.ascontiguousarray(S, dtype=np.float64)
        if self.S.ndim != 2 or min(self.S.shape) < 2:
            raise ValueError("S(Q,E) grid must be 2D with at least 2 x 2 nodes")
        self.Qmin, self.Qmax, self.Emin, self.Emax = map(float, (Qmin, Qmax, Emin, Emax))
        self.absorption_coefficient = float(absorption_coefficient)
        self.scattering_coefficient = float(scattering_coefficient)

    def kernel_params(self):
        p = np.zeros(N_KERNEL_PARAMS)
        p[:4] = self.Qmin, self.Qmax, self.Emin, self.Emax
        return p

    def kernel_data(self):
        return self.S.reshape(-1)

    def kernel_dims(self):
        return self.S.shape

    def identify(self, visitor):
        return visitor.onSQEGridKernel(self)


class PhononDOSKernel(Kernel):
    kernel_type = MCV_KERNEL_PHONON_DOS

    def __init__(self, amplitudes, centers, widths, Qmin, Qmax, Emin, Emax,
                 absorption_coefficient, scattering_coefficient):
        self.amplitudes = np.asarray(amplitudes, dtype=float)
        self.centers = np.asarray(centers, dtype=float)
        self.widths = np.asarray(widths, dtype=float)
        if not (self.amplitudes.size == self.centers.size == self.widths.size == 3):
            raise ValueError("PhononDOSKernel takes exactly three peaks")
        self.Qmin, self.Qmax, self.Emin, self.Emax = map(float, (Qmin, Qmax, Emin, Emax))
        self.absorption_coefficient = float(absorption_coefficient)
        self.scattering_coefficient = float(scattering_coefficient)

    def kernel_params(self):
        p = np.zeros(N_KERNEL_PARAMS)
        p[:4] = self.Qmin, self.Qmax, self.Emin, self.Emax
        for k in range(3):
            p[4 + 3 * k: 7 + 3 * k] = self.amplitudes[k], self.centers[k], self.widths[k]
        return p

    def S(self, Q, E):
        g = sum(A * np.exp(-0.5 * ((E - c) / w) ** 2)
                for A, c, w in zip(self.amplitudes, self.centers, self.widths))
        return Q * Q * g

    def identify(self, visitor):
        return visitor.onPhononDOSKernel(self)
