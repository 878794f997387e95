"""Monitor base (acc/components/monitors/MonitorBase.py:4-12).

A monitor owns `self.out`, a (3, ...) float64 array of N, sum p, sum p^2 that
the engine fills in place after a run; `getHistogram(scale_factor)` and
`save(scale_factor)` turn it into a histogram file.
"""
import numpy as np

from ..ComponentBase import ComponentBase


class MonitorBase(ComponentBase):

    filename = None  # set in ctor

    def save(self, scale_factor=1.):
        h = self.getHistogram(scale_factor=scale_factor)
        return h.save(self.filename)

    def _set_out_from_device(self, out3):
        """Store a (3, nbins) block read back from the engine into self.out."""
        self.out[...] = np.asarray(out3).reshape(self.out.shape)

    def clear(self):
        self.out[...] = 0
