"""Monte-Carlo normalizer of the I(Q,E) monitor on the device
(acc/components/monitors/iqe_monitor.py:115-174).

The reference shoots N = nQ*nE*1e4 neutrons (E uniform over the monitor's
energy range, isotropic directions, weight vf/vi) at a copy of the monitor
through the buffered engine, drawing from numpy's unseeded generator.  Here
the same neutrons come from a device source (mcv::iqe_normalizer_source)
fused with a copy of the monitor, on seeded xoroshiro128+ streams.
"""
import numpy as np

from ._abi import MCV_SRC_IQE_NORMALIZER
from .components.ComponentBase import ComponentBase, DeviceFunction
from .instrument import Instrument


class IQENormalizerSource(ComponentBase):
    mcv_type = MCV_SRC_IQE_NORMALIZER

    def __init__(self, name, Ei, Emin, Emax):
        self.name = name
        self.propagate_params = (float(Ei), float(Emin), float(Emax))

    propagate = DeviceFunction("mcv::iqe_normalizer_source(rng, {n}, P + A.off[{i}]);", stochastic=True)

    def _mcv_desc(self):
        return dict(type=self.mcv_type, params=np.array(self.propagate_params))


def iqe_normalizer(monitor, N, seed=0):
    """Histogram of N normalizer neutrons binned by a copy of `monitor`."""
    from .run_script import run_instrument
    mcopy = monitor.copy()
    instr = Instrument("iqe-normalizer")
    instr.append(IQENormalizerSource("normalizer", monitor.Ei, monitor.Emin, monitor.Emax), position=(0, 0, 0))
    instr.append(mcopy, position=(0, 0, 0))
    run_instrument(instr, N, seed=seed, gate_dead_neutrons=False)
    return mcopy._getHistogram(1.)
