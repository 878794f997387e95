"""Source replaying neutrons from an .mcv file
(acc/components/sources/neutronfromstorage.py:42-89).

Neutron i of a run is row i % N of the stored array (the reference's
`neutron_index % N`, :78-81); nothing is drawn.  Device routine:
mcv::storage_source.
"""
import numpy as np

from ..._abi import MCV_SRC_STORAGE
from ...neutron_storage import read
from ..ComponentBase import DeviceFunction
from .SourceBase import SourceBase

category = 'sources'


class NeutronFromStorage(SourceBase):

    requires_neutron_index_in_processing = True
    mcv_type = MCV_SRC_STORAGE
    sets_spins = True  # stored rows carry s1, s2 (the fused kernels' queues then keep them)

    def __init__(self, name, path):
        """name: component name; path: neutron file (.mcv)."""
        self.name = name
        neutrons = read(path)
        if neutrons.shape[0] == 0:
            raise RuntimeError("%s holds no neutrons" % path)
        self.propagate_params = (neutrons,)

    propagate = DeviceFunction("mcv::storage_source({n}, P + A.off[{i}], P + A.doff[{i}], nidx);")

    def _mcv_desc(self):
        (neutrons,) = self.propagate_params
        return dict(type=self.mcv_type, params=np.zeros(0), iparams=np.array([neutrons.shape[0]], dtype=np.int64),
                    data=np.ascontiguousarray(neutrons, dtype=np.float64).reshape(-1))
