"""Segmented (tapered) guide read from a Guide_tapering segment file
(acc/components/optics/guide_tapering.py:21-291).

The file holds `h1 h2 w1 w2` rows; lines starting with "c" are comments.  As
in the reference loader (:84-115) the last data row is dropped, and the option
string is parsed with str.lstrip("file=") (:45-46).  The per-segment table the
reference derives on the device (prep_inputs_kernel, :117-139) is derived here
on the host with the same float64 expressions.  Device routine:
mcv::guide_tapering.
"""
import numpy as np

from ..._abi import MCV_GUIDE_TAPERING
from ..ComponentBase import ComponentBase, DeviceFunction

category = 'optics'

max_bounces = 100000


def prep_inputs(l_seg, w1, w2, h1, h2):
    """(10, nseg) table: w1, w2, h1, h2, ww, hh, whalf, hhalf, lwhalf, lhhalf."""
    data = np.zeros((10, len(w1)), dtype=np.float64)
    data[0] = w1
    data[1] = w2
    data[2] = h1
    data[3] = h2
    data[4] = 0.5 * (data[1] - data[0])
    data[5] = 0.5 * (data[3] - data[2])
    data[6] = 0.5 * data[0]
    data[7] = 0.5 * data[2]
    data[8] = l_seg * (0.5 * data[0])
    data[9] = l_seg * (0.5 * data[2])
    return data


class Guide(ComponentBase):

    mcv_type = MCV_GUIDE_TAPERING

    def __init__(self, name, option, l, mx, my, R0=0.99, Qcx=0.021, Qcy=0.021,
                 alphax=6.07, alphay=6.07, W=0.003, **kwargs):
        """
        name (str): the name of this component
        option (str): component options, segment file as "file={filename}"
        l (m): total length of guide
        mx: m-value for the vertical (left/right) mirrors; my: for the horizontal ones
        R0: low-angle reflectivity
        Qcx, Qcy: critical scattering vectors
        alphax, alphay: slopes of reflectivity
        W: width of supermirror cutoff
        """
        self.name = name
        self.filename = None
        if option.startswith("file="):
            self.filename = option.lstrip("file=")
        if not self.filename:
            raise RuntimeError("Expected segment file passed to guide")
        self.h1, self.h2, self.w1, self.w2 = self.load_segments()
        self.nseg = len(self.h1)
        self.l_seg = l / self.nseg
        self.data = prep_inputs(self.l_seg, self.w1, self.w2, self.h1, self.h2)
        scalars = np.array([self.l_seg, R0, Qcx, Qcy, alphax, alphay, mx, my, W], dtype=np.float64)
        self.propagate_params = (self.nseg, scalars, self.data)

    def load_segments(self):
        """Loads parameters from file for each guide segment. Returns: [h1, h2, w1, w2]"""
        h1, h2, w1, w2 = [], [], [], []
        with open(self.filename, "r") as file:
            while line := file.readline():
                if line.startswith("c"):
                    continue
                x, y, z, w = line.split(" ", 3)
                h1.append(x)
                h2.append(y)
                w1.append(z)
                w2.append(w)
        h1 = np.asarray(h1[:-1], dtype=np.float64)
        h2 = np.asarray(h2[:-1], dtype=np.float64)
        w1 = np.asarray(w1[:-1], dtype=np.float64)
        w2 = np.asarray(w2[:-1], dtype=np.float64)
        return h1, h2, w1, w2

    propagate = DeviceFunction(
        "mcv::guide_tapering({n}, P + A.off[{i}], P + A.doff[{i}]);",
        adjoint="mcv::guide_tapering_t({n}, P + A.off[{i}], P + A.doff[{i}], mcv::TaperScore({s}));",
        dparams=lambda comp: [("R0", 1), ("Qcx", 2), ("Qcy", 3), ("alphax", 4), ("alphay", 5),
                              ("mx", 6), ("my", 7), ("W", 8)],
        opening=("mcv::taper_admits({n}, P + A.off[{i}], P + A.doff[{i}])", "{n}.p = 0;"))

    def _mcv_desc(self):
        nseg, scalars, data = self.propagate_params
        return dict(type=self.mcv_type, params=scalars, iparams=np.array([nseg], dtype=np.int64),
                    data=np.ascontiguousarray(data).reshape(-1))
