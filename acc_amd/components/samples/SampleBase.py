"""Base of sample components (acc/components/samples/SampleBase.py)."""
import numpy as np

from ..._abi import MCV_SAMPLE_HSS
from ...geometry import encode, cpp_type, max_roots
from ..ComponentBase import StochasticComponentBase


class SampleBase(StochasticComponentBase):
    """A homogeneous sample: a shape and a scattering kernel.

    Packs as MCV_SAMPLE_HSS / MCV_SAMPLE_HMS (include/mcv_acc.h):
    params = mu, sigma, minimum_neutron_event_probability, kernel params[13];
    iparams = n_nodes, kernel_type, max_scattered_neutrons, max_ms_loops,
    max_ms_loops_path1, kernel_data_offset, n0, n1;
    data = CSG node table then the kernel table.
    """

    mcv_type = MCV_SAMPLE_HSS
    shape = None
    kernel = None
    max_scattered_neutrons = 1
    max_ms_loops = 0
    max_ms_loops_path1 = 0
    minimum_neutron_event_probability = 0.0

    def _sample_params(self):
        k = self.kernel
        return np.concatenate([[k.absorption_coefficient, k.scattering_coefficient,
                                self.minimum_neutron_event_probability], k.kernel_params()])

    def _mcv_desc(self):
        nodes = encode(self.shape).reshape(-1)
        kdata = np.asarray(self.kernel.kernel_data(), dtype=float).reshape(-1)
        n0, n1 = self.kernel.kernel_dims()
        iparams = np.array([len(nodes) // 6, self.kernel.kernel_type, self.max_scattered_neutrons,
                            self.max_ms_loops, self.max_ms_loops_path1, len(nodes), n0, n1], dtype=np.int64)
        return dict(type=self.mcv_type, params=self._sample_params(), iparams=iparams,
                    data=np.concatenate([nodes, kdata]))

    @staticmethod
    def differentiable_params(comp):
        """Score slots of a sample in the adjoint kernel: the absorption
        coefficient (at 2200 m/s) and the scattering coefficient, then for the
        phonon-DOS kernel its peaks A_k, c_k, w_k (packed P[7..15])."""
        from ..._abi import MCV_KERNEL_PHONON_DOS
        out = [("absorption_coefficient", 0), ("scattering_coefficient", 1)]
        if comp.kernel.kernel_type == MCV_KERNEL_PHONON_DOS:
            for k in range(3):
                out += [("A%d" % (k + 1), 7 + 3 * k), ("c%d" % (k + 1), 8 + 3 * k), ("w%d" % (k + 1), 9 + 3 * k)]
        return out

    def _render_context(self, i):
        return dict(shape=f"Shape{i}", kind=self.kernel.kernel_type)

    @classmethod
    def _shape_preamble(cls, comp, i):
        return f"using Shape{i} = {cpp_type(comp.shape)};  // max roots {max_roots(comp.shape)}\n"
