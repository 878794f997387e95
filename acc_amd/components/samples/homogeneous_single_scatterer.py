"""Homogeneous single scatterer (acc/components/samples/homogeneous_single_scatterer.py).

`factory(shape, kernel)` returns a sample class, as in the reference (:26-102):
forced single scattering at a point drawn uniformly along the in-shape path,
weight L exp(-(mu+sigma) x) times the kernel's, exit attenuation
exp(-(mu+sigma) v_i t_exit).  Device routine: mcv::hss_propagate<Shape, kind>
with the shape rendered as a C++ CSG type.
"""
from ..ComponentBase import DeviceFunction
from .SampleBase import SampleBase

category = 'samples'


def factory(shape, kernel):
    """
    Usage (as in the reference):
      HSSbase = factory(shape=shape, kernel=kernel)
      class HSS(HSSbase): pass
      sample = HSS('sample')
    """
    class HomogeneousSingleScattererBase(SampleBase):

        def __init__(self, name):
            self.name = name
            self.propagate_params = ()

        propagate = DeviceFunction(
            "mcv::hss_propagate<{shape}, {kind}>(rng, {n}, P + A.off[{i}], P + A.doff[{i}]);",
            stochastic=True, preamble=SampleBase._shape_preamble,
            adjoint="mcv::hss_propagate_t<{shape}, {kind}>(rng, {n}, P + A.off[{i}], P + A.doff[{i}], "
                    "mcv::SampleScore({s}, &{dz}));",
            dparams=SampleBase.differentiable_params)

    HomogeneousSingleScattererBase.shape = shape
    HomogeneousSingleScattererBase.kernel = kernel
    return HomogeneousSingleScattererBase
