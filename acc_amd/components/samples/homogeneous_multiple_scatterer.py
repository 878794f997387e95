"""Homogeneous multiple scatterer
(acc/components/samples/homogeneous_multiple_scatterer.py:22-342).

`factory(shape, kernel, max_scattered_neutrons, max_ms_loops,
max_ms_loops_path1, minimum_neutron_event_probability)` returns a sample
class whose propagate emits up to max_scattered_neutrons output neutrons per
incident neutron (transmitted and scattered branches, bounded breadth-first
loops); the fuser runs every downstream component on each output
(the loop acc/run_script.py:87-92 generates).
"""
from ..._abi import MCV_SAMPLE_HMS
from ..ComponentBase import DeviceFunction
from .SampleBase import SampleBase

category = 'samples'


def factory(
        shape, kernel,
        max_scattered_neutrons=10,
        max_ms_loops=3, max_ms_loops_path1=2,
        minimum_neutron_event_probability=1e-20):
    """
    max_scattered_neutrons: output buffer size per incident neutron
    max_ms_loops: loops of interactM_path1
    max_ms_loops_path1: loops in the first encounter
    """
    # worst case it can be (max_ms_loops_path1+1) ^ max_ms_loops (reference :46-47)
    assert max_scattered_neutrons > max_ms_loops * (max_ms_loops_path1 + 1)

    class HomogeneousMultipleScattererBase(SampleBase):

        mcv_type = MCV_SAMPLE_HMS
        is_multiplescattering = True
        NUM_MULTIPLE_SCATTER = max_scattered_neutrons

        def __init__(self, name):
            self.name = name
            self.propagate_params = ()

        propagate = DeviceFunction(
            "mcv::hms_scatterM<{shape}, {kind}, {M}>(rng, {n}, P + A.off[{i}], P + A.doff[{i}], out)",
            stochastic=True, preamble=SampleBase._shape_preamble)
        # the same scatterer handing each output to the callable {emit} as it
        # is produced (for chains whose downstream components draw nothing);
        # {checked}: true where the kernel has already found that the neutron
        # is not exiting the shape (the scatterer's first test)
        propagate_emit = ("mcv::hms_scatter<{shape}, {kind}, {M}, {checked}>"
                          "(rng, {n}, P + A.off[{i}], P + A.doff[{i}], {emit})")

        def _render_context(self, i):
            ctx = SampleBase._render_context(self, i)
            ctx["M"] = self.max_scattered_neutrons
            return ctx

    C = HomogeneousMultipleScattererBase
    C.shape = shape
    C.kernel = kernel
    C.max_scattered_neutrons = max_scattered_neutrons
    C.max_ms_loops = max_ms_loops
    C.max_ms_loops_path1 = max_ms_loops_path1
    C.minimum_neutron_event_probability = minimum_neutron_event_probability
    return C
