"""Base of source components (acc/components/sources/SourceBase.py).

`process_no_buffer(N)` runs this source alone on the device (the reference's
single-component no-buffer kernel, :22-34 and :62-72) and returns the stats.
"""
from ..ComponentBase import StochasticComponentBase


class SourceBase(StochasticComponentBase):

    def process_no_buffer(self, N, threads_per_block=None, ntotalthreads=None, seed=None):
        from ...instrument import Instrument
        from ...run_script import run_instrument
        instr = Instrument("source-only")
        instr.append(self, position=(0, 0, 0))
        return run_instrument(instr, N, ntotalthreads=ntotalthreads, threads_per_block=threads_per_block,
                              seed=seed)
