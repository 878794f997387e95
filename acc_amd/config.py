"""Launch defaults (acc/config.py:1-4).

`rng_seed` is read at run time, not bound at import as the reference does
(acc/components/sources/SourceBase.py:18), and `run()` also takes an explicit
`seed`.  `gate_dead_neutrons` selects the histogram semantics: True drops a
neutron once its weight p <= 0 (the mcvine engine's, and the reference's
buffered mode, acc/components/ComponentBase.py:100-102); False reproduces the
reference fused kernel, where an absorbed neutron keeps flowing downstream
(acc/run_script.py:182-196, acc/neutron.py:25-35).  `kernel_block` is the
GPU workgroup size the fuser compiles for.
"""
ntotalthreads = int(1e6)
threads_per_block = 512
# workgroup size of the fused kernels (results do not depend on it; the
# reference's threads_per_block only shapes the stream partition)
kernel_block = 256
# occupancy target (waves per SIMD) of the fused kernels; None = compiler's choice
kernel_waves_per_eu = None
# random streams per lane in compacted kernels whose later stages draw
kernel_streams_per_lane = 4
# workgroup size of plain (uncompacted) chains, whose monitors add through the
# LDS bin cache and whose grid holds the resident workgroups (c1, 1e7
# neutrons, 2048 cache slots: 0.96 ms at 256, 0.90 at 512, 1.07 at 1024;
# 1.93 ms without the cache)
plain_kernel_block = 512
# launch shape of chains with a multiple-scattering sample: their scatterer
# needs the whole register file (occupancy 1 unconstrained); a target of 2
# waves per SIMD trades spills for latency hiding (c3, 1e8 neutrons: 78 ms at
# 128 / 2 / 4 streams per lane vs 98 ms at 256 / compiler's choice / 4, when
# the scatterer was out of line; with it inlined, 34.5 ms at 64 / 2 / 4, 36.3
# at 128 / 2 / 4, 37.6 at 256 / 2 / 4, 43.4 at 128 / 3 / 2, 53.5 at 128 / 1 / 4)
ms_kernel_block = 64
ms_kernel_waves_per_eu = 2
# launch shape of the compacted adjoint kernels (their queues carry the score
# payload; on the c5 fit chain 64 x 4 measured fastest of 64/128/256 x 1/2/4)
adjoint_block = 64
adjoint_streams_per_lane = 4
floattype = "float64"
rng_seed = 1
gate_dead_neutrons = True


def get_numpy_floattype():
    import numpy as np
    return getattr(np, floattype)
