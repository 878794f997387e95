"""Buffered-mode kernels: a chain (usually one component) over the rows of a
neutron array, as the reference's `component.process(neutrons)`
(acc/components/ComponentBase.py:79-107) runs a component's process_kernel
(acc/components/StochasticComponentBase.py:67-100)."""
from .fuser import _generate, compile_source


def generate_process_source(instrument, block=None):
    """(source text, outputs per neutron: 0 = rows updated in place, M = a lone
    multiple-scattering sample writing M rows per neutron)."""
    from . import config
    comps = instrument.components
    if not comps:
        raise ValueError("empty chain")
    blk = int(block or config.kernel_block)
    ms = [c for c in comps if c.is_multiplescattering]
    if ms:
        if len(comps) != 1:
            raise NotImplementedError("a multiple-scattering sample is processed on its own")
        c = comps[0]
        M = int(c.max_scattered_neutrons)
        pre = c.propagate.render_preamble(c, 0)
        call = c.propagate.render(c, 0)
        src = f"""// buffered mode, multiple scattering: {type(c).__name__}
#define MCV_BLOCK {blk}
#include "mcv_kernel.h"
#include "mcv_hms.h"

namespace {{
{pre}
struct Chain {{
  __device__ static __forceinline__ void scatter(mcv::Rng& rng, mcv::Neutron& n, const mcv::LaunchArgs& A,
                                                 double* rows) {{
  const double* __restrict__ P = A.P;
  // the sample writes its outputs straight into this neutron's M output rows
  mcv::Neutron* out = reinterpret_cast<mcv::Neutron*>(rows);
  const int num = {call};
  mcv::Neutron zero;
  zero.x = zero.y = zero.z = zero.vx = zero.vy = zero.vz = zero.s1 = zero.s2 = zero.t = zero.p = 0;
  for (int k = num; k < {M}; ++k) mcv::store_row(rows + 10 * k, zero);
  }}
}};
}}  // namespace

MCV_DEFINE_PROCESS_MS_KERNEL(KERNEL_NAME, Chain)
"""
        return src, M
    src, _ = _generate(instrument, compact=False)
    src = src.replace("MCV_DEFINE_KERNELS_PLAIN(KERNEL_NAME, Chain)", "MCV_DEFINE_PROCESS_KERNEL(KERNEL_NAME, Chain)")
    src = src.replace('#include "mcv_kernel.h"', '#define MCV_BLOCK %d\n#include "mcv_kernel.h"' % blk, 1)
    return "// buffered mode\n" + src, 0


def fuse_process(instrument, block=None, verbose=False):
    """(code object path, kernel name, outputs per neutron) of the buffered-mode kernel."""
    src, m = generate_process_source(instrument, block=block)
    path, name = compile_source(src, verbose=verbose)
    return path, name, m
