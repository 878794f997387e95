"""Adjoint (vector-Jacobian product) kernels of a component chain.

The generated kernel re-runs the chain exactly as the forward kernel does
(same device routines, same random streams, one neutron after another per
stream), but each component with weight-carrying parameters runs in its
score-tracking form and each monitor, instead of adding to its histogram, adds
(G1[bin] p + 2 G2[bin] p^2) times the neutron's score to per-lane sums
(mcv_adjoint.h).  The reference has no counterpart; the forward chain it
differentiates is the one of acc/run_script.py:175-277.
"""
from .fuser import compile_source


def score_layout(instrument):
    """[(component index, component name, [(param name, packed index)])] of
    the chain's differentiable components, in score-slot order."""
    out = []
    for i, c in enumerate(instrument.components):
        dp = c.propagate.differentiable(c)
        if dp:
            out.append((i, c.name, dp))
    return out


def generate_adjoint_source(instrument, block=None):
    """(source text, score layout) of the adjoint kernel."""
    from . import config
    comps = instrument.components
    if not comps:
        raise ValueError("empty instrument")
    if any(c.is_multiplescattering for c in comps):
        raise NotImplementedError("adjoint kernels cover single-scattering chains")
    layout = score_layout(instrument)
    if not layout:
        raise ValueError("the chain has no differentiable parameters")
    from ._abi import MCV_KERNEL_SQE_GRID
    grids = [c for c in comps if getattr(getattr(c, "kernel", None), "kernel_type", None) == MCV_KERNEL_SQE_GRID]
    if len(grids) > 1:
        raise NotImplementedError("at most one S(Q,E)-grid sample per differentiated chain")
    slot = {}
    base = 0
    decl = []
    for i, name, dp in layout:
        slot[i] = (base, len(dp))
        decl.append("double s%d[%d] = {%s};" % (i, len(dp), ", ".join(["0"] * len(dp))))
        base += len(dp)
    ns = base
    pre = "".join(c.propagate.render_preamble(c, i) for i, c in enumerate(comps))
    lines = list(decl) + ["mcv::DenseTrack dz;", "dz.set = false;", "(void)rng; (void)dz;"]
    for i, c in enumerate(comps):
        if i > 0:
            lines.append(f"mcv::apply_transformation(n, P + A.xoff[{i - 1}]);")
        if c.propagate.monitor:
            if c.propagate.binning is None:
                raise NotImplementedError(f"monitor {c.name} has no binning form")
            upd = []
            for j, (b0, nj) in sorted(slot.items()):
                for k in range(nj):
                    upd.append(f"acc[{b0 + k}] += c * s{j}[{k}];")
            lines.append("{\n    int64_t nb;\n    const int64_t b = %s;\n    if (b >= 0) {\n"
                         "      const double c = mcv::adjoint_weight(J.G + A.hoff[%d], nb, b, n.p);\n      %s\n"
                         "      mcv::dense_add(J, dz, c);\n    }\n  }"
                         % (c.propagate.render_binning(c, i), i, "\n      ".join(upd)))
        elif i in slot:
            lines.append(c.propagate.render_adjoint(c, i, "s%d" % i))
        else:
            lines.append(c.propagate.render(c, i))
        lines.append("if (A.gate && !(n.p > 0)) return;")
    body = "\n  ".join(lines)
    blk = int(block or config.kernel_block)
    src = f"""// adjoint of the fused instrument kernel: {" -> ".join(type(c).__name__ for c in comps)}
#define MCV_BLOCK {blk}
#include "mcv_adjoint.h"

namespace {{
{pre}
struct AdjChain {{
  static constexpr int NS = {ns};
  __device__ static __forceinline__ void body(mcv::Rng& rng, mcv::Neutron& n, const mcv::LaunchArgs& A,
                                               const mcv::AdjointArgs& J, double* acc, const int64_t nidx) {{
  const double* __restrict__ P = A.P;
  (void)nidx;
  {body}
  }}
}};
}}  // namespace

MCV_DEFINE_ADJOINT_KERNEL(KERNEL_NAME, AdjChain)
"""
    return src, layout


def fuse_adjoint(instrument, block=None, verbose=False):
    """(code object path, kernel name, score layout) of the adjoint kernel."""
    src, layout = generate_adjoint_source(instrument, block=block)
    path, name = compile_source(src, verbose=verbose)
    return path, name, layout
