"""Adjoint (vector-Jacobian product) kernels of a component chain.

The generated kernel re-runs the chain exactly as the forward kernel does
(same device routines, same random streams, one neutron after another per
stream), but each component with weight-carrying parameters runs in its
score-tracking form and each monitor, instead of adding to its histogram, adds
(G1[bin] p + 2 G2[bin] p^2) times the neutron's score to per-lane sums
(mcv_adjoint.h).  The reference has no counterpart; the forward chain it
differentiates is the one of acc/run_script.py:37-139 (templates :165-289).
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


def _monitor_adjoint(c, i, arrays):
    """Adjoint statements of monitor i: its bin's weight times every score array."""
    if c.propagate.binning is None:
        raise NotImplementedError(f"monitor {c.name} has no binning form")
    upd = []
    for name, base, size in arrays:
        for k in range(size):
            upd.append(f"acc[{base + k}] += c * {name}[{k}];")
    return ("{\n    int64_t nb;\n    const int64_t b = %s;\n    if (b >= 0) {\n"
            "      const double c = mcv::adjoint_weight(J.G + A.hoff[%d], nb, b, n.p);\n      %s\n"
            "      mcv::dense_add(J, dz, c);\n    }\n  }"
            % (c.propagate.render_binning(c, i), i, "\n      ".join(upd)))


def _adjoint_lines(comps, lo, hi, slot, arrays, first_transform, ret):
    lines = []
    for i in range(lo, hi):
        c = comps[i]
        if i > 0 and (i > lo or first_transform):
            lines.append(f"mcv::apply_transformation(n, P + A.xoff[{i - 1}]);")
        if c.propagate.monitor:
            lines.append(_monitor_adjoint(c, i, arrays))
        elif i in slot:
            lines.append(c.propagate.render_adjoint(c, i, "s%d" % i))
        else:
            lines.append(c.propagate.render(c, i))
        lines.append("if (A.gate && !(n.p > 0)) return%s;" % (" " + ret if ret else ""))
    return lines


def generate_adjoint_source(instrument, block=None, compact=True, streams_per_lane=None):
    """(source text, score layout) of the adjoint kernel: the stream-carrying
    compacted form (phase-A scores travel with the queued neutron) when the
    chain has a compaction point, else one lane per stream."""
    from . import config
    from .fuser import split_index
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
    blk = int(block or config.kernel_block)
    k = split_index(comps) if compact else None
    if k is not None:
        spl = int(streams_per_lane or config.kernel_streams_per_lane)
        # phase-A score arrays are views into the payload xa (xb in phase B)
        a_slots = [(i, b0, nj) for i, (b0, nj) in sorted(slot.items()) if i <= k]
        nx, xoff = 0, {}
        for i, b0, nj in a_slots:
            xoff[i] = nx
            nx += nj
        arrays_a = [("s%d" % i, slot[i][0], slot[i][1]) for i in sorted(slot)]
        a_views = ["double* s%d = xa + %d;" % (i, xoff[i]) for i, _, _ in a_slots]
        a_decl = ["double s%d[%d] = {%s};" % (i, nj, ", ".join(["0"] * nj))
                  for i, (b0, nj) in sorted(slot.items()) if i <= k and i not in xoff]
        b_views = ["double* s%d = xb + %d;" % (i, xoff[i]) for i, _, _ in a_slots]
        b_decl = ["double s%d[%d] = {%s};" % (i, nj, ", ".join(["0"] * nj))
                  for i, (b0, nj) in sorted(slot.items()) if i > k]
        # scores of phase-B components do not exist yet during phase A
        arrays_pa = [a for a in arrays_a if int(a[0][1:]) <= k]
        la = a_views + a_decl + ["mcv::DenseTrack dz;", "dz.set = false;", "(void)rng; (void)dz; (void)nidx;"]
        lb = b_views + b_decl + ["mcv::DenseTrack dz;", "dz.set = false;", "(void)rng; (void)dz;"]
        c = comps[k]
        if c.propagate.split is not None:
            # phase A ends at component k's entrance when gating (as the
            # forward kernel does); its scores then start in phase B
            la += _adjoint_lines(comps, 0, k, slot, arrays_pa, False, "0")
            if k > 0:
                la.append(f"mcv::apply_transformation(n, P + A.xoff[{k - 1}]);")
            full = (c.propagate.render_adjoint(c, k, "s%d" % k) if k in slot else c.propagate.render(c, k))
            rest = (c.propagate.render_split(c, k, 2, s="s%d" % k) if k in slot
                    else c.propagate.render_split(c, k, 1))
            la.append(f"""if (A.gate) {{
    if (!{c.propagate.render_split(c, k, 0)}) return 0;
  }} else {{
    {full}
  }}""")
            la.append("if (A.gate && !(n.p > 0)) return 0;")
            lb.append(f"if (A.gate) {{ {rest} }}")
            lb.append("if (A.gate && !(n.p > 0)) return;")
        else:
            la += _adjoint_lines(comps, 0, k + 1, slot, arrays_pa, False, "0")
        la.append("return 1;")
        lb += _adjoint_lines(comps, k + 1, len(comps), slot, arrays_a, True, "")
        src = f"""// adjoint of the fused instrument kernel, compacted after component {k}:
// {" -> ".join(type(c).__name__ for c in comps)}
#define MCV_BLOCK {blk}
#define MCV_SPL {spl}
#include "mcv_adjoint.h"

namespace {{
{pre}
struct AdjChain {{
  static constexpr int NS = {ns};
  static constexpr int NX = {max(nx, 1)};
  __device__ static __forceinline__ int phase_a(mcv::Rng& rng, mcv::Neutron& n, const mcv::LaunchArgs& A,
                                                 const mcv::AdjointArgs& J, double* acc, double* xa,
                                                 const int64_t nidx) {{
  const double* __restrict__ P = A.P;
  {chr(10).join("  " + l for l in la).strip()}
  }}

  __device__ static __forceinline__ void phase_b(mcv::Rng& rng, mcv::Neutron& n, const mcv::LaunchArgs& A,
                                                  const mcv::AdjointArgs& J, double* acc, double* xb) {{
  const double* __restrict__ P = A.P;
  {chr(10).join("  " + l for l in lb).strip()}
  }}
}};
}}  // namespace

MCV_DEFINE_ADJOINT_KERNEL_COMPACTED(KERNEL_NAME, AdjChain)
"""
        return src, layout
    arrays = [("s%d" % i, b0, nj) for i, (b0, nj) in sorted(slot.items())]
    lines = list(decl) + ["mcv::DenseTrack dz;", "dz.set = false;", "(void)rng; (void)dz;"]
    lines += _adjoint_lines(comps, 0, len(comps), slot, arrays, False, "")
    body = "\n  ".join(lines)
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


def fuse_adjoint(instrument, block=None, verbose=False, compact=True, streams_per_lane=None):
    """(code object path, kernel name, score layout) of the adjoint kernel."""
    src, layout = generate_adjoint_source(instrument, block=block, compact=compact,
                                          streams_per_lane=streams_per_lane)
    path, name = compile_source(src, verbose=verbose)
    return path, name, layout
