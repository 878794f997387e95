"""Fused kernels for chains with a multiple-scattering sample.

The reference's generated kernel calls the multiple scatterer with a per-thread
out_neutrons buffer and then runs every downstream component, with its frame
change, on each output neutron (acc/run_script.py:72-92, template :245-260).
The same structure is emitted here; the dump records output 0's final state and
the number of outputs.  As in the reference, one multiple-scattering component
per chain is supported (run_script.py:73).  When no downstream component
draws random numbers, the downstream components run on each output as the
scatterer produces it, and no output list is kept (every draw and every
output is the reference's; only the order of histogram additions changes).

Compacted form (the default): in a wavefront most neutrons miss the sample
(a hollow cylinder 3 cm across in a 10 cm beam), and the scatterer's bounded
breadth-first loops would run with the few lanes that hit it.  The scatterer's
first step decides a miss without drawing (its neutron passes as the single
output, homogeneous_multiple_scatterer.py:218-220 via is_exiting), so phase A
runs the components before the sample, that test, and for a miss the
downstream components on the one output; a neutron that will interact is
queued with its random stream (stream-carrying compaction: the stream pauses
until the neutron is through), and phase B runs the scatterer and the
downstream components on each output with whole wavefronts of interacting
neutrons.
"""


def _tail(comps, m, n="o"):
    tail = []
    for i in range(m + 1, len(comps)):
        c = comps[i]
        tail.append(f"mcv::apply_transformation({n}, P + A.xoff[{i - 1}]);")
        tail.append(c.propagate.render(c, i, n=n))
        tail.append(f"if (A.gate && !({n}.p > 0)) break;")
    return "\n      ".join(tail)


def _scatter_and_tail(comps, m, checked=False):
    """Statements running the scatterer and, on each output, the downstream
    components; the first output's final state is left in `n`.  When nothing
    downstream draws, each output goes down the chain as the scatterer hands
    it over (no output list: the draw order per stream is the reference's
    either way); otherwise the outputs are kept in a list and run afterwards,
    as the reference does."""
    c = comps[m]
    M = c.max_scattered_neutrons
    tail = _tail(comps, m)
    if not any(d.propagate.stochastic for d in comps[m + 1:]):
        call = c.propagate_emit.format(i=m, n="n", emit="emit", checked="true" if checked else "false",
                                       **c._render_context(m))
        return [f"""int k = 0;
  auto emit = [&](const mcv::Neutron& out) {{
    mcv::Neutron o = out;
    do {{
      if (A.gate && !(o.p > 0)) break;
      {tail}
    }} while (0);
    if (k == 0) n = o;
    ++k;
  }};""",
                f"const int num = {call};",
                "return num;"]
    return [f"mcv::NeutronS out[{M}];  // outputs without spins: they keep the incident neutron's",
            "const double s1 = n.s1, s2 = n.s2;",
            "const int num = " + c.propagate.render(c, m) + ";",
            f"""for (int k = 0; k < num; ++k) {{
    mcv::Neutron o = mcv::load_s(out[k], s1, s2);
    do {{
      if (A.gate && !(o.p > 0)) break;
      {tail}
    }} while (0);
    if (k == 0) n = o;
  }}""",
            "return num;"]


def _generate_compacted(comps, m, pre, structure):
    a = []
    for i in range(m):
        c = comps[i]
        if i > 0:
            a.append(f"mcv::apply_transformation(n, P + A.xoff[{i - 1}]);")
        a.append(c.propagate.render(c, i))
        a.append("if (A.gate && !(n.p > 0)) return 0;")
    if m > 0:
        a.append(f"mcv::apply_transformation(n, P + A.xoff[{m - 1}]);")
    shape = comps[m]._render_context(m)["shape"]
    a.append(f"""if (!mcv::is_exiting<{shape}>(P + A.doff[{m}], n)) return 1;  // interacts: phase B
  {{  // a miss: the neutron is the scatterer's one output
    mcv::Neutron o = n;
    do {{
      if (A.gate && !(o.p > 0)) break;
      {_tail(comps, m)}
    }} while (0);
    n = o;
  }}
  return -1;""")
    b = _scatter_and_tail(comps, m, checked=True)  # phase A queued only neutrons not exiting the shape
    a_text = "\n  ".join(a)
    b_text = "\n  ".join(b)
    src = f"""// fused instrument kernel (multiple scattering), compacted at the sample: {structure}
#include "mcv_kernel.h"
#include "mcv_hms.h"

namespace {{
{pre}
struct Chain {{
  __device__ static __forceinline__ int phase_a(mcv::Rng& rng, mcv::Neutron& n, const mcv::LaunchArgs& A,
                                                 const int64_t nidx) {{
  const double* __restrict__ P = A.P;
  double* __restrict__ H = A.H;
  (void)rng; (void)H; (void)nidx;
  {a_text}
  }}

  __device__ static __forceinline__ int phase_b(mcv::Rng& rng, mcv::Neutron& n, const mcv::LaunchArgs& A,
                                                 const bool skip) {{
  const double* __restrict__ P = A.P;
  double* __restrict__ H = A.H;
  (void)H; (void)skip;
  {b_text}
  }}
}};
}}  // namespace

MCV_DEFINE_KERNELS_COMPACTED_RNG(KERNEL_NAME, Chain)
"""
    return src, structure + " [compacted at the sample, streams carried]"


def generate_ms_source(instrument, compact=True):
    comps = instrument.components
    ms = [i for i, c in enumerate(comps) if c.is_multiplescattering]
    if len(ms) != 1:
        raise NotImplementedError("one multiple-scattering component per instrument is supported")
    m = ms[0]
    structure = " -> ".join(type(c).__name__ for c in comps)
    pre = "".join(c.propagate.render_preamble(c, i) for i, c in enumerate(comps))
    if compact:
        return _generate_compacted(comps, m, pre, structure)
    lines = []
    for i in range(m):
        c = comps[i]
        if i > 0:
            lines.append(f"mcv::apply_transformation(n, P + A.xoff[{i - 1}]);")
        lines.append(c.propagate.render(c, i))
        lines.append("if (A.gate && !(n.p > 0)) return 0;")
    if m > 0:
        lines.append(f"mcv::apply_transformation(n, P + A.xoff[{m - 1}]);")
    lines += _scatter_and_tail(comps, m)
    body = "\n  ".join(lines)
    src = f"""// fused instrument kernel (multiple scattering): {structure}
#include "mcv_kernel.h"
#include "mcv_hms.h"

namespace {{
{pre}
struct Chain {{
  __device__ static __forceinline__ int body(mcv::Rng& rng, mcv::Neutron& n, const mcv::LaunchArgs& A,
                                              const int64_t nidx) {{
  const double* __restrict__ P = A.P;
  double* __restrict__ H = A.H;
  (void)rng; (void)H; (void)nidx;
  {body}
  }}
}};
}}  // namespace

MCV_DEFINE_KERNELS_PLAIN(KERNEL_NAME, Chain)
"""
    return src, structure
