"""Instrument fuser: one component chain -> one gfx950 kernel.

The reference text-generates a Python module that inlines every component's
Numba `propagate` into one @cuda.jit kernel, with applyTransformation between
components (acc/run_script.py:37-139, templates :165-289).  This fuser emits the C++ body of the
same chain (each component's device routine, frame changes between them, and a
dead-neutron test when MCV_FLAG_GATE_DEAD is set), wraps it in the per-stream
loop of mcv_kernel.h, and compiles it with hipcc into a code object cached by
the hash of its source.  Only the chain's structure (component kinds, sample
shapes and kernel kinds) enters the source: parameter values, tables and
histogram sizes are device data bound at run time, so one code object serves
any parameter set of that structure.
"""
import glob
import hashlib
import os
import subprocess
import tempfile
import threading

_here = os.path.dirname(os.path.abspath(__file__))
DEVICE_DIR = os.path.join(_here, "csrc", "device")
CACHE_DIR = os.environ.get("MCV_KERNEL_CACHE", os.path.join(_here, "_kcache"))
OFFLOAD_ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: no FMA contraction, so every float64 operation rounds as
# the reference's (Python/Numba source order) does.  The reference computes
# path lengths as differences of far-away intersection times; a fused
# multiply-add there moves the result by ~1e-11 relative for grazing paths,
# while without contraction the kernels match the oracle to ~1e-13.
HIPCC_FLAGS = [
    "--genco", f"--offload-arch={OFFLOAD_ARCH}", "-O3", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# The max-ILP machine scheduler (instruction order only; the results are the
# same): 1.3 % faster on the guide chains (c2 19.96 -> 19.71 ms, c2s 25.2 ->
# 24.9 ms per 1e8), neutral on c1, 1.3 % slower on the multiple-scattering c3,
# whose kernels keep the default scheduler.
ILP_FLAGS = ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]


def compile_flags(src):
    """hipcc flags of a generated kernel source."""
    return HIPCC_FLAGS + ([] if '#include "mcv_hms.h"' in src else ILP_FLAGS)


def _device_headers_digest():
    h = hashlib.sha1()
    for p in sorted(glob.glob(os.path.join(DEVICE_DIR, "*.h"))):
        with open(p, "rb") as f:
            h.update(os.path.basename(p).encode())
            h.update(f.read())
    return h.hexdigest()


def chain_body(components):
    """C++ statements of the per-neutron chain body (plain chains)."""
    lines = []
    for i, comp in enumerate(components):
        if i > 0:
            lines.append(f"mcv::apply_transformation(n, P + A.xoff[{i - 1}]);")
        lines.append(comp.propagate.render(comp, i))
        lines.append("if (A.gate && !(n.p > 0)) return 0;")
    lines.append("return 1;")
    return lines


def split_index(components):
    """Index of the last component of phase A (the per-stream part) of the
    compacted kernel, or None.  Phase A ends at the first component after the
    source (or the first drawing component) that is neither a monitor nor an Arm: the
    first point where neutrons die in numbers (a guide entrance, a sample).
    Compaction needs a phase B holding a component that is not a monitor
    (otherwise there is nothing to run densely)."""
    from .components.optics.arm import Arm
    from .components.sources.SourceBase import SourceBase
    first = [i for i, c in enumerate(components) if c.propagate.stochastic or isinstance(c, SourceBase)]
    if not first:
        return None
    k = None
    for j in range(first[0] + 1, len(components)):
        c = components[j]
        if not c.propagate.monitor and not isinstance(c, Arm):
            k = j
            break
    if k is None or k >= len(components) - 1:
        return None
    if all(c.propagate.monitor or isinstance(c, Arm) for c in components[k + 1:]):
        return None
    return k


def phase_b_draws(components, k):
    """Whether phase B (components after k) draws random numbers: its
    neutrons then carry their stream's state through the queue."""
    return any(c.propagate.stochastic for c in components[k + 1:])


def _phase_lines(components, lo, hi, first_transform):
    lines = []
    for i in range(lo, hi):
        comp = components[i]
        if i > 0 and (i > lo or first_transform):
            lines.append(f"mcv::apply_transformation(n, P + A.xoff[{i - 1}]);")
        lines.append(comp.propagate.render(comp, i))
        lines.append("if (A.gate && !(n.p > 0)) return 0;")
    return lines


def _phase(components, lo, hi, first_transform):
    return "\n  ".join(_phase_lines(components, lo, hi, first_transform) + ["return 1;"])


def resolve_shape(instrument, block=None, waves_per_eu=None, streams_per_lane=None):
    """(workgroup size, occupancy target or None, streams per lane) of the
    fused kernel, the given values first, then the chain's defaults:
    config.ms_kernel_block / ms_kernel_waves_per_eu for chains with a
    multiple-scattering sample, config.plain_kernel_block for chains that are
    not compacted, config.kernel_block / kernel_waves_per_eu otherwise, and
    config.kernel_streams_per_lane."""
    from . import config
    if any(c.is_multiplescattering for c in instrument.components):
        block = block or config.ms_kernel_block
        if waves_per_eu is None:
            waves_per_eu = config.ms_kernel_waves_per_eu
    elif split_index(instrument.components) is None:
        block = block or config.plain_kernel_block
    block = int(block or config.kernel_block)
    wpe = waves_per_eu if waves_per_eu is not None else config.kernel_waves_per_eu
    spl = int(streams_per_lane or config.kernel_streams_per_lane)
    return block, wpe, spl


def generate_source(instrument, compact=True, block=None, waves_per_eu=None, streams_per_lane=None):
    """(source text, structural description) of the fused kernel, for
    workgroups of `block` threads (default config.kernel_block), if given an
    occupancy target of `waves_per_eu` waves per SIMD (default
    config.kernel_waves_per_eu; None leaves it to the compiler), and
    `streams_per_lane` random streams per lane in stream-carrying compacted
    kernels (default config.kernel_streams_per_lane).  Chains with a
    multiple-scattering sample default to config.ms_kernel_block and
    config.ms_kernel_waves_per_eu instead."""
    block, wpe, spl = resolve_shape(instrument, block, waves_per_eu, streams_per_lane)
    if block % 64 or not 64 <= block <= 1024:
        raise ValueError("workgroup size must be a multiple of 64 in [64, 1024], got %d" % block)
    defs = "#define MCV_BLOCK %d\n" % block
    if wpe:
        if not 1 <= int(wpe) <= 8:
            raise ValueError("waves per SIMD must be in [1, 8], got %r" % wpe)
        defs += "#define MCV_WAVES_PER_EU %d\n" % int(wpe)
    src, structure = _generate(instrument, compact)
    if any(getattr(c, "sets_spins", False) for c in instrument.components):
        defs += "#define MCV_SPINS 1\n"  # queued neutrons keep their spins
    if "MCV_DEFINE_KERNELS_PLAIN(" in src:
        defs += "#define MCV_LDS_CACHE 1\n"
    if "MCV_DEFINE_KERNELS_COMPACTED_RNG" in src:
        if not 1 <= spl <= 16:
            raise ValueError("streams per lane must be in [1, 16], got %d" % spl)
        defs += "#define MCV_SPL %d\n" % spl
    return src.replace('#include "mcv_kernel.h"', defs + '#include "mcv_kernel.h"', 1), structure


def _ungated_opening(comps, k):
    """Statements ending phase A of an ungated run for a neutron that
    component k + 1 turns away at its opening: that component's reject (the
    only change it would make), then the rest of the chain, in phase A; an
    admitted neutron falls through to the caller's queueing (phase B applies
    the frame change and runs component k + 1 in full).  Without gating every
    neutron runs through every component, so phase B would otherwise hold
    every neutron, most of which the next component's opening stops at
    once."""
    if k + 1 >= len(comps) or comps[k + 1].propagate.opening is None:
        return ""
    nxt = comps[k + 1]
    tail = _phase_lines(comps, k + 2, len(comps), True)
    return f"""
    {{  // the next component's opening
      mcv::Neutron m = n;
      mcv::apply_transformation(m, P + A.xoff[{k}]);
      if (!{nxt.propagate.render_opening(nxt, k + 1, 0, n="m")}) {{
        n = m;
        {nxt.propagate.render_opening(nxt, k + 1, 1)}
        {chr(10).join("        " + line for line in tail).strip()}
        return -1;
      }}
    }}"""


def _split_phases(comps, k):
    """Phases A and B when component k can stop after its entrance: phase A
    ends with the entrance and phase B starts with the rest of component k,
    so that its loop runs with whole wavefronts of neutrons that entered.
    With gating (dead neutrons dropped) a neutron the entrance absorbs is
    dropped.  Without gating (every neutron runs through every component,
    absorbed or not) it runs on from component k + 1: when that component
    has an opening test and turns it away, in phase A; otherwise it is queued
    with the skip flag (phase_a returns 2), and phase B skips the rest of
    component k for it."""
    c = comps[k]
    enter = c.propagate.render_split(c, k, 0)
    a = _phase_lines(comps, 0, k, False)
    if k > 0:
        a.append(f"mcv::apply_transformation(n, P + A.xoff[{k - 1}]);")
    a.append(f"""if (A.gate) {{
    if (!{enter}) return 0;
  }} else if (!{enter}) {{  // absorbed at the entrance: the rest of the chain runs on it{_ungated_opening(comps, k)}
    return 2;  // phase B skips the rest of component {k}
  }}""")
    a.append("if (A.gate && !(n.p > 0)) return 0;")
    a.append("return 1;")
    b = [f"if (!skip) {{ {c.propagate.render_split(c, k, 1)} }}",
         "if (A.gate && !(n.p > 0)) return 0;",
         _phase(comps, k + 1, len(comps), True)]
    return "\n  ".join(a), "\n  ".join(b)


def _generate(instrument, compact):
    comps = instrument.components
    if not comps:
        raise ValueError("empty instrument")
    from .components.ComponentBase import ComponentBase
    for c in comps:
        assert isinstance(c, ComponentBase), f"{c} is not an acc_amd component"
    if any(c.is_multiplescattering for c in comps):
        from .fuser_ms import generate_ms_source
        return generate_ms_source(instrument, compact=compact)
    structure = " -> ".join(type(c).__name__ for c in comps)
    pre = "".join(c.propagate.render_preamble(c, i) for i, c in enumerate(comps))
    k = split_index(comps) if compact else None
    if k is not None:
        a = _phase(comps, 0, k + 1, False)
        b = _phase(comps, k + 1, len(comps), True)
        if comps[k].propagate.split is not None:
            a, b = _split_phases(comps, k)
        draws = phase_b_draws(comps, k)
        b_sig = "mcv::Rng& rng, mcv::Neutron& n" if draws else "mcv::Neutron& n"
        b_void = "(void)rng; (void)H;" if draws else "(void)H;"
        macro = "MCV_DEFINE_KERNELS_COMPACTED_RNG" if draws else "MCV_DEFINE_KERNELS_COMPACTED"
        src = f"""// fused instrument kernel, compacted after component {k}: {structure}
#include "mcv_kernel.h"

namespace {{
{pre}
struct Chain {{
  __device__ static __forceinline__ int phase_a(mcv::Rng& rng, mcv::Neutron& n, const mcv::LaunchArgs& A,
                                                 const int64_t nidx) {{
  const double* __restrict__ P = A.P;
  double* __restrict__ H = A.H;
  (void)rng; (void)H; (void)nidx;
  {a}
  }}

  __device__ static __forceinline__ int phase_b({b_sig}, const mcv::LaunchArgs& A, const bool skip) {{
  const double* __restrict__ P = A.P;
  double* __restrict__ H = A.H;
  {b_void} (void)skip;
  {b}
  }}
}};
}}  // namespace

{macro}(KERNEL_NAME, Chain)
"""
        tag = " [compacted after %d%s]" % (k, ", streams carried" if draws else "")
        return src, structure + tag
    body = "\n  ".join(chain_body(comps))
    src = f"""// fused instrument kernel: {structure}
#include "mcv_kernel.h"

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


def kernel_key(src):
    h = hashlib.sha1()
    h.update(src.encode())
    h.update(" ".join(compile_flags(src)).encode())
    h.update(_device_headers_digest().encode())
    return h.hexdigest()[:16]


def compile_source(src, verbose=False):
    """Compile a generated chain source; returns (code object path, kernel name)."""
    key = kernel_key(src)
    name = f"mcv_fused_{key}"
    os.makedirs(CACHE_DIR, exist_ok=True)
    out = os.path.join(CACHE_DIR, name + ".hsaco")
    if os.path.exists(out):
        return out, name
    text = src.replace("KERNEL_NAME", name)
    with tempfile.TemporaryDirectory() as td:
        srcpath = os.path.join(td, name + ".hip")
        with open(srcpath, "w") as f:
            f.write(text)
        tmp_out = out + f".tmp{os.getpid()}_{threading.get_ident()}"
        cmd = [HIPCC] + compile_flags(src) + ["-I", DEVICE_DIR, srcpath, "-o", tmp_out]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("hipcc failed for fused kernel %s:\n%s\n%s" % (name, " ".join(cmd), r.stderr[-4000:]))
        os.replace(tmp_out, out)
        with open(os.path.join(CACHE_DIR, name + ".hip"), "w") as f:
            f.write(text)
    if verbose:
        print("compiled", out)
    return out, name


def fuse(instrument, verbose=False, compact=True, block=None, waves_per_eu=None, streams_per_lane=None):
    """Compile (or fetch from the cache) the fused kernel of an instrument."""
    src, _ = generate_source(instrument, compact=compact, block=block, waves_per_eu=waves_per_eu,
                             streams_per_lane=streams_per_lane)
    return compile_source(src, verbose=verbose)
