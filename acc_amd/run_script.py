"""Run an instrument script on the GPU (acc/run_script.py).

`run(script, workdir, ncount, ...)` keeps the reference's signature
(run_script.py:10-35): it loads the script (which defines `instrument`, an
Instrument or a function returning one), fuses the component chain into one
gfx950 kernel, runs ncount neutrons with the reference's stream partition and
seeding, and saves each monitor's histogram scaled by 1/ncount
(saveMonitorOutputs, :291-295).  Differences: the instrument is loaded once
(the reference loads it twice, :211), `seed` is an explicit argument, and
`buffer_size`, which the reference only passes on to the instrument function
(where it is reported as an unrecognized keyword, :316-328), splits the run
into launches of whole streams of about that many neutrons; the streams and
their neutrons are the single launch's, so the results do not change.
"""
import importlib.util
import inspect
import os
import shutil
import warnings

import numpy as np

from . import config
from ._abi import Handle, MCV_FLAG_GATE_DEAD
from .instrument import Instrument, calcTransformations  # noqa: F401  (re-exported)


def _check_workdir(workdir, overwrite_datafiles):
    if os.path.exists(workdir):
        if not overwrite_datafiles:
            raise RuntimeError(f"{workdir} already exists. Specify overwrite_datafiles=True to overwrite")
        shutil.rmtree(workdir)


def run(script, workdir, ncount,
        overwrite_datafiles=True,
        ntotalthreads=None, threads_per_block=None,
        seed=None, buffer_size=None, **kwds):
    """Run an instrument script on one GPU.

    script: path to the instrument script; workdir: output directory;
    ncount: neutron count; buffer_size: neutrons per launch (None: one
    launch).  Extra keywords go to the script's `instrument` function
    (unknown ones are warned about, as in the reference)."""
    _check_workdir(workdir, overwrite_datafiles)
    os.makedirs(workdir)
    gate = kwds.pop("gate_dead_neutrons", config.gate_dead_neutrons)
    instrument = loadInstrument(script, **kwds)
    curdir = os.path.abspath(os.curdir)
    os.chdir(workdir)
    try:
        stats = run_instrument(instrument, ncount, ntotalthreads=ntotalthreads,
                               threads_per_block=threads_per_block, seed=seed, gate_dead_neutrons=gate,
                               buffer_size=buffer_size)
        saveMonitorOutputs(instrument, scale_factor=1.0 / ncount)
    finally:
        os.chdir(curdir)
    return stats


def compile(script, compiled_script=None, **kwds):
    """Fuse and compile an instrument script; returns the code object path
    (the reference returns the generated module path, run_script.py:37-139)."""
    from .fuser import generate_source, compile_source
    kwds.pop("gate_dead_neutrons", None)
    instrument = loadInstrument(os.path.abspath(script), **kwds)
    src, _ = generate_source(instrument)
    path, name = compile_source(src)
    if compiled_script is not None:
        with open(compiled_script, "w") as f:
            f.write(src.replace("KERNEL_NAME", name))
    return path


def get_handle(instrument, block=None, waves_per_eu=None, streams_per_lane=None):
    """The device handle of an instrument for workgroups of `block` threads,
    an optional occupancy target and, for stream-carrying compacted chains,
    `streams_per_lane` (each defaulting to the chain's, fuser.resolve_shape);
    built on first use, then reused."""
    from .fuser import fuse, resolve_shape
    block, wpe, spl = resolve_shape(instrument, block, waves_per_eu, streams_per_lane)
    handles = instrument.__dict__.setdefault("_mcv_handles", {})
    h = handles.get((block, wpe, spl))
    if h is not None:
        return h
    code_object, kernel_name = fuse(instrument, block=block, waves_per_eu=wpe, streams_per_lane=spl)
    offsets, rotmats = calcTransformations(instrument)
    descs = [c.mcv_desc() for c in instrument.components]
    h = Handle(descs, offsets, rotmats, code_object, kernel_name)
    handles[(block, wpe, spl)] = h
    return h


def run_instrument(instrument, ncount, ntotalthreads=None, threads_per_block=None, seed=None,
                   gate_dead_neutrons=None, accumulate=True, buffer_size=None):
    """Run ncount neutrons through an in-memory instrument; monitor `out`
    arrays receive the device histograms (added, as the reference's numba
    round trip of `out` accumulates).  buffer_size: neutrons per launch
    (None or 0: one launch; the results are those of one launch)."""
    from .components.monitors.MonitorBase import MonitorBase
    h = get_handle(instrument)
    seed = config.rng_seed if seed is None else seed
    gate = config.gate_dead_neutrons if gate_dead_neutrons is None else gate_dead_neutrons
    h.clear()
    stats = h.run(int(ncount), seed, ntotalthreads or config.ntotalthreads,
                  threads_per_block or config.threads_per_block, MCV_FLAG_GATE_DEAD if gate else 0,
                  buffer_size=int(buffer_size or 0))
    for i, c in enumerate(instrument.components):
        if isinstance(c, MonitorBase):
            direct = (not accumulate and c.out.dtype == np.float64 and c.out.flags.c_contiguous
                      and c.out.flags.writeable and c.out.size == 3 * h.nbins[i])
            if direct:  # straight into the monitor's array
                h.read_monitor(i, out=c.out)
                continue
            out = h.read_monitor(i).reshape(c.out.shape)
            if accumulate:
                c.out += out
            else:
                c.out[...] = out
    return stats


def saveMonitorOutputs(instrument, scale_factor=1.0):
    from .components.monitors.MonitorBase import MonitorBase
    for comp in instrument.components:
        if isinstance(comp, MonitorBase):
            comp.save(scale_factor=scale_factor)


def _load_module(script):
    spec = importlib.util.spec_from_file_location("mcvinesim", script)
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def loadInstrument(script, **kwds):
    """The instrument an instrument script defines: its `instrument` object,
    or what calling `instrument` returns, given those of `kwds` that its
    signature names (the others are reported as unrecognized)."""
    module = _load_module(script)
    if not hasattr(module, "instrument"):
        raise AssertionError("%s does not define `instrument`" % script)
    target = module.instrument
    if isinstance(target, Instrument):
        return target
    if not callable(target):
        raise AssertionError("`instrument` in %s is neither an Instrument nor a function" % script)
    return target(**_take_accepted(target, kwds))


def _take_accepted(fn, kwds):
    """Removes from `kwds`, and returns, the keywords fn's signature names;
    warns about each one it does not name."""
    accepted = set(inspect.signature(fn).parameters)
    for name in kwds:
        if name not in accepted:
            warnings.warn("Unrecognized kwd: {!r}".format(name))
    return {name: kwds.pop(name) for name in list(kwds) if name in accepted}


def main(argv=None):
    """CLI mirroring acc/run_script.py:330-355."""
    import argparse
    import yaml
    ap = argparse.ArgumentParser(description="run an acc_amd instrument script")
    ap.add_argument("script")
    ap.add_argument("--workdir", default="output")
    ap.add_argument("--ncount", type=float, default=1e6)
    ap.add_argument("--overwrite_datafiles", action="store_true")
    ap.add_argument("--total_threads", type=int, default=int(1e6))
    ap.add_argument("--threads_per_block", type=int, default=512)
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--buffer_size", type=float, default=None)
    ap.add_argument("--additional-kargs", dest="additional_kargs", default=None)
    a = ap.parse_args(argv)
    kwds = yaml.safe_load(open(a.additional_kargs)) if a.additional_kargs else {}
    stats = run(a.script, a.workdir, int(a.ncount), overwrite_datafiles=a.overwrite_datafiles,
                ntotalthreads=a.total_threads, threads_per_block=a.threads_per_block, seed=a.seed,
                buffer_size=int(a.buffer_size) if a.buffer_size else None, **kwds)
    print(stats)
    return 0


if __name__ == '__main__':
    raise SystemExit(main())
