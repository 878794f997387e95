"""Buffered mode: components applied to arrays of neutrons.

The reference's `component.process(neutrons)` (acc/components/ComponentBase.py
:79-107) copies an (N, 10) neutron buffer to the GPU, runs the component's
process kernel with the partition of N and streams seeded from the configured
seed (acc/components/StochasticComponentBase.py:21-55), copies it back and
keeps the rows with p > 0; monitors add to their `out` histograms.  Here the
component (or a whole chain) is compiled into a buffered-mode kernel
(acc_amd/fuser_proc.py) and run through mcv_instr_process.  A lone
multiple-scattering sample returns the outputs of every input neutron (the
reference's process kernel hands every neutron of a thread the same output
rows, :79-93, so with several neutrons per thread it keeps only the last
one's; here each input has its own M rows).
"""
import hashlib

import numpy as np

from . import config
from ._abi import Handle, MCV_FLAG_GATE_DEAD
from .fuser_proc import fuse_process
from .instrument import Instrument, calcTransformations


def _desc_key(instr):
    h = hashlib.sha1()
    for c in instr.components:
        d = c.mcv_desc()
        for k in ("params", "iparams", "data"):
            h.update(np.ascontiguousarray(d[k]).tobytes())
        h.update(str(int(d["type"])).encode())
    return h.hexdigest()


def _handle(instr):
    """(Handle, code object, kernel name, outputs per neutron), rebuilt when a
    component's parameters changed since the last call."""
    key = _desc_key(instr)
    cached = instr.__dict__.get("_mcv_process")
    if cached is not None and cached[0] == key:
        return cached[1]
    path, name, m = fuse_process(instr)
    off, rot = calcTransformations(instr)
    h = Handle([c.mcv_desc() for c in instr.components], off, rot, path, name)
    instr._mcv_process = (key, (h, path, name, m))
    return h, path, name, m


def process_chain(instrument, neutrons, seed=None, ntotalthreads=None, threads_per_block=None,
                  gate_dead_neutrons=None, keep_dead=False):
    """Run the chain over an (N, 10) array; returns the rows with p > 0 (all
    rows with keep_dead=True) in the last component's frame.  Monitors of the
    chain add to their `out` arrays."""
    from .components.monitors.MonitorBase import MonitorBase
    arr = np.array(neutrons, dtype=np.float64, order="C").reshape(-1, 10)
    h, path, name, m = _handle(instrument)
    seed = config.rng_seed if seed is None else seed
    gate = config.gate_dead_neutrons if gate_dead_neutrons is None else gate_dead_neutrons
    h.clear()
    out, _ = h.process(path, name, arr, seed, m, ntotalthreads or config.ntotalthreads,
                       threads_per_block or config.threads_per_block, MCV_FLAG_GATE_DEAD if gate else 0)
    for i, c in enumerate(instrument.components):
        if isinstance(c, MonitorBase):
            c.out += h.read_monitor(i).reshape(c.out.shape)
    return out if keep_dead else out[out[:, 9] > 0]


def process_component(component, neutrons, seed=None, **kw):
    """`component.process(neutrons)`: the component alone, in its own frame."""
    instr = component.__dict__.get("_mcv_process_instr")
    if instr is None:
        instr = Instrument("process-" + str(getattr(component, "name", "component")))
        instr.append(component, position=(0, 0, 0))
        component._mcv_process_instr = instr
    return process_chain(instr, neutrons, seed=seed, **kw)
