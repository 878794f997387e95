"""Batched scans: many instances of one instrument chain in one launch.

The reference runs one instrument per `run_script.run` call
(acc/run_script.py:10-35); a scan over Ei, sample rotation or any other
parameter is a Python loop of separate runs, each a full kernel launch that
leaves the GPU mostly idle when its neutron count is small.  `BatchedScan`
takes instruments that share one chain structure (same component types,
shapes, kernels, table sizes and histogram sizes; any parameter values) and
runs them all in one persistent kernel: workgroups pull (instance, block of
streams) work items from one atomic counter, so small and large instances
share the chip.  A work item is one workgroup's worth of streams, or sixteen
for chains whose later stages draw random numbers (their workgroups refill
stream slots from the item, so the queue is drained once per item).  Items
are handed out longest first (instances with more neutrons per stream
first), so that the launch ends on short items.  Instance k with ncounts[k] neutrons uses exactly the
partition and random streams a separate `run_instrument(instr_k, ncounts[k],
seed=seed)` would, so each instance's histograms are identical (up to the
order of atomic additions) to its one-by-one run.  The launch (stream
seeding, counter reset, batch kernel) is captured once in a hipGraph and
replayed while the workload shape is unchanged.
"""
import numpy as np

from . import config
from ._abi import BatchHandle, MCV_FLAG_GATE_DEAD
from .instrument import calcTransformations


class BatchedScan:
    """Instances of one chain structure, run together on the GPU."""

    def __init__(self, instruments, compact=True, block=None, waves_per_eu=None, streams_per_lane=None):
        from .fuser import generate_source, compile_source, resolve_shape
        if not instruments:
            raise ValueError("empty scan")
        srcs = [generate_source(instr, compact=compact, block=block, waves_per_eu=waves_per_eu,
                                streams_per_lane=streams_per_lane)[0]
                for instr in instruments]
        for k, s in enumerate(srcs):
            if s != srcs[0]:
                raise ValueError("instance %d has a different chain structure than instance 0" % k)
        self.code_object, self.kernel_name = compile_source(srcs[0])
        self.block, self.waves_per_eu, self.streams_per_lane = resolve_shape(
            instruments[0], block, waves_per_eu, streams_per_lane)
        self.instruments = list(instruments)
        descs, offs, rots = [], [], []
        for instr in self.instruments:
            descs.append([c.mcv_desc() for c in instr.components])
            o, r = calcTransformations(instr)
            offs.append(np.asarray(o, dtype=np.float64).reshape(-1, 3))
            rots.append(np.asarray(r, dtype=np.float64).reshape(-1, 3, 3))
        self.handle = BatchHandle(descs, np.concatenate(offs), np.concatenate(rots),
                                  self.code_object, self.kernel_name)

    @property
    def ninst(self):
        return len(self.instruments)

    def _args(self, ncounts, seed, ntotalthreads, threads_per_block, gate_dead_neutrons):
        n = np.asarray(ncounts, dtype=np.uint64).reshape(-1)
        if n.size != self.ninst:
            raise ValueError("need %d neutron counts, got %d" % (self.ninst, n.size))
        seed = config.rng_seed if seed is None else seed
        gate = config.gate_dead_neutrons if gate_dead_neutrons is None else gate_dead_neutrons
        return (n, seed, ntotalthreads or config.ntotalthreads, threads_per_block or config.threads_per_block,
                MCV_FLAG_GATE_DEAD if gate else 0)

    def run(self, ncounts, seed=None, ntotalthreads=None, threads_per_block=None, gate_dead_neutrons=None,
            grid=0, use_graph=True):
        """All instances in one persistent launch; histograms are cleared first."""
        n, seed, nt, tpb, flags = self._args(ncounts, seed, ntotalthreads, threads_per_block, gate_dead_neutrons)
        self.handle.clear()
        return self.handle.run(n, seed, nt, tpb, flags, grid=grid, use_graph=use_graph)

    def run_sequential(self, ncounts, seed=None, ntotalthreads=None, threads_per_block=None,
                       gate_dead_neutrons=None):
        """The instances one launch each, in order (what a loop of runs does)."""
        n, seed, nt, tpb, flags = self._args(ncounts, seed, ntotalthreads, threads_per_block, gate_dead_neutrons)
        self.handle.clear()
        return self.handle.run_sequential(n, seed, nt, tpb, flags)

    def monitor(self, inst, comp):
        """(3, nbins) histogram (counts, sum p, sum p^2) of component `comp`
        (index or name) of instance `inst`."""
        if not isinstance(comp, (int, np.integer)):
            names = [c.name for c in self.instruments[inst].components]
            comp = names.index(comp)
        return self.handle.read_monitor(inst, comp)

    def store(self, accumulate=False):
        """Copy (or add, accumulate=True) every instance's histograms into its
        monitors' `out` arrays: one device-to-host copy of all of them, then
        each block written straight into its array by the runtime."""
        from .components.monitors.MonitorBase import MonitorBase
        nb = self.handle.nbins
        H = 3 * sum(nb)
        blocks, staged = [], []
        for k, instr in enumerate(self.instruments):
            off = k * H
            for i, c in enumerate(instr.components):
                if nb[i] > 0 and isinstance(c, MonitorBase):
                    if c.out.size != 3 * nb[i]:
                        raise ValueError("%s: out holds %d values, the device histogram %d" %
                                         (c.name, c.out.size, 3 * nb[i]))
                    if c.out.dtype == np.float64 and c.out.flags.c_contiguous and c.out.flags.writeable:
                        blocks.append((off, c.out))
                    else:  # an out array the runtime cannot write: through a temporary
                        tmp = np.zeros(3 * nb[i])  # added to, or written: the block's values either way
                        blocks.append((off, tmp))
                        staged.append((c, tmp))
                off += 3 * nb[i]
        self.handle.deliver_histograms(blocks, accumulate=accumulate)
        for c, tmp in staged:
            if accumulate:
                c.out += tmp.reshape(c.out.shape)
            else:
                c.out[...] = tmp.reshape(c.out.shape)

    def close(self):
        self.handle.close()
