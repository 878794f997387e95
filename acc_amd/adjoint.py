"""Differentiable runs: histograms as functions of weight-carrying parameters.

No reference counterpart (the reference simulates forward only).  For fixed
random streams every neutron's path is independent of the supermirror
coating, the sample's cross-sections and its S(Q,E): they scale the neutron's
weight.  `DifferentiableRun` evaluates the chosen monitors' sum-p histograms
for a parameter vector theta with the forward kernel, and their
vector-Jacobian product g^T dH/dtheta with the chain's adjoint kernel
(mcv_instr_adjoint), both over the same ncount neutrons, seed and partition
(common random numbers).  `acc_amd.autograd` wraps the pair in a
torch.autograd.Function.

Parameters are named "component.param" after the component's score slots
(Guide: R0, Qc, alpha, m, W; tapered Guide: R0, Qcx, Qcy, alphax, alphay, mx,
my, W; samples: absorption_coefficient, scattering_coefficient and, for the
phonon-DOS kernel, A1, c1, w1, A2, c2, w2, A3, c3, w3).  One theta entry may
drive several of them (a shared coating): give a list of names.
"""
import numpy as np

from . import config
from ._abi import MCV_FLAG_GATE_DEAD
from .fuser_adj import fuse_adjoint
from .run_script import get_handle


class DifferentiableRun:

    def __init__(self, instrument, params, ncount, seed=0, monitors=None, ntotalthreads=None,
                 threads_per_block=None, gate_dead_neutrons=True, block=None, waves_per_eu=None,
                 streams_per_lane=None, adjoint_block=None, adjoint_streams_per_lane=None, adjoint_compact=True):
        self.instrument = instrument
        self.ncount, self.seed = int(ncount), int(seed)
        self.ntotalthreads = ntotalthreads or config.ntotalthreads
        self.tpb = threads_per_block or config.threads_per_block
        self.flags = MCV_FLAG_GATE_DEAD if gate_dead_neutrons else 0
        self.h = get_handle(instrument, block=block, waves_per_eu=waves_per_eu, streams_per_lane=streams_per_lane)
        self.code_object, self.kernel_name, layout = fuse_adjoint(
            instrument, block=adjoint_block or config.adjoint_block, compact=adjoint_compact,
            streams_per_lane=adjoint_streams_per_lane or config.adjoint_streams_per_lane)
        comps = instrument.components
        names = [c.name for c in comps]
        # score slots
        self.slots = {}
        k = 0
        for ci, cname, dp in layout:
            for pname, pidx in dp:
                self.slots[(cname, pname)] = (ci, pidx, k)
                k += 1
        self.nscore = k
        # theta entries -> targets
        self.params = []
        for p in params:
            targets = [p] if isinstance(p, str) else list(p)
            tl = []
            for t in targets:
                cname, pname = t.split(".", 1)
                if (cname, pname) not in self.slots:
                    raise KeyError("%s is not a differentiable parameter of this chain (have %s)"
                                   % (t, sorted("%s.%s" % s for s in self.slots)))
                tl.append(self.slots[(cname, pname)])
            self.params.append(tl)
        self._packed = {i: np.array(c.mcv_desc()["params"], dtype=np.float64) for i, c in enumerate(comps)}
        mons = monitors or [c.name for c in comps if c.propagate.monitor]
        self.monitors = [names.index(m) for m in mons]
        self.sizes = [self.h.nbins[i] for i in self.monitors]
        self.grid_nodes = 0
        from ._abi import MCV_KERNEL_SQE_GRID
        for c in comps:
            if getattr(getattr(c, "kernel", None), "kernel_type", None) == MCV_KERNEL_SQE_GRID:
                self.grid_nodes = int(np.asarray(c.kernel.S).size)
        self.last_forward_stats = self.last_adjoint_stats = None

    # ---------------------------------------------------------------- parameters
    def theta0(self):
        """Current value of each theta entry (its first target's)."""
        return np.array([self._packed[tl[0][0]][tl[0][1]] for tl in self.params])

    def set_theta(self, theta):
        theta = np.asarray(theta, dtype=np.float64).reshape(-1)
        if theta.size != len(self.params):
            raise ValueError("theta has %d entries, expected %d" % (theta.size, len(self.params)))
        touched = set()
        for v, tl in zip(theta, self.params):
            for ci, pidx, _ in tl:
                self._packed[ci][pidx] = v
                touched.add(ci)
        for ci in sorted(touched):
            self.h.set_params(ci, self._packed[ci])

    # ---------------------------------------------------------------- evaluation
    def forward(self, theta=None):
        """Concatenated sum-p histograms of the selected monitors."""
        if theta is not None:
            self.set_theta(theta)
        self.h.clear()
        self.last_forward_stats = self.h.run(self.ncount, self.seed, self.ntotalthreads, self.tpb, self.flags)
        return np.concatenate([self.h.read_monitor(i)[1] for i in self.monitors])

    def counts(self):
        """Concatenated neutron-count histograms of the selected monitors after
        the last forward(): the weight-carrying parameters do not move any
        neutron, so equal counts at two parameter values certify that no
        neutron changed its fate (a kill decision at the reflectivity cutoff,
        which with dead neutrons dropped also changes the random numbers of the
        stream's later neutrons) between them."""
        return np.concatenate([self.h.read_monitor(i)[0] for i in self.monitors])

    def vjp(self, g, theta=None, dense=False):
        """g^T dH/dtheta for an upstream gradient g of forward()'s output
        (and, with dense=True, the gradient with respect to the S(Q,E) grid)."""
        if theta is not None:
            self.set_theta(theta)
        g = np.asarray(g, dtype=np.float64).reshape(-1)
        if g.size != sum(self.sizes):
            raise ValueError("upstream gradient has %d values, output %d" % (g.size, sum(self.sizes)))
        G = np.zeros(self.h.hist_total)
        o = 0
        for i, nb in zip(self.monitors, self.sizes):
            G[self.h.hoff[i] + nb: self.h.hoff[i] + 2 * nb] = g[o: o + nb]
            o += nb
        nd = self.grid_nodes if dense else 0
        slots, dgrad, st = self.h.adjoint(self.code_object, self.kernel_name, self.ncount, self.seed, G,
                                          self.nscore, self.ntotalthreads, self.tpb, self.flags, ndense=nd)
        self.last_adjoint_stats = st
        grad = np.array([sum(slots[k] for _, _, k in tl) for tl in self.params])
        return (grad, dgrad) if dense else grad
