"""Component plugin API (acc/components/ComponentBase.py).

As in the reference, a concrete component class (name not ending in "Base")
must define `propagate`, and its constructor sets `self.name` and
`self.propagate_params` (same tuple contents as the reference's).  The
reference's `propagate` is a Numba device function; here it is a
`DeviceFunction`: the name of the gfx950 device routine the fuser inlines into
the fused kernel, plus how the component packs its parameters for the C ABI
(include/mcv_acc.h).  Subclasses implement `_mcv_desc()`.

`process(neutrons)` is the buffered mode of the reference
(ComponentBase.py:79-107): an (N, 10) float64 array in, the surviving rows
(p > 0) out.
"""
import numpy as np


class DeviceFunction:
    """Marker for a component's device routine.

    `call` is a format string producing the C++ statement(s) for component
    index {i} acting on the neutron variable {n}; the fused kernel exposes
    `rng`, `P` (packed parameters), `H` (histograms) and the launch block `A`.
    """

    def __init__(self, call, stochastic=False, monitor=False, preamble=None, adjoint=None, dparams=None,
                 binning=None, split=None, opening=None):
        self.call = call
        self.stochastic = stochastic
        self.monitor = monitor
        self.preamble = preamble
        # adjoint kernels: `adjoint` is the score-tracking form of `call`
        # ({s} = the component's score array, {dz} = the dense tracker),
        # `dparams(comp)` lists (name, packed parameter index) in score-slot
        # order, and a monitor's `binning` computes (bin, nbins) as `call` does
        self.adjoint = adjoint
        self.dparams = dparams
        self.binning = binning
        # a component whose first part kills most neutrons (a guide entrance):
        # (`enter`, a bool expression that is false when the neutron dies,
        # `rest`, the statement(s) completing `call`, and `rest` in score-tracking
        # form, as `adjoint` is of `call`); a compacted kernel may end its
        # per-stream phase after `enter`
        self.split = split
        # a component that turns many neutrons away at its opening without
        # changing them otherwise: (`admits`, a bool expression that is false
        # exactly when the component would only apply `reject`, the
        # statement(s) it then applies, e.g. p = 0); without dead-neutron
        # gating a compacted kernel queues only the neutrons it admits
        self.opening = opening

    def render(self, comp, i, n="n"):
        return self.call.format(i=i, n=n, **comp._render_context(i))

    def render_adjoint(self, comp, i, s, n="n"):
        return self.adjoint.format(i=i, n=n, s=s, dz="dz", **comp._render_context(i))

    def render_split(self, comp, i, part, n="n", s=None):
        return self.split[part].format(i=i, n=n, s=s, **comp._render_context(i))

    def render_opening(self, comp, i, part, n="n"):
        return self.opening[part].format(i=i, n=n, **comp._render_context(i))

    def render_binning(self, comp, i, n="n"):
        return self.binning.format(i=i, n=n, **comp._render_context(i))

    def differentiable(self, comp):
        """(name, packed parameter index) of the weight-carrying parameters."""
        return list(self.dparams(comp)) if self.dparams is not None else []

    def render_preamble(self, comp, i):
        if self.preamble is None:
            return ""
        return self.preamble(comp, i)


class Curator(type):
    """Registers concrete components; requires `propagate`
    (acc/components/ComponentBase.py:21-32)."""

    components = []

    def __new__(cls, name, bases, dct):
        x = super().__new__(cls, name, bases, dct)
        if not name.endswith("Base"):
            if not hasattr(x, "propagate") or not isinstance(x.propagate, DeviceFunction):
                raise TypeError(f"{name} should define `propagate` method")
            Curator.components.append(x)
        return x


class ComponentBase(metaclass=Curator):

    propagate_params = ()

    # flag whether this is a multiple scattering component
    is_multiplescattering = False
    NUM_MULTIPLE_SCATTER = 10

    mcv_type = None

    def _render_context(self, i):
        return {}

    def _mcv_desc(self):
        raise NotImplementedError

    def mcv_desc(self):
        """dict(type, params, iparams, data, nbins) for include/mcv_acc.h's mcv_comp_desc."""
        d = self._mcv_desc()
        d.setdefault("iparams", np.zeros(0, dtype=np.int64))
        d.setdefault("data", np.zeros(0))
        d.setdefault("nbins", 0)
        return d

    def process(self, neutrons, seed=None):
        """Buffered mode: propagate an (N, 10) array through this component and
        return the rows with p > 0 (acc/components/ComponentBase.py:79-107)."""
        from ..buffered import process_component
        return process_component(self, neutrons, seed=seed)


class StochasticComponentBase(ComponentBase):
    """Components drawing random numbers (acc/components/StochasticComponentBase.py)."""
