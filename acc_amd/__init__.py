"""acc_amd -- MI355X-native engine for mcvine.acc's per-neutron propagate path.

Component classes mirror mcvine.acc's (acc/components/...), `run_script.run`
keeps the reference signature, and the hot path is a fused gfx950 HIP kernel
per instrument reached through the C ABI in include/mcv_acc.h.
"""
from . import config  # noqa: F401
from .instrument import Instrument, instrument  # noqa: F401

__version__ = "0.1.0"
