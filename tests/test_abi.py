"""The C ABI library loads and exports every entry point include/mcv_acc.h
declares; calls that need no device behave without one."""
import ctypes
import os
import re

import numpy as np
import pytest

from acc_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)  # drop comments
            names += re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(mcv_\w+)\s*\(", text, re.M)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    L = _abi.lib()
    names = declared_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), n
    assert set(_abi.EXPORTED) <= set(names)


def test_abi_version_and_constants():
    L = _abi.lib()
    assert L.mcv_abi_version() == _abi.ABI_VERSION == 2
    text = open(os.path.join(ROOT, "include", "mcv_acc.h")).read()
    for name in ("MCV_SOURCE_SIMPLE", "MCV_GUIDE", "MCV_GUIDE_TAPERING", "MCV_SAMPLE_HSS", "MCV_SAMPLE_HMS",
                 "MCV_MON_PSD", "MCV_MON_WAVELENGTH", "MCV_MON_IQE", "MCV_DET_TUBES", "MCV_SRC_IQE_NORMALIZER",
                 "MCV_KERNEL_ISOTROPIC", "MCV_KERNEL_SQE_GRID", "MCV_KERNEL_PHONON_DOS", "MCV_CSG_SPHERE",
                 "MCV_CSG_CYLINDER", "MCV_CSG_BLOCK", "MCV_CSG_UNION", "MCV_CSG_DIFFERENCE",
                 "MCV_CSG_INTERSECTION", "MCV_CSG_NODE_SIZE", "MCV_ERR_COUNT", "MCV_ERR_HIP"):
        m = re.search(r"#define\s+%s\s+(-?\d+)" % name, text)
        assert m, name
        assert int(m.group(1)) == getattr(_abi, name), name


def test_no_device_errors_cleanly():
    if _abi.device_count() > 0:
        pytest.skip("a GPU is visible")
    from acc_amd import configs
    from acc_amd.instrument import calcTransformations
    instr = configs.c2()
    off, rot = calcTransformations(instr)
    with pytest.raises(_abi.McvError, match="no HIP device"):
        _abi.Handle([c.mcv_desc() for c in instr.components], off, rot, "/nonexistent.hsaco", "k")
    with pytest.raises(_abi.McvError):
        _abi.rng_states(1, 10)


def test_scratch_fit_arithmetic():
    """Workgroups whose lanes' scratch fits a budget: budget // (bytes x lanes),
    at most the resident count, 0 when one does not fit, unbounded without
    scratch."""
    fit = _abi.scratch_fit
    assert fit(0, 256, 512, 0) == 512
    assert fit(25632, 64, 10 ** 6, 1 << 30) == (1 << 30) // (25632 * 64)
    assert fit(25632, 64, 100, 1 << 30) == 100
    assert fit(25632, 64, 100, 25632 * 64 - 1) == 0
    assert fit(25632, 64, 100, 25632 * 64) == 1
    assert fit(100, 128, 0, 1 << 30) == 0
    assert fit(100, 0, 5, 1 << 30) == 0


def test_scratch_guard_on_ms_kernel_metadata():
    """The multiple-scattering kernel with the reference's 100-neutron lists
    keeps tens of KB per lane in scratch (code-object metadata of the built
    kernel): at the device's full residency (256 CUs, 2 one-wave workgroups
    per SIMD) that is several GB, so a small budget narrows the grid and the
    forward kernel runs in more launches."""
    from acc_amd import prebuilt
    from acc_amd.fuser import fuse
    from codeobj_tools import private_segment_sizes
    path, name = fuse(prebuilt.instruments()["hms-ref"]())
    sizes = private_segment_sizes(path)
    per_lane = sizes[name]
    assert sizes[name + "_batch"] > 0 and per_lane > 16384, sizes
    resident = 256 * 4 * 2
    full = per_lane * 64 * resident
    assert _abi.scratch_fit(per_lane, 64, resident, full) == resident
    assert _abi.scratch_fit(per_lane, 64, resident, full // 8) == resident // 8
