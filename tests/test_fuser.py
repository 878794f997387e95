"""Fuser: chain sources, split points and gfx950 compilation (hipcc runs here
without a GPU)."""
import os

import pytest

from acc_amd import prebuilt, fuser


@pytest.mark.parametrize("name", sorted(prebuilt.instruments()))
def test_every_packaged_chain_compiles(name):
    instr = prebuilt.instruments()[name]()
    path, kernel = fuser.fuse(instr)
    assert os.path.exists(path) and kernel.startswith("mcv_fused_")


def test_split_points():
    I = prebuilt.instruments()
    c2 = I["c2"]().components
    assert fuser.split_index(c2) == 1 and not fuser.phase_b_draws(c2, 1)  # after the straight guide
    assert fuser.split_index(I["c2-straight"]().components) is None  # only monitors after the guide
    assert fuser.split_index(I["c1"]().components) is None        # only monitors after the sample
    c2s = I["c2s"]().components
    assert fuser.split_index(c2s) == 1 and fuser.phase_b_draws(c2s, 1)  # the sample draws in phase B
    src = fuser.generate_source(I["c2s"]())[0]
    assert "MCV_DEFINE_KERNELS_COMPACTED_RNG" in src and "#define MCV_BLOCK 256" in src
    assert "MCV_DEFINE_KERNELS_REFILL_RNG" in fuser.generate_source(I["c2s"](), refill=True)[0]
    assert "MCV_DEFINE_KERNELS_COMPACTED(" in fuser.generate_source(I["c2"]())[0]
    assert "MCV_DEFINE_KERNELS_REFILL(" in fuser.generate_source(I["c2"](), refill=True)[0]


def test_structure_only_enters_the_source():
    from acc_amd import configs
    a = fuser.generate_source(configs.c3(ntubes=40, npixels=16, ntof=50))[0]
    b = fuser.generate_source(configs.c3())[0]
    assert a == b  # sizes and values are runtime data


def test_multiple_ms_components_rejected():
    from acc_amd.instrument import instrument
    from acc_amd.configs import _source, hollow_cylinder, sqe_grid_kernel
    from acc_amd.components.samples.homogeneous_multiple_scatterer import factory
    i = instrument()
    i.append(_source(60., 10.))
    H = factory(hollow_cylinder(), sqe_grid_kernel())
    i.append(H("a"), position=(0, 0, 10))
    i.append(H("b"), position=(0, 0, 11))
    with pytest.raises(NotImplementedError):
        fuser.generate_source(i)
