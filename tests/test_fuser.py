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
    assert "MCV_DEFINE_KERNELS_COMPACTED(" in fuser.generate_source(I["c2"]())[0]


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


def _body(src, fn):
    i = src.index(fn + "(")
    return src[i:src.index("\n  }\n", i)]


def test_entrance_split():
    """Phase A ends at the straight guide's entrance and phase B starts with
    its bounces.  With gating an absorbed neutron is dropped; without, it runs
    on from the next component, in phase A when that component's opening
    turns it away, else queued with the skip flag (phase_a returns 2), phase B
    then skipping the guide's bounces for it.  The adjoint kernel (gated
    runs) splits the same way, the guide's scores starting in phase B; its
    ungated phase A runs the whole guide."""
    from acc_amd.fuser_adj import generate_adjoint_source
    from acc_amd import configs
    src = fuser.generate_source(prebuilt.instruments()["c2"]())[0]
    a, b = _body(src, "phase_a"), _body(src, "phase_b")
    assert "if (A.gate) {\n    if (!mcv::guide_enter(n, P + A.off[1])) return 0;\n  } else if (!mcv::guide_enter(" in a
    assert "mcv::taper_admits(m," in a and "return 2;" in a
    assert "mcv::guide_bounces" not in a and "mcv::guide_enter" not in b
    assert b.index("if (!skip) { mcv::guide_bounces") < b.index("mcv::guide_tapering")
    src = generate_adjoint_source(configs.c5())[0]
    a, b = _body(src, "phase_a"), _body(src, "phase_b")
    assert "if (A.gate) {\n    if (!mcv::guide_enter(n, P + A.off[1])) return 0;\n  } else {" in a
    assert "mcv::guide_enter" not in b
    assert b.index("if (A.gate) { mcv::guide_bounces") < b.index("mcv::guide_tapering")


def test_ms_outputs_streamed():
    """A multiple-scattering chain whose downstream components draw nothing
    hands each output down the chain as it is produced; the compacted kernel
    tells the scatterer that phase A has found the neutron not exiting."""
    from acc_amd import configs
    src = fuser.generate_source(configs.c3())[0]
    b = _body(src, "phase_b")
    assert "mcv::hms_scatter<Shape1, 2, 10, true>(rng, n, P + A.off[1], P + A.doff[1], emit)" in b
    assert "NeutronS out[" not in b
    plain = fuser.generate_source(configs.c3(), compact=False)[0]
    assert "mcv::hms_scatter<Shape1, 2, 10, false>" in plain


def test_ms_outputs_listed_before_a_drawing_component():
    """When a component after the scatterer draws random numbers, the outputs
    are kept in a list and run down the chain after the scatterer, as the
    reference does (the draw order per stream depends on it)."""
    b = _body(fuser.generate_source(prebuilt.instruments()["hms-sample"]())[0], "phase_b")
    assert "mcv::NeutronS out[8];" in b and "mcv::hms_scatterM<Shape1, 1, 8>" in b
    assert b.index("hms_scatterM") < b.index("hss_propagate")


def test_launch_defaults_resolved_per_chain():
    """Multiple-scattering chains take their own default launch shape (one
    wavefront per workgroup, 2 waves per SIMD); the other chains the general
    one; given values win.  run_script.get_handle resolves its shape here."""
    from acc_amd import config, configs
    ms = configs.c3(ntubes=40, npixels=16, ntof=50)
    assert fuser.resolve_shape(ms) == (config.ms_kernel_block, config.ms_kernel_waves_per_eu,
                                       config.kernel_streams_per_lane)
    assert fuser.resolve_shape(prebuilt.instruments()["c2"]()) == (config.kernel_block, config.kernel_waves_per_eu,
                                                                 config.kernel_streams_per_lane)
    assert fuser.resolve_shape(ms, block=256, waves_per_eu=3, streams_per_lane=2) == (256, 3, 2)
    assert "#define MCV_BLOCK %d" % config.ms_kernel_block in fuser.generate_source(ms)[0]


def test_scheduler_flags_per_chain():
    """Guide chains compile with the max-ILP scheduler, multiple-scattering
    chains with the default one (measured faster for each)."""
    from acc_amd import configs
    c2 = fuser.compile_flags(fuser.generate_source(prebuilt.instruments()["c2"]())[0])
    c3 = fuser.compile_flags(fuser.generate_source(configs.c3())[0])
    assert "-amdgpu-sched-strategy=max-ilp" in c2 and "-amdgpu-sched-strategy=max-ilp" not in c3
    assert all(f in c3 for f in fuser.HIPCC_FLAGS) and "-ffp-contract=off" in c2
