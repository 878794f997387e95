"""Instruments whose fused kernels are compiled ahead of time by the build step.

Every chain structure the packaged workloads, the smoke run and the GPU tests
use is listed here, so a machine that only runs (and does not compile) finds
its code objects in the kernel cache.
"""
import os


def _c2_no_taper():
    from .instrument import instrument
    from .configs import _source
    from .components.optics.guide import Guide
    from .components.monitors.wavelength_monitor import Wavelength_monitor
    from .components.monitors.psd_monitor import PSD_monitor
    i = instrument("c2-straight")
    i.append(_source(52.5, 47.5), position=(0, 0, 0))
    i.append(Guide('guide', w1=0.05, h1=0.05, w2=0.05, h2=0.05, l=10.), position=(0, 0, 10.0))
    i.append(Wavelength_monitor('wavelength', Lmin=1., Lmax=10., nchan=200, xwidth=0.06, yheight=0.06),
             position=(0, 0, 20.001))
    i.append(PSD_monitor('psd', xwidth=0.06, yheight=0.06, nx=100, ny=100), position=(0, 0, 20.001))
    return i


def _source_only():
    from .instrument import instrument
    from .configs import _source
    from .components.monitors.psd_monitor import PSD_monitor
    i = instrument("source-psd")
    i.append(_source(60.0, 10.0), position=(0, 0, 0))
    i.append(PSD_monitor('psd', xwidth=0.2, yheight=0.2, nx=40, ny=40), position=(0, 0, 10.0))
    return i


def _hollow_cylinder_hss():
    from .instrument import instrument
    from .configs import _source, isotropic_al_kernel
    from .components.samples.homogeneous_single_scatterer import factory
    from .components.monitors.psd_monitor import PSD_monitor
    from .components.monitors.iqe_monitor import IQE_monitor
    from .geometry import Cylinder, Difference
    i = instrument("hollow-cylinder-hss")
    i.append(_source(60.0, 10.0), position=(0, 0, 0))
    shape = Difference(Cylinder(0.015, 0.05), Cylinder(0.01, 0.06))
    HSS = factory(shape=shape, kernel=isotropic_al_kernel())
    s = HSS('sample')
    i.append(s, position=(0, 0, 10.0), orientation=(90, 0, 0))
    i.append(IQE_monitor('iqe', Ei=60., Qmin=0., Qmax=12., nQ=60, Emin=-50., Emax=50., nE=50),
             position=(0, 0, 10.0))
    i.append(PSD_monitor('psd', xwidth=0.5, yheight=0.5, nx=50, ny=50), position=(0, 0, 11.0))
    return i


def _verdi_taper():
    """Reference tapered-guide test geometry (tests/components/optics/
    acc_tapered_guide_instrument.py:17-54) with its VERDI segment file."""
    from .instrument import instrument
    from .components.sources.source_simple import Source_simple
    from .components.optics.guide_tapering import Guide
    from .components.monitors.psd_monitor import PSD_monitor
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    path = os.path.join(here, "tests", "golden", "optics", "VERDI_V01_guide1.1")
    i = instrument("verdi-taper")
    i.append(Source_simple('source', radius=0., width=0.03, height=0.03, dist=6.35, xw=0.35, yh=0.35,
                           Lambda0=10., dLambda=9.5, E0=0, dE=0), position=(0, 0, 0.))
    i.append(Guide('guide', option="file={}".format(path), l=10.99, mx=6., my=6.), position=(0, 0, 6.35))
    i.append(PSD_monitor('Ixy', nx=250, ny=250, xwidth=0.25, yheight=0.25), position=(0, 0, 6.35 + 10.99 + 1e-3))
    return i


def _hms_reference():
    """The reference's multiple-scattering test sample
    (tests/components/samples/HMS_isotropic_hollowcylinder.py: max_ms_loops=2,
    max_ms_loops_path1=5, max_scattered_neutrons=100; hollow cylinder r 2/1 cm,
    h 6/7 cm, coefficients 10/m) behind a small source, with a PSD after it."""
    from .instrument import instrument
    from .components.sources.source_simple import Source_simple
    from .components.samples import loadFirstHomogeneousScatterer
    from .components.samples.homogeneous_multiple_scatterer import factory
    from .components.monitors.psd_monitor import PSD_monitor
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    hs = loadFirstHomogeneousScatterer(os.path.join(here, "tests", "golden", "sampleassemblies",
                                                    "isotropic_hollowcylinder", "sampleassembly.xml"))
    HMS = factory(shape=hs.shape(), kernel=hs.kernel(), max_ms_loops=2, max_ms_loops_path1=5,
                  max_scattered_neutrons=100)
    i = instrument("hms-ref")
    i.append(Source_simple('source', radius=0., width=0.01, height=0.01, dist=0.9, xw=0.008, yh=0.008,
                           Lambda0=10., dLambda=9.5, E0=0, dE=0), position=(0, 0, 0))
    i.append(HMS('sample'), position=(0, 0, 1.0), orientation=(90, 0, 0))
    i.append(PSD_monitor('psd', xwidth=0.1, yheight=0.1, nx=50, ny=50), position=(0, 0, 1.5))
    return i


def _hms_then_sample():
    """A multiple-scattering sample followed by a component that draws (a
    single-scattering sphere): the scatterer's outputs are kept in a list and
    run down the chain afterwards, in the reference's order of draws."""
    from .instrument import instrument
    from .configs import _source, hollow_cylinder, isotropic_al_kernel
    from .geometry import Sphere
    from .components.samples.homogeneous_multiple_scatterer import factory as ms_factory
    from .components.samples.homogeneous_single_scatterer import factory as ss_factory
    from .components.monitors.psd_monitor import PSD_monitor
    HMS = ms_factory(shape=hollow_cylinder(), kernel=isotropic_al_kernel(50.0, 5.0), max_ms_loops=2,
                     max_ms_loops_path1=2, max_scattered_neutrons=8)
    HSS = ss_factory(shape=Sphere(0.05), kernel=isotropic_al_kernel(50.0, 5.0))
    i = instrument("hms-sample")
    i.append(_source(60.0, 10.0), position=(0, 0, 0))
    i.append(HMS('ms'), position=(0, 0, 10.0))
    i.append(HSS('ss'), position=(0, 0, 10.3))
    i.append(PSD_monitor('psd', xwidth=0.4, yheight=0.4, nx=40, ny=40), position=(0, 0, 10.6))
    return i


def _csg_chain(xmlname, orientation=(0, 0, 0)):
    """Single-scatterer sample with a shape from the reference's geometry
    tests (tests/geometry/*.xml), for the union/intersection/difference code."""
    from .instrument import instrument
    from .configs import _source, isotropic_al_kernel
    from .components.samples.homogeneous_single_scatterer import factory
    from .components.monitors.psd_monitor import PSD_monitor
    from .geometry import parse_file
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    shape = parse_file(os.path.join(here, "tests", "golden", "geometry", xmlname))[0]
    i = instrument("csg-" + xmlname)
    i.append(_source(60.0, 10.0), position=(0, 0, 0))
    HSS = factory(shape=shape, kernel=isotropic_al_kernel(50.0, 2.0))
    i.append(HSS('sample'), position=(0, 0, 10.0), orientation=orientation)
    i.append(PSD_monitor('psd', xwidth=0.5, yheight=0.5, nx=50, ny=50), position=(0, 0, 10.5))
    return i


def _iqe_normalizer():
    from .instrument import instrument
    from .normalization import IQENormalizerSource
    from .components.monitors.iqe_monitor import IQE_monitor
    i = instrument("iqe-normalizer")
    i.append(IQENormalizerSource("normalizer", 60.0, -50.0, 50.0), position=(0, 0, 0))
    i.append(IQE_monitor('iqe', Ei=60., Qmin=0., Qmax=12., nQ=120, Emin=-50., Emax=50., nE=100),
             position=(0, 0, 0))
    return i


def grid_sphere(S=None):
    """Source (30 +- 10 meV) -> sphere with an S(Q,E)-grid kernel (`S`, default
    the 40 x 40 ridge grid) -> I(Q,E) monitor over all angles: the chain of the
    dense-gradient test."""
    from .instrument import instrument
    from .configs import _source, sqe_grid_kernel, sqe_ridge_grid
    from .components.samples.homogeneous_single_scatterer import factory
    from .components.monitors.iqe_monitor import IQE_monitor
    from .geometry import Sphere
    from .kernels import SQEGridKernel
    S0, (Qmin, Qmax, Emin, Emax) = sqe_ridge_grid(nQ=40, nE=40)
    S = S0 if S is None else S
    k = sqe_grid_kernel()
    kern = SQEGridKernel(S, Qmin, Qmax, Emin, Emax, k.absorption_coefficient, k.scattering_coefficient)
    i = instrument("grid-sphere")
    i.append(_source(30.0, 10.0), position=(0, 0, 0))
    HSS = factory(shape=Sphere(0.01), kernel=kern)
    i.append(HSS('sample'), position=(0, 0, 10.0))
    i.append(IQE_monitor('iqe', Ei=30., Qmin=0., Qmax=10., nQ=50, Emin=-40., Emax=40., nE=40,
                         min_angle_in_plane=-180, max_angle_in_plane=180,
                         min_angle_out_of_plane=-90, max_angle_out_of_plane=90), position=(0, 0, 10.0))
    return i


# chains whose adjoint kernels are built ahead of time
ADJOINT = ("c5", "grid-sphere")


def _storage_psd():
    """Stored-neutron source (the reference's singletestneutron2.mcv) -> PSD."""
    from .instrument import instrument
    from .components.sources.neutronfromstorage import NeutronFromStorage
    from .components.monitors.psd_monitor import PSD_monitor
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    i = instrument("storage-psd")
    i.append(NeutronFromStorage('src', os.path.join(here, "tests", "golden", "sources", "singletestneutron2.mcv")),
             position=(0, 0, 0))
    i.append(PSD_monitor('psd', xwidth=0.1, yheight=0.1, nx=5, ny=5), position=(0, 0, 1.0))
    return i


def spin_beam_path():
    """A stored beam of 4096 neutrons with non-zero spins (seeded, written
    once into the kernel cache directory): x, y within 2 cm of the axis,
    directions within 1 % of z, speeds 2-4 km/s, s1 and s2 in (-1, 1)."""
    import numpy as np
    from .fuser import CACHE_DIR
    from .neutron_storage import write
    path = os.path.join(CACHE_DIR, "spin_beam_4096.mcv")
    if not os.path.exists(path):
        g = np.random.default_rng(4096)
        n = 4096
        rows = np.zeros((n, 10))
        rows[:, 0:2] = g.uniform(-0.02, 0.02, (n, 2))
        v = g.uniform(2000.0, 4000.0, n)
        rows[:, 3:5] = g.uniform(-0.01, 0.01, (n, 2)) * v[:, None]
        rows[:, 5] = v
        rows[:, 6:8] = g.uniform(-1.0, 1.0, (n, 2))
        rows[:, 9] = 1.0
        os.makedirs(CACHE_DIR, exist_ok=True)
        tmp = path + ".tmp%d" % os.getpid()
        write(tmp, rows)
        os.replace(tmp, path)
    return path


def _storage_beam_psd():
    """The stored 4096-neutron beam with spins -> PSD (a plain kernel whose
    source reads row i % 4096 for neutron i)."""
    from .instrument import instrument
    from .components.sources.neutronfromstorage import NeutronFromStorage
    from .components.monitors.psd_monitor import PSD_monitor
    i = instrument("storage-beam-psd")
    i.append(NeutronFromStorage('src', spin_beam_path()), position=(0, 0, 0))
    i.append(PSD_monitor('psd', xwidth=0.1, yheight=0.1, nx=50, ny=50), position=(0, 0, 1.0))
    return i


def _storage_guide(sample):
    """Stored neutrons with spins -> straight guide -> (sphere, or a second
    guide) -> PSD: the spins go through the compacted kernel's queue (the
    stream-carrying one with the sphere, the plain one without)."""
    from .instrument import instrument
    from .configs import isotropic_al_kernel
    from .components.sources.neutronfromstorage import NeutronFromStorage
    from .components.optics.guide import Guide
    from .components.samples.homogeneous_single_scatterer import factory
    from .components.monitors.psd_monitor import PSD_monitor
    from .geometry import Sphere
    i = instrument("storage-guide" + ("-sample" if sample else ""))
    i.append(NeutronFromStorage('src', spin_beam_path()), position=(0, 0, 0))
    i.append(Guide('guide', w1=0.03, h1=0.03, w2=0.03, h2=0.03, l=2.0), position=(0, 0, 1.0))
    if sample:
        i.append(factory(shape=Sphere(0.02), kernel=isotropic_al_kernel())('sample'), position=(0, 0, 3.2))
    else:
        i.append(Guide('guide2', w1=0.03, h1=0.03, w2=0.02, h2=0.02, l=0.3), position=(0, 0, 3.1))
    i.append(PSD_monitor('psd', xwidth=0.2, yheight=0.2, nx=20, ny=20), position=(0, 0, 3.5))
    return i


def _rotated():
    """Source -> arm turned 0.4 deg about y and 0.2 deg about x -> guide
    placed relative to the arm -> PSD tilted by 10 deg, relative to the arm:
    multi-axis orientations and positions relative to another component."""
    from .instrument import instrument
    from .configs import _source
    from .components.optics.arm import Arm
    from .components.optics.guide import Guide
    from .components.monitors.psd_monitor import PSD_monitor
    i = instrument("rotated")
    i.append(_source(60.0, 10.0), position=(0, 0, 0))
    arm = Arm('arm')
    i.append(arm, position=(0.01, 0.0, 9.5), orientation=(0.2, 0.4, 0))
    i.append(Guide('guide', w1=0.08, h1=0.08, w2=0.06, h2=0.06, l=1.0), position=(0, 0, 0.5), relativeTo=arm)
    i.append(PSD_monitor('psd', xwidth=0.4, yheight=0.4, nx=40, ny=40), position=(0, 0, 2.0), orientation=(0, 10, 0),
             relativeTo=arm)
    return i


def _storage_ms():
    """Stored neutrons with spins -> multiple scatterer -> PSD: the outputs
    keep the incident neutron's spins."""
    from .instrument import instrument
    from .configs import hollow_cylinder, isotropic_al_kernel
    from .components.sources.neutronfromstorage import NeutronFromStorage
    from .components.samples.homogeneous_multiple_scatterer import factory
    from .components.monitors.psd_monitor import PSD_monitor
    HMS = factory(shape=hollow_cylinder(), kernel=isotropic_al_kernel(50.0, 5.0), max_ms_loops=2,
                  max_ms_loops_path1=2, max_scattered_neutrons=7)
    i = instrument("storage-ms")
    i.append(NeutronFromStorage('src', spin_beam_path()), position=(0, 0, 0))
    i.append(HMS('sample'), position=(0, 0, 1.0), orientation=(90, 0, 0))
    i.append(PSD_monitor('psd', xwidth=0.4, yheight=0.4, nx=40, ny=40), position=(0, 0, 1.5))
    return i


def _taper_coatings(mx, my):
    """Source -> 100-segment tapered guide with coating indices mx, my -> PSD:
    m = 0 (a wall hit ends the segment, the reference's extra channel shift)
    and m < 0 (m scaled by the entrance width over the segment's)."""
    from .instrument import instrument
    from .configs import _source, taper_file
    from .components.optics.guide_tapering import Guide as TaperedGuide
    from .components.monitors.psd_monitor import PSD_monitor
    i = instrument("taper-m%g-%g" % (mx, my))
    i.append(_source(52.5, 47.5), position=(0, 0, 0))
    i.append(TaperedGuide('taper', option="file=" + taper_file(), l=10., mx=mx, my=my, R0=0.99, Qcx=0.0219,
                          Qcy=0.0219, alphax=6.07, alphay=6.07, W=0.003), position=(0, 0, 10.0))
    i.append(PSD_monitor('psd', xwidth=0.1, yheight=0.1, nx=50, ny=50), position=(0, 0, 20.5))
    return i


def _hms_shape(kind):
    """Source -> multiple scatterer on a solid sphere or a plate (the
    scatterer's geometry on convex shapes: one entry, one exit) -> PSD."""
    from .instrument import instrument
    from .configs import _source, isotropic_al_kernel
    from .geometry import Sphere, Block
    from .components.samples.homogeneous_multiple_scatterer import factory
    from .components.monitors.psd_monitor import PSD_monitor
    shape = Sphere(0.02) if kind == "sphere" else Block(0.04, 0.04, 0.01)
    HMS = factory(shape=shape, kernel=isotropic_al_kernel(80.0, 5.0), max_ms_loops=3, max_ms_loops_path1=2,
                  max_scattered_neutrons=10)
    i = instrument("hms-" + kind)
    i.append(_source(60.0, 10.0), position=(0, 0, 0))
    i.append(HMS('sample'), position=(0, 0, 10.0), orientation=(0, 30, 0))
    i.append(PSD_monitor('psd', xwidth=0.4, yheight=0.4, nx=40, ny=40), position=(0, 0, 10.5))
    return i


def batch_plain(k):
    """Instance k of a batch of plain chains: source (E0 40 + 15 k meV) ->
    sphere (r 8 + 2 k mm) -> I(Q,E) -> PSD 250 x 250 (the LDS bin cache)."""
    from .instrument import instrument
    from .configs import _source, isotropic_al_kernel
    from .geometry import Sphere
    from .components.samples.homogeneous_single_scatterer import factory
    from .components.monitors.psd_monitor import PSD_monitor
    from .components.monitors.iqe_monitor import IQE_monitor
    E0 = 40.0 + 15.0 * k
    i = instrument("batch-plain-%d" % k)
    i.append(_source(E0, 8.0), position=(0, 0, 0))
    sample = factory(shape=Sphere(0.008 + 0.002 * k), kernel=isotropic_al_kernel())('sample')
    i.append(sample, position=(0, 0, 10.0))
    i.append(IQE_monitor('iqe', Ei=E0, Qmin=0., Qmax=12., nQ=120, Emin=-50., Emax=50., nE=100),
             position=(0, 0, 0), relativeTo=sample)
    i.append(PSD_monitor('psd', xwidth=0.5, yheight=0.5, nx=250, ny=250), position=(0, 0, 11.0))
    return i


def batch_compacted(k):
    """Instance k of a batch of compacted chains that draw nothing after the
    culling point: source (E0 30 + 20 k meV) -> guides -> wavelength + PSD."""
    from .instrument import instrument
    from .configs import _source, _guides
    from .components.monitors.psd_monitor import PSD_monitor
    from .components.monitors.wavelength_monitor import Wavelength_monitor
    i = instrument("batch-compacted-%d" % k)
    i.append(_source(30.0 + 20.0 * k, 10.0), position=(0, 0, 0))
    _guides(i)
    i.append(Wavelength_monitor('wavelength', Lmin=1., Lmax=10., nchan=200, xwidth=0.06, yheight=0.06),
             position=(0, 0, 30.001))
    i.append(PSD_monitor('psd', xwidth=0.06, yheight=0.06, nx=100, ny=100), position=(0, 0, 30.001))
    return i


def batch_ms(k):
    """Instance k of a batch of multiple-scattering chains: source (E0 50 +
    10 k meV) -> hollow cylinder (scattering coefficient 40 + 20 k /m) -> PSD."""
    from .instrument import instrument
    from .configs import _source, hollow_cylinder, isotropic_al_kernel
    from .components.samples.homogeneous_multiple_scatterer import factory
    from .components.monitors.psd_monitor import PSD_monitor
    HMS = factory(shape=hollow_cylinder(), kernel=isotropic_al_kernel(40.0 + 20.0 * k, 5.0), max_ms_loops=2,
                  max_ms_loops_path1=2, max_scattered_neutrons=7)
    i = instrument("batch-ms-%d" % k)
    i.append(_source(50.0 + 10.0 * k, 10.0), position=(0, 0, 0))
    i.append(HMS('sample'), position=(0, 0, 10.0), orientation=(90, 0, 0))
    i.append(PSD_monitor('psd', xwidth=0.4, yheight=0.4, nx=40, ny=40), position=(0, 0, 10.5))
    return i


def _two_samples():
    """Source -> two drawing samples in a row -> PSD: two stochastic
    components after the culling point share each neutron's carried stream."""
    from .instrument import instrument
    from .configs import _source, isotropic_al_kernel
    from .components.samples.homogeneous_single_scatterer import factory
    from .components.monitors.psd_monitor import PSD_monitor
    from .geometry import Sphere, Block
    i = instrument("two-samples")
    i.append(_source(60.0, 10.0), position=(0, 0, 0))
    i.append(factory(shape=Sphere(0.02), kernel=isotropic_al_kernel())('s1'), position=(0, 0, 10.0))
    i.append(factory(shape=Block(0.05, 0.05, 0.01), kernel=isotropic_al_kernel())('s2'), position=(0, 0, 10.2))
    i.append(PSD_monitor('psd', xwidth=0.5, yheight=0.5, nx=50, ny=50), position=(0, 0, 11.0))
    return i


def _three_guides():
    """Source -> three straight guides with gaps -> PSD: phase A ends at the
    first guide's entrance; the later guides run whole in phase B."""
    from .instrument import instrument
    from .configs import _source
    from .components.optics.guide import Guide
    from .components.monitors.psd_monitor import PSD_monitor
    i = instrument("three-guides")
    i.append(_source(52.5, 47.5), position=(0, 0, 0))
    for k, (w, m) in enumerate(((0.05, 2.0), (0.045, 3.0), (0.04, 1.0))):
        i.append(Guide('guide%d' % k, w1=w, h1=w, w2=w - 0.005, h2=w - 0.005, l=1.0, m=m),
                 position=(0, 0, 10.0 + 1.1 * k))
    i.append(PSD_monitor('psd', xwidth=0.06, yheight=0.06, nx=30, ny=30), position=(0, 0, 13.5))
    return i


def _mixed_chain():
    """Source -> PSD -> slit -> guide -> tapered guide -> sphere -> beamstop
    -> I(Q,E) -> wavelength monitor: a monitor and an aperture before the
    culling point, a drawing sample after it, an aperture and two monitors
    after the sample."""
    from .instrument import instrument
    from .configs import _source, _guides, isotropic_al_kernel
    from .geometry import Sphere
    from .components.samples.homogeneous_single_scatterer import factory
    from .components.optics.slit import Slit
    from .components.optics.beamstop import Beamstop
    from .components.monitors.psd_monitor import PSD_monitor
    from .components.monitors.iqe_monitor import IQE_monitor
    from .components.monitors.wavelength_monitor import Wavelength_monitor
    i = instrument("mixed")
    i.append(_source(52.5, 47.5), position=(0, 0, 0))
    i.append(PSD_monitor('psd0', xwidth=0.2, yheight=0.2, nx=20, ny=20), position=(0, 0, 9.0))
    i.append(Slit('slit', width=0.06, height=0.06, cut=0.0), position=(0, 0, 9.9))
    _guides(i)
    i.append(factory(shape=Sphere(0.01), kernel=isotropic_al_kernel())('sample'), position=(0, 0, 30.5))
    i.append(Beamstop('stop', radius=0.005), position=(0, 0, 31.0))
    i.append(IQE_monitor('iqe', Ei=52.5, Qmin=0., Qmax=12., nQ=60, Emin=-50., Emax=50., nE=50),
             position=(0, 0, 30.5))
    i.append(Wavelength_monitor('wavelength', Lmin=1., Lmax=10., nchan=100, xwidth=0.2, yheight=0.2),
             position=(0, 0, 31.5))
    return i


def _guide_ms():
    """Source -> straight guide -> hollow cylinder with multiple scattering ->
    PSD: a culling guide before the multiple scatterer."""
    from .instrument import instrument
    from .configs import _source, hollow_cylinder, isotropic_al_kernel
    from .components.optics.guide import Guide
    from .components.samples.homogeneous_multiple_scatterer import factory
    from .components.monitors.psd_monitor import PSD_monitor
    HMS = factory(shape=hollow_cylinder(), kernel=isotropic_al_kernel(50.0, 5.0), max_ms_loops=2,
                  max_ms_loops_path1=3, max_scattered_neutrons=9)
    i = instrument("guide-ms")
    i.append(_source(60.0, 10.0), position=(0, 0, 0))
    i.append(Guide('guide', w1=0.05, h1=0.05, w2=0.03, h2=0.03, l=2.0), position=(0, 0, 10.0))
    i.append(HMS('sample'), position=(0, 0, 12.5), orientation=(90, 0, 0))
    i.append(PSD_monitor('psd', xwidth=0.4, yheight=0.4, nx=40, ny=40), position=(0, 0, 13.0))
    return i


def _aperture_monitors():
    """Source -> slit -> circular beamstop -> PosDiv monitor -> PSD_monitor_4Pi:
    the cheap optics and monitors of the reference's sgm/sphere scripts."""
    from .instrument import instrument
    from .configs import _source
    from .components.optics.slit import Slit
    from .components.optics.beamstop import Beamstop
    from .components.monitors.posdiv_monitor import PosDiv_monitor
    from .components.monitors.psd_monitor_4pi import PSD_monitor_4Pi
    i = instrument("aperture-monitors")
    i.append(_source(60.0, 10.0), position=(0, 0, 0))
    i.append(Slit('slit', width=0.08, height=0.06, cut=0.0), position=(0, 0, 9.8))
    i.append(Beamstop('stop', radius=0.01), position=(0, 0, 9.9))
    i.append(PosDiv_monitor('posdiv', xwidth=0.1, yheight=0.1, maxdiv=0.5, npos=40, ndiv=40), position=(0, 0, 10.0))
    i.append(PSD_monitor_4Pi('m4pi', radius=0.5, nphi=60, ntheta=30), position=(0, 0, 10.0))
    return i


def single(comp):
    """A one-component chain (the buffered-mode unit)."""
    from .instrument import instrument
    i = instrument("process-" + comp.name)
    i.append(comp, position=(0, 0, 0))
    return i


def process_units():
    """Components the buffered-mode tests run alone."""
    from . import configs
    from .components.optics.guide import Guide
    from .components.monitors.psd_monitor import PSD_monitor
    c2 = configs.c2()
    hms = _hms_reference()
    return {
        "source": lambda: single(configs._source(60.0, 10.0)),
        "guide": lambda: single(Guide('guide', w1=0.05, h1=0.05, w2=0.03, h2=0.03, l=5.)),
        "taper": lambda: single(c2.components[2]),
        "psd": lambda: single(PSD_monitor('psd', xwidth=0.1, yheight=0.1, nx=20, ny=20)),
        "hms": lambda: single(hms.components[1]),
        "storage": lambda: single(_storage_psd().components[0]),
        "slit": lambda: single(__import__("acc_amd.components.optics.slit", fromlist=["Slit"]).Slit("s", width=1)),
        "beamstop": lambda: single(
            __import__("acc_amd.components.optics.beamstop", fromlist=["Beamstop"]).Beamstop("b", width=1)),
        "c2-chain": configs.c2,
        "c5-chain": configs.c5,
    }


def instruments():
    from . import configs
    return {
        "c1": configs.c1, "c2": configs.c2, "c2s": configs.c2s,
        "c2-straight": _c2_no_taper, "source-psd": _source_only,
        "hollow-cylinder-hss": _hollow_cylinder_hss, "verdi-taper": _verdi_taper,
        "c3s": configs.c3_single, "c3s-small": lambda: configs.c3_single(ntubes=40, npixels=16, ntof=50),
        "c5s": configs.c5_single,
        "c3": configs.c3, "c3-small": lambda: configs.c3(ntubes=40, npixels=16, ntof=50),
        "hms-ref": _hms_reference, "hms-sample": _hms_then_sample,
        "csg-union2": lambda: _csg_chain("union_example2.xml", (90, 0, 0)),
        "csg-intersection": lambda: _csg_chain("intersection_example1.xml", (90, 0, 0)),
        "csg-difference": lambda: _csg_chain("difference_example1.xml", (90, 0, 0)),
        "iqe-normalizer": _iqe_normalizer,
        "scan": lambda: configs.scan_instance(60.0, 30.0),
        "scan-plate": lambda: configs.scan_instance(60.0, 30.0, sample="plate"),
        "c5": configs.c5,
        "grid-sphere": grid_sphere,
        "storage-psd": _storage_psd,
        "aperture-monitors": _aperture_monitors,
        "two-samples": _two_samples,
        "three-guides": _three_guides,
        "mixed": _mixed_chain,
        "batch-plain": lambda: batch_plain(0),
        "batch-compacted": lambda: batch_compacted(0),
        "batch-ms": lambda: batch_ms(0),
        "hms-sphere": lambda: _hms_shape("sphere"),
        "hms-plate": lambda: _hms_shape("plate"),
        "taper-m0": lambda: _taper_coatings(0.0, -2.0),
        "taper-mneg": lambda: _taper_coatings(-1.5, 0.0),
        "rotated": _rotated,
        "storage-ms": _storage_ms,
        "storage-beam-psd": _storage_beam_psd,
        "storage-guide": lambda: _storage_guide(False),
        "storage-guide-sample": lambda: _storage_guide(True),
        "guide-ms": _guide_ms,
    }


# chains also built for every launch shape the autotuner tries
TUNED = ("c2", "c2s", "c5")
# multiple-scattering chains also built uncompacted (their compaction invariance test)
UNCOMPACTED = ("c3-small", "hms-ref")


def build_all(verbose=False, jobs=None):
    """Compile every listed chain (hipcc runs in parallel, one per job)."""
    from concurrent.futures import ThreadPoolExecutor
    from .fuser import generate_source, compile_source
    from .fuser_adj import generate_adjoint_source
    srcs = {name: generate_source(make())[0] for name, make in instruments().items()}
    for name in ADJOINT:
        srcs[name + "@adjoint"] = generate_adjoint_source(instruments()[name]())[0]
    from .fuser_proc import generate_process_source
    for name, make in process_units().items():
        srcs[name + "@process"] = generate_process_source(make())[0]
    from .autotune import DEFAULT_BLOCKS, DEFAULT_WAVES, DEFAULT_SPL, _spls
    for name in TUNED:
        instr = instruments()[name]()
        for block in DEFAULT_BLOCKS:
            for w in DEFAULT_WAVES:
                for spl in _spls(instr, DEFAULT_SPL):
                    srcs["%s@%d/%s/%s" % (name, block, w, spl)] = generate_source(
                        instr, block=block, waves_per_eu=w, streams_per_lane=spl)[0]
    for name in UNCOMPACTED:
        srcs[name + "@plain"] = generate_source(instruments()[name](), compact=False)[0]
    unique = sorted(set(srcs.values()))
    jobs = jobs or max(1, min(8, os.cpu_count() or 1))
    with ThreadPoolExecutor(jobs) as ex:
        built = dict(zip(unique, ex.map(lambda s: compile_source(s, verbose=verbose), unique)))
    return {name: built[s] for name, s in srcs.items()}


def prune(keep):
    """Delete cached chain kernels other than `keep` ({name: (path, kernel)})."""
    import glob
    from .fuser import CACHE_DIR
    names = {k for _, k in keep.values()}
    for path in glob.glob(os.path.join(CACHE_DIR, "mcv_fused_*")):
        if os.path.basename(path).split(".")[0] not in names:
            os.remove(path)
