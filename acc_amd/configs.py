"""Instruments of the benchmark workloads, built from the component classes.

 * c1: Source_simple (disk r 5 cm, target 10 cm x 10 cm at 10 m, E 60 +- 10 meV)
   -> isotropic sphere r 1 cm (sigma_scatt 5 b, sigma_abs 2 b, Al cell density)
   -> I(Q,E) monitor at the sample (Q 0-12 /AA x 120, E -50..50 meV x 100)
   -> 250 x 250 PSD (0.5 m x 0.5 m) at 11 m.
 * c2: Source_simple (E uniform 5-100 meV) -> 10 m straight guide 5 x 5 cm
   (m=2, R0 .99, Qc .0219, alpha 6.07, W .003) -> 10 m, 100-segment elliptic
   guide narrowing 6 -> 3 cm -> wavelength monitor (1-10 AA, 200 bins) and
   100 x 100 PSD at its exit.
 * c2s: the c2 guides followed by the c1 sample and monitors (the
   source -> guide -> sample -> monitor chain).
 * scan: 64 instances of a c2s-like chain (Ei and sample rotation scanned,
   uneven neutron counts) for the batched launch.
"""
import os

import numpy as np

from .instrument import instrument as new_instrument

# number density of the Al cell used by the reference's isotropic samples
# (tests/components/samples/sampleassemblies/isotropic_sphere/Al.xyz: 4 atoms in 4.04^3 AA^3)
AL_NUMBER_DENSITY = 4 / (4.04e-10) ** 3
BARN = 1e-28


def isotropic_al_kernel(sigma_scatt_barn=5.0, sigma_abs_barn=2.0):
    from .kernels import IsotropicKernel
    return IsotropicKernel(absorption_coefficient=AL_NUMBER_DENSITY * sigma_abs_barn * BARN,
                           scattering_coefficient=AL_NUMBER_DENSITY * sigma_scatt_barn * BARN)


def _source(E0, dE):
    from .components.sources.source_simple import Source_simple
    return Source_simple('source', radius=0.05, height=0, width=0, dist=10.0, xw=0.1, yh=0.1,
                         E0=E0, dE=dE, flux=1, N=1)


def taper_file(directory=None):
    from .components.optics.ellipse import config2_taper_file
    directory = directory or os.path.join(os.path.dirname(os.path.abspath(__file__)), "_kcache")
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, "taper_6to3cm_100seg.dat")
    if not os.path.exists(path):
        config2_taper_file(path + ".tmp%d" % os.getpid())
        os.replace(path + ".tmp%d" % os.getpid(), path)
    return path


def c1(sample_r=0.01):
    from .components.samples.homogeneous_single_scatterer import factory
    from .components.monitors.psd_monitor import PSD_monitor
    from .components.monitors.iqe_monitor import IQE_monitor
    from .geometry import Sphere
    instr = new_instrument("c1")
    instr.append(_source(60.0, 10.0), position=(0, 0, 0))
    HSS = factory(shape=Sphere(sample_r), kernel=isotropic_al_kernel())
    sample = HSS('sample')
    instr.append(sample, position=(0, 0, 10.0))
    instr.append(IQE_monitor('iqe', Ei=60., Qmin=0., Qmax=12., nQ=120, Emin=-50., Emax=50., nE=100),
                 position=(0, 0, 0), relativeTo=sample)
    instr.append(PSD_monitor('psd', xwidth=0.5, yheight=0.5, nx=250, ny=250), position=(0, 0, 11.0))
    return instr


def _guides(instr):
    from .components.optics.guide import Guide
    from .components.optics.guide_tapering import Guide as TaperedGuide
    instr.append(Guide('guide', w1=0.05, h1=0.05, w2=0.05, h2=0.05, l=10.,
                       R0=0.99, Qc=0.0219, alpha=6.07, m=2, W=0.003), position=(0, 0, 10.0))
    instr.append(TaperedGuide('taper', option="file=" + taper_file(), l=10., mx=2., my=2.,
                              R0=0.99, Qcx=0.0219, Qcy=0.0219, alphax=6.07, alphay=6.07, W=0.003),
                 position=(0, 0, 20.0))


def c2():
    from .components.monitors.psd_monitor import PSD_monitor
    from .components.monitors.wavelength_monitor import Wavelength_monitor
    instr = new_instrument("c2")
    instr.append(_source(52.5, 47.5), position=(0, 0, 0))
    _guides(instr)
    instr.append(Wavelength_monitor('wavelength', Lmin=1., Lmax=10., nchan=200, xwidth=0.06, yheight=0.06),
                 position=(0, 0, 30.001))
    instr.append(PSD_monitor('psd', xwidth=0.06, yheight=0.06, nx=100, ny=100), position=(0, 0, 30.001))
    return instr


def c2s():
    from .components.samples.homogeneous_single_scatterer import factory
    from .components.monitors.psd_monitor import PSD_monitor
    from .components.monitors.iqe_monitor import IQE_monitor
    from .geometry import Sphere
    instr = new_instrument("c2s")
    instr.append(_source(52.5, 47.5), position=(0, 0, 0))
    _guides(instr)
    HSS = factory(shape=Sphere(0.01), kernel=isotropic_al_kernel())
    sample = HSS('sample')
    instr.append(sample, position=(0, 0, 30.5))
    instr.append(IQE_monitor('iqe', Ei=52.5, Qmin=0., Qmax=12., nQ=120, Emin=-50., Emax=50., nE=100),
                 position=(0, 0, 0), relativeTo=sample)
    instr.append(PSD_monitor('psd', xwidth=0.5, yheight=0.5, nx=250, ny=250), position=(0, 0, 31.5))
    return instr


def sqe_ridge_grid(nQ=200, nE=200, Qmin=0.0, Qmax=10.0, Emin=-40.0, Emax=40.0, seed=0):
    """S(Q,E) node values (Q_i = Qmin + i dQ, dQ = (Qmax-Qmin)/nQ; same for E):
    a Gaussian ridge E(Q) = 30 |sin(Q/2)| meV with sigma 2 meV, plus a flat 5%
    background, with 2% multiplicative noise from default_rng(seed)."""
    Q = Qmin + (Qmax - Qmin) / nQ * np.arange(nQ)
    E = Emin + (Emax - Emin) / nE * np.arange(nE)
    QQ, EE = np.meshgrid(Q, E, indexing="ij")
    ridge = np.exp(-(EE - 30.0 * np.abs(np.sin(QQ / 2))) ** 2 / (2 * 2.0 ** 2))
    S = ridge + 0.05 * ridge.max()
    noise = 1 + 0.02 * np.random.default_rng(seed).standard_normal(S.shape)
    return np.clip(S * noise, 0, None), (Qmin, Qmax, Emin, Emax)


def sqe_grid_kernel(**kw):
    from .kernels import SQEGridKernel
    S, (Qmin, Qmax, Emin, Emax) = sqe_ridge_grid(**kw)
    mu = AL_NUMBER_DENSITY * 2.0 * BARN
    sigma = AL_NUMBER_DENSITY * 5.0 * BARN
    return SQEGridKernel(S, Qmin, Qmax, Emin, Emax, absorption_coefficient=mu, scattering_coefficient=sigma)


def phonon_dos_kernel(seed=0, amplitudes=None, centers=(10.0, 25.0, 40.0), widths=None,
                      sigma_scatt_barn=5.0, sigma_abs_barn=2.0):
    """Three Gaussian DOS peaks at 10/25/40 meV, widths U(1, 4) meV, amplitudes
    U(0.5, 1.5), both from default_rng(seed) unless given."""
    from .kernels import PhononDOSKernel
    rng = np.random.default_rng(seed)
    w = rng.uniform(1.0, 4.0, 3)
    a = rng.uniform(0.5, 1.5, 3)
    widths = w if widths is None else widths
    amplitudes = a if amplitudes is None else amplitudes
    mu = AL_NUMBER_DENSITY * sigma_abs_barn * BARN
    sigma = AL_NUMBER_DENSITY * sigma_scatt_barn * BARN
    return PhononDOSKernel(amplitudes, list(centers), widths, 0.0, 12.0, -50.0, 50.0, mu, sigma)


def tube_array(name="detector", ntubes=400, npixels=128, ntof=500, radius=3.0, height=1.28,
               phi_min=5.0, phi_max=135.0, tof_min=2.5e-3, tof_max=6.0e-3):
    from .components.monitors.tube_detector import TubeDetectorArray
    return TubeDetectorArray(name, radius=radius, height=height, phi_min=phi_min, phi_max=phi_max,
                             tof_min=tof_min, tof_max=tof_max, ntubes=ntubes, npixels=npixels, ntof=ntof)


def hollow_cylinder():
    """r_in 1.0 cm, r_out 1.5 cm, h 5 cm; the inner cylinder is taller so the
    hole goes through (axis = local z; the sample is turned to make it vertical)."""
    from .geometry import Cylinder, Difference
    return Difference(Cylinder(0.015, 0.05), Cylinder(0.01, 0.06))


def c3_single(ntubes=400, npixels=128, ntof=500):
    """c1 source -> hollow cylinder with the S(Q,E)-grid kernel, single
    scattering -> tube array."""
    from .components.samples.homogeneous_single_scatterer import factory
    instr = new_instrument("c3-single")
    instr.append(_source(60.0, 10.0), position=(0, 0, 0))
    HSS = factory(shape=hollow_cylinder(), kernel=sqe_grid_kernel())
    instr.append(HSS('sample'), position=(0, 0, 10.0), orientation=(90, 0, 0))
    instr.append(tube_array(ntubes=ntubes, npixels=npixels, ntof=ntof), position=(0, 0, 10.0))
    return instr


def c5_single():
    """c1 source -> sphere with the phonon-DOS kernel -> I(Q,E) monitor."""
    from .components.samples.homogeneous_single_scatterer import factory
    from .components.monitors.iqe_monitor import IQE_monitor
    from .geometry import Sphere
    instr = new_instrument("c5-single")
    instr.append(_source(60.0, 10.0), position=(0, 0, 0))
    HSS = factory(shape=Sphere(0.01), kernel=phonon_dos_kernel())
    s = HSS('sample')
    instr.append(s, position=(0, 0, 10.0))
    instr.append(IQE_monitor('iqe', Ei=60., Qmin=0., Qmax=12., nQ=120, Emin=-50., Emax=50., nE=100,
                             min_angle_in_plane=-180, max_angle_in_plane=180,
                             min_angle_out_of_plane=-90, max_angle_out_of_plane=90),
                 position=(0, 0, 0), relativeTo=s)
    return instr


def c5(Ei=60.0, m=2.0, W=0.003, alpha=6.07, **kernel):
    """The differentiable-fit chain: one scan instance (c2 source at Ei +- 5%,
    straight + tapered guide with a shared coating m/W/alpha) with the
    c1-sized sphere holding the phonon-DOS kernel, then an I(Q,E) monitor over
    all angles at the sample and the 250 x 250 PSD.  `kernel` goes to
    phonon_dos_kernel (amplitudes, centers, widths, cross-sections)."""
    from .components.samples.homogeneous_single_scatterer import factory
    from .components.monitors.psd_monitor import PSD_monitor
    from .components.monitors.iqe_monitor import IQE_monitor
    from .components.optics.guide import Guide
    from .components.optics.guide_tapering import Guide as TaperedGuide
    from .geometry import Sphere
    instr = new_instrument("c5")
    instr.append(_source(Ei, 0.05 * Ei), position=(0, 0, 0))
    instr.append(Guide('guide', w1=0.05, h1=0.05, w2=0.05, h2=0.05, l=10.,
                       R0=0.99, Qc=0.0219, alpha=alpha, m=m, W=W), position=(0, 0, 10.0))
    instr.append(TaperedGuide('taper', option="file=" + taper_file(), l=10., mx=m, my=m,
                              R0=0.99, Qcx=0.0219, Qcy=0.0219, alphax=alpha, alphay=alpha, W=W),
                 position=(0, 0, 20.0))
    HSS = factory(shape=Sphere(0.01), kernel=phonon_dos_kernel(**kernel))
    instr.append(HSS('sample'), position=(0, 0, 30.5))
    instr.append(IQE_monitor('iqe', Ei=Ei, Qmin=0., Qmax=12., nQ=120, Emin=-50., Emax=50., nE=100,
                             min_angle_in_plane=-180, max_angle_in_plane=180,
                             min_angle_out_of_plane=-90, max_angle_out_of_plane=90),
                 position=(0, 0, 30.5))
    instr.append(PSD_monitor('psd', xwidth=0.5, yheight=0.5, nx=250, ny=250), position=(0, 0, 31.5))
    return instr


# theta of the differentiable fit: the nine peak parameters, the two
# cross-sections and the shared guide coating
C5_PARAMS = (["sample.A1", "sample.c1", "sample.w1", "sample.A2", "sample.c2", "sample.w2",
              "sample.A3", "sample.c3", "sample.w3", "sample.scattering_coefficient",
              "sample.absorption_coefficient", ["guide.m", "taper.mx", "taper.my"], ["guide.W", "taper.W"],
              ["guide.alpha", "taper.alphax", "taper.alphay"]])


def c3(ntubes=400, npixels=128, ntof=500, max_ms_loops=2, max_ms_loops_path1=3, max_scattered_neutrons=10):
    """c1 source -> hollow cylinder with the S(Q,E)-grid kernel and multiple
    scattering -> 400 x 128 x 500 tube array.  "Up to 3 scattering orders"
    maps onto the reference scatterer's loop bounds as 3 interactions per
    encounter with the sample (max_ms_loops_path1) and 2 encounters
    (max_ms_loops: the two walls of the hollow cylinder)."""
    from .components.samples.homogeneous_multiple_scatterer import factory
    instr = new_instrument("c3")
    instr.append(_source(60.0, 10.0), position=(0, 0, 0))
    HMS = factory(shape=hollow_cylinder(), kernel=sqe_grid_kernel(), max_scattered_neutrons=max_scattered_neutrons,
                  max_ms_loops=max_ms_loops, max_ms_loops_path1=max_ms_loops_path1)
    instr.append(HMS('sample'), position=(0, 0, 10.0), orientation=(90, 0, 0))
    instr.append(tube_array(ntubes=ntubes, npixels=npixels, ntof=ntof), position=(0, 0, 10.0))
    return instr


def scan_instance(Ei, rotation_deg, name="scan", sample="sphere"):
    """One instance of the batched scan: the c2s chain with source energy Ei
    (uniform Ei +- 5%), the c1 isotropic sample (sample="sphere"; or a
    4 x 4 x 0.5 cm plate, sample="plate", whose path lengths do change with
    the rotation) rotated by `rotation_deg` about the vertical axis, an I(Q,E)
    monitor for that Ei at the sample position (beam frame) and the
    250 x 250 PSD."""
    from .components.samples.homogeneous_single_scatterer import factory
    from .components.monitors.psd_monitor import PSD_monitor
    from .components.monitors.iqe_monitor import IQE_monitor
    from .geometry import Block, Sphere
    instr = new_instrument(name)
    instr.append(_source(Ei, 0.05 * Ei), position=(0, 0, 0))
    _guides(instr)
    shape = {"sphere": lambda: Sphere(0.01), "plate": lambda: Block(0.04, 0.04, 0.005)}[sample]()
    HSS = factory(shape=shape, kernel=isotropic_al_kernel())
    instr.append(HSS('sample'), position=(0, 0, 30.5), orientation=(0, rotation_deg, 0))
    instr.append(IQE_monitor('iqe', Ei=Ei, Qmin=0., Qmax=12., nQ=120, Emin=-50., Emax=50., nE=100),
                 position=(0, 0, 30.5))
    instr.append(PSD_monitor('psd', xwidth=0.5, yheight=0.5, nx=250, ny=250), position=(0, 0, 31.5))
    return instr


def scan_plan(ninst=64, seed=0, nmin=1e5, nmax=1e7):
    """(Ei, rotation, ncount) of each scan instance: Ei log-uniform on
    20-200 meV, rotation uniform on 0-180 deg, counts log-uniform on
    [nmin, nmax], all from numpy.random.default_rng(seed)."""
    g = np.random.default_rng(seed)
    Ei = np.exp(g.uniform(np.log(20.0), np.log(200.0), ninst))
    rot = g.uniform(0.0, 180.0, ninst)
    n = np.exp(g.uniform(np.log(nmin), np.log(nmax), ninst)).astype(np.int64)
    return Ei, rot, n


def scan(ninst=64, seed=0, nmin=1e5, nmax=1e7, sample="sphere"):
    """(instruments, ncounts) of the 64-instance scan."""
    Ei, rot, n = scan_plan(ninst, seed, nmin, nmax)
    return [scan_instance(float(e), float(r), "scan%d" % k, sample)
            for k, (e, r) in enumerate(zip(Ei, rot))], n


BUILDERS = {"c1": c1, "c2": c2, "c2s": c2s, "c3": c3, "c3s": c3_single, "c5s": c5_single, "c5": c5}
