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


BUILDERS = {"c1": c1, "c2": c2, "c2s": c2s}
