"""Sample components and the sample-assembly loader (acc/components/samples/__init__.py).

`loadFirstHomogeneousScatterer(xml)` reads a sampleassembly.xml and the
`<name>-scatterer.xml` of its first sample (mcvine's sampleassembly +
mccomponents parsers in the reference, :4-48): the shape and an
IsotropicKernel with `absorption_coefficient` / `scattering_coefficient`
(1/m).  Coefficients from cross sections and a unit cell are taken when the
kernel gives none: mu = n sigma_abs, sigma = n sigma_scatt with n the number
density of the <xyzfile> cell and the cross sections (barn) from the kernel
element's `absorption_cross_section` / `scattering_cross_section` attributes.
"""
import os
import xml.etree.ElementTree as ET

from ...geometry import shape_from_element, eval_quantity
from ...kernels import IsotropicKernel


class HomogeneousScatterer:
    def __init__(self, shape, kernel, attrs=None):
        self._shape = shape
        self._kernel = kernel
        self.attrs = attrs or {}

    def shape(self):
        return self._shape

    def kernel(self):
        return self._kernel


def _number_density(xyzpath):
    """atoms per m^3 of a .xyz cell file: natoms, lattice vectors (AA), atoms."""
    import numpy as np
    with open(xyzpath) as f:
        lines = [l for l in f.read().splitlines() if l.strip()]
    natoms = int(lines[0].split()[0])
    cell = np.array([float(v) for v in lines[1].split()[:9]]).reshape(3, 3)
    vol = abs(np.linalg.det(cell)) * 1e-30
    return natoms / vol


def loadFirstHomogeneousScatterer(sampleassembly_xml):
    path = os.path.abspath(sampleassembly_xml)
    d = os.path.dirname(path)
    root = ET.parse(path).getroot()
    sample = None
    for el in root:
        if isinstance(el.tag, str) and el.tag not in ("LocalGeometer",):
            sample = el
            break
    if sample is None:
        raise ValueError(f"no sample in {sampleassembly_xml}")
    name = sample.attrib.get("name", "sample")
    shape_el = sample.find("Shape")
    shape = shape_from_element([c for c in shape_el if isinstance(c.tag, str)][0])
    xyz = None
    phase = sample.find("Phase")
    if phase is not None and phase.find("xyzfile") is not None:
        xyz = os.path.join(d, phase.find("xyzfile").text.strip())
    sc = ET.parse(os.path.join(d, f"{name}-scatterer.xml")).getroot()
    kel = [c for c in sc if isinstance(c.tag, str)][0]
    if kel.tag != "IsotropicKernel":
        raise NotImplementedError(f"kernel <{kel.tag}> is not supported by the XML loader")
    a = kel.attrib
    if "absorption_coefficient" in a and "scattering_coefficient" in a:
        mu = eval_quantity(a["absorption_coefficient"])
        sigma = eval_quantity(a["scattering_coefficient"])
    else:
        n = _number_density(xyz)
        barn = 1e-28
        mu = n * float(a["absorption_cross_section"]) * barn
        sigma = n * float(a["scattering_cross_section"]) * barn
    return HomogeneousScatterer(shape, IsotropicKernel(mu, sigma), dict(sc.attrib))
