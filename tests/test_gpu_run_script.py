"""run_script.run / compile / CLI on an instrument script, as the reference's
tests/test_run_script.py does with its sgm script."""
import os
import subprocess
import sys

import numpy as np
import pytest

from acc_amd import run_script
from acc_amd.histogram import load

pytestmark = pytest.mark.gpu

SCRIPT = '''
import acc_amd
from acc_amd.components.sources.source_simple import Source_simple
from acc_amd.components.optics.guide import Guide
from acc_amd.components.monitors.psd_monitor import PSD_monitor
from acc_amd.components.monitors.wavelength_monitor import Wavelength_monitor

def instrument(guide_m=2.):
    instrument = acc_amd.instrument()
    source = Source_simple('source', radius=0.05, dist=10.0, xw=0.1, yh=0.1, E0=52.5, dE=47.5)
    instrument.append(source, position=(0, 0, 0.))
    guide = Guide('guide', w1=0.05, h1=0.05, w2=0.05, h2=0.05, l=10., m=guide_m)
    instrument.append(guide, position=(0, 0, 10.))
    instrument.append(Wavelength_monitor('wavelength', Lmin=1., Lmax=10., nchan=200, xwidth=0.06, yheight=0.06,
                                         filename="IL.h5"), position=(0, 0, 20.001))
    instrument.append(PSD_monitor('psd', xwidth=0.06, yheight=0.06, nx=100, ny=100, filename="psd.h5"),
                      position=(0, 0, 20.001))
    return instrument
'''


@pytest.fixture
def script(tmp_path):
    p = tmp_path / "guide_instrument.py"
    p.write_text(SCRIPT)
    return str(p)


def test_run_writes_scaled_histograms(script, tmp_path):
    workdir = str(tmp_path / "out")
    ncount = 2000000
    stats = run_script.run(script, workdir, ncount=ncount, guide_m=2.)
    assert stats["processed"] == ncount
    psd = load(os.path.join(workdir, "psd.npz"))
    il = load(os.path.join(workdir, "IL.npz"))
    assert psd.I.shape == (100, 100) and il.I.shape == (200,)
    assert psd.I.sum() > 0
    # both monitors sit at the same plane; the wavelength window is narrower
    assert il.I.sum() <= psd.I.sum() * (1 + 1e-12)
    with pytest.raises(RuntimeError):
        run_script.run(script, workdir, ncount=10, overwrite_datafiles=False)


def test_compile_returns_code_object(script):
    path = run_script.compile(script)
    assert path.endswith(".hsaco") and os.path.exists(path)


def test_cli(script, tmp_path):
    workdir = str(tmp_path / "cli")
    r = subprocess.run([sys.executable, "-m", "acc_amd.run_script", script, "--workdir", workdir,
                        "--ncount", "100000"], capture_output=True, text=True,
                       cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    assert r.returncode == 0, r.stderr[-2000:]
    assert os.path.exists(os.path.join(workdir, "psd.npz"))
