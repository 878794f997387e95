"""run_script.run / compile / CLI on an instrument script, as the reference's
tests/test_run_script.py does with its sgm script."""
import os
import subprocess
import sys

import numpy as np
import pytest

import pyoracle
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


C1_SCRIPT = """
from acc_amd import configs

def instrument():
    return configs.c1()
"""


def test_saved_histograms_are_the_scaled_sums(tmp_path, monkeypatch):
    """What run() saves is the reference's finalization of the raw sums
    (acc/run_script.py:291-295 saveMonitorOutputs with scale_factor 1/ncount;
    psd_monitor.py:58-64 getHistogram: I = sum p * s, E2 = sum p^2 * s^2;
    iqe_monitor.py:67-70 getHistogram: that histogram divided by the
    Monte-Carlo normalizer, errors propagated as for a quotient), checked
    against the CPU oracle's raw sums for the same neutrons; a run split into
    launches by buffer_size saves the same counts."""
    from acc_amd.components.monitors.MonitorBase import MonitorBase
    from acc_amd.components.monitors.iqe_monitor import IQE_monitor, getNormalization
    from acc_amd.instrument import calcTransformations
    p = tmp_path / "c1_instrument.py"
    p.write_text(C1_SCRIPT)
    ncount, seed, ntot = 300000, 3, 20000
    seen = []
    orig = run_script.saveMonitorOutputs

    def spy(instrument, scale_factor=1.0):
        seen.append((instrument, scale_factor))
        return orig(instrument, scale_factor=scale_factor)

    monkeypatch.setattr(run_script, "saveMonitorOutputs", spy)
    saved, instrs = {}, {}
    for label, buf in (("one", None), ("chunked", 37000)):
        workdir = str(tmp_path / label)
        st = run_script.run(str(p), workdir, ncount, ntotalthreads=ntot, seed=seed, buffer_size=buf)
        streams = -(-ncount // st["npt"])
        assert st["processed"] == ncount
        # c1's chain compiles to a plain kernel: a buffered run is split into
        # slices of max(1, buf // streams) neutrons of every stream
        assert st["nchunks"] == (1 if buf is None else -(-st["npt"] // max(1, buf // streams)))
        instr, sf = seen[-1]
        assert sf == 1.0 / ncount
        mons = [(i, c) for i, c in enumerate(instr.components) if isinstance(c, MonitorBase)]
        saved[label] = {c.name: load(os.path.join(workdir, os.path.splitext(c.filename)[0] + ".npz"))
                        for _, c in mons}
        instrs[label] = instr
    instr = instrs["one"]
    mons = [(i, c) for i, c in enumerate(instr.components) if isinstance(c, MonitorBase)]
    descs = [c.mcv_desc() for c in instr.components]
    off, rot = calcTransformations(instr)
    hists, processed, _, _ = pyoracle.run(descs, off, rot, ncount, seed, ntotalthreads=ntot, tpb=512, gate=True)
    assert processed == ncount
    for i, c in mons:
        raw = hists[i].reshape(c.out.shape)
        assert raw[0].sum() > 100, c.name
        # the monitor's raw sums are the oracle's (counts exact)
        np.testing.assert_array_equal(c.out[0], raw[0])
        np.testing.assert_allclose(c.out[1:], raw[1:], rtol=1e-9, atol=1e-300)
        h = saved["one"][c.name]
        np.testing.assert_array_equal(instrs["chunked"].components[i].out[0], c.out[0])
        np.testing.assert_array_equal(saved["chunked"][c.name].I > 0, h.I > 0)
        if isinstance(c, IQE_monitor):
            n = getNormalization(c)
            a, ea = raw[1] / ncount, raw[2] / ncount ** 2
            exp_I = a / n.I
            exp_E2 = ea / n.I ** 2 + a ** 2 * n.E2 / n.I ** 4
            assert (n.I > 0).all() and (h.I != raw[1]).any()
        else:
            exp_I, exp_E2 = raw[1] / ncount, raw[2] / ncount ** 2
            # exactly the monitor's sums times 1/ncount
            np.testing.assert_array_equal(h.I, c.out[1] * (1.0 / ncount))
            np.testing.assert_array_equal(h.E2, c.out[2] * (1.0 / ncount) * (1.0 / ncount))
        np.testing.assert_allclose(h.I, exp_I, rtol=1e-9, atol=1e-300, err_msg=c.name)
        np.testing.assert_allclose(h.E2, exp_E2, rtol=1e-9, atol=1e-300, err_msg=c.name)
