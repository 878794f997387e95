"""Buffered mode on the GPU: `component.process(neutrons)` and chain
processing against the oracle's buffered mode (same partition of N, same
seed), the reference's NeutronFromStorage expectations
(tests/components/sources/test_neutronfromstorage.py:28-36), and monitors
filling their `out` arrays."""
import os

import numpy as np
import pytest

import pyoracle
from acc_amd import prebuilt
from acc_amd.buffered import process_chain
from acc_amd.instrument import calcTransformations
from parity_tools import TOL, compare_states, ill_conditioned, outlier_budget

pytestmark = pytest.mark.gpu


def _beam(n, seed=0, z=-0.1, spread=0.03):
    rng = np.random.default_rng(seed)
    arr = np.zeros((n, 10))
    arr[:, 0:2] = rng.uniform(-spread, spread, (n, 2))
    arr[:, 2] = z
    arr[:, 3:5] = rng.normal(0, 15, (n, 2))
    arr[:, 5] = rng.uniform(500, 3000, n)
    arr[:, 8] = rng.uniform(0, 1e-3, n)
    arr[:, 9] = rng.uniform(0.5, 1.5, n)
    return arr


def _oracle(instr, arr, seed, ntot=0, gate=True):
    descs = [c.mcv_desc() for c in instr.components]
    off, rot = calcTransformations(instr)
    out, hists, processed = pyoracle.process(descs, off, rot, arr, seed, ntotalthreads=ntot, tpb=512, gate=gate)
    assert processed == len(arr)
    return out, hists


def _check(instr, arr, got, seed, ntot):
    """Per-neutron parity with the oracle's buffered mode, with the
    conditioning probe of the chain parity tests (oracle reruns with every
    transcendental moved by one ulp)."""
    exp, _ = _oracle(instr, arr, seed, ntot)
    nudged = []
    try:
        for mode in (1, -1, 2):
            pyoracle.set_trans_nudge(mode)
            nudged.append(_oracle(instr, arr, seed, ntot)[0])
    finally:
        pyoracle.set_trans_nudge(0)
    ga, oa = (got[:, 9] > 0).astype(np.int32), (exp[:, 9] > 0).astype(np.int32)
    err, flips, ngood, nout, err_rest = compare_states(got, exp, ga, oa, nudged)
    n = len(got)
    n_ill = ill_conditioned(exp, oa, nudged)
    print("[buffered] %d live rows, max err %.2e (%.2e without %d conditioning outliers), flips %d"
          % (ngood, err, err_rest, nout, flips), flush=True)
    assert err_rest <= TOL and err <= 1e-9
    assert nout <= outlier_budget(n, n_ill) and flips <= max(2, 1e-4 * n)
    return exp


@pytest.mark.parametrize("unit,n,ntot", [("guide", 50000, 4096), ("taper", 50000, 4096), ("source", 30000, 2048),
                                         ("c2-chain", 30000, 1000000)])
def test_process_matches_oracle(unit, n, ntot):
    print(f"\n[buffered] {unit} n={n}", flush=True)
    instr = prebuilt.process_units()[unit]()
    arr = _beam(n, seed=3, z=0.0 if unit == "taper" else -0.1, spread=0.025)
    got = process_chain(instr, arr, seed=4, ntotalthreads=ntot, keep_dead=True)
    _check(instr, arr, got, 4, ntot)
    # the public filter keeps exactly the live rows, in order
    kept = process_chain(instr, arr, seed=4, ntotalthreads=ntot)
    np.testing.assert_array_equal(kept, got[got[:, 9] > 0])


def test_component_process_api():
    """component.process(neutrons) is the filtered buffered run of that component."""
    from acc_amd.components.optics.guide import Guide
    g = Guide('g', w1=0.05, h1=0.05, w2=0.03, h2=0.03, l=5.)
    arr = _beam(20000, seed=5)
    out = g.process(arr)
    full = process_chain(prebuilt.single(Guide('g2', w1=0.05, h1=0.05, w2=0.03, h2=0.03, l=5.)), arr, seed=1,
                         keep_dead=True)
    np.testing.assert_array_equal(out, full[full[:, 9] > 0])
    _check(prebuilt.single(Guide('g3', w1=0.05, h1=0.05, w2=0.03, h2=0.03, l=5.)), arr, full, 1, 0)


def test_multiple_scattering_process():
    instr = prebuilt.process_units()["hms"]()
    M = instr.components[0].max_scattered_neutrons
    # across the hollow cylinder (axis z in the sample frame), through its shell
    arr = np.zeros((4000, 10))
    arr[:, 0] = -0.05
    arr[:, 1:3] = np.random.default_rng(6).uniform(-0.02, 0.02, (4000, 2))
    arr[:, 3] = 2000.0
    arr[:, 9] = 1.0
    got = process_chain(instr, arr, seed=2, ntotalthreads=1024, keep_dead=True)
    assert got.shape == (4000 * M, 10)
    exp = _check(instr, arr, got, 2, 1024)
    assert (exp[:, 9] > 0).sum() > 4000


def test_monitor_process_fills_out():
    instr = prebuilt.process_units()["psd"]()
    mon = instr.components[0]
    mon.out[...] = 0
    arr = _beam(40000, seed=7, z=0.0, spread=0.06)
    rows = process_chain(instr, arr, seed=1)
    np.testing.assert_array_equal(rows, arr[arr[:, 9] > 0])  # monitors leave neutrons as they are
    _, hists = _oracle(instr, arr, 1)
    np.testing.assert_array_equal(mon.out[0].reshape(-1), hists[0][0])
    np.testing.assert_allclose(mon.out[1:].reshape(2, -1), hists[0][1:], rtol=1e-12)
    process_chain(instr, arr, seed=1)
    np.testing.assert_array_equal(mon.out[0].reshape(-1), 2 * hists[0][0])  # accumulates


def test_neutron_from_storage(golden_dir):
    """The reference's test_component: processing 10 rows gives the stored
    neutron (0..9) in every row."""
    from acc_amd.components.sources.neutronfromstorage import NeutronFromStorage
    src = NeutronFromStorage('src', os.path.join(golden_dir, "sources", "singletestneutron.mcv"))
    out = src.process(np.zeros((10, 10)))
    np.testing.assert_array_equal(out, np.tile(np.arange(10.), (10, 1)))


def test_storage_chain_run(tmp_path):
    """NeutronFromStorage -> PSD through run_instrument: every neutron lands in
    the centre pixel (the reference's test_component_with_monitors setup)."""
    from acc_amd.run_script import run_instrument
    instr = prebuilt.instruments()["storage-psd"]()
    psd = instr.components[1]
    psd.out[...] = 0
    st = run_instrument(instr, 1000, seed=0, accumulate=False)
    assert st["processed"] == 1000
    I = psd.out[1]
    assert I.sum() == 1000 and I[2, 2] == 1000


def test_slit_beamstop_known_answers():
    """The reference's exhaustive slit/beamstop known answers
    (tests/components/optics/test_slit_beamstop.py:91-140) through
    component.process on the GPU."""
    from slit_kat import cases, neutrons
    from acc_amd.components.optics.slit import Slit
    from acc_amd.components.optics.beamstop import Beamstop
    for name, kw, invert, passes in cases():
        comp = (Beamstop if invert else Slit)("c", **kw)
        start, expected = neutrons(passes)
        got = comp.process(start)
        assert got.shape == expected.shape, name
        assert np.allclose(got, expected, atol=1e-15), name
