"""Neutron files (.mcv) and the oracle's buffered mode, on the CPU.

The .mcv fixtures are the reference's own
(tests/components/sources/singletestneutron*.mcv, written by
create_neutron_storage_for_test.py:12-17 with r=(0,1,2), v=(3,4,5), s=(6,7),
t=8, p=9 and r=0, v=(0,0,1000), p=1); the expected rows are those the
reference's test_neutronfromstorage.py:28-36 asserts."""
import os

import numpy as np

from acc_amd import neutron_storage as ns


def test_reads_reference_fixtures(golden_dir):
    a = ns.read(os.path.join(golden_dir, "sources", "singletestneutron.mcv"))
    np.testing.assert_array_equal(a, [np.arange(10.)])
    b = ns.read(os.path.join(golden_dir, "sources", "singletestneutron2.mcv"))
    np.testing.assert_array_equal(b, [[0, 0, 0, 0, 0, 1000, 0, 0, 0, 1]])


def test_writes_reference_layout(golden_dir, tmp_path):
    p = str(tmp_path / "one.mcv")
    ns.write(p, np.arange(10.).reshape(1, 10))
    with open(p, "rb") as f, open(os.path.join(golden_dir, "sources", "singletestneutron.mcv"), "rb") as g:
        assert f.read() == g.read()


def test_round_trip_and_append(tmp_path):
    p = str(tmp_path / "many.mcv")
    rng = np.random.default_rng(0)
    a = rng.normal(size=(1000, 10))
    ns.write(p, a)
    np.testing.assert_array_equal(ns.read(p), a)
    b = rng.normal(size=(17, 10))
    ns.append(p, b)
    np.testing.assert_array_equal(ns.read(p), np.vstack([a, b]))
    st = ns.storage(str(tmp_path / "s.mcv"), "w")
    st.write(a[:5])
    st.write(a[5:9])
    st.close()
    np.testing.assert_array_equal(ns.storage(str(tmp_path / "s.mcv")).read(), a[:9])


def test_rejects_bad_files(tmp_path):
    import pytest
    p = tmp_path / "bad.mcv"
    p.write_bytes(b"x" * 1200)
    with pytest.raises(ValueError):
        ns.read(str(p))
    ns.write(str(p), np.zeros((2, 10)))
    with open(p, "ab") as f:
        f.write(b"\0" * 8)
    with pytest.raises(ValueError):
        ns.read(str(p))


def test_oracle_process_matches_component_by_component():
    """The oracle's buffered mode over N neutrons equals the unit-level
    component entry point applied neutron by neutron (deterministic guide)."""
    import pyoracle
    from acc_amd.components.optics.guide import Guide
    g = Guide('g', w1=0.05, h1=0.05, w2=0.03, h2=0.03, l=5.)
    d = g.mcv_desc()
    rng = np.random.default_rng(1)
    n = 2000
    arr = np.zeros((n, 10))
    arr[:, 0:2] = rng.uniform(-0.03, 0.03, (n, 2))
    arr[:, 3:5] = rng.normal(0, 20, (n, 2))
    arr[:, 5] = rng.uniform(500, 3000, n)
    arr[:, 2] = -0.1
    arr[:, 9] = 1.0
    out, _, processed = pyoracle.process([d], np.zeros((0, 3)), np.zeros((0, 3, 3)), arr, seed=1)
    assert processed == n
    for i in range(n):
        ref = pyoracle.component(d, arr[i])[0]
        np.testing.assert_array_equal(out[i], ref)


def test_oracle_storage_source_replays_rows(golden_dir):
    import pyoracle
    from acc_amd.components.sources.neutronfromstorage import NeutronFromStorage
    src = NeutronFromStorage('src', os.path.join(golden_dir, "sources", "singletestneutron.mcv"))
    out, _, processed = pyoracle.process([src.mcv_desc()], np.zeros((0, 3)), np.zeros((0, 3, 3)),
                                         np.zeros((7, 10)), seed=1)
    assert processed == 7
    np.testing.assert_array_equal(out, np.tile(np.arange(10.), (7, 1)))


def test_oracle_slit_beamstop_known_answers():
    """The reference's slit/beamstop known answers through the oracle's
    buffered mode (the same check runs on the GPU in test_gpu_buffered.py)."""
    import pyoracle
    from slit_kat import cases, neutrons
    from acc_amd.components.optics.slit import Slit
    from acc_amd.components.optics.beamstop import Beamstop
    for name, kw, invert, passes in cases():
        comp = (Beamstop if invert else Slit)("c", **kw)
        start, expected = neutrons(passes)
        out, _, processed = pyoracle.process([comp.mcv_desc()], np.zeros((0, 3)), np.zeros((0, 3, 3)), start,
                                             seed=1)
        assert processed == len(start)
        got = out[out[:, 9] > 0]
        assert got.shape == expected.shape, name
        assert np.allclose(got, expected, atol=1e-15), name
