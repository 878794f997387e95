"""Device deviations traced to single libm ulps.

Over other seeds and partitions, a few neutrons differ from the oracle by a
few 1e-12 while the oracle's uniform one-ulp nudges (every transcendental up,
every one down, or a hashed sign per call) do not move them: the amplifying
path depends on which call differs and in which direction.  For the two such
neutrons found, moving one or two transcendental results of that neutron's
history by one ulp in the oracle reproduces the device's final state bit for bit, so
those deviations are exactly a host-vs-device libm difference of one ulp in
that call, and not a difference in the arithmetic of the kernels.

Each case runs the chain on the device, then the oracle on the neutron's own
stream with and without the pinned nudge.  The device state must equal one of
the two bit for bit (the plain oracle too would mean the two libms now agree
on that call)."""
import numpy as np
import pytest

import pyoracle
from acc_amd import prebuilt
from acc_amd._abi import MCV_FLAG_GATE_DEAD
from acc_amd.instrument import calcTransformations
from acc_amd.run_script import get_handle
from parity_tools import TOL, rel_errors

pytestmark = pytest.mark.gpu

# (chain, neutron index, N, reference thread count, seed, per-call ulp moves)
# csg-union2: call 1 is the source's sin(chi) (acc/components/sources/
# source_simple.py:114-161), which alone moves p by 4.26e-12, and call 3 the
# sample's sin(phi) (acc/kernels/isotropic.py:5-21), which moves the position
# by an ulp; hms-ref: call 3 is cos(phi) of the first scattering's direction.
CASES = [("csg-union2", 146706, 200000, 1000000, 123456789, (0, +1, 0, +1)),
         ("hms-ref", 13424, 30000, 6161, 123456789, (0, 0, 0, -1))]


@pytest.mark.parametrize("name,idx,N,ntot,seed,dirs", CASES, ids=[c[0] for c in CASES])
def test_deviation_is_one_libm_ulp(name, idx, N, ntot, seed, dirs):
    print(f"\n[ulp-repro] {name} neutron {idx} N={N} ntot={ntot} seed={seed}", flush=True)
    instr = prebuilt.instruments()[name]()
    h = get_handle(instr)
    h.set_dump(idx, 1)
    h.clear()
    st = h.run(N, seed, ntot, 512, MCV_FLAG_GATE_DEAD)
    assert st["processed"] == N
    g, ga = h.read_dump()
    descs = [c.mcv_desc() for c in instr.components]
    off, rot = calcTransformations(instr)
    npt = pyoracle.partition(N, ntot, 512)[2]
    tid = idx // npt
    kw = dict(ntotalthreads=ntot, tpb=512, gate=True, tid_range=(tid, tid + 1), dump=(idx, 1))
    _, _, o, oa = pyoracle.run(descs, off, rot, N, seed, **kw)
    try:
        pyoracle.set_pinned_nudge(idx, dirs)
        _, _, op, opa = pyoracle.run(descs, off, rot, N, seed, **kw)
    finally:
        pyoracle.set_pinned_nudge()
    assert ga[0] == oa[0] == opa[0] >= 1
    print("[ulp-repro] gpu %s" % np.array2string(g[0], precision=17), flush=True)
    print("[ulp-repro] orc %s (max rel err %.2e)" % (np.array2string(o[0], precision=17),
                                                    rel_errors(g, o).max()), flush=True)
    # the pinned ulp is what moves this neutron: by more than the parity tolerance
    assert rel_errors(op, o).max() > TOL
    assert np.array_equal(g[0], op[0]) or np.array_equal(g[0], o[0])
