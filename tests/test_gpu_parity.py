"""GPU parity: the fused HIP kernels against the oracle on identical streams.

Each chain runs the same neutrons (reference partition, numba-seeded
xoroshiro128+ streams) on the device through the C ABI and on the oracle.
Per neutron, the final state must agree to 1e-12 relative (positions with a
1 mm floor, velocities 1 m/s, time 1 us; weights relative).  A neutron whose
discrete path differs (a branch decided on values equal to the last ulp,
e.g. a t2 < t1 tie) is a "flip": flips must be rare (<= 1e-4 of the
neutrons) and are reported; with no flip, histograms agree bin by bin (counts
exactly, sums to 1e-9 relative: only the atomic summation order differs).
A neutron off by more than 1e-12 (but at most 1e-9) passes only as a
conditioning outlier: the oracle itself, rerun with its transcendentals moved
by one ulp, moves it by at least 1/8 of the deviation.  Their number is
bounded by a few per 1e4 neutrons or 5 % of the neutrons so sensitive.  Flips
are bounded likewise by a few per 1e4 neutrons or a tenth of the flips that
those 1-ulp perturbations cause in the oracle itself (a multiple scatterer in
a box, whose reference locate() has no margin for `inside`, shows thousands).
"""
import numpy as np
import pytest

import pyoracle
from acc_amd import prebuilt
from acc_amd._abi import MCV_FLAG_GATE_DEAD, rng_states
from acc_amd.instrument import calcTransformations
from acc_amd.run_script import get_handle
from parity_tools import TOL, compare_states, flip_budget, ill_conditioned, nudge_flips, outlier_budget

pytestmark = pytest.mark.gpu


def test_stream_seeding_matches_sequential_jumps():
    """Every stream of the device seeding (threads seed up to 8 streams 2^k
    apart) equals the oracle's sequential jump chain, for counts around the
    chunk boundaries and for the reference's default launch."""
    for seed in (0, 1, 12345):
        d = rng_states(seed, 3000)
        o = pyoracle.rng_states(seed, 3000)
        assert np.array_equal(d, o), seed
    for n in (1, 2, 7, 8, 9, 16, 17, 255, 257, 4097):
        assert np.array_equal(rng_states(7, n), pyoracle.rng_states(7, n)), n
    n = 1000448  # streams of the reference's default launch (1954 blocks x 512)
    d = rng_states(1, n)
    o = pyoracle.rng_states(1, n)
    assert np.array_equal(d, o)


CASES = [
    # name, gate, N, ntotalthreads
    ("c2", True, 120000, 2048),
    ("c2", False, 60000, 1024),
    ("c2-straight", True, 100000, 1000000),
    ("c2-straight", False, 100000, 4096),
    ("source-psd", True, 50000, 512),
    ("c1", True, 300000, 8192),
    ("c1", False, 100000, 8192),
    ("c2s", True, 200000, 4096),
    ("c2s", False, 100000, 4096),
    ("scan-plate", True, 200000, 1000000),
    ("c5", True, 150000, 4096),
    ("grid-sphere", True, 150000, 4096),
    ("storage-psd", True, 5000, 512),
    ("aperture-monitors", True, 300000, 8192),
    ("aperture-monitors", False, 100000, 4096),
    ("hollow-cylinder-hss", True, 300000, 8192),
    ("verdi-taper", True, 60000, 2048),
    ("c3s-small", True, 300000, 8192),
    ("c5s", True, 300000, 8192),
    ("c5s", False, 100000, 4096),
    ("csg-union2", True, 200000, 4096),
    ("csg-intersection", True, 200000, 4096),
    ("csg-difference", True, 200000, 4096),
    ("c3-small", True, 100000, 4096),
    # the configured 400 x 128 x 500 tube array (a 614 MB histogram, not replicated)
    ("c3", True, 100000, 8192),
    ("c3-small", False, 50000, 2048),
    ("hms-ref", True, 30000, 2048),
    ("hms-sample", True, 30000, 2048),
    ("hms-sample", False, 20000, 1024),
    ("iqe-normalizer", False, 200000, 4096),
    ("two-samples", True, 200000, 4096),
    ("two-samples", False, 100000, 2048),
    ("three-guides", True, 200000, 4096),
    ("three-guides", False, 100000, 2048),
    ("mixed", True, 200000, 4096),
    ("mixed", False, 100000, 2048),
    ("guide-ms", True, 200000, 4096),
    ("guide-ms", False, 100000, 2048),
    # stored neutrons with spins through the queue (plain and stream-carrying)
    ("storage-guide", True, 50000, 4096),
    ("storage-guide-sample", True, 50000, 4096),
    ("storage-guide-sample", False, 50000, 2048),
    ("storage-ms", True, 50000, 4096),
    ("rotated", True, 200000, 4096),
    ("rotated", False, 100000, 2048),
    ("taper-m0", True, 200000, 4096),
    ("taper-m0", False, 100000, 2048),
    ("taper-mneg", True, 200000, 4096),
    ("hms-sphere", True, 100000, 4096),
    ("hms-plate", True, 100000, 4096),
    ("hms-plate", False, 50000, 2048),
]


@pytest.mark.parametrize("name,gate,N,ntot", CASES)
def test_chain_parity(name, gate, N, ntot):
    print(f"\n[parity] {name} gate={gate} N={N} ntot={ntot}", flush=True)
    instr = prebuilt.instruments()[name]()
    h = get_handle(instr)
    h.set_dump(0, N)
    seed = 0
    flags = MCV_FLAG_GATE_DEAD if gate else 0
    h.clear()
    stats = h.run(N, seed, ntot, 512, flags)
    assert stats["processed"] == N
    g, ga = h.read_dump()
    descs = [c.mcv_desc() for c in instr.components]
    off, rot = calcTransformations(instr)
    hists, processed, o, oa = pyoracle.run(descs, off, rot, N, seed, ntotalthreads=ntot, tpb=512, gate=gate,
                                           dump=(0, N))
    assert processed == N
    nudged = []
    try:
        for mode in (1, -1, 2):
            pyoracle.set_trans_nudge(mode)
            nudged.append(pyoracle.run(descs, off, rot, N, seed, ntotalthreads=ntot, tpb=512, gate=gate,
                                       dump=(0, N))[2:])
    finally:
        pyoracle.set_trans_nudge(0)
    on = [s for s, _ in nudged]
    err, flips, ngood, nout, err_rest = compare_states(g, o, ga, oa, on)
    n_ill = ill_conditioned(o, oa, on)
    n_nf = nudge_flips(o, oa, nudged)
    print(f"[parity] {name}: {ngood} alive neutrons, max err {err:.2e} ({err_rest:.2e} without "
          f"{nout} conditioning outliers; {n_ill} ill-conditioned), flips {flips} "
          f"(the oracle under 1-ulp nudges: {n_nf})", flush=True)
    assert err_rest <= TOL
    assert nout <= outlier_budget(N, n_ill)
    assert err <= 1e-9
    assert flips <= flip_budget(N, n_nf)
    for i, c in enumerate(instr.components):
        if hists[i] is None:
            continue
        gh = h.read_monitor(i)
        oh = hists[i]
        if flips == 0:
            assert np.array_equal(gh[0], oh[0]), c.name
            np.testing.assert_allclose(gh[1], oh[1], rtol=1e-9, atol=1e-300, err_msg=c.name)
            np.testing.assert_allclose(gh[2], oh[2], rtol=1e-9, atol=1e-300, err_msg=c.name)
        else:  # a flipped neutron changes at most its own outputs' counts, on either side
            assert abs(gh[0].sum() - oh[0].sum()) <= 2 * flips * max(1, int(oa.max()), int(ga.max()))
        assert oh[0].sum() > 0, f"{c.name} saw no neutron: test chain is not exercising it"


# Partition edges: one neutron, fewer neutrons than threads, workgroup sizes
# (the reference's threads_per_block) that are not powers of two, the last
# stream short, ntotalthreads below the workgroup size.
PARTITIONS = [(1, 1000000, 512), (7, 1000000, 100), (1000, 333, 128), (12345, 4096, 1000), (100000, 99999, 77),
              (5000, 50, 64)]


@pytest.mark.parametrize("name", ["c1", "c2s", "c3-small"])
@pytest.mark.parametrize("N,ntot,tpb", PARTITIONS)
def test_partition_edges(name, N, ntot, tpb):
    print(f"\n[partition] {name} N={N} ntot={ntot} tpb={tpb}", flush=True)
    instr = prebuilt.instruments()[name]()
    h = get_handle(instr)
    h.set_dump(0, N)
    h.clear()
    stats = h.run(N, 2, ntot, tpb, MCV_FLAG_GATE_DEAD)
    assert stats["processed"] == N
    g, ga = h.read_dump()
    descs = [c.mcv_desc() for c in instr.components]
    off, rot = calcTransformations(instr)
    hists, processed, o, oa = pyoracle.run(descs, off, rot, N, 2, ntotalthreads=ntot, tpb=tpb, gate=True,
                                           dump=(0, N))
    assert processed == N
    nblocks = -(-min(N, ntot) // tpb)
    assert stats["nblocks"] == nblocks and stats["nthreads"] == nblocks * tpb
    assert stats["npt"] == -(-N // (nblocks * tpb))
    nudged = []
    try:
        for mode in (1, -1, 2):
            pyoracle.set_trans_nudge(mode)
            nudged.append(pyoracle.run(descs, off, rot, N, 2, ntotalthreads=ntot, tpb=tpb, gate=True,
                                       dump=(0, N))[2])
    finally:
        pyoracle.set_trans_nudge(0)
    err, flips, ngood, nout, err_rest = compare_states(g, o, ga, oa, nudged)
    assert flips == 0 and err_rest <= TOL and err <= 1e-9, (err, err_rest, flips)
    assert nout <= outlier_budget(N, ill_conditioned(o, oa, nudged))
    for i, c in enumerate(instr.components):
        if hists[i] is not None:
            gh = h.read_monitor(i)
            assert np.array_equal(gh[0], hists[i][0]), c.name
            np.testing.assert_allclose(gh[1:], hists[i][1:], rtol=1e-9, atol=1e-300, err_msg=c.name)
