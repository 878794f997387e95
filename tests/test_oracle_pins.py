"""Closed-form pins of the oracle functions the reference's own tests leave
open: the S(Q,E)-grid lookup, the phonon-DOS weight, the tube-detector and
wavelength binning, the I(Q,E) normalizer's neutrons and a reflection off a
sloped wall of the segmented guide.

Each expectation comes from geometry or kinematics computed here from the
neutron's own state (Q and E recomputed from the velocities, hit points from
the ray equation), so that a swapped axis, a shifted node index, a swapped
parameter or a wrong normal fails: none of them retypes the oracle's formula.
Patterns: acc/kernels/SANS2D_ongrid.py:42-55 (the grid lookup),
tests/kernels/test_E_Q.py:25-34 and tests/kernels/test_constant_qe.py:55-65
((Q, E) recovered from the final velocity), acc/components/monitors/
wavelength_monitor.py:52-74, iqe_monitor.py:146-174,
acc/components/optics/guide_tapering.py:189-284."""
import math

import numpy as np
import pytest

import pyoracle as orc
from acc_amd import conversion as C
from acc_amd.components.monitors.tube_detector import TubeDetectorArray
from acc_amd.components.monitors.wavelength_monitor import Wavelength_monitor
from acc_amd.components.optics.guide_tapering import Guide as TaperedGuide
from acc_amd.components.samples.homogeneous_single_scatterer import factory as hss_factory
from acc_amd.geometry import Sphere
from acc_amd.kernels import PhononDOSKernel, SQEGridKernel


def neutron(r, v, t=0.0, p=1.0):
    return np.array(list(r) + list(v) + [0, 0, t, p], dtype=float)


def reflectivity(Q, R0, Qc, alpha, m, W):
    # McStas supermirror reflectivity as a function of Q
    if Q <= Qc:
        return R0
    x = (Q - m * Qc) / W
    if x >= 10:
        return 0.0
    return R0 * (1 - math.tanh(x)) * (1 - alpha * (Q - Qc)) / 2


# ---------------------------------------------------------------- S(Q,E) kernels
R_SPHERE, SIGMA, EI = 0.01, 1e-8, 60.0
QMIN, QMAX, EMIN, EMAX = 0.0, 10.0, -40.0, 40.0


def _scatter(kernel, n, seed):
    """(Q, E, weight) of n neutrons scattered once by a weak sphere, the beam
    through its centre (path 2R); Q and E from the final velocities."""
    S = hss_factory(Sphere(R_SPHERE), kernel)('s').mcv_desc()
    vi = C.e2v(EI)
    st = orc.rng_states(seed, 1)[0]
    out = []
    for _ in range(n):
        f, st, *_ = orc.component(S, neutron((0, 0, -1.0), (0, 0, vi)), st)
        if f[9] == 0:
            continue  # kinematically closed out
        vf = math.sqrt(f[3] ** 2 + f[4] ** 2 + f[5] ** 2)
        Q = C.V2K * math.sqrt(f[3] ** 2 + f[4] ** 2 + (f[5] - vi) ** 2)
        E = C.v2e(vi) - C.v2e(vf)
        out.append((Q, E, f[9], vf))
    return np.array(out), vi


def _weight(S, Q, E, vf, vi):
    """Path x sigma x S(Q,E) x the Jacobian of uniform (Q, E) sampling."""
    ki, kf = C.V2K * vi, C.V2K * vf
    return 2 * R_SPHERE * SIGMA * S * (vf / vi) * Q * (QMAX - QMIN) * (EMAX - EMIN) / (2 * ki * kf)


def test_grid_reproduces_a_bilinear_surface():
    """Node values a + bQ + cE + dQE on a 20 x 16 grid (nodes at Qmin + i dQ,
    Emin + j dE): bilinear interpolation is exact for such a surface, so each
    weight must carry S(Q, E) of the recovered (Q, E).  A transposed grid, a
    node index off by one or an interpolation weight on the wrong axis moves
    S by a whole b dQ or c dE."""
    nQ, nE = 20, 16
    a, b, c, d = 1.0, 0.3, 0.02, 0.005
    Qn = QMIN + np.arange(nQ) * (QMAX - QMIN) / nQ
    En = EMIN + np.arange(nE) * (EMAX - EMIN) / nE
    grid = a + b * Qn[:, None] + c * En[None, :] + d * Qn[:, None] * En[None, :]
    res, vi = _scatter(SQEGridKernel(grid, QMIN, QMAX, EMIN, EMAX, 0.0, SIGMA), 1500, 11)
    assert len(res) > 900
    Q, E, w, vf = res.T
    S = a + b * Q + c * E + d * Q * E
    np.testing.assert_allclose(w, _weight(S, Q, E, vf, vi), rtol=1e-6)
    # the surface varies enough over one cell that a shifted node would show
    assert b * (QMAX - QMIN) / nQ / S.max() > 1e-3 and c * (EMAX - EMIN) / nE / S.max() > 1e-3


def test_grid_delta_node_recovers_its_q_and_e():
    """One non-zero node: the interpolated S is the product of two hat
    functions around that node, so every neutron with weight lies within one
    cell of its (Q, E), and the weight follows the hats."""
    nQ, nE, k0, j0 = 20, 16, 7, 10
    dQ, dE = (QMAX - QMIN) / nQ, (EMAX - EMIN) / nE
    grid = np.zeros((nQ, nE))
    grid[k0, j0] = 1.0
    Q0, E0 = QMIN + k0 * dQ, EMIN + j0 * dE  # 3.5 1/A, 10 meV
    res, vi = _scatter(SQEGridKernel(grid, QMIN, QMAX, EMIN, EMAX, 0.0, SIGMA), 40000, 12)
    Q, E, w, vf = res.T
    hit = w != 0
    assert hit.sum() > 150
    assert np.all(np.abs(Q[hit] - Q0) < dQ * (1 + 1e-9)) and np.all(np.abs(E[hit] - E0) < dE * (1 + 1e-9))
    hat = np.clip(1 - np.abs(Q - Q0) / dQ, 0, None) * np.clip(1 - np.abs(E - E0) / dE, 0, None)
    # (hats near zero amplify the ~1e-9 relative difference between recovering E
    # through VS2E and the sampling through SE2V: absolute floor at 1e-6 of the peak)
    np.testing.assert_allclose(w, _weight(hat, Q, E, vf, vi), rtol=1e-6, atol=1e-6 * w.max())
    qm, em = np.average(Q[hit], weights=w[hit]), np.average(E[hit], weights=w[hit])
    assert abs(qm - Q0) < 0.25 * dQ and abs(em - E0) < 0.25 * dE


@pytest.mark.parametrize("peak", [0, 1, 2])
def test_phonon_dos_single_narrow_peak(peak):
    """One narrow Gaussian peak of the DOS: the weight is Q^2 A exp(-(E-c)^2 /
    (2 w^2)) of the recovered (Q, E) times the sampling Jacobian, and the
    weighted energy transfer sits at the peak centre c."""
    A, c, wdt = [0.0] * 3, [10.0, 25.0, 40.0], [1.0, 2.0, 3.0]
    A[peak] = 2.5
    centre, width = [12.0, 25.0, 33.0][peak], 0.6
    c[peak], wdt[peak] = centre, width
    K = PhononDOSKernel(A, c, wdt, QMIN, QMAX, EMIN, EMAX, 0.0, SIGMA)
    res, vi = _scatter(K, 20000, 20 + peak)
    Q, E, w, vf = res.T
    S = Q ** 2 * 2.5 * np.exp(-0.5 * ((E - centre) / width) ** 2)
    exp_w = _weight(S, Q, E, vf, vi)
    # far out in the tail the exponent amplifies the ~1e-9 relative difference
    # between recovering E through VS2E and sampling through SE2V
    near = np.abs(E - centre) < 4 * width
    assert near.sum() > 200
    np.testing.assert_allclose(w[near], exp_w[near], rtol=1e-6)
    np.testing.assert_allclose(w, exp_w, rtol=1e-6, atol=1e-6 * w.max())
    assert abs(np.average(E, weights=w) - centre) < 0.2


# ---------------------------------------------------------------- tube detector
def _tube_bin(desc, n):
    _, _, h, _ = orc.component(desc, n)
    h = h.reshape(3, -1)
    nz = np.flatnonzero(h[0])
    return h, (int(nz[0]) if nz.size == 1 else (None if nz.size == 0 else -1))


def test_tube_detector_bins_analytic_rays():
    """A ray from a point inside the cylinder reaches radius R at the positive
    root of |(x, z) + (vx, vz) t| = R: its azimuth, height and arrival time
    pick the tube, pixel and time bin (tube-major, then pixel, then time)."""
    R, H, phi0, phi1, t0, t1 = 3.0, 1.28, 5.0, 135.0, 2.5e-3, 6.0e-3
    nt, npix, ntof = 40, 16, 50
    det = TubeDetectorArray('d', radius=R, height=H, phi_min=phi0, phi_max=phi1, tof_min=t0, tof_max=t1,
                            ntubes=nt, npixels=npix, ntof=ntof)
    d = det.mcv_desc()
    dphi, dy, dT = (phi1 - phi0) / nt, H / npix, (t1 - t0) / ntof
    rng = np.random.default_rng(3)
    for _ in range(300):
        j, k, m = rng.integers(nt), rng.integers(npix), rng.integers(ntof)
        phi = math.radians(phi0 + (j + rng.uniform(0.05, 0.95)) * dphi)
        yh = -H / 2 + (k + rng.uniform(0.05, 0.95)) * dy
        tof = t0 + (m + rng.uniform(0.05, 0.95)) * dT
        start = np.array([rng.uniform(-0.3, 0.3), rng.uniform(-0.1, 0.1), rng.uniform(-0.3, 0.3)])
        tstart = rng.uniform(0, 1e-3)
        hit = np.array([R * math.sin(phi), yh, R * math.cos(phi)])
        v = (hit - start) / (tof - tstart)
        p = rng.uniform(0.1, 2.0)
        h, b = _tube_bin(d, neutron(start, v, t=tstart, p=p))
        assert b == (j * npix + k) * ntof + m, (j, k, m)
        assert h[1, b] == p and h[2, b] == p * p
    # outside the height, the azimuth range or the time window: not counted
    for hit, tof in (([0.0, 0.7, R], 4e-3), ([R * math.sin(math.radians(140)), 0, R * math.cos(math.radians(140))],
                                              4e-3), ([0, 0, R], 7e-3), ([0, 0, R], 2e-3)):
        h, b = _tube_bin(d, neutron((0, 0, 0), np.array(hit) / tof))
        assert b is None
    # moving away from the cylinder wall along the axis: no crossing
    h, b = _tube_bin(d, neutron((0, 0, 0), (0, 100.0, 0)))
    assert b is None


def test_tube_detector_edges():
    """Azimuth and time bins include their lower edge and exclude the upper
    one ([phi_min, phi_max), [tof_min, tof_max)); the height window is open at
    both ends.  Rays along +z / +x meet phi = 0 / 90 deg exactly, and a speed
    of 32 m/s over R = 2 m arrives at exactly 1/16 s."""
    def det(**kw):
        a = dict(radius=2.0, height=1.0, phi_min=0.0, phi_max=90.0, tof_min=0.0, tof_max=1.0, ntubes=9,
                 npixels=4, ntof=10)
        a.update(kw)
        return TubeDetectorArray('d', **a).mcv_desc()
    _, b = _tube_bin(det(), neutron((0, 0, 0), (0, 0, 32.0)))  # phi = 0 exactly, y = 0, tof = 1/16
    assert b == (0 * 4 + 2) * 10 + 0
    _, b = _tube_bin(det(), neutron((0, 0, 0), (32.0, 0, 0)))  # phi = 90 exactly
    assert b is None
    _, b = _tube_bin(det(phi_max=90.5), neutron((0, 0, 0), (32.0, 0, 0)))
    assert b == (8 * 4 + 2) * 10 + 0
    _, b = _tube_bin(det(tof_min=0.0625), neutron((0, 0, 0), (0, 0, 32.0)))  # tof = tof_min
    assert b == (0 * 4 + 2) * 10 + 0
    _, b = _tube_bin(det(tof_max=0.0625), neutron((0, 0, 0), (0, 0, 32.0)))  # tof = tof_max
    assert b is None
    for y0 in (-1.5, 0.5):  # arrives at y = +0.5 or -0.5 (= +-height/2) exactly
        vy = 16.0 if y0 < 0 else -16.0
        _, b = _tube_bin(det(), neutron((0, y0, 0), (0, vy, 32.0)))
        assert b is None, y0
    _, b = _tube_bin(det(), neutron((0, -1.4375, 0), (0, 16.0, 32.0)))  # y = -0.4375: pixel 0
    assert b == (0 * 4 + 0) * 10 + 0


# ---------------------------------------------------------------- wavelength monitor
def test_wavelength_binning():
    """lambda = 2 pi / (V2K |v|): a neutron at the centre of channel i is
    counted in channel i; the window and the wavelength range are strict, and
    a neutron that has passed the plane is not counted."""
    Lmin, Lmax, nchan = 1.0, 10.0, 200
    mon = Wavelength_monitor('w', Lmin=Lmin, Lmax=Lmax, nchan=nchan, xwidth=0.06, yheight=0.06)
    d = mon.mcv_desc()
    dL = (Lmax - Lmin) / nchan
    rng = np.random.default_rng(5)
    for i in list(rng.integers(0, nchan, 60)) + [0, nchan - 1]:
        lam = Lmin + (i + rng.uniform(0.02, 0.98)) * dL
        speed = 2 * math.pi / (lam * C.V2K)
        u = np.array([rng.normal(0, 0.01), rng.normal(0, 0.01), 1.0])
        u /= np.linalg.norm(u)
        p = rng.uniform(0.1, 3.0)
        _, _, h, _ = orc.component(d, neutron((rng.uniform(-0.02, 0.02), rng.uniform(-0.02, 0.02), -0.3),
                                              speed * u, p=p))
        h = h.reshape(3, nchan)
        assert h[0].sum() == 1 and h[0, i] == 1, i
        assert h[1, i] == p and h[2, i] == p * p
    for lam, z, x in ((0.5, -0.3, 0.0), (10.5, -0.3, 0.0), (5.0, 0.3, 0.0), (5.0, -0.3, 0.031)):
        speed = 2 * math.pi / (lam * C.V2K)
        _, _, h, _ = orc.component(d, neutron((x, 0, z), (0, 0, speed)))
        assert h.sum() == 0, (lam, z, x)


# ---------------------------------------------------------------- I(Q,E) normalizer neutrons
def test_iqe_normalizer_neutrons():
    """The normalizer's neutrons start at the origin at t = 0 with an energy
    transfer E uniform in [Emin, Emax] (|v_f| = SE2V sqrt(Ei - E)), weight
    v_f / v_i, and isotropic directions (cos theta and phi uniform)."""
    Ei, Emin, Emax = 60.0, -20.0, 45.0  # an asymmetric range: E and -E differ
    desc = dict(type=20, params=np.array([Ei, Emin, Emax]))
    vi = C.e2v(Ei)
    st = orc.rng_states(4, 1)[0]
    E, cos_t, phi = [], [], []
    for _ in range(20000):
        n, st, *_ = orc.component(desc, np.zeros(10), st)
        assert n[0] == n[1] == n[2] == 0 and n[8] == 0
        vf = math.sqrt(n[3] ** 2 + n[4] ** 2 + n[5] ** 2)
        assert n[9] == pytest.approx(vf / vi, rel=1e-12)
        e = Ei - C.v2e(vf)
        assert Emin - 1e-6 <= e <= Emax + 1e-6
        E.append(e)
        cos_t.append(n[5] / vf)
        phi.append(math.atan2(n[4], n[3]) % (2 * math.pi))

    def ks(x, lo, hi):
        x = np.sort((np.array(x) - lo) / (hi - lo))
        return np.max(np.abs(x - np.arange(1, x.size + 1) / x.size))

    lim = 2 * 1.63 / math.sqrt(len(E))
    assert ks(E, Emin, Emax) < lim and ks(cos_t, -1, 1) < lim and ks(phi, 0, 2 * math.pi) < lim


# ---------------------------------------------------------------- sloped wall of the segmented guide
def _one_segment_guide(tmp_path, h1, h2, w1, w2, l, **kw):
    p = tmp_path / "one.dat"
    p.write_text("c one linearly tapering segment\n%r %r %r %r\n0.1 0.1 0.1 0.1\n" % (h1, h2, w1, w2))
    return TaperedGuide('t', option="file=" + str(p), l=l, **kw)


COAT = dict(R0=0.99, mx=2.0, my=3.0, Qcx=0.0219, Qcy=0.02, alphax=6.07, alphay=5.0, W=0.003)


@pytest.mark.parametrize("axis,side", [(0, 1), (0, -1), (1, 1), (1, -1)])
def test_taper_bounce_off_sloped_wall(tmp_path, axis, side):
    """One bounce off a wall of a linearly tapering segment (w 6 -> 5 cm,
    h 5 -> 4 cm over 2 m): the neutron meets the plane side*x = w1/2 + s z
    (s = (w2 - w1) / 2l) where the ray equation says, leaves with its velocity
    mirrored about that wall's tilted normal (|v| kept), at the hit point (the
    guide leaves a neutron at its last reflection), with the weight times
    R(Q), Q = 2 V2K |v.n|, from the coating of that wall's axis."""
    h1, h2, w1, w2, l = 0.05, 0.04, 0.06, 0.05, 2.0
    g = _one_segment_guide(tmp_path, h1, h2, w1, w2, l, **COAT)
    assert g.nseg == 1
    half0, s = ((w1 / 2, (w2 - w1) / (2 * l)) if axis == 0 else (h1 / 2, (h2 - h1) / (2 * l)))
    vz, vt = 2000.0, side * 5.0
    z_in, t_in, p_in = -0.4, 1e-3, 0.7
    r0 = np.zeros(3)
    r0[axis] = side * 0.025 if axis == 0 else side * 0.02  # at z = 0
    v = np.zeros(3)
    v[axis], v[2] = vt, vz
    start = r0 - v * (-z_in / vz)  # so that it crosses z = 0 at r0
    # hit: side*(r0[axis] + vt t) = half0 + s vz t
    t_hit = (half0 - side * r0[axis]) / (side * vt - s * vz)
    assert 0 < t_hit < l / vz
    hit = r0 + v * t_hit
    nrm = np.zeros(3)
    nrm[axis], nrm[2] = side, -s
    nrm /= np.linalg.norm(nrm)
    vdn = v @ nrm
    assert vdn > 0  # moving into the wall
    v_out = v - 2 * vdn * nrm
    Q = 2 * C.V2K * vdn
    R = (reflectivity(Q, COAT["R0"], COAT["Qcx"], COAT["alphax"], COAT["mx"], COAT["W"]) if axis == 0 else
         reflectivity(Q, COAT["R0"], COAT["Qcy"], COAT["alphay"], COAT["my"], COAT["W"]))
    assert 0.05 < R < 0.98  # in the supermirror's sloped range
    out, *_ = orc.component(g.mcv_desc(), neutron(start, v, t=t_in, p=p_in))
    np.testing.assert_allclose(out[:3], hit, rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(out[3:6], v_out, rtol=1e-10, atol=1e-9)
    assert np.linalg.norm(out[3:6]) == pytest.approx(np.linalg.norm(v), rel=1e-13)
    assert out[8] == pytest.approx(t_in + (-z_in / vz) + t_hit, rel=1e-12)
    assert out[9] == pytest.approx(p_in * R, rel=1e-10)
    # the mirrored ray converges on the axis: it leaves through the segment end
    assert abs(hit[axis] + v_out[axis] * (l - hit[2]) / v_out[2]) < half0 + s * l


# ---------------------------------------------------------------- Source_simple draw mapping
def test_source_simple_draw_mapping():
    """The five draws of a stream, in order, fix a Source_simple neutron
    (acc/components/sources/source_simple.py:114-161): r1 the azimuth and r2
    the radius (float32 sqrt) of the start point on the disk, r3 and r4 the
    target point, r5 the energy.  The target point lies at -dx, -dy for
    offsets dx = xw (2 r3 - 1) / 2, dy = yh (2 r4 - 1) / 2: the target-plane
    axes cross((0, 0, d), (0, 1, 0)) / d = (-1, 0, 0) and
    cross((0, 0, d), (-1, 0, 0)) / d = (0, -1, 0).  A mirrored target, a
    different draw order or a float64 radius fails."""
    from acc_amd.components.sources.source_simple import Source_simple
    radius, dist, xw, yh, E0, dE = 0.05, 10.0, 0.1, 0.06, 60.0, 10.0
    src = Source_simple('s', radius=radius, dist=dist, xw=xw, yh=yh, E0=E0, dE=dE)
    d = src.mcv_desc()
    pmul = src.propagate_params[-1]
    st = orc.rng_states(8, 1)[0]
    for _ in range(500):
        _, r, _ = orc.rng_draws(st, 5)
        n, st, *_ = orc.component(d, np.zeros(10), st)
        r = [float(v) for v in r]
        rr = float(np.sqrt(np.float32(r[1]))) * radius
        assert math.hypot(n[0], n[1]) == pytest.approx(rr, rel=1e-13, abs=1e-17)
        if rr > 1e-6:
            assert math.atan2(n[1], n[0]) % (2 * math.pi) == pytest.approx((2 * math.pi * r[0]) % (2 * math.pi),
                                                                             abs=1e-9)
        dx, dy = xw * (r[2] * 2 - 1) / 2.0, yh * (r[3] * 2 - 1) / 2.0
        s = dist / n[5]
        assert n[0] + n[3] * s == pytest.approx(-dx, abs=1e-12)
        assert n[1] + n[4] * s == pytest.approx(-dy, abs=1e-12)
        assert math.sqrt(n[3] ** 2 + n[4] ** 2 + n[5] ** 2) == pytest.approx(C.e2v(E0 + dE * (r[4] * 2 - 1)),
                                                                              rel=1e-13)
        d2 = math.sqrt(dx * dx + dy * dy + dist * dist)
        assert n[9] == pytest.approx(pmul * xw * yh * dist / d2 ** 3, rel=1e-13)
        assert n[2] == 0 and n[8] == 0
