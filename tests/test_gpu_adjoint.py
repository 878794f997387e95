"""Adjoint kernel: g^T dH/dtheta of the monitor histograms against common-
random-number central finite differences of the forward kernel (h = 1e-6
relative) and of the CPU oracle, the dense S(Q,E)-grid gradient, the
torch.autograd wrapper, and a fit of perturbed parameters back to a target.

The adjoint has no reference counterpart; what pins it is that for fixed
random streams the histograms are smooth functions of the weight-carrying
parameters wherever no neutron's fate changes, so finite differences of the
forward simulation (GPU kernel and CPU oracle, each a restatement of the
reference's forward algorithm) give the same derivatives.  At 1e7 neutrons a
few neutrons sit within a 1e-6 step of the supermirror cutoff, where a
reflectivity drops to zero and the neutron dies (and, with dead neutrons
dropped, its stream's later neutrons draw other numbers): the difference
steps are certified by equal monitor counts and cut when a fate flips."""
import copy

import numpy as np
import pytest

from acc_amd import configs
from acc_amd.adjoint import DifferentiableRun
from acc_amd.instrument import calcTransformations

pytestmark = pytest.mark.gpu

H_REL = 1e-6
TOL = 1e-6


def _scale(f0, theta):
    return np.maximum(np.abs(f0) / np.maximum(np.abs(theta), 1e-300), 1e-300)


def _fd(fun, theta, j, h_rel=H_REL):
    h = h_rel * abs(theta[j])
    tp, tm = theta.copy(), theta.copy()
    tp[j] += h
    tm[j] -= h
    return (fun(tp) - fun(tm)) / (tp[j] - tm[j])


def _fd_certified(run, g, theta, j):
    """Central difference of g^T H in theta[j] whose two runs have identical
    monitor counts.  The histograms are smooth in the weight-carrying
    parameters except where a neutron's fate flips: a reflectivity that crosses
    the supermirror cutoff ((Q - m Qc) / W = 10, R -> 0) kills the neutron,
    which then draws no more random numbers, so with dead neutrons dropped the
    rest of its stream runs on other numbers.  Such a flip inside +-h changes
    the counts; the step is then cut tenfold.  Returns (difference, h_rel)."""
    h_rel = H_REL
    for _ in range(4):
        h = h_rel * abs(theta[j])
        tp, tm = theta.copy(), theta.copy()
        tp[j] += h
        tm[j] -= h
        fp = float(g @ run.forward(tp))
        cp = run.counts()
        fm = float(g @ run.forward(tm))
        cm = run.counts()
        if np.array_equal(cp, cm):
            return (fp - fm) / (tp[j] - tm[j]), h_rel
        h_rel /= 10
    raise AssertionError("parameter %d: neutron fates differ at every step down to %g" % (j, h_rel * 10))


@pytest.fixture(scope="module")
def c5run():
    print("\n[adjoint] building c5", flush=True)
    # the configured 1e7 neutrons
    run = DifferentiableRun(configs.c5(), configs.C5_PARAMS, ncount=int(1e7), seed=7)
    yield run


def test_adjoint_matches_gpu_crn_fd(c5run):
    run = c5run
    theta = run.theta0()
    g = np.random.default_rng(1).uniform(0.5, 1.5, sum(run.sizes))
    adj = run.vjp(g, theta)
    f0 = float(g @ run.forward(theta))
    assert f0 > 0
    fd, steps = zip(*[_fd_certified(run, g, theta, j) for j in range(len(theta))])
    fd = np.array(fd)
    run.set_theta(theta)
    rel = np.abs(adj - fd) / np.maximum(np.abs(fd), 1e-9 * _scale(f0, theta))
    for j, p in enumerate(configs.C5_PARAMS):
        print("[adjoint] %-40s adj % .9e fd % .9e (h %.0e) rel %.2e" % (p, adj[j], fd[j], steps[j], rel[j]),
              flush=True)
    assert np.all(rel <= TOL), rel


def test_adjoint_cost(c5run):
    run = c5run
    theta = run.theta0()
    g = np.ones(sum(run.sizes))
    run.forward(theta)
    run.vjp(g, theta)
    fwd, adj = [], []
    for _ in range(3):
        run.forward(theta)
        fwd.append(run.last_forward_stats["ms_kernel"])
        run.vjp(g, theta)
        adj.append(run.last_adjoint_stats["ms_kernel"])
    fwd, adj = min(fwd), min(adj)
    print("[adjoint] forward %.3f ms, adjoint %.3f ms, ratio %.2f" % (fwd, adj, adj / fwd), flush=True)
    assert adj > 0 and fwd > 0


def test_adjoint_matches_oracle_fd():
    """GPU adjoint vs CRN finite differences of the CPU oracle (same N, seed
    and partition, so the same neutrons)."""
    import pyoracle
    instr = configs.c5()
    N, ntot, seed = 20000, 2048, 3
    run = DifferentiableRun(instr, configs.C5_PARAMS, ncount=N, seed=seed, ntotalthreads=ntot)
    theta = run.theta0()
    g = np.random.default_rng(2).uniform(0.5, 1.5, sum(run.sizes))
    adj = run.vjp(g, theta)
    descs0 = [c.mcv_desc() for c in instr.components]
    off, rot = calcTransformations(instr)

    def f(t):
        descs = copy.deepcopy(descs0)
        for v, tl in zip(t, run.params):
            for ci, pidx, _ in tl:
                descs[ci]["params"] = np.array(descs[ci]["params"], dtype=float)
                descs[ci]["params"][pidx] = v
        hists, processed, _, _ = pyoracle.run(descs, off, rot, N, seed, ntotalthreads=ntot, tpb=512, gate=True,
                                              nthreads=1)
        assert processed == N
        return float(g @ np.concatenate([hists[i][1] for i in run.monitors]))

    f0 = f(theta)
    fd = np.array([_fd(f, theta, j) for j in range(len(theta))])
    rel = np.abs(adj - fd) / np.maximum(np.abs(fd), 1e-9 * _scale(f0, theta))
    print("[adjoint] oracle FD max rel %.2e" % rel.max(), flush=True)
    assert np.all(rel <= TOL), rel


def test_dense_grid_gradient():
    """d(g^T H)/d(S node) of an S(Q,E)-grid sample against finite differences
    of the forward kernel with that node perturbed."""
    from acc_amd.prebuilt import grid_sphere

    S0 = configs.sqe_ridge_grid(nQ=40, nE=40)[0]
    build = grid_sphere

    N = 300000
    run = DifferentiableRun(build(S0), ["sample.scattering_coefficient"], ncount=N, seed=5)
    g = np.random.default_rng(3).uniform(0.5, 1.5, sum(run.sizes))
    grad, dense = run.vjp(g, dense=True)
    assert dense.shape == (S0.size,) and np.abs(dense).max() > 0
    order = np.argsort(-np.abs(dense))[:4]
    for node in order:
        # g^T H is linear in each node value (bilinear interpolation, weight
        # proportional to S), so any step is exact up to rounding; a node holds
        # a small share of g^T H, and 1e-2 keeps the difference well above the
        # rounding noise of the sums
        h = 1e-2 * abs(S0.flat[node])
        vals = []
        for sgn in (1, -1):
            S = S0.copy()
            S.flat[node] += sgn * h
            r = DifferentiableRun(build(S), ["sample.scattering_coefficient"], ncount=N, seed=5)
            vals.append(float(g @ r.forward()))
        fd = (vals[0] - vals[1]) / (2 * h)
        print("[adjoint] grid node %d adj %.9e fd %.9e" % (node, dense[node], fd), flush=True)
        assert abs(dense[node] - fd) <= TOL * abs(fd)


def test_autograd_and_fit():
    """torch gradients equal the adjoint's; a fit of all 14 parameters (nine
    peak parameters, the two cross-sections, the shared guide coating m, W,
    alpha) from +-10% perturbed values returns to the target's values (common
    random numbers)."""
    import torch
    from acc_amd.autograd import histograms
    params = configs.C5_PARAMS
    N = 1000000
    target_run = DifferentiableRun(configs.c5(), params, ncount=N, seed=11)
    truth = target_run.theta0()
    target = torch.from_numpy(target_run.forward(truth))
    run = DifferentiableRun(configs.c5(), params, ncount=N, seed=11)
    x = torch.zeros(len(truth), dtype=torch.float64)
    x.data += torch.from_numpy(np.random.default_rng(4).uniform(-0.1, 0.1, len(truth)))
    x.requires_grad_(True)
    t0 = torch.from_numpy(truth)
    scale = float((target ** 2).sum())

    def loss_fn():
        H = histograms(t0 * (1 + x), run)
        return ((H - target) ** 2).sum() / scale

    loss = loss_fn()
    loss.backward()
    # the same gradient from the adjoint directly
    H = run.forward((t0 * (1 + x)).detach().numpy())
    gH = 2 * (H - target.numpy()) / scale
    direct = run.vjp(gH, (t0 * (1 + x)).detach().numpy()) * truth
    np.testing.assert_allclose(x.grad.numpy(), direct, rtol=1e-10, atol=1e-300)
    first = float(loss.detach())
    # the coating, the cross-sections and the peak amplitudes are nearly
    # degenerate in the histograms (sigma_s and A_k enter as products; sigma_a
    # differs from sigma_s only by the 1/v of absorption), so the fit needs a
    # long curvature history
    opt = torch.optim.LBFGS([x], lr=1, max_iter=3000, max_eval=6000, tolerance_grad=0, tolerance_change=0,
                            history_size=100, line_search_fn="strong_wolfe")

    def closure():
        opt.zero_grad()
        lo = loss_fn()
        lo.backward()
        return lo

    opt.step(closure)
    final = float(loss_fn().detach())
    print("[adjoint] fit loss %.3e -> %.3e, max |x| %.2e" % (first, final, float(x.detach().abs().max())), flush=True)
    for name, v in zip(params, x.detach().numpy()):
        print("[adjoint]   %-45s relative offset %.2e" % (name, v), flush=True)
    assert final < 1e-20 * first
    assert float(x.detach().abs().max()) < 1e-8
