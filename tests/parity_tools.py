"""Per-neutron state comparison shared by the GPU parity tests (not a test module)."""
import numpy as np

FLOOR = np.array([1e-3] * 3 + [1.0] * 3 + [1.0, 1.0, 1e-6, 0.0])
TOL = 1e-12


def rel_errors(g, o):
    """Per-neutron max relative error of the 10-vector (floors for components
    near zero: 1 mm, 1 m/s, 1 us; the weight p relative)."""
    scale = np.maximum(np.abs(o), FLOOR)
    scale[:, 9] = np.maximum(np.abs(o[:, 9]), 1e-300)
    with np.errstate(invalid="ignore"):
        rel = np.abs(g - o) / scale
    rel[np.isnan(g) & np.isnan(o)] = 0
    rel = np.where(np.isnan(rel), np.inf, rel)
    return rel


def ill_conditioned(o, oa, o_nudged):
    """How many of the oracle's neutrons with an output move by more than TOL
    when every transcendental is moved by one ulp: the neutrons that a 1-ulp
    difference between host and device libm can put beyond the tolerance."""
    kappa = np.max([rel_errors(on, o).max(axis=1) for on in o_nudged], axis=0)
    return int(((oa >= 1) & (kappa > TOL)).sum())


def nudge_flips(o, oa, nudged):
    """How many neutrons the oracle itself sends down another discrete path
    (other output count, or an O(1) state change) when its transcendentals
    are moved by one ulp; `nudged` holds (states, output counts) per rerun.
    Chains whose branch decisions hang on the last bit of a position (the
    reference's box locate() has no margin for `inside`, so a neutron moved
    onto a face lands inside or on it by rounding) show many; for most chains
    there are none."""
    m = np.zeros(len(o), dtype=bool)
    for on, ona in nudged:
        rel = rel_errors(on, o).max(axis=1)
        rel[(oa == 0) & (ona == 0)] = 0
        m |= (ona != oa) | (rel > 1e-6)
    return int(m.sum())


def flip_budget(n, n_nudge_flips):
    """Discrete path differences a run of n neutrons may show against the
    oracle: a few per 1e4 neutrons, or a tenth of those that 1-ulp libm
    perturbations cause in the oracle itself (host and device libm differ in
    the last bit on a small fraction of calls)."""
    return max(2, 1e-4 * n, 0.1 * n_nudge_flips)


def outlier_budget(n, n_ill):
    """Conditioning outliers a run of n neutrons may show: a few per 1e4
    neutrons, or 5 % of the ill-conditioned ones, where libm results that
    differ in their last bit (a small fraction of the calls) reach them."""
    return max(2, 1e-4 * n, 0.05 * n_ill)


def compare_states(g, o, ga, oa, o_nudged=None):
    """(max error, flips, agreeing, conditioning outliers, max error of the rest).

    A flip is a neutron whose discrete path differs (alive flag or an O(1)
    state difference).  With `o_nudged` (the oracle rerun with every
    transcendental moved by one ulp), a neutron off by more than TOL counts as a
    conditioning outlier when its own oracle state moves by at least 1/8 of
    that error under the nudge: its final state amplifies 1-ulp libm
    differences, and host and device libm legitimately differ by that much."""
    same_alive = ga == oa
    both = same_alive & (oa >= 1)  # >= 1: multiple-scattering chains flag the output count
    rel = rel_errors(g, o)
    per = rel.max(axis=1)
    per[~both] = 0
    flips = (~same_alive) | (per > 1e-6)
    good = both & ~flips
    outliers = np.zeros_like(good)
    if o_nudged is not None:
        kappa = np.max([rel_errors(on, o).max(axis=1) for on in o_nudged], axis=0)
        outliers = good & (per > TOL) & (per <= 8 * kappa)
    rest = good & ~outliers
    if good.any():
        fields = rel[good].max(axis=0)
        worst = np.flatnonzero(good)[np.argmax(per[good])]
        print("[parity] per-field max rel err:", " ".join("%.1e" % v for v in fields))
        print("[parity] worst neutron %d gpu %s" % (worst, np.array2string(g[worst], precision=17)))
        print("[parity] worst neutron %d orc %s" % (worst, np.array2string(o[worst], precision=17)))
        q = np.quantile(per[good], [0.5, 0.99, 0.9999])
        print("[parity] error quantiles 50%% %.1e 99%% %.1e 99.99%% %.1e; conditioning outliers %d" % (
            q[0], q[1], q[2], int(outliers.sum())))
    return ((per[good].max() if good.any() else 0.0), int(flips.sum()), int(good.sum()), int(outliers.sum()),
            (per[rest].max() if rest.any() else 0.0))
