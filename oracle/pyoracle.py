"""ctypes wrapper of oracle/liboracle.so -- the CPU restatement of the
reference's propagate path.

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg may import this module.  It is the checker, never the
product path.
"""
import ctypes
import os
import subprocess

import numpy as np

_here = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_here, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int64)
_up = ctypes.POINTER(ctypes.c_uint64)


class CompDesc(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_int32),
        ("n_params", ctypes.c_int32),
        ("params", _dp),
        ("n_iparams", ctypes.c_int32),
        ("iparams", _ip),
        ("n_data", ctypes.c_int64),
        ("data", _dp),
        ("n_bins", ctypes.c_int64),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _here], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.orc_rng_states.argtypes = [ctypes.c_uint64, ctypes.c_int64, _up]
        L.orc_splitmix64.restype = ctypes.c_uint64
        L.orc_splitmix64.argtypes = [ctypes.c_uint64]
        L.orc_rng_draws.argtypes = [_up, ctypes.c_int64, _up, ctypes.POINTER(ctypes.c_float)]
        L.orc_rng_jump.argtypes = [_up]
        L.orc_locate.restype = ctypes.c_int
        L.orc_locate.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_intersect.restype = ctypes.c_int
        L.orc_intersect.argtypes = [_dp, ctypes.c_int] + [ctypes.c_double] * 6 + [_dp]
        L.orc_insert_into_sorted_list.restype = ctypes.c_int
        L.orc_insert_into_sorted_list.argtypes = [ctypes.c_double, _dp, ctypes.c_int, ctypes.c_int]
        L.orc_hss_times.argtypes = [_dp, ctypes.c_int, ctypes.c_double, _dp]
        L.orc_propagation.restype = ctypes.c_double
        L.orc_propagation.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp]
        L.orc_transform.argtypes = [_dp, _dp, _dp]
        L.orc_component.restype = ctypes.c_int
        L.orc_component.argtypes = [ctypes.POINTER(CompDesc), _dp, _up, _dp, _dp, ctypes.c_int]
        L.orc_partition.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int32, _ip]
        L.orc_set_trans_nudge.argtypes = [ctypes.c_int]
        L.orc_set_pinned_nudge.argtypes = [ctypes.c_int64, ctypes.POINTER(ctypes.c_int32), ctypes.c_int64]
        L.orc_process.restype = ctypes.c_int64
        L.orc_process.argtypes = [ctypes.POINTER(CompDesc), ctypes.c_int, _dp, _dp, _dp, ctypes.c_int64,
                                  ctypes.c_uint64, ctypes.c_int64, ctypes.c_int32, ctypes.c_uint32, _dp, _dp,
                                  ctypes.c_int]
        L.orc_run.restype = ctypes.c_int64
        L.orc_run.argtypes = [ctypes.POINTER(CompDesc), ctypes.c_int, _dp, _dp, ctypes.c_uint64, ctypes.c_uint64,
                              ctypes.c_int64, ctypes.c_int32, ctypes.c_uint32, ctypes.c_int64, ctypes.c_int64,
                              ctypes.c_int, _dp, ctypes.c_int64, ctypes.c_int64, _dp,
                              ctypes.POINTER(ctypes.c_int32)]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None and a.size else None


class _Descs:
    def __init__(self, descs):
        self.keep = []
        self.arr = (CompDesc * len(descs))()
        for k, d in enumerate(descs):
            p = np.ascontiguousarray(d["params"], dtype=np.float64)
            ip = np.ascontiguousarray(d.get("iparams", np.zeros(0)), dtype=np.int64)
            data = np.ascontiguousarray(d.get("data", np.zeros(0)), dtype=np.float64)
            self.keep += [p, ip, data]
            self.arr[k].type = int(d["type"])
            self.arr[k].n_params = p.size
            self.arr[k].params = _d(p)
            self.arr[k].n_iparams = ip.size
            self.arr[k].iparams = ip.ctypes.data_as(_ip) if ip.size else None
            self.arr[k].n_data = data.size
            self.arr[k].data = _d(data)
            self.arr[k].n_bins = int(d.get("nbins", 0))
        self.n = len(descs)


# ---------------------------------------------------------------- RNG
def rng_states(seed, n):
    out = np.zeros((int(n), 2), dtype=np.uint64)
    lib().orc_rng_states(int(seed), int(n), out.ctypes.data_as(_up))
    return out


def splitmix64(seed):
    return lib().orc_splitmix64(int(seed))


def rng_draws(state, k):
    """(raw uint64 outputs, float32 draws, new state) of k draws from a state."""
    st = np.array(state, dtype=np.uint64).copy()
    raw = np.zeros(int(k), dtype=np.uint64)
    f = np.zeros(int(k), dtype=np.float32)
    lib().orc_rng_draws(st.ctypes.data_as(_up), int(k), raw.ctypes.data_as(_up),
                        f.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return raw, f, st


def rng_jump(state):
    st = np.array(state, dtype=np.uint64).copy()
    lib().orc_rng_jump(st.ctypes.data_as(_up))
    return st


# ---------------------------------------------------------------- geometry
def locate(nodes, x, y, z):
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    return lib().orc_locate(_d(nodes), nodes.size // 6, x, y, z)


def intersect(nodes, x, y, z, vx, vy, vz):
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    ts = np.zeros(10)
    n = lib().orc_intersect(_d(nodes), nodes.size // 6, x, y, z, vx, vy, vz, _d(ts))
    return ts[:n].copy()


def insert_into_sorted_list(d, l, N):
    l = np.ascontiguousarray(l, dtype=np.float64)
    n = lib().orc_insert_into_sorted_list(float(d), _d(l), int(N), l.size)
    return n, l


def hss_times(ts, ninter, rand):
    ts = np.ascontiguousarray(ts, dtype=np.float64)
    out = np.zeros(5)
    lib().orc_hss_times(_d(ts), int(ninter), float(rand), _d(out))
    return out


PROPAGATION_OPS = {"propagate_out": 0, "tof_before_first_exit": 1, "propagate_to_next_incident_surface": 2,
                   "propagate_to_next_exiting_surface": 3, "is_exiting": 4}


def propagation(op, nodes, neutron):
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    n = np.ascontiguousarray(neutron, dtype=np.float64).copy()
    r = lib().orc_propagation(PROPAGATION_OPS[op], _d(nodes), nodes.size // 6, _d(n))
    return r, n


def transform(neutron, rotmat, offset):
    n = np.ascontiguousarray(neutron, dtype=np.float64).copy()
    R = np.ascontiguousarray(rotmat, dtype=np.float64).reshape(-1)
    o = np.ascontiguousarray(offset, dtype=np.float64)
    lib().orc_transform(_d(n), _d(R), _d(o))
    return n


def component(desc, neutron, state=(0, 0), which=0):
    """Run one component on one neutron; returns (neutron, state, hist, outputs)."""
    D = _Descs([desc])
    n = np.ascontiguousarray(neutron, dtype=np.float64).copy()
    st = np.array(state, dtype=np.uint64).copy()
    hist = np.zeros(3 * max(int(desc.get("nbins", 0)), 1))
    M = 1
    if desc["type"] == 5:  # MCV_SAMPLE_HMS
        M = int(desc["iparams"][2])
    out = np.zeros((max(M, 2), 10))
    ret = lib().orc_component(D.arr, _d(n), st.ctypes.data_as(_up), _d(hist), _d(out), int(which))
    if ret < 0:
        raise ValueError("orc_component does not run component type %d" % desc["type"])
    return n, st, hist, out[:ret] if desc["type"] == 5 else None


def partition(N, ntotalthreads=0, tpb=0):
    out = np.zeros(3, dtype=np.int64)
    lib().orc_partition(int(N), int(ntotalthreads), int(tpb), out.ctypes.data_as(_ip))
    return tuple(int(v) for v in out)


def set_trans_nudge(ulps):
    """Move every transcendental result by `ulps` (sign only) ulp: conditioning probe."""
    lib().orc_set_trans_nudge(int(ulps))


def set_pinned_nudge(nidx=-1, dirs=()):
    """For neutron `nidx` only, move its k-th transcendental result by dirs[k]
    (-1, 0, +1) ulp; nidx=-1 clears it.  Reproduction probe: a device state this
    matches bit for bit differs from the oracle's by exactly those ulps."""
    d = np.ascontiguousarray(dirs, dtype=np.int32)
    lib().orc_set_pinned_nudge(int(nidx), d.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), int(d.size))


def run(descs, offsets, rotmats, N, seed, ntotalthreads=0, tpb=0, gate=True, tid_range=(0, -1), nthreads=0,
        dump=None):
    """Run the chain over streams [tid0, tid1) of the reference partition.

    Returns (hist list of (3, nbins) arrays per component (None for
    non-monitors), processed, dump states, dump alive)."""
    D = _Descs(descs)
    off = np.ascontiguousarray(offsets, dtype=np.float64).reshape(-1)
    rot = np.ascontiguousarray(rotmats, dtype=np.float64).reshape(-1)
    nb = [int(d.get("nbins", 0)) for d in descs]
    total = 3 * sum(nb)
    hist = np.zeros(max(total, 1))
    d0, dc = (0, 0) if dump is None else (int(dump[0]), int(dump[1]))
    dstates = np.zeros((max(dc, 1), 10))
    dalive = np.full(max(dc, 1), -1, dtype=np.int32)
    processed = lib().orc_run(D.arr, D.n, _d(off), _d(rot), int(N), int(seed), int(ntotalthreads), int(tpb),
                              1 if gate else 0, int(tid_range[0]), int(tid_range[1]), int(nthreads), _d(hist),
                              d0, dc, _d(dstates) if dc else None,
                              dalive.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) if dc else None)
    hists, o = [], 0
    for n in nb:
        if n:
            hists.append(hist[o:o + 3 * n].reshape(3, n).copy())
            o += 3 * n
        else:
            hists.append(None)
    return hists, processed, dstates[:dc], dalive[:dc]


def process(descs, offsets, rotmats, neutrons, seed, ntotalthreads=0, tpb=0, gate=True, nthreads=0):
    """Buffered mode: the chain over an (n, 10) array (a copy is returned)
    with the partition of n and streams seeded from `seed`.  Returns
    (neutrons after, hist list, processed); a single multiple-scattering
    sample returns its (n*M, 10) output rows instead of the neutrons."""
    D = _Descs(descs)
    off = np.ascontiguousarray(offsets, dtype=np.float64).reshape(-1)
    rot = np.ascontiguousarray(rotmats, dtype=np.float64).reshape(-1)
    arr = np.array(neutrons, dtype=np.float64, order="C").reshape(-1, 10)
    nb = [int(d.get("nbins", 0)) for d in descs]
    hist = np.zeros(max(3 * sum(nb), 1))
    ms = [d for d in descs if d["type"] == 5]
    out = None
    if ms:
        M = int(np.asarray(ms[0]["iparams"])[2])
        out = np.zeros((arr.shape[0] * M, 10))
    processed = lib().orc_process(D.arr, D.n, _d(off), _d(rot), _d(arr), arr.shape[0], int(seed),
                                  int(ntotalthreads), int(tpb), 1 if gate else 0, _d(hist), _d(out), int(nthreads))
    if processed < 0:
        raise ValueError("a multiple-scattering sample is processed on its own")
    hists, o = [], 0
    for n in nb:
        if n:
            hists.append(hist[o:o + 3 * n].reshape(3, n).copy())
            o += 3 * n
        else:
            hists.append(None)
    return (out if ms else arr), hists, processed


def run_instrument(instrument, N, seed, **kw):
    """Convenience: run an acc_amd Instrument on the oracle."""
    from acc_amd.instrument import calcTransformations
    descs = [c.mcv_desc() for c in instrument.components]
    off, rot = calcTransformations(instrument)
    return run(descs, off, rot, N, seed, **kw)
