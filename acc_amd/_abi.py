"""ctypes binding of the C ABI in include/mcv_acc.h (libmcvacc.so).

The reference launches Numba kernels directly from Python
(acc/components/sources/SourceBase.py:36-52); this engine crosses into native
code here instead.  Loading fails loudly: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_here = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_here, "lib", "libmcvacc.so")

# include/mcv_acc.h constants
MCV_OK = 0
MCV_ERR_ARG = -1
MCV_ERR_HIP = -2
MCV_ERR_MODULE = -3
MCV_ERR_COUNT = -4
MCV_ERR_NOMEM = -5

MCV_ARM = 0
MCV_SOURCE_SIMPLE = 1
MCV_GUIDE = 2
MCV_GUIDE_TAPERING = 3
MCV_SAMPLE_HSS = 4
MCV_SAMPLE_HMS = 5
MCV_SLIT = 6
MCV_BEAMSTOP = 7
MCV_MON_PSD = 10
MCV_MON_WAVELENGTH = 11
MCV_MON_IQE = 12
MCV_DET_TUBES = 13
MCV_MON_PSD_4PI = 14
MCV_MON_POSDIV = 15
MCV_SRC_IQE_NORMALIZER = 20
MCV_SRC_STORAGE = 21

MCV_KERNEL_ISOTROPIC = 1
MCV_KERNEL_SQE_GRID = 2
MCV_KERNEL_PHONON_DOS = 3

MCV_CSG_NODE_SIZE = 6
MCV_CSG_SPHERE = 1
MCV_CSG_CYLINDER = 2
MCV_CSG_BLOCK = 3
MCV_CSG_UNION = 10
MCV_CSG_DIFFERENCE = 11
MCV_CSG_INTERSECTION = 12

MCV_FLAG_GATE_DEAD = 1

EXPORTED = (
    "mcv_abi_version", "mcv_last_error", "mcv_device_count", "mcv_instr_create",
    "mcv_instr_set_params", "mcv_instr_run", "mcv_monitor_read", "mcv_monitor_clear",
    "mcv_instr_set_dump", "mcv_instr_read_dump", "mcv_rng_states", "mcv_instr_destroy",
    "mcv_batch_create", "mcv_batch_run", "mcv_batch_run_sequential", "mcv_batch_monitor_read",
    "mcv_batch_histograms_read",
    "mcv_batch_histograms_deliver",
    "mcv_batch_clear", "mcv_batch_destroy", "mcv_instr_adjoint", "mcv_instr_process", "mcv_scratch_fit",
    "mcv_instr_run_noalloc",
)
ABI_VERSION = 2


class McvCompDesc(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_int32),
        ("n_params", ctypes.c_int32),
        ("params", ctypes.POINTER(ctypes.c_double)),
        ("n_iparams", ctypes.c_int32),
        ("iparams", ctypes.POINTER(ctypes.c_int64)),
        ("n_data", ctypes.c_int64),
        ("data", ctypes.POINTER(ctypes.c_double)),
        ("n_bins", ctypes.c_int64),
    ]


class McvRunStats(ctypes.Structure):
    _fields_ = [
        ("processed", ctypes.c_uint64),
        ("nblocks", ctypes.c_int64),
        ("nthreads", ctypes.c_int64),
        ("npt", ctypes.c_int64),
        ("ms_rng_init", ctypes.c_double),
        ("ms_kernel", ctypes.c_double),
        ("ms_total", ctypes.c_double),
        ("nchunks", ctypes.c_int64),
        ("scratch_per_lane", ctypes.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class McvError(RuntimeError):
    """Error raised by the native engine (status code + message)."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


_lib = None


def lib():
    """Load libmcvacc.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"native engine {LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(LIB_PATH)
    dp = ctypes.POINTER(ctypes.c_double)
    L.mcv_abi_version.restype = ctypes.c_int
    L.mcv_last_error.restype = ctypes.c_char_p
    L.mcv_device_count.restype = ctypes.c_int
    L.mcv_instr_create.restype = ctypes.c_void_p
    L.mcv_instr_create.argtypes = [ctypes.POINTER(McvCompDesc), ctypes.c_int32, dp, dp,
                                   ctypes.c_char_p, ctypes.c_char_p]
    L.mcv_instr_set_params.restype = ctypes.c_int
    L.mcv_instr_set_params.argtypes = [ctypes.c_void_p, ctypes.c_int32, dp, ctypes.c_int32]
    L.mcv_instr_run.restype = ctypes.c_int
    L.mcv_instr_run.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64,
                                ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(McvRunStats)]
    L.mcv_instr_run_noalloc.restype = ctypes.c_int
    L.mcv_instr_run_noalloc.argtypes = L.mcv_instr_run.argtypes
    L.mcv_monitor_read.restype = ctypes.c_int
    L.mcv_monitor_read.argtypes = [ctypes.c_void_p, ctypes.c_int32, dp]
    L.mcv_monitor_clear.restype = ctypes.c_int
    L.mcv_monitor_clear.argtypes = [ctypes.c_void_p]
    L.mcv_instr_set_dump.restype = ctypes.c_int
    L.mcv_instr_set_dump.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64]
    L.mcv_instr_read_dump.restype = ctypes.c_int
    L.mcv_instr_read_dump.argtypes = [ctypes.c_void_p, dp, ctypes.POINTER(ctypes.c_int32)]
    L.mcv_rng_states.restype = ctypes.c_int
    L.mcv_rng_states.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(ctypes.c_uint64)]
    L.mcv_scratch_fit.restype = ctypes.c_int64
    L.mcv_scratch_fit.argtypes = [ctypes.c_int64, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64]
    L.mcv_instr_destroy.restype = None
    L.mcv_instr_destroy.argtypes = [ctypes.c_void_p]
    L.mcv_batch_create.restype = ctypes.c_void_p
    L.mcv_batch_create.argtypes = [ctypes.POINTER(McvCompDesc), ctypes.c_int32, ctypes.c_int32, dp, dp,
                                   ctypes.c_char_p, ctypes.c_char_p]
    L.mcv_batch_run.restype = ctypes.c_int
    L.mcv_batch_run.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64,
                                ctypes.c_int64, ctypes.c_int32, ctypes.c_uint32, ctypes.c_int32, ctypes.c_int32,
                                ctypes.POINTER(McvRunStats)]
    L.mcv_batch_run_sequential.restype = ctypes.c_int
    L.mcv_batch_run_sequential.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64,
                                           ctypes.c_int64, ctypes.c_int32, ctypes.c_uint32,
                                           ctypes.POINTER(McvRunStats)]
    L.mcv_batch_monitor_read.restype = ctypes.c_int
    L.mcv_batch_monitor_read.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, dp]
    L.mcv_batch_histograms_read.restype = ctypes.c_int
    L.mcv_batch_histograms_read.argtypes = [ctypes.c_void_p, dp, ctypes.c_int64]
    L.mcv_batch_histograms_deliver.restype = ctypes.c_int
    L.mcv_batch_histograms_deliver.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64),
                                               ctypes.POINTER(ctypes.c_int64),
                                               ctypes.POINTER(ctypes.POINTER(ctypes.c_double)), ctypes.c_int32]
    L.mcv_batch_clear.restype = ctypes.c_int
    L.mcv_batch_clear.argtypes = [ctypes.c_void_p]
    L.mcv_batch_destroy.restype = None
    L.mcv_batch_destroy.argtypes = [ctypes.c_void_p]
    L.mcv_instr_process.restype = ctypes.c_int
    L.mcv_instr_process.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, dp, ctypes.c_int64,
                                    ctypes.c_uint64, ctypes.c_int64, ctypes.c_int32, ctypes.c_uint32, dp,
                                    ctypes.c_int32, ctypes.POINTER(McvRunStats)]
    L.mcv_instr_adjoint.restype = ctypes.c_int
    L.mcv_instr_adjoint.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint64,
                                    ctypes.c_uint64, ctypes.c_int64, ctypes.c_int32, ctypes.c_uint32, dp,
                                    ctypes.c_int32, dp, ctypes.c_int64, dp, ctypes.POINTER(McvRunStats)]
    if L.mcv_abi_version() != ABI_VERSION:
        raise ImportError("libmcvacc.so ABI version mismatch")
    _lib = L
    return L


def check(rc):
    if rc != MCV_OK:
        msg = lib().mcv_last_error().decode()
        if rc == MCV_ERR_COUNT:
            raise AssertionError(msg)
        raise McvError(rc, msg)
    return rc


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None and a.size else None


def _iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)) if a is not None and a.size else None


class CompDescs:
    """Keeps the numpy arrays behind an array of McvCompDesc alive."""

    def __init__(self, descs):
        self._keep = []
        arr = (McvCompDesc * len(descs))()
        for k, d in enumerate(descs):
            p = np.ascontiguousarray(d["params"], dtype=np.float64)
            ip = np.ascontiguousarray(d["iparams"], dtype=np.int64)
            data = np.ascontiguousarray(d["data"], dtype=np.float64)
            self._keep += [p, ip, data]
            arr[k].type = int(d["type"])
            arr[k].n_params = p.size
            arr[k].params = _dptr(p)
            arr[k].n_iparams = ip.size
            arr[k].iparams = _iptr(ip)
            arr[k].n_data = data.size
            arr[k].data = _dptr(data)
            arr[k].n_bins = int(d["nbins"])
        self.array = arr
        self.n = len(descs)


class Handle:
    """An mcv_instr: device-resident instrument with one fused kernel."""

    def __init__(self, descs, offsets, rotmats, code_object, kernel_name):
        L = lib()
        self._descs = CompDescs(descs)
        self._offsets = np.ascontiguousarray(offsets, dtype=np.float64).reshape(-1)
        self._rotmats = np.ascontiguousarray(rotmats, dtype=np.float64).reshape(-1)
        h = L.mcv_instr_create(self._descs.array, self._descs.n, _dptr(self._offsets), _dptr(self._rotmats),
                               code_object.encode(), kernel_name.encode())
        if not h:
            raise McvError(MCV_ERR_ARG, L.mcv_last_error().decode())
        self.h = h
        self.nbins = [int(d["nbins"]) for d in descs]
        # histogram layout: (3, nbins) blocks of the components with bins, in chain order
        self.hoff = list(np.cumsum([0] + [3 * nb for nb in self.nbins])[:-1])
        self.hist_total = int(3 * sum(self.nbins))

    def run(self, ncount, seed, ntotalthreads=0, threads_per_block=0, flags=MCV_FLAG_GATE_DEAD, buffer_size=0):
        """mcv_instr_run; buffer_size > 0 splits the run into launches of whole
        streams of about that many neutrons (results as one launch)."""
        st = McvRunStats()
        check(lib().mcv_instr_run(self.h, int(ncount), int(seed), int(ntotalthreads or 0),
                                  int(threads_per_block or 0), int(flags), int(buffer_size or 0), ctypes.byref(st)))
        return st.as_dict()

    def run_noalloc(self, ncount, seed, ntotalthreads=0, threads_per_block=0, flags=MCV_FLAG_GATE_DEAD,
                    buffer_size=0):
        """mcv_instr_run_noalloc: as run(), with no device allocation."""
        st = McvRunStats()
        check(lib().mcv_instr_run_noalloc(self.h, int(ncount), int(seed), int(ntotalthreads or 0),
                                          int(threads_per_block or 0), int(flags), int(buffer_size or 0),
                                          ctypes.byref(st)))
        return st.as_dict()

    def set_params(self, comp, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        check(lib().mcv_instr_set_params(self.h, int(comp), _dptr(p), p.size))

    def read_monitor(self, comp, out=None):
        """Monitor comp's (3, nbins) histogram: a new array, or written into
        `out` (a writeable C-contiguous float64 array of 3 x nbins values)."""
        if out is None:
            out = np.empty(3 * self.nbins[comp])
        elif (out.dtype != np.float64 or not out.flags.c_contiguous or not out.flags.writeable
              or out.size != 3 * self.nbins[comp]):
            raise McvError(MCV_ERR_ARG, "monitor %d: out must be a writeable C-contiguous float64 array of %d "
                           "values" % (comp, 3 * self.nbins[comp]))
        check(lib().mcv_monitor_read(self.h, int(comp), _dptr(out)))
        return out.reshape(3, -1)

    def clear(self):
        check(lib().mcv_monitor_clear(self.h))

    def set_dump(self, first, count):
        check(lib().mcv_instr_set_dump(self.h, int(first), int(count)))
        self._dump_count = int(count)

    def adjoint(self, code_object, kernel_name, ncount, seed, g_hist, nscore, ntotalthreads=0,
                threads_per_block=0, flags=MCV_FLAG_GATE_DEAD, ndense=0):
        """(grad[nscore], dense gradient or None, run stats) of mcv_instr_adjoint."""
        g = np.ascontiguousarray(g_hist, dtype=np.float64).reshape(-1)
        if g.size != self.hist_total:
            raise ValueError("upstream gradient has %d values, histograms %d" % (g.size, self.hist_total))
        grad = np.zeros(int(nscore))
        dense = np.zeros(int(ndense)) if ndense else None
        st = McvRunStats()
        check(lib().mcv_instr_adjoint(self.h, code_object.encode(), kernel_name.encode(), int(ncount), int(seed),
                                      int(ntotalthreads or 0), int(threads_per_block or 0), int(flags), _dptr(g),
                                      int(nscore), _dptr(grad), int(ndense),
                                      _dptr(dense) if dense is not None else None, ctypes.byref(st)))
        return grad, dense, st.as_dict()

    def process(self, code_object, kernel_name, neutrons, seed, m=0, ntotalthreads=0, threads_per_block=0,
                flags=MCV_FLAG_GATE_DEAD):
        """mcv_instr_process: (rows after, or the (n*m, 10) outputs when m > 0; run stats)."""
        arr = np.array(neutrons, dtype=np.float64, order="C").reshape(-1, 10)
        out = np.zeros((arr.shape[0] * m, 10)) if m else None
        st = McvRunStats()
        check(lib().mcv_instr_process(self.h, code_object.encode(), kernel_name.encode(), _dptr(arr), arr.shape[0],
                                      int(seed), int(ntotalthreads or 0), int(threads_per_block or 0), int(flags),
                                      _dptr(out) if out is not None else None, int(m), ctypes.byref(st)))
        return (out if m else arr), st.as_dict()

    def read_dump(self):
        states = np.zeros((self._dump_count, 10))
        alive = np.zeros(self._dump_count, dtype=np.int32)
        check(lib().mcv_instr_read_dump(self.h, _dptr(states),
                                        alive.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return states, alive

    def close(self):
        if getattr(self, "h", None):
            lib().mcv_instr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rng_states(seed, n, first=0):
    """Device-seeded (n, 2) uint64 states of streams [first, first + n)
    (numba create_xoroshiro128p_states)."""
    out = np.zeros((int(n), 2), dtype=np.uint64)
    check(lib().mcv_rng_states(int(seed), int(first), int(n), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
    return out


def scratch_fit(scratch_per_lane, block, resident_groups, budget_bytes):
    """mcv_scratch_fit: workgroups that may be resident within a scratch budget."""
    return int(lib().mcv_scratch_fit(int(scratch_per_lane), int(block), int(resident_groups), int(budget_bytes)))


def device_count():
    return lib().mcv_device_count()


class BatchHandle:
    """An mcv_batch: instances of one chain structure, one persistent launch."""

    def __init__(self, descs_per_instance, offsets, rotmats, code_object, kernel_name):
        L = lib()
        ninst = len(descs_per_instance)
        ncomp = len(descs_per_instance[0])
        flat = [d for inst in descs_per_instance for d in inst]
        self._descs = CompDescs(flat)
        self._offsets = np.ascontiguousarray(offsets, dtype=np.float64).reshape(-1)
        self._rotmats = np.ascontiguousarray(rotmats, dtype=np.float64).reshape(-1)
        h = L.mcv_batch_create(self._descs.array, ninst, ncomp, _dptr(self._offsets), _dptr(self._rotmats),
                               code_object.encode(), kernel_name.encode())
        if not h:
            raise McvError(MCV_ERR_ARG, L.mcv_last_error().decode())
        self.h = h
        self.ninst, self.ncomp = ninst, ncomp
        self.nbins = [int(d["nbins"]) for d in descs_per_instance[0]]

    def run(self, ncounts, seed, ntotalthreads=0, threads_per_block=0, flags=MCV_FLAG_GATE_DEAD, grid=0,
            use_graph=True):
        n = np.ascontiguousarray(ncounts, dtype=np.uint64)
        st = McvRunStats()
        check(lib().mcv_batch_run(self.h, n.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), int(seed),
                                  int(ntotalthreads or 0), int(threads_per_block or 0), int(flags), int(grid),
                                  1 if use_graph else 0, ctypes.byref(st)))
        return st.as_dict()

    def run_sequential(self, ncounts, seed, ntotalthreads=0, threads_per_block=0, flags=MCV_FLAG_GATE_DEAD):
        n = np.ascontiguousarray(ncounts, dtype=np.uint64)
        st = McvRunStats()
        check(lib().mcv_batch_run_sequential(self.h, n.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                             int(seed), int(ntotalthreads or 0), int(threads_per_block or 0),
                                             int(flags), ctypes.byref(st)))
        return st.as_dict()

    def read_monitor(self, inst, comp):
        out = np.zeros(3 * self.nbins[comp])
        check(lib().mcv_batch_monitor_read(self.h, int(inst), int(comp), _dptr(out)))
        return out.reshape(3, -1)

    def read_histograms(self):
        """(ninst, H) array: every instance's (3, nbins) monitor blocks in chain order, one copy."""
        out = np.empty((self.ninst, 3 * sum(self.nbins)))
        check(lib().mcv_batch_histograms_read(self.h, _dptr(out), out.size))
        return out

    def deliver_histograms(self, blocks, accumulate=False):
        """Writes (or adds) histogram blocks into arrays: blocks is a list of
        (offset, dst) with offset in read_histograms' flattened layout and dst a
        C-contiguous float64 array receiving dst.size doubles."""
        nb = len(blocks)
        src = (ctypes.c_int64 * nb)(*[int(o) for o, _ in blocks])
        ln = (ctypes.c_int64 * nb)()
        dst = (ctypes.POINTER(ctypes.c_double) * nb)()
        for j, (_, d) in enumerate(blocks):
            if d.dtype != np.float64 or not d.flags.c_contiguous or not d.flags.writeable:
                raise McvError(MCV_ERR_ARG, "block %d: destination must be a writeable C-contiguous float64 array" % j)
            ln[j] = d.size
            dst[j] = _dptr(d)
        check(lib().mcv_batch_histograms_deliver(self.h, nb, src, ln, dst, 1 if accumulate else 0))

    def clear(self):
        check(lib().mcv_batch_clear(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().mcv_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
