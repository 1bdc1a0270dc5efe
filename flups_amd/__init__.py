"""Python host interface of the MI355X FLUPS solve path.

Mirrors the reference's Python binding (python_api/pyflups/flups.py:23-500:
``Topology``, ``Solver`` and the enums) over the C interface of
``libflups.so`` (include/flups.h). Arrays passed to ``Solver.solve`` may be
numpy arrays (host memory, copied over PCIe inside the call), ``DeviceArray``
buffers allocated by the library in HBM (the fast path), or torch tensors on
the GPU (import torch before the first solver so that its HIP runtime starts
first). There is no CPU fallback: importing works without a GPU, but solver
construction needs the HIP library and a GPU.
"""
import ctypes
import enum
import os
from ctypes import POINTER, byref, c_bool, c_char_p, c_double, c_int, c_int64, c_long, c_size_t, c_void_p

import numpy as np

__all__ = ["BoundaryType", "GreenType", "SolverType", "DiffType", "CenterType", "Topology", "Solver", "Profiler",
           "DeviceArray", "locID", "print_data", "lib", "lib_path", "last_error", "FLUPS_ALIGNMENT", "FLUPS_FORWARD",
           "FLUPS_BACKWARD", "FLUPS_BACKWARD_DIFF"]

FLUPS_FORWARD = -1
FLUPS_BACKWARD = 1
FLUPS_BACKWARD_DIFF = 2
FLUPS_ALIGNMENT = 8


class BoundaryType(enum.IntEnum):
    EVEN = 0
    ODD = 1
    PER = 3
    UNB = 4
    NONE = 9


class GreenType(enum.IntEnum):
    CHAT_2 = 0
    LGF_2 = 1
    HEJ_2 = 2
    HEJ_4 = 3
    HEJ_6 = 4
    HEJ_8 = 5
    HEJ_10 = 6
    HEJ_0 = 7
    LGF_4 = 8
    LGF_6 = 9
    LGF_8 = 10
    MEHR_4L = 11
    MEHR_6L = 12
    MEHR_4F = 13
    MEHR_6F = 14


class SolverType(enum.IntEnum):
    STD = 0
    ROT = 1


class DiffType(enum.IntEnum):
    NOD = 0
    SPE = 1
    FD2 = 2
    FD4 = 4
    FD6 = 6


class CenterType(enum.IntEnum):
    NODE_CENTER = 0
    CELL_CENTER = 1


_HERE = os.path.dirname(os.path.abspath(__file__))
lib_path = os.path.join(_HERE, "libflups.so")
_lib = None


def _init_torch_first():
    """PyTorch-ROCm bundles its own HIP runtime; in a process that uses both, torch's
    runtime must claim the device before libflups' does, so when torch is already
    imported it is initialised first."""
    import sys
    torch = sys.modules.get("torch")
    if torch is None:
        return
    try:
        torch.cuda.is_available()
    except Exception:
        pass


def lib():
    """Load libflups.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        _init_torch_first()
        if not os.path.exists(lib_path):
            raise RuntimeError("libflups.so is missing: build it with `make -C flups_amd` "
                               "(or __graft_entry__.build())")
        L = ctypes.CDLL(lib_path)
        L.flups_topo_new.restype = c_void_p
        L.flups_topo_new.argtypes = [c_int, c_int, POINTER(c_int), POINTER(c_int), c_bool, POINTER(c_int), c_int, c_int64]
        L.flups_topo_free.argtypes = [c_void_p]
        for name in ["flups_topo_get_nglob", "flups_topo_get_nloc", "flups_topo_get_nmem", "flups_topo_get_nproc"]:
            getattr(L, name).argtypes = [c_void_p, c_int]
            getattr(L, name).restype = c_int
        for name in ["flups_topo_get_lda", "flups_topo_get_axis"]:
            getattr(L, name).argtypes = [c_void_p]
            getattr(L, name).restype = c_int
        L.flups_topo_get_isComplex.argtypes = [c_void_p]
        L.flups_topo_get_isComplex.restype = c_bool
        L.flups_topo_get_istartGlob.argtypes = [c_void_p, POINTER(c_int)]
        L.flups_topo_get_memsize.argtypes = [c_void_p]
        L.flups_topo_get_memsize.restype = c_size_t
        L.flups_topo_get_locsize.argtypes = [c_void_p]
        L.flups_topo_get_locsize.restype = c_size_t
        L.flups_init_timed.restype = c_void_p
        L.flups_init_timed.argtypes = [c_void_p, POINTER(POINTER(c_int)), POINTER(c_double), POINTER(c_double), c_int,
                                       POINTER(c_int), c_void_p]
        L.flups_init.restype = c_void_p
        L.flups_init.argtypes = [c_void_p, POINTER(POINTER(c_int)), POINTER(c_double), POINTER(c_double), c_int,
                                 POINTER(c_int)]
        L.flups_cleanup_solver.argtypes = [c_void_p]
        L.flups_set_greenType.argtypes = [c_void_p, c_int]
        L.flups_set_alpha.argtypes = [c_void_p, c_double]
        L.flups_set_stream.argtypes = [c_void_p, c_void_p]
        L.flups_set_precision.argtypes = [c_void_p, c_int]
        L.flups_setup.argtypes = [c_void_p, c_bool]
        L.flups_solve.argtypes = [c_void_p, c_void_p, c_void_p, c_int]
        L.flups_solve_noalloc.argtypes = [c_void_p, c_void_p, c_void_p, c_int]
        L.flups_get_allocSize.argtypes = [c_void_p]
        L.flups_get_allocSize.restype = c_size_t
        L.flups_get_spectralInfo.argtypes = [c_void_p, POINTER(c_double), POINTER(c_double), POINTER(c_double)]
        L.flups_get_innerBuffer.restype = c_void_p
        L.flups_get_innerBuffer.argtypes = [c_void_p]
        for name in ["flups_get_innerTopo_physical", "flups_get_innerTopo_spectral"]:
            getattr(L, name).restype = c_void_p
            getattr(L, name).argtypes = [c_void_p]
        L.flups_skip_firstSwitchtopo.argtypes = [c_void_p]
        L.flups_do_copy.argtypes = [c_void_p, c_void_p, c_void_p, c_int]
        L.flups_do_FFT.argtypes = [c_void_p, c_void_p, c_int]
        L.flups_do_mult.argtypes = [c_void_p, c_void_p, c_int]
        L.flups_switchtopo_info.argtypes = [c_void_p]
        L.flups_set_autotune.argtypes = [c_void_p, c_int]
        L.flups_get_schedule.argtypes = [c_void_p, POINTER(c_int)]
        L.flups_get_solve_bytes.restype = c_double
        L.flups_get_solve_bytes.argtypes = [c_void_p, c_int]
        L.flups_get_path.argtypes = [c_void_p]
        L.flups_get_path.restype = c_int
        L.flups_get_pass_kernels.argtypes = [c_void_p, POINTER(c_int), c_int]
        L.flups_get_pass_kernels.restype = c_int
        L.flups_get_last_solve_ns.argtypes = [c_void_p]
        L.flups_get_last_solve_ns.restype = c_double
        L.flups_get_last_error.restype = c_char_p
        L.flups_set_abort_on_error.argtypes = [c_int]
        L.flups_device_malloc.restype = c_void_p
        L.flups_device_malloc.argtypes = [c_size_t]
        L.flups_device_free.argtypes = [c_void_p]
        L.flups_memcpy.argtypes = [c_void_p, c_void_p, c_size_t]
        L.flups_memcpy.restype = c_int
        L.flups_pointer_is_device.argtypes = [c_void_p]
        L.flups_pointer_is_device.restype = c_int
        L.flups_stream_create.restype = c_void_p
        L.flups_stream_destroy.argtypes = [c_void_p]
        L.flups_stream_synchronize.argtypes = [c_void_p]
        L.flups_stream_synchronize.restype = c_int
        L.flups_server_new.restype = c_void_p
        L.flups_server_new.argtypes = [c_size_t, c_int, c_int]
        L.flups_server_free.argtypes = [c_void_p]
        L.flups_server_submit.restype = c_long
        L.flups_server_submit.argtypes = [c_void_p, POINTER(ServeRequest), c_void_p, c_void_p]
        L.flups_server_wait.restype = c_int
        L.flups_server_wait.argtypes = [c_void_p, c_long, POINTER(c_double)]
        L.flups_server_drain.argtypes = [c_void_p]
        L.flups_server_stats.argtypes = [c_void_p, POINTER(c_double)]
        L.flups_server_reset_stats.argtypes = [c_void_p]
        L.flups_server_request_bytes.restype = c_size_t
        L.flups_server_request_bytes.argtypes = [POINTER(ServeRequest)]
        L.flups_batch_new.restype = c_void_p
        L.flups_batch_new.argtypes = [c_int, POINTER(ServeRequest), POINTER(c_void_p), POINTER(c_void_p)]
        L.flups_batch_free.argtypes = [c_void_p]
        L.flups_batch_run.restype = c_int
        L.flups_batch_run.argtypes = [c_void_p]
        L.flups_batch_run_branches.restype = c_int
        L.flups_batch_run_branches.argtypes = [c_void_p]
        L.flups_batch_launches.restype = c_int
        L.flups_batch_launches.argtypes = [c_void_p, c_int]
        L.flups_batch_run_sequential.restype = c_int
        L.flups_batch_run_sequential.argtypes = [c_void_p]
        L.flups_batch_last_ms.restype = c_double
        L.flups_batch_last_ms.argtypes = [c_void_p]
        L.flups_batch_device_bytes.restype = c_size_t
        L.flups_batch_device_bytes.argtypes = [c_void_p]
        L.flups_profiler_new.restype = c_void_p
        L.flups_profiler_new.argtypes = []
        L.flups_profiler_new_n.restype = c_void_p
        L.flups_profiler_new_n.argtypes = [c_char_p]
        L.flups_profiler_free.argtypes = [c_void_p]
        L.flups_profiler_disp.argtypes = [c_void_p]
        L.flups_profiler_get.restype = c_int
        L.flups_profiler_get.argtypes = [c_void_p, c_int, c_char_p, c_int, POINTER(c_double), POINTER(c_double),
                                         POINTER(c_long)]
        L.flups_print_data.argtypes = [c_void_p, c_void_p]
        L.flups_topo_cmpt_rank_fromid.argtypes = [c_void_p, c_int, c_int]
        L.flups_topo_cmpt_rank_fromid.restype = c_int
        L.flups_topo_cmpt_start_id_from_rank.argtypes = [c_void_p, c_int, c_int]
        L.flups_topo_cmpt_start_id_from_rank.restype = c_int
        L.flups_topo_get_comm.argtypes = [c_void_p]
        L.flups_topo_get_comm.restype = c_int64
        L.flups_topo_ranksplit.argtypes = [c_void_p, c_int, POINTER(c_int)]
        L.flups_topo_rankindex.argtypes = [c_void_p, POINTER(c_int)]
        L.flups_topo_rankindex.restype = c_int
        _lib = L
    return _lib


class ServeRequest(ctypes.Structure):
    """one whole problem for the serving loop / batches (FLUPS_ServeRequest)"""
    _fields_ = [("nglob", c_int * 3), ("center", c_int), ("bc", (c_int * 2) * 3), ("lda", c_int),
                ("h", c_double * 3), ("L", c_double * 3), ("green", c_int), ("alpha", c_double), ("fp32", c_int),
                ("type", c_int), ("order_diff", c_int)]

    @classmethod
    def make(cls, nglob, bc, h, L, green=GreenType.HEJ_2, lda=1, center=CenterType.CELL_CENTER, fp32=False,
             solver_type=SolverType.STD, order_diff=DiffType.NOD, alpha=0.0):
        r = cls()
        for d in range(3):
            r.nglob[d] = int(nglob[d])
            r.h[d] = float(h[d])
            r.L[d] = float(L[d])
            for side in range(2):
                r.bc[d][side] = int(bc[d][side])
        r.center, r.lda, r.green, r.alpha = int(center), int(lda), int(green), float(alpha)
        r.fp32, r.type, r.order_diff = int(bool(fp32)), int(solver_type), int(order_diff)
        return r

    def device_bytes(self):
        """device memory the serving loop reserves to admit this request when its solver is not cached"""
        return lib().flups_server_request_bytes(ctypes.byref(self))

    def shape(self):
        """caller array shape: (lda, n2, n1, n0), dense"""
        return (self.lda, self.nglob[2], self.nglob[1], self.nglob[0])


SERVER_STATS = ("completed", "failed", "p50_ms", "p99_ms", "mean_ms", "solves_per_s", "cache_hits", "cache_misses",
                "evictions", "peak_device_bytes", "device_bytes", "cached_solvers")


class Server:
    """serving loop of the library: worker threads, solver cache keyed by the request, HBM-budget admission"""

    def __init__(self, hbm_budget_bytes, workers=4, solvers_per_key=1):
        self._lib = lib()   # kept for the release at collection (also at interpreter exit)
        self._s = lib().flups_server_new(int(hbm_budget_bytes), int(workers), int(solvers_per_key))
        _check(self._s, "flups_server_new")

    def submit(self, req, rhs, field):
        t = lib().flups_server_submit(self._s, ctypes.byref(req), _data_ptr(rhs), _data_ptr(field))
        _check(t > 0, "flups_server_submit")
        return t

    def wait(self, ticket):
        lat = c_double()
        _check(lib().flups_server_wait(self._s, ticket, ctypes.byref(lat)), "flups_server_wait")
        return lat.value

    def drain(self):
        lib().flups_server_drain(self._s)

    def stats(self):
        out = (c_double * 12)()
        lib().flups_server_stats(self._s, out)
        return dict(zip(SERVER_STATS, list(out)))

    def reset_stats(self):
        lib().flups_server_reset_stats(self._s)

    def close(self):
        if getattr(self, "_s", None):
            self._lib.flups_server_free(self._s)
            self._s = None

    def __del__(self):
        self.close()


class Batch:
    """independent problems on device buffers, set up once, solved by one hipGraph launch of grouped
    kernels (flups_batch_run) or of one branch per problem (run_branches)"""

    def __init__(self, reqs, rhs, fields):
        self._lib = lib()   # kept for the release at collection (also at interpreter exit)
        n = len(reqs)
        self._reqs = (ServeRequest * n)(*reqs)
        self._rhs = (c_void_p * n)(*[_data_ptr(a) for a in rhs])
        self._fields = (c_void_p * n)(*[_data_ptr(a) for a in fields])
        self._b = lib().flups_batch_new(n, self._reqs, self._rhs, self._fields)
        _check(self._b, "flups_batch_new")

    def run(self):
        """grouped launches (one launch per kernel and step over the problems); device ms"""
        _check(lib().flups_batch_run(self._b), "flups_batch_run")
        return lib().flups_batch_last_ms(self._b)

    def run_branches(self):
        """one graph branch per problem with its own launches; device ms"""
        _check(lib().flups_batch_run_branches(self._b), "flups_batch_run_branches")
        return lib().flups_batch_last_ms(self._b)

    def launches(self, grouped=True):
        """kernel launches of one run (grouped) or of one run_branches"""
        return lib().flups_batch_launches(self._b, 1 if grouped else 0)

    def run_sequential(self):
        _check(lib().flups_batch_run_sequential(self._b), "flups_batch_run_sequential")

    def device_bytes(self):
        return lib().flups_batch_device_bytes(self._b)

    def close(self):
        if getattr(self, "_b", None):
            self._lib.flups_batch_free(self._b)
            self._b = None

    def __del__(self):
        self.close()


def last_error():
    return lib().flups_get_last_error().decode()


def _check(ok, what):
    if not ok:
        raise RuntimeError("%s failed: %s" % (what, last_error()))


class Topology:
    """User topology (python_api/pyflups/flups.py:23-183); one rank, nproc = (1,1,1)."""

    def __init__(self, axis, lda, nglob, nproc=(1, 1, 1), is_complex=False, axproc=None,
                 alignment=FLUPS_ALIGNMENT, comm=0):
        self._lib = lib()   # kept for the release at collection (also at interpreter exit)
        L = lib()
        L.flups_clear_last_error()
        ax = (c_int * 3)(*axproc) if axproc is not None else None
        self._topo = L.flups_topo_new(int(axis), int(lda), (c_int * 3)(*nglob), (c_int * 3)(*nproc), bool(is_complex),
                                      ax, int(alignment), int(comm))
        _check(self._topo, "flups_topo_new")

    def __del__(self):
        if getattr(self, "_topo", None) and getattr(self, "_owned", True):
            self._lib.flups_topo_free(self._topo)
        self._topo = None

    @property
    def is_complex(self):
        return bool(lib().flups_topo_get_isComplex(self._topo))

    @property
    def lda(self):
        return lib().flups_topo_get_lda(self._topo)

    @property
    def axis(self):
        return lib().flups_topo_get_axis(self._topo)

    def get_nglob(self, dim):
        return lib().flups_topo_get_nglob(self._topo, dim)

    def get_nloc(self, dim):
        return lib().flups_topo_get_nloc(self._topo, dim)

    def get_nmem(self, dim):
        return lib().flups_topo_get_nmem(self._topo, dim)

    def get_nproc(self, dim):
        return lib().flups_topo_get_nproc(self._topo, dim)

    def get_istart(self):
        a = (c_int * 3)()
        lib().flups_topo_get_istartGlob(self._topo, a)
        return np.array(a[:])

    @property
    def locsize(self):
        return lib().flups_topo_get_locsize(self._topo)

    @property
    def memsize(self):
        return lib().flups_topo_get_memsize(self._topo)

    def cmpt_rank_from_id(self, global_id, direction):
        """rank owning global index global_id along direction (always 0 on one GPU)"""
        return lib().flups_topo_cmpt_rank_fromid(self._topo, int(global_id), int(direction))

    def cmpt_start_id_from_rank(self, rank_id, direction):
        return lib().flups_topo_cmpt_start_id_from_rank(self._topo, int(rank_id), int(direction))

    def get_comm(self):
        """the communicator handle given at construction (kept, never used on one GPU)"""
        return lib().flups_topo_get_comm(self._topo)

    def ranksplit(self, rank):
        out = (c_int * 3)()
        lib().flups_topo_ranksplit(self._topo, int(rank), out)
        return np.array(out[:])

    def rankindex(self, rankd):
        return lib().flups_topo_rankindex(self._topo, (c_int * 3)(*[int(r) for r in rankd]))

    @property
    def _ptr(self):
        return self._topo

    @classmethod
    def _own(cls, handle, solver=None):
        """wrap a topology handle owned by a solver (the inner topologies): not freed by the wrapper,
        which keeps the solver alive"""
        _check(handle, "inner topology")
        t = cls.__new__(cls)
        t._lib = lib()
        t._topo = handle
        t._owned = False
        t._solver_ref = solver
        return t


FORWARD, BACKWARD, BACKWARD_DIFF = -1, 1, 2   # flups_interface.h signs


class DeviceView:
    """float64 device memory owned by the library (e.g. the solver's inner buffer)"""

    def __init__(self, ptr, n, owner):
        self.ptr, self.shape, self.nbytes, self._owner = int(ptr), (int(n),), int(n) * 8, owner

    def numpy(self):
        out = np.empty(self.shape)
        _check(lib().flups_memcpy(out.ctypes.data, self.ptr, self.nbytes), "flups_memcpy")
        return out

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        if a.size > self.shape[0]:
            raise ValueError("array larger than the device buffer")
        _check(lib().flups_memcpy(self.ptr, a.ctypes.data, a.size * 8), "flups_memcpy")

    def data_ptr(self):
        return self.ptr


class DeviceArray:
    """float64 array in HBM allocated by libflups (shape in C order, dimension 0 of FLUPS last)."""

    def __init__(self, shape):
        self._lib = lib()   # kept for the release at collection (also at interpreter exit)
        self.shape = tuple(int(s) for s in shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        self.ptr = lib().flups_device_malloc(self.nbytes)
        _check(self.ptr, "flups_device_malloc")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.shape)
        _check(lib().flups_memcpy(d.ptr, a.ctypes.data, d.nbytes), "flups_memcpy")
        return d

    def numpy(self):
        out = np.empty(self.shape)
        _check(lib().flups_memcpy(out.ctypes.data, self.ptr, self.nbytes), "flups_memcpy")
        return out

    def data_ptr(self):
        return self.ptr

    def __del__(self):
        if getattr(self, "ptr", None):
            self._lib.flups_device_free(self.ptr)
            self.ptr = None


def _data_ptr(a):
    """device pointer (DeviceArray / torch tensor) or host pointer (numpy array)"""
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]):
        raise TypeError("arrays must be C-contiguous float64 numpy arrays or torch tensors")
    return a.ctypes.data


class Stream:
    """HIP stream of libflups' runtime (flups_stream_create), for Solver.set_stream."""

    def __init__(self):
        self._lib = lib()   # kept for the release at collection (also at interpreter exit)
        self.handle = lib().flups_stream_create()
        _check(self.handle, "flups_stream_create")

    def synchronize(self):
        _check(lib().flups_stream_synchronize(self.handle), "flups_stream_synchronize")

    def __del__(self):
        if getattr(self, "handle", None):
            self._lib.flups_stream_destroy(self.handle)
            self.handle = None


class Profiler:
    """Timers of the solvers created with it (python_api/pyflups/flups.py:185-218): init, setup and
    every solve with its stages (HIP events on the solve stream)."""

    def __init__(self, name="default"):
        self._lib = lib()   # kept for the release at collection (also at interpreter exit)
        L = lib()
        self._prof = L.flups_profiler_new() if name == "default" else L.flups_profiler_new_n(name.encode("utf-8"))
        _check(self._prof, "flups_profiler_new")

    def __del__(self):
        if getattr(self, "_prof", None):
            self._lib.flups_profiler_free(self._prof)
            self._prof = None

    def display(self):
        """print the timer table (flups_profiler_disp)"""
        lib().flups_profiler_disp(self._prof)

    def timers(self):
        """{name: (calls, host seconds, device milliseconds)} of every timer recorded so far"""
        L = lib()
        out = {}
        n = L.flups_profiler_get(self._prof, 0, None, 0, None, None, None)
        for i in range(max(n, 0)):
            buf = ctypes.create_string_buffer(256)
            hs, dm, calls = c_double(), c_double(), c_long()
            L.flups_profiler_get(self._prof, i, buf, 256, byref(hs), byref(dm), byref(calls))
            out[buf.value.decode()] = (calls.value, hs.value, dm.value)
        return out

    @property
    def _ptr(self):
        return self._prof


class Solver:
    """Poisson solver (python_api/pyflups/flups.py:221-500) on one MI355X."""

    def __init__(self, topology, boundary_conditions, h, L, order_diff=DiffType.NOD, center_type=None,
                 profiler=None):
        self._lib = lib()   # kept for the release at collection (also at interpreter exit)
        Lb = lib()
        Lb.flups_clear_last_error()
        self._topology = topology
        lda = topology.lda
        self._bc_arrays = []
        for i in range(3):
            for j in range(2):
                v = boundary_conditions[i][j]
                vals = list(v) if hasattr(v, "__len__") else [v] * lda
                self._bc_arrays.append((c_int * lda)(*[int(x) for x in vals]))
        self._bc_c = (POINTER(c_int) * 6)(*[ctypes.cast(a, POINTER(c_int)) for a in self._bc_arrays])
        if center_type is None:
            center_type = [CenterType.CELL_CENTER] * 3
        self._profiler = profiler   # kept alive as long as the solver records into it
        self._solver = Lb.flups_init_timed(topology._ptr, self._bc_c, (c_double * 3)(*h), (c_double * 3)(*L),
                                           int(order_diff), (c_int * 3)(*[int(c) for c in center_type]),
                                           profiler._ptr if profiler is not None else None)
        _check(self._solver, "flups_init")

    def __del__(self):
        if getattr(self, "_solver", None):
            self._lib.flups_cleanup_solver(self._solver)
            self._solver = None

    def set_green_type(self, green_type):
        lib().flups_set_greenType(self._solver, int(green_type))

    def set_alpha(self, alpha):
        lib().flups_set_alpha(self._solver, float(alpha))

    def set_stream(self, stream):
        """run on `stream` (a Stream, a raw hipStream_t of the library's runtime, or None for the
        default stream). With a stream, solve() on device arrays returns once the work is queued;
        synchronise the stream before reading the result."""
        handle = stream.handle if isinstance(stream, Stream) else stream
        self._stream = stream   # kept alive while the solver uses it
        lib().flups_set_stream(self._solver, handle)

    def set_autotune(self, on=True):
        lib().flups_clear_last_error()
        lib().flups_set_autotune(self._solver, 1 if on else 0)
        _check(not lib().flups_get_last_error(), "flups_set_autotune")

    def solve_bytes(self, solver_type=SolverType.STD):
        """minimum HBM bytes of one solve (roofline numerator)"""
        return lib().flups_get_solve_bytes(self._solver, int(solver_type))

    @property
    def schedule(self):
        """(z pass, y passes, x order) schedule of the fused pipeline; -1s on the general path"""
        out = (c_int * 3)()
        lib().flups_get_schedule(self._solver, out)
        return tuple(out)

    def set_precision(self, fp32):
        lib().flups_set_precision(self._solver, 1 if fp32 else 0)

    def setup(self, change_comm=False):
        Lb = lib()
        Lb.flups_clear_last_error()
        Lb.flups_setup(self._solver, bool(change_comm))
        err = last_error()
        if err:
            raise RuntimeError("flups_setup failed: " + err)

    def solve(self, field, rhs, solver_type=SolverType.STD):
        Lb = lib()
        Lb.flups_clear_last_error()
        Lb.flups_solve(self._solver, _data_ptr(field), _data_ptr(rhs), int(solver_type))
        err = last_error()
        if err:
            raise RuntimeError("flups_solve failed: " + err)
        return field

    def get_alloc_size(self):
        return lib().flups_get_allocSize(self._solver)

    # ---- advanced interface (python_api/pyflups/flups.py, src/flups.h:467-578) ----
    def inner_buffer(self):
        L = lib()
        L.flups_clear_last_error()
        p = L.flups_get_innerBuffer(self._solver)
        _check(p, "flups_get_innerBuffer")
        return DeviceView(p, self.get_alloc_size(), self)

    def inner_topo_physical(self):
        return Topology._own(lib().flups_get_innerTopo_physical(self._solver), self)

    def inner_topo_spectral(self):
        return Topology._own(lib().flups_get_innerTopo_spectral(self._solver), self)

    # pyflups names (python_api/pyflups/flups.py:439-470): the inner buffer lives in HBM, so it is
    # returned as a device view (numpy() / upload() copy it) rather than a host numpy array
    def get_inner_buffer(self):
        return self.inner_buffer()

    def get_inner_topo_physical(self):
        return self.inner_topo_physical()

    def get_inner_topo_spectral(self):
        return self.inner_topo_spectral()

    def skip_first_switchtopo(self):
        lib().flups_skip_firstSwitchtopo(self._solver)

    def switchtopo_info(self):
        lib().flups_switchtopo_info(self._solver)

    @staticmethod
    def cleanup_backend():
        lib().flups_cleanup_backend()

    def _adv(self, fn, *args):
        L = lib()
        L.flups_clear_last_error()
        fn(self._solver, *args)
        _check(not L.flups_get_last_error(), fn.__name__)

    def do_copy(self, topology, data, sign):
        self._adv(lib().flups_do_copy, topology._ptr, _data_ptr(data), int(sign))

    def do_fft(self, data, sign):
        self._adv(lib().flups_do_FFT, _data_ptr(data), int(sign))

    def do_mult(self, data, solver_type=SolverType.STD):
        self._adv(lib().flups_do_mult, _data_ptr(data), int(solver_type))

    def get_spectral_info(self):
        kf, ko, ss = (c_double * 3)(), (c_double * 3)(), (c_double * 3)()
        lib().flups_get_spectralInfo(self._solver, kf, ko, ss)
        return np.array(kf[:]), np.array(ko[:]), np.array(ss[:])

    @property
    def path(self):
        """0 = general pencil path (FP64), 1 = fused pipeline FP64, 2 = fused FP32, 3 = general + fused last direction"""
        return lib().flups_get_path(self._solver)

    @property
    def pass_kernels(self):
        """kernel of each general-path pass (forward passes, then backward): 0 generic, 1 contiguous r2r,
        2 paired real columns, 3 contiguous r2c / c2r, 4 strided r2r, 5 strided c2c, 6 contiguous c2c, -1 none, -2 fused with the multiply"""
        buf = (c_int * 8)()
        n = lib().flups_get_pass_kernels(self._solver, buf, 8)
        return list(buf[:max(n, 0)])

    @property
    def last_solve_ns(self):
        return lib().flups_get_last_solve_ns(self._solver)


def locID(axsrc, i0, i1, i2, lia, axtrg, size, nf):
    """memory index of (i0, i1, i2) given in the axsrc frame, component lia, in an array whose fastest
    direction is axtrg (src/flups.h:109-119, python_api/pyflups/flups.py:534)"""
    i = (int(i0), int(i1), int(i2))
    d0 = (3 + axtrg - axsrc) % 3
    d1, d2 = (d0 + 1) % 3, (d0 + 2) % 3
    a0 = axtrg
    a1, a2 = (a0 + 1) % 3, (a0 + 2) % 3
    return i[d0] * nf + size[a0] * nf * (i[d1] + size[a1] * (i[d2] + lia * size[a2]))


def print_data(topology, data):
    """print data laid out by the topology (flups_print_data; host or device memory)"""
    lib().flups_print_data(topology._ptr, _data_ptr(data))
