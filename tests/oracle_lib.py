"""ctypes access to the CPU oracle (oracle/liboracle.so), used only as the checker."""
import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KDIR = os.path.join(ROOT, "flups_amd", "kernel").encode()
_lib = None

dp = POINTER(c_double)
ip = POINTER(c_int)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(ROOT, "oracle", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.oracle_last_error.restype = c_char_p
        L.oracle_validation.argtypes = [c_int, ip, c_int, c_int, c_int, c_char_p, dp]
        L.oracle_validation_fields.argtypes = [c_int, ip, c_int, c_int, ip, dp, dp, dp]
        L.oracle_vtube.argtypes = [c_int] * 6 + [c_double, c_char_p, dp]
        L.oracle_vtube_fields.argtypes = [c_int] * 4 + [c_double, ip, dp, dp]
        L.oracle_solve.argtypes = [ip, ip, c_int, dp, dp, c_int, c_int, c_int, c_double, c_char_p, c_int, dp, dp]
        L.oracle_solve_timed.argtypes = [ip, ip, c_int, dp, dp, c_int, c_int, c_int, c_double, c_char_p, c_int, dp, dp,
                                         c_int, c_int, dp]
        L.oracle_solve_timed_for.argtypes = [ip, ip, c_int, dp, dp, c_int, c_int, c_int, c_double, c_char_p, c_int, dp,
                                             dp, c_int, c_double, c_int, c_int, ip, dp]
        L.oracle_spectral_size.argtypes = [ip, ip, c_int, dp, dp, c_int, ip, ip]
        L.oracle_green.argtypes = [ip, ip, dp, dp, c_int, c_int, c_double, c_char_p, dp]
        L.oracle_plan_info.argtypes = [ip, ip, c_int, dp, dp, c_int, dp, dp, dp, dp, dp, ip]
        L.oracle_c2c.argtypes = [c_int, c_int, dp, dp]
        L.oracle_r2c.argtypes = [c_int, dp, dp]
        L.oracle_c2r.argtypes = [c_int, dp, dp]
        L.oracle_r2r.argtypes = [c_int, c_int, dp, dp]
        for n in ["oracle_lgf1d"]:
            getattr(L, n).argtypes = [c_int, c_int, c_double]
            getattr(L, n).restype = c_double
        L.oracle_lgf_symbol.argtypes = [c_int, c_double]
        L.oracle_lgf_symbol.restype = c_double
        for n in ["oracle_bessel_k0", "oracle_bessel_k1", "oracle_expint_e1", "oracle_sine_integral", "oracle_ji0c"]:
            getattr(L, n).argtypes = [c_double]
            getattr(L, n).restype = c_double
        L.oracle_mehr_1d.argtypes = [c_int, c_int, c_double, c_double]
        L.oracle_mehr_1d.restype = c_double
        _lib = L
    return _lib


def _p(a, t=c_double):
    return a.ctypes.data_as(POINTER(t))


def check(rc):
    if rc < 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return rc


def flat_bc(bc, lda):
    """bc[d][side] -> scalar or per-component list; returns int32 [(d*2+side)*lda + c]"""
    out = np.zeros(6 * lda, dtype=np.int32)
    for d in range(3):
        for s in range(2):
            v = bc[d][s]
            vals = list(v) if hasattr(v, "__len__") else [v] * lda
            out[(d * 2 + s) * lda:(d * 2 + s + 1) * lda] = [int(x) for x in vals]
    return out


def solve(nglob, bc, lda, h, L, center, kernel, rhs, typ=0, odiff=0, alpha=2.0):
    """rhs: float64 array shaped (lda, n2, n1, n0) (dim 0 fastest)"""
    nglob = np.asarray(nglob, dtype=np.int32)
    fb = flat_bc(bc, lda)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    field = np.zeros_like(rhs)
    check(lib().oracle_solve(_p(nglob, c_int), _p(fb, c_int), lda, _p(np.asarray(h, float)), _p(np.asarray(L, float)),
                             center, odiff, kernel, alpha, KDIR, typ, _p(rhs), _p(field)))
    return field


def validation_fields(center, bc6, N, lda=1):
    b = np.asarray(bc6, dtype=np.int32)
    ng = np.zeros(3, dtype=np.int32)
    h = np.zeros(3)
    check(lib().oracle_validation_fields(center, _p(b, c_int), N, lda, _p(ng, c_int), _p(h), None, None))
    shape = (lda, ng[2], ng[1], ng[0])
    rhs = np.zeros(shape)
    sol = np.zeros(shape)
    check(lib().oracle_validation_fields(center, _p(b, c_int), N, lda, _p(ng, c_int), _p(h), _p(rhs), _p(sol)))
    return ng, h, rhs, sol


def validation_errors(center, bc6, kernel, N, lda=1):
    b = np.asarray(bc6, dtype=np.int32)
    out = np.zeros(2 * lda)
    check(lib().oracle_validation(center, _p(b, c_int), kernel, N, lda, KDIR, _p(out)))
    return out


def vtube_fields(vdir, symx, symy, N, sigma=0.1):
    bc = np.zeros(18, dtype=np.int32)
    rhs = np.zeros((3, N, N, N))
    sol = np.zeros((3, N, N, N))
    check(lib().oracle_vtube_fields(vdir, symx, symy, N, sigma, _p(bc, c_int), _p(rhs), _p(sol)))
    bcl = [[list(bc[(d * 2 + s) * 3:(d * 2 + s + 1) * 3]) for s in range(2)] for d in range(3)]
    return bcl, rhs, sol


def vtube_errors(vdir, symx, symy, N, kernel=0, order=1, sigma=0.1):
    out = np.zeros(6)
    check(lib().oracle_vtube(vdir, symx, symy, N, kernel, order, sigma, KDIR, _p(out)))
    return out
