"""flups_solve_noalloc (the solve on the library's static device scratch) against flups_solve.

Both entry points run the same kernels on the same inputs; only where the solve's stage arrays
live differs (the solver's own allocation vs the static scratch), so the fields must be
bitwise equal. The calls run back to back, each smaller than the one before, then the first
again, so every call after the first finds the previous call's data in the scratch:
  * cell 64^3, all faces unbounded, HEJ_2, FP64 (fused pipeline);
  * cell 40x36x32, x EVEN-ODD, y periodic, z unbounded, HEJ_2 (general passes, last direction
    fused with the multiply);
  * cell 24x20x16, lda 3 ROT (SPE), x EVEN-EVEN, y ODD-ODD, z periodic (general passes and the
    rotational multiply);
  * cell 32^3, all faces unbounded, HEJ_2, FP32 (fused pipeline in FP32);
  * the 64^3 case again.
A 192^3 unbounded FP64 solve needs more stage memory than the scratch holds and must fail
with an error naming the limit.
"""
import numpy as np
import pytest

try:   # torch's HIP runtime must start before libflups' one in a process using both
    import torch  # noqa: F401
except Exception:   # pragma: no cover
    torch = None

import flups_amd as fl

pytestmark = pytest.mark.gpu

UNB, PER, EVEN, ODD = 4, 3, 0, 1
CASES = [
    ("unb-64-fp64", (64, 64, 64), [[UNB, UNB]] * 3, 1, 0, False),
    ("eo-per-unb-40x36x32", (40, 36, 32), [[EVEN, ODD], [PER, PER], [UNB, UNB]], 1, 0, False),
    ("rot-ee-oo-per-24x20x16", (24, 20, 16), [[EVEN, EVEN], [ODD, ODD], [PER, PER]], 3, 1, False),
    ("unb-32-fp32", (32, 32, 32), [[UNB, UNB]] * 3, 1, 0, True),
    ("unb-64-fp64-again", (64, 64, 64), [[UNB, UNB]] * 3, 1, 0, False),
]


def make_solver(n, bc, lda, rot, fp32):
    topo = fl.Topology(0, lda, list(n), [1, 1, 1], False)
    h = [1.0 / n[0]] * 3
    L = [n[d] * h[d] for d in range(3)]
    kw = {"order_diff": fl.DiffType.SPE} if rot else {}
    s = fl.Solver(topo, bc, h, L, **kw)   # scalar boundary types apply to every component
    s.set_green_type(fl.GreenType.HEJ_2)
    if fp32:
        s.set_precision(True)
    s.setup()
    return topo, s


def rhs_for(n, lda, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(lda, n[2], n[1], n[0]))


def test_noalloc_matches_solve_on_calls_in_a_row():
    L = fl.lib()
    results = {}
    for i, (name, n, bc, lda, typ, fp32) in enumerate(CASES):
        print("noalloc case", name, flush=True)
        _, s = make_solver(n, bc, lda, typ == 1, fp32)
        rhs = rhs_for(n, lda, 100 + (i % 4))
        d_rhs = fl.DeviceArray.from_numpy(rhs)
        d_ref = fl.DeviceArray(rhs.shape)
        d_out = fl.DeviceArray(rhs.shape)
        s.solve(d_ref, d_rhs, typ)
        L.flups_clear_last_error()
        L.flups_solve_noalloc(s._solver, d_out.ptr, d_rhs.ptr, typ)
        assert fl.last_error() == "", fl.last_error()
        ref = d_ref.numpy()     # synchronous copies on the default stream (the solver's stream)
        got = d_out.numpy()
        assert np.all(np.isfinite(ref))
        assert np.array_equal(got, ref), (name, float(np.max(np.abs(got - ref))))
        results[name] = got
    # the repeated first case equals its first run
    assert np.array_equal(results["unb-64-fp64"], results["unb-64-fp64-again"])


def test_noalloc_rejects_stage_arrays_beyond_the_scratch():
    L = fl.lib()
    n = (192, 192, 192)
    _, s = make_solver(n, [[UNB, UNB]] * 3, 1, False, False)
    d_rhs = fl.DeviceArray((1, n[2], n[1], n[0]))
    d_out = fl.DeviceArray((1, n[2], n[1], n[0]))
    L.flups_clear_last_error()
    L.flups_solve_noalloc(s._solver, d_out.ptr, d_rhs.ptr, 0)
    err = fl.last_error()
    assert "static scratch" in err and "134217728" in err, err


def test_noalloc_rejects_host_pointers():
    L = fl.lib()
    n = (16, 16, 16)
    _, s = make_solver(n, [[UNB, UNB]] * 3, 1, False, False)
    a = np.zeros((1, 16, 16, 16))
    b = np.zeros((1, 16, 16, 16))
    L.flups_clear_last_error()
    L.flups_solve_noalloc(s._solver, a.ctypes.data, b.ctypes.data, 0)
    assert "device pointers only" in fl.last_error()
