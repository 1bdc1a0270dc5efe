"""GPU tests of the interface around the solve: the profiler handle, the pyflups-named Python
surface and a compiled C caller.

  * a solver made with a FLUPS_Profiler records init / setup / solve and the solve's stages
    (copy rhs, transforms, multiply, copy field) as HIP-event device times whose stages add up
    to the solve (src/flups.cpp:110-114, 219-237; src/Solver.cpp:1218, 1286, 1411-1419, 1475);
  * python_api/tests/test_validation_simple.py:12-68 ported: a periodic LGF_2 solve of
    sin x + sin y + sin z filled through locID, within 1e-2 of -rhs;
  * tests/c_driver/solve_c, built against include/flups.h and MPICH's <mpi.h>, runs a timed
    solve on host arrays and the advanced steps with a spectral filter, and exits 0.
"""
import os
import subprocess

import numpy as np
import pytest

try:   # torch's HIP runtime must start before libflups' one in a process using both
    import torch  # noqa: F401
except Exception:   # pragma: no cover
    torch = None

import flups_amd as fl
from flups_amd import BoundaryType, CenterType, DiffType, GreenType, SolverType

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _stages(timers, prefix):
    return {k: v for k, v in timers.items() if k.startswith(prefix) and "/" not in k[len(prefix):]}


@pytest.mark.parametrize("case", ["fused", "general"])
def test_profiler_records_solve_stages(case):
    prof = fl.Profiler("pytest")
    n = 32
    bc = [[4, 4]] * 3 if case == "fused" else [[0, 0], [1, 1], [0, 0]]
    topo = fl.Topology(0, 1, [n] * 3, [1, 1, 1], False)
    s = fl.Solver(topo, bc, [1.0 / n] * 3, [1.0] * 3, profiler=prof)
    s.set_green_type(GreenType.HEJ_2 if case == "fused" else GreenType.CHAT_2)
    s.setup()
    assert s.path == (1 if case == "fused" else 0)
    rhs = np.random.default_rng(1).normal(size=(1, n, n, n))
    field = np.zeros_like(rhs)
    for _ in range(3):
        s.solve(field, rhs)   # host arrays: the copy stages are timed too
    t = prof.timers()
    for k in ("init", "setup", "setup/green", "solve", "solve/copy rhs", "solve/copy field"):
        assert k in t, (k, sorted(t))
    assert t["solve"][0] == 3 and t["init"][0] == 1 and t["setup"][0] == 1
    assert t["solve"][2] > 0 and t["solve"][1] > 0
    stages = _stages(t, "solve/")
    if case == "fused":
        assert set(stages) == {"solve/copy rhs", "solve/fused pipeline (fftw + domagic)", "solve/copy field"}
        kern = _stages(t, "solve/fused pipeline (fftw + domagic)/")
        assert {k.rsplit("/", 1)[1] for k in kern} == {"x_r2c", "y_fwd", "z_fused", "y_inv", "x_c2r"}
        assert all(v[0] == 3 and v[2] > 0 for v in kern.values())
        assert sum(v[2] for v in kern.values()) <= 1.02 * t["solve/fused pipeline (fftw + domagic)"][2] + 0.05
    else:
        assert set(stages) == {"solve/copy rhs", "solve/fftw forward", "solve/domagic", "solve/fftw backward",
                               "solve/copy field"}
    # the stages tile the solve: consecutive events from its first to its last mark
    total = sum(v[2] for v in stages.values())
    assert abs(total - t["solve"][2]) <= 1e-6 * t["solve"][2] + 1e-6
    prof.display()


def test_validation_simple_port():
    """python_api/tests/test_validation_simple.py:12-68 with this package"""
    nproc = [1, 1, 1]
    nglob = [64, 64, 64]
    L = [2.0 * np.pi] * 3
    h = [L[d] / nglob[d] for d in range(3)]
    topo = fl.Topology(axis=0, lda=1, nglob=nglob, nproc=nproc, is_complex=False)
    bc = [[BoundaryType.PER, BoundaryType.PER]] * 3
    solver = fl.Solver(topo, bc, h, L, order_diff=DiffType.NOD, center_type=[CenterType.CELL_CENTER] * 3)
    solver.set_green_type(GreenType.LGF_2)
    solver.setup()
    nloc = [topo.get_nloc(d) for d in range(3)]
    size = [topo.get_nmem(d) for d in range(3)]
    istart = topo.get_istart()
    ax = topo.axis
    rhs = np.zeros(topo.memsize)
    field = np.zeros_like(rhs)
    for i0 in range(nloc[0]):
        x = (istart[0] + i0) * h[0]
        for i1 in range(nloc[1]):
            y = (istart[1] + i1) * h[1]
            for i2 in range(nloc[2]):
                z = (istart[2] + i2) * h[2]
                rhs[fl.locID(ax, i0, i1, i2, 0, ax, size, 1)] = np.sin(x) + np.sin(y) + np.sin(z)
    solver.solve(field, rhs, solver_type=SolverType.STD)
    assert np.all(np.isfinite(field))
    assert np.max(np.abs(field)) > 1e-6
    err = np.linalg.norm(field + rhs) / max(1e-12, np.linalg.norm(rhs))
    assert err < 1e-2, err
    # the pyflups names of the advanced interface
    buf = solver.get_inner_buffer()
    assert buf.shape[0] == solver.get_alloc_size()
    phys, spec = solver.get_inner_topo_physical(), solver.get_inner_topo_spectral()
    assert [phys.get_nglob(d) for d in range(3)] == nglob and spec.is_complex
    assert topo.cmpt_rank_from_id(5, 0) == 0 and list(topo.ranksplit(0)) == [0, 0, 0]


def test_c_driver():
    exe = os.path.join(ROOT, "tests", "c_driver", "solve_c")
    assert os.path.exists(exe), "tests/c_driver/solve_c is not built (run __graft_entry__.build())"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert "c driver:" in r.stdout and "FLUPS profiler" in r.stdout


@pytest.mark.parametrize("case", ["fused", "general", "fused-fp32"])
def test_solve_on_caller_stream_is_asynchronous_and_equal(case):
    """flups_set_stream with a stream: flups_solve on device arrays queues the solve and returns; after the
    stream is synchronised every field equals the synchronous solve's bitwise. Host arrays stay
    synchronous on the same solver (the result is complete when the call returns)."""
    n = 48
    bc = [[4, 4]] * 3 if case.startswith("fused") else [[0, 1], [3, 3], [4, 4]]
    topo = fl.Topology(0, 1, [n] * 3, [1, 1, 1], False)
    s = fl.Solver(topo, bc, [1.0 / n] * 3, [1.0] * 3)
    s.set_green_type(GreenType.HEJ_2)
    if case == "fused-fp32":
        s.set_precision(True)
    s.setup()
    rng = np.random.default_rng(5)
    rhs = [rng.normal(size=(1, n, n, n)) for _ in range(3)]
    want = []
    for r in rhs:   # synchronous reference (default stream)
        d_r, d_f = fl.DeviceArray.from_numpy(r), fl.DeviceArray(r.shape)
        s.solve(d_f, d_r)
        want.append(d_f.numpy())
    st = fl.Stream()
    s.set_stream(st)
    d_in = [fl.DeviceArray.from_numpy(r) for r in rhs]
    d_out = [fl.DeviceArray(r.shape) for r in rhs]
    for _ in range(2):   # two rounds of three queued solves
        for a, b in zip(d_out, d_in):
            s.solve(a, b)
    st.synchronize()
    for a, w in zip(d_out, want):
        assert np.array_equal(a.numpy(), w)
    assert s.last_solve_ns > 0
    h_f = np.zeros_like(rhs[0])
    s.solve(h_f, rhs[0])   # host arrays: complete on return, no stream wait needed
    assert np.array_equal(h_f, want[0])
    s.set_stream(None)
