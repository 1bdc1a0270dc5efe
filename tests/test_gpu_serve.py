"""GPU tests of the serving loop and of batched (hipGraph) solves.

Each problem's result must be the same alone and in the mix: the kernels are
deterministic (no atomics, no order-dependent reductions), so the check is
bitwise. FP64 results are also checked against the oracle (relative L2 / Linf
<= 1e-12) and FP32 ones against it at <= 1e-5.
"""
import numpy as np
import pytest

try:   # torch's HIP runtime must start before libflups' one in a process using both
    import torch  # noqa: F401
except Exception:   # pragma: no cover
    torch = None

import flups_amd as fl
from tests import oracle_lib as orc

pytestmark = pytest.mark.gpu

# BC templates of the mixed workloads
TEMPLATES = {
    "unb": [[4, 4], [4, 4], [4, 4]],
    "perx": [[3, 3], [4, 4], [4, 4]],
    "symx-pery": [[0, 1], [3, 3], [4, 4]],
    "per": [[3, 3], [3, 3], [3, 3]],
}


def blob_field(shape, seed):
    """lda components of 4 Gaussian blobs on the cell grid of [0, n_d h]"""
    lda, n2, n1, n0 = shape
    rng = np.random.default_rng(seed)
    n = max(n0, n1, n2)
    x = [(np.arange(m) + 0.5) / n for m in (n0, n1, n2)]
    out = np.zeros(shape)
    for c in range(lda):
        for _ in range(4):
            cen = rng.uniform(0.3, 0.7, 3) * np.array([n0, n1, n2]) / n
            s = rng.uniform(0.03, 0.06)
            g = [np.exp(-(x[d] - cen[d]) ** 2 / (2 * s * s)) for d in range(3)]
            out[c] += rng.normal() * g[2][:, None, None] * g[1][None, :, None] * g[0][None, None, :]
    return out


def problems():
    """a mix of shapes, BC templates, kernels, lda, precision and a 2-D grid (the last problem runs the general
    path's row r2r and paired-column kernels)"""
    P = []
    spec = [((32, 32, 32), "unb", 2, 1, False), ((48, 40, 32), "unb", 4, 1, True), ((32, 32, 32), "perx", 2, 1, False),
            ((32, 48, 32), "symx-pery", 8, 1, False), ((32, 32, 32), "per", 4, 1, False), ((64, 64, 64), "unb", 2, 1, True),
            ((32, 32, 32), "unb", 4, 3, False), ((64, 64, 1), "unb2d", 2, 1, False), ((40, 32, 48), "perx", 8, 1, False),
            ((32, 32, 32), "unb", 2, 1, False), ((64, 64, 32), "symx-pery", 2, 1, False)]
    for k, (ng, tpl, green, lda, fp32) in enumerate(spec):
        n = max(ng)
        h = [1.0 / n] * 3
        L = [h[0] * m for m in ng]
        bc = [[4, 4], [4, 4], [9, 9]] if tpl == "unb2d" else TEMPLATES[tpl]
        if tpl == "unb2d":
            L[2] = h[2]
        r = fl.ServeRequest.make(ng, bc, h, L, green=green, lda=lda, fp32=fp32)
        rhs = blob_field(r.shape(), 100 + k)
        if tpl in ("per",):
            rhs -= rhs.mean(axis=(1, 2, 3), keepdims=True)
        if tpl == "symx-pery":
            rhs -= rhs.mean(axis=2, keepdims=True)
        P.append((r, rhs, ng, bc, h, L, green, lda, fp32))
    return P


def solve_alone(r, rhs, ng, bc, h, L, green, lda, fp32):
    topo = fl.Topology(0, lda, list(ng), [1, 1, 1], False, alignment=8)
    s = fl.Solver(topo, bc, h, L)
    s.set_green_type(green)
    if fp32:
        s.set_precision(True)
    s.setup()
    f = np.zeros_like(rhs)
    s.solve(f, np.ascontiguousarray(rhs))
    return f


def rel(a, b):
    return (np.sqrt(np.sum((a - b) ** 2) / max(np.sum(b * b), 1e-300)), np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_server_alone_vs_mix_and_oracle():
    P = problems()
    alone = [solve_alone(*p) for p in P]
    for (r, rhs, ng, bc, h, L, green, lda, fp32), fa in zip(P, alone):
        fo = orc.solve(list(ng), bc, lda, h, L, 1, green, rhs)
        e2, ei = rel(fa, fo)
        tol = 1e-5 if fp32 else 1e-12
        assert e2 <= tol and ei <= tol, (ng, bc, green, lda, fp32, e2, ei)
    srv = fl.Server(8 << 30, workers=4, solvers_per_key=2)
    dev_in = [fl.DeviceArray.from_numpy(p[1]) for p in P]
    dev_out = [fl.DeviceArray(p[1].shape) for p in P]
    # every problem in flight together
    tickets = [srv.submit(p[0], dev_in[i], dev_out[i]) for i, p in enumerate(P)]
    for t in tickets:
        srv.wait(t)
    for i, fa in enumerate(alone):
        assert np.array_equal(dev_out[i].numpy(), fa), ("mix differs from alone", i)
    # second round on cached solvers, host buffers this time
    outs = [np.zeros_like(p[1]) for p in P]
    ts = [srv.submit(p[0], p[1], outs[i]) for i, p in enumerate(P)]
    for t in ts:
        srv.wait(t)
    for i, fa in enumerate(alone):
        assert np.array_equal(outs[i], fa), ("cached solver differs", i)
    st = srv.stats()
    assert st["completed"] == 2 * len(P) and st["failed"] == 0
    assert st["cache_hits"] >= len(P)
    assert st["p99_ms"] >= st["p50_ms"] > 0
    srv.close()


def test_server_budget_evicts_and_completes():
    P = problems()[:6]
    alone = [solve_alone(*p) for p in P]
    # room for the largest request plus a little: caching the others forces evictions
    budget = max(p[0].device_bytes() for p in P) + (4 << 20)
    srv = fl.Server(budget, workers=3)
    outs = [np.zeros_like(p[1]) for p in P]
    ts = [srv.submit(p[0], p[1], outs[i]) for i, p in enumerate(P)] + \
         [srv.submit(p[0], p[1], outs[i]) for i, p in enumerate(P)]
    for t in ts:
        srv.wait(t)
    st = srv.stats()
    assert st["failed"] == 0 and st["completed"] == len(ts)
    assert st["evictions"] > 0
    assert st["peak_device_bytes"] <= budget
    for i, fa in enumerate(alone):
        assert np.array_equal(outs[i], fa), i
    srv.close()


def test_server_rejects_oversized_request():
    P = problems()
    srv = fl.Server(1 << 20, workers=1)
    with pytest.raises(RuntimeError):
        srv.submit(P[0][0], P[0][1], np.zeros_like(P[0][1]))
    assert "budget" in fl.last_error()
    srv.close()


def rot_problems():
    """Biot-Savart (ROT) problems with three components: the fused pipeline's rotational z pass (all
    unbounded) and the general path's passes with the rotational multiply (spectral and centred-difference
    derivative symbols, FP64 and FP32 storage)"""
    P = []
    spec = [((32, 40, 48), "unb", 2, False, 1), ((48, 32, 40), "perx", 2, False, 1), ((40, 48, 32), "symx-pery", 2, True, 2),
            ((32, 32, 48), "per", 4, False, 4), ((48, 40, 32), "unb", 8, True, 1)]
    for k, (ng, tpl, green, fp32, odiff) in enumerate(spec):
        n = max(ng)
        h = [1.0 / n] * 3
        L = [h[0] * m for m in ng]
        bc = TEMPLATES[tpl]
        r = fl.ServeRequest.make(ng, bc, h, L, green=green, lda=3, fp32=fp32, solver_type=1, order_diff=odiff)
        rhs = blob_field(r.shape(), 300 + k)
        if tpl == "per":
            rhs -= rhs.mean(axis=(1, 2, 3), keepdims=True)
        if tpl == "symx-pery":
            rhs -= rhs.mean(axis=2, keepdims=True)
        P.append((r, rhs, ng, bc, h, L, green, fp32, odiff))
    return P


def solve_alone_rot(r, rhs, ng, bc, h, L, green, fp32, odiff):
    topo = fl.Topology(0, 3, list(ng), [1, 1, 1], False, alignment=8)
    s = fl.Solver(topo, bc, h, L, order_diff=odiff)
    s.set_green_type(green)
    if fp32:
        s.set_precision(True)
    s.setup()
    f = np.zeros_like(rhs)
    s.solve(f, np.ascontiguousarray(rhs), 1)
    return f


def test_rot_problems_alone_vs_oracle():
    for (r, rhs, ng, bc, h, L, green, fp32, odiff) in rot_problems():
        fa = solve_alone_rot(r, rhs, ng, bc, h, L, green, fp32, odiff)
        fo = orc.solve(list(ng), bc, 3, h, L, 1, green, rhs, typ=1, odiff=odiff)
        e2, ei = rel(fa, fo)
        tol = 1e-5 if fp32 else 1e-12
        assert e2 <= tol and ei <= tol, (ng, bc, green, fp32, odiff, e2, ei)


def test_batch_graph_replays_match_eager():
    """grouped launches, one branch per problem and one problem after the other all give each problem's
    result solved alone, bit for bit, over replays with eager work between them"""
    P = problems()
    # copies of two problems with other right-hand sides: launches of equal grids are grouped across problems
    for k, seed in ((0, 500), (5, 501)):
        q = list(P[k])
        q[1] = blob_field(q[0].shape(), seed)
        P.append(tuple(q))
    # two FP32 problems of the general path (FP32 stage arrays; symmetric x, periodic y, unbounded z), grouped
    ng, bc = (32, 48, 40), TEMPLATES["symx-pery"]
    h = [1.0 / 48] * 3
    L = [h[0] * m for m in ng]
    for seed in (503, 504):
        r = fl.ServeRequest.make(ng, bc, h, L, green=2, lda=1, fp32=True)
        rhs = blob_field(r.shape(), seed)
        rhs -= rhs.mean(axis=2, keepdims=True)
        P.append((r, rhs, ng, bc, h, L, 2, 1, True))
    R = rot_problems()
    # a copy of the FP32 rotational problem of the fused path (grouped rotational z passes)
    q = list(R[4])
    q[1] = blob_field(q[0].shape(), 502)
    R.append(tuple(q))
    alone = [solve_alone(*p) for p in P] + [solve_alone_rot(*p) for p in R]
    reqs = [p[0] for p in P] + [p[0] for p in R]
    rhs = [p[1] for p in P] + [p[1] for p in R]
    dev_in = [fl.DeviceArray.from_numpy(x) for x in rhs]
    dev_out = [fl.DeviceArray(x.shape) for x in rhs]
    b = fl.Batch(reqs, dev_in, dev_out)
    ng, nb = b.launches(True), b.launches(False)
    assert 0 < ng < nb, (ng, nb)

    def clear():   # eager work between replays: clear the outputs through the library
        for d in dev_out:
            zero = fl.DeviceArray.from_numpy(np.zeros(d.shape))
            assert fl.lib().flups_memcpy(d.ptr, zero.ptr, d.nbytes) == 1
            del zero

    for mode in ("grouped", "branches", "grouped"):
        for rep in range(2):
            clear()
            ms = b.run() if mode == "grouped" else b.run_branches()
            assert ms > 0
            for i, fa in enumerate(alone):
                assert np.array_equal(dev_out[i].numpy(), fa), (mode, rep, i)
    clear()
    b.run_sequential()
    for i, fa in enumerate(alone):
        assert np.array_equal(dev_out[i].numpy(), fa), ("sequential", i)
    b.close()


def test_batch_rejects_host_buffers_and_recovers():
    """a batch whose second problem has host buffers fails at set-up (the first problem's solver, stream and
    event are released) and a valid batch built afterwards runs"""
    P = problems()[:2]
    dev_in = [fl.DeviceArray.from_numpy(P[0][1])]
    dev_out = [fl.DeviceArray(P[0][1].shape)]
    host_in, host_out = np.ascontiguousarray(P[1][1]), np.zeros_like(P[1][1])
    fl.lib().flups_clear_last_error()
    with pytest.raises(RuntimeError):
        fl.Batch([P[0][0], P[1][0]], dev_in + [host_in], dev_out + [host_out])
    assert "device memory" in fl.last_error()
    b = fl.Batch([P[0][0]], dev_in, dev_out)
    b.run()
    assert np.array_equal(dev_out[0].numpy(), solve_alone(*P[0]))
    b.close()
