"""GPU parity at the benchmarked sizes: the kernel instantiations bench.py times.

  * cell 256^3, all faces unbounded, HEJ_2 (the headline workload): the fused pipeline with
    M = 512 columns (x_r2c<512/2>, y_fwd / y_inv at 512, z_fused at 512), FP64 against the
    oracle at 1e-12 relative L2 / Linf, and the FP32 pipeline against the same oracle field at
    1e-5;
  * cell 256^3, x EVEN-ODD, y periodic, z unbounded, LGF_4 (the mixed-boundary workload): the
    general path's 256-point row real-to-real and paired real-column kernels and its fused
    512-point last direction, against the oracle at 1e-12;
  * node-centred 33 x 33 x 513, all faces unbounded, HEJ_6, lda 3, ROT with the spectral curl: the
    rotational z pass at the 1024-point length of the 513^3 ring workload (z_rot<double, 1024, 4>
    and z_rot<float, 1024, 4>), FP64 at 1e-12 and FP32 at 1e-5 against the oracle.
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
N = 256


def rel_err(a, b):
    return (np.sqrt(np.sum((a - b) ** 2)) / np.sqrt(np.sum(b * b)), np.max(np.abs(a - b)) / np.max(np.abs(b)))


def blobs(n, seed):
    """sum of 4 Gaussian blobs: centres U[0.3, 0.7]^3, sigma U[0.03, 0.06], amplitudes N(0, 1)"""
    rng = np.random.default_rng(seed)
    x = (np.arange(n) + 0.5) / n
    f = np.zeros((n, n, n))
    for _ in range(4):
        c = rng.uniform(0.3, 0.7, 3)
        s = rng.uniform(0.03, 0.06)
        a = rng.normal()
        gx = np.exp(-(x - c[0]) ** 2 / (2 * s * s))
        gy = np.exp(-(x - c[1]) ** 2 / (2 * s * s))
        gz = np.exp(-(x - c[2]) ** 2 / (2 * s * s))
        f += a * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    return f[None]


@pytest.fixture(scope="module")
def unbounded_case():
    rhs = blobs(N, 2024)
    h = [1.0 / N] * 3
    ref = orc.solve([N] * 3, [[4, 4]] * 3, 1, h, [1.0] * 3, 1, 2, rhs)
    return rhs, h, ref


@pytest.mark.parametrize("fp32", [False, True], ids=["fp64", "fp32"])
def test_headline_256_unbounded(unbounded_case, fp32):
    rhs, h, ref = unbounded_case
    topo = fl.Topology(0, 1, [N] * 3, [1, 1, 1], False)
    s = fl.Solver(topo, [[4, 4]] * 3, h, [1.0] * 3)
    s.set_green_type(fl.GreenType.HEJ_2)
    if fp32:
        s.set_precision(True)
    s.setup()
    assert s.path == (2 if fp32 else 1)
    assert s.schedule == (20, 0, 12)   # the bench's default: half-width, XCD-grouped z tiles, x passes backwards
    d_rhs = fl.DeviceArray.from_numpy(rhs)
    d_f = fl.DeviceArray(rhs.shape)
    s.solve(d_f, d_rhs)
    e2, ei = rel_err(d_f.numpy(), ref)
    tol = 1e-5 if fp32 else 1e-12
    assert e2 <= tol and ei <= tol, (e2, ei)
    if fp32:
        assert e2 > 1e-12


def test_mixed_256_even_odd_periodic_unbounded():
    rng = np.random.default_rng(31)
    rhs = rng.normal(size=(1, N, N, N))
    rhs -= rhs.mean(axis=2, keepdims=True)   # zero mean along the periodic y
    bc = [[0, 1], [3, 3], [4, 4]]
    h = [1.0 / N] * 3
    topo = fl.Topology(0, 1, [N] * 3, [1, 1, 1], False)
    s = fl.Solver(topo, bc, h, [1.0] * 3)
    s.set_green_type(fl.GreenType.LGF_4)
    s.setup()
    assert s.path == 3
    kinds = s.pass_kernels
    assert 0 not in kinds and -2 in kinds, kinds   # no generic pass; the last direction fused
    d_rhs = fl.DeviceArray.from_numpy(rhs)
    d_f = fl.DeviceArray(rhs.shape)
    s.solve(d_f, d_rhs)
    ref = orc.solve([N] * 3, bc, 1, h, [1.0] * 3, 1, 8, rhs)
    e2, ei = rel_err(d_f.numpy(), ref)
    assert e2 <= 1e-12 and ei <= 1e-12, (e2, ei)


@pytest.mark.parametrize("fp32", [False, True], ids=["fp64", "fp32"])
def test_rot_1024_node(fp32):
    nglob = (33, 33, 513)
    h = [1.0 / 512] * 3
    L = [h[0] * (n - 1) for n in nglob]
    rng = np.random.default_rng(77)
    rhs = rng.normal(size=(3, nglob[2], nglob[1], nglob[0]))
    topo = fl.Topology(0, 3, list(nglob), [1, 1, 1], False)
    s = fl.Solver(topo, [[4, 4]] * 3, h, L, order_diff=fl.DiffType.SPE, center_type=[fl.CenterType.NODE_CENTER] * 3)
    s.set_green_type(fl.GreenType.HEJ_6)
    if fp32:
        s.set_precision(True)
    s.setup()
    assert s.path == (2 if fp32 else 1)
    d_rhs = fl.DeviceArray.from_numpy(rhs)
    d_f = fl.DeviceArray(rhs.shape)
    s.solve(d_f, d_rhs, fl.SolverType.ROT)
    ref = orc.solve(list(nglob), [[4, 4]] * 3, 3, h, L, 0, 4, rhs, typ=1, odiff=1)
    e2, ei = rel_err(d_f.numpy(), ref)
    tol = 1e-5 if fp32 else 1e-12
    assert e2 <= tol and ei <= tol, (e2, ei)
