"""FP32 requests on the general path (boundary sets the fused pipeline does not run): FP32 stage arrays
and Green's function, FP64 arithmetic in the kernels (flups_get_path 4, or 5 with the fused last
direction), against the FP64 oracle at relative L2 and Linf <= 1e-5 and above 1e-12 (the values really
went through single precision). The cases reach every kernel family of the general path with float
storage on each side: contiguous and strided real-to-real passes, paired real columns, contiguous
r2c / c2r rows, strided and contiguous c2c, thread teams of 5 and 7 times a power of two, the
generic kernel (a length with a factor 11), the multiply (real and complex, STD and ROT), and the
fused last directions (c2c z, r2c columns, r2c rows, 2-D).
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
EVEN, ODD, PER, UNB, NONE = 0, 1, 3, 4, 9
TOL = 1e-5


def rel_err(a, b):
    return (np.sqrt(np.sum((a - b) ** 2)) / np.sqrt(np.sum(b * b)), np.max(np.abs(a - b)) / np.max(np.abs(b)))


def rhs_for(nglob, lda, seed):
    """Gaussian blobs plus a little noise on the grid (dim 0 fastest)"""
    rng = np.random.default_rng(seed)
    n0, n1, n2 = nglob
    n = max(nglob)
    xs = [(np.arange(m) + 0.5) / n for m in (n0, n1, n2)]
    out = 1e-2 * rng.normal(size=(lda, n2, n1, n0))
    for c in range(lda):
        for _ in range(3):
            cen = rng.uniform(0.3, 0.7, 3) * np.array([n0, n1, n2]) / n
            g = [np.exp(-(xs[d] - cen[d]) ** 2 / (2 * 0.05 ** 2)) for d in range(3)]
            out[c] += rng.normal() * g[2][:, None, None] * g[1][None, :, None] * g[0][None, None, :]
    return out


# (id, nglob, bc, center, lda, kernel, type, expected path, expected pass kernels present)
CASES = [
    ("sym-x/per-y/unb-z", (64, 64, 64), [[EVEN, ODD], [PER, PER], [UNB, UNB]], 1, 1, 2, 0, 5, {1, 2}),
    ("all-symmetric", (48, 40, 32), [[EVEN, EVEN], [ODD, ODD], [EVEN, ODD]], 1, 1, 2, 0, 4, {1, 4}),
    ("semi-unbounded-y", (64, 64, 32), [[UNB, UNB], [EVEN, UNB], [PER, PER]], 1, 1, 2, 0, None, set()),
    ("node-ee-oo-unb", (33, 33, 33), [[EVEN, EVEN], [ODD, ODD], [UNB, UNB]], 0, 1, 2, 0, 5, set()),
    ("x-last-r2c-rows", (64, 48, 32), [[PER, PER], [ODD, ODD], [EVEN, EVEN]], 1, 1, 2, 0, 5, set()),
    ("mixed-radix-unb", (80, 48, 112), [[UNB, UNB]] * 3, 1, 1, 2, 0, 5, set()),
    ("node-periodic", (33, 65, 33), [[PER, PER], [UNB, UNB], [PER, PER]], 0, 1, 2, 0, None, set()),
    ("generic-length", (44, 22, 26), [[EVEN, ODD], [PER, PER], [UNB, UNB]], 1, 1, 2, 0, None, {0}),
    # the solver has a fused last direction for STD solves (path 5); ROT solves run the passes and mult_rot
    ("rot-symmetric", (32, 24, 40), [[EVEN, ODD], [PER, PER], [UNB, UNB]], 1, 3, 2, 1, 5, set()),
    ("lgf4-sym-x", (64, 32, 64), [[ODD, EVEN], [PER, PER], [UNB, UNB]], 1, 1, 8, 0, 5, set()),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_general_path_fp32_vs_oracle(case):
    name, nglob, bc, center, lda, kernel, typ, path, kinds = case
    n0, n1, n2 = nglob
    h = [1.0 / (max(nglob) - (1 - center))] * 3
    L = [h[0] * (n - (1 - center)) for n in nglob]
    rhs = rhs_for(nglob, lda, 17)
    topo = fl.Topology(0, lda, list(nglob), [1, 1, 1], False)
    odiff = fl.DiffType.SPE if typ == 1 else fl.DiffType.NOD
    s = fl.Solver(topo, bc, h, L, order_diff=odiff, center_type=[center] * 3)
    s.set_green_type(kernel)
    s.set_precision(True)
    s.setup()
    assert s.path in (4, 5), (name, s.path)
    if path is not None:
        assert s.path == path, (name, s.path)
    got_kinds = set(s.pass_kernels)
    assert kinds <= got_kinds, (name, s.pass_kernels)
    d_rhs = fl.DeviceArray.from_numpy(rhs)
    d_f = fl.DeviceArray(rhs.shape)
    s.solve(d_f, d_rhs, typ)
    f32 = d_f.numpy()
    ref = orc.solve(list(nglob), bc, lda, h, L, center, kernel, rhs, typ=typ, odiff=int(odiff))
    e2, ei = rel_err(f32, ref)
    assert e2 <= TOL and ei <= TOL, (name, e2, ei)
    assert e2 > 1e-12, (name, e2)   # the stage arrays really are single precision


def test_general_fp32_host_arrays_and_repeat():
    """host arrays and repeated solves give the same FP32 result; the advanced interface refuses FP32 solvers"""
    nglob = (48, 40, 32)
    bc = [[EVEN, ODD], [PER, PER], [UNB, UNB]]
    h = [1.0 / 48] * 3
    L = [h[0] * n for n in nglob]
    rhs = rhs_for(nglob, 1, 3)
    topo = fl.Topology(0, 1, list(nglob), [1, 1, 1], False)
    s = fl.Solver(topo, bc, h, L)
    s.set_green_type(fl.GreenType.HEJ_2)
    s.set_precision(True)
    s.setup()
    a = np.zeros_like(rhs)
    b = np.zeros_like(rhs)
    s.solve(a, rhs)
    s.solve(b, rhs)
    assert np.array_equal(a, b)
    d_rhs = fl.DeviceArray.from_numpy(rhs)
    d_f = fl.DeviceArray(rhs.shape)
    s.solve(d_f, d_rhs)
    assert np.array_equal(d_f.numpy(), a)
    fl.lib().flups_clear_last_error()
    with pytest.raises(RuntimeError):
        s.inner_buffer()
