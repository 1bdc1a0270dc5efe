"""GPU parity of the HIP solve path against the CPU oracle and the reference's golden data.

Every case solves on the GPU through the C interface (flups_amd -> libflups.so)
and compares:
  * the field with the oracle's field on the same input: relative L2 and Linf
    errors <= 1e-12 (FP64 path);
  * for the reference's validation cases, the error norms against the exact
    solution with samples/validation/data_ref at the reference's own tolerance
    (1e-10 relative, absolute below 1e-14; check_res.py:9,40-47).
"""
import json
import os

import numpy as np
import pytest

try:   # torch's HIP runtime must start before libflups' one in a process using both
    import torch
except Exception:   # pragma: no cover
    torch = None

import flups_amd as fl
from tests import oracle_lib as orc

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-12

BC1 = [(a, b) for a in (0, 1, 4) for b in (0, 1, 4)] + [(3, 3)]
BC2 = BC1 + [(9, 9)]
ALL_BCS = [x + y + z for x in BC1 for y in BC1 for z in BC2]


def gpu_solve(nglob, bc, lda, h, L, center, kernel, rhs, typ=0, odiff=0, alpha=2.0):
    topo = fl.Topology(0, lda, list(nglob), [1, 1, 1], False)
    s = fl.Solver(topo, bc, list(h), list(L), order_diff=odiff, center_type=[center] * 3)
    s.set_green_type(kernel)
    if alpha != 2.0:
        s.set_alpha(alpha)
    s.setup()
    field = np.zeros_like(rhs)
    s.solve(field, np.ascontiguousarray(rhs), typ)
    return field


def rel_err(a, b):
    den2 = np.sqrt(np.sum(b * b))
    deni = np.max(np.abs(b))
    return np.sqrt(np.sum((a - b) ** 2)) / max(den2, 1e-300), np.max(np.abs(a - b)) / max(deni, 1e-300)


def err_norms(field, sol, h, bc6):
    vol = 1.0
    for d in range(3):
        if bc6[2 * d] != 9:
            vol *= h[d]
    out = []
    for c in range(field.shape[0]):
        e = sol[c] - field[c]
        out += [np.sqrt(np.sum(e * e) * vol), np.max(np.abs(e))]
    return np.array(out)


def golden_ok(got, want):
    """relative difference < 1e-10, absolute where the reference value is below 1e-14 (check_res.py:40-47)"""
    got, want = np.asarray(got, float), np.asarray(want, float)
    small = np.abs(want) < 1e-14
    r = np.where(small, np.abs(got - want), np.abs(got - want) / np.where(small, 1.0, np.abs(want)))
    return bool(np.all(r < 1e-10))


def golden_table():
    ref = json.load(open(os.path.join(HERE, "golden", "validation_ref.json")))
    return {(e["center"], e["bc"], e["kernel"], e["N"], len(e["err"]) // 2): e["err"] for e in ref}


def run_validation_case(center, bc6, kernel, N, lda=1, golden=None):
    ng, h, rhs, sol = orc.validation_fields(center, bc6, N, lda)
    bc = [[bc6[2 * d], bc6[2 * d + 1]] for d in range(3)]
    f_gpu = gpu_solve(ng, bc, lda, h, [1, 1, 1], center, kernel, rhs)
    f_ora = orc.solve(ng, bc, lda, h, [1, 1, 1], center, kernel, rhs)
    e2, ei = rel_err(f_gpu, f_ora)
    assert e2 <= TOL and ei <= TOL, (center, bc6, kernel, N, e2, ei)
    if golden is not None:
        key = ("CellCenter" if center == 1 else "NodeCenter", "".join(str(b) for b in bc6), kernel, N, lda)
        if key in golden:
            got = err_norms(f_gpu, sol, h, bc6)
            assert golden_ok(got, golden[key]), (key, got, golden[key])
            return 1
    return 0


@pytest.mark.parametrize("center", [1, 0])
def test_validation_sweep_chat2(center):
    """all 1100 boundary combinations of test_3D_bcs.py, CHAT_2, N = 8 (cell) / 9 (node)"""
    golden = golden_table()
    N = 8 if center == 1 else 9
    matched = 0
    for bc6 in ALL_BCS:
        matched += run_validation_case(center, bc6, 0, N, 1, golden)
    assert matched == len(ALL_BCS)


KERN_BCS = [(4, 4, 4, 4, 4, 4), (4, 4, 0, 4, 4, 4), (4, 1, 1, 4, 4, 4), (4, 1, 4, 4, 4, 4), (4, 0, 1, 4, 4, 4),
            (4, 0, 1, 4, 4, 1), (3, 3, 3, 3, 3, 3), (4, 0, 1, 4, 9, 9), (3, 3, 3, 3, 9, 9)]
DEVICE_KERNELS = list(range(15))   # every kernel of the reference


@pytest.mark.parametrize("center", [1, 0])
def test_kernel_sweep(center):
    """9 BC sets x 15 kernels of test_3D_kerns.py at N = 16 (cell) / 17 (node), against the oracle
    and the golden norms; the reference's data hold no entry for HEJ_0 on the 2-D unbounded set
    (its Ji0c kernel), which is checked against the oracle only"""
    golden = golden_table()
    N = 16 if center == 1 else 17
    matched = 0
    for bc6 in KERN_BCS:
        for k in DEVICE_KERNELS:
            matched += run_validation_case(center, bc6, k, N, 1, golden)
    assert matched == len(KERN_BCS) * len(DEVICE_KERNELS) - 1


# the reference's tests hold no case of these: HEJ_0 with two unbounded directions (Ji0c at r = 0
# and on the k = 0 plane, zero_ = 0 elsewhere, no spectral rewrite) and the Mehrstellen kernels
# with one unbounded direction (1-D lattice Green's functions); oracle parity at 1e-12
TWO_ONE_CASES = [
    # (nglob, bc, center, kernel)
    ((32, 32, 16), [[4, 4], [4, 4], [3, 3]], 1, 7),
    ((33, 33, 16), [[4, 4], [1, 4], [3, 3]], 0, 7),
    ((24, 32, 32), [[3, 3], [4, 0], [4, 4]], 1, 7),
    ((64, 64, 1), [[4, 4], [4, 4], [9, 9]], 1, 7),
    ((32, 16, 16), [[4, 4], [3, 3], [3, 3]], 1, 11),
    ((32, 16, 16), [[4, 4], [3, 3], [3, 3]], 1, 12),
    ((32, 16, 16), [[4, 4], [3, 3], [0, 1]], 1, 13),
    ((17, 17, 17), [[4, 4], [3, 3], [3, 3]], 0, 14),
    ((16, 32, 16), [[3, 3], [0, 4], [3, 3]], 1, 14),
]


@pytest.mark.parametrize("case", TWO_ONE_CASES, ids=["hej0-unb2-per", "hej0-node-mix", "hej0-per-yz", "hej0-2d",
                                                     "mehr4l-unb1", "mehr6l-unb1", "mehr4f-unb1-sym", "mehr6f-node",
                                                     "mehr6f-semi-y"])
def test_hej0_two_unbounded_and_mehr_one_unbounded(case):
    nglob, bc, center, kernel = case
    rng = np.random.default_rng(5)
    n0, n1, n2 = nglob
    rhs = rng.normal(size=(1, n2, n1, n0))
    if any(b == [3, 3] for b in bc):
        rhs -= rhs.mean(axis=(1, 2, 3), keepdims=True)
    n = max(nglob)
    h = [1.0 / (n - (1 - center))] * 3
    L = [h[0] * (m - (1 - center)) for m in nglob]
    if nglob[2] == 1:
        L[2] = h[2]
    f = gpu_solve(list(nglob), bc, 1, h, L, center, kernel, rhs)
    ora = orc.solve(list(nglob), bc, 1, h, L, center, kernel, rhs)
    e2, ei = rel_err(f, ora)
    assert e2 <= TOL and ei <= TOL, (e2, ei)


def test_hej0_two_unbounded_device_green_planes():
    """the device's transformed HEJ_0 Green's function with x, y unbounded and z periodic, read back
    through flups_do_mult on a spectral buffer of ones (data = normfact G): equal to the oracle's
    G at 1e-12, every kz != 0 plane the constant volfact G(r = 0) = h^2 (log(2 h / pi) - gamma) /
    (2 pi), as zero_ = 0 off r = 0 and k = 0 implies (src/green_kernels.hpp:430-432)"""
    import ctypes
    n, hh = 32, 1.0 / 32
    nglob, bc, L = [n, n, 16], [[4, 4], [4, 4], [3, 3]], [1.0, 1.0, 0.5]
    topo = fl.Topology(0, 1, nglob, [1, 1, 1], False)
    s = fl.Solver(topo, bc, [hh] * 3, L)
    s.set_green_type(7)
    s.setup()
    spec_t = s.inner_topo_spectral()
    nl = [spec_t.get_nloc(d) for d in range(3)]
    assert nl == [2 * n, 2 * n, 9] and spec_t.is_complex
    inner = s.inner_buffer()
    ones = np.zeros(inner.shape[0])
    ones[0: 2 * int(np.prod(nl)): 2] = 1.0
    inner.upload(ones)
    s.do_mult(inner, fl.SolverType.STD)
    got = inner.numpy()[: 2 * int(np.prod(nl))].reshape(nl[2], nl[1], nl[0], 2)
    c_int, dp = ctypes.c_int, ctypes.POINTER(ctypes.c_double)
    fb = orc.flat_bc(bc, 1)
    ng = np.asarray(nglob, np.int32)
    nf_, vf_ = np.zeros(1), np.zeros(1)
    kf, ko, ss, order = np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(3, np.int32)
    orc.check(orc.lib().oracle_plan_info(orc._p(ng, c_int), orc._p(fb, c_int), 1, orc._p(np.array([hh] * 3)),
                                         orc._p(np.array(L)), 1, orc._p(nf_), orc._p(vf_), orc._p(kf), orc._p(ko),
                                         orc._p(ss), orc._p(order, c_int)))
    g = np.zeros(2 * int(np.prod(nl)))
    orc.check(orc.lib().oracle_green(orc._p(ng, c_int), orc._p(fb, c_int), orc._p(np.array([hh] * 3)),
                                     orc._p(np.array(L)), 1, 7, 2.0, orc.KDIR, orc._p(g)))
    want = nf_[0] * g.reshape(got.shape)
    assert np.max(np.abs(got - want)) <= TOL * np.max(np.abs(want))
    c = nf_[0] * hh * hh * (np.log(2 * hh / np.pi) - np.euler_gamma) / (2 * np.pi)
    for kz in range(1, 9):
        assert np.max(np.abs(got[kz, :, :, 0] - c)) <= TOL * abs(c), kz
        assert np.max(np.abs(got[kz, :, :, 1])) <= TOL * abs(c), kz


HEJ_ONE_UNB = [(bc6, k) for bc6 in [(4, 4, 3, 3, 3, 3), (0, 1, 3, 3, 4, 4)] for k in (3, 4, 5, 6)]


@pytest.mark.parametrize("case", HEJ_ONE_UNB, ids=["%s-k%d" % ("".join(map(str, c[0])), c[1]) for c in HEJ_ONE_UNB])
def test_hej_one_unbounded_vs_oracle(case):
    """HEJ_4..10 with one unbounded direction (green_kernels.hpp:461-566) at N = 32: GPU against the
    oracle at 1e-12 (HEJ_2 is in test_mixed_bc_lgf)"""
    bc6, kernel = case
    run_validation_case(1, bc6, kernel, 32)


def test_vtube_rot():
    """the 27 Biot-Savart (lda = 3, ROT, spectral derivative) cases of solve_vtube"""
    ref = json.load(open(os.path.join(HERE, "golden", "vtube_ref.json")))
    N = 16
    h = [1.0 / N] * 3
    for e in ref:
        bc, rhs, sol = orc.vtube_fields(e["dir"], e["symx"], e["symy"], N)
        f_gpu = gpu_solve([N] * 3, bc, 3, h, [1, 1, 1], 1, 0, rhs, typ=1, odiff=1)
        f_ora = orc.solve([N] * 3, bc, 3, h, [1, 1, 1], 1, 0, rhs, typ=1, odiff=1)
        e2, ei = rel_err(f_gpu, f_ora)
        assert e2 <= TOL and ei <= TOL, (e, e2, ei)
        vol = h[0] ** 3
        got = []
        for c in range(3):
            d = sol[c] - f_gpu[c]
            got += [np.sqrt(np.sum(d * d) * vol), np.max(np.abs(d))]
        assert golden_ok(got, e["err"]), (e, got)


def blobs_rhs(n, seed, L=1.0):
    """sum of 4 Gaussian blobs, centres U[0.3,0.7]^3, sigma U[0.03,0.06], amplitudes N(0,1)"""
    rng = np.random.default_rng(seed)
    h = L / n
    x = (np.arange(n) + 0.5) * h
    X, Y, Z = x[None, None, :], x[None, :, None], x[:, None, None]
    f = np.zeros((n, n, n))
    for _ in range(4):
        c = rng.uniform(0.3, 0.7, 3) * L
        s = rng.uniform(0.03, 0.06) * L
        a = rng.normal()
        f += a * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / (2 * s * s))
    return f[None]


@pytest.mark.parametrize("kernel", [2, 4, 8, 0])
def test_unbounded_blobs_64(kernel):
    """config 0 shape: cell 64^3 fully unbounded, Gaussian-blob rhs"""
    n = 64
    rhs = blobs_rhs(n, 1234 + kernel)
    bc = [[4, 4]] * 3
    h = [1.0 / n] * 3
    f_gpu = gpu_solve([n] * 3, bc, 1, h, [1, 1, 1], 1, kernel, rhs)
    f_ora = orc.solve([n] * 3, bc, 1, h, [1, 1, 1], 1, kernel, rhs)
    e2, ei = rel_err(f_gpu, f_ora)
    assert e2 <= TOL and ei <= TOL, (e2, ei)


@pytest.mark.parametrize("kernel", [8, 2])
def test_mixed_bc_lgf(kernel):
    """config 1 shape at 32^3: x EVEN-ODD, y periodic, z unbounded (1-unbounded LGF_4 / HEJ_2)"""
    n = 32
    rng = np.random.default_rng(7)
    rhs = rng.normal(size=(1, n, n, n))
    rhs -= rhs.mean(axis=2, keepdims=True)
    bc = [[0, 1], [3, 3], [4, 4]]
    h = [1.0 / n] * 3
    f_gpu = gpu_solve([n] * 3, bc, 1, h, [1, 1, 1], 1, kernel, rhs)
    f_ora = orc.solve([n] * 3, bc, 1, h, [1, 1, 1], 1, kernel, rhs)
    e2, ei = rel_err(f_gpu, f_ora)
    assert e2 <= TOL and ei <= TOL, (e2, ei)


FAST_CASES = [
    # (nglob, center, lda, kernel)
    ((64, 64, 64), 1, 1, 2),
    ((128, 128, 128), 1, 1, 4),
    ((64, 32, 128), 1, 1, 2),
    ((65, 65, 65), 0, 1, 2),
    ((64, 64, 64), 1, 3, 8),
    ((256, 32, 64), 1, 1, 0),
]


@pytest.mark.parametrize("case", FAST_CASES, ids=lambda c: "x".join(map(str, c[0])) + "-c%d-l%d-k%d" % c[1:])
def test_fused_pipeline_vs_oracle(case):
    """the fused unbounded pipeline (path 1) against the oracle, cell and node, scalar and vector"""
    nglob, center, lda, kernel = case
    rng = np.random.default_rng(sum(nglob) + lda)
    n0, n1, n2 = nglob
    rhs = rng.normal(size=(lda, n2, n1, n0))
    h = [1.0 / (n - (1 - center)) for n in nglob]
    if kernel in (1, 2, 3, 4, 5, 6, 8, 9, 10):
        h = [h[0]] * 3
    L = [h[d] * (nglob[d] - (1 - center)) for d in range(3)]
    bc = [[4, 4]] * 3
    topo = fl.Topology(0, lda, list(nglob), [1, 1, 1], False)
    s = fl.Solver(topo, bc, h, L, center_type=[center] * 3)
    s.set_green_type(kernel)
    s.setup()
    assert s.path == 1
    f_gpu = np.zeros_like(rhs)
    s.solve(f_gpu, rhs)
    f_ora = orc.solve(list(nglob), bc, lda, h, L, center, kernel, rhs)
    e2, ei = rel_err(f_gpu, f_ora)
    assert e2 <= TOL and ei <= TOL, (e2, ei)


def test_fused_vs_general_and_device_pointers(monkeypatch):
    """same problem through the general path, through library device buffers and torch device tensors"""
    n = 64
    rhs = blobs_rhs(n, 99)
    bc = [[4, 4]] * 3
    h = [1.0 / n] * 3
    outs = []
    for force in (False, True):
        if force:
            monkeypatch.setenv("FLUPS_FORCE_GENERIC", "1")
        topo = fl.Topology(0, 1, [n] * 3, [1, 1, 1], False)
        s = fl.Solver(topo, bc, h, [1, 1, 1])
        s.set_green_type(2)
        s.setup()
        assert s.path == (0 if force else 1)
        f = np.zeros_like(rhs)
        s.solve(f, rhs)
        outs.append(f)
        if not force:
            d_rhs = fl.DeviceArray.from_numpy(rhs)
            d_f = fl.DeviceArray(rhs.shape)
            assert fl.lib().flups_pointer_is_device(d_rhs.ptr) == 1
            s.solve(d_f, d_rhs)
            assert np.array_equal(d_f.numpy(), f)
            s.solve(d_rhs, d_rhs)   # in place: field and rhs alias
            assert np.array_equal(d_rhs.numpy(), f)
            if torch is not None and torch.cuda.is_available():
                t_rhs = torch.from_numpy(rhs).cuda()
                t_f = torch.zeros_like(t_rhs)
                s.solve(t_f, t_rhs)
                torch.cuda.synchronize()
                assert np.array_equal(t_f.cpu().numpy(), f)
    e2, ei = rel_err(outs[0], outs[1])
    assert e2 <= TOL and ei <= TOL, (e2, ei)


@pytest.mark.parametrize("odiff", [1, 2, 4, 6])
@pytest.mark.parametrize("bc", [[[3, 3]] * 3, [[3, 3], [0, 1], [4, 4]], [[4, 4]] * 3], ids=["per", "mix", "unb"])
def test_rot_diff_orders(odiff, bc):
    """lda = 3 ROT with spectral and FD2/4/6 derivative symbols against the oracle"""
    n = 16
    rng = np.random.default_rng(odiff)
    rhs = rng.normal(size=(3, n, n, n))
    h = [1.0 / n] * 3
    f_gpu = gpu_solve([n] * 3, bc, 3, h, [1, 1, 1], 1, 2, rhs, typ=1, odiff=odiff)
    f_ora = orc.solve([n] * 3, bc, 3, h, [1, 1, 1], 1, 2, rhs, typ=1, odiff=odiff)
    e2, ei = rel_err(f_gpu, f_ora)
    assert e2 <= TOL and ei <= TOL, (e2, ei)


FP32_TOL = 1e-5


@pytest.mark.parametrize("case", [((64, 64, 64), 1, 1, 2), ((128, 64, 32), 1, 1, 4), ((65, 65, 65), 0, 1, 2),
                                  ((64, 64, 64), 1, 3, 8)], ids=["64-hej2", "128x64x32", "node65", "lda3-lgf4"])
def test_fused_fp32_vs_oracle(case):
    """FP32 transforms + multiply (FP64-computed Green's function) against the FP64 oracle: relative
    L2 and Linf errors <= 1e-5"""
    nglob, center, lda, kernel = case
    n0, n1, n2 = nglob
    h = [1.0 / (max(nglob) - (1 - center))] * 3
    L = [h[0] * (n - (1 - center)) for n in nglob]
    rhs = np.concatenate([blobs_rhs(max(nglob), 5 + c)[:, :n2, :n1, :n0] for c in range(lda)])
    bc = [[4, 4]] * 3
    topo = fl.Topology(0, lda, list(nglob), [1, 1, 1], False)
    s = fl.Solver(topo, bc, h, L, center_type=[center] * 3)
    s.set_green_type(kernel)
    s.set_precision(True)
    s.setup()
    assert s.path == 2
    f32 = np.zeros_like(rhs)
    s.solve(f32, rhs)
    f_ora = orc.solve(list(nglob), bc, lda, h, L, center, kernel, rhs)
    e2, ei = rel_err(f32, f_ora)
    assert e2 <= FP32_TOL and ei <= FP32_TOL, (e2, ei)
    assert e2 > 1e-12   # really computed in single precision


@pytest.mark.parametrize("nglob", [(64, 64, 64), (128, 64, 32)])
def test_schedules_are_bitwise_equal_and_autotune(monkeypatch, nglob):
    """every schedule of the fused pipeline computes the same arithmetic: results are bitwise
    equal to the default schedule's; the autotuner picks one of them"""
    rng = np.random.default_rng(3)
    n0, n1, n2 = nglob
    rhs = rng.normal(size=(1, n2, n1, n0))
    h = [1.0 / max(nglob)] * 3
    L = [h[0] * n for n in nglob]

    def run(env, tune=False):
        for k in ("FLUPS_ZSCHED", "FLUPS_YSCHED", "FLUPS_ORDER"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, str(v))
        topo = fl.Topology(0, 1, list(nglob), [1, 1, 1], False)
        s = fl.Solver(topo, [[4, 4]] * 3, h, L)
        s.set_green_type(2)
        if tune:
            s.set_autotune(True)
        s.setup()
        assert s.path == 1
        f = np.zeros_like(rhs)
        s.solve(f, rhs)
        return f, s.schedule

    ref, sched = run({})
    assert sched == (20, 0, 12)
    for z in (0, 1, 2, 4, 16, 20, 48, 52):
        for y, o in ((0, 0), (16, 15), (4, 4)):
            f, sc = run({"FLUPS_ZSCHED": z, "FLUPS_YSCHED": y, "FLUPS_ORDER": o})
            assert sc == (z, y, o)
            assert np.array_equal(f, ref), (z, y, o)
    f, sc = run({}, tune=True)
    assert sc[0] in (0, 4, 16, 20) and sc[1] in (0, 16) and sc[2] in (4, 12)
    assert np.array_equal(f, ref)
    f_ora = orc.solve(list(nglob), [[4, 4]] * 3, 1, h, L, 1, 2, rhs)
    e2, ei = rel_err(ref, f_ora)
    assert e2 <= TOL and ei <= TOL


@pytest.mark.parametrize("fp32", [False, True], ids=["fp64", "fp32"])
def test_nontemporal_accesses_are_bitwise_neutral(monkeypatch, fp32):
    """the fused pipeline's non-temporal accesses (stage arrays from FLUPS_NT_MB MiB on) change the
    cache policy only: with every array non-temporal (threshold 0) the result is bitwise the one
    with none (a threshold above the problem), on aligned user rows (16-byte row modes)"""
    nglob = (64, 32, 128)
    rng = np.random.default_rng(11)
    rhs = rng.normal(size=(1, nglob[2], nglob[1], nglob[0]))
    h = [1.0 / max(nglob)] * 3
    L = [h[0] * n for n in nglob]
    out = {}
    for mb in ("0", "100000"):
        monkeypatch.setenv("FLUPS_NT_MB", mb)
        topo = fl.Topology(0, 1, list(nglob), [1, 1, 1], False)
        s = fl.Solver(topo, [[4, 4]] * 3, h, L)
        s.set_green_type(2)
        if fp32:
            s.set_precision(1)
        s.setup()
        assert s.path == (2 if fp32 else 1)
        f = np.zeros_like(rhs)
        s.solve(f, rhs)
        out[mb] = f
    assert np.array_equal(out["0"], out["100000"])
    f_ora = orc.solve(list(nglob), [[4, 4]] * 3, 1, h, L, 1, 2, rhs)
    e2, ei = rel_err(out["0"], f_ora)
    tol = FP32_TOL if fp32 else TOL
    assert e2 <= tol and ei <= tol, (e2, ei)


@pytest.mark.parametrize("case", [((64, 32, 32), 1, 1), ((64, 32, 64), 1, 2), ((65, 33, 33), 0, 4), ((64, 64, 32), 1, 6)],
                         ids=["64x32x32-spe", "64x32x64-fd2", "node65x33x33-fd4", "64x64x32-fd6"])
def test_fused_rot_vs_oracle(case):
    """lda = 3 rotational solve through the fused pipeline (curl in registers between the z transforms)"""
    nglob, center, odiff = case
    rng = np.random.default_rng(sum(nglob) + odiff)
    n0, n1, n2 = nglob
    rhs = rng.normal(size=(3, n2, n1, n0))
    h = [1.0 / (max(nglob) - (1 - center))] * 3
    L = [h[0] * (n - (1 - center)) for n in nglob]
    bc = [[4, 4]] * 3
    topo = fl.Topology(0, 3, list(nglob), [1, 1, 1], False)
    s = fl.Solver(topo, bc, h, L, order_diff=odiff, center_type=[center] * 3)
    s.set_green_type(4)
    s.setup()
    assert s.path == 1
    f_gpu = np.zeros_like(rhs)
    s.solve(f_gpu, rhs, 1)
    f_ora = orc.solve(list(nglob), bc, 3, h, L, center, 4, rhs, typ=1, odiff=odiff)
    e2, ei = rel_err(f_gpu, f_ora)
    assert e2 <= TOL and ei <= TOL, (e2, ei)
    # the STD solve of the same solver still runs the fused STD pass
    f_std = np.zeros_like(rhs)
    s.solve(f_std, rhs, 0)
    e2, ei = rel_err(f_std, orc.solve(list(nglob), bc, 3, h, L, center, 4, rhs))
    assert e2 <= TOL and ei <= TOL, (e2, ei)


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("odiff", [1, 4])
def test_fused_rot_long_z_vs_oracle(fp32, odiff):
    """the rotational pass at M2 = 1024 (tiles narrower than the Green's layout, symbols from the
    host tables): oracle parity at 1e-12 (FP32: 1e-5)"""
    nglob = (16, 16, 512)
    rng = np.random.default_rng(77 + odiff)
    rhs = rng.normal(size=(3, 512, 16, 16))
    h = [1.0 / 512] * 3
    L = [h[0] * n for n in nglob]
    bc = [[4, 4]] * 3
    topo = fl.Topology(0, 3, list(nglob), [1, 1, 1], False)
    s = fl.Solver(topo, bc, h, L, order_diff=odiff)
    s.set_green_type(4)
    if fp32:
        s.set_precision(True)
    s.setup()
    assert s.path == (2 if fp32 else 1)
    f_gpu = np.zeros_like(rhs)
    s.solve(f_gpu, rhs, 1)
    f_ora = orc.solve(list(nglob), bc, 3, h, L, 1, 4, rhs, typ=1, odiff=odiff)
    e2, ei = rel_err(f_gpu, f_ora)
    tol = 1e-5 if fp32 else TOL
    assert e2 <= tol and ei <= tol, (e2, ei)


PER_CASES = [
    # (nglob, bc, kernel, lda)
    ((64, 32, 32), [[3, 3], [4, 4], [4, 4]], 2, 1),
    ((64, 64, 32), [[3, 3], [4, 4], [4, 4]], 8, 1),
    ((32, 64, 64), [[3, 3], [3, 3], [4, 4]], 4, 1),
    ((64, 32, 64), [[3, 3], [3, 3], [3, 3]], 0, 1),
    ((32, 32, 32), [[4, 4], [4, 4], [4, 4]], 2, 3),
    ((128, 32, 32), [[3, 3], [3, 3], [3, 3]], 13, 1),
]


@pytest.mark.parametrize("case", PER_CASES, ids=["perx-hej2", "perx-lgf4", "perxy-hej6", "per-chat", "unb32-lda3",
                                                 "per-mehr4f"])
def test_fused_periodic_and_small_vs_oracle(case):
    """the fused pipeline with periodic directions (no padding) and with 32-point transforms"""
    nglob, bc, kernel, lda = case
    rng = np.random.default_rng(sum(nglob) + kernel)
    n0, n1, n2 = nglob
    rhs = rng.normal(size=(lda, n2, n1, n0))
    rhs -= rhs.mean(axis=(1, 2, 3), keepdims=True)
    h = [1.0 / max(nglob)] * 3
    L = [h[0] * n for n in nglob]
    topo = fl.Topology(0, lda, list(nglob), [1, 1, 1], False)
    s = fl.Solver(topo, bc, h, L)
    s.set_green_type(kernel)
    s.setup()
    assert s.path == 1
    f_gpu = np.zeros_like(rhs)
    s.solve(f_gpu, rhs)
    f_ora = orc.solve(list(nglob), bc, lda, h, L, 1, kernel, rhs)
    e2, ei = rel_err(f_gpu, f_ora)
    assert e2 <= TOL and ei <= TOL, (e2, ei)


HYBRID_CASES = [
    # (nglob, bc, center, kernel, lda): the last direction (z) c2c of a power-of-two length
    ((64, 64, 64), [[0, 1], [3, 3], [4, 4]], 1, 8, 1),      # x EVEN-ODD, y PER, z UNB (mixed config)
    ((48, 40, 32), [[4, 4], [0, 4], [4, 4]], 1, 2, 1),      # y semi-unbounded first, x r2c, z fused
    ((33, 17, 33), [[1, 4], [4, 4], [4, 4]], 0, 4, 3),      # node, x semi-unbounded, lda 3
    ((32, 24, 64), [[1, 1], [3, 3], [3, 3]], 1, 0, 1),      # x ODD-ODD, y PER (r2c), z PER
    # 2-D: the last direction is y (columns of up to 2048 points)
    ((256, 512, 1), [[4, 4], [4, 4], [9, 9]], 1, 2, 1),     # UNB 2-D, y 1024 padded
    ((128, 1024, 1), [[3, 3], [4, 4], [9, 9]], 1, 4, 1),    # x PER, y UNB: columns of 2048
    ((96, 128, 1), [[3, 3], [3, 3], [9, 9]], 1, 0, 2),      # fully periodic 2-D, lda 2
    ((80, 64, 1), [[4, 4], [4, 4], [9, 9]], 1, 8, 1),       # UNB 2-D, LGF_4, x of 160 (generic pass)
    # last direction of 3 times a power of two (mixed-radix fused column)
    ((64, 32, 48), [[0, 1], [3, 3], [3, 3]], 1, 0, 1),      # z PER 48
    ((48, 40, 48), [[4, 4], [4, 4], [4, 4]], 1, 2, 1),      # all UNB, paired y rows, z of 96
    ((32, 24, 96), [[0, 0], [4, 4], [4, 4]], 1, 4, 3),      # lda 3, z of 192
    ((64, 48, 1), [[4, 4], [4, 4], [9, 9]], 1, 2, 1),       # 2-D, y of 96
    # last direction of 5 or 7 times a power of two (thread teams over LDS-staged columns)
    ((48, 32, 80), [[0, 1], [3, 3], [3, 3]], 1, 0, 1),      # z PER 80
    ((32, 24, 112), [[4, 4], [4, 4], [4, 4]], 1, 2, 1),     # all UNB, z of 224
    ((32, 32, 40), [[0, 0], [3, 3], [4, 4]], 1, 2, 3),      # lda 3, z of 80
    ((32, 24, 224), [[4, 4], [4, 4], [4, 4]], 1, 2, 1),     # z of 448 (7 x 64): four-column tiles
    ((16, 16, 448), [[3, 3], [3, 3], [4, 4]], 1, 2, 1),     # z of 896 (7 x 128): four-column tiles, 56 KB of LDS
    ((64, 80, 1), [[3, 3], [4, 4], [9, 9]], 1, 2, 1),       # 2-D, y of 160 (5 x 32)
    # r2c last direction after real-to-real ones: r2c, G and c2r of paired real columns in one launch
    ((64, 32, 32), [[0, 1], [0, 0], [4, 4]], 1, 8, 1),      # x EO, y EE, z UNB
    ((33, 48, 64), [[1, 1], [0, 1], [3, 3]], 1, 0, 1),      # odd x count (an unpaired column), z PER
    ((32, 32, 32), [[0, 0], [1, 1], [4, 4]], 1, 2, 3),      # lda 3
    ((64, 128, 1), [[0, 0], [4, 4], [9, 9]], 1, 2, 1),      # 2-D, y UNB
    ((33, 33, 33), [[0, 0], [1, 1], [4, 4]], 0, 2, 1),      # node-centred
    # x last (r2c rows after real-to-real y, z): r2c, G and c2r of each row in LDS
    ((64, 32, 32), [[3, 3], [1, 1], [0, 0]], 1, 2, 1),      # x PER
    ((48, 32, 16), [[4, 4], [0, 0], [1, 1]], 1, 2, 1),      # x UNB, rows of 96 (3 x 32)
    ((80, 32, 1), [[3, 3], [0, 1], [9, 9]], 1, 0, 2),       # 2-D, lda 2, rows of 80 (5 x 16)
    ((33, 17, 17), [[4, 4], [0, 0], [1, 1]], 0, 2, 1),      # node-centred x UNB
    ((80, 16, 16), [[3, 3], [1, 1], [0, 0]], 1, 2, 1),      # x PER, rows of 80 (5 x 8): c2r built in LDS first
    ((112, 16, 16), [[3, 3], [0, 0], [1, 1]], 1, 0, 1),     # x PER, rows of 112 (7 x 8)
]


@pytest.mark.parametrize("case", HYBRID_CASES, ids=["mixed64", "semiunb-y", "node-lda3", "sym-per-per", "2d-unb-1024",
                                                    "2d-per-unb-2048", "2d-per-lda2", "2d-unb-lgf4", "pf3-per48", "pf3-unb-pairs",
                                                    "pf3-lda3-192", "pf3-2d-96", "pf5-per80", "pf7-unb224",
                                                    "pf5-lda3", "pf7-448-tx4", "pf7-896-tx4", "pf5-2d-160",
                                                    "rz-eo-ee-unb", "rz-odd-per",
                                                    "rz-lda3", "rz-2d", "rz-node", "xrow-per", "xrow-unb-3x32",
                                                    "xrow-2d-lda2", "xrow-node", "xrow-per-5x8", "xrow-per-7x8"])
def test_hybrid_last_direction_vs_general_and_oracle(case, monkeypatch):
    """general passes + the last direction fused with the multiply, against the all-general path and the oracle"""
    nglob, bc, center, kernel, lda = case
    rng = np.random.default_rng(sum(nglob))
    n0, n1, n2 = nglob
    rhs = rng.normal(size=(lda, n2, n1, n0))
    rhs -= rhs.mean(axis=(1, 2, 3), keepdims=True)
    h = [1.0 / (max(nglob) - (1 - center))] * 3
    L = [h[0] * (n - (1 - center)) for n in nglob]
    outs = []
    for force in (False, True):
        if force:
            monkeypatch.setenv("FLUPS_FORCE_GENERIC", "1")
        topo = fl.Topology(0, lda, list(nglob), [1, 1, 1], False)
        s = fl.Solver(topo, bc, h, L, center_type=[center] * 3)
        s.set_green_type(kernel)
        s.setup()
        assert s.path == (0 if force else 3)
        f = np.zeros_like(rhs)
        s.solve(f, rhs)
        outs.append(f)
    e2, ei = rel_err(outs[0], outs[1])
    assert e2 <= TOL and ei <= TOL, (e2, ei)
    f_ora = orc.solve(list(nglob), bc, lda, h, L, center, kernel, rhs)
    e2, ei = rel_err(outs[0], f_ora)
    assert e2 <= TOL and ei <= TOL, (e2, ei)


SLAB_CASES = [
    # (nglob, bc, lda, kernel, typ, odiff, fp32)
    ((128, 64, 64), [[4, 4]] * 3, 1, 2, 0, 0, False),
    ((128, 64, 32), [[3, 3], [4, 4], [4, 4]], 1, 8, 0, 0, False),
    ((128, 64, 64), [[4, 4]] * 3, 3, 4, 1, 1, False),
    ((128, 64, 64), [[4, 4]] * 3, 1, 2, 0, 0, True),
]


@pytest.mark.parametrize("case", SLAB_CASES, ids=["unb", "perx-lgf4", "rot-lda3", "unb-fp32"])
def test_fused_x_slabs_bitwise_and_oracle(case, monkeypatch):
    """the fused pipeline split into x slabs (FLUPS_SLAB_MB: the y-transformed spectrum in
    slabs of about that many MiB, run one after the other) computes the same arithmetic per
    column as one slab: bitwise equal results, and the oracle's to 1e-12 (FP32: 1e-5)"""
    nglob, bc, lda, kernel, typ, odiff, fp32 = case
    rng = np.random.default_rng(21)
    n0, n1, n2 = nglob
    rhs = rng.normal(size=(lda, n2, n1, n0))
    if any(b == [3, 3] for b in bc):
        rhs -= rhs.mean(axis=3, keepdims=True)
    h = [1.0 / max(nglob)] * 3
    L = [h[0] * n for n in nglob]

    def run(slab_mb):
        monkeypatch.setenv("FLUPS_SLAB_MB", str(slab_mb))
        topo = fl.Topology(0, lda, list(nglob), [1, 1, 1], False)
        s = fl.Solver(topo, bc, h, L, order_diff=odiff)
        s.set_green_type(kernel)
        if fp32:
            s.set_precision(True)
        s.setup()
        assert s.path == (2 if fp32 else 1)
        f = np.zeros_like(rhs)
        s.solve(f, rhs, typ)
        return f

    one = run(0)
    # one x column of S1: M1 * n2 complex values per component; 0.3 MiB slabs give several slabs
    for mb in (0.3, 4):
        assert np.array_equal(run(mb), one), mb
    ora = orc.solve(list(nglob), bc, lda, h, L, 1, kernel, rhs, typ=typ, odiff=odiff)
    e2, ei = rel_err(one, ora)
    tol = 1e-5 if fp32 else TOL
    assert e2 <= tol and ei <= tol, (e2, ei)


ROW_CASES = [
    # (nglob, bc, center, kernel, fast kernel that must run: flups_get_pass_kernels codes)
    # symmetric x (cell-centred: REDFT10/01, RODFT10/01, REDFT11, RODFT11)
    ((64, 32, 32), [[0, 0], [3, 3], [4, 4]], 1, 2, 1),
    ((64, 32, 32), [[1, 1], [3, 3], [4, 4]], 1, 2, 1),
    ((256, 16, 16), [[0, 1], [3, 3], [4, 4]], 1, 8, 1),
    ((32, 32, 32), [[1, 0], [4, 4], [3, 3]], 1, 0, 1),
    ((128, 16, 16), [[0, 1], [1, 0], [0, 0]], 1, 0, 1),
    ((64, 24, 16), [[1, 1], [0, 1], [4, 4]], 1, 4, 1),
    # periodic y or z after symmetric directions: r2c / c2r of paired real columns
    ((32, 64, 32), [[0, 0], [3, 3], [4, 4]], 1, 2, 2),
    ((33, 64, 16), [[1, 1], [3, 3], [4, 4]], 1, 0, 2),
    ((64, 32, 128), [[0, 1], [1, 0], [3, 3]], 1, 0, 2),
    ((256, 256, 8), [[0, 1], [3, 3], [0, 0]], 1, 0, 2),
    # semi-unbounded x (transforms of length 2N over the zero-padded, possibly shifted field)
    ((32, 32, 32), [[0, 4], [3, 3], [4, 4]], 1, 2, 1),
    ((64, 16, 32), [[4, 1], [4, 4], [3, 3]], 1, 4, 1),
    ((32, 64, 16), [[1, 4], [0, 0], [4, 4]], 1, 0, 4),
    # unbounded y after a symmetric x: zero-padded r2c / truncated c2r of paired real columns
    ((32, 64, 32), [[0, 0], [4, 4], [3, 3]], 1, 2, 2),
    ((64, 32, 32), [[1, 0], [4, 4], [4, 4]], 1, 8, 2),
    # symmetric y first, then a contiguous r2c / c2r along x (packed half-length FFT)
    ((64, 32, 32), [[4, 4], [0, 0], [4, 4]], 1, 2, 4),
    ((64, 32, 32), [[3, 3], [1, 1], [4, 4]], 1, 0, 4),
    # node-centred periodic: r2c with the periodic copy
    ((65, 33, 17), [[3, 3], [3, 3], [3, 3]], 0, 0, 3),
    # symmetric y or z on strided pencils: tiles of neighbouring pencils (ragged last tile,
    # semi-unbounded shift, long pencils with narrow tiles)
    ((32, 32, 128), [[4, 4], [3, 3], [0, 1]], 1, 2, 4),
    ((33, 64, 16), [[3, 3], [1, 1], [4, 4]], 1, 0, 4),
    ((32, 64, 32), [[3, 3], [0, 4], [4, 4]], 1, 2, 4),
    ((20, 16, 64), [[4, 4], [4, 1], [1, 0]], 1, 4, 4),
    ((16, 32, 512), [[4, 4], [4, 4], [1, 1]], 1, 0, 4),
    ((8, 2048, 4), [[3, 3], [0, 0], [3, 3]], 1, 0, 4),
    # complex-to-complex along strided directions (zero-padded / truncated, plain periodic, 2-D)
    ((64, 128, 24), [[4, 4], [4, 4], [4, 4]], 1, 2, 5),
    ((48, 32, 20), [[3, 3], [3, 3], [3, 3]], 1, 0, 5),
    ((256, 512, 1), [[4, 4], [4, 4], [9, 9]], 1, 2, 5),
    ((128, 1024, 1), [[3, 3], [4, 4], [9, 9]], 1, 4, 5),
    # strided c2c of 3 or 5 times a power of two (small DFTs, then power-of-two column FFTs)
    ((32, 48, 40), [[3, 3], [4, 4], [3, 3]], 1, 2, 5),
    ((96, 56, 24), [[4, 4], [3, 3], [4, 4]], 1, 8, 5),
    ((96, 80, 1), [[4, 4], [4, 4], [9, 9]], 1, 2, 5),
    ((16, 80, 160), [[3, 3], [3, 3], [3, 3]], 1, 0, 5),
    # contiguous r2c / c2r of 3, 5, 7 times a power of two (teams of threads, one per factor)
    ((48, 32, 32), [[3, 3], [0, 0], [4, 4]], 1, 2, 3),
    ((80, 32, 16), [[4, 4], [1, 1], [4, 4]], 1, 0, 3),
    ((112, 16, 16), [[4, 4], [0, 0], [0, 0]], 1, 2, 3),
    ((49, 17, 17), [[3, 3], [3, 3], [3, 3]], 0, 0, 3),
    # real-to-real pencils of 3, 5, 7 times a power of two: rows (x) and strided tiles (y, z)
    ((48, 32, 32), [[0, 1], [3, 3], [4, 4]], 1, 8, 1),
    ((96, 32, 16), [[0, 0], [4, 4], [3, 3]], 1, 2, 1),
    ((80, 16, 16), [[1, 1], [3, 3], [4, 4]], 1, 0, 1),
    ((112, 16, 16), [[1, 0], [4, 4], [4, 4]], 1, 4, 1),
    ((48, 32, 16), [[0, 4], [3, 3], [3, 3]], 1, 2, 1),
    ((32, 48, 32), [[4, 4], [0, 0], [3, 3]], 1, 2, 4),
    ((16, 32, 160), [[3, 3], [4, 4], [1, 0]], 1, 0, 4),
    ((32, 56, 16), [[3, 3], [4, 1], [4, 4]], 1, 4, 4),
    # paired real columns of 3 or 5 times a power of two (periodic / unbounded y, odd pair count)
    ((32, 48, 32), [[0, 0], [3, 3], [4, 4]], 1, 2, 2),
    ((32, 40, 32), [[1, 1], [4, 4], [3, 3]], 1, 2, 2),
    ((33, 96, 16), [[1, 0], [3, 3], [4, 4]], 1, 0, 2),
    # strided c2c of 7 times a power of two (thread teams over the LDS-staged columns)
    ((16, 56, 32), [[3, 3], [3, 3], [4, 4]], 1, 2, 5),
    ((16, 32, 112), [[4, 4], [3, 3], [4, 4]], 1, 0, 5),
    # node-centred symmetric and semi-unbounded directions: DCT-I / DST-I through the packed
    # real FFT of the 2(n -+ 1)-point extension (rows and strided tiles)
    ((33, 17, 17), [[0, 0], [3, 3], [4, 4]], 0, 2, 1),
    ((17, 33, 17), [[4, 4], [1, 1], [3, 3]], 0, 0, 4),
    ((65, 33, 17), [[1, 1], [0, 0], [4, 4]], 0, 4, 1),
    ((33, 17, 17), [[0, 4], [3, 3], [3, 3]], 0, 2, 1),
    ((33, 17, 17), [[4, 1], [3, 3], [3, 3]], 0, 2, 1),
    ((49, 17, 25), [[0, 0], [3, 3], [1, 1]], 0, 0, 1),
    # c2c along x after an r2c along y or z (contiguous complex pencils, zero padding / truncation)
    ((64, 32, 16), [[4, 4], [3, 3], [0, 0]], 1, 2, 6),
    ((48, 32, 32), [[4, 4], [3, 3], [4, 4]], 1, 2, 6),
    ((40, 64, 16), [[4, 4], [4, 4], [3, 3]], 1, 0, 6),
    ((56, 64, 1), [[4, 4], [3, 3], [9, 9]], 1, 2, 6),
    # node-centred periodic c2c along y or z: the periodic copy (position n = X[0])
    ((65, 65, 17), [[3, 3], [3, 3], [4, 4]], 0, 2, 5),
    ((65, 33, 33), [[4, 4], [3, 3], [3, 3]], 0, 2, 5),
    ((65, 33, 1), [[3, 3], [3, 3], [9, 9]], 0, 0, 5),
    # node-centred periodic r2c along y or z (paired real columns with the periodic copy)
    ((17, 65, 33), [[0, 0], [3, 3], [3, 3]], 0, 2, 2),
    ((33, 17, 65), [[1, 1], [4, 4], [3, 3]], 0, 0, 2),
    # mixed radix with one thread per team (Q = 8): the c2r and REDFT01 sequences built in LDS first
    ((80, 16, 16), [[3, 3], [0, 0], [4, 4]], 1, 2, 3),
    ((112, 16, 16), [[3, 3], [1, 1], [4, 4]], 1, 0, 3),
    ((112, 16, 16), [[0, 0], [3, 3], [4, 4]], 1, 2, 1),
    ((16, 80, 16), [[3, 3], [0, 0], [4, 4]], 1, 2, 4),
]


@pytest.mark.parametrize("case", ROW_CASES, ids=["ee", "oo", "eo-256", "oe", "sym3", "oo-mix", "ee-pery", "odd-x-pery",
                                                 "perz-128", "pery-256", "even-unb", "unb-odd", "odd-unb-ee", "ee-unby-perz",
                                                 "oe-unby-unbz", "unbx-eey", "perx-ooy", "node-per", "eoz-128", "ragged-ooy",
                                                 "eunb-y", "unbo-y-oez", "ooz-512", "eey-2048", "c2c-y-256", "c2c-y-32",
                                                 "c2c-2d-1024", "c2c-2d-2048", "c2c-5x8-3x32", "c2c-3x16", "2d-5x32",
                                                 "c2c-5x16-5x32", "rc-3x8", "rc-5x16", "rc-7x16", "rc-node-3x8",
                                                 "r2r-eo-3x8", "r2r-ee-3x16", "r2r-oo-5x8", "r2r-oe-7x8", "r2r-eu-3x16",
                                                 "col-ee-3x8", "col-oe-5x16", "col-uo-7x8", "real-3x16", "real-unb-5x16",
                                                 "real-odd-3x32", "c2c-7x8", "c2c-unb-7x32", "node-ee-32",
                                                 "node-oo-col-32", "node-oo-ee-64", "node-eu-64", "node-uo-64",
                                                 "node-ee-3x16-oo-3x8", "rowc2c-128", "rowc2c-3x32", "rowc2c-5x16",
                                                 "rowc2c-2d-7x16", "node-per-c2c-y", "node-per-c2c-z", "node-per-2d",
                                                 "node-per-real-y", "node-per-real-z", "rc-5x8", "rc-7x8",
                                                 "r2r-ee-7x8", "col-ee-5x8"])
@pytest.mark.parametrize("misalign", [False, True])
def test_fast_pass_kernels_vs_generic_and_oracle(case, misalign, monkeypatch):
    """the fast kernels of the general path (real-to-real pencils and tiles, paired real columns,
    packed r2c / c2r rows, complex columns) against the generic pass kernel (FLUPS_GENERIC_PASS)
    and the oracle at 1e-12; misaligned user rows take the scalar loads and stores"""
    nglob, bc, center, kernel, want = case
    rng = np.random.default_rng(31)
    n0, n1, n2 = nglob
    rhs = rng.normal(size=(1, n2, n1, n0))
    rhs -= rhs.mean()
    h = [1.0 / (max(nglob) - (1 - center))] * 3
    L = [h[0] * (m - (1 - center)) for m in nglob]

    # every direction through the pass kernels (no fused last direction)
    monkeypatch.setenv("FLUPS_FORCE_GENERIC", "1")

    def run(generic):
        if generic:
            monkeypatch.setenv("FLUPS_GENERIC_PASS", "1")
        else:
            monkeypatch.delenv("FLUPS_GENERIC_PASS", raising=False)
        topo = fl.Topology(0, 1, list(nglob), [1, 1, 1], False)
        s = fl.Solver(topo, bc, h, L, center_type=[center] * 3)
        s.set_green_type(kernel)
        s.setup()
        kinds = s.pass_kernels
        if generic:
            assert all(k <= 0 for k in kinds), kinds
        else:
            assert want in kinds, (want, kinds)
            # a direction's forward and backward passes are both fast or both generic
            nd = len(kinds) // 2
            assert all((kinds[i] == 0) == (kinds[nd + i] == 0) for i in range(nd)), kinds
        n = rhs.size
        off = 1 if misalign else 0   # one double in: rows start 8 bytes off a 16-byte boundary
        flat = np.zeros(n + 1)
        flat[off:off + n] = rhs.ravel()
        buf_in = fl.DeviceArray.from_numpy(flat)
        buf_out = fl.DeviceArray((n + 1,))
        fl.lib().flups_clear_last_error()
        fl.lib().flups_solve(s._solver, buf_out.ptr + 8 * off, buf_in.ptr + 8 * off, 0)
        assert fl.last_error() == "", fl.last_error()
        return buf_out.numpy()[off:off + n].reshape(rhs.shape)

    fast = run(False)
    gen = run(True)
    e2, ei = rel_err(fast, gen)
    assert e2 <= TOL and ei <= TOL, ("fast vs generic", e2, ei)
    ora = orc.solve(list(nglob), bc, 1, h, L, center, kernel, rhs)
    e2, ei = rel_err(fast, ora)
    assert e2 <= TOL and ei <= TOL, ("fast vs oracle", e2, ei)
