"""GPU tests of the advanced interface (src/flups.h:467-578): the solve's steps one by one.

  * do_copy + do_FFT + do_mult + do_FFT(backward) + do_copy reproduce flups_solve
    (bitwise on the general path, <= 1e-12 against the fused one and the oracle);
  * the spectral layout described by flups_get_innerTopo_spectral, indexed with
    flups_symID / flups_get_spectralInfo as in samples/solve_advanced_C/
    main_solve_advanced.c:143-196, is the one the transforms use: a Gaussian
    filter applied entry by entry gives the analytic filtered field.
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


def symid(i, symstart):
    """flups_symID for one axis: index mirrored about symstart, at least 1 (src/flups.h:170-184)"""
    i = np.asarray(i)
    if symstart == 0.0:
        return i
    m = np.maximum(np.abs(2.0 * symstart - i).astype(int), 1)
    return np.where(i <= symstart, i, m)


CASES = [
    # (nglob, bc, lda, center, kernel, type, odiff)
    ((32, 24, 16), [[0, 1], [3, 3], [4, 4]], 1, 1, 8, 0, 0),
    ((32, 32, 32), [[4, 4], [4, 4], [4, 4]], 1, 1, 2, 0, 0),
    ((17, 17, 17), [[1, 4], [4, 4], [3, 3]], 1, 0, 4, 0, 0),
    ((16, 16, 16), [[3, 3], [4, 4], [4, 4]], 3, 1, 2, 1, 1),
    ((16, 16, 16), [[3, 3], [3, 3], [3, 3]], 3, 1, 0, 1, 4),
]


@pytest.mark.parametrize("case", CASES, ids=["mixed-lgf4", "unb-fused", "node-mixed", "rot-spe", "rot-fd4"])
def test_advanced_steps_reproduce_solve(case):
    nglob, bc, lda, center, kernel, typ, odiff = case
    rng = np.random.default_rng(11)
    n0, n1, n2 = nglob
    rhs = rng.normal(size=(lda, n2, n1, n0))
    if any(b == [3, 3] for b in bc):
        rhs -= rhs.mean(axis=(1, 2, 3), keepdims=True)
    h = [1.0 / (max(nglob) - (1 - center))] * 3
    L = [h[0] * (n - (1 - center)) for n in nglob]
    topo = fl.Topology(0, lda, list(nglob), [1, 1, 1], False)
    s = fl.Solver(topo, bc, h, L, order_diff=odiff, center_type=[center] * 3)
    s.set_green_type(kernel)
    s.setup()
    ref = np.zeros_like(rhs)
    s.solve(ref, rhs, typ)
    inner = s.inner_buffer()
    phys = s.inner_topo_physical()
    assert [phys.get_nglob(d) for d in range(3)] == list(nglob)
    s.skip_first_switchtopo()   # accepted; the physical layout stays the user topology
    s.do_copy(phys, rhs, fl.FORWARD)
    s.do_fft(inner, fl.FORWARD)
    s.do_mult(inner, typ)
    s.do_fft(inner, fl.BACKWARD if typ == 0 else fl.BACKWARD_DIFF)
    out = np.zeros_like(rhs)
    s.do_copy(phys, out, fl.BACKWARD)
    if s.path == 0:
        assert np.array_equal(out, ref)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(out - ref)) <= 1e-12 * scale
    ora = orc.solve(list(nglob), bc, lda, h, L, center, kernel, rhs, typ=typ, odiff=odiff)
    assert np.max(np.abs(out - ora)) <= 1e-12 * np.max(np.abs(ora))


def test_spectral_layout_gaussian_filter():
    n, eps = 16, 1e-3
    h = 1.0 / n
    x = (np.arange(n) + 0.5) * h
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    m1 = np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y) * np.sin(2 * np.pi * Z)
    m2 = np.sin(4 * np.pi * X) * np.cos(2 * np.pi * Y) * np.sin(6 * np.pi * Z)
    rhs = (m1 + 0.5 * m2)[None]
    topo = fl.Topology(0, 1, [n] * 3, [1, 1, 1], False)
    s = fl.Solver(topo, [[3, 3]] * 3, [h] * 3, [1.0] * 3)
    s.set_green_type(0)
    s.setup()
    inner = s.inner_buffer()
    spec_t = s.inner_topo_spectral()
    ax0 = spec_t.axis
    assert ax0 == 0 and spec_t.is_complex
    nmem = [spec_t.get_nmem(d) for d in range(3)]
    nloc = [spec_t.get_nloc(d) for d in range(3)]
    kfact, koffset, symstart = s.get_spectral_info()
    np.testing.assert_allclose(kfact, 2 * np.pi)
    # forward + backward without touching the spectrum: the field times the transform lengths
    inner.upload(rhs)
    s.do_fft(inner, fl.FORWARD)
    s.do_fft(inner, fl.BACKWARD)
    plain = inner.numpy()[: rhs.size].reshape(rhs.shape)
    np.testing.assert_allclose(plain, rhs * n ** 3, atol=1e-10 * n ** 3, rtol=0)
    # filter exp(-eps |k|^2) entry by entry, wave numbers from symID / spectral info
    inner.upload(rhs)
    s.do_fft(inner, fl.FORWARD)
    spec = inner.numpy()
    cnt = nmem[0] * nmem[1] * nmem[2] * 2
    a = spec[:cnt].reshape(nmem[2], nmem[1], nmem[0], 2)
    i0, i1, i2 = np.arange(nloc[0]), np.arange(nloc[1]), np.arange(nloc[2])
    k0 = (symid(i0, symstart[0]) + koffset[0]) * kfact[0]
    k1 = (symid(i1, symstart[1]) + koffset[1]) * kfact[1]
    k2 = (symid(i2, symstart[2]) + koffset[2]) * kfact[2]
    ksq = k2[:, None, None] ** 2 + k1[None, :, None] ** 2 + k0[None, None, :] ** 2
    a[: nloc[2], : nloc[1], : nloc[0], :] *= np.exp(-eps * ksq)[..., None]
    spec[:cnt] = a.ravel()
    inner.upload(spec)
    s.do_fft(inner, fl.BACKWARD)
    filt = inner.numpy()[: rhs.size].reshape(rhs.shape)
    want = (m1 * np.exp(-eps * 3 * (2 * np.pi) ** 2) + 0.5 * m2 * np.exp(-eps * 56 * np.pi ** 2))[None] * n ** 3
    np.testing.assert_allclose(filt, want, atol=1e-10 * n ** 3, rtol=0)


def test_advanced_rejects_host_buffers():
    topo = fl.Topology(0, 1, [16] * 3, [1, 1, 1], False)
    s = fl.Solver(topo, [[4, 4]] * 3, [1 / 16] * 3, [1.0] * 3)
    s.set_green_type(2)
    s.setup()
    with pytest.raises(RuntimeError):
        s.do_fft(np.zeros(s.get_alloc_size()), fl.FORWARD)
    assert "device" in fl.last_error()
