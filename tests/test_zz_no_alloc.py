"""The allocation-free apply (dielectric_fft_apply_noalloc) against the apply it
is a variant of (dielectric_fft_apply_ex), through the C-ABI: FP64 and FP32,
with and without the fused Jacobi scaling, two calls in a row on systems of
decreasing size (the second call reuses the per-device scratch the first one
filled), a volume-only system, and the refusals (a system too large for the
scratch, stale device copies)."""
import ctypes

import numpy as np
import pytest

from nec2c_amd import vie, workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def system(shape, conductors=True, holes=False, seed=3):
    pitch = 0.5 / 64
    centers = W.cube_cells(shape, pitch)
    eps = W.random_permittivity(shape, seed).reshape(-1)
    if holes:
        keep = np.ones(len(centers), bool)
        keep[::7] = False
        centers, eps = centers[keep], eps[keep]
    k0 = W.k0_of(300e6)
    dip = W.Dipole(center=(-0.1, 0.25, 0.25))
    c = dip.segments if conductors else 0
    s = vie.System(centers, [vie.MaterialSpec(1, 1.5)], c, k0, pitch=pitch)
    s.set_permittivities(eps)
    if conductors:
        B, D = W.coupling_blocks(dip, centers, pitch, eps, k0)
        s.set_blocks(C=dip.impedance_matrix(k0), B=B, D=D)
    return s


def noalloc(s, x, precision, jacobi):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.full(s.order, np.nan + 0j)
    print(f"noalloc apply: order {s.order}, precision {precision}, jacobi {jacobi}", flush=True)
    ok = vie.lib().dielectric_fft_apply_noalloc(s.s, int(precision), int(jacobi), x.ctypes.data, y.ctypes.data)
    return bool(ok), y


def vec(n, seed):
    g = np.random.default_rng(seed)
    return g.standard_normal(n) + 1j * g.standard_normal(n)


@pytest.mark.parametrize("precision,tol", [(vie.FP64, 1e-13), (vie.MIXED, 1e-6)])
def test_noalloc_matches_apply_ex_on_two_calls(precision, tol):
    big = system((14, 12, 10), holes=True)
    small = system((6, 5, 4))
    for s, seed in ((big, 1), (small, 2)):
        x = vec(s.order, seed)
        for jacobi in (False, True):
            ref = s.apply_ex(x, precision, jacobi)          # also makes the device copies current
            need = vie.lib().dielectric_fft_apply_noalloc_bytes(s.s, int(precision))
            assert 0 < need <= 96 << 20
            ok, y = noalloc(s, x, precision, jacobi)
            assert ok, vie.last_error()
            assert rel(y, ref) < tol, (s.order, jacobi, rel(y, ref))
    big.close()
    small.close()


def test_noalloc_volume_only_and_second_smaller_call():
    a = system((10, 9, 8), conductors=False)
    b = system((4, 4, 3), conductors=False)
    for s in (a, b, a):
        x = vec(s.order, 5)
        ref = s.apply_ex(x, vie.FP64)
        ok, y = noalloc(s, x, vie.FP64, False)
        assert ok, vie.last_error()
        assert rel(y, ref) < 1e-13
    a.close()
    b.close()


def test_noalloc_refusals():
    s = system((6, 5, 4))
    x = vec(s.order, 7)
    s.apply_ex(x, vie.FP64)
    vie.lib().dielectric_fft_entry(s.s, 0, 0)   # a block write makes the device copies stale
    ok, _ = noalloc(s, x, vie.FP64, False)
    assert not ok and "not current" in vie.last_error()
    s.apply_ex(x, vie.FP64)           # refreshes them
    ok, _ = noalloc(s, x, vie.FP64, False)
    assert ok, vie.last_error()
    s.close()
    big = system((64, 64, 48), conductors=False)
    assert vie.lib().dielectric_fft_apply_noalloc_bytes(big.s, int(vie.FP64)) > 96 << 20
    ok, _ = noalloc(big, vec(big.order, 8), vie.FP64, False)
    assert not ok and "maximum" in vie.last_error()
    big.close()
