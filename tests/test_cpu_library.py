"""CPU-side checks of libvie_hip.so: it loads, exports every symbol the
public headers declare, keeps the reference's names/constants, rejects bad
arguments before touching a device, and leaves callback operators that are
not its own to the host program's vie_gmres.  No device compute happens here."""
import ctypes
import glob
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from nec2c_amd import vie

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M):
            line = text[text.rfind("\n", 0, m.start()) + 1:m.end()]
            if line.lstrip().startswith(("typedef", "#", "return", "if")):
                continue
            names.add(m.group(1))
        for m in re.finditer(r"extern\s+\w+\s+(\w+)\s*;", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = vie.lib()
    names = declared_symbols()
    assert "vie_gmres" in names and "dielectric_fft_prepare" in names and "dielectric_gmres_settings" in names
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert set(vie.EXPORTS) <= names | {"dielectric_gmres_settings"}


def test_reference_constants_and_names():
    L = vie.lib()
    o = vie.gmres_settings()
    assert (o.restart, o.maximum_iterations, o.relative_tolerance, o.absolute_tolerance) == (30, 300, 1e-12, 0.0)
    d = vie.GmresOptions()
    L.vie_gmres_default_options(ctypes.byref(d))
    assert (d.restart, d.maximum_iterations, d.relative_tolerance, d.absolute_tolerance) == (30, 300, 1e-12, 0.0)
    assert [L.vie_iterative_status_name(i).decode() for i in range(7)] == list(vie.STATUS)
    assert L.vie_iterative_status_name(99) == b"unknown status"
    assert L.dielectric_fft_backend(None) == b"none"
    assert L.vie_fft_fftw_available() == 0


def test_backend_selection():
    assert not vie.backend_set("fftw") and not vie.backend_set("builtin") and not vie.backend_set("invalid")
    assert vie.backend_set("hip")
    assert vie.lib().vie_fft_backend_requested_name() == b"hip"
    assert vie.backend_set("auto") and vie.backend_set(None)
    assert vie.lib().vie_fft_backend_requested_name() == b"auto"
    vie.lib().vie_fft_backend_clear_last_used()
    assert vie.lib().vie_fft_backend_last_used_name() is None


@pytest.mark.parametrize("args", [
    ((5, 3, 2), [0, 0, 0, 4, 2, 1, 1, 0, 0, 0, 1, 0, 0, 0, 1, 1, 1, 1], .25, .2, 3, 1),     # budget
    ((5, 3, 2), [0, 0, 0, 0, 0, 0], .25, .2, 3, 10 ** 9),                                   # duplicate
    ((1, 1, 1), [0, 0, 0, 4, 2, 1], .25, .2, 3, 10 ** 9),                                   # outside
    ((1, 1, 1), [0, 0, 0], .25, 5.0, 3, 10 ** 9),                                           # k h > 1
    ((1, 1, 1), [0, 0, 0], .25, .2, 3 + .4j, 10 ** 9),                                      # active
    ((1, 1, 1), [0, 0, 0], .25, .2, 0.5, 10 ** 9),                                          # Re < 1
    ((1, 1, 1), [0, 0, 0], -.25, .2, 3, 10 ** 9),                                           # pitch
    ((0, 2, 2), [0, 0, 0], .25, .2, 3, 10 ** 9),                                            # empty axis
])
def test_create_rejections_before_any_device_use(args):
    # tests/vie_fft_test.c:50-54,69 -- all of these fail on argument checks
    shape, xyz, h, k, er, limit = args
    with pytest.raises(ValueError):
        vie.FFTPlan(shape, xyz, h, k, er, byte_limit=limit)


def test_create_overflow_shape_rejected():
    L = vie.lib()
    p = ctypes.c_void_p(1234)
    shape = (ctypes.c_size_t * 3)(2 ** 64 - 1, 2, 2)
    xyz = (ctypes.c_size_t * 3)(0, 0, 0)
    assert not L.vie_fft_create(ctypes.byref(p), shape, 1, xyz, .25, .2, vie.CDouble.of(3), 2 ** 64 - 1)
    assert not p.value


def test_prepare_rejections_before_any_device_use():
    M = [vie.MaterialSpec(1, 3.0), vie.MaterialSpec(2, 2.0)]
    C = vie.Cell
    cases = [
        ([C((0, 0, 0), (.25,) * 3, 1), C((.3, 0, 0), (.25,) * 3, 1)], "common lattice"),
        ([C((0, 0, 0), (.25,) * 3, 1), C((.25, 0, 0), (.25, .25, .125), 1)], "equal cubic"),
        ([C((0, 0, 0), (.25,) * 3, 9)], "passive material"),
        ([C((0, 0, 0), (.25,) * 3, 1), C((float("nan"), 0, 0), (.25,) * 3, 1)], "equal cubic"),
    ]
    for cells, msg in cases:
        with pytest.raises(ValueError, match=msg):
            vie.System(cells, M, 0, .5)
    # PEC: cells below z = 0, and a gap off the half-cell lattice (dielectric_fft_test.c:127-129)
    with pytest.raises(ValueError, match="z >= 0"):
        vie.System([C((0, 0, 0), (.25,) * 3, 1)], M, 0, .5, pec=True)
    with pytest.raises(ValueError, match="half-cell lattice"):
        vie.System([C((0, 0, .155), (.25,) * 3, 1), C((.25, -.25, .405), (.25,) * 3, 2)], M, 0, .5, pec=True)
    # budget (dielectric_fft_test.c:159)
    L = vie.lib()
    old = L.dielectric_fft_budget()
    try:
        L.dielectric_fft_set_budget(512 * 1024 * 1024)
        with pytest.raises(ValueError, match="budget"):
            vie.System([C((0, 0, 0), (.25,) * 3, 1)], M, 2, .5, extra_bytes=512 * 1024 * 1024)
    finally:
        L.dielectric_fft_set_budget(old)
    # every diagnostic carries the reference's "FFT: " prefix (dielectric_fft.c:21-26)
    try:
        vie.System([C((0, 0, 0), (.25,) * 3, 9)], M, 0, .5)
    except ValueError as e:
        assert str(e).startswith("FFT: ")


# ---- vie_gmres argument checks and foreign operators ----
def test_gmres_refuses_foreign_operators_and_bad_arguments():
    """vie_gmres solves only this library's operators (on the device); any other
    callback operator is refused with VIE_ITERATIVE_INVALID_INPUT before a
    callback runs, so the host program's own vie_gmres serves it.  Bad
    options and aliasing are refused first, as the reference refuses them
    (vie_iterative.c:96-108)."""
    n = 3
    calls = []
    ident = vie.CallbackOperator(n, lambda v: (calls.append(1), v)[1])
    e0 = np.array([1, 0, 0], complex)
    ok, x, rep = vie.gmres(ident.op, e0)
    assert not ok and rep.status == 4 and not calls
    assert "own vie_gmres" in vie.last_error()
    assert np.all(x == 0)
    ok, x, rep = vie.gmres(ident.op, e0, restart=0)
    assert not ok and rep.status == 4
    ok, x, rep = vie.gmres(ident.op, e0, relative_tolerance=0.0, absolute_tolerance=0.0)
    assert not ok and rep.status == 4
    L = vie.lib()
    rep = vie.GmresReport()
    bb = e0.copy()
    assert not L.vie_gmres(ctypes.byref(ident.op), bb.ctypes.data, bb.ctypes.data, None, ctypes.byref(rep))
    assert rep.status == 4
    with pytest.raises(AttributeError):
        L.vie_dense_gmres   # the dense adapter stays nec2c's (vie_iterative.c:201-244)


def test_struct_layouts_match_c():
    assert ctypes.sizeof(vie.MeshCell) == 80
    assert ctypes.sizeof(vie.MeshResult) == 224
    assert ctypes.sizeof(vie.Diagnostic) == 1448
    assert ctypes.sizeof(vie.Model) == 112
    assert ctypes.sizeof(vie.Material) == 64
    assert ctypes.sizeof(vie.GmresReport) == 48
    assert ctypes.sizeof(vie.SolveOptions) == 24
    assert ctypes.sizeof(vie.Clearance) == 72


def _arnoldi_hessenberg(A, b, m):
    n = len(b)
    V = np.zeros((n, m + 1), complex)
    H = np.zeros((m + 1, m), complex)
    V[:, 0] = b / np.linalg.norm(b)
    for j in range(m):
        w = A @ V[:, j]
        for _ in range(2):
            h = V[:, :j + 1].conj().T @ w
            w = w - V[:, :j + 1] @ h
            H[:j + 1, j] += h
        H[j + 1, j] = np.linalg.norm(w)
        V[:, j + 1] = w / H[j + 1, j]
    return H


@pytest.mark.parametrize("m,k", [(12, 4), (60, 10), (120, 20)])
def test_harmonic_ritz_vectors(m, k):
    """Host step of the recycled subspace against its definition (Morgan 2002):
    each column is a unit eigenvector of H_m + |h_{m+1,m}|^2 H_m^{-H} e_m e_m^T,
    and the eigenvalues are its k of smallest modulus (numpy's eig as the check).
    The Hessenberg comes from an Arnoldi run on a shifted random operator with a
    few small outlying eigenvalues, the shape the solver's cycles produce."""
    rng = np.random.default_rng(m)
    n = 400
    A = (2 + 0.5j) * np.eye(n) + 0.3 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    A[:6, :6] += np.diag([-1.9, -1.8, -1.95, -1.7, -1.85, -1.6])
    H = _arnoldi_hessenberg(A, rng.standard_normal(n) + 0j, m)
    P = vie.harmonic_ritz(H, k)
    Hm = H[:m]
    em = np.zeros(m)
    em[-1] = 1
    f = np.linalg.solve(Hm.conj().T, em)
    M = Hm + abs(H[m, m - 1]) ** 2 * np.outer(f, em)
    ref = np.sort(np.abs(np.linalg.eigvals(M)))[:k]
    got = []
    for c in range(k):
        v = P[:, c]
        assert abs(np.linalg.norm(v) - 1) < 1e-12
        lam = v.conj() @ M @ v
        assert np.linalg.norm(M @ v - lam * v) <= 1e-10 * np.linalg.norm(M)
        got.append(abs(lam))
    assert np.allclose(np.sort(got), ref, rtol=1e-9)
    # a wrong selection (largest modulus) or a dropped rank-one term fails:
    assert not np.allclose(np.sort(got), np.sort(np.abs(np.linalg.eigvals(Hm)))[:k], rtol=1e-6)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_far_field_rejections_before_any_device_use():
    L = vie.lib()
    d = np.array([0.0, 0.0, 1.0])
    out = np.zeros(3, complex)
    c = np.zeros(3)
    J = np.zeros(3, complex)
    # unknown ground, missing Fresnel coefficients, bad k0, null cell arrays
    assert not L.vie_hip_radiation_amplitudes(1, c.ctypes.data, c.ctypes.data, J.ctypes.data, 1.0, 5, None, None, 1,
                                              d.ctypes.data, out.ctypes.data)
    assert not L.vie_hip_radiation_amplitudes(1, c.ctypes.data, c.ctypes.data, J.ctypes.data, 1.0, 2, None, None, 1,
                                              d.ctypes.data, out.ctypes.data)
    assert not L.vie_hip_radiation_amplitudes(1, c.ctypes.data, c.ctypes.data, J.ctypes.data, float("nan"), 0, None,
                                              None, 1, d.ctypes.data, out.ctypes.data)
    assert not L.vie_hip_radiation_amplitudes(1, None, None, None, 1.0, 0, None, None, 1, d.ctypes.data,
                                              out.ctypes.data)
    p = ctypes.c_double()
    assert not L.vie_hip_scattered_power(1, c.ctypes.data, c.ctypes.data, J.ctypes.data, 1.0, 2, ctypes.byref(p))
    # no directions: nothing to do, no device needed
    assert L.vie_hip_radiation_amplitudes(1, c.ctypes.data, c.ctypes.data, J.ctypes.data, 1.0, 0, None, None, 0,
                                          None, None)


def test_conductor_row_rejections_before_any_device_use():
    L = vie.lib()
    c = np.zeros(3)
    out = np.zeros(3, complex)
    ms = ctypes.c_float()
    # bad k0, negative count, null cell or wire arrays, ld below 3 cells, a finite ground
    assert not L.vie_hip_volume_conductor_rows(float("nan"), 1, c.ctypes.data, c.ctypes.data, 1, c.ctypes.data,
                                               c.ctypes.data, 0, None, None, None, None, 0, out.ctypes.data, 3)
    assert not L.vie_hip_volume_conductor_rows(1.0, 1, c.ctypes.data, c.ctypes.data, -1, c.ctypes.data,
                                               c.ctypes.data, 0, None, None, None, None, 0, out.ctypes.data, 3)
    assert not L.vie_hip_volume_conductor_rows(1.0, 1, None, None, 1, c.ctypes.data, c.ctypes.data, 0, None, None,
                                               None, None, 0, out.ctypes.data, 3)
    assert not L.vie_hip_volume_conductor_rows(1.0, 1, c.ctypes.data, c.ctypes.data, 1, None, None, 0, None, None,
                                               None, None, 0, out.ctypes.data, 3)
    assert not L.vie_hip_volume_conductor_rows(1.0, 1, c.ctypes.data, c.ctypes.data, 0, None, None, 1,
                                               c.ctypes.data, c.ctypes.data, None, c.ctypes.data, 0,
                                               out.ctypes.data, 3)
    assert not L.vie_hip_volume_conductor_rows(1.0, 1, c.ctypes.data, c.ctypes.data, 1, c.ctypes.data,
                                               c.ctypes.data, 0, None, None, None, None, 0, out.ctypes.data, 2)
    assert not L.vie_hip_volume_conductor_rows(1.0, 1, c.ctypes.data, c.ctypes.data, 1, c.ctypes.data,
                                               c.ctypes.data, 0, None, None, None, None, 2, out.ctypes.data, 3)
    assert "finite ground" in vie.last_error()
    assert not L.vie_hip_volume_conductor_rows(1.0, 1, c.ctypes.data, c.ctypes.data, 1, c.ctypes.data,
                                               c.ctypes.data, 0, None, None, None, None, 0, None, 3)
    assert not L.vie_hip_time_volume_conductor_rows(1.0, 1, c.ctypes.data, c.ctypes.data, 1, c.ctypes.data,
                                                    c.ctypes.data, 0, None, None, None, None, 0, 0,
                                                    ctypes.byref(ms))
    # nothing to assemble: success without a device
    assert L.vie_hip_volume_conductor_rows(1.0, 0, None, None, 1, c.ctypes.data, c.ctypes.data, 0, None, None,
                                           None, None, 0, None, 0)
    assert L.vie_hip_volume_conductor_rows(1.0, 1, c.ctypes.data, c.ctypes.data, 0, None, None, 0, None, None, None,
                                           None, 0, None, 0)


def test_conductor_volume_rejections_before_any_device_use():
    from nec2c_amd import vie as Vh
    c = np.zeros(3)
    sz = np.full(3, 0.01)
    chi = np.ones(1, complex)
    seg = [np.array([v]) for v in (0.0, 0.0, 0.0, 0.05, 0.001, 0.0, 0.0, 1.0)]
    src = np.zeros((1, 3), complex)
    src[0, 0] = 1.0
    call = lambda **kw: Vh.conductor_volume_columns(kw.get("k0", 1.0), kw.get("wlam", 1.0), c, kw.get("sizes", sz),
                                                    chi, kw.get("seg", seg), src, rkh=kw.get("rkh", 1.0),
                                                    ground=kw.get("ground", 0))
    for kw, msg in ((dict(k0=float("nan")), "invalid"), (dict(wlam=0.0), "invalid"), (dict(rkh=float("nan")), "invalid"),
                    (dict(ground=3), "invalid"), (dict(ground=2), "finite ground"),
                    (dict(sizes=np.array([0.01, np.inf, 0.01])), "non-finite"),
                    (dict(seg=seg[:3] + [np.array([-0.05])] + seg[4:]), "segment length")):
        with pytest.raises(RuntimeError, match=msg):
            call(**kw)
    # no columns: nothing to do, no device needed
    assert Vh.conductor_volume_columns(1.0, 1.0, c, sz, chi, seg, np.zeros((0, 3), complex)).shape == (0, 3)


def test_release_workspaces_without_device_work():
    assert vie.lib().vie_hip_release_workspaces() == 1


def test_mesh_check_rejections_before_any_device_use():
    a, b = ctypes.c_longlong(7), ctypes.c_longlong(7)
    assert vie.lib().vie_hip_box_overlap(None, None, 5, ctypes.byref(a), ctypes.byref(b)) == 0
    assert "invalid" in vie.last_error()
    lo, hi = np.zeros((2, 3)), np.ones((2, 3))
    out = vie.Clearance()
    assert vie.lib().vie_hip_conductor_clearance(lo.ctypes.data, hi.ctypes.data, None, 2, 0, *([None] * 8), 0,
                                                 None, None, ctypes.byref(out)) == 0
    with pytest.raises(ValueError):
        vie.box_overlap(lo, hi[:1])
    with pytest.raises(ValueError):
        vie.conductor_clearance(lo, hi, [0.1])
    # no conductors: nothing is evaluated and no device is needed (mesh->has_conductors = 0)
    r = vie.conductor_clearance(lo, hi, [0.1, 0.1])
    assert r["has_conductors"] == 0 and r["insufficient"] == 0 and r["worst_box"] == 0
