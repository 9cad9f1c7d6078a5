"""CPU oracle for the dielectric VIE FFT-GMRES path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module, and only as the checker.  The product library never uses it.

Contents, each restating the reference (file:line):

* kernel quadrature -> oracle/vie_oracle.c (C, via ctypes)
* ``offset_table``        unique sorted offsets + usefulness test   vie_fft.c:338-353
* ``KernelSpectrum``      sign/permutation spread, padding, FFT     vie_fft.c:363-376, 386-391
* ``VolumeOperator``      scatter / 3x3 spectral multiply / gather  vie_fft.c:239-273, 445-455
* ``HybridSystem``        [C B; D V] blocks, PEC image embedding    dielectric_fft.c:34-175
* ``gmres``               right-preconditioned restarted GMRES(m)   vie_iterative.c:9-199
* ``ReferenceLibrary``    ctypes over oracle/_ref/libvieref.so, the reference's
                          own C sources compiled by oracle/Makefile

The FFT here is numpy's (a textbook DFT); the reference's radix-2 DFT
(vie_fft.c:94-115) computes the same transform, so the two differ only by
rounding (checked against oracle/_ref in tests/test_oracle_operator.py).

Parity status: every function above is pinned by a test under tests/ (golden
tensors from the reference's tests/vie-reference.csv, the static cube
diag(2/3), dense brute force on tiny lattices, the reference's own known-
solution GMRES mocks and, where it builds, oracle/_ref entry for entry).
"""
from __future__ import annotations

import ctypes
import itertools
import math
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

PERMUTATIONS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))
# tensor component order xx, yy, zz, xy, xz, yz (vie_fft.c:53-54)
TENSOR_PAIRS = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))


def tensor_index(i: int, j: int) -> int:
    """Packed index of symmetric component (i, j) (vie_fft.c:53-54)."""
    if i == j:
        return i
    i, j = min(i, j), max(i, j)
    return (3 if j == 1 else 4) if i == 0 else 5


def lib() -> ctypes.CDLL:
    """Load oracle/_build/liboracle.so (built by oracle/Makefile)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "_build", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: run make -C {HERE}")
        L = ctypes.CDLL(path)
        d3 = ctypes.POINTER(ctypes.c_double)
        L.oracle_cuboid_n.argtypes = [d3, d3, d3, d3, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int, ctypes.c_int, d3,
                                      ctypes.POINTER(ctypes.c_int)]
        L.oracle_cuboid_n.restype = ctypes.c_int
        L.oracle_relation.argtypes = [d3, d3, d3, d3]
        L.oracle_relation.restype = ctypes.c_int
        L.oracle_far_kernel.argtypes = [ctypes.POINTER(ctypes.c_long), ctypes.c_double,
                                        ctypes.c_double, d3]
        L.oracle_far_kernel.restype = ctypes.c_int
        L.oracle_offset_table.argtypes = [ctypes.POINTER(ctypes.c_long), ctypes.c_long,
                                          ctypes.c_double, ctypes.c_double, d3]
        L.oracle_offset_table.restype = ctypes.c_long
        _LIB = L
    return _LIB


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def cuboid_n(tc, tsz, sc, ssz, k, rtol=2e-8, atol=2e-11, max_order=64, far_order=8):
    """Galerkin tensor between two cuboids (vie_kernel.c:550-641); None on failure."""
    args = [np.ascontiguousarray(v, dtype=np.float64) for v in (tc, tsz, sc, ssz)]
    out = np.zeros(18)
    order = ctypes.c_int(0)
    ok = lib().oracle_cuboid_n(*[_dp(a) for a in args], float(k), rtol, atol, max_order,
                               far_order, _dp(out), ctypes.byref(order))
    if not ok:
        return None
    return out.view(np.complex128).reshape(3, 3).copy()


def relation(tc, tsz, sc, ssz) -> int:
    """0 self, 1 face, 2 edge, 3 corner, 4 near, 5 far (vie_kernel.c:85-123)."""
    args = [np.ascontiguousarray(v, dtype=np.float64) for v in (tc, tsz, sc, ssz)]
    return lib().oracle_relation(*[_dp(a) for a in args])


def far_kernel(v, h, k):
    """Far-pair ladder (vie_fft.c:319-335); None when not far / not settled."""
    vv = np.ascontiguousarray(v, dtype=np.int64)
    out = np.zeros(18)
    ok = lib().oracle_far_kernel(vv.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), h, k, _dp(out))
    return out.view(np.complex128).reshape(3, 3).copy() if ok else None


def padded_extent(n: int) -> int:
    """Smallest power of two >= 2n-1 (vie_fft.c:386-391)."""
    p = 1
    while p < 2 * n - 1:
        p *= 2
    return p


def useful_offsets(shape) -> np.ndarray:
    """Sorted |offset| triples a>=b>=c used by some permutation inside shape
    (vie_fft.c:347-353), in the reference's loop order."""
    mx = max(shape)
    out = []
    for a in range(mx):
        for b in range(a + 1):
            for c in range(b + 1):
                v = (a, b, c)
                if any(v[q[0]] < shape[0] and v[q[1]] < shape[1] and v[q[2]] < shape[2]
                       for q in PERMUTATIONS):
                    out.append(v)
    return np.array(out, dtype=np.int64).reshape(-1, 3)


def offset_table(shape, h, k):
    """(offsets, tensors[count,3,3]) with the reference's far/near choice and
    off-diagonal symmetrisation (vie_fft.c:354-359)."""
    offs = useful_offsets(shape)
    out = np.zeros((len(offs), 18))
    n = lib().oracle_offset_table(offs.ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
                                  len(offs), h, k, _dp(out))
    if n != len(offs):
        raise RuntimeError(f"kernel quadrature failed at offset {offs[n]}")
    return offs, out.view(np.complex128).reshape(-1, 3, 3)


class KernelSpectrum:
    """Padded 6-component Green tensor and its FFT (vie_fft.c:338-377)."""

    def __init__(self, shape, h, k):
        self.shape = tuple(int(s) for s in shape)
        self.padded = tuple(padded_extent(s) for s in self.shape)
        self.h, self.k = float(h), float(k)
        offs, table = offset_table(self.shape, h, k)
        P = self.padded
        T = np.zeros((6,) + P, dtype=np.complex128)
        self.self_term = np.zeros(3, dtype=np.complex128)
        for v, n in zip(offs, table):
            if v[0] == 0:
                self.self_term = np.array([n[0, 0], n[1, 1], n[2, 2]])
            for q in PERMUTATIONS:
                d = (v[q[0]], v[q[1]], v[q[2]])
                if d[0] >= self.shape[0] or d[1] >= self.shape[1] or d[2] >= self.shape[2]:
                    continue
                for mask in range(8):
                    sign = [(-1 if (mask >> i) & 1 else 1) for i in range(3)]
                    pos = tuple((P[i] - d[i]) if (sign[i] < 0 and d[i]) else d[i] for i in range(3))
                    for c, (i, j) in enumerate(TENSOR_PAIRS):
                        T[(c,) + pos] = sign[i] * sign[j] * n[q[i], q[j]]
        self.spatial = T
        self.spectrum = np.fft.fftn(T, axes=(1, 2, 3))


class VolumeOperator:
    """FFT volume operator y = eps x - (eps-1) IFFT(T^ FFT(x)) (vie_fft.c:239-273)."""

    def __init__(self, shape, xyz, h, k, eps, spectrum: KernelSpectrum | None = None):
        self.shape = tuple(int(s) for s in shape)
        self.xyz = np.asarray(xyz, dtype=np.int64).reshape(-1, 3)
        self.cells = len(self.xyz)
        if np.any(self.xyz < 0) or np.any(self.xyz >= np.array(self.shape)):
            raise ValueError("cell outside lattice")
        self.spec = spectrum if spectrum is not None else KernelSpectrum(self.shape, h, k)
        P = self.spec.padded
        flat = (self.xyz[:, 0] * P[1] + self.xyz[:, 1]) * P[2] + self.xyz[:, 2]
        if len(np.unique(flat)) != len(flat):
            raise ValueError("duplicate lattice site")
        self.flat = flat
        eps = np.broadcast_to(np.asarray(eps, dtype=np.complex128), (self.cells,)).copy()
        self.eps = eps

    @property
    def order(self):
        return 3 * self.cells

    def convolve(self, x):
        """dst[i] = sum_j IFFT(T^_ij FFT(src_j)) at occupied sites, [cells,3]."""
        P = self.spec.padded
        G = P[0] * P[1] * P[2]
        src = np.zeros((3, G), dtype=np.complex128)
        xv = np.asarray(x, dtype=np.complex128).reshape(-1, 3)
        for i in range(3):
            src[i, self.flat] = xv[:, i]
        S = np.fft.fftn(src.reshape((3,) + P), axes=(1, 2, 3))
        Th = self.spec.spectrum
        D = np.empty_like(S)
        for i in range(3):
            D[i] = sum(Th[tensor_index(i, j)] * S[j] for j in range(3))
        dst = np.fft.ifftn(D, axes=(1, 2, 3)).reshape(3, G)
        return dst[:, self.flat].T

    def apply(self, x):
        xv = np.asarray(x, dtype=np.complex128).reshape(-1, 3)
        dst = self.convolve(xv)
        e = self.eps[:, None]
        return (e * xv - (e - 1) * dst).reshape(-1)

    def precondition(self, x):
        xv = np.asarray(x, dtype=np.complex128).reshape(-1, 3)
        e = self.eps[:, None]
        return (xv / (e - (e - 1) * self.spec.self_term[None, :])).reshape(-1)


@dataclass
class Cell:
    center: tuple
    size: tuple
    material_id: int


@dataclass
class Material:
    id: int
    epsilon_r: float
    epsilon_loss: float = 0.0
    conductivity: float = 0.0

    def epsilon(self, k0: float) -> complex:
        """Untabulated branch of dielectric_material_epsilon (dielectric.h:36-56)."""
        s = self.conductivity
        extra = 0.0 if s == 0 else s / (k0 * 299792458.0 * 8.8541878128e-12)
        return complex(self.epsilon_r, -(self.epsilon_loss + extra))


class PrepareError(ValueError):
    pass


class HybridSystem:
    """Hybrid [C B; D V] operator with optional PEC image (dielectric_fft.c:9-175)."""

    def __init__(self, cells: list, materials: list, conductors: int, k0: float, pec=False):
        self.c = int(conductors)
        N = len(cells)
        self.v = 3 * N
        self.pec = bool(pec)
        self.blocks_cc = np.zeros((self.c, self.c), dtype=np.complex128)
        self.blocks_cv = np.zeros((self.c, self.v), dtype=np.complex128)   # B
        self.blocks_vc = np.zeros((self.v, self.c), dtype=np.complex128)   # D
        self.volume = None
        if N == 0:
            return
        h = cells[0].size[0]
        origin = [min(cl.center[d] for cl in cells) for d in range(3)]
        mats = {m.id: m for m in materials}
        eps = []
        for cl in cells:
            if cl.material_id not in mats:
                raise PrepareError("FFT: unknown or invalid passive material")
            e = mats[cl.material_id].epsilon(k0)
            if not (np.isfinite(e) and e.real >= 1 and e.imag <= 0):
                raise PrepareError("FFT: unknown or invalid passive material")
            if pec and cl.center[2] - 0.5 * cl.size[2] < 0:
                raise PrepareError("FFT: PEC FFT cells must lie in z >= 0")
            for d in range(3):
                if abs(cl.size[d] - h) > 1e-10 * h:
                    raise PrepareError("FFT: requires equal cubic cells")
            eps.append(e)
        xyz = np.zeros((N, 3), dtype=np.int64)
        for i, cl in enumerate(cells):
            for d in range(3):
                idx = (cl.center[d] - origin[d]) / h
                near = round(idx)
                if not math.isfinite(idx) or near < 0 or abs(idx - near) > 1e-8:
                    raise PrepareError("FFT: cell centers do not share a common lattice")
                xyz[i, d] = near
        shape = [int(xyz[:, d].max()) + 1 for d in range(3)]
        eps = np.array(eps)
        if pec:
            gap = 2 * (origin[2] - 0.5 * h) / h
            near = round(gap)
            if abs(gap - near) > 1e-8 or near < 0:
                raise PrepareError("FFT: PEC FFT requires nonnegative bottom gap on the half-cell lattice")
            nz = shape[2]
            shape[2] = 2 * nz + near
            image = xyz.copy()
            image[:, 2] = nz - 1 - xyz[:, 2]
            xyz[:, 2] += nz + near
            xyz = np.concatenate([xyz, image])
            eps = np.concatenate([eps, eps])
        self.shape = tuple(shape)
        self.volume = VolumeOperator(shape, xyz, h, k0, eps)

    @property
    def order(self):
        return self.c + self.v

    def _volume_apply(self, xv):
        if self.pec:
            sign = np.tile(np.array([-1, -1, 1]), self.v // 3)
            full = np.concatenate([xv, sign * xv])
            return self.volume.apply(full)[: self.v]
        return self.volume.apply(xv)

    def apply(self, x):
        x = np.asarray(x, dtype=np.complex128)
        y = np.zeros(self.order, dtype=np.complex128)
        xc, xv = x[: self.c], x[self.c:]
        if self.v:
            y[self.c:] = self._volume_apply(xv)
        y[: self.c] = self.blocks_cc @ xc + self.blocks_cv @ xv
        y[self.c:] += self.blocks_vc @ xc
        return y

    def precondition(self, x):
        x = np.asarray(x, dtype=np.complex128)
        y = np.zeros(self.order, dtype=np.complex128)
        if self.v:
            if self.pec:
                full = np.concatenate([x[self.c:], np.zeros(self.v)])
                y[self.c:] = self.volume.precondition(full)[: self.v]
            else:
                y[self.c:] = self.volume.precondition(x[self.c:])
        d = np.diag(self.blocks_cc)
        if np.any(~np.isfinite(np.abs(d))) or np.any(np.abs(d) < np.finfo(float).tiny):
            raise ZeroDivisionError("conductor diagonal")
        y[: self.c] = x[: self.c] / d
        return y


# -------------------------------------------------------------------------
# GMRES restatement (vie_iterative.c)
# -------------------------------------------------------------------------
STATUS = ("converged", "iteration limit", "Krylov breakdown",
          "operator/preconditioner failure", "invalid input",
          "workspace allocation/size failure", "nonfinite arithmetic")
CONVERGED, ITERATION_LIMIT, BREAKDOWN, CALLBACK_FAILED, INVALID_INPUT, ALLOCATION_FAILED, NONFINITE = range(7)
DBL_EPS = np.finfo(np.float64).eps


@dataclass
class Report:
    status: int = INVALID_INPUT
    iterations: int = 0
    operator_applications: int = 0
    restarts: int = 0
    residual_norm: float = math.inf
    relative_residual: float = math.inf


def _norm(x):
    """hypot-accumulated norm with a NaN on any nonfinite entry (vie_iterative.c:9-19)."""
    if not np.all(np.isfinite(x)):
        return math.nan
    return float(np.sqrt(np.sum(x.real ** 2 + x.imag ** 2)))


def _rotation(a, b):
    """Complex Givens [c s; -conj(s) c] (vie_iterative.c:66-78)."""
    aa, bb = abs(a), abs(b)
    if bb == 0.0:
        return 1.0, 0.0
    if aa == 0.0:
        return 0.0, np.conj(b) / bb
    ln = math.hypot(aa, bb)
    return aa / ln, (a / aa) * (np.conj(b) / ln)


def gmres(apply, precondition, b, x, restart=30, maximum_iterations=300,
          relative_tolerance=1e-12, absolute_tolerance=0.0):
    """Restatement of vie_gmres (vie_iterative.c:80-199).  ``apply`` and
    ``precondition`` return arrays or None (callback failure).  x is updated
    in place.  Returns (success, Report)."""
    rep = Report()
    b = np.asarray(b, dtype=np.complex128)
    n = len(b)
    if (apply is None or n == 0 or x is b or restart == 0 or maximum_iterations == 0
            or not math.isfinite(relative_tolerance) or relative_tolerance < 0
            or not math.isfinite(absolute_tolerance) or absolute_tolerance < 0
            or (relative_tolerance == 0 and absolute_tolerance == 0)):
        return False, rep
    m = min(restart, n, maximum_iterations)
    bnorm = _norm(b)
    target = max(absolute_tolerance, relative_tolerance * bnorm)
    if not (math.isfinite(bnorm) and math.isfinite(_norm(x)) and math.isfinite(target)):
        rep.status = NONFINITE
        return False, rep

    def residual():
        rep.operator_applications += 1
        ax = apply(x)
        if ax is None:
            rep.status = CALLBACK_FAILED
            return None
        r = b - ax
        rep.residual_norm = _norm(r)
        if not math.isfinite(rep.residual_norm):
            rep.status = NONFINITE
            return None
        return r

    r = residual()
    if r is None:
        return False, rep
    while True:
        beta = rep.residual_norm
        rep.relative_residual = beta / bnorm if bnorm > 0 else beta
        if beta <= target:
            rep.status = CONVERGED
            return True, rep
        if rep.iterations >= maximum_iterations:
            rep.status = ITERATION_LIMIT
            return False, rep
        H = np.zeros((m + 1, m), dtype=np.complex128)
        g = np.zeros(m + 1, dtype=np.complex128)
        g[0] = beta
        V = np.zeros((m + 1, n), dtype=np.complex128)
        Z = np.zeros((m, n), dtype=np.complex128)
        cs = np.zeros(m)
        sn = np.zeros(m, dtype=np.complex128)
        V[0] = r / beta
        used = 0
        breakdown = False
        for j in range(m):
            if rep.iterations >= maximum_iterations:
                break
            if precondition is not None:
                z = precondition(V[j])
                if z is None:
                    rep.status = CALLBACK_FAILED
                    return False, rep
                Z[j] = z
            else:
                Z[j] = V[j]
            if not math.isfinite(_norm(Z[j])):
                rep.status = NONFINITE
                return False, rep
            rep.operator_applications += 1
            w = apply(Z[j])
            if w is None:
                rep.status = CALLBACK_FAILED
                return False, rep
            w = np.array(w, dtype=np.complex128)
            before = _norm(w)
            if not math.isfinite(before):
                rep.status = NONFINITE
                return False, rep
            for _ in range(2):
                for k in range(j + 1):
                    p = np.vdot(V[k], w)
                    H[k, j] += p
                    w -= p * V[k]
            nxt = _norm(w)
            if not math.isfinite(nxt):
                rep.status = NONFINITE
                return False, rep
            H[j + 1, j] = nxt
            breakdown = nxt <= 32 * DBL_EPS * before
            if not breakdown:
                V[j + 1] = w / nxt
            for k in range(j):
                f, s = H[k, j], H[k + 1, j]
                H[k, j] = cs[k] * f + sn[k] * s
                H[k + 1, j] = -np.conj(sn[k]) * f + cs[k] * s
            cs[j], sn[j] = _rotation(H[j, j], H[j + 1, j])
            H[j, j] = cs[j] * H[j, j] + sn[j] * H[j + 1, j]
            H[j + 1, j] = 0
            g[j + 1] = -np.conj(sn[j]) * g[j]
            g[j] *= cs[j]
            rep.iterations += 1
            used = j + 1
            if breakdown or abs(g[j + 1]) <= target:
                break
        y = np.zeros(used, dtype=np.complex128)
        for i in range(used - 1, -1, -1):
            val = g[i] - np.dot(H[i, i + 1:used], y[i + 1:used])
            if abs(H[i, i]) == 0.0:
                rep.status = BREAKDOWN
                return False, rep
            y[i] = val / H[i, i]
        x += y @ Z[:used]
        r = residual()
        if r is None:
            return False, rep
        rep.relative_residual = rep.residual_norm / bnorm if bnorm > 0 else rep.residual_norm
        if breakdown and rep.residual_norm > target:
            rep.status = BREAKDOWN
            return False, rep
        if rep.residual_norm > target and rep.iterations < maximum_iterations:
            rep.restarts += 1


# -------------------------------------------------------------------------
# The reference's own sources, compiled by oracle/Makefile (oracle/_ref)
# -------------------------------------------------------------------------
class ReferenceLibrary:
    """ctypes over oracle/_ref/libvieref.so (reference sources + ref_shim.c)."""

    def __init__(self, path: str | None = None):
        path = path or os.path.join(HERE, "_ref", "libvieref.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = ctypes.CDLL(path)
        d = ctypes.POINTER(ctypes.c_double)
        sz = ctypes.POINTER(ctypes.c_size_t)
        vp = ctypes.c_void_p
        L.refshim_threads.argtypes = [ctypes.c_int]
        L.refshim_cuboid_n.argtypes = [d, d, d, d, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_int, d, ctypes.POINTER(ctypes.c_int)]
        L.refshim_cuboid_pec_n.argtypes = [d, d, d, d, ctypes.c_double, d]
        L.refshim_backend_set.argtypes = [ctypes.c_char_p]
        L.refshim_halfspace_cuboids.argtypes = [d, d, d, d, ctypes.c_double, ctypes.c_double, ctypes.c_double, d,
                                                ctypes.POINTER(ctypes.c_int)]
        L.refshim_halfspace_fresnel.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, d]
        L.refshim_create.argtypes = [sz, ctypes.c_size_t, sz, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_size_t]
        L.refshim_create.restype = vp
        L.refshim_set_eps.argtypes = [vp, ctypes.c_size_t, ctypes.c_double, ctypes.c_double]
        L.refshim_destroy.argtypes = [vp]
        L.refshim_bytes.argtypes = [vp]
        L.refshim_bytes.restype = ctypes.c_size_t
        L.refshim_apply.argtypes = [vp, d, d]
        L.refshim_precondition.argtypes = [vp, d, d]
        L.refshim_gmres.argtypes = [vp, d, d, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_double,
                                    ctypes.c_double, d]
        L.refshim_system_prepare.argtypes = [ctypes.c_size_t, d, d, ctypes.POINTER(ctypes.c_int),
                                             ctypes.c_size_t, ctypes.POINTER(ctypes.c_int), d,
                                             ctypes.c_size_t, ctypes.c_double, ctypes.c_size_t,
                                             ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
        L.refshim_system_prepare.restype = vp
        L.refshim_system_prepare_tab.argtypes = [ctypes.c_size_t, d, d, ctypes.POINTER(ctypes.c_int),
                                                 ctypes.c_size_t, ctypes.POINTER(ctypes.c_int), d, sz, d,
                                                 ctypes.c_size_t, ctypes.c_double, ctypes.c_size_t,
                                                 ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
        L.refshim_system_prepare_tab.restype = vp
        L.refshim_system_set_blocks.argtypes = [vp, ctypes.c_size_t, ctypes.c_size_t, d, d, d]
        L.refshim_system_free.argtypes = [vp]
        L.refshim_system_set.argtypes = [vp, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_double,
                                         ctypes.c_double]
        L.refshim_system_storage.argtypes = [vp]
        L.refshim_system_storage.restype = ctypes.c_size_t
        L.refshim_system_apply.argtypes = [vp, d, d]
        L.refshim_system_precondition.argtypes = [vp, d, d]
        L.refshim_system_gmres.argtypes = [vp, d, d, ctypes.c_size_t, ctypes.c_size_t,
                                           ctypes.c_double, ctypes.c_double, d]
        L.refshim_dense_gmres.argtypes = [d, ctypes.c_size_t, d, d, ctypes.c_size_t, ctypes.c_size_t,
                                          ctypes.c_double, ctypes.c_double, d]
        L.refshim_coupled_create.argtypes = [vp, ctypes.c_size_t, d, d, d]
        L.refshim_coupled_create.restype = vp
        L.refshim_coupled_free.argtypes = [vp]
        L.refshim_coupled_apply.argtypes = [vp, d, d]
        L.refshim_coupled_gmres.argtypes = [vp, d, d, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_double,
                                            ctypes.c_double, d]
        self.L = L

    def threads(self, n=0):
        return self.L.refshim_threads(int(n))

    def cuboid_n(self, tc, tsz, sc, ssz, k, rtol=2e-8, atol=2e-11, max_order=64, far_order=8):
        args = [np.ascontiguousarray(v, dtype=np.float64) for v in (tc, tsz, sc, ssz)]
        out = np.zeros(18)
        order = ctypes.c_int(0)
        ok = self.L.refshim_cuboid_n(*[_dp(a) for a in args], float(k), rtol, atol, max_order,
                                     far_order, _dp(out), ctypes.byref(order))
        return out.view(np.complex128).reshape(3, 3).copy() if ok else None

    def halfspace_cuboids(self, tc, tsz, sc, ssz, k, eps):
        """Reflected tensor of a lower half-space of permittivity eps (vie_halfspace_cuboids)."""
        args = [np.ascontiguousarray(v, dtype=np.float64) for v in (tc, tsz, sc, ssz)]
        out = np.zeros(18)
        order = ctypes.c_int(0)
        eps = complex(eps)
        ok = self.L.refshim_halfspace_cuboids(*[_dp(a) for a in args], float(k), eps.real, eps.imag, _dp(out),
                                              ctypes.byref(order))
        return (out.view(np.complex128).reshape(3, 3).copy(), order.value) if ok else (None, order.value)

    def halfspace_fresnel(self, eps, u):
        out = np.zeros(4)
        eps = complex(eps)
        ok = self.L.refshim_halfspace_fresnel(eps.real, eps.imag, float(u), _dp(out))
        return (complex(out[0], out[1]), complex(out[2], out[3])) if ok else None

    def cuboid_pec_n(self, tc, tsz, sc, ssz, k):
        args = [np.ascontiguousarray(v, dtype=np.float64) for v in (tc, tsz, sc, ssz)]
        out = np.zeros(18)
        ok = self.L.refshim_cuboid_pec_n(*[_dp(a) for a in args], float(k), _dp(out))
        return out.view(np.complex128).reshape(3, 3).copy() if ok else None

    def plan(self, shape, xyz, h, k, er, limit=1 << 62):
        return RefPlan(self, shape, xyz, h, k, er, limit)

    def system(self, cells, materials, conductors, k0, pec=False, extra=0):
        return RefSystem(self, cells, materials, conductors, k0, pec, extra)

    def dense_gmres(self, a, b, restart=30, maxit=300, rtol=1e-12, atol=0.0):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        b = np.ascontiguousarray(b, dtype=np.complex128)
        x = np.zeros_like(b)
        rep = np.zeros(6)
        ok = self.L.refshim_dense_gmres(_dp(a.view(np.float64)), len(b), _dp(b.view(np.float64)),
                                        _dp(x.view(np.float64)), restart, maxit, rtol, atol, _dp(rep))
        return bool(ok), x, _report(rep)


def _report(rep):
    return Report(int(rep[0]), int(rep[1]), int(rep[2]), int(rep[3]), float(rep[4]), float(rep[5]))


class RefCoupled:
    """[C B; D V] over a reference volume plan (ref_shim.c refshim_coupled_*)."""

    def __init__(self, plan: "RefPlan", C, B, D):
        self.plan = plan
        self.C = np.ascontiguousarray(C, dtype=np.complex128)
        self.B = np.ascontiguousarray(B, dtype=np.complex128)
        self.D = np.ascontiguousarray(D, dtype=np.complex128)
        self.c = self.C.shape[0]
        L = plan.R.L
        self.r = L.refshim_coupled_create(plan.p, self.c, _dp(self.C.view(np.float64)),
                                          _dp(self.B.view(np.float64)), _dp(self.D.view(np.float64)))
        self.order = self.c + 3 * plan.cells

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros_like(x)
        self.plan.R.L.refshim_coupled_apply(self.r, _dp(x.view(np.float64)), _dp(y.view(np.float64)))
        return y

    def gmres(self, b, restart=30, maxit=300, rtol=1e-12, atol=0.0):
        b = np.ascontiguousarray(b, dtype=np.complex128)
        x = np.zeros_like(b)
        rep = np.zeros(6)
        ok = self.plan.R.L.refshim_coupled_gmres(self.r, _dp(b.view(np.float64)), _dp(x.view(np.float64)),
                                                 restart, maxit, rtol, atol, _dp(rep))
        return bool(ok), x, _report(rep)

    def __del__(self):
        if getattr(self, "r", None):
            self.plan.R.L.refshim_coupled_free(self.r)
            self.r = None


class RefPlan:
    def __init__(self, R: ReferenceLibrary, shape, xyz, h, k, er, limit):
        self.R = R
        shp = np.ascontiguousarray(shape, dtype=np.uintp)
        xyz = np.ascontiguousarray(np.asarray(xyz).reshape(-1, 3), dtype=np.uintp)
        self.cells = len(xyz)
        er = complex(er)
        szp = ctypes.POINTER(ctypes.c_size_t)
        self.p = R.L.refshim_create(shp.ctypes.data_as(szp), self.cells, xyz.ctypes.data_as(szp),
                                    float(h), float(k), er.real, er.imag, int(limit))
        if not self.p:
            raise ValueError("vie_fft_create rejected the plan")

    def set_eps(self, cell, er):
        er = complex(er)
        return bool(self.R.L.refshim_set_eps(self.p, cell, er.real, er.imag))

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros_like(x)
        if not self.R.L.refshim_apply(self.p, _dp(x.view(np.float64)), _dp(y.view(np.float64))):
            return None
        return y

    def precondition(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros_like(x)
        self.R.L.refshim_precondition(self.p, _dp(x.view(np.float64)), _dp(y.view(np.float64)))
        return y

    def gmres(self, b, x=None, restart=30, maxit=300, rtol=1e-12, atol=0.0):
        b = np.ascontiguousarray(b, dtype=np.complex128)
        x = np.zeros_like(b) if x is None else np.ascontiguousarray(x, dtype=np.complex128).copy()
        rep = np.zeros(6)
        ok = self.R.L.refshim_gmres(self.p, _dp(b.view(np.float64)), _dp(x.view(np.float64)),
                                    restart, maxit, rtol, atol, _dp(rep))
        return bool(ok), x, _report(rep)

    def bytes(self):
        return self.R.L.refshim_bytes(self.p)

    def set_permittivities(self, eps):
        for i, e in enumerate(np.asarray(eps).reshape(-1)):
            if not self.set_eps(i, e):
                raise ValueError(f"reference rejected permittivity of cell {i}")

    def coupled(self, C, B, D):
        return RefCoupled(self, C, B, D)

    def __del__(self):
        if getattr(self, "p", None):
            self.R.L.refshim_destroy(self.p)
            self.p = None


class RefSystem:
    def __init__(self, R, cells, materials, conductors, k0, pec, extra):
        self.R = R
        n = len(cells)
        centers = np.ascontiguousarray([c.center for c in cells] or np.zeros((0, 3)), dtype=np.float64)
        sizes = np.ascontiguousarray([c.size for c in cells] or np.zeros((0, 3)), dtype=np.float64)
        mids = np.ascontiguousarray([c.material_id for c in cells] or [0], dtype=np.int32)
        matid = np.ascontiguousarray([m.id for m in materials] or [0], dtype=np.int32)
        props = np.ascontiguousarray([[m.epsilon_r, m.epsilon_loss, m.conductivity] for m in materials]
                                     or np.zeros((1, 3)), dtype=np.float64)
        msg = ctypes.create_string_buffer(512)
        ip = ctypes.POINTER(ctypes.c_int)
        tabs = [tuple(getattr(m, "samples", ()) or ()) for m in materials]
        if any(tabs):
            counts = np.ascontiguousarray([len(t) for t in tabs], dtype=np.uintp)
            flat = np.ascontiguousarray([v for t in tabs for smp in t for v in smp] or [0.0], dtype=np.float64)
            self.s = R.L.refshim_system_prepare_tab(n, _dp(centers), _dp(sizes), mids.ctypes.data_as(ip),
                                                    len(materials), matid.ctypes.data_as(ip), _dp(props),
                                                    counts.ctypes.data_as(ctypes.POINTER(ctypes.c_size_t)),
                                                    _dp(flat), conductors, float(k0), int(extra), int(pec),
                                                    msg, 512)
        else:
            self.s = R.L.refshim_system_prepare(n, _dp(centers), _dp(sizes), mids.ctypes.data_as(ip),
                                                len(materials), matid.ctypes.data_as(ip), _dp(props),
                                                conductors, float(k0), int(extra), int(pec), msg, 512)
        self.message = msg.value.decode()
        if not self.s:
            raise PrepareError(self.message)
        self.c = int(conductors)
        self.v = 3 * n

    def set(self, row, col, value):
        value = complex(value)
        return bool(self.R.L.refshim_system_set(self.s, row, col, value.real, value.imag))

    def set_blocks(self, C, B, D):
        """C (c x c), B (c x v), D (v x c) through dielectric_fft_entry."""
        C, B, D = [np.ascontiguousarray(a, dtype=np.complex128) for a in (C, B, D)]
        self.R.L.refshim_system_set_blocks(self.s, self.c, self.v, _dp(C.view(np.float64)),
                                           _dp(B.view(np.float64)), _dp(D.view(np.float64)))

    def storage(self):
        return self.R.L.refshim_system_storage(self.s)

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros_like(x)
        if not self.R.L.refshim_system_apply(self.s, _dp(x.view(np.float64)), _dp(y.view(np.float64))):
            return None
        return y

    def precondition(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros_like(x)
        if not self.R.L.refshim_system_precondition(self.s, _dp(x.view(np.float64)),
                                                    _dp(y.view(np.float64))):
            return None
        return y

    def gmres(self, b, restart=30, maxit=300, rtol=1e-12, atol=0.0):
        b = np.ascontiguousarray(b, dtype=np.complex128)
        x = np.zeros_like(b)
        rep = np.zeros(6)
        ok = self.R.L.refshim_system_gmres(self.s, _dp(b.view(np.float64)), _dp(x.view(np.float64)),
                                           restart, maxit, rtol, atol, _dp(rep))
        return bool(ok), x, _report(rep)

    def __del__(self):
        if getattr(self, "s", None):
            self.R.L.refshim_system_free(self.s)
            self.s = None


def reference_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libvieref.so"))
