"""CPU oracle for the conductor-to-volume coupling -- TEST INFRASTRUCTURE ONLY.

Only tests/ and bench.py's CPU-baseline leg import this module, and only as
the checker; the product library never uses it.

ctypes front of oracle/wire_volume.c, the C restatement of NEC's classic
near field of frozen basis coefficients (nefld / efld / eksc / intx / gx /
unere, fields.c and calculations.c) averaged over each volume cell
(dw_average_wire_field, dielectric_wire.c:316-357) and scaled into the
volume rows of the conductor columns (dw_wire_to_volume_range,
dielectric_wire.c:399-420).  Segment and patch data are in wavelengths, as
NEC holds them (data.x, data.si, data.bi, data.cab ... and data.px, data.pbi,
t1, t2); cells in metres; wlam is NEC's wavelength (299.8 / f_MHz m).

Parity status: pinned in tests/test_wire_volume_oracle.py (see the C file's
header for the list of checks).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

C0 = 299792458.0
EPSILON_0 = 8.8541878128e-12
NEC_CONST1 = 4.771341189j
NEC_ETA = 376.73
NEC_TP = 6.283185308

_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: run make -C {os.path.dirname(path)}")
        L = ctypes.CDLL(path)
        vp, d, i, sz = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_size_t
        L.oracle_conductor_volume.restype = i
        L.oracle_conductor_volume.argtypes = [d, d, sz, vp, vp, vp, i] + [vp] * 8 + [i, vp, vp, vp, vp, sz, vp, d, i,
                                                                               vp, i]
        L.oracle_nec_near_field.restype = i
        L.oracle_nec_near_field.argtypes = [i] + [vp] * 8 + [i, vp, vp, vp, vp, d, i, d, vp, vp, vp]
        L.oracle_nec_eksc.restype = None
        L.oracle_nec_eksc.argtypes = [d, d, d, vp]
        L.oracle_nec_intx.restype = None
        L.oracle_nec_intx.argtypes = [d, d, d, d, ctypes.POINTER(d), ctypes.POINTER(d)]
        _LIB = L
    return _LIB


def _f64(a, width=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(-1, width) if width else a.reshape(-1)


@dataclass
class Conductors:
    """NEC's segment and patch arrays, in wavelengths."""
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    si: np.ndarray
    bi: np.ndarray
    cab: np.ndarray
    sab: np.ndarray
    salp: np.ndarray
    patch_points: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    patch_area: np.ndarray = field(default_factory=lambda: np.zeros(0))
    t1: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    t2: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))

    def arrays(self):
        seg = [_f64(a) for a in (self.x, self.y, self.z, self.si, self.bi, self.cab, self.sab, self.salp)]
        pat = [_f64(self.patch_points, 3), _f64(self.patch_area), _f64(self.t1, 3), _f64(self.t2, 3)]
        return seg, pat

    @property
    def n(self):
        return len(np.atleast_1d(self.x))

    @property
    def m(self):
        return len(np.atleast_1d(self.patch_area))


def straight_wire(start, end, segments, radius):
    """NEC's segmentation of a straight wire (GW card): equal segments,
    centres, lengths, radius and direction cosines, in the units given."""
    start, end = np.asarray(start, float), np.asarray(end, float)
    d = (end - start) / segments
    centers = start + (np.arange(segments)[:, None] + 0.5) * d
    length = np.linalg.norm(d)
    t = d / length
    return Conductors(centers[:, 0], centers[:, 1], centers[:, 2], np.full(segments, length),
                      np.full(segments, radius), np.full(segments, t[0]), np.full(segments, t[1]),
                      np.full(segments, t[2]))


def scaled(c: Conductors, wlam):
    """The same conductors with lengths divided by wlam (metres -> wavelengths)."""
    return Conductors(c.x / wlam, c.y / wlam, c.z / wlam, c.si / wlam, c.bi / wlam, c.cab, c.sab, c.salp,
                      np.asarray(c.patch_points) / wlam, np.asarray(c.patch_area) / wlam ** 2, c.t1, c.t2)


def near_field(cond: Conductors, sources, point, wlam, rkh=1.0, ground=0):
    """NEC's near field E (x, y, z) at a point in metres of one source state
    (sources: 3 per segment (a, b, c), then 3 per patch)."""
    seg, pat = cond.arrays()
    src = np.ascontiguousarray(sources, dtype=np.complex128)
    p = _f64(point)
    e = np.zeros(3, dtype=np.complex128)
    _lib().oracle_nec_near_field(cond.n, *[a.ctypes.data for a in seg], cond.m, *[a.ctypes.data for a in pat],
                                 float(rkh), int(ground), float(wlam), src.ctypes.data, p.ctypes.data, e.ctypes.data)
    return e


def eksc(s, z, rh):
    """{ezs, ers, ezc, erc, ezk, erk} of a segment of length s (wavelengths)."""
    e = np.zeros(6, dtype=np.complex128)
    _lib().oracle_nec_eksc(float(s), float(z), float(rh), e.ctypes.data)
    return e


def intx(zpk, rkb2, el1, el2):
    c, s = ctypes.c_double(), ctypes.c_double()
    _lib().oracle_nec_intx(float(zpk), float(rkb2), float(el1), float(el2), ctypes.byref(c), ctypes.byref(s))
    return c.value + 1j * s.value


def conductor_volume(k0, wlam, centers, sizes, chi, cond: Conductors, sources, rkh=1.0, ground=0, threads=1):
    """[columns, 3 cells] volume rows of each conductor column
    (dielectric_wire.c:399-420), from the C restatement with `threads`
    OpenMP threads."""
    c, s = _f64(centers, 3), _f64(sizes, 3)
    ch = np.ascontiguousarray(np.broadcast_to(chi, (len(c),)), dtype=np.complex128)
    seg, pat = cond.arrays()
    stride = 3 * cond.n + 3 * cond.m
    src = np.ascontiguousarray(sources, dtype=np.complex128).reshape(-1, stride)
    out = np.zeros((len(src), 3 * len(c)), dtype=np.complex128)
    ok = _lib().oracle_conductor_volume(float(k0), float(wlam), len(c), c.ctypes.data, s.ctypes.data, ch.ctypes.data,
                                        cond.n, *[a.ctypes.data for a in seg], cond.m,
                                        *[a.ctypes.data for a in pat], len(src), src.ctypes.data, float(rkh),
                                        int(ground), out.ctypes.data, int(threads))
    if not ok:
        raise RuntimeError("oracle conductor-volume assembly failed")
    return out
