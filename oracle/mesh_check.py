"""CPU oracle of the mesher's relationship checks -- TEST INFRASTRUCTURE ONLY.

Only tests/ and bench.py's CPU-baseline leg import this module, and only as
the checker; the product library never uses it.

ctypes front of oracle/mesh_check.c (the loop-for-loop restatement of
validate_box_relationships, dielectric_mesh.c:75-123, and evaluate_clearance,
dielectric_mesh.c:125-316) and of oracle/_ref/libmeshref.so (the reference's
own dielectric_mesh_build compiled from its sources, where that was built).

Parity status: pinned in tests/test_mesh_check_oracle.py (see there).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_L = None
_R = None
vp, sz, i, d = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_double
ll = ctypes.POINTER(ctypes.c_longlong)


def _lib():
    global _L
    if _L is None:
        path = os.path.join(HERE, "_build", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/_build/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(path)
        L.oracle_box_overlap.argtypes = [vp, vp, sz, sz, sz, ll, ll]
        L.oracle_box_overlap.restype = None
        L.oracle_clearance.argtypes = [vp, vp, vp, sz, sz, sz, i] + [vp] * 8 + [i, vp, vp, vp, vp]
        L.oracle_clearance.restype = None
        L.oracle_segment_box_d2.argtypes = [vp, vp, vp, vp]
        L.oracle_segment_box_d2.restype = d
        _L = L
    return _L


def ref_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libmeshref.so"))


def _ref():
    global _R
    if _R is None:
        R = ctypes.CDLL(os.path.join(HERE, "_ref", "libmeshref.so"))
        R.ref_mesh_check.argtypes = [sz, vp, vp, d, d, sz, i] + [vp] * 8 + [i] + [vp] * 4 + \
            [vp, vp, ctypes.POINTER(sz), vp, ctypes.POINTER(i), vp]
        R.ref_mesh_check.restype = i
        R.ref_mesh_cells.argtypes = [sz, vp, vp, d, d, sz, vp, vp, vp, sz, ctypes.POINTER(sz), vp, vp]
        R.ref_mesh_cells.restype = i
        _R = R
    return _R


def _f64(a, shape=(-1,)):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(shape)


def box_overlap(minimum, maximum, first_left=0, last_left=None):
    """0-based first overlapping pair (left, right) or None; left restricted
    to [first_left, last_left) for a bounded sample."""
    mn, mx = _f64(minimum, (-1, 3)), _f64(maximum, (-1, 3))
    a, b = ctypes.c_longlong(), ctypes.c_longlong()
    _lib().oracle_box_overlap(mn.ctypes.data, mx.ctypes.data, len(mn), first_left,
                              len(mn) if last_left is None else last_left, ctypes.byref(a), ctypes.byref(b))
    return None if a.value < 0 else (a.value, b.value)


def segment_box_d2(start, end, center, size):
    s, e, c, h = (_f64(v) for v in (start, end, center, size))
    return _lib().oracle_segment_box_d2(s.ctypes.data, e.ctypes.data, c.ctypes.data, h.ctypes.data)


def clearance(minimum, maximum, cell_longest, segments=None, patches=None, first_box=0, last_box=None):
    """Dict of worst_margin, worst_gap, worst_required, minimum_gap,
    required_gap, worst_box, worst_conductor, worst_is_patch."""
    mn, mx = _f64(minimum, (-1, 3)), _f64(maximum, (-1, 3))
    cl = _f64(cell_longest)
    seg = [_f64(a) for a in (segments or [[]] * 8)]
    pts = _f64(patches[0] if patches else np.zeros((0, 3)), (-1, 3))
    area = _f64(patches[1] if patches else np.zeros(0))
    out = np.zeros(5)
    idx = np.zeros(3, dtype=np.int64)
    _lib().oracle_clearance(mn.ctypes.data, mx.ctypes.data, cl.ctypes.data, len(mn), first_box,
                            len(mn) if last_box is None else last_box, len(seg[0]), *[a.ctypes.data for a in seg],
                            len(area), pts.ctypes.data, area.ctypes.data, out.ctypes.data, idx.ctypes.data)
    return {"worst_margin": out[0], "worst_gap": out[1], "worst_required": out[2], "minimum_gap": out[3],
            "required_gap": out[4], "worst_box": int(idx[0]), "worst_conductor": int(idx[1]),
            "worst_is_patch": int(idx[2])}


def reference_mesh(minimum, maximum, frequency_mhz=1.0, max_kh=0.5, max_cells=1_000_000, segments=None,
                   patches=None):
    """The reference's dielectric_mesh_build on one material and one region
    per box: (ok, record, value, line, minimum_gap, required_gap,
    has_conductors, cell_longest per box)."""
    mn, mx = _f64(minimum, (-1, 3)), _f64(maximum, (-1, 3))
    seg = [_f64(a).copy() for a in (segments or [[]] * 8)]
    pts = _f64(patches[0] if patches else np.zeros((0, 3)), (-1, 3))
    area = _f64(patches[1] if patches else np.zeros(0)).copy()
    px, py, pz = (np.ascontiguousarray(pts[:, a]) for a in range(3))
    rec, val = ctypes.create_string_buffer(32), ctypes.create_string_buffer(128)
    line, has = ctypes.c_size_t(), ctypes.c_int()
    gaps = np.zeros(2)
    cl = np.zeros(len(mn))
    ok = _ref().ref_mesh_check(len(mn), mn.ctypes.data, mx.ctypes.data, frequency_mhz, max_kh, max_cells,
                               len(seg[0]), *[a.ctypes.data for a in seg], len(area), px.ctypes.data,
                               py.ctypes.data, pz.ctypes.data, area.ctypes.data, rec, val, ctypes.byref(line),
                               gaps.ctypes.data, ctypes.byref(has), cl.ctypes.data)
    return (bool(ok), rec.value.decode(), val.value.decode(), line.value, gaps[0], gaps[1], has.value, cl)


def reference_cells(minimum, maximum, frequency_mhz=1.0, max_kh=0.5, max_cells=1_000_000):
    """Cells of the reference's dielectric_mesh_build for one region per box:
    (ok, record, value, centers [n, 3], sizes [n, 3], box index [n])."""
    mn, mx = _f64(minimum, (-1, 3)), _f64(maximum, (-1, 3))
    cap = max_cells
    c, s, b = np.zeros((cap, 3)), np.zeros((cap, 3)), np.zeros(cap, dtype=np.uintp)
    n = ctypes.c_size_t()
    rec, val = ctypes.create_string_buffer(32), ctypes.create_string_buffer(128)
    ok = _ref().ref_mesh_cells(len(mn), mn.ctypes.data, mx.ctypes.data, frequency_mhz, max_kh, max_cells,
                               c.ctypes.data, s.ctypes.data, b.ctypes.data, cap, ctypes.byref(n), rec, val)
    k = min(n.value, cap)
    return bool(ok), rec.value.decode(), val.value.decode(), c[:k], s[:k], b[:k]


def wire(start, end, radius, segments=1):
    """NEC segment arrays (x, y, z, si, cab, sab, salp, bi) of a straight wire
    in metres, as the GW card divides it."""
    a, b = np.asarray(start, float), np.asarray(end, float)
    t = np.arange(segments + 1) / segments
    ends = a + np.outer(t, b - a)
    c = 0.5 * (ends[:-1] + ends[1:])
    dv = ends[1:] - ends[:-1]
    si = np.linalg.norm(dv, axis=1)
    u = dv / si[:, None]
    return [c[:, 0], c[:, 1], c[:, 2], si, u[:, 0], u[:, 1], u[:, 2], np.full(segments, float(radius))]
