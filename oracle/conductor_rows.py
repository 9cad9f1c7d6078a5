"""CPU oracle for the volume-to-conductor rows -- TEST INFRASTRUCTURE ONLY.

Only tests/ and bench.py's CPU-baseline leg import this module, and only as
the checker; the product library never uses it.

Numpy restatement of the reference's point-tested fields of a constant
cuboid current and the conductor rows they fill (each function cites the
lines it follows):

* ``gauss``                 dw_gauss, dielectric_wire.c:70-98 (Newton on P_n)
* ``point_cuboid_n``        dw_point_cuboid_n_free, dielectric_wire.c:101-142
* ``point_cuboid_h``        dw_point_cuboid_h_free, dielectric_wire.c:144-181
* ``point_background``      dw_point_background, dielectric_wire.c:183-214
                            (free space, or the PEC image with gnd.iperf == 1)
* ``conductor_rows``        wire rows dielectric_wire.c:1164-1213 and patch
                            rows :1215-1262 (row order: wires, then two rows
                            per patch)
* ``conductor_rows_c``      the same rows from the C restatement
                            oracle/conductor_rows.c (OpenMP over the tasks, as
                            the reference runs them; the CPU baseline)

Parity status: pinned in tests/test_conductor_rows_oracle.py by numpy's
Gauss-Legendre rule, finite differences of the scalar Green function for
the E dyadic (k^2 + grad grad) g and for grad g x e_q, the point-source limit
of a small cell far away, image theory over a PEC plane, the wire rows
against the mixed-potential field (vector potential plus face charges) of the
cell current, the patch rows against NEC's patch equations (matrix.c cmss)
formed from the Biot-Savart field; the C restatement entry for entry against
this module.
"""
from __future__ import annotations

import math

import numpy as np

EPSILON_0 = 8.8541878128e-12  # dielectric_wire.c:33-35
C0 = 299792458.0
PI = 3.141592653589793238462643383279502884


def gauss(order):
    """Gauss-Legendre nodes (ascending) and weights by the reference's Newton
    iteration (dielectric_wire.c:70-98)."""
    nodes = np.zeros(order)
    weights = np.zeros(order)
    for root in range((order + 1) // 2):
        z = math.cos(PI * (root + 0.75) / (order + 0.5))
        for _ in range(32):
            p0, p1 = 1.0, z
            for degree in range(2, order + 1):
                p0, p1 = p1, ((2.0 * degree - 1.0) * z * p1 - (degree - 1.0) * p0) / degree
            derivative = order * (z * p1 - p0) / (z * z - 1.0)
            previous = z
            z = previous - p1 / derivative
            if abs(z - previous) <= 4.0 * np.finfo(float).eps:
                break
        # the weight uses the derivative of the last Newton step, as the reference does
        nodes[root] = -z
        nodes[order - 1 - root] = z
        weights[root] = 2.0 / ((1.0 - z * z) * derivative * derivative)
        weights[order - 1 - root] = weights[root]
    return nodes, weights


def _points(center, size, order):
    x, w = gauss(order)
    pts = [center[a] + 0.5 * size[a] * x for a in range(3)]
    wts = [0.5 * size[a] * w for a in range(3)]
    P = np.stack(np.meshgrid(*pts, indexing="ij"), -1).reshape(-1, 3)
    W = np.einsum("i,j,k->ijk", *wts).reshape(-1)
    return P, W


def point_cuboid_n(observation, center, size, k0, order=10):
    """3x3 E dyadic of a constant unit current over the cuboid, point-tested
    (dielectric_wire.c:101-142): sum w g [(k^2 - ik/r - 1/r^2) I
    + (-k^2 + 3ik/r + 3/r^2) rr/r^2], g = exp(-ikr)/(4 pi r)."""
    P, W = _points(center, size, order)
    d = np.asarray(observation, dtype=np.float64) - P
    r2 = np.sum(d * d, axis=1)
    if not (r2 > 0).all():
        raise ValueError("observation point on a quadrature node")
    r = np.sqrt(r2)
    g = np.exp(-1j * k0 * r) / (4.0 * PI * r)
    diag = g * (k0 * k0 - 1j * k0 / r - 1.0 / r2)
    radial = g * (-k0 * k0 + 3j * k0 / r + 3.0 / r2)
    out = np.eye(3) * np.sum(W * diag)
    out = out + np.einsum("n,np,nq->pq", W * radial / r2, d, d)
    return out


def point_cuboid_h(observation, center, size, k0, order=10):
    """3x3 H dyadic, column q = integral grad g x e_q (dielectric_wire.c:144-181)."""
    P, W = _points(center, size, order)
    d = np.asarray(observation, dtype=np.float64) - P
    r2 = np.sum(d * d, axis=1)
    if not (r2 > 0).all():
        raise ValueError("observation point on a quadrature node")
    r = np.sqrt(r2)
    g = np.exp(-1j * k0 * r) / (4.0 * PI * r)
    factor = W * g * (-1j * k0 - 1.0 / r) / r
    grad = np.sum(factor[:, None] * d, axis=0)
    out = np.zeros((3, 3), dtype=np.complex128)
    out[1, 0] += grad[2]
    out[2, 0] -= grad[1]
    out[0, 1] -= grad[2]
    out[2, 1] += grad[0]
    out[0, 2] += grad[1]
    out[1, 2] -= grad[0]
    return out


def point_background(observation, center, size, k0, magnetic=False, ground=0):
    """dw_point_background (dielectric_wire.c:183-214): free space, or over a
    PEC plane the image cell with column signs (q == 2 ? +1 : -1)."""
    f = point_cuboid_h if magnetic else point_cuboid_n
    out = f(observation, center, size, k0)
    if ground == 1:
        image = np.array([center[0], center[1], -center[2]])
        refl = f(observation, image, size, k0)
        out = out + refl * np.array([-1.0, -1.0, 1.0])[None, :]
    return out


def conductor_rows(k0, centers, sizes, wire_points=(), wire_tangents=(), patch_points=(), patch_t1=(),
                   patch_t2=(), patch_scale=(), ground=0):
    """Rows [wires + 2 patches, 3 cells] of the volume-to-conductor block
    (dielectric_wire.c:1164-1262)."""
    omega = k0 * C0
    nw, npch = len(wire_points), len(patch_points)
    rows = np.zeros((nw + 2 * npch, 3 * len(centers)), dtype=np.complex128)
    for w in range(nw):
        t = np.asarray(wire_tangents[w])
        for c in range(len(centers)):
            block = point_background(wire_points[w], centers[c], sizes[c], k0, False, ground)
            rows[w, 3 * c:3 * c + 3] = (t @ block) / (1j * omega * EPSILON_0)
    for p in range(npch):
        t1, t2 = np.asarray(patch_t1[p]), np.asarray(patch_t2[p])
        for c in range(len(centers)):
            block = point_background(patch_points[p], centers[c], sizes[c], k0, True, ground)
            rows[nw + 2 * p, 3 * c:3 * c + 3] = -patch_scale[p] * (t2 @ block)
            rows[nw + 2 * p + 1, 3 * c:3 * c + 3] = -patch_scale[p] * (t1 @ block)
    return rows


_CLIB = None


def _clib():
    """oracle/_build/liboracle.so (built by oracle/Makefile), which holds the
    C restatement oracle/conductor_rows.c."""
    global _CLIB
    if _CLIB is None:
        import ctypes
        import os
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: run make -C {os.path.dirname(path)}")
        L = ctypes.CDLL(path)
        vp, d, i, sz = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_size_t
        L.oracle_conductor_rows.restype = i
        L.oracle_conductor_rows.argtypes = [d, sz, vp, vp, i, vp, vp, i, vp, vp, vp, vp, i, vp, sz, i,
                                            ctypes.POINTER(ctypes.c_longlong)]
        _CLIB = L
    return _CLIB


def conductor_rows_c(k0, centers, sizes, wire_points=(), wire_tangents=(), patch_points=(), patch_t1=(),
                     patch_t2=(), patch_scale=(), ground=0, threads=1):
    """The same rows from the C restatement (oracle/conductor_rows.c), OpenMP
    over the (conductor, cell) tasks with `threads` threads.  Returns (rows,
    first failing task * cells + cell or -1)."""
    import ctypes

    def arr(x, width):
        a = np.ascontiguousarray(np.zeros((0, width)) if len(x) == 0 else x, dtype=np.float64)
        return a.reshape(-1, width) if width > 1 else a.reshape(-1)
    c, s = arr(centers, 3), arr(sizes, 3)
    wp, wt, pp = arr(wire_points, 3), arr(wire_tangents, 3), arr(patch_points, 3)
    t1, t2, ps = arr(patch_t1, 3), arr(patch_t2, 3), arr(patch_scale, 1)
    rows = np.zeros((len(wp) + 2 * len(pp), 3 * len(c)), dtype=np.complex128)
    failed = ctypes.c_longlong(-1)
    _clib().oracle_conductor_rows(float(k0), len(c), c.ctypes.data, s.ctypes.data, len(wp), wp.ctypes.data,
                                  wt.ctypes.data, len(pp), pp.ctypes.data, t1.ctypes.data, t2.ctypes.data,
                                  ps.ctypes.data, int(ground), rows.ctypes.data, rows.shape[1], int(threads),
                                  ctypes.byref(failed))
    return rows, failed.value
