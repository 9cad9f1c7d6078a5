"""Reflected (real-ground) tensor of a lattice of cubes above a lossy
half-space: a numpy restatement of the reference's vie_halfspace_cuboids
(vie_halfspace.c:73-203), evaluated for a whole lattice at once.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the device table
(vie_hip_ground_table); never imported by the product.

The reference integrates, per pair of cuboids, over the transverse
wavenumber u (propagating u = sin t, w = cos t on [0, pi/2]; evanescent
u = hypot(1, ev), w = -i ev on ev = t / scale, t in [0, 1, 4, 16, 64] and the
substrate branch point) and the azimuth phi (periodic trapezoid):
  N = -i k^3/(8 pi^2) V sum radial * zfac * shape * exp(-i k u rho.phi) *
      (rs te te^T + rp up down^T)
with the Fresnel coefficients rs, rp of the substrate (vie_halfspace.c:14-28),
zfac the product of the two cells' z averages (:61-69, exp(-i k w (za + zb))
sinc^2(k w h / 2) for cubes), shape the four lateral sinc factors.  For a
lattice every factor but exp(-i k u (c dx + s dy) h) and exp(-i k w S) is
independent of the pair, so one node set serves all pairs; the node set is
the reference's for the lattice's smallest gap, which integrates every
larger gap (a faster-decaying integrand) as well.  Per node, the azimuthal
integrals of {s^2, c^2, s c, c, s} give the six components
  xx = rs J_ss - rp w^2 J_cc, yy = rs J_cc - rp w^2 J_ss, zz = rp u^2 (J_ss + J_cc),
  xy = -(rs + rp w^2) J_sc, xz = -rp u w J_c, yz = -rp u w J_s   (zx = -xz, zy = -yz).
"""
from __future__ import annotations

import math

import numpy as np


def fresnel(eps, u):
    """rs, rp of the lower half-space at transverse wavenumber u (units of k0),
    on the outgoing / decaying branch (vie_halfspace.c:14-28)."""
    eps = complex(eps)
    if eps == 1:
        return 0j, 0j
    w = math.sqrt((1 - u) * (1 + u)) if u <= 1 else -1j * math.sqrt((u - 1) * (u + 1))
    v = np.sqrt(eps - u * u)
    if v.imag > 0:
        v = -v
    return (w - v) / (w + v), (eps * w - v) / (eps * w + v)


def _sinc(v):
    v = np.asarray(v, dtype=float)
    out = np.ones_like(v)
    big = np.abs(v) >= 1e-5
    out[big] = np.sin(v[big]) / v[big]
    small = ~big
    out[small] = 1 - v[small] ** 2 / 6 + v[small] ** 4 / 120
    return out


def table(shape, h, z0, k, eps, order=48, lateral=None):
    """T[6, nx, ny, 2 nz - 1]: the reflected tensor between cells dx, dy
    apart laterally (test minus source, in cells, >= 0) whose centre heights
    add to 2 z0 + s h; z0 is the lowest layer's centre height.  With
    ``lateral`` = (dx array, dy array) only those offsets are evaluated:
    T[6, len(dx), len(dy), 2 nz - 1] (the node set is still the lattice's)."""
    nx, ny, nz = shape
    eps = complex(eps)
    gap = 2 * z0 - h
    scale = k * gap
    if not (gap > 0 and scale > 0):
        raise ValueError("cells must lie above the ground")
    cut = [0.0, 1.0, 4.0, 16.0, 64.0]
    branch = scale * math.sqrt(max(eps.real - 1, 0.0))
    if 0 < branch < 64:
        cut.append(branch)
    cut.sort()
    gx, gw = np.polynomial.legendre.leggauss(order)
    rho = h * math.hypot(max(nx - 1, 0), max(ny - 1, 0))
    pref = -1j * k ** 3 / (8 * math.pi ** 2) * h ** 3
    L = 2 * nz - 1
    if lateral is None:
        dx, dy = np.arange(nx), np.arange(ny)
    else:
        dx, dy = np.asarray(lateral[0]), np.asarray(lateral[1])
    T = np.zeros((6, len(dx), len(dy), L), complex)
    dx = dx[:, None]
    dy = dy[None, :]
    S = 2 * z0 + np.arange(L) * h
    for seg in range(-1, len(cut) - 1):
        lo = 0.0 if seg < 0 else cut[seg]
        hi = math.pi / 2 if seg < 0 else cut[seg + 1]
        if hi == lo:
            continue
        for j in range(order):
            t = 0.5 * ((hi - lo) * gx[j] + hi + lo)
            rw = 0.5 * (hi - lo) * gw[j]
            if seg < 0:
                u, w, radial = math.sin(t), math.cos(t), math.sin(t) * rw
            else:
                ev = t / scale
                u, w, radial = math.hypot(1.0, ev), -1j * ev, 1j * rw / scale
            rs, rp = fresnel(eps, u)
            nphi = max(int(math.ceil(1.25 * k * u * rho + 48.0)), 2 * order)
            nphi = (nphi + 3) // 4 * 4
            phi = 2 * math.pi * np.arange(nphi) / nphi
            c, s = np.cos(phi), np.sin(phi)
            q = (4 * np.arange(nphi)) % nphi == 0          # exact quadrant points
            c[q] = np.round(c[q])
            s[q] = np.round(s[q])
            shp = _sinc(0.5 * k * u * c * h) ** 2 * _sinc(0.5 * k * u * s * h) ** 2 * (2 * math.pi / nphi)
            ph = np.exp(-1j * k * u * h * (c[:, None, None] * dx[None] + s[:, None, None] * dy[None]))
            J = [np.tensordot(shp * f, ph, axes=(0, 0)) for f in (s * s, c * c, s * c, c, s)]
            kw = k * w
            a = 0.5 * kw * h
            sc = np.sin(a) / a if abs(a) >= 1e-4 else 1 - a * a / 6 + a ** 4 / 120
            Z = np.exp(-1j * kw * S) * sc * sc
            G = [rs * J[0] - rp * w * w * J[1], rs * J[1] - rp * w * w * J[0], rp * u * u * (J[0] + J[1]),
                 -(rs + rp * w * w) * J[2], -rp * u * w * J[3], -rp * u * w * J[4]]
            wgt = pref * radial
            for comp in range(6):
                T[comp] += wgt * G[comp][:, :, None] * Z[None, None, :]
    return T


COMPONENTS = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))


def tensor(T, dx, dy, s):
    """The 3 x 3 reflected tensor at a signed lateral offset (dx, dy) and
    layer sum s from the table (parities: xy odd in x and y, xz odd in x, yz
    odd in y; zx = -xz, zy = -yz)."""
    sx, sy = (1 if dx >= 0 else -1), (1 if dy >= 0 else -1)
    v = T[:, abs(dx), abs(dy), s]
    xx, yy, zz, xy, xz, yz = v[0], v[1], v[2], sx * sy * v[3], sx * v[4], sy * v[5]
    return np.array([[xx, xy, xz], [xy, yy, yz], [-xz, -yz, zz]])


def reflected_matrix(T, xyz):
    """Dense [3N, 3N] reflected part N_R for cells at integer lattice
    coordinates xyz (rows: test cell, columns: source cell; 3 per cell)."""
    xyz = np.asarray(xyz, dtype=int).reshape(-1, 3)
    n = len(xyz)
    M = np.zeros((3 * n, 3 * n), complex)
    for a in range(n):
        for b in range(n):
            d = xyz[a] - xyz[b]
            M[3 * a:3 * a + 3, 3 * b:3 * b + 3] = tensor(T, d[0], d[1], xyz[a][2] + xyz[b][2])
    return M
