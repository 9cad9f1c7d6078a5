"""CPU oracle for the far-field post-processing row -- TEST INFRASTRUCTURE ONLY.

Only tests/ and bench.py's CPU-baseline leg import this module, and only as
the checker; the product library never uses it.

Numpy restatement of the reference's scattered far field of a volume
polarisation current (each function cites the lines it follows):

* ``sinc_real``                 dielectric_solver.c:47-55
* ``cell_phase_average``        dielectric_solver.c:95-107
* ``ground_reflect_matrix``     dielectric_ground.c:133-144 (the reflection of
                                dielectric_ground_reflect as a 3x3 matrix per direction)
* ``radiation_amplitudes``      dielectric_solver.c:124-158, vectorised over
                                directions (cells looped in blocks)
* ``golden_directions``         dielectric_solver.c:180-187
* ``scattered_power``           dielectric_solver.c:160-217

Parity status: pinned in tests/test_farfield_oracle.py by closed forms (the
Hertzian-dipole far field and radiated power of one small cell, the image
pair of a dipole over a PEC plane), by the translation phase law, by the
sinc cell factor against brute-force quadrature, and by a pure-Python loop
over cells that follows the reference statement by statement.
"""
from __future__ import annotations

import math

import numpy as np

EPSILON_0 = 8.8541878128e-12  # dielectric_solver.c:20-23
MU_0 = 1.25663706212e-6
C0 = 299792458.0
PI = 3.141592653589793238462643383279502884
ETA0 = math.sqrt(MU_0 / EPSILON_0)


def sinc_real(x):
    """sin(x)/x with the reference's series below |x| < 1e-5 (dielectric_solver.c:47-55)."""
    x = np.asarray(x, dtype=np.float64)
    x2 = x * x
    small = np.abs(x) < 1.0e-5
    safe = np.where(small, 1.0, x)
    return np.where(small, 1.0 - x2 / 6.0 + x2 * x2 / 120.0, np.sin(safe) / safe)


def cell_phase_average(centers, sizes, directions, k0, sign):
    """[D, N] cell averages of exp(i sign k0 d.r) (dielectric_solver.c:95-107)."""
    d = np.asarray(directions, dtype=np.float64)
    avg = np.ones((len(d), len(centers)))
    for axis in range(3):
        avg = avg * sinc_real(0.5 * k0 * np.outer(d[:, axis], sizes[:, axis]))
    return avg * np.exp(1j * sign * k0 * (d @ centers.T))


def ground_reflect_matrix(direction, rs=-1.0, rp=1.0):
    """R with reflected = R field for dielectric_ground_reflect(down, field)
    (dielectric_ground.c:133-144); ``direction`` is the upward observation
    direction, down = (dx, dy, -dz) as dielectric_solver.c:149 passes it.
    rs = -1, rp = +1 is the PEC ground (gnd.iperf == 1)."""
    down = (direction[0], direction[1], -direction[2])
    u = math.hypot(down[0], down[1])
    c = down[0] / u if u > 1e-14 else 1.0
    s = down[1] / u if u > 1e-14 else 0.0
    w = abs(down[2])
    te = np.array([-s, c, 0.0])
    pd = np.array([-w * c, -w * s, -u])
    pu = np.array([w * c, w * s, -u])
    return rs * np.outer(te, te) + rp * np.outer(pu, pd)


def radiation_amplitudes(centers, sizes, current, k0, directions, ground=0, rs=None, rp=None, block=4096):
    """amplitude[D, 3] of dielectric_radiation_amplitude for every direction
    (dielectric_solver.c:124-158).  ground 0: free space (gnd.ksymp == 1);
    1: PEC plane (image cells, rs = -1, rp = +1); 2: per-direction rs, rp."""
    centers = np.asarray(centers, dtype=np.float64).reshape(-1, 3)
    sizes = np.asarray(sizes, dtype=np.float64).reshape(-1, 3)
    J = np.asarray(current, dtype=np.complex128).reshape(-1, 3)
    d = np.asarray(directions, dtype=np.float64).reshape(-1, 3)
    volume = np.prod(sizes, axis=1)
    image_centers = centers * np.array([1.0, 1.0, -1.0])
    direct = np.zeros((len(d), 3), dtype=np.complex128)
    image = np.zeros((len(d), 3), dtype=np.complex128)
    for c0 in range(0, len(centers), block):
        sl = slice(c0, c0 + block)
        w = volume[sl] * cell_phase_average(centers[sl], sizes[sl], d, k0, 1.0)
        direct += w @ J[sl]
        if ground:
            wi = volume[sl] * cell_phase_average(image_centers[sl], sizes[sl], d, k0, 1.0)
            image += wi @ J[sl]
    integral = direct
    if ground:
        refl = np.empty_like(image)
        for q in range(len(d)):
            r_s = -1.0 if ground == 1 else rs[q]
            r_p = 1.0 if ground == 1 else rp[q]
            refl[q] = ground_reflect_matrix(d[q], r_s, r_p) @ image[q]
        integral = integral + refl
    crossed = np.cross(integral, d)
    transverse = np.cross(crossed, d)
    return 1j * k0 * C0 * MU_0 * transverse / (4.0 * PI)


def radiation_amplitude_loop(centers, sizes, current, k0, direction, ground=0):
    """One direction, cell by cell, statement for statement as
    dielectric_solver.c:124-158 (small cases only)."""
    integral = [0j, 0j, 0j]
    for cell in range(len(centers)):
        size = sizes[cell]
        volume = size[0] * size[1] * size[2]
        avg = 1.0
        for axis in range(3):
            avg *= float(sinc_real(0.5 * k0 * direction[axis] * size[axis]))
        dot = sum(direction[a] * centers[cell][a] for a in range(3))
        weight = volume * avg * complex(math.cos(k0 * dot), math.sin(k0 * dot))
        for comp in range(3):
            integral[comp] += current[3 * cell + comp] * weight
        if ground:
            img = (centers[cell][0], centers[cell][1], -centers[cell][2])
            dot = sum(direction[a] * img[a] for a in range(3))
            weight = volume * avg * complex(math.cos(k0 * dot), math.sin(k0 * dot))
            refl = ground_reflect_matrix(direction) @ np.asarray(current[3 * cell:3 * cell + 3])
            for comp in range(3):
                integral[comp] += refl[comp] * weight
    integral = np.asarray(integral)
    transverse = np.cross(np.cross(integral, direction), direction)
    return 1j * k0 * C0 * MU_0 * transverse / (4.0 * PI)


def golden_directions(count=974, ground=0):
    """The golden-spiral quadrature nodes of scattered_power
    (dielectric_solver.c:180-187): the whole sphere, or the upper hemisphere
    over a ground."""
    golden = PI * (3.0 - math.sqrt(5.0))
    i = np.arange(count, dtype=np.float64)
    z = 1.0 - (1.0 if ground else 2.0) * (i + 0.5) / count
    radial = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    az = golden * i
    return np.stack([radial * np.cos(az), radial * np.sin(az), z], axis=1)


def scattered_power(centers, sizes, current, k0, ground=0, count=974):
    """Scattered power in W (dielectric_solver.c:160-217)."""
    d = golden_directions(count, ground)
    a = radiation_amplitudes(centers, sizes, current, k0, d, ground=1 if ground else 0)
    total = float(np.sum(np.real(np.sum(a * np.conj(a), axis=1))))
    return ((2.0 if ground else 4.0) * PI / count) * total / (2.0 * ETA0)
