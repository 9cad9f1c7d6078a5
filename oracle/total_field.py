"""CPU oracle for the coupled far field -- TEST INFRASTRUCTURE ONLY.

Only tests/ and bench.py's CPU-baseline leg import this module, and only as
the checker; the product library never uses it.

Numpy restatement of the far field of the solved wires and patches plus the
volume current, and the radiated power of the coupled solve (each function
cites the lines it follows):

* ``conductor_far_field``  ffld, radiation.c:36-338 (classic engine: free
                           space, or the PEC image with rrv = rrh = -1;
                           no cliff or radial screen), with fflds,
                           radiation.c:340-373, for the patches
* ``total_far_field``      dw_far_field, dielectric_wire.c:768-791 (the
                           conductor field times NEC's wavelength plus the
                           volume amplitude of oracle/farfield.py projected on
                           theta and phi)
* ``radiated_power``       dw_radiated_power, dielectric_wire.c:793-840
                           (golden-spiral rule over the sphere, or the upper
                           hemisphere over a PEC ground)

NEC's rounded constants (nec2c.h: PI, TP, ETA, CONST3, CONST4) are kept as
NEC defines them.  Segment data and currents are NEC's: positions and lengths
in wavelengths, (a, b, c) the constant, sine and cosine coefficients of each
segment (crnt.air + i crnt.aii, ...), the patch currents 3 per patch
(crnt.cur[n + 3 j ..]), patch areas in square wavelengths.

Parity status: pinned in tests/test_total_field_oracle.py (see there).
"""
from __future__ import annotations

import math

import numpy as np

from oracle import farfield as F

NEC_PI = 3.141592654
NEC_TP = 6.283185308
NEC_ETA = 376.73
NEC_CONST3 = -29.97922085j
NEC_CONST4 = 188.365j
DW_PI = 3.141592653589793238462643383279502884


def _unit_vectors(theta, phi):
    phx, phy = -math.sin(phi), math.cos(phi)
    roz = math.cos(theta)
    thx, thy, thz = roz * phy, -roz * phx, -math.sin(theta)
    rox, roy = -thz * phy, thz * phx
    return phx, phy, roz, thx, thy, thz, rox, roy


def _segment_sum(seg, cur, rox, roy, roz):
    """sum over segments of exp(i 2 pi d.r) (rr + i ri) t (radiation.c:102-140)."""
    x, y, z, si, cab, sab, salp = seg
    ci = np.zeros(3, dtype=np.complex128)
    for i in range(len(x)):
        omega = -(rox * cab[i] + roy * sab[i] + roz * salp[i])
        el = NEC_PI * si[i]
        sill = omega * el
        top, bot = el + sill, el - sill
        a = 2.0 * math.sin(sill) / omega if abs(omega) >= 1.0e-7 else (2.0 - omega * omega * el * el / 3.0) * el
        too = math.sin(top) / top if abs(top) >= 1.0e-7 else 1.0 - top * top / 6.0
        boo = math.sin(bot) / bot if abs(bot) >= 1.0e-7 else 1.0 - bot * bot / 6.0
        b = el * (boo - too)
        c = el * (boo + too)
        A, B, C = cur[3 * i], cur[3 * i + 1], cur[3 * i + 2]
        rr = a * A.real + b * B.imag + c * C.real
        ri = a * A.imag - b * B.real + c * C.imag
        arg = NEC_TP * (x[i] * rox + y[i] * roy + z[i] * roz)
        exa = complex(math.cos(arg), math.sin(arg)) * complex(rr, ri)
        ci += exa * np.array([cab[i], sab[i], salp[i]])
    return ci


def _patch_sum(patches, pcur, rox, roy, roz):
    """fflds (radiation.c:340-373)."""
    pts, area = patches
    e = np.zeros(3, dtype=np.complex128)
    for j in range(len(area)):
        arg = NEC_TP * (rox * pts[j][0] + roy * pts[j][1] + roz * pts[j][2])
        ct = complex(math.cos(arg) * area[j], math.sin(arg) * area[j])
        e += np.asarray(pcur[3 * j:3 * j + 3]) * ct
    r = np.array([rox, roy, roz])
    return NEC_CONST4 * ((e @ r) * r - e)


def conductor_far_field(theta, phi, seg=None, cur=None, patches=None, pcur=None, ground=0):
    """(E_theta, E_phi) of the wires and patches in NEC's units (ffld)."""
    phx, phy, roz, thx, thy, thz, rox, roy = _unit_vectors(theta, phi)
    n = 0 if seg is None else len(seg[0])
    m = 0 if patches is None else len(patches[1])
    ci = np.zeros(3, dtype=np.complex128)
    if n:
        ci = _segment_sum(seg, cur, rox, roy, roz)
        if ground == 1:
            # image segments: roz reversed, rrv = rrh = -1 (so rrh - rrv = 0)
            cim = _segment_sum(seg, cur, rox, roy, -roz)
            ci = np.array([ci[0] - cim[0], ci[1] - cim[1], ci[2] + cim[2]])
        if m == 0:
            return ((ci[0] * thx + ci[1] * thy + ci[2] * thz) * NEC_CONST3,
                    (ci[0] * phx + ci[1] * phy) * NEC_CONST3)
    e = _patch_sum(patches, pcur, rox, roy, roz) if m else np.zeros(3, complex)
    if m and ground == 1:
        g = -_patch_sum(patches, pcur, rox, roy, -roz)
        e = np.array([e[0] + g[0], e[1] + g[1], e[2] - g[2]])
    e = e + ci * NEC_CONST3
    return e[0] * thx + e[1] * thy + e[2] * thz, e[0] * phx + e[1] * phy


def total_far_field(theta, phi, wlam, k0, seg=None, cur=None, patches=None, pcur=None, centers=None, sizes=None,
                    current=None, ground=0):
    """dw_far_field: r E_theta, r E_phi of conductors and volume together."""
    et, ep = conductor_far_field(theta, phi, seg, cur, patches, pcur, ground)
    et, ep = et * wlam, ep * wlam
    if centers is not None and len(centers):
        d = np.array([[math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)]])
        a = F.radiation_amplitudes(centers, sizes, current, k0, d, ground=ground)[0]
        t = np.array([math.cos(theta) * math.cos(phi), math.cos(theta) * math.sin(phi), -math.sin(theta)])
        p = np.array([-math.sin(phi), math.cos(phi), 0.0])
        et, ep = et + a @ t, ep + a @ p
    return et, ep


def spiral(count, ground=0):
    """(theta, phi) of the golden-spiral rule of dw_radiated_power."""
    out = []
    for i in range(count):
        z = 1.0 - (1.0 if ground else 2.0) * (i + 0.5) / count
        out.append((math.acos(z), DW_PI * (3.0 - math.sqrt(5.0)) * i))
    return out


def radiated_power(count, wlam, k0, ground=0, **sources):
    """dw_radiated_power (dielectric_wire.c:793-840)."""
    total = 0.0
    for theta, phi in spiral(count, ground):
        et, ep = total_far_field(theta, phi, wlam, k0, ground=ground, **sources)
        total += (et * et.conjugate() + ep * ep.conjugate()).real
    return total * ((2.0 if ground else 4.0) * DW_PI / count) / (2.0 * NEC_ETA)
