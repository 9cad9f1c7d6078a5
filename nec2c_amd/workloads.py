"""Synthetic inputs of the benchmark configurations (host-side, numpy).

* ``random_permittivity`` -- the Gaussian random field of the 64^3 sweep:
  white noise filtered by exp(-r^2 / 2 l^2) (l = 4 voxels), standardised, put
  through the normal CDF onto [lo, hi]; loss tangent uniform on [0, tmax];
  eps = eps_r (1 - j tan d), passive under exp(+j w t) (dielectric.h:55).
* ``Dipole`` -- a z-directed thin-wire dipole with a 1 V delta gap at its centre
  segment, discretised by Harrington's pulse-basis point-matching method of
  moments (R. F. Harrington, Field Computation by Moment Methods, 1968, ch. 4):
  Z_mn = j w mu D^2 psi(m,n) + [psi(m+,n+) - psi(m-,n+) - psi(m+,n-) + psi(m-,n-)] / (j w eps),
  psi the segment-averaged reduced kernel exp(-jkR)/(4 pi R).
* ``coupling_blocks`` -- the cross blocks of the reference's hybrid operator
  [C B; D V] (dielectric_fft.c:41-66) for that dipole and a cube lattice, from
  the free-space dyadic (k^2 I + grad grad) g evaluated at cell centres:
  the volume unknown is the normalised polarisation (eps - 1) E, so
  D = (eps_i - 1) (j w mu D / k^2) G_{i z} (incident field of each wire
  segment, moved to the left-hand side) and B = -D h^3 G_{z j} (voltage induced
  on each segment by the cell's polarisation current).

These are stand-ins of the right shape and physical scale for the blocks the
reference assembles with NEC's wire code (dielectric_wire.c:422-508,
1164-1213), which this project does not rebuild; parity is always stated
against the reference solving the same [C B; D V] system.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

C0 = 299792458.0
MU0 = 4e-7 * math.pi
EPS0 = 1.0 / (MU0 * C0 * C0)


def k0_of(freq_hz: float) -> float:
    return 2 * math.pi * freq_hz / C0


def _norm_cdf(x):
    from scipy.special import ndtr
    return ndtr(x)


def random_permittivity(shape, seed, corr=4.0, lo=1.5, hi=10.0, tan_max=0.05):
    """Per-voxel complex permittivity [nx, ny, nz] (x slowest, z fastest)."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(shape)
    freqs = [np.fft.fftfreq(n) for n in shape]
    kx, ky, kz = np.meshgrid(*freqs, indexing="ij")
    # FFT of exp(-r^2/2l^2) is proportional to exp(-2 pi^2 l^2 |f|^2)
    filt = np.exp(-2 * math.pi ** 2 * corr ** 2 * (kx ** 2 + ky ** 2 + kz ** 2))
    field = np.real(np.fft.ifftn(np.fft.fftn(noise) * filt))
    field = (field - field.mean()) / field.std()
    er = lo + (hi - lo) * _norm_cdf(field)
    tan = rng.uniform(0.0, tan_max, shape)
    return er * (1 - 1j * tan)


def _gauss01(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1), 0.5 * w


@dataclass
class Dipole:
    center: tuple      # (x, y, z) of the feed, metres
    length: float = 0.48
    segments: int = 21
    radius: float = 1e-3

    @property
    def delta(self):
        return self.length / self.segments

    def points(self):
        """Segment centres along z."""
        z0 = self.center[2] - 0.5 * self.length + 0.5 * self.delta
        return np.array([[self.center[0], self.center[1], z0 + i * self.delta] for i in range(self.segments)])

    def impedance_matrix(self, k0: float) -> np.ndarray:
        """Harrington point-matching Z (ohms) of the straight wire."""
        N, D, a = self.segments, self.delta, self.radius
        omega = k0 * C0
        z = self.points()[:, 2]
        xg, wg = _gauss01(16)

        def psi(zm, zc):
            # (1/D) int_{zc-D/2}^{zc+D/2} exp(-jkR)/(4 pi R) dz', R = sqrt((zm-z')^2 + a^2)
            zp = zc - 0.5 * D + D * xg
            R = np.sqrt((zm - zp) ** 2 + a * a)
            val = np.sum(wg * np.exp(-1j * k0 * R) / (4 * math.pi * R))
            if abs(zm - zc) < 1e-12 * D:
                # self term of the reduced kernel (Harrington eq. 4-44)
                val = math.log(D / a) / (2 * math.pi * D) - 1j * k0 / (4 * math.pi)
            return val

        Z = np.zeros((N, N), dtype=np.complex128)
        for m in range(N):
            for n in range(N):
                t1 = 1j * omega * MU0 * D * D * psi(z[m], z[n])
                t2 = (psi(z[m] + D / 2, z[n] + D / 2) - psi(z[m] - D / 2, z[n] + D / 2)
                      - psi(z[m] + D / 2, z[n] - D / 2) + psi(z[m] - D / 2, z[n] - D / 2)) / (1j * omega * EPS0)
                Z[m, n] = t1 + t2
        return Z

    def feed_index(self):
        return self.segments // 2

    def rhs(self, volt=1.0):
        v = np.zeros(self.segments, dtype=np.complex128)
        v[self.feed_index()] = volt
        return v


def dyadic(rvec, k0):
    """(k^2 I + grad grad) exp(-jkR)/(4 pi R) for an array of separations [..., 3]."""
    R = np.linalg.norm(rvec, axis=-1)
    g = np.exp(-1j * k0 * R) / (4 * math.pi * R)
    diag = g * (k0 * k0 - 1j * k0 / R - 1 / R ** 2)
    rad = g * (-k0 * k0 + 3j * k0 / R + 3 / R ** 2) / R ** 2
    G = rad[..., None, None] * rvec[..., :, None] * rvec[..., None, :]
    G[..., 0, 0] += diag
    G[..., 1, 1] += diag
    G[..., 2, 2] += diag
    return G


@dataclass
class WireArray:
    """Parallel z-directed thin wires (a Yagi-Uda array, say), each cut into
    the same number of segments, discretised like ``Dipole`` (Harrington's
    pulse-basis point matching); between wires the reduced kernel uses the
    axis separation in place of the radius.  ``wires`` holds (x, y, z_centre,
    length) per wire; the feed is the centre segment of wire ``feed_wire``."""
    wires: tuple
    segments_per_wire: int = 21
    radius: float = 1e-3
    feed_wire: int = 1

    @property
    def segments(self):
        return len(self.wires) * self.segments_per_wire

    def deltas(self):
        return np.repeat([w[3] / self.segments_per_wire for w in self.wires], self.segments_per_wire)

    def points(self):
        out = []
        for x, y, zc, length in self.wires:
            d = length / self.segments_per_wire
            z0 = zc - 0.5 * length + 0.5 * d
            out += [[x, y, z0 + i * d] for i in range(self.segments_per_wire)]
        return np.array(out)

    def impedance_matrix(self, k0: float) -> np.ndarray:
        pts, dl = self.points(), self.deltas()
        N, a, omega = len(pts), self.radius, k0 * C0
        xg, wg = _gauss01(16)

        def psi(pm, zm, n):
            # segment-averaged exp(-jkR)/(4 pi R) over segment n seen from (pm_xy, zm)
            dn, zc = dl[n], pts[n, 2]
            rho = np.hypot(pm[0] - pts[n, 0], pm[1] - pts[n, 1])
            same = rho < 1e-12
            r0 = a if same else rho
            if same and abs(zm - zc) < 1e-12 * dn:
                return math.log(dn / a) / (2 * math.pi * dn) - 1j * k0 / (4 * math.pi)
            zp = zc - 0.5 * dn + dn * xg
            R = np.sqrt((zm - zp) ** 2 + r0 * r0)
            return np.sum(wg * np.exp(-1j * k0 * R) / (4 * math.pi * R))

        Z = np.zeros((N, N), dtype=np.complex128)
        for m in range(N):
            pm, zm, dm = pts[m], pts[m, 2], dl[m]
            for n in range(N):
                dn = dl[n]
                t1 = 1j * omega * MU0 * dm * dn * psi(pm, zm, n)
                # the charge terms: psi between the ends of segment m and the
                # half-segment-shifted segment n (the Dipole recipe with a shift)
                def psi_s(zs, sn):
                    zc = pts[n, 2] + sn * 0.5 * dn
                    rho = np.hypot(pm[0] - pts[n, 0], pm[1] - pts[n, 1])
                    same = rho < 1e-12
                    if same and abs(zs - zc) < 1e-12 * dn:
                        return math.log(dn / a) / (2 * math.pi * dn) - 1j * k0 / (4 * math.pi)
                    r0 = a if same else rho
                    zp = zc - 0.5 * dn + dn * xg
                    R = np.sqrt((zs - zp) ** 2 + r0 * r0)
                    return np.sum(wg * np.exp(-1j * k0 * R) / (4 * math.pi * R))
                t2 = (psi_s(zm + dm / 2, 1) - psi_s(zm - dm / 2, 1) - psi_s(zm + dm / 2, -1)
                      + psi_s(zm - dm / 2, -1)) / (1j * omega * EPS0)
                Z[m, n] = t1 + t2
        return Z

    def feed_index(self):
        return self.feed_wire * self.segments_per_wire + self.segments_per_wire // 2

    def rhs(self, volt=1.0):
        v = np.zeros(self.segments, dtype=np.complex128)
        v[self.feed_index()] = volt
        return v


def coupling_blocks(wire, centers: np.ndarray, pitch: float, eps: np.ndarray, k0: float, threads: int = 1):
    """B (c x 3N) and D (3N x c) in the reference's cell-major layout for a
    ``Dipole`` or a ``WireArray`` (segment lengths from ``deltas()``)."""
    pts = wire.points()
    c = len(pts)
    N = len(centers)
    omega = k0 * C0
    dl = wire.deltas() if hasattr(wire, "deltas") else np.full(c, wire.delta)
    B = np.empty((c, 3 * N), dtype=np.complex128)
    D = np.empty((3 * N, c), dtype=np.complex128)
    contrast = (np.broadcast_to(eps, (N,)) - 1.0).reshape(N)

    def one(n):
        G = dyadic(centers - pts[n][None, :], k0)            # [N, 3, 3]
        inc = (1j * omega * MU0 * dl[n] / (k0 * k0)) * G[:, :, 2]   # field per amp, [N, 3]
        D[:, n] = (contrast[:, None] * inc).reshape(-1)
        B[n, :] = (-(pitch ** 3) * dl[n] * G[:, 2, :]).reshape(-1)

    if threads > 1:
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(one, range(c)))
    else:
        for n in range(c):
            one(n)
    return B, D


SLAB_PITCH = 0.01      # m: 1.28 m edge at 128^3
SLAB_FREQ = 200e6
EPS0_REF = 8.8541878128e-12   # the reference's eps0 in dielectric_material_epsilon (dielectric.h:55)


SLAB_GROUND_HEIGHT = 0.1     # m: the slab's bottom face above the real ground
SLAB_GROUND = (13.0, 0.005)  # eps_r and sigma (S/m) of the ground


def nec_ground_epsilon(epsr, sigma, frequency_mhz):
    """NEC's complex ground permittivity for a GN card (dielectric_ground.c:97-100:
    epsr - i sigma (CVEL / f) 59.96 with NEC's CVEL = 299.8, or the given
    imaginary part when sigma is negative)."""
    return complex(epsr, sigma if sigma < 0 else -sigma * (299.8 / frequency_mhz) * 59.96)


def yagi_slab(n, ground=False):
    """The slab workload: an n^3 block of 1 cm voxels (eps_r 6, 0.01 S/m)
    driven by a 4-element Yagi (reflector 0.50 m, driven 0.47 m, directors
    0.44 m and 0.43 m, 21 segments each, 1 mm radius, 0.2 m spacing) 0.1 m
    from its x = 0 face, at 200 MHz.  Returns (yagi, k0, eps_cell) with the
    cell permittivity eps_r - i sigma / (omega eps0) evaluated as the
    reference evaluates a conductive material.  With ground=True the block
    sits 0.1 m above the real ground (its cells from z = 0.1 m up: use
    slab_origin) and the Yagi at the block's mid-height."""
    edge = n * SLAB_PITCH
    yc = zc = 0.5 * edge
    if ground:
        zc += SLAB_GROUND_HEIGHT
    yagi = WireArray(tuple((-0.1, yc + (i - 1.5) * 0.2, zc, L) for i, L in enumerate((0.50, 0.47, 0.44, 0.43))),
                     segments_per_wire=21, radius=1e-3, feed_wire=1)
    k0 = k0_of(SLAB_FREQ)
    eps_cell = 6.0 - 1j * 0.01 / (k0 * C0 * EPS0_REF)
    return yagi, k0, eps_cell


def slab_origin(ground=False):
    return (0.0, 0.0, SLAB_GROUND_HEIGHT if ground else 0.0)


def cube_cells(n, pitch, origin=(0.0, 0.0, 0.0)):
    """Centres of an n^3 (or nx x ny x nz) block, x slowest, z fastest."""
    shape = (n, n, n) if np.isscalar(n) else tuple(n)
    g = np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).reshape(-1, 3)
    return np.asarray(origin)[None, :] + pitch * (g + 0.5)


@dataclass
class BatchCase:
    """One solve of the batched workload: a homogeneous block, a dipole and
    its feed ports (one right-hand side per port)."""
    shape: tuple
    eps_r: float
    freq: float
    pitch: float
    dipole: Dipole
    ports: tuple        # feed segment of each right-hand side

    @property
    def k0(self):
        return k0_of(self.freq)

    def centers(self):
        return cube_cells(self.shape, self.pitch)

    def rhs(self, port):
        v = np.zeros(self.dipole.segments, dtype=np.complex128)
        v[self.ports[port]] = 1.0
        return v


def batch_cases(count, seed, lo=8, hi=48, max_ports=8, seg_lo=11, seg_hi=31, f_lo=50e6, f_hi=500e6,
                eps_lo=2.0, eps_hi=12.0, kh=0.3):
    """Draws of the batched workload: per case the grid extent per axis is
    uniform on [lo, hi] (a cube or a cuboid with equal odds), eps_r uniform on
    [eps_lo, eps_hi], the frequency uniform on [f_lo, f_hi], the pitch chosen
    so k0 sqrt(eps_r) h = kh (kh <= 0.5), and a z-directed dipole of 0.47
    free-space wavelengths with seg_lo..seg_hi segments, a tenth of a
    wavelength (at least two cells) from the block's x = 0 face, fed at
    1..max_ports distinct segments."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        if rng.random() < 0.5:
            n = int(rng.integers(lo, hi + 1))
            shape = (n, n, n)
        else:
            shape = tuple(int(v) for v in rng.integers(lo, hi + 1, 3))
        er = float(rng.uniform(eps_lo, eps_hi))
        f = float(rng.uniform(f_lo, f_hi))
        k0 = k0_of(f)
        h = kh / (k0 * math.sqrt(er))
        lam = 2 * math.pi / k0
        segs = int(rng.integers(seg_lo, seg_hi + 1))
        gap = max(0.1 * lam, 2 * h)
        dip = Dipole(center=(-gap, 0.5 * shape[1] * h, 0.5 * shape[2] * h), length=0.47 * lam, segments=segs,
                     radius=lam / 1000)
        p = int(rng.integers(1, max_ports + 1))
        ports = tuple(int(v) for v in rng.choice(segs, size=min(p, segs), replace=False))
        out.append(BatchCase(shape, er, f, h, dip, ports))
    return out


def nec_straight_wire(start, end, segments, radius, wlam):
    """NEC's arrays for a straight wire cut into equal segments (a GW card):
    (x, y, z, si, bi, cab, sab, salp), lengths in wavelengths of wlam metres."""
    start, end = np.asarray(start, float), np.asarray(end, float)
    d = (end - start) / segments
    centers = start + (np.arange(segments)[:, None] + 0.5) * d
    length = float(np.linalg.norm(d))
    t = d / length
    full = lambda v: np.full(segments, v)
    return (centers[:, 0] / wlam, centers[:, 1] / wlam, centers[:, 2] / wlam, full(length / wlam),
            full(radius / wlam), full(t[0]), full(t[1]), full(t[2]))
