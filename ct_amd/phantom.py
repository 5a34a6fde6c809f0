"""Synthetic scans: seeded ellipsoid phantoms in the reference's geometry model.

Ct reads its projections from disk (Ct/CtVolume.cpp:136-330) and ships no data, so
the scans used for parity and measurement are generated here.  The random stream
is xoshiro256** seeded through splitmix64, uniforms are (x >> 11) * 2^-53, so a
seed gives the same phantom on every machine.  Ellipsoids are rows
{cx, cy, cz, a, b, c, phi, theta, psi, mu} in voxel units with ZYZ Euler angles;
line integrals and uint16 intensities are computed on the GPU by
``ctg_phantom_images``.
"""
from __future__ import annotations

import math

_M64 = (1 << 64) - 1


def _rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & _M64


class Xoshiro256ss:
    """xoshiro256** (Blackman & Vigna), state seeded by splitmix64."""

    def __init__(self, seed: int):
        s = seed & _M64
        self.s = []
        for _ in range(4):
            s = (s + 0x9E3779B97F4A7C15) & _M64
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
            self.s.append(z ^ (z >> 31))

    def next(self) -> int:
        s = self.s
        result = (_rotl((s[1] * 5) & _M64, 7) * 9) & _M64
        t = (s[1] << 17) & _M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = _rotl(s[3], 45)
        return result

    def uniform(self, lo=0.0, hi=1.0) -> float:
        return lo + (hi - lo) * ((self.next() >> 11) * 2.0 ** -53)


def phantom_unit_radius(radius: float, height: int, fcd: float) -> float:
    """Radius (voxels) of the unit sphere that keeps every object inside the fully
    sampled region: 0.9 * min(R, (H/2 - 1.1) * (FCD - R) / FCD)."""
    return 0.9 * min(radius, (height / 2.0 - 1.1) * (fcd - radius) / fcd)


def random_ellipsoids(n: int, seed: int, unit: float, mu_range=(0.1, 1.0), axes=(0.05, 0.4)):
    rng = Xoshiro256ss(seed)
    out = []
    for _ in range(n):
        c = [rng.uniform(-0.5, 0.5) * unit for _ in range(3)]
        ax = [rng.uniform(*axes) * unit for _ in range(3)]
        ang = [rng.uniform(0, 2 * math.pi), rng.uniform(0, math.pi), rng.uniform(0, 2 * math.pi)]
        mu = rng.uniform(*mu_range)
        out.append(c + ax + ang + [mu])
    return out


# 3-D Shepp-Logan (Kak & Slaney, "Principles of Computerized Tomographic Imaging",
# table 3.1 extended to 3-D): centre, semi-axes, rotation about z (degrees), density.
_SHEPP_LOGAN_3D = [
    ((0.0, 0.0, 0.0), (0.69, 0.92, 0.90), 0.0, 2.0),
    ((0.0, 0.0, 0.0), (0.6624, 0.874, 0.88), 0.0, -0.98),
    ((-0.22, 0.0, -0.25), (0.41, 0.16, 0.21), 108.0, -0.02),
    ((0.22, 0.0, -0.25), (0.31, 0.11, 0.22), 72.0, -0.02),
    ((0.0, 0.35, -0.25), (0.21, 0.25, 0.50), 0.0, 0.02),
    ((0.0, 0.1, -0.25), (0.046, 0.046, 0.046), 0.0, 0.02),
    ((-0.08, -0.65, -0.25), (0.046, 0.023, 0.02), 0.0, 0.01),
    ((0.06, -0.65, -0.25), (0.046, 0.023, 0.02), 90.0, 0.01),
    ((0.06, -0.105, 0.625), (0.056, 0.04, 0.1), 90.0, 0.02),
    ((0.0, 0.1, 0.625), (0.056, 0.056, 0.1), 0.0, -0.02),
]


def shepp_logan_3d(unit: float, scale: float = 1.0):
    out = []
    for (cx, cy, cz), (a, b, c), rot, rho in _SHEPP_LOGAN_3D:
        out.append([cx * unit, cy * unit, cz * unit, a * unit, b * unit, c * unit,
                    math.radians(rot), 0.0, 0.0, rho * scale])
    return out


def circular_angles(nproj: int):
    """Equiangular full scan, 360*i/P degrees (ccw, as Ct stores them)."""
    return [360.0 * i / nproj for i in range(nproj)]


def _rotation(phi, theta, psi):
    """R = Rz(phi) Ry(theta) Rz(psi); its columns are the ellipsoid's axes."""
    c1, s1, c2, s2, c3, s3 = (math.cos(phi), math.sin(phi), math.cos(theta), math.sin(theta),
                              math.cos(psi), math.sin(psi))
    return [[c1 * c2 * c3 - s1 * s3, -c1 * c2 * s3 - s1 * c3, c1 * s2],
            [s1 * c2 * c3 + c1 * s3, -s1 * c2 * s3 + c1 * c3, s1 * s2],
            [-s2 * c3, s2 * s3, c2]]


def needles(n: int, seed: int, unit: float, mu: float, radius: float = 2.0, half_length: float = 0.6):
    """Thin high-contrast rods: ellipsoids with two semi-axes of `radius` voxels and a
    long semi-axis of half_length * unit, at random centres and orientations."""
    rng = Xoshiro256ss(seed)
    out = []
    for _ in range(n):
        c = [rng.uniform(-0.4, 0.4) * unit for _ in range(3)]
        ang = [rng.uniform(0, 2 * math.pi), rng.uniform(0, math.pi), rng.uniform(0, 2 * math.pi)]
        out.append(c + [radius, radius, half_length * unit] + ang + [mu])
    return out


def voxelize(ellipsoids, geometry, device, dtype=None):
    """Attenuation of the ellipsoid phantom sampled at the voxels of `geometry`
    (voxel (i, j, k) at world (i - x_pre, j - y_pre, k - z_pre), the backprojection's
    model), as an (nx, ny, nz) tensor on `device`.  Used to score reconstructions."""
    import torch
    dtype = dtype or torch.float32
    g = geometry
    out = torch.zeros((g.nx, g.ny, g.nz), dtype=dtype, device=device)
    pre = (g.x_pre, g.y_pre, g.z_pre)
    size = (g.nx, g.ny, g.nz)
    for e in ellipsoids:
        c, ax, (phi, theta, psi), mu = e[0:3], e[3:6], e[6:9], e[9]
        R = _rotation(phi, theta, psi)
        reach = max(ax) + 1.0
        lo = [max(0, int(math.floor(c[d] + pre[d] - reach))) for d in range(3)]
        hi = [min(size[d], int(math.ceil(c[d] + pre[d] + reach)) + 1) for d in range(3)]
        if any(h <= l for l, h in zip(lo, hi)):
            continue
        coords = [torch.arange(lo[d], hi[d], device=device, dtype=torch.float64) - pre[d] - c[d] for d in range(3)]
        X, Y, Z = torch.meshgrid(*coords, indexing="ij")
        q = 0.0
        for a in range(3):  # body coordinate a = (R^T (p - c))_a / ax_a
            qa = (R[0][a] * X + R[1][a] * Y + R[2][a] * Z) / ax[a]
            q = q + qa * qa
        out[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] += (q < 1.0).to(dtype) * mu
    return out
