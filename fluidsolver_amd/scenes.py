"""Seeded synthetic scenes for the benchmark configurations.

All randomness comes from SplitMix64 (Steele, Lea, Flood 2014): output k of
stream `seed` is mix(seed + (k + 1) * 0x9E3779B97F4A7C15) with the standard
finaliser; a uniform float in [0, 1) takes the top 24 bits.  The same stream is
simple to restate in C++ or HIP, so every implementation sees bit-identical
inputs.

Geometry conventions (lattice spacing = particle_size = 1):
  * a fluid block of nx x ny x nz particles sits at integer coordinates
    0..n-1 in the corner of the box interior, each coordinate jittered by
    U(-jitter, +jitter);
  * the box interior holds lattice points 0..L-1 per axis, and the boundary is
    the `layers`-thick closed shell of lattice points around it, so a
    300 x 150 x 150 interior with 3 layers has 306*156*156 - 300*150*150 =
    696,816 boundary particles and a 200 x 100 2D interior has 1,836.
"""
import math

import numpy as np

from .abi import TYPE_BOUNDARY, TYPE_NORMAL
from .collection import ParticleCollection

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, count, offset=0):
    """Outputs offset..offset+count-1 of the SplitMix64 stream `seed` (uint64)."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform24(seed, count, offset=0):
    """float32 in [0, 1): top 24 bits of SplitMix64 outputs, scaled by 2^-24."""
    z = splitmix64(seed, count, offset)
    return ((z >> np.uint64(40)).astype(np.float32) * np.float32(1.0 / 16777216.0)).astype(np.float32)


def jitter(seed, count, amplitude, offset=0):
    """U(-amplitude, amplitude) in float32: (2u - 1) * amplitude."""
    u = uniform24(seed, count, offset)
    return ((u * np.float32(2.0) - np.float32(1.0)) * np.float32(amplitude)).astype(np.float32)


def lattice(dims, origin=None):
    """Integer lattice points (float32) of a box of `dims`, x fastest."""
    axes = [np.arange(d, dtype=np.float32) for d in dims]
    grids = np.meshgrid(*axes[::-1], indexing="ij")[::-1]
    pts = np.stack([g.reshape(-1) for g in grids], axis=1).astype(np.float32)
    if origin is not None:
        pts += np.asarray(origin, dtype=np.float32)
    return pts


def box_shell(interior, layers=3):
    """Lattice points of a closed `layers`-thick shell around interior lattice
    points 0..L-1 per axis."""
    dim = len(interior)
    outer = [L + 2 * layers for L in interior]
    pts = lattice(outer, origin=[-layers] * dim)
    inside = np.ones(pts.shape[0], dtype=bool)
    for a in range(dim):
        inside &= (pts[:, a] >= 0) & (pts[:, a] <= interior[a] - 1)
    return pts[~inside]


def fluid_block(dims, seed, amplitude=1e-3, origin=None):
    pts = lattice(dims, origin)
    j = jitter(seed, pts.size, amplitude).reshape(pts.shape)
    return (pts + j).astype(np.float32)


def dam_scene(fluid_dims, interior, layers=3, seed=1, amplitude=1e-3):
    """Breaking dam: fluid block in the corner of a closed box (fluid first,
    then boundary)."""
    dim = len(fluid_dims)
    f = fluid_block(fluid_dims, seed, amplitude)
    b = box_shell(interior, layers)
    pos = np.concatenate([f, b]).astype(np.float32)
    types = np.concatenate(
        [np.full(f.shape[0], TYPE_NORMAL, np.uint8), np.full(b.shape[0], TYPE_BOUNDARY, np.uint8)]
    )
    return ParticleCollection.from_arrays(dim, pos, types)


# ---- benchmark configurations -------------------------------------------------
def config1_dam2d(seed=1):
    """2D IISPH breaking dam: 64x64 fluid in a 200x100 box, 3-layer shell."""
    return dam_scene((64, 64), (200, 100), 3, seed)


def config2_dam3d(seed=2, side=100, interior=(300, 150, 150)):
    """3D IISPH breaking dam: side^3 fluid in the corner of a 300x150x150 box."""
    return dam_scene((side, side, side), interior, 3, seed)


def scaled_dam3d(n_fluid, seed, amplitude=1e-3):
    """Cube-ish fluid block of about n_fluid particles in a box with the
    proportions of configuration 2 (3a x 1.5a x 1.5a for block side a)."""
    a = max(2, int(round(n_fluid ** (1.0 / 3.0))))
    nx = a
    ny = a
    nz = max(1, int(round(n_fluid / float(a * a))))
    interior = (3 * a, max(ny + 2, int(round(1.5 * a))), max(nz + 2, int(round(1.5 * a))))
    return dam_scene((nx, ny, nz), interior, 3, seed, amplitude)


def loguniform_counts(seed, count, lo, hi):
    u = uniform24(seed, count).astype(np.float64)
    return np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo))).astype(np.int64)


def config4_sweep(seed=4, n_scenes=64, lo=5000, hi=200000):
    """Batched parameter sweep: scene sizes log-uniform in [lo, hi], per-scene
    gamma U[1,7] and viscosity U[0.001, 0.05]."""
    counts = loguniform_counts(seed, n_scenes, lo, hi)
    ug = uniform24(seed + 1000, n_scenes).astype(np.float64)
    uv = uniform24(seed + 2000, n_scenes).astype(np.float64)
    scenes = []
    for s in range(n_scenes):
        scenes.append(
            {
                "collection": scaled_dam3d(int(counts[s]), seed * 1000 + s),
                "gamma": float(1.0 + 6.0 * ug[s]),
                "viscosity": float(0.001 + 0.049 * uv[s]),
            }
        )
    return scenes


def config5_arrivals(seed=5, duration_s=600.0, rate=20.0, lo=10000, hi=500000, steps_lo=100, steps_hi=2000):
    """Poisson arrival stream: (arrival time, fluid count, steps) per job."""
    n_max = int(rate * duration_s * 2 + 100)
    u = uniform24(seed, 3 * n_max).astype(np.float64).reshape(n_max, 3)
    t = np.cumsum(-np.log1p(-u[:, 0]) / rate)
    keep = t < duration_s
    sizes = np.exp(np.log(lo) + u[:, 1] * (np.log(hi) - np.log(lo))).astype(np.int64)
    steps = (steps_lo + np.floor(u[:, 2] * (steps_hi - steps_lo + 1))).astype(np.int64)
    return [(float(a), int(b), int(c)) for a, b, c in zip(t[keep], sizes[keep], steps[keep])]


def config3_obstacles(seed=3, side=200, n_spheres=80, n_tori=40, interior=None):
    """Fluid block plus seeded sphere and torus obstacles surface-sampled at
    spacing ~1 (sphere radius U[5,20]); boundary typed 1.  The box keeps the
    proportions of configuration 2 (3a x 1.5a x 1.5a for block side a): at
    side 200 that is 8,000,000 fluid and 2,991,211 boundary particles (the
    3-layer shell of a 600 x 300 x 300 box plus the obstacles)."""
    if interior is None:
        interior = (3 * side, 3 * side // 2, 3 * side // 2)
    base = dam_scene((side, side, side), interior, 3, seed)
    u = uniform24(seed + 77, 8 * (n_spheres + n_tori)).astype(np.float64).reshape(-1, 8)
    pts = []
    for k in range(n_spheres):
        r = 5.0 + 15.0 * u[k, 0]
        c = np.array([side + 20 + u[k, 1] * (interior[0] - side - 40), r + u[k, 2] * (interior[1] - 2 * r),
                      r + u[k, 3] * (interior[2] - 2 * r)])
        m = max(8, int(4 * math.pi * r * r))
        i = np.arange(m, dtype=np.float64) + 0.5
        phi = np.arccos(1 - 2 * i / m)
        th = math.pi * (1 + 5 ** 0.5) * i
        pts.append(np.stack([c[0] + r * np.cos(th) * np.sin(phi), c[1] + r * np.cos(phi),
                             c[2] + r * np.sin(th) * np.sin(phi)], axis=1))
    for k in range(n_tori):
        q = u[n_spheres + k]
        R = 8.0 + 10.0 * q[0]
        rr = 2.0 + 3.0 * q[1]
        c = np.array([side + 20 + q[2] * (interior[0] - side - 40), R + rr + q[3] * (interior[1] - 2 * (R + rr)),
                      R + rr + q[4] * (interior[2] - 2 * (R + rr))])
        nu = max(8, int(2 * math.pi * R))
        nv = max(6, int(2 * math.pi * rr))
        a, b = np.meshgrid(np.arange(nu) * 2 * math.pi / nu, np.arange(nv) * 2 * math.pi / nv, indexing="ij")
        a = a.reshape(-1)
        b = b.reshape(-1)
        pts.append(np.stack([c[0] + (R + rr * np.cos(b)) * np.cos(a), c[1] + rr * np.sin(b),
                             c[2] + (R + rr * np.cos(b)) * np.sin(a)], axis=1))
    obs = np.concatenate(pts).astype(np.float32)
    oc = ParticleCollection.from_arrays(3, obs, np.full(obs.shape[0], TYPE_BOUNDARY, np.uint8))
    return ParticleCollection.concatenate([base, oc])
