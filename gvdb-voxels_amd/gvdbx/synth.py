"""Seeded synthetic inputs of the five benchmark workloads (FP32 positions, numpy PCG64).

The reference ships no point fixture (its points.dat is absent), so every workload is
generated from its description: sizes, distributions and seeds are fixed here.
"""
import numpy as np


def sphere_shell(n=262_144, seed=0, center=(128.0, 128.0, 128.0), radius=100.0, sigma=1.5):
    """Points on a sphere of radius 100 voxels about (128,128,128), Gaussian radial jitter."""
    g = np.random.default_rng(seed)
    d = g.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rho = radius + sigma * g.standard_normal(n)
    return (np.asarray(center) + rho[:, None] * d).astype(np.float32)


def sph_block(n=4_194_304, seed=1, box=256.0, fill=0.4):
    """Jittered lattice filling the lower `fill` of a box^3 domain; velocities ~ N(0,1)."""
    g = np.random.default_rng(seed)
    h = (box * box * box * fill / n) ** (1.0 / 3.0)
    nx = int(np.ceil(box / h))
    ny = int(np.ceil(box * fill / h))
    nz = nx
    while nx * ny * nz < n:
        ny += 1
    idx = np.arange(n, dtype=np.int64)
    ix = idx // (ny * nz)
    iy = (idx // nz) % ny
    iz = idx % nz
    lat = (np.stack([ix, iy, iz], axis=1).astype(np.float64) + 0.5) * h
    pos = lat + g.uniform(-0.5, 0.5, size=(n, 3))
    vel = g.standard_normal((n, 3))
    return pos.astype(np.float32), vel.astype(np.float32)


def sphere_surfaces(n=16_777_216, seed=3, count=64, rmin=8.0, rmax=40.0, domain=1024.0):
    """Points on the surfaces of random spheres, allocated in proportion to sphere area."""
    g = np.random.default_rng(seed)
    radii = g.uniform(rmin, rmax, count)
    centers = g.uniform(0.0, domain, (count, 3))
    area = radii ** 2
    alloc = np.floor(n * area / area.sum()).astype(np.int64)
    alloc[: n - alloc.sum()] += 1
    sid = np.repeat(np.arange(count), alloc)
    d = g.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos = centers[sid] + radii[sid, None] * d
    return pos.astype(np.float32)


def uniform_queries(n=1_048_576, seed=3, domain=1024.0):
    g = np.random.default_rng(seed + 7919)
    return g.uniform(0.0, domain, (n, 3)).astype(np.float32)


def uniform_massive(n=2_097_152, seed=4, box=256.0):
    """Uniform points in box^3 with masses U(0.5, 1.5)."""
    g = np.random.default_rng(seed)
    pos = g.uniform(0.0, box, (n, 3)).astype(np.float32)
    mass = g.uniform(0.5, 1.5, n).astype(np.float32)
    return pos, mass


def blob_clouds(count=256, seed=5, nmin=4096, nmax=1_048_576, blobs=8, smin=4.0, smax=32.0, cmin=96.0, cmax=416.0):
    """Independent clouds with log-uniform sizes, each a mixture of Gaussian blobs."""
    g = np.random.default_rng(seed)
    sizes = np.round(np.exp(g.uniform(np.log(nmin), np.log(nmax), count))).astype(np.int64)
    clouds = []
    for n in sizes:
        cen = g.uniform(cmin, cmax, (blobs, 3))
        sig = g.uniform(smin, smax, blobs)
        which = g.integers(0, blobs, n)
        p = cen[which] + sig[which, None] * g.standard_normal((n, 3))
        clouds.append(p.astype(np.float32))
    return clouds


# ---------------------------------------------------------------- shared with C++
# The same bits as gvdbx_synth_uniform / _sphere / _lattice of the library (csrc/synth.cpp): draw
# k of stream `seed` is the SplitMix64 finalizer of seed * 0xD1B54A32D192ED03 + (k + 1) *
# 0x9E3779B97F4A7C15 (wrapping uint64), a uniform its top 53 bits times 2^-53, a normal the
# in-order sum of 12 uniforms minus 6.  Elementwise float64 operations in the C order, then FP32.

_M1, _M2 = np.uint64(0xBF58476D1CE4E5B9), np.uint64(0x94D049BB133111EB)
_SEEDMUL, _GOLD = np.uint64(0xD1B54A32D192ED03), np.uint64(0x9E3779B97F4A7C15)


def _draw(seed, k):
    with np.errstate(over="ignore"):
        z = np.uint64(seed) * _SEEDMUL + (np.asarray(k, np.uint64) + np.uint64(1)) * _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def _uniform(seed, k):
    return (_draw(seed, k) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def _normal(seed, k):
    s = _uniform(seed, k)
    for j in range(1, 12):
        s = s + _uniform(seed, k + np.uint64(j))
    return s - 6.0


def shared_uniform(n, seed, lo, hi):
    """gvdbx_synth_uniform: (n, 3) FP32 points uniform in the box [lo, hi)."""
    k = np.arange(3 * n, dtype=np.uint64).reshape(n, 3)
    u = _uniform(seed, k)
    lo = np.asarray(lo, np.float32).astype(np.float64)
    hi = np.asarray(hi, np.float32).astype(np.float64)
    return (lo + (hi - lo) * u).astype(np.float32)


def shared_sphere(n, seed, center, radius, sigma):
    """gvdbx_synth_sphere: (n, 3) FP32 points on a sphere with normal radial jitter."""
    k = np.arange(n, dtype=np.uint64) * np.uint64(48)
    d0, d1, d2 = _normal(seed, k), _normal(seed, k + np.uint64(12)), _normal(seed, k + np.uint64(24))
    ln = np.sqrt((d0 * d0 + d1 * d1) + d2 * d2)
    rho = np.float64(np.float32(radius)) + np.float64(np.float32(sigma)) * _normal(seed, k + np.uint64(36))
    s = np.where(ln > 0.0, rho / np.where(ln > 0.0, ln, 1.0), 0.0)
    c = np.asarray(center, np.float32).astype(np.float64)
    return np.stack([c[0] + d0 * s, c[1] + d1 * s, c[2] + d2 * s], axis=1).astype(np.float32)


def shared_lattice(dims, seed, spacing, jitter, vel_sigma):
    """gvdbx_synth_lattice: the jittered lattice of dims[0] x dims[1] x dims[2] points and their
    normal velocities, both (n, 3) FP32."""
    nx, ny, nz = (int(d) for d in dims)
    n = nx * ny * nz
    i = np.arange(n, dtype=np.int64)
    c = np.stack([i // (ny * nz), (i // nz) % ny, i % nz], axis=1).astype(np.float64)
    k = i.astype(np.uint64) * np.uint64(39)
    u = np.stack([_uniform(seed, k + np.uint64(a)) for a in range(3)], axis=1)
    h, jit, vs = (np.float64(np.float32(x)) for x in (spacing, jitter, vel_sigma))
    pos = ((c + 0.5) * h + (u - 0.5) * jit).astype(np.float32)
    vel = np.stack([vs * _normal(seed, k + np.uint64(3 + 12 * a)) for a in range(3)], axis=1).astype(np.float32)
    return pos, vel
