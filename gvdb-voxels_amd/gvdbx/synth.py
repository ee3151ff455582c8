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
