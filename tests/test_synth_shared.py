"""The seeded point sets of the library (csrc/synth.cpp, C ABI) and of gvdbx/synth.py are the same
bits: sphere shells, jittered lattices with velocities and uniform boxes, for several seeds and
sizes; and the generator's statistics are what its definition gives (Irwin-Hall normals have
mean 0 and variance 1, uniforms fill [lo, hi))."""
import ctypes as C

import numpy as np
import pytest

from gvdbx import synth
from gvdbx.lib import load_library


def fptr(a):
    return a.ctypes.data


@pytest.mark.parametrize("seed,n", [(0, 1), (0, 5000), (7, 777), (2 ** 63 + 5, 1000)])
def test_sphere_same_bits(seed, n):
    L = load_library()
    c = np.array([64.0, 70.5, 33.25], np.float32)
    out = np.zeros((n, 3), np.float32)
    assert L.gvdbx_synth_sphere(C.c_uint64(seed), n, fptr(c), C.c_float(20.0), C.c_float(1.5), fptr(out)) == 0
    want = synth.shared_sphere(n, seed, c, 20.0, 1.5)
    assert out.tobytes() == want.tobytes()


@pytest.mark.parametrize("seed,dims", [(1, (5, 4, 3)), (11, (40, 17, 33))])
def test_lattice_same_bits(seed, dims):
    L = load_library()
    d = np.array(dims, np.int32)
    n = int(np.prod(dims))
    pos, vel = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32)
    assert L.gvdbx_synth_lattice(C.c_uint64(seed), fptr(d), C.c_float(1.17), C.c_float(1.0), C.c_float(1.0),
                                 fptr(pos), fptr(vel)) == 0
    wp, wv = synth.shared_lattice(dims, seed, 1.17, 1.0, 1.0)
    assert pos.tobytes() == wp.tobytes() and vel.tobytes() == wv.tobytes()


def test_uniform_same_bits_and_statistics():
    L = load_library()
    lo, hi = np.array([0.0, -3.0, 100.0], np.float32), np.array([256.0, 3.0, 101.0], np.float32)
    n = 200_000
    out = np.zeros((n, 3), np.float32)
    assert L.gvdbx_synth_uniform(C.c_uint64(3), n, fptr(lo), fptr(hi), fptr(out)) == 0
    assert out.tobytes() == synth.shared_uniform(n, 3, lo, hi).tobytes()
    assert (out >= lo).all() and (out <= hi).all()
    assert np.allclose(out.astype(np.float64).mean(0), (lo + hi) / 2, atol=0.01 * (hi - lo))
    g = synth._normal(9, np.arange(n, dtype=np.uint64) * np.uint64(12))
    assert abs(g.mean()) < 0.01 and abs(g.var() - 1.0) < 0.01 and np.abs(g).max() <= 6.0


def test_bad_arguments():
    L = load_library()
    out = np.zeros(3, np.float32)
    c = np.zeros(3, np.float32)
    assert L.gvdbx_synth_sphere(C.c_uint64(0), -1, fptr(c), C.c_float(1.0), C.c_float(0.0), fptr(out)) != 0
    assert L.gvdbx_synth_lattice(C.c_uint64(0), None, C.c_float(1.0), C.c_float(0.0), C.c_float(0.0), fptr(out),
                                 None) != 0
