"""The C++ VolumeGVDB facade (gvdbx/volume_gvdb.hpp) driven by a compiled caller
(examples/points_to_grid.cpp), checked against the oracle: the reference's per-frame call
sequence from C++ gives the same leaf pool and the same channel as the oracle."""
import os
import subprocess

import numpy as np
import pytest

import gpu_util as U
from gvdbx import NODE_DTYPE
from gvdbx.lib import PKG_DIR

EXE = os.path.join(PKG_DIR, "lib", "points_to_grid")


@pytest.mark.gpu
@pytest.mark.parametrize("levelset", [0, 1])
def test_cpp_facade_points_to_grid(oracle_mod, tmp_path, levelset):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built")
    U.log(f"case c++ facade levelset={levelset}")
    p = np.random.default_rng(40 + levelset).uniform(30, 120, (5000, 3)).astype(np.float32)
    radius = 1.5
    fin, fout = tmp_path / "p.f32", tmp_path / "out.bin"
    p.tofile(fin)
    r = subprocess.run([EXE, str(fin), str(len(p)), str(radius), "4", str(levelset), str(fout)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = fout.read_bytes()
    total, slots, br = np.frombuffer(raw[:24], np.int64)
    nodes = np.frombuffer(raw[24:24 + 64 * total], NODE_DTYPE)
    chan = np.frombuffer(raw[24 + 64 * total:], np.float32).reshape(slots, br, br, br)

    t = U.new_tree(oracle_mod)
    t.rebuild_topology(p, radius)
    t.update_atlas()
    t.insert_points_subcell(4, p, radius)
    if levelset:
        t.gather_levelset(radius, 0)
    else:
        t.gather_density(radius, 0)
    t.update_apron(0, 3.0 if levelset else 0.0)
    assert nodes.tobytes() == t.nodes(0).tobytes()
    want = t.channel(0)[:slots]
    err = np.abs(chan.astype(np.float64) - want) / np.maximum(np.abs(want), 1.0)
    assert err.max() <= U.TOL, err.max()
    assert f"leaves {total}" in r.stdout


@pytest.mark.gpu
def test_cpp_facade_shared_generator_golden():
    """The C++ caller generates the golden sphere itself (gvdbx_synth_sphere, the same bits as
    gvdbx/synth.py) and must reproduce the committed fixture: the leaf pool byte for byte and the
    density within 1e-5 of the stored field."""
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built")
    U.log("case c++ facade, shared generator, golden sphere")
    with np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sphere.npz")) as z:
        want_nodes, want = z["nodes0"], z["density"]
        n = len(z["points"])
    fout = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"gvdbx_golden_{os.getpid()}.bin")
    r = subprocess.run([EXE, "synth:0", str(n), "1.0", "4", "0", fout], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    with open(fout, "rb") as f:
        raw = f.read()
    os.remove(fout)
    total, slots, br = np.frombuffer(raw[:24], np.int64)
    assert raw[24:24 + 64 * total] == want_nodes.tobytes()
    chan = np.frombuffer(raw[24 + 64 * total:], np.float32).reshape(slots, br, br, br)
    err = np.abs(chan.astype(np.float64) - want) / np.maximum(np.abs(want.astype(np.float64)), 1.0)
    assert err.max() <= U.TOL, err.max()
