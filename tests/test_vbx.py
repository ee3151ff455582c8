"""The VBX file codec (GVDB_FILESPEC.txt; SaveVBX V.cpp:1626-1768, LoadVBX V.cpp:507-685) on the CPU.

An independent writer below builds the file field by field from the specification and the field
order of SaveVBX (version 1.11: transform, one grid, no bitmask byte; grid header; per-level
dims; pool 0 rows of every level, then pool 1 rows; per channel type, stride and the atlas
volume with brick b at (b % ax, b / ax % ay, b / (ax ay))).  The library's writer must produce
the same bytes from the same pools and brick-linear channels, and its reader must return them
unchanged.  Pools come from the oracle's tree of a real cloud.
"""
import ctypes as C
import struct

import numpy as np
import pytest

from gvdbx.lib import VBX_MAXCHAN, VbxDesc, load_library


def spec_file(ld, root, nodes, clists, apron, axis, chan_types, chans, slots, transform):
    out = bytearray()
    out += struct.pack("<BB", 1, 11)
    out += struct.pack("<12f", *transform)
    out += struct.pack("<i", 1)
    table = len(out)
    out += struct.pack("<Q", 0)
    grid = len(out)
    res0 = 1 << ld[0]
    br = res0 + 2 * apron
    axisres = [a * br for a in axis]
    out += b"\0" * 256 + struct.pack("<BBB", ord("f"), 1, 0) + struct.pack("<3f", 1, 1, 1)
    out += struct.pack("<i3ii", len(nodes[0]), res0, res0, res0, apron)
    out += struct.pack("<iQ", len(chan_types), 4 * int(np.prod(axisres)) if chan_types else 0)
    out += struct.pack("<BiB", 2, 0, 0) + struct.pack("<3i", *axis) + struct.pack("<3i", *axisres)
    out += struct.pack("<iQ", len(ld), root)
    rng = 0
    for l, d in enumerate(ld):
        rng += d
        out += struct.pack("<9i", d, 1 << d, 1 << rng, 1 << rng, 1 << rng, len(nodes[l]), 64,
                           len(clists[l]) if l else 0, (8 << (3 * d)) if l else 0)
    for l in range(len(ld)):
        out += nodes[l].tobytes()
    for l in range(1, len(ld)):
        out += clists[l].tobytes()
    for t, data in zip(chan_types, chans):
        out += struct.pack("<ii", t, 4)
        vol = np.zeros((axisres[2], axisres[1], axisres[0]), np.uint32)
        for b in range(slots):
            bx, by, bz = b % axis[0], b // axis[0] % axis[1], b // (axis[0] * axis[1])
            vol[bz * br:(bz + 1) * br, by * br:(by + 1) * br, bx * br:(bx + 1) * br] = data[b]
        out += vol.tobytes()
    struct.pack_into("<Q", out, table, grid)
    return bytes(out)


def oracle_grid(oracle_mod):
    from gvdbx import synth
    p = synth.sphere_shell(n=20_000, seed=3)
    t = oracle_mod.Tree()
    t.rebuild_topology(p, 1.0)
    t.add_channel(0)
    t.update_atlas()
    ld = list(t.log2dims)
    nodes = [t.nodes(l) for l in range(len(ld))]
    clists = [None] + [t.childlists(l) for l in range(1, len(ld))]
    slots = int((nodes[0]["flags"] == 1).sum())
    assert slots == t.num_slots()
    return ld, t.depth() << 8, nodes, clists, slots


def make_desc(ld, root, nodes, clists, apron, axis, chan_types, chans, slots, transform):
    d = VbxDesc()
    d.levels = len(ld)
    for l, x in enumerate(ld):
        d.log2dim[l] = x
        d.node_count[l] = len(nodes[l])
        d.nodes[l] = nodes[l].ctypes.data if len(nodes[l]) else None
        if l:
            d.clist_count[l] = len(clists[l])
            d.clists[l] = clists[l].ctypes.data if len(clists[l]) else None
    d.apron = apron
    for a in range(3):
        d.axis_count[a] = axis[a]
    d.nchan = len(chan_types)
    for c, (t, data) in enumerate(zip(chan_types, chans)):
        d.chan_type[c] = t
        d.chan_data[c] = data.ctypes.data
    for k in range(12):
        d.transform[k] = transform[k]
    d.root = root
    d.slots = slots
    return d


@pytest.fixture(scope="module")
def grid(oracle_mod):
    ld, root, nodes, clists, slots = oracle_grid(oracle_mod)
    apron, br = 1, (1 << ld[0]) + 2
    axis = (16, 16, (slots + 255) // 256)
    g = np.random.default_rng(9)
    dens = g.standard_normal((slots, br, br, br)).astype(np.float32).view(np.uint32)
    clr = g.integers(0, 2 ** 32, (slots, br, br, br), dtype=np.uint64).astype(np.uint32)
    transform = [0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0]
    return dict(ld=ld, root=root, nodes=nodes, clists=clists, apron=apron, axis=axis, chan_types=[3, 2],
                chans=[dens, clr], slots=slots, transform=transform)


def test_writer_matches_the_specification(grid, tmp_path):
    L = load_library()
    want = spec_file(**grid)
    d = make_desc(**grid)
    path = tmp_path / "g.vbx"
    assert L.gvdbx_vbx_write(str(path).encode(), C.byref(d)) == 0
    got = path.read_bytes()
    assert len(got) == len(want)
    assert got == want


def test_reader_round_trip(grid, tmp_path):
    L = load_library()
    path = tmp_path / "g.vbx"
    path.write_bytes(spec_file(**grid))
    d = VbxDesc()
    assert L.gvdbx_vbx_probe(str(path).encode(), C.byref(d)) == 0
    ld = grid["ld"]
    assert d.levels == len(ld) and [d.log2dim[l] for l in range(len(ld))] == ld
    assert d.root == grid["root"] and d.apron == 1 and tuple(d.axis_count) == grid["axis"]
    assert d.nchan == 2 and d.chan_type[0] == 3 and d.chan_type[1] == 2 and d.slots == grid["slots"]
    nodes = [np.zeros(d.node_count[l], grid["nodes"][l].dtype) for l in range(len(ld))]
    clists = [None] + [np.zeros((d.clist_count[l], 1 << (3 * ld[l])), np.uint64) for l in range(1, len(ld))]
    chans = [np.zeros_like(c) for c in grid["chans"]]
    for l in range(len(ld)):
        d.nodes[l] = nodes[l].ctypes.data
        if l:
            d.clists[l] = clists[l].ctypes.data
    for c in range(2):
        d.chan_data[c] = chans[c].ctypes.data
    assert L.gvdbx_vbx_read(str(path).encode(), C.byref(d)) == 0
    for l in range(len(ld)):
        assert nodes[l].tobytes() == grid["nodes"][l].tobytes()
        if l:
            assert np.array_equal(clists[l], grid["clists"][l])
    for c in range(2):
        assert chans[c].tobytes() == grid["chans"][c].tobytes()


def test_reader_refuses_what_loadvbx_cannot_read(grid, tmp_path):
    """Bitmask nodes (width0 != 64, V.cpp:626-630), compression, brick layout, topology reuse."""
    L = load_library()
    good = spec_file(**grid)
    name_at = 2 + 48 + 4 + 8  # version, transform, grid count, one grid offset
    fields = {"compress": name_at + 256 + 2, "topotype": name_at + 256 + 3 + 12 + 4 + 12 + 4 + 4 + 8,
              "layout": name_at + 256 + 3 + 12 + 4 + 12 + 4 + 4 + 8 + 1 + 4}
    width0 = name_at + 256 + 3 + 12 + 4 + 12 + 4 + 4 + 8 + 1 + 4 + 1 + 24 + 4 + 8 + 4 * 6
    assert struct.unpack_from("<i", good, width0)[0] == 64
    cases = [(fields["compress"], b"\x01"), (fields["topotype"], b"\x01"), (fields["layout"], b"\x01"),
             (width0, struct.pack("<i", 72))]
    for off, val in cases:
        bad = bytearray(good)
        bad[off:off + len(val)] = val
        path = tmp_path / "bad.vbx"
        path.write_bytes(bytes(bad))
        d = VbxDesc()
        assert L.gvdbx_vbx_probe(str(path).encode(), C.byref(d)) == 1, off
    d = VbxDesc()
    assert L.gvdbx_vbx_probe(str(tmp_path / "missing.vbx").encode(), C.byref(d)) == 1
    assert VBX_MAXCHAN == 32
