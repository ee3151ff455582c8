"""torch.autograd.Function check harness for the splat and sampling adjoints.

Runs as its own process (started by tests/test_gpu_autograd_harness.py): torch is imported
before libgvdbx.so so that both bind the one HIP runtime torch ships.  torch is only the
autograd driver here: every tensor lives on the CPU, the forward and backward passes are the
library's HIP kernels called through the C ABI, and the reference gradient comes from torch
autograd over a dense FP64 grid (voxel v centred at v + 0.5, zero outside active bricks),
with the topology recomputed from the points rather than stored.

Workload: the 2,097,152-point adjoint case (masses U(0.5, 1.5), target splatted from a second
draw, loss = mean over active voxels of (rho - rho*)^2), and the sampling backward for
1,048,576 of the points over the target field.  Prints one JSON line; exit status 0 when
every gradient is within 1e-5 of the dense reference (relative to its largest magnitude).
"""
import json
import os
import sys
import time

import numpy as np
import torch  # first: the library then binds torch's HIP runtime

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "gvdb-voxels_amd"))

from gvdbx import VolumeGVDB, synth  # noqa: E402

TOL = 1e-5


def interior(a):
    return a[:, 1:-1, 1:-1, 1:-1]


class Splat(torch.autograd.Function):
    """rho = S^T m (GPU splat); backward = GPU gather of dL/drho."""

    @staticmethod
    def forward(ctx, q, m, vol, chan, chan_grad):
        ctx.vol, ctx.chan_grad = vol, chan_grad
        ctx.qd, ctx.md = vol.to_device(q.numpy()), vol.to_device(m.numpy())
        vol.SplatTrilinear(chan, ctx.qd, len(q), ctx.md)
        return torch.from_numpy(interior(vol.get_channel(chan)).copy())

    @staticmethod
    def backward(ctx, grad):
        vol = ctx.vol
        s, br = vol.num_slots()
        G = np.zeros((s, br, br, br), np.float32)
        G[:, 1:-1, 1:-1, 1:-1] = grad.numpy()
        vol.set_channel(ctx.chan_grad, G)
        n = ctx.qd.nbytes // 12
        gq, gm = vol.alloc(12 * n), vol.alloc(4 * n)
        vol.SplatBackward(ctx.chan_grad, ctx.qd, n, ctx.md, gq, gm)
        dq = torch.from_numpy(gq.download(np.float32).reshape(-1, 3).copy())
        dm = torch.from_numpy(gm.download(np.float32).copy())
        return dq, dm, None, None, None


class Sample(torch.autograd.Function):
    """s = S f (GPU sampler over the padded bricks); backward = GPU dq and splat of dL/ds."""

    @staticmethod
    def forward(ctx, q, field, vol, chan, chan_grad):
        ctx.vol, ctx.chan, ctx.chan_grad = vol, chan, chan_grad
        s, br = vol.num_slots()
        F = np.zeros((s, br, br, br), np.float32)
        F[:, 1:-1, 1:-1, 1:-1] = field.detach().numpy()
        vol.set_channel(chan, F)
        vol.UpdateApron(chan, 0.0)
        ctx.qd = vol.to_device(q.numpy())
        out = vol.alloc(4 * len(q))
        vol.SampleTrilinear(chan, ctx.qd, len(q), out)
        return torch.from_numpy(out.download(np.float32).copy())

    @staticmethod
    def backward(ctx, dout):
        vol = ctx.vol
        n = ctx.qd.nbytes // 12
        dd = vol.to_device(dout.numpy().astype(np.float32))
        gq = vol.alloc(12 * n)
        vol.SampleTrilinearBackward(ctx.chan, ctx.qd, dd, n, gq, chan_grad=ctx.chan_grad)
        dq = torch.from_numpy(gq.download(np.float32).reshape(-1, 3).copy())
        dfield = torch.from_numpy(interior(vol.get_channel(ctx.chan_grad)).copy())
        return dq, dfield, None, None, None


class Dense:
    """Dense FP64 grid over the active bricks, for torch-autograd reference gradients."""

    def __init__(self, vol):
        nodes = vol.get_nodes(0)
        pos = nodes["pos"][nodes["flags"] == 1].astype(np.int64)  # slot order = pool order of flagged leaves
        self.res = 8
        self.dim = int(pos.max()) + self.res + 2
        d = self.dim
        self.active = torch.zeros(d ** 3, dtype=torch.bool)
        lz, ly, lx = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
        gx = pos[:, 0, None, None, None] + lx
        gy = pos[:, 1, None, None, None] + ly
        gz = pos[:, 2, None, None, None] + lz
        self.slot_index = torch.from_numpy(((gz * d + gy) * d + gx).reshape(-1))  # [slot, z, y, x] order
        self.active[self.slot_index] = True

    def lin(self, x, y, z):
        d = self.dim
        return (z * d + y) * d + x

    def corners(self, q):
        """Per point: home-voxel validity, 8 corner indices and differentiable weights."""
        c = torch.floor(q.detach() - 0.5)
        al = q - 0.5 - c
        ci = c.long()
        d = self.dim
        home = torch.floor(q.detach()).long()
        inside = (q.detach() >= 0).all(1) & (home < d).all(1)
        home = home.clamp(0, d - 1)
        ok = inside & self.active[self.lin(home[:, 0], home[:, 1], home[:, 2])]
        idx, w = [], []
        for k in range(8):
            dx, dy, dz = k & 1, (k >> 1) & 1, k >> 2
            gx, gy, gz = ci[:, 0] + dx, ci[:, 1] + dy, ci[:, 2] + dz
            inb = (gx >= 0) & (gy >= 0) & (gz >= 0) & (gx < d) & (gy < d) & (gz < d)
            li = self.lin(gx.clamp(0, d - 1), gy.clamp(0, d - 1), gz.clamp(0, d - 1))
            valid = ok & inb & self.active[li]
            wx = al[:, 0] if dx else 1 - al[:, 0]
            wy = al[:, 1] if dy else 1 - al[:, 1]
            wz = al[:, 2] if dz else 1 - al[:, 2]
            idx.append(li)
            w.append(wx * wy * wz * valid)
        return idx, w

    def splat(self, q, m):
        rho = torch.zeros(self.dim ** 3, dtype=torch.float64)
        idx, w = self.corners(q)
        for li, wk in zip(idx, w):
            rho = rho.index_add(0, li, m * wk)
        return rho

    def sample(self, q, field):
        idx, w = self.corners(q)
        return sum(wk * field[li] for li, wk in zip(idx, w))


def max_rel(got, want):
    got, want = got.double(), want.double()
    return float((got - want).abs().max() / max(float(want.abs().max()), 1e-30))


def main():
    t0 = time.time()
    a, ma = synth.uniform_massive(seed=4)
    b, mb = synth.uniform_massive(seed=4 + 1000)
    vol = VolumeGVDB(0)
    vol.configure_levels((3, 3, 3, 3, 5))
    vol.SetChannelDefault(16, 16, 1)
    for c in range(4):
        vol.AddChannel(c, apron=1)
    allp = np.concatenate([a, b])
    dall = vol.to_device(allp)
    vol.SetPoints(dall, n=len(allp))
    vol.RebuildTopology(len(allp), 1.0)
    vol.UpdateAtlas()
    print(f"topology done {time.time() - t0:.1f}s", flush=True)

    # ---- splat adjoint through the autograd Function (GPU) and the dense reference (CPU)
    qa = torch.from_numpy(a).requires_grad_(True)
    mta = torch.from_numpy(ma).requires_grad_(True)
    rho = Splat.apply(qa, mta, vol, 0, 2)
    with torch.no_grad():
        tgt = Splat.apply(torch.from_numpy(b), torch.from_numpy(mb), vol, 1, 2)
    loss = ((rho.double() - tgt.double()) ** 2).mean()
    loss.backward()
    print(f"gpu splat adjoint done {time.time() - t0:.1f}s", flush=True)

    dense = Dense(vol)
    nvox = dense.slot_index.numel()
    qr = torch.from_numpy(a).double().requires_grad_(True)
    mr = torch.from_numpy(ma).double().requires_grad_(True)
    rho_r = dense.splat(qr, mr)
    tgt_r = dense.splat(torch.from_numpy(b).double(), torch.from_numpy(mb).double()).detach()
    diff = (rho_r - tgt_r)[dense.slot_index]
    loss_r = (diff * diff).sum() / nvox
    loss_r.backward()
    print(f"dense reference done {time.time() - t0:.1f}s", flush=True)
    res = {
        "points": len(a),
        "slots": int(vol.num_slots()[0]),
        "rho_vs_dense": max_rel(rho.reshape(-1), rho_r.detach()[dense.slot_index]),
        "loss_rel": abs(float(loss) - float(loss_r)) / abs(float(loss_r)),
        "dq_vs_dense": max_rel(qa.grad, qr.grad),
        "dm_vs_dense": max_rel(mta.grad, mr.grad),
    }

    # ---- sampling backward: s = S f over 1,048,576 of the points, L = sum(s * r) with r random
    nq = 1_048_576
    g = np.random.default_rng(11)
    wts = torch.from_numpy(g.standard_normal(nq))
    qs = torch.from_numpy(a[:nq]).requires_grad_(True)
    field = tgt.clone().requires_grad_(True)
    s = Sample.apply(qs, field, vol, 3, 2)
    (s.double() * wts).sum().backward()
    qsr = torch.from_numpy(a[:nq]).double().requires_grad_(True)
    fr = torch.zeros(dense.dim ** 3, dtype=torch.float64)
    fr[dense.slot_index] = tgt.reshape(-1).double()
    fr.requires_grad_(True)
    sr = dense.sample(qsr, fr)
    (sr * wts).sum().backward()
    res["sample_vs_dense"] = max_rel(s, sr.detach())
    res["sample_dq_vs_dense"] = max_rel(qs.grad, qsr.grad)
    res["sample_dfield_vs_dense"] = max_rel(field.grad.reshape(-1), fr.grad[dense.slot_index])
    res["seconds"] = round(time.time() - t0, 1)
    print(json.dumps(res), flush=True)
    bad = [k for k, v in res.items() if k.endswith("dense") and not v <= TOL] + \
          (["loss_rel"] if not res["loss_rel"] <= TOL else [])
    vol.close()
    if bad:
        print("FAILED:", bad, flush=True)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
