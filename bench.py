"""Benchmark of the point-to-sparse-grid step on one MI355X.

Workload (the SPH-style step): 4,194,304 particles on a jittered lattice filling the
lower 40% of a 256^3 box, velocities ~N(0,1), dt = 0.5, seed 1.  One step is
  q += v*dt (on the device, gvdbx_advance_points; not on step 0)
  + RebuildTopology(r=1) + UpdateAtlas + InsertPointsSubcell(sc=4, r=1)
  + GatherDensity(ch0) + velocity means (ch1-3) + UpdateApron(ch0..3)
with the points and velocities resident in HBM.  The rebuild state machine runs as in the
reference: steps 0 and 1 are full rebuilds (inside the warmup), later steps are incremental.
The headline times the steps after the warmup; `late_steps` times steps 40-59 of the same
60-step trajectory (the steps in between run untimed).

Prints one JSON line with the throughput in Mpoints/s, the HBM roofline of the dominant
kernel (HIP events on the library stream around that kernel) and a CPU baseline (the
oracle restatement on the host: topology on one thread, binning / gathers / apron on one
thread and on all threads).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gvdb-voxels_amd"))

HBM_PEAK_GBS = 8000.0


def progress(msg):
    """Stage marker on stderr (the JSON line alone goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)
N_POINTS = 4_194_304
DT = np.float32(0.5)
RADIUS = 1.0
SUBCELL = 4


LATE_FIRST, LATE_LAST = 40, 59  # the late window of the 60-step trajectory


def run_step(v, n, advance):
    if advance:
        v.AdvancePoints(n, DT)  # q += v dt on the device
    v.RebuildTopology(n, RADIUS)
    v.UpdateAtlas()
    ln = v.InsertPointsSubcell(SUBCELL, n, RADIUS)
    # GatherDensity into channel 0 with the velocity means (channels 1-3) in the same pass
    v.GatherDensityVelocity(SUBCELL, n, RADIUS, (0, 0, 0), ln, 0, (1, 2, 3))
    v.UpdateApron()  # all four channels, boundary value 0
    return ln


def algorithmic_bytes(n, s_len, nsc, bricks, chans=4, vel=True):
    """Compulsory HBM bytes of one insert+gather+apron step (lower bound):
    points read once, binned copies written and read once, subcell count+offset,
    interior written, apron read+written."""
    return (n * (12 + 12 * vel) + s_len * (24 + 24 * vel) + 8 * nsc + 4 * chans * bricks * 512
            + 8 * chans * bricks * 488)


APRON_VOX = 10 ** 3 - 8 ** 3  # apron voxels of a padded 8^3 brick with apron 1


def design_step_bytes(n, s_len, nsc, bricks, chans=4):
    """HBM bytes this design must move per step: positions and velocities read once, the binned
    positions written and read once (no velocity lists: the velocity means come from per-subcell
    sums), per-subcell count and offset, the 64-bit velocity sums written and read, and every
    channel's padded brick written once (the gather writes the apron too, so no apron copy)."""
    return n * 24 + s_len * 24 + nsc * 8 + nsc * 48 + chans * bricks * 1000 * 4


def gather_bytes(s_len, nsc, bricks, vel=True, apron=True):
    """Algorithmic bytes of one gather launch: list positions read, per-subcell
    count/offset/corner/leaf read, interior voxels written; with the velocity means in the same
    pass, per-subcell count and 64-bit sums read and 3 more channels' interior written; with the
    apron computed in the same pass, its voxels written once per channel (the 4-byte write of
    the apron term 8 C B 488 of the step's bytes, whose copy the pass makes unnecessary)."""
    chans = 4 if vel else 1
    b = s_len * 12 + nsc * (4 + 4 + 12 + 4) + bricks * 512 * 4
    if vel:
        b += nsc * (4 + 24) + 3 * bricks * 512 * 4
    if apron:
        b += chans * bricks * APRON_VOX * 4
    return b


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def omp_threads():
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return max(1, min(16, len(os.sched_getaffinity(0))))


def median_s(fn, reps=5, warm=1):
    """Median wall time of `reps` calls after `warm` untimed ones."""
    for _ in range(warm):
        fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def cpu_baseline(p0, vel, sample):
    """The oracle's C restatement on the host cores, on the first `sample` points of steps 0-2
    of the trajectory: RebuildTopology + UpdateAtlas on one thread (the reference's CPU topology
    is serial), then binning + density + 3 velocity means + 4 aprons on one thread and on all
    threads; medians of 5 after 1 warm-up.  `value` = the incremental step 2 with the binning,
    gathers and apron on all threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    ps0, vs = p0[:sample], vel[:sample]
    ps1 = (ps0 + vs * DT).astype(np.float32)
    ps2 = (ps1 + vs * DT).astype(np.float32)
    threads = omp_threads()

    def tree():
        t = O.Tree((3, 3, 3, 3, 5))
        for c in range(4):
            t.add_channel(c)
        return t

    def full():
        t = tree()
        t.rebuild_topology(ps0, RADIUS)
        t.update_atlas()
    t_full = median_s(full)

    inc = []
    for k in range(6):  # steps 0 and 1 (full) untimed, step 2 (incremental) timed; the first is warm-up
        t = tree()
        t.rebuild_topology(ps0, RADIUS)
        t.rebuild_topology(ps1, RADIUS)
        t0 = time.perf_counter()
        t.rebuild_topology(ps2, RADIUS)
        t.update_atlas()
        if k:
            inc.append(time.perf_counter() - t0)
    t_inc = float(np.median(inc))

    def bga():
        t.insert_points_subcell(SUBCELL, ps2, RADIUS)
        t.gather_density(RADIUS, 0)
        t.gather_velocity(vs, (1, 2, 3))
        for c in range(4):
            t.update_apron(c, 0.0)
    O.set_threads(1)
    t_bga1 = median_s(bga)
    O.set_threads(threads)
    try:
        t_bgan = median_s(bga)
    finally:
        O.set_threads(1)
    del t
    t_act = median_s(lambda: O.cpu_activate_space(ps0))
    mps = lambda sec: round(sample / sec / 1e6, 3)
    return {"value": mps(t_inc + t_bgan), "unit": "Mpoints/s", "cores": threads, "kind": "port",
            "sample": f"first {sample} points of steps 0-2 of the trajectory; value = incremental step 2: "
                      f"RebuildTopology + UpdateAtlas on 1 thread, InsertPointsSubcell + density + 3 velocity "
                      f"means + 4 aprons on {threads} threads (OpenMP); oracle C restatement; medians of 5 after "
                      f"1 warm-up",
            "cpu_model": cpu_model(), "hardware_concurrency": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)), "omp_threads": threads,
            "ms": {"full_topology_step0_1_thread": round(t_full * 1e3, 2),
                   "incremental_topology_step2_1_thread": round(t_inc * 1e3, 2),
                   "bin_gather_apron_1_thread": round(t_bga1 * 1e3, 2),
                   f"bin_gather_apron_{threads}_threads": round(t_bgan * 1e3, 2)},
            "Mpoints_s": {"incremental_step_1_thread": mps(t_inc + t_bga1),
                          f"incremental_step_{threads}_threads": mps(t_inc + t_bgan),
                          "full_step0_1_thread": mps(t_full + t_bga1),
                          f"full_step0_{threads}_threads": mps(t_full + t_bgan)},
            "reference_cpu_activate_space_Mpoints_s": mps(t_act)}


def adjoint_bench(reps):
    """The 2,097,152-point adjoint workload: forward splat of the points (masses U(0.5, 1.5))
    against the backward gather of dL/drho into dL/dq and dL/dm, on the union topology of the
    points and a second draw (the target).  Inputs resident in HBM; wall clock over `reps`
    synchronised repetitions of each."""
    from gvdbx import VolumeGVDB, synth
    a, ma = synth.uniform_massive(seed=4)
    b, _ = synth.uniform_massive(seed=4 + 1000)
    v = VolumeGVDB(0)
    v.Configure(5, 3, 3, 3, 3)
    v.SetChannelDefault(16, 16, 1)
    for c in range(2):
        v.AddChannel(c, apron=1)
    allp = np.concatenate([a, b])
    dall = v.to_device(allp)
    v.SetPoints(dall, n=len(allp))
    v.RebuildTopology(len(allp), RADIUS)
    v.UpdateAtlas()
    n = len(a)
    da, dm = v.to_device(a), v.to_device(ma)
    s, br = v.num_slots()
    G = np.random.default_rng(0).standard_normal((s, br, br, br)).astype(np.float32) * np.float32(1e-7)
    v.set_channel(1, G)
    gq, gm = v.alloc(12 * n), v.alloc(4 * n)
    for _ in range(2):
        v.SplatTrilinear(0, da, n, dm)
        v.SplatBackward(1, da, n, dm, gq, gm)
    v.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        v.SplatTrilinear(0, da, n, dm)
    v.synchronize()
    t1 = time.perf_counter()
    for _ in range(reps):
        v.SplatBackward(1, da, n, dm, gq, gm)
    v.synchronize()
    t2 = time.perf_counter()
    ev = {}
    for name in ("splat", "splat_sort", "splat_brick"):
        v.kernel_timer(name, True)
        for _ in range(reps):
            v.SplatTrilinear(0, da, n, dm)
        v.synchronize()
        k_ms, k_n = v.kernel_timer_read()
        v.kernel_timer(None, False)
        ev[name] = k_ms / max(k_n, 1)
    v.kernel_timer("splat_backward", True)
    for _ in range(reps):
        v.SplatBackward(1, da, n, dm, gq, gm)
    v.synchronize()
    k_ms, k_n = v.kernel_timer_read()
    v.kernel_timer(None, False)
    ev["splat_backward"] = k_ms / max(k_n, 1)
    fwd = (t1 - t0) * 1e3 / reps
    bwd = (t2 - t1) * 1e3 / reps
    v.close()
    # forward algorithmic bytes: positions + masses read, interior of every brick written
    fbytes = n * 16 + int(s) * 512 * 4
    return {"workload": "adjoint_2M (2,097,152 points U(0,256)^3, masses U(0.5,1.5), r=1 union topology)",
            "bricks": int(s), "forward_ms": round(fwd, 4), "backward_ms": round(bwd, 4),
            "backward_over_forward": round(bwd / fwd, 3),
            "forward_Mpoints_s": round(n / (fwd * 1e-3) / 1e6, 1),
            "event_ms": {k: round(x, 4) for k, x in ev.items()},
            "forward_algorithmic_bytes": fbytes,
            "forward_hbm_frac": round(fbytes / (fwd * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}


# wave64 FP32 VALU instructions per second that MI355X's published FP32 vector peak implies:
# 157.3 TFLOP/s / (64 lanes x 2 flops per FMA) = 256 CUs x 4 SIMDs x 2.4 GHz / 2 cycles
VALU_WAVE_INSTR_PER_S = 256 * 4 * 2.4e9 / 2


def levelset_valu_per_entry():
    """VALU wave-instructions per list entry of the level-set kernel (k_levelset_sc4), from the
    committed counter summary (profiles/pmc_levelset.json: SQ_INSTS_VALU of the same workload
    over its list length), or None."""
    f = os.path.join(ROOT, "profiles", "pmc_levelset.json")
    if not os.path.exists(f):
        return None
    e = json.load(open(f)).get("kernels", {}).get("k_levelset_sc4")
    return e["valu_per_entry"] if e else None


def levelset_bench(reps):
    """The level-set workload: 16,777,216 points on 64 sphere surfaces in 1024^3, full
    RebuildTopology(r=2) + UpdateAtlas + InsertPointsSubcell(4, r=2) + ScatterReduceLevelSet
    (r=2, init 3) + UpdateApron(3) + 1,048,576 trilinear queries; per-stage wall clock with a
    synchronisation after each stage, best of `reps` passes."""
    from gvdbx import VolumeGVDB, synth
    p = synth.sphere_surfaces()
    qs = synth.uniform_queries()
    n = len(p)
    v = VolumeGVDB(0)
    v.Configure(5, 3, 3, 3, 3)
    v.SetChannelDefault(16, 16, 1)
    v.AddChannel(0, apron=1)
    d, dq = v.to_device(p), v.to_device(qs)
    out = v.alloc(4 * len(qs))
    stages = ("rebuild", "atlas", "insert", "scatter_reduce", "apron", "sample")
    best = {k: float("inf") for k in stages}
    for _ in range(reps + 1):
        t = [time.perf_counter()]
        v.SetPoints(d, n=n)
        v.RequestFullRebuild(True)
        v.RebuildTopology(n, 2.0)
        v.synchronize()
        t.append(time.perf_counter())
        v.UpdateAtlas()
        v.synchronize()
        t.append(time.perf_counter())
        ln = v.InsertPointsSubcell(4, n, 2.0)
        v.synchronize()
        t.append(time.perf_counter())
        v.ScatterReduceLevelSet(2.0, 0, 3.0)
        v.synchronize()
        t.append(time.perf_counter())
        v.UpdateApron(0, 3.0)
        v.synchronize()
        t.append(time.perf_counter())
        v.SampleTrilinear(0, dq, len(qs), out)
        v.synchronize()
        t.append(time.perf_counter())
        for i, k in enumerate(stages):
            best[k] = min(best[k], (t[i + 1] - t[i]) * 1e3)
    bricks = v.getNumUsedNodes(0)
    v.close()
    total = sum(best.values()) - best["sample"]
    sr = best["scatter_reduce"] * 1e-3
    vpe = levelset_valu_per_entry()  # measured by SQ_INSTS_VALU (the kernel's instruction count)
    return {"workload": "levelset_16M (16,777,216 points on 64 spheres in 1024^3, r=2, sc=4)",
            "bricks": int(bricks), "subcell_entries": int(ln), "ms": {k: round(x, 4) for k, x in best.items()},
            "points_to_sdf_ms": round(total, 4), "Mpoints_s": round(n / (total * 1e-3) / 1e6, 1),
            "pair_evals_per_s": round(64 * ln / sr, 1),
            "scatter_valu_frac": (round(vpe * ln / sr / VALU_WAVE_INSTR_PER_S, 4) if vpe else None),
            "scatter_valu_per_entry": vpe,
            "queries_Mper_s": round(len(qs) / (best["sample"] * 1e-3) / 1e6, 1)}


def batched_bench(reps, tune=True, per_cloud=True):
    """The batched workload: 256 Gaussian-blob clouds with log-uniform sizes in
    [4,096, 1,048,576] (8 blobs each, sigma U(4, 32), centres U(96, 416)^3, seed 5), each with
    its own grid, r = 1.  The autotuner picks brick size, subcell size and launch shape (threads
    per brick workgroup, grid or persistent form) on the graph-launched insert + gather + apron.
    Timed, best of `reps` synchronised passes each:
      - graph: InsertPointsSubcell + GatherDensity + UpdateApron for all clouds, recorded once on
        the batch's topology and run as one graph launch (the configured form);
      - eager: the same calls issued one by one on the batch;
      - per cloud: the same calls on each cloud's own context (256 contexts, default launch
        shape, topology built and every call run once beforehand), one after the other;
      - full steps: RebuildTopology (full) + UpdateAtlas + the three calls, batched and per cloud."""
    from gvdbx import autotune, synth
    cs = synth.blob_clouds()
    counts = [len(c) for c in cs]
    allp = np.concatenate(cs)
    n = len(allp)
    mc = autotune.max_valid_coord(cs)
    if tune:
        tuned = autotune.tune(lambda v: v.to_device(allp), counts, mc, radius=RADIUS, reps=3, log=progress)
    else:  # a configuration the autotuner picks for this batch
        tuned = {"leaf_log2": 3, "subcell": 4, "block": 0, "persistent": False,
                 "cell_log2": autotune.cell_for(3, mc, RADIUS), "fixed": True}
    leaf, sc = tuned["leaf_log2"], tuned["subcell"]

    v = autotune.new_batch_volume(leaf)
    d = v.to_device(allp)
    autotune.bind(v, leaf, d, counts, RADIUS, mc)
    v.SetLaunch(tuned["block"], tuned["persistent"])
    best_full = float("inf")
    for _ in range(reps + 1):
        v.synchronize()
        t0 = time.perf_counter()
        v.RequestFullRebuild(True)
        autotune.run_batch(v, n, RADIUS, sc)
        v.synchronize()
        best_full = min(best_full, (time.perf_counter() - t0) * 1e3)
    best_eager = float("inf")
    for _ in range(reps + 1):
        v.synchronize()
        t0 = time.perf_counter()
        autotune.insert_gather_apron(v, n, RADIUS, sc)
        v.synchronize()
        best_eager = min(best_eager, (time.perf_counter() - t0) * 1e3)
    autotune.record_step(v, n, RADIUS, sc)
    v.GraphLaunch()
    best_graph = autotune.time_launches(v, reps + 1)
    _, ln, _ = v.subcell_info()
    bricks = v.getNumUsedNodes(0)
    v.GraphRelease()
    v.close()
    del d
    out = {"workload": "batched_256 (256 blob clouds, sizes log-uniform in [4096, 1048576], own grid each, r=1)",
           "points": n, "clouds": len(cs), "bricks": int(bricks), "subcell_entries": int(ln), "autotune": tuned,
           "graph_ms": round(best_graph, 3), "graph_Mpoints_s": round(n / (best_graph * 1e-3) / 1e6, 1),
           "eager_ms": round(best_eager, 3), "full_step_ms": round(best_full, 3),
           "full_step_Mpoints_s": round(n / (best_full * 1e-3) / 1e6, 1)}
    if not per_cloud:
        return out

    vols, devs = [], []
    for c in cs:
        w = autotune.new_batch_volume(leaf)
        dc = w.to_device(c)
        w.SetPoints(dc, n=len(c))
        autotune.run_batch(w, len(c), RADIUS, sc)
        vols.append(w)
        devs.append(dc)
    best_p = float("inf")
    for _ in range(reps + 1):
        for w in vols:
            w.synchronize()
        t0 = time.perf_counter()
        for w, c in zip(vols, cs):
            autotune.insert_gather_apron(w, len(c), RADIUS, sc)
        for w in vols:
            w.synchronize()
        best_p = min(best_p, (time.perf_counter() - t0) * 1e3)
    best_pf = float("inf")
    for _ in range(reps + 1):
        for w in vols:
            w.synchronize()
        t0 = time.perf_counter()
        for w, c in zip(vols, cs):
            w.RequestFullRebuild(True)
            autotune.run_batch(w, len(c), RADIUS, sc)
        for w in vols:
            w.synchronize()
        best_pf = min(best_pf, (time.perf_counter() - t0) * 1e3)
    del devs
    for w in vols:
        w.close()
    out.update({"per_cloud_ms": round(best_p, 3), "per_cloud_Mpoints_s": round(n / (best_p * 1e-3) / 1e6, 1),
                "speedup": round(best_p / best_graph, 3),
                "per_cloud_full_step_ms": round(best_pf, 3), "full_step_speedup": round(best_pf / best_full, 3)})
    return out


def calibrate_copy():
    """A plain device-to-device copy of 256 MiB, three times: the reference reading of the HBM
    counters (FETCH_SIZE / WRITE_SIZE) that scales the kernel's counter values to bytes.
    Uses torch only (the library is not loaded in this mode)."""
    import torch
    n = 64 * 1024 * 1024
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibration_copy_bytes": n * 4, "copies": 3}))


def measured_traffic_ratio(kernel):
    """HBM bytes over algorithmic bytes of one launch of the roofline kernel, from the committed
    counter summary (profiles/pmc_traffic.json: the memory-side read and write requests of the
    same workload's last step, by request size), or None."""
    f = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(f):
        return None
    e = json.load(open(f)).get("kernels", {}).get(kernel)
    return round(e["bytes"] / e["algorithmic_bytes"], 4) if e and e.get("algorithmic_bytes") else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--points", type=int, default=N_POINTS)
    ap.add_argument("--cpu-sample", type=int, default=1_048_576)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-levelset", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-cloud batch against per-cloud runs")
    ap.add_argument("--batched-only", action="store_true",
                    help="only time the 256-cloud batch (8^3 bricks, subcell 4; no autotune, no per-cloud runs), "
                         "for kernel profiles")
    ap.add_argument("--no-late", action="store_true", help="skip the late window (steps 40-59)")
    ap.add_argument("--block", type=int, default=0,
                    help="threads per brick workgroup of the count, insert and density kernels (0: chosen from "
                         "the items per brick)")
    ap.add_argument("--levelset-only", action="store_true",
                    help="only time the 16M-point level set (rebuild, insert, ScatterReduceLevelSet, apron, "
                         "queries), for kernel profiles")
    ap.add_argument("--kernel", default="gather", help="kernel family timed for the roofline")
    ap.add_argument("--calibrate-copy", action="store_true", help="only run the counter-calibration copy")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.batched_only:
        print(json.dumps({"batched": batched_bench(args.steps, tune=False, per_cloud=False)}))
        return
    if args.levelset_only:
        print(json.dumps({"levelset": levelset_bench(args.steps)}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "described in DESIGN.md, not implemented)"}))
        sys.exit(2)
    if args.warmup < 2:
        args.warmup = 2  # the two full rebuilds of the reference state machine stay outside timing

    from gvdbx import VolumeGVDB, synth

    p0, vel = synth.sph_block(n=args.points, seed=1)
    n = len(p0)
    total = args.warmup + args.steps

    v = VolumeGVDB(0)
    v.Configure(5, 3, 3, 3, 3)
    v.SetChannelDefault(16, 16, 1)
    for c in range(4):
        v.AddChannel(c, apron=1)
    dpos = v.to_device(p0)
    dvel = v.to_device(vel)
    v.SetPoints(dpos, dvel, n=n)
    v.SetLaunch(args.block, False)

    for k in range(args.warmup):
        run_step(v, n, k > 0)
    v.synchronize()

    # timed region: K steps, kernel-family events recorded on the library stream
    v.kernel_timer(args.kernel, True)
    t0 = time.perf_counter()
    shapes = []  # (list length, subcells, bricks) of every timed step: host-side values, no wait
    for k in range(args.warmup, total):
        ln = run_step(v, n, True)
        shapes.append((ln, v.subcell_info()[0], v.getNumUsedNodes(0)))
    v.synchronize()
    t1 = time.perf_counter()
    k_ms, k_launches = v.kernel_timer_read()
    v.kernel_timer(None, False)
    s_len, nsc, bricks = shapes[-1]

    late = None
    if total <= LATE_FIRST and not args.no_late:
        for k in range(total, LATE_FIRST):
            run_step(v, n, True)
        v.synchronize()
        t2 = time.perf_counter()
        for k in range(LATE_FIRST, LATE_LAST + 1):
            run_step(v, n, True)
        v.synchronize()
        t3 = time.perf_counter()
        late_ms = (t3 - t2) * 1e3 / (LATE_LAST - LATE_FIRST + 1)
        late = {"steps": [LATE_FIRST, LATE_LAST], "ms_per_step": round(late_ms, 4),
                "Mpoints_s": round(n / (late_ms * 1e-3) / 1e6, 2),
                "bricks_last_step": int(v.getNumUsedNodes(0)),
                "subcell_entries_last_step": int(v.subcell_info()[1])}
        progress(f"late steps {LATE_FIRST}-{LATE_LAST}: {late_ms:.4f} ms")

    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = n / (ms_per_step * 1e-3) / 1e6
    # byte counts are averaged over the timed steps (the topology grows from step to step)
    mean = lambda f: sum(f(*x) for x in shapes) / len(shapes)
    step_bytes = mean(lambda sl, ns, b: algorithmic_bytes(n, sl, ns, b))
    design_bytes = mean(lambda sl, ns, b: design_step_bytes(n, sl, ns, b))
    gather = args.kernel == "gather"
    kb = mean(gather_bytes) if gather else None
    kb_interior = mean(lambda sl, ns, b: gather_bytes(sl, ns, b, apron=False)) if gather else None
    k_avg_ms = k_ms / max(k_launches, 1)
    achieved = kb / (k_avg_ms * 1e-3) / 1e9 if kb else None
    ratio = measured_traffic_ratio("k_density_brick") if gather else None

    out = {
        "metric": "Mpoints/s insert+gather+apron (SPH step, 4M particles)",
        "value": round(value, 2),
        "unit": "Mpoints/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered lattice, N(0,1) velocities; displaced on the device every step)",
        "config": {"workload": "sph_step_4M", "points": n, "box": 256, "fill": 0.4, "dt": 0.5, "block": args.block,
                   "radius": RADIUS, "subcell": SUBCELL, "log2dims": [3, 3, 3, 3, 5], "channels": 4,
                   "apron": 1, "rebuild": "reference state machine (incremental after 2 full)",
                   "bricks_last_step": int(bricks), "subcell_entries_last_step": int(s_len)},
        "step_algorithmic_bytes": int(step_bytes),
        "step_hbm_frac": round(step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "step_design_bytes": int(design_bytes),
        "step_design_hbm_frac": round(design_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": {"bound": "hbm", "kernel": args.kernel,
                     "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": int(kb * ratio) if ratio else None,
                     "traffic_over_algorithmic": ratio,
                     "avg_launch_ms": round(k_avg_ms, 5), "launches": int(k_launches),
                     "algorithmic_bytes_per_launch": int(kb) if kb else None,
                     "algorithmic_bytes_interiors_only": int(kb_interior) if kb_interior else None},
    }
    out["late_steps"] = late
    del dpos, dvel
    v.close()
    progress(f"sph step done: {ms_per_step:.4f} ms")
    if not args.no_adjoint:
        out["adjoint"] = adjoint_bench(max(args.steps, 5))
        progress("adjoint done")
    if not args.no_levelset:
        out["levelset"] = levelset_bench(3)
        progress("level set done")
    if not args.no_batched:
        out["batched"] = batched_bench(2)
        progress("batched done")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p0, vel, min(args.cpu_sample, n))
        progress("cpu baseline done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
