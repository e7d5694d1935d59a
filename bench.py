"""Benchmark: points -> sparse volume on one MI355X (plain GPU workload).

Workload (the reference's plain GPU configuration): 2^24 particles of a
dam-break block, uniform in [64,576)x[64,320)x[64,576) with +-0.5 voxel
jitter, seeded; ten frames advected by a seeded divergence-free ABC flow
(each velocity component <= 2 voxels/frame).  One step = one frame of the
hot path on resident positions:
    RebuildTopology(N, 2r=3) -> FinishTopology -> UpdateAtlas
    -> InsertPointsSubcell(4, N, 3) -> GatherLevelSet(4, N, r=1.5) -> UpdateApron(0, 3.0)
The ten frames are generated and advected on the GPU before timing.

A second, differentiable leg (2^21 points in 64 anisotropic Gaussian
clusters, per-point weights, r = 2 cubic splat, 512x512 orthographic
emission-absorption raycast, step 0.5, extinction 0.05, loss 1/2|I - I*|^2)
times the forward and the forward+backward and reports their ratio.

Prints one JSON line (contract: metric/value/unit/.../roofline/cpu_baseline).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gvdb-voxels_amd"))

METRIC = "Mpoints/s topology+splat (HBM % of peak); fwd+bwd/fwd time ratio"
HBM_PEAK = 8.0e12
SEED = 20261203 + 1          # configs[1]


def make_frames(torch, gvdbx, n, nframes, dev):
    from gvdbx import workloads
    return workloads.dambreak_frames(torch, gvdbx.gen_box, SEED, n, nframes, dev)


def frame_bytes_alg(n, nleaf, node_cnt, res, depth, br=10, r0=8, passes=3):
    """Algorithmic bytes of one frame (N points, B leaves), per the measurement plan:
    N*(24 + 4 + 16P) + B*(4*br^3 + 4*(br^3 - r0^3) + 64 + 16) + sum_internal(64 + 8*res_l^3)."""
    b = n * (24 + 4 + 16 * passes)
    b += nleaf * (4 * br ** 3 + 4 * (br ** 3 - r0 ** 3) + 64 + 16)
    for l in range(1, depth + 1):
        b += node_cnt[l] * (64 + 8 * res[l] ** 3)
    return b


def event_pool(torch, stream, n):
    """n timing events, each recorded once so that its device event exists
    (torch creates it on first record): creating events inside the timed
    loop would add host time at points where the device waits on the host."""
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(n)]
    for e in evs:
        e.record(stream)
    torch.cuda.synchronize()
    return evs


def measured_value(kernel_profile, key):
    """A `key: value` line of a committed rocprofv3 counter summary, or None."""
    path = os.path.join(ROOT, "profiles", kernel_profile)
    try:
        for line in open(path):
            if line.startswith(key + ":"):
                return int(float(line.split(":")[1]))
    except OSError:
        return None
    return None


def measured_traffic(kernel_profile):
    """HBM bytes per launch of the dominant kernel from its committed rocprofv3
    counter summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    return measured_value(kernel_profile, "traffic_bytes_per_launch")


def stage_counters(profile):
    """Per-stage kernel time and copy-calibrated HBM bytes of the frame from a
    committed rocprofv3 summary (its first table: stage, kernel ms, FETCH GB,
    WRITE GB, ...), or None."""
    path = os.path.join(ROOT, "profiles", profile)
    out = {}
    try:
        for line in open(path):
            c = [x.strip() for x in line.strip().strip("|").split("|")]
            if len(c) == 6 and c[0] in ("topology", "subcell", "gather", "apron", "frame"):
                out[c[0]] = {"kernel_ms": float(c[1]), "fetch_GB": float(c[2]), "write_GB": float(c[3])}
    except (OSError, ValueError):
        return None
    return out or None


def valu_ceiling(stream):
    """Vector-issue ceilings measured on this GPU (libgvdbx_calib.so): wave64
    VALU instructions per second of eight independent FMA chains, and of the
    level-set walk's own instruction mix (readlanes, subtractions, FMAs,
    unsigned min), each on 2048 workgroups of 256 threads."""
    import ctypes as C
    L = C.CDLL(os.path.join(ROOT, "gvdb-voxels_amd", "libgvdbx_calib.so"))
    L.gvdbx_calib_valu.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]
    out = {}
    for kind, name in ((0, "fma_chains"), (1, "walk_mix")):
        ms, valu = C.c_double(0), C.c_double(0)
        code = L.gvdbx_calib_valu(kind, 2048, 20000 if kind == 0 else 8000, C.c_void_p(stream.cuda_stream),
                                  C.byref(ms), C.byref(valu))
        if code != 0:
            raise RuntimeError(f"VALU calibration kernel {kind} failed with status {code}")
        out[name] = {"valu_per_s": valu.value / (ms.value / 1e3), "ms": round(ms.value, 4)}
    return out


def make_clusters(seed, n_clusters=64, per=32768):
    """64 anisotropic Gaussian clusters: means U[64,320)^3, per-axis sigma
    U(4,20), truncated at 3 sigma; weights U(0.1,1); and the target cloud
    with every cluster mean shifted by N(0, 2^2)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    means = rng.uniform(64, 320, size=(n_clusters, 3))
    sig = rng.uniform(4, 20, size=(n_clusters, 3))
    z = rng.standard_normal(size=(n_clusters * per, 3))
    bad = np.abs(z) > 3
    while bad.any():
        z[bad] = rng.standard_normal(size=int(bad.sum()))
        bad = np.abs(z) > 3
    lab = np.repeat(np.arange(n_clusters), per)
    pts = (means[lab] + z * sig[lab]).astype(np.float32)
    w = rng.uniform(0.1, 1.0, size=n_clusters * per).astype(np.float32)
    shift = rng.normal(0.0, 2.0, size=(n_clusters, 3))
    tgt = (means[lab] + shift[lab] + z * sig[lab]).astype(np.float32)
    return pts, w, tgt


def bench_differentiable(torch, gvdbx, dev, steps, warmup):
    """Forward and forward+backward of splat -> apron -> EA raycast."""
    pts, w, tgt = make_clusters(20261203 + 2)
    n = len(pts)
    P = torch.from_numpy(pts).to(dev)
    W = torch.from_numpy(w).to(dev)
    PT = torch.from_numpy(tgt).to(dev)
    stream = torch.cuda.current_stream()
    gd = gvdbx.VolumeGVDB(0, stream=stream)
    gd.configure([3, 3, 3, 3, 5])
    gd.SetChannelDefault(16, 16, 1)
    gd.AddChannel(0, gvdbx.T_FLOAT, 1)
    cam = gvdbx.camera_ortho(512, 512, (64.0, 64.0, 0.0), (256.0, 0, 0), (0, 256.0, 0), (0, 0, 1))
    rp = gvdbx.ea_params(step=0.5, kappa=0.05)
    radius, topo_r = 2.0, 4.0
    ev = {}
    pool = event_pool(torch, stream, 8 * (steps + 1))

    def mark(name, on):
        if on:
            e = pool.pop()
            e.record(stream)
            ev.setdefault(name, []).append(e)

    def forward(pos, record, timed=False):
        gd.SetPoints(pos)
        gd.SetPointWeights(W)
        mark("f0", timed)
        gd.RebuildTopology(n, topo_r)
        gd.FinishTopology(False, True)
        gd.UpdateAtlas()
        mark("f1", timed)
        gd.SplatCubic(n, radius)
        mark("f2", timed)
        gd.UpdateApron(0, 0.0)
        mark("f3", timed)
        I, T = gd.RaycastEA(0, cam, rp, record=record)
        mark("f4", timed)
        return I, T

    Itgt, _ = forward(PT, False)
    Itgt = Itgt.clone()

    def backward(I, T, timed=False):
        gI = I - Itgt
        grad = gd.RaycastEABackward(0, cam, rp, I, T, gI)
        mark("b1", timed)
        gd.ApronAdjoint(grad)
        mark("b2", timed)
        dpos, dw = gd.SplatCubicBackward(n, radius, (0.0, 0.0, 0.0), grad)
        mark("b3", timed)
        return dpos, dw

    for _ in range(warmup):
        forward(P, False)
        I, T = forward(P, True)
        backward(I, T)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        forward(P, False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        I, T = forward(P, True, timed=True)
        backward(I, T, timed=True)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    fwd_ms = (t1 - t0) * 1e3 / steps
    train_ms = (t2 - t1) * 1e3 / steps

    def avg(a, b):
        return sum(x.elapsed_time(y) for x, y in zip(ev[a], ev[b])) / len(ev[a])

    st = {"topology_ms": avg("f0", "f1"), "splat_cubic_ms": avg("f1", "f2"), "apron_ms": avg("f2", "f3"),
          "raycast_ea_record_ms": avg("f3", "f4"), "raycast_ea_backward_ms": avg("f4", "b1"),
          "apron_adjoint_ms": avg("b1", "b2"), "splat_backward_ms": avg("b2", "b3")}
    inf = gd.info()
    return {
        "workload": "2^21 points in 64 Gaussian clusters (means U[64,320)^3, sigma U(4,20) per axis, 3-sigma "
                    "truncation), weights U(0.1,1), cubic splat r=2 (topology radius 4), UpdateApron(0), "
                    "512x512 orthographic +z EA raycast over [64,320)^2, step 0.5, kappa 0.05; "
                    "loss 1/2|I-I*|^2, I* = image of the clusters shifted by N(0,2^2)",
        "n_points": n, "leaf_bricks": int(inf.node_cnt[0]), "fwd_ms": round(fwd_ms, 4),
        "fwd_plus_bwd_ms": round(train_ms, 4), "fwd_bwd_over_fwd": round(train_ms / fwd_ms, 3),
        "stages_ms": {k: round(v, 4) for k, v in st.items()},
    }


def bench_depth_fusion(torch, gvdbx, dev, steps):
    """Depth-frame fusion: all frames in one pass vs one pass per frame."""
    from gvdbx import workloads
    F, W, H = 64, 640, 480
    d, poses, intr = workloads.depth_scene(20261203 + 3, nframes=F, width=W, height=H)
    origin, vox = (-6.5, -2.0, -6.5), 0.02
    dt = torch.from_numpy(d).to(dev)
    pt = torch.from_numpy(poses).to(dev)
    per = W * H
    n = F * per
    stream = torch.cuda.current_stream()
    pts = torch.empty((n, 3), dtype=torch.float32, device=dev)
    gv = gvdbx.VolumeGVDB(0, stream=stream)
    gv.configure([3, 3, 3, 3, 5])
    gv.SetChannelDefault(16, 16, 1)
    gv.AddChannel(0, gvdbx.T_FLOAT, 1)

    def batched():
        gvdbx.backproject_depth(dt, intr, pt, vox, origin, out=pts)
        gv.SetPoints(pts)
        gv.RebuildTopology(n, 2.0)
        gv.FinishTopology(False, True)
        gv.UpdateAtlas()
        ln = gv.InsertPointsSubcell(4, n, 2.0)
        gv.GatherLevelSet(4, n, 1.0, (0.0, 0.0, 0.0), ln, 0)
        gv.UpdateApron(0, 0.0)

    def per_frame():
        gvdbx.backproject_depth(dt, intr, pt, vox, origin, out=pts)
        gv.SetPoints(pts)
        gv.RebuildTopology(n, 2.0)
        gv.FinishTopology(False, True)
        gv.UpdateAtlas()
        gv.FillChannel(0, 3.0)
        for f in range(F):
            fp = pts[f * per:(f + 1) * per]
            gvdbx.backproject_depth(dt[f:f + 1], intr, pt[f:f + 1], vox, origin, out=fp)
            gv.SetPoints(fp)
            lf = gv.InsertPointsSubcell(4, per, 2.0)
            gv.GatherLevelSet(4, per, 1.0, (0.0, 0.0, 0.0), lf, 0, accum=True)
        gv.UpdateApron(0, 0.0)

    def timed(fn, k):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / k

    tb = timed(batched, steps)
    tp = timed(per_frame, max(1, steps // 2))

    # per-frame accumulation on the fixed topology of all frames, eager calls
    # vs one captured frame graph replayed per frame (same kernels, same bits)
    fbuf = torch.empty((per, 3), dtype=torch.float32, device=dev)
    gvdbx.backproject_depth(dt, intr, pt, vox, origin, out=pts)
    gv.SetPoints(pts)
    gv.RebuildTopology(n, 2.0)
    gv.FinishTopology(False, True)
    gv.UpdateAtlas()
    gvdbx.backproject_depth(dt[0:1], intr, pt[0:1], vox, origin, out=fbuf)
    gv.SetPoints(fbuf)
    graph = gv.FrameGraph(per, 4, 2.0, 1.0, chan=0, accum=True, apron_chan=-1)

    def frames_eager():
        gv.FillChannel(0, 3.0)
        for f in range(F):
            gvdbx.backproject_depth(dt[f:f + 1], intr, pt[f:f + 1], vox, origin, out=fbuf)
            lf = gv.InsertPointsSubcell(4, per, 2.0)
            gv.GatherLevelSet(4, per, 1.0, (0.0, 0.0, 0.0), lf, 0, accum=True)
        gv.UpdateApron(0, 0.0)

    def frames_graph():
        gv.FillChannel(0, 3.0)
        for f in range(F):
            gvdbx.backproject_depth(dt[f:f + 1], intr, pt[f:f + 1], vox, origin, out=fbuf)
            graph.launch()
        gv.UpdateApron(0, 0.0)

    te = timed(frames_eager, max(1, steps // 2))
    tgr = timed(frames_graph, max(1, steps // 2))
    del graph
    valid = int((torch.isfinite(dt) & (dt >= 0.5) & (dt <= 8.0)).sum().item())
    inf = gv.info()
    return {
        "workload": f"{F} depth frames {W}x{H} (20 seeded planes/spheres, 4 m orbit, depth [0.5,8] m, "
                    "U(5%,60%) invalid masks), fx=fy=525, 2 cm voxels, level-set fusion r=1 (topology and "
                    "subcell radius 2), UpdateApron(0)",
        "frames": F, "valid_points": valid, "leaf_bricks": int(inf.node_cnt[0]),
        "batched_ms": round(tb, 4), "per_frame_ms": round(tp, 4),
        "batched_Mpoints_per_s": round(valid / (tb / 1e3) / 1e6, 2),
        "batched_speedup": round(tp / tb, 3),
        "fixed_topology_per_frame_eager_ms": round(te, 4), "fixed_topology_per_frame_graph_ms": round(tgr, 4),
        "graph_speedup": round(te / tgr, 3),
    }


def bench_batched_clouds(torch, gvdbx, dev, steps):
    """256 independent clouds (log-uniform sizes 1e4..2e6): splat, apron,
    256x256 EA raycast and the full backward, all clouds per launch vs one
    cloud at a time."""
    import numpy as np
    from gvdbx import workloads
    clouds = workloads.cloud_batch(20261203 + 4, nclouds=256)
    off = np.concatenate([[0], np.cumsum([len(p) for p, w, k in clouds])]).astype(np.int64)
    n = int(off[-1])
    P = torch.from_numpy(np.concatenate([p for p, w, k in clouds])).to(dev)
    Wt = torch.from_numpy(np.concatenate([w for p, w, k in clouds])).to(dev)
    W, H = 256, 256
    cams = []
    for p, w, k in clouds:
        lo, hi = workloads.cloud_camera_box(p)
        cams.append(gvdbx.camera_ortho(W, H, (lo[0] + 0.25, lo[1] + 0.25, -0.3), (hi[0] - lo[0], 0, 0),
                                       (0, hi[1] - lo[1], 0), (0, 0, 1)))
    rp = gvdbx.ea_params(step=0.5, kappa=0.05)
    gI = torch.randn((len(clouds), H, W), generator=torch.Generator().manual_seed(5)).to(dev)
    stream = torch.cuda.current_stream()

    def make():
        g = gvdbx.VolumeGVDB(0, stream=stream)
        g.configure([3, 3, 3, 3, 5])
        g.SetChannelDefault(16, 16, 1)
        g.AddChannel(0, gvdbx.T_FLOAT, 1)
        return g

    gb, gs = make(), make()

    def batched():
        gb.SetPointsBatched(P, off)
        gb.SetPointWeights(Wt)
        gb.RebuildTopology(n, 4.0)
        gb.FinishTopology(False, True)
        gb.UpdateAtlas()
        gb.SplatCubic(n, 2.0)
        gb.UpdateApron(0, 0.0)
        I, T = gb.RaycastEABatched(0, cams, rp, record=True)
        grad = gb.RaycastEABackwardBatched(0, cams, rp, I, T, gI)
        gb.ApronAdjoint(grad)
        gb.SplatCubicBackward(n, 2.0, (0.0, 0.0, 0.0), grad)

    def per_cloud():
        for c in range(len(clouds)):
            a, b = int(off[c]), int(off[c + 1])
            gs.SetPoints(P[a:b])
            gs.SetPointWeights(Wt[a:b])
            gs.RebuildTopology(b - a, 4.0)
            gs.FinishTopology(False, True)
            gs.UpdateAtlas()
            gs.SplatCubic(b - a, 2.0)
            gs.UpdateApron(0, 0.0)
            I, T = gs.RaycastEA(0, cams[c], rp, record=True)
            grad = gs.RaycastEABackward(0, cams[c], rp, I, T, gI[c])
            gs.ApronAdjoint(grad)
            gs.SplatCubicBackward(b - a, 2.0, (0.0, 0.0, 0.0), grad)

    def timed(fn, k):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / k

    tb = timed(batched, steps)
    tp = timed(per_cloud, steps)
    return {
        "workload": "256 clouds, sizes floor(exp(U(ln 1e4, ln 2e6))), seeded shell/box/cluster shapes, weights "
                    "U(0.1,1), each in its own index space; cubic splat r=2 (topology radius 4), UpdateApron(0), "
                    "256x256 orthographic EA raycast per cloud (step 0.5, kappa 0.05, segment checkpoints), "
                    "backward to positions and weights",
        "clouds": len(clouds), "n_points": n, "leaf_bricks": int(gb.info().node_cnt[0]),
        "batched_ms": round(tb, 3), "per_cloud_ms": round(tp, 3),
        "batched_Mpoints_per_s": round(n / (tb / 1e3) / 1e6, 2), "batched_speedup": round(tp / tb, 3),
    }


def host_cpu():
    """nproc and the CPU model name of this host (lscpu's 'Model name')."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return os.cpu_count(), model


def cpu_baseline(pts):
    """The CPU oracle (single-threaded C restatement of the reference
    algorithm) on one host core:
      - the shell workload in full (2^20 points, topology and subcell
        radius 4, GatherDensity r = 2, UpdateApron(0));
      - the dam-break frame 0 on a full-density slab x < 96 (1/16 of the
        points, every slab point of the full frame), topology + subcell lists
        + GatherLevelSet + UpdateApron, extrapolated to the whole frame by its
        point count (the work per point is the same throughout the block).
    value: the dam-break frame's extrapolated Mpoints/s (the unit of the GPU
    line)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import gvdbx
    nproc, model = host_cpu()
    # dam-break frame 0, slab x < 96
    sub = pts[pts[:, 0] < 96.0]
    ov = oracle.OracleVolume((3, 3, 3, 3, 5))
    ov.add_channel(0, 1)
    t0 = time.perf_counter()
    ov.rebuild_topology(sub, 3.0)
    ov.update_atlas()
    ov.insert_points_subcell(4, sub, 3.0)
    ov.gather_level_set(1.5)
    ov.update_apron(0, 3.0)
    dt = time.perf_counter() - t0
    del ov
    frame_s = dt * len(pts) / len(sub)
    # shell workload in full
    shell = gvdbx.gen_shell(20261203, 1 << 20, (256.0, 256.0, 256.0), 180.0, 200.0)
    ov = oracle.OracleVolume((3, 3, 3, 3, 5))
    ov.add_channel(0, 1)
    t0 = time.perf_counter()
    ov.rebuild_topology(shell, 4.0)
    ov.update_atlas()
    ov.insert_points_subcell(4, shell, 4.0)
    ov.gather_density(2.0)
    ov.update_apron(0, 0.0)
    ds = time.perf_counter() - t0
    leaves = ov.info().node_cnt[0]
    del ov
    return {"value": round(len(pts) / frame_s / 1e6, 4), "unit": "Mpoints/s", "cores": 1, "kind": "port",
            "nproc": nproc, "cpu_model": model, "cores_used": 1,
            "sample": f"dam-break frame 0: slab x<96 ({len(sub)} of {len(pts)} points, full density), topology + "
                      f"subcell lists + level set + apron in {dt:.2f} s on 1 host core, extrapolated by point count "
                      f"to {frame_s:.1f} s per frame",
            "shell_workload": {"points": len(shell), "leaf_bricks": int(leaves), "seconds": round(ds, 2),
                               "Mpoints_per_s": round(len(shell) / ds / 1e6, 4),
                               "steps": "topology r4 + subcell lists + GatherDensity r2 + apron, run in full"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--points", type=int, default=1 << 24)
    ap.add_argument("--frames", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default kernel configuration")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiable splat+render leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched depth-fusion and batched-clouds legs")
    ap.add_argument("--only", choices=["diff", "depth", "clouds"], default=None,
                    help="run only this leg and print its JSON (for profiling; not the benchmark line)")
    ap.add_argument("--gather-variant", type=int, default=None,
                    help="0: LDS-staged candidates, 1: ballot/readlane walk (default), 2: point-centric LDS minima")
    ap.add_argument("--xcd-remap", type=int, default=None, help="1: contiguous brick runs per XCD (default 0)")
    ap.add_argument("--topology", choices=["auto", "dense", "radix"], default="auto",
                    help="topology builder (gvdbx_set_topology_mode) of the config-2 frame")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="set a kernel tuning knob (gvdbx_set_tuning) before the run; repeatable")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run three 1 GiB device-to-device copies (HBM counter calibration) and exit")
    ap.add_argument("--calibrate-valu", action="store_true",
                    help="only measure the vector-issue ceilings (libgvdbx_calib.so) and print them")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the path has no multi-GPU form here); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import gvdbx

    dev = torch.device("cuda:0")
    if args.calibrate_copy:
        a = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": a.numel() * 4}))
        return
    if args.calibrate_valu:
        print(json.dumps(valu_ceiling(torch.cuda.current_stream())), flush=True)
        return
    if args.only == "diff":
        print(json.dumps(bench_differentiable(torch, gvdbx, dev, max(3, min(args.steps, 10)),
                                              max(1, min(args.warmup, 2)))), flush=True)
        return
    if args.only == "depth":
        print(json.dumps(bench_depth_fusion(torch, gvdbx, dev, max(2, min(args.steps, 5)))), flush=True)
        return
    if args.only == "clouds":
        print(json.dumps(bench_batched_clouds(torch, gvdbx, dev, 2)), flush=True)
        return
    n = args.points
    frames, pts0 = make_frames(torch, gvdbx, n, args.frames, dev)
    stream = torch.cuda.current_stream()
    gv = gvdbx.VolumeGVDB(0, stream=stream)
    gv.Configure(5, 3, 3, 3, 3)
    gv.SetChannelDefault(16, 16, 1)
    gv.AddChannel(0, gvdbx.T_FLOAT, 1)
    if args.gather_variant is not None:
        gv.SetTuning("gather_variant", args.gather_variant)
    if args.xcd_remap is not None:
        gv.SetTuning("xcd_remap", args.xcd_remap)
    for kv in args.tune:
        k, _, val = kv.partition("=")
        gv.SetTuning(k, int(val))
    gv.SetTopologyMode({"auto": gvdbx.TOPO_AUTO, "dense": gvdbx.TOPO_DENSE, "radix": gvdbx.TOPO_RADIX}[args.topology])
    torch.cuda.synchronize()

    ev = {}
    pool = event_pool(torch, stream, 5 * (args.steps + 1))

    def mark(name):
        e = pool.pop()
        e.record(stream)
        ev.setdefault(name, []).append(e)

    def step(i, timed):
        p = frames[i % len(frames)]
        gv.SetPoints(p)
        if timed:
            mark("a")
        gv.RebuildTopology(n, 3.0, (0.0, 0.0, 0.0))
        gv.FinishTopology(False, True)
        gv.UpdateAtlas()
        if timed:
            mark("b")
        ln = gv.InsertPointsSubcell(4, n, 3.0, (0.0, 0.0, 0.0))
        if timed:
            mark("c")
        gv.GatherLevelSet(4, n, 1.5, (0.0, 0.0, 0.0), ln, 0)
        if timed:
            mark("d")
        gv.UpdateApron(0, 3.0)
        if timed:
            mark("e")
        return ln

    for i in range(args.warmup):
        step(i, False)
    tune = None
    if not args.no_autotune:
        # the library's autotuner on the last warm-up frame: every candidate
        # must reproduce the default's bits, the fastest is kept
        tb = gv.AutotuneBinning(n, 4, 3.0, reps=2)
        tg = gv.AutotuneGather(1, 4, 1.5, chan=0, reps=2)
        tune = {"binning": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in tb.items()},
                "gather": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in tg.items()}}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(args.warmup + i, True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps

    def avg(a, b):
        return sum(x.elapsed_time(y) for x, y in zip(ev[a], ev[b])) / len(ev[a])

    stage = {"topology_ms": avg("a", "b"), "insert_subcell_ms": avg("b", "c"), "gather_levelset_ms": avg("c", "d"),
             "apron_ms": avg("d", "e")}
    inf = gv.info()
    nleaf = inf.node_cnt[0]
    res = [1 << d for d in gv.log2dims]
    binned = n   # every valid point is binned once
    gather_bytes = 16 * binned + 4 * 512 * nleaf + 8 * nleaf
    gather_s = stage["gather_levelset_ms"] / 1e3
    achieved = gather_bytes / gather_s
    fb = frame_bytes_alg(n, nleaf, list(inf.node_cnt), res, inf.depth)
    value = n / (ms / 1e3) / 1e6
    traffic = measured_traffic("gather_levelset_config2_counters.md") if n == 1 << 24 else None
    valu = measured_value("gather_levelset_config2_counters.md", "valu_insts_per_launch") if n == 1 << 24 else None
    ceil = valu_ceiling(stream)
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mpoints/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "dam-break level set: 2^24 points in [64,576)x[64,320)x[64,576) +-0.5 jitter, "
                               "tree {3,3,3,3,5} (8^3 bricks, apron 1), r=1.5 (topology/subcell radius 3.0), "
                               "UpdateApron(3.0), 10 ABC-advected frames, full rebuild per frame",
                   "n_points": n, "frames": args.frames, "leaf_bricks": int(nleaf), "depth": inf.depth,
                   "sc_len": int(inf.sc_len)},
        "roofline": {"bound": "hbm",
                     "kernel": "GatherLevelSet call (k_gather_classes + k_gather_rl + k_gather_items + k_gather_finish)", "achieved": round(achieved / 1e9, 2),
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 5),
                     "traffic": traffic, "alg_bytes_per_launch": gather_bytes,
                     "avg_launch_ms": round(stage["gather_levelset_ms"], 4),
                     "traffic_source": "profiles/gather_levelset_config2_counters.md" if traffic else None,
                     # vector issue: wave64 VALU instructions of the call's kernels (same counter
                     # summary) per second of the event-timed call, against the issue rate of the
                     # walk's instruction mix measured on this GPU by a calibration kernel
                     "valu_issue": {
                         "insts_per_launch": valu,
                         "achieved_per_s": round(valu / gather_s, -6) if valu else None,
                         "ceiling_walk_mix_per_s": round(ceil["walk_mix"]["valu_per_s"], -6),
                         "ceiling_fma_chains_per_s": round(ceil["fma_chains"]["valu_per_s"], -6),
                         "frac": round(valu / gather_s / ceil["walk_mix"]["valu_per_s"], 3) if valu else None}},
        "frame_roofline": {"alg_bytes": fb, "achieved_GBps": round(fb / (ms / 1e3) / 1e9, 2),
                           "frac": round(fb / (ms / 1e3) / HBM_PEAK, 5)},
        "stages_ms": {k: round(v, 4) for k, v in stage.items()},
        # HBM bytes per stage beside the stage times (steady-state counter runs of this workload)
        "stage_traffic": stage_counters("config2_frame_stage_counters.md") if n == 1 << 24 else None,
        "stage_traffic_source": "profiles/config2_frame_stage_counters.md" if n == 1 << 24 else None,
        "autotune": tune,
        "tuning": {k: gv.GetTuning(k) for k in ("gather_variant", "gather_classes", "gather_split", "xcd_remap", "apron_order",
                                                "apron_push", "gather_quarters", "bin_records", "sort_tiles")},
        "topology_mode": {1: "dense", 2: "radix"}.get(inf.topology_mode_used, "?"),
        "fwd_bwd_over_fwd": None,
    }
    if not args.no_diff:
        diff = bench_differentiable(torch, gvdbx, dev, max(3, min(args.steps, 10)), max(1, min(args.warmup, 2)))
        out["differentiable"] = diff
        out["fwd_bwd_over_fwd"] = diff["fwd_bwd_over_fwd"]
    if not args.no_batched:
        out["batched_depth_fusion"] = bench_depth_fusion(torch, gvdbx, dev, max(2, min(args.steps, 5)))
        out["batched_clouds"] = bench_batched_clouds(torch, gvdbx, dev, 2)
        out["batched_speedup"] = out["batched_clouds"]["batched_speedup"]
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(pts0)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
