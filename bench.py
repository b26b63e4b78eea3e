#!/usr/bin/env python3
"""Benchmark: Mvoxels/s to exact max-flow on one MI355X.

Default workload (the large 2D cut): 15000x10000 RGB uint8 image from the
synthetic generator (40 colour blobs, noise sigma 25, scribble seeds of
capacity 2^30), integer k-means data term (K=5 per side), 8-connected grid with
pairwise round(50 exp(-|dI|^2 / 2 s2)) (diagonals / sqrt 2), int32 capacities.
A step is one full solve from the capacities resident in HBM to the exact int64
flow and the canonical labels resident in HBM.

Prints one JSON line.  --gpus N > 1 exits: this build runs on one GPU.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "jf-cut_amd"))

import numpy as np  # noqa: E402

import jfcut  # noqa: E402

METRIC = "Mvoxels/s to exact max-flow on 1 GPU; HBM GB/s; batched fwd+bwd vs plain"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "c2": dict(W=15000, H=10000, conn=8, nblobs=40, noise=25.0, lam=50.0, k=5,
               name="15000x10000 RGB 8-connected grid cut, k-means K=5, lambda 50, int32"),
    "c1": dict(W=512, H=512, conn=4, nblobs=8, noise=25.0, lam=50.0, k=5,
               name="512x512 RGB 4-connected grid cut, k-means K=5, lambda 50, int32"),
}


def hip_sync():
    hip = ctypes.CDLL("libamdhip64.so")
    rc = hip.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_graph(wl, seed=0, W=None, H=None):
    W = W or wl["W"]
    H = H or wl["H"]
    rgb, sd = jfcut.synth_image(W, H, nblobs=wl["nblobs"], noise=wl["noise"], seed=seed)
    g = jfcut.GridGraph.grid(W, H, 1, wl["conn"])
    g.model_image(rgb, sd, k_fg=wl["k"], k_bg=wl["k"], lambda_xy=wl["lam"])
    return g


class DeviceBuffer:
    """Raw HIP allocation (no torch needed); exposes data_ptr() like a tensor."""

    def __init__(self, nbytes):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.p = ctypes.c_void_p()
        rc = self.hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(max(1, nbytes)))
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed: {rc}")

    def data_ptr(self):
        return self.p.value

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = ctypes.c_void_p()


def batched_forward_backward(n=256, reps=3):
    """Config 5: the 256 uneven cuts (4C/8C, lambda 20-80) solved together and
    differentiated in one backward launch, against the same cuts solved and
    differentiated one graph at a time.  Flows must agree exactly."""
    specs = jfcut.batch_specs(n, 0)
    shapes = [(s["W"], s["H"], 1, s["conn"]) for s in specs]
    G = jfcut.GridGraph(shapes)
    singles = []
    for i, s in enumerate(specs):
        img, sd = jfcut.synth_image(s["W"], s["H"], nblobs=8, noise=25.0, seed=s["seed"])
        G.model_image(img, sd, inst=i, k_fg=5, k_bg=5, lambda_xy=s["lam"])
        g = jfcut.GridGraph([shapes[i]])
        g.model_image(img, sd, k_fg=5, k_bg=5, lambda_xy=s["lam"])
        singles.append(g)
    ntot, kmax = G.size()
    gb = [DeviceBuffer(ntot * 4), DeviceBuffer(ntot * 4), DeviceBuffer(ntot * kmax * 4)]
    maxn = max(s["W"] * s["H"] for s in specs)
    gs = [DeviceBuffer(maxn * 4), DeviceBuffer(maxn * 4), DeviceBuffer(maxn * 8 * 4)]
    # tile shape per size class, chosen by the autotuner (outside the timing)
    tile = G.autotune(reps=2)
    G.solve()
    for g in singles:
        g.solve()
    hip_sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        st = G.solve()
        G.backward_all(out=gb)
    hip_sync()
    t_batch = (time.perf_counter() - t0) / reps
    flows_b = G.flows()
    t0 = time.perf_counter()
    for _ in range(reps):
        for g in singles:
            g.solve()
            g.backward(out=gs)
    hip_sync()
    t_seq = (time.perf_counter() - t0) / reps
    flows_s = np.array([g.flow() for g in singles])
    assert np.array_equal(flows_b, flows_s), "batched and one-at-a-time flows differ"
    for g in singles:
        g.close()
    G.close()
    for b in gb + gs:
        b.free()
    return {"instances": n, "nodes": int(ntot),
            "batched_images_per_s": round(n / t_batch, 1),
            "sequential_images_per_s": round(n / t_seq, 1),
            "speedup": round(t_seq / t_batch, 2),
            "batched_mvoxels_per_s": round(ntot / t_batch / 1e6, 1),
            "batched_ms": round(t_batch * 1e3, 2), "sequential_ms": round(t_seq * 1e3, 2),
            "batched_sweeps": st.sweeps, "tile_shape": tile, "flows_equal": True}


def cpu_baseline(wl, sample_w, sample_h):
    """The reference's CPU max-flow (BK303 built from the reference sources)
    on a bounded sample of the workload, plus the GPU solve of the same sample
    for parity in the same run."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    if not ol.have_bk():
        return None
    g = build_graph(wl, seed=1, W=sample_w, H=sample_h)
    cs, ct = g.terminal_caps()
    caps = g.neighbor_caps()
    st = g.solve()
    gflow, glab = g.flow(), g.labels()
    g.close()
    t0 = time.perf_counter()
    flow, lab, (ti, tm, to) = ol.bk303(sample_w, sample_h, 1, wl["conn"], cs, ct, caps,
                                       timings=True)
    wall = time.perf_counter() - t0
    n = sample_w * sample_h
    return {
        "value": round(n / (ti + tm + to) / 1e6, 3),
        "unit": "Mvoxels/s",
        "cores": 1,
        "kind": "reference",
        "sample": f"BK303 (benchmark/GridCut/code/BK303, graph init + maxflow + labels) on a "
                  f"{sample_w}x{sample_h} instance of the same generator and data term "
                  f"({wl['conn']}-connected, seed 1)",
        "seconds": {"init": round(ti, 3), "maxflow": round(tm, 3), "output": round(to, 3),
                    "wall": round(wall, 3)},
        "gpu_same_sample_ms": round(st.ms_total, 3),
        "parity_flow": bool(flow == gflow),
        "parity_labels": bool(np.array_equal(lab, glab)),
    }


def fastcut_comparator(reps=5):
    """The paper's in-tree GPU comparison on this GPU: the same 4-connected
    512x512 cut (config 1) solved by the tile push-relabel (JF-Cut) and by
    the Fast-Cut line pushes (jfc_options.method = 1); device ms per solve,
    flows equal."""
    wl = WORKLOADS["c1"]
    g = build_graph(wl, seed=0)
    out = {"workload": wl["name"]}
    flows = {}
    for name, m in (("jfcut", 0), ("fastcut", 1)):
        g.solve(method=m)
        ms = [g.solve(method=m).ms_total for _ in range(reps)]
        flows[name] = g.flow()
        out[name + "_ms"] = round(float(np.median(ms)), 3)
    g.close()
    assert flows["jfcut"] == flows["fastcut"], flows
    out["jfcut_speedup"] = round(out["fastcut_ms"] / out["jfcut_ms"], 2)
    out["flows_equal"] = True
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-comparator", action="store_true")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB, the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--cpu-sample", default="7500x5000")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); multi-GPU "
              "slab sharding is a design note, not part of this build", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy_mib > 0:
        nb = args.calibrate_copy_mib << 20
        a, b = DeviceBuffer(nb), DeviceBuffer(nb)
        hip = ctypes.CDLL("libamdhip64.so")
        rc = hip.hipMemcpy(ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(a.data_ptr()),
                           ctypes.c_size_t(nb), 3)  # hipMemcpyDeviceToDevice
        hip_sync()
        a.free()
        b.free()
        log(f"calibration copy of {args.calibrate_copy_mib} MiB: rc {rc}")

    wl = WORKLOADS[args.workload]
    N = wl["W"] * wl["H"]
    t0 = time.time()
    g = build_graph(wl, seed=0)
    log(f"workload built in {time.time() - t0:.1f}s")
    for i in range(args.warmup):
        st = g.solve()
        log(f"warmup {i}: {st.ms_total:.1f} ms, sweeps {st.sweeps}, flow {st.flow}")
    flows, stats = [], []
    hip_sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = g.solve(profile=1)
        flows.append(st.flow)
        stats.append(st)
        log(f"step {i}: {st.ms_total:.1f} ms device, push {st.ms_push:.1f} ms, "
            f"sweeps {st.sweeps}, relabels {st.global_relabels}")
    hip_sync()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    assert len(set(flows)) == 1, flows
    assert all(s.converged == 1 for s in stats)
    energy = g.energy()
    assert energy == flows[0], (energy, flows[0])  # weak duality: cut == flow
    push_bytes = sum(s.push_bytes for s in stats)
    push_ms = sum(s.ms_push for s in stats)
    push_launches = sum(s.push_launches for s in stats)
    achieved = push_bytes / (push_ms * 1e-3) / 1e9 if push_ms > 0 else 0.0
    g.close()
    # HBM bytes per k_push dispatch from the calibrated rocprofv3 FETCH_SIZE /
    # WRITE_SIZE runs of this workload (profiles/k_push_c2_15000x10000_8c_traffic.md)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_push_c2_traffic.json")
    if args.workload == "c2" and os.path.exists(tpath):
        with open(tpath) as f:
            kt = [v for k, v in json.load(f)["kernels"].items() if "k_push" in k]
        if kt:
            traffic = kt[0]["read_bytes_per_dispatch"] + kt[0]["write_bytes_per_dispatch"]

    batched = None
    if not args.no_batched:
        batched = batched_forward_backward()
        log(f"batched fwd+bwd: {batched}")

    comparator = None
    if not args.no_comparator:
        comparator = fastcut_comparator()
        log(f"fast-cut comparator: {comparator}")

    cpu = None
    if not args.no_cpu_baseline:
        sw, sh = [int(v) for v in args.cpu_sample.split("x")]
        cpu = cpu_baseline(wl, sw, sh)

    line = {
        "metric": METRIC,
        "value": round(N / (ms_per_step * 1e-3) / 1e6, 3),
        "unit": "Mvoxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (deterministic generator, seed 0)",
        "config": {"workload": wl["name"], "nodes": N, "connectivity": wl["conn"],
                   "flow": int(flows[0]), "device_ms_per_step": round(
                       sum(s.ms_total for s in stats) / len(stats), 3),
                   "sweeps": stats[-1].sweeps, "global_relabels": stats[-1].global_relabels,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "kernel": "k_push (block-wise push-relabel)",
                     "bytes_per_launch": int(push_bytes / max(1, push_launches)),
                     "avg_launch_ms": round(push_ms / max(1, push_launches), 4)},
        "cpu_baseline": cpu,
        "batched_fwd_bwd": batched,
        "fastcut_comparator": comparator,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
