#!/usr/bin/env python3
"""Benchmark: Mvoxels/s to exact max-flow on one MI355X.

Default workload (the large 2D cut): 15000x10000 RGB uint8 image from the
synthetic generator (40 colour blobs, noise sigma 25, scribble seeds of
capacity 2^30), integer k-means data term (K=5 per side), 8-connected grid with
pairwise round(50 exp(-|dI|^2 / 2 s2)) (diagonals / sqrt 2), int32 capacities.

A step is one full solve of freshly written inputs: the capacities are
resident in HBM (written in place, jfc_inputs_written), and the timed solve
rebuilds everything derived from them (residual-arc masks, folded excess) and
runs to the exact int64 flow and the canonical labels in HBM, as a caller
runs it (no timing events between launches).  The push kernel's launch
durations come from a second timed run of the same fresh-input solve with a
HIP event pair around every push launch.  Re-solves of unchanged inputs (the
first of them writes the input mask words, later ones copy them) and a step
that also copies the inputs in from caller buffers are reported beside it.

Prints one JSON line.  --gpus N > 1 exits: this build runs on one GPU.
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "jf-cut_amd"))

import numpy as np  # noqa: E402

import jfcut  # noqa: E402

METRIC = "Mvoxels/s to exact max-flow on 1 GPU; HBM GB/s; batched fwd+bwd vs plain"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "c2": dict(W=15000, H=10000, conn=8, nblobs=40, noise=25.0, lam=50.0, k=5, unary=100,
               name="15000x10000 RGB 8-connected grid cut, k-means K=5, lambda 50, int32"),
    "c1": dict(W=512, H=512, conn=4, nblobs=8, noise=25.0, lam=50.0, k=5, unary=100,
               name="512x512 RGB 4-connected grid cut, k-means K=5, lambda 50, int32"),
}
# the published max-flows of the reference's own graphs
# (latex/sources/section-5-results.tex:31,36,41) and the reference's JF-Cut
# total for flower on a GTX TITAN (benchmark/performance.xlsx, sheet CPU, O3)
FIXTURE_FLOWS = {"flower": 1208089, "person": 880461, "sponge": 343591}
FLOWER_TITAN_MS = 5.70


def hip_sync():
    hip = ctypes.CDLL("libamdhip64.so")
    rc = hip.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def data_term_ms(wl, reps=3):
    """The data term of the workload (integer k-means per side, unary and
    pairwise capacities, jfc_model_image) with the image and seeds resident
    in HBM: wall ms per call, the call's host waits included."""
    W, H = wl["W"], wl["H"]
    rgb, sd = jfcut.synth_image(W, H, nblobs=wl["nblobs"], noise=wl["noise"], seed=0)
    L = jfcut.lib()
    hip = ctypes.CDLL("libamdhip64.so")
    bi, bs = DeviceBuffer(rgb.nbytes), DeviceBuffer(sd.nbytes)
    hip.hipMemcpy(ctypes.c_void_p(bi.data_ptr()), rgb.ctypes.data_as(ctypes.c_void_p),
                  ctypes.c_size_t(rgb.nbytes), 1)
    hip.hipMemcpy(ctypes.c_void_p(bs.data_ptr()), sd.ctypes.data_as(ctypes.c_void_p),
                  ctypes.c_size_t(sd.nbytes), 1)
    g = jfcut.GridGraph.grid(W, H, 1, wl["conn"])
    prm = jfcut.ModelParams(wl["k"], wl["k"], 32, wl["unary"], 255, float(wl["lam"]),
                            float(wl["lam"]), 0)
    ms = []
    for i in range(reps + 1):
        hip_sync()
        t0 = time.perf_counter()
        jfcut._check(L.jfc_model_image(g._h, 0, ctypes.c_void_p(bi.data_ptr()), 3, 1,
                                       ctypes.c_void_p(bs.data_ptr()), 1, ctypes.byref(prm), None))
        hip_sync()
        if i:
            ms.append((time.perf_counter() - t0) * 1e3)
    g.close()
    bi.free()
    bs.free()
    return round(float(np.median(ms)), 3)


def build_graph(wl, seed=0, W=None, H=None):
    W = W or wl["W"]
    H = H or wl["H"]
    rgb, sd = jfcut.synth_image(W, H, nblobs=wl["nblobs"], noise=wl["noise"], seed=seed)
    g = jfcut.GridGraph.grid(W, H, 1, wl["conn"])
    g.model_image(rgb, sd, k_fg=wl["k"], k_bg=wl["k"], lambda_xy=wl["lam"],
                  unary_scale=wl["unary"])
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


def timed_steps(g, steps, fresh, setter=None, **opt):
    """Time `steps` solves (wall clock, device synchronised on both sides).
    fresh: mark the inputs written before each solve; setter: callable that
    copies the inputs in before each solve (inside the timed region)."""
    stats = []
    hip_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        if setter is not None:
            setter()
        elif fresh:
            g.inputs_written()
        st = g.solve(**opt)
        stats.append(st)
    hip_sync()
    ms = (time.perf_counter() - t0) / steps * 1e3
    return ms, stats


def device_setter(g, n, K):
    """Copies of the graph's inputs in device buffers and a setter that
    copies them back in through the C-ABI (device pointers), validation
    included."""
    L = jfcut.lib()
    bufs = [DeviceBuffer(n * 4), DeviceBuffer(n * 4)] + [DeviceBuffer(n * 4) for _ in range(K)]
    vp = lambda b: ctypes.c_void_p(b.data_ptr())
    jfcut._check(L.jfc_get_terminal_caps(g._h, 0, vp(bufs[0]), vp(bufs[1]), 1))
    for d in range(K):
        jfcut._check(L.jfc_get_neighbor_caps(g._h, 0, d, vp(bufs[2 + d]), 1))

    def setter():
        jfcut._check(L.jfc_set_terminal_caps(g._h, 0, vp(bufs[0]), vp(bufs[1]), 1))
        for d in range(K):
            jfcut._check(L.jfc_set_neighbor_caps(g._h, 0, d, vp(bufs[2 + d]), 1))

    return setter, bufs


def batched_forward_backward(n=256, reps=3):
    """The 256 uneven cuts of preset c5 (4C/8C, lambda 20-80) solved together and
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
    L = jfcut.lib()
    vp = lambda b: ctypes.c_void_p(b.data_ptr())
    # tile shape per size class, chosen by the autotuner (outside the timing)
    tile = G.autotune(reps=2)
    G.solve()
    for g in singles:
        g.solve()
    hip_sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        st = G.solve()
        jfcut._check(L.jfc_backward_all(G._h, vp(gb[0]), vp(gb[1]), vp(gb[2]), 1))
    hip_sync()
    t_batch = (time.perf_counter() - t0) / reps
    flows_b = G.flows()
    t0 = time.perf_counter()
    for _ in range(reps):
        for g in singles:
            g.solve()
            jfcut._check(L.jfc_backward(g._h, 0, vp(gs[0]), vp(gs[1]), vp(gs[2]), 1))
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


def host_info():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    avail = 0
    try:
        for line in open("/proc/meminfo"):
            if line.startswith("MemAvailable"):
                avail = int(line.split()[1]) * 1024
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model or platform.processor(),
            "mem_available_gb": round(avail / 2**30, 1)}


def cpu_baseline(wl, gflow, glab, cs, ct, caps, want_full):
    """The reference's CPU max-flow (BK303 built from the reference sources,
    one thread) on the workload itself when host memory allows (BK303 needs
    about 300 bytes per node on an 8-connected grid), else on a bounded
    sample of it; flow and labels compared with the GPU's in the same run."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    if not ol.have_bk():
        return None
    host = host_info()
    W, H = wl["W"], wl["H"]
    need_gb = W * H * 320 / 2**30 + 8
    full = want_full and host["mem_available_gb"] >= need_gb
    reason = None
    if full:
        n = W * H
        sample = f"the full {W}x{H} workload (the same capacities as the GPU solve)"
        t0 = time.perf_counter()
        flow, lab, (ti, tm, to) = ol.bk303(W, H, 1, wl["conn"], cs, ct, caps, timings=True)
        wall = time.perf_counter() - t0
        gpu_sample_ms = None
        parity = (bool(flow == gflow), bool(np.array_equal(lab, glab)))
    else:
        reason = (f"host memory: {host['mem_available_gb']} GB available, BK303 needs about "
                  f"{need_gb:.0f} GB for the full graph") if want_full else "sample requested"
        sw, sh = W // 2, H // 2
        n = sw * sh
        g = build_graph(wl, seed=1, W=sw, H=sh)
        scs, sct = g.terminal_caps()
        scaps = g.neighbor_caps()
        st = g.solve()
        sflow, slab = g.flow(), g.labels()
        g.close()
        sample = f"a {sw}x{sh} instance of the same generator and data term (seed 1)"
        t0 = time.perf_counter()
        flow, lab, (ti, tm, to) = ol.bk303(sw, sh, 1, wl["conn"], scs, sct, scaps, timings=True)
        wall = time.perf_counter() - t0
        gpu_sample_ms = round(st.ms_total, 3)
        parity = (bool(flow == sflow), bool(np.array_equal(lab, slab)))
    out = {
        "value": round(n / (ti + tm + to) / 1e6, 3),
        "unit": "Mvoxels/s",
        "cores": 1,
        "kind": "reference",
        "sample": "BK303 (benchmark/GridCut/code/BK303, single thread: graph init + maxflow + "
                  "labels, the phases of bench-BK303.cpp:34-57) on " + sample,
        "seconds": {"init": round(ti, 3), "maxflow": round(tm, 3), "output": round(to, 3),
                    "wall": round(wall, 3)},
        "host": host,
        "parity_flow": parity[0],
        "parity_labels": parity[1],
    }
    if reason:
        out["not_full_reason"] = reason
        out["gpu_same_sample_ms"] = gpu_sample_ms
    return out


def fixtures_leg(reps=5):
    """The reference's own graphs (flower 600x450, person 600x450, sponge
    640x480, 4-connected; tests/golden/ref_*.npz hold the reference's
    .ef/.c1/.c2 arrays): GPU ms per fresh solve and BK303 ms on the same
    graphs in the same run, flows equal to the published values.  The three
    graphs share one size class, autotuned once on flower (outside the
    timing) as a caller would for its image size."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    out = {"published_jfcut_flower_ms_gtx_titan": FLOWER_TITAN_MS}
    tuned = None
    for name, flow in FIXTURE_FLOWS.items():
        z = np.load(os.path.join(ROOT, "tests", "golden", f"ref_{name}.npz"))
        nx, ny = int(z["nx"]), int(z["ny"])
        ef = z["ef"].astype(np.int32)
        caps = z["caps"].astype(np.int32)
        g = jfcut.GridGraph([(nx, ny, 1, 4)])
        g.set_excess(ef)
        for d in range(4):
            g.set_neighbor_caps(d, caps[d])
        if tuned is None:
            g.autotune(reps=2)
            tuned = g.tuning()
            out["tuning"] = tuned
        g.solve()
        ms, stats = timed_steps(g, reps, fresh=True)
        assert g.flow() == flow, (name, g.flow(), flow)
        dev = float(np.median([s.ms_total for s in stats]))
        g.close()
        rec = {"nodes": nx * ny, "gpu_ms": round(ms, 3), "gpu_device_ms": round(dev, 3),
               "sweeps": stats[-1].sweeps, "global_relabels": stats[-1].global_relabels,
               "flow": flow, "flow_equal_published": True}
        if ol.have_bk():
            cs, ct = np.maximum(ef, 0), np.maximum(-ef, 0)
            bt = []
            for _ in range(3):
                f, lab, t = ol.bk303(nx, ny, 1, 4, cs, ct, caps, timings=True)
                bt.append(sum(t))
            assert f == flow
            rec["bk303_ms_1_core"] = round(min(bt) * 1e3, 2)
        out[name] = rec
    return out


def config_leg(name, steps=3):
    """A 3D configuration (video volume c3 or 512^3 volume c4): autotuned,
    then fresh-input solves; Mvoxels/s and the flow (checked equal to the cut
    energy)."""
    t0 = time.time()
    g, img, sd = jfcut.build_config(name, seed=0)
    del img, sd
    c = jfcut.CONFIGS[name]
    n = c["W"] * c["H"] * c.get("T", c.get("D", 1))
    g.autotune(reps=1)
    ms, stats = timed_steps(g, steps, fresh=True)
    flows = {s.flow for s in stats}
    assert len(flows) == 1 and all(s.converged == 1 for s in stats)
    assert g.energy() == stats[0].flow
    st = stats[-1]
    out = {"nodes": n, "ms_per_solve": round(ms, 3),
           "mvoxels_per_s": round(n / (ms * 1e-3) / 1e6, 1),
           "flow": int(stats[0].flow), "tuning": g.tuning(), "sweeps": st.sweeps,
           "global_relabels": st.global_relabels, "relabel_passes": st.relabel_passes,
           "ms_relabel": round(st.ms_relabel, 3),
           "whole_solve_frac": round(st.solve_bytes / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
           "build_s": round(time.time() - t0, 1)}
    g.close()
    return out


def seeds_only_leg(W, H, steps=2):
    """A harder 2D cut: the 15000x10000 workload's image and pairwise terms with hard seeds
    only (unary scale 0), so flow has to travel from the foreground scribbles
    across the image to the background scribbles."""
    wl = dict(WORKLOADS["c2"], W=W, H=H, unary=0)
    g = build_graph(wl, seed=0)
    g.solve()
    ms, stats = timed_steps(g, steps, fresh=True)
    st = stats[-1]
    assert all(s.converged == 1 for s in stats) and len({s.flow for s in stats}) == 1
    assert g.energy() == st.flow
    out = {"workload": f"{W}x{H} RGB 8-connected, seeds only (unary scale 0), lambda 50",
           "nodes": W * H, "ms_per_solve": round(ms, 3),
           "mvoxels_per_s": round(W * H / (ms * 1e-3) / 1e6, 1), "flow": int(st.flow),
           "sweeps": st.sweeps, "global_relabels": st.global_relabels,
           "push_tiles": int(st.push_tiles),
           "whole_solve_frac": round(st.solve_bytes / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)}
    g.close()
    return out


def fastcut_comparator(reps=5):
    """The paper's in-tree GPU comparison on this GPU: the same 4-connected
    512x512 cut (preset c1) solved by the tile push-relabel (JF-Cut) and by
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
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", action="store_true",
                    help="time BK303 on a quarter-size sample instead of the full workload")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-comparator", action="store_true")
    ap.add_argument("--no-3d", action="store_true")
    ap.add_argument("--no-fixtures", action="store_true")
    ap.add_argument("--no-seeds-only", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tuning", default="",
                    help="tile_shape,max_local_iters,initial_relabel,stall_certify,"
                         "sparse_local_iters: solve the headline workload with these options "
                         "instead of autotuning (profiling runs repeat a bench run's choice)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB, the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE counters")
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
    K = wl["conn"]
    t0 = time.time()
    g = build_graph(wl, seed=0)
    log(f"workload built in {time.time() - t0:.1f}s")
    topt = {}
    if args.tuning:
        sh, mli, ir, sc, sp = args.tuning.split(",")
        topt = dict(tile_shape=sh, max_local_iters=int(mli), initial_relabel=int(ir),
                    stall_certify=int(sc), sparse_local_iters=int(sp))
        tuning = dict(topt)
    else:
        if not args.no_autotune:
            g.autotune(reps=2)
        tuning = g.tuning()
    log(f"tuning: {tuning}")
    for i in range(args.warmup):
        g.inputs_written()
        st = g.solve(**topt)
        log(f"warmup {i}: {st.ms_total:.2f} ms, sweeps {st.sweeps}, flow {st.flow}")

    # headline: fresh inputs, the solve as a caller runs it (sweeps after the
    # first replay a captured graph, no timing events between launches)
    ms_per_step, stats = timed_steps(g, args.steps, fresh=True, **topt)
    for i, st in enumerate(stats):
        log(f"step {i}: {st.ms_total:.2f} ms device, sweeps {st.sweeps}, "
            f"relabels {st.global_relabels}, fresh {st.inputs_fresh}")
    flows = [s.flow for s in stats]
    assert len(set(flows)) == 1, flows
    assert all(s.converged == 1 and s.inputs_fresh == 1 for s in stats)
    energy = g.energy()
    assert energy == flows[0], (energy, flows[0])  # weak duality: cut == flow
    solve_bytes = sum(s.solve_bytes for s in stats) / len(stats)
    dev_ms = sum(s.ms_total for s in stats) / len(stats)

    # the same fresh-input solve with every push launch timed by HIP events on
    # the solve stream (profile = 1: sweeps launched eagerly, an event pair
    # around each push launch); the push roofline comes from these launches
    ev_ms, ev_stats = timed_steps(g, args.steps, fresh=True, profile=1, **topt)
    for i, st in enumerate(ev_stats):
        log(f"timed-launch step {i}: {st.ms_total:.2f} ms device, push {st.ms_push:.2f} ms, "
            f"sweeps {st.sweeps}")
    assert all(s.flow == flows[0] and s.converged == 1 for s in ev_stats)
    push_bytes = sum(s.push_bytes for s in ev_stats)
    push_ms = sum(s.ms_push for s in ev_stats)
    push_launches = sum(s.push_launches for s in ev_stats)
    achieved = push_bytes / (push_ms * 1e-3) / 1e9 if push_ms > 0 else 0.0
    # re-solves of unchanged inputs: the first writes mask0 (a fresh solve
    # does not), the others copy it
    warm_ms, wst = timed_steps(g, args.steps, fresh=False, **topt)
    assert all(s.inputs_fresh == 0 for s in wst)
    # inputs copied in from caller device buffers (validation included)
    setter, bufs = device_setter(g, N, K)
    set_ms, _ = timed_steps(g, max(2, args.steps // 2), fresh=False, setter=setter, **topt)
    for b in bufs:
        b.free()
    log(f"fresh {ms_per_step:.3f} ms (with push events {ev_ms:.3f}), warm {warm_ms:.3f} ms, "
        f"with device setters {set_ms:.3f} ms")
    glab = g.labels()
    cs = ct = caps = None
    if not args.no_cpu_baseline:
        cs, ct = g.terminal_caps()
        caps = g.neighbor_caps()
    g.close()

    # HBM bytes per k_push dispatch, from the calibrated rocprofv3 FETCH_SIZE /
    # WRITE_SIZE runs of this workload committed under profiles/ (counters are
    # not readable from inside this process)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_push_c2_traffic.json")
    if args.workload == "c2" and os.path.exists(tpath):
        with open(tpath) as f:
            kt = [v for k, v in json.load(f)["kernels"].items() if "k_push" in k]
        if kt:
            traffic = kt[0]["read_bytes_per_dispatch"] + kt[0]["write_bytes_per_dispatch"]

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(wl, flows[0], glab, cs, ct, caps, want_full=not args.cpu_sample)
        log(f"cpu baseline: {cpu}")
        del cs, ct, caps

    fixtures = None
    if not args.no_fixtures:
        fixtures = fixtures_leg()
        log(f"fixtures: {fixtures}")

    dterm = data_term_ms(wl)
    log(f"data term: {dterm} ms")

    seeds_only = None
    if not args.no_seeds_only and args.workload == "c2":
        seeds_only = seeds_only_leg(wl["W"] // 4, wl["H"] // 4)
        log(f"seeds only: {seeds_only}")

    volumes = None
    if not args.no_3d:
        volumes = {name: config_leg(name) for name in ("c3", "c4")}
        log(f"3D configs: {volumes}")

    batched = None
    if not args.no_batched:
        batched = batched_forward_backward()
        log(f"batched fwd+bwd: {batched}")

    comparator = None
    if not args.no_comparator:
        comparator = fastcut_comparator()
        log(f"fast-cut comparator: {comparator}")

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
        "config": {"workload": wl["name"], "nodes": N, "connectivity": K,
                   "step": "fresh inputs: capacities written in HBM, every derived "
                           "structure rebuilt inside the timed solve",
                   "flow": int(flows[0]), "device_ms_per_step": round(dev_ms, 3),
                   "ms_with_push_events": round(ev_ms, 3),
                   "ms_warm_resolve": round(warm_ms, 3),
                   "ms_with_device_setters": round(set_ms, 3),
                   "mvoxels_per_s_warm": round(N / (warm_ms * 1e-3) / 1e6, 1),
                   "data_term_ms": dterm,
                   "tuning": tuning, "sweeps": stats[-1].sweeps,
                   "global_relabels": stats[-1].global_relabels,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "traffic_source": "profiles/k_push_c2_traffic.json (rocprofv3 --pmc "
                                       "FETCH_SIZE / WRITE_SIZE of this workload, calibrated)",
                     "kernel": "k_push (block-wise push-relabel)",
                     "bytes_per_launch": int(push_bytes / max(1, push_launches)),
                     "avg_launch_ms": round(push_ms / max(1, push_launches), 4),
                     "whole_solve": {
                         "bytes": int(solve_bytes),
                         "achieved": round(solve_bytes / (ms_per_step * 1e-3) / 1e9, 1),
                         "frac": round(solve_bytes / (ms_per_step * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)}},
        "cpu_baseline": cpu,
        "fixtures": fixtures,
        "seeds_only_2d": seeds_only,
        "volumes_3d": volumes,
        "batched_fwd_bwd": batched,
        "fastcut_comparator": comparator,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
