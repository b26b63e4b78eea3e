"""Writes tests/golden/config_goldens.json: full-size goldens of the benchmark
configurations 2-5, computed on the CPU by the oracle (oracle/jfc_oracle.c,
itself pinned against BK303 and scipy in tests/test_oracle.py) and, where the
graph fits this container's memory, cross-checked with BK303 built from the
reference sources (oracle/_ref).

For each configuration: sha256 of the synthetic image and seeds (libjfcut's
host generators), sha256 of the capacity planes from the oracle's integer
data term (source, sink, then the neighbour planes), the max-flow value, the
size of the canonical source set and the sha256 of the canonical labels.

Usage: python tests/golden/make_config_goldens.py [c2 c3 c4 c5 ...] [--bk c4]
Each configuration is independent; results are merged into the JSON file.
"""
import gc
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "jf-cut_amd"))
import jfcut  # noqa: E402
import oracle_lib as ol  # noqa: E402

OUT = os.path.join(HERE, "config_goldens.json")


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def synth(name, seed):
    c = jfcut.CONFIGS[name]
    if c["kind"] == "image":
        img, sd = jfcut.synth_image(c["W"], c["H"], nblobs=c["nblobs"], noise=c["noise"], seed=seed)
        shape = (c["W"], c["H"], 1)
    elif c["kind"] == "video":
        img, sd = jfcut.synth_video(c["W"], c["H"], c["T"], nblobs=c["nblobs"], noise=c["noise"],
                                    seed=seed)
        shape = (c["W"], c["H"], c["T"])
    else:
        img, sd = jfcut.synth_volume(c["W"], c["H"], c["D"], nell=c["nell"], seed=seed)
        shape = (c["W"], c["H"], c["D"])
    return c, img, sd, shape


def golden_one(name, seed=0, use_bk=False):
    t0 = time.time()
    c, img, sd, (nx, ny, nz) = synth(name, seed)
    rec = {"config": name, "seed": seed, "shape": [nx, ny, nz], "conn": c["conn"],
           "image_sha256": sha(img), "seeds_sha256": sha(sd)}
    cs, ct, caps, info = ol.model_image(img, sd, c["conn"], k=c["k"], lam=c["lam"],
                                        lam_z=c["lam_z"], z_sep=c["z_sep"], R=c["rng"])
    del img
    gc.collect()
    rec["caps_sha256"] = sha(cs, ct, caps)
    rec["centroids_fg"] = info["cfg"].tolist()
    rec["centroids_bg"] = info["cbg"].tolist()
    flow, lab, ok = ol.maxflow(nx, ny, nz, c["conn"], cs, ct, caps)
    assert ok == 1
    energy = ol.energy(nx, ny, nz, c["conn"], cs, ct, caps, lab)
    assert energy == flow, (energy, flow)
    rec.update(flow=int(flow), source_set=int(lab.sum()), labels_sha256=sha(lab))
    if use_bk:
        fb, lb = ol.bk303(nx, ny, nz, c["conn"], cs, ct, caps)
        assert fb == flow and np.array_equal(lb, lab)
        rec["checked_by"] = "oracle+bk303"
    else:
        rec["checked_by"] = "oracle"
    print(name, rec["shape"], flow, rec["source_set"], rec["checked_by"],
          f"{time.time() - t0:.1f}s", flush=True)
    return rec


def golden_batch(n=256, seed=0):
    t0 = time.time()
    recs = []
    for sp in jfcut.batch_specs(n, seed):
        img, sd = jfcut.synth_image(sp["W"], sp["H"], nblobs=8, noise=25.0, seed=sp["seed"])
        cs, ct, caps, info = ol.model_image(img, sd, sp["conn"], k=5, lam=sp["lam"])
        flow, lab, ok = ol.maxflow(sp["W"], sp["H"], 1, sp["conn"], cs, ct, caps)
        fb, lb = ol.bk303(sp["W"], sp["H"], 1, sp["conn"], cs, ct, caps)
        assert ok == 1 and fb == flow and np.array_equal(lb, lab)
        recs.append(dict(sp, flow=int(flow), source_set=int(lab.sum()), labels_sha256=sha(lab),
                         caps_sha256=sha(cs, ct, caps)))
    print("c5", n, "instances", sum(r["W"] * r["H"] for r in recs), "nodes",
          f"{time.time() - t0:.1f}s", flush=True)
    return {"config": "c5", "seed": seed, "instances": recs, "checked_by": "oracle+bk303"}


def main():
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    bk = set(sys.argv[sys.argv.index("--bk") + 1].split(",")) if "--bk" in sys.argv else set()
    data = {}
    if os.path.exists(OUT):
        data = json.load(open(OUT))
    for name in args or ["c2", "c3", "c4", "c5"]:
        data[name] = golden_batch() if name == "c5" else golden_one(name, use_bk=name in bk)
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1)
        gc.collect()


if __name__ == "__main__":
    main()
