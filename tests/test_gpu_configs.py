"""Full-size benchmark configurations on the GPU against goldens computed on
the CPU by the oracle (and BK303 where the graph fits the build container):
same image, same int32 capacity planes, same int64 flow, same canonical
labels (sha256).  Plus the differentiable forward/backward of config 4 and
the batched forward+backward of config 5."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def goldens():
    path = os.path.join(GOLDEN, "config_goldens.json")
    if not os.path.exists(path):
        pytest.skip("config goldens not generated")
    return json.load(open(path))


@pytest.mark.parametrize("name", ["c2", "c3", "c4"])
def test_config_full_size(gpu, name):
    gd = goldens().get(name)
    if gd is None:
        pytest.skip(f"no golden for {name}")
    say("case", name, gd["shape"])
    g, img, sd = gpu.build_config(name, seed=gd["seed"])
    assert sha(img) == gd["image_sha256"]
    assert sha(sd) == gd["seeds_sha256"]
    del img
    cs, ct = g.terminal_caps()
    caps = g.neighbor_caps()
    assert sha(cs, ct, caps) == gd["caps_sha256"], "data term differs from the oracle"
    del cs, ct, caps
    st = g.solve()
    assert st.converged == 1
    assert g.flow() == gd["flow"]
    lab = g.labels()
    assert int(lab.sum()) == gd["source_set"]
    assert sha(lab) == gd["labels_sha256"]
    assert g.energy() == gd["flow"]
    say("ok", name, "flow", gd["flow"], "device ms", round(st.ms_total, 2))
    g.close()


def test_config5_batched_forward_backward(gpu):
    gd = goldens().get("c5")
    if gd is None:
        pytest.skip("no golden for c5")
    recs = gd["instances"]
    shapes = [(r["W"], r["H"], 1, r["conn"]) for r in recs]
    g = gpu.GridGraph(shapes)
    for i, r in enumerate(recs):
        img, sd = gpu.synth_image(r["W"], r["H"], nblobs=8, noise=25.0, seed=r["seed"])
        g.model_image(img, sd, inst=i, k_fg=5, k_bg=5, lambda_xy=r["lam"])
    st = g.solve()
    assert st.converged == 1
    flows = g.flows()
    dcs, dct, dnb = g.backward_all()
    off = 0
    for i, r in enumerate(recs):
        n = r["W"] * r["H"]
        assert flows[i] == r["flow"], i
        lab = g.labels(i)
        assert sha(lab) == r["labels_sha256"], i
        assert np.array_equal(dct[off:off + n], lab.astype(np.float32))
        assert np.array_equal(dcs[off:off + n], 1 - lab.astype(np.float32))
        off += n
    # the cut-edge planes agree with the per-instance backward
    for i in (0, len(recs) // 2, len(recs) - 1):
        a, b, c = g.backward(i)
        o = sum(r["W"] * r["H"] for r in recs[:i])
        n = recs[i]["W"] * recs[i]["H"]
        assert np.array_equal(c, dnb[:recs[i]["conn"], o:o + n])
    g.close()


def test_config4_autograd_full_size(gpu):
    """Forward energy and backward gradients of the 512^3 volume through the
    torch.autograd.Function; the gradient equals the one derived from the CPU
    cut (golden labels)."""
    torch = pytest.importorskip("torch")
    gd = goldens().get("c4")
    if gd is None:
        pytest.skip("no golden for c4")
    g, img, sd = gpu.build_config("c4", seed=gd["seed"])
    cs, ct = g.terminal_caps()
    caps = g.neighbor_caps()
    g.close()
    nx, ny, nz = gd["shape"]
    dev = torch.device("cuda:0")
    tcs = torch.from_numpy(cs.astype(np.float32)).to(dev).requires_grad_()
    tct = torch.from_numpy(ct.astype(np.float32)).to(dev).requires_grad_()
    tw = torch.from_numpy(caps.astype(np.float32)).to(dev).requires_grad_()
    del cs, ct
    E = gpu.min_cut_energy(tcs, tct, tw, (nx, ny, nz), 6)
    assert int(E.item()) == gd["flow"]
    E.backward()
    lab = tct.grad.to(torch.uint8).cpu().numpy()
    assert sha(lab) == gd["labels_sha256"]
    assert torch.equal(tcs.grad, 1 - tct.grad)
    # cut-edge indicator: 1[u in S and u+off(d) in T] from the golden labels
    L = torch.from_numpy(lab.reshape(nz, ny, nx)).to(dev).bool()
    for d, (dx, dy, dz) in enumerate(gpu.DIRS_3D):
        exp = torch.zeros_like(L)
        src = L[max(0, -dz):nz - max(0, dz), max(0, -dy):ny - max(0, dy), max(0, -dx):nx - max(0, dx)]
        dst = L[max(0, dz):nz - max(0, -dz) or None, max(0, dy):ny - max(0, -dy) or None,
                max(0, dx):nx - max(0, -dx) or None]
        exp[max(0, -dz):nz - max(0, dz), max(0, -dy):ny - max(0, dy),
            max(0, -dx):nx - max(0, dx)] = src & ~dst
        assert torch.equal(tw.grad[d].reshape(nz, ny, nx).bool(), exp), d


def test_config4_finite_differences(gpu):
    """+-1 integer finite differences of the min-cut energy on 10^4 non-tie
    arcs of a reduced config-4 volume (24^3, same generator and data term),
    each E(w+-1) an exact re-solve, batched 512 instances per launch."""
    c = gpu.CONFIGS["c4"]
    n3 = 24
    vol, sd = gpu.synth_volume(n3, n3, n3, nell=c["nell"], seed=5)
    g = gpu.GridGraph([(n3, n3, n3, 6)])
    g.model_image(vol, sd, k_fg=c["k"], k_bg=c["k"], rng=c["rng"], lambda_xy=c["lam"])
    cs, ct = g.terminal_caps()
    caps = g.neighbor_caps()
    g.solve()
    E0 = g.flow()
    lab = g.labels()
    _, _, grad = g.backward()
    g.close()
    N = n3 ** 3
    rng = np.random.default_rng(0)
    # half the sample on cut arcs (gradient 1), half anywhere
    idx = np.arange(N).reshape(n3, n3, n3)
    cand = []
    for d, (dx, dy, dz) in enumerate(gpu.DIRS_3D):
        src = idx[max(0, -dz):n3 - max(0, dz), max(0, -dy):n3 - max(0, dy), max(0, -dx):n3 - max(0, dx)].ravel()
        ok = caps[d][src] >= 1
        cand.append(np.stack([np.full(ok.sum(), d), src[ok]], 1))
    cand = np.concatenate(cand)
    cut = cand[grad[cand[:, 0], cand[:, 1]] == 1]
    pick = np.concatenate([cut[rng.integers(0, len(cut), 6500)],
                           cand[rng.integers(0, len(cand), 6500)]])
    B = 512
    checked = ties = 0
    for b0 in range(0, len(pick), B):
        chunk = pick[b0:b0 + B]
        G = gpu.GridGraph([(n3, n3, n3, 6)] * (2 * len(chunk)))
        for i, (d, v) in enumerate(chunk):
            for s, delta in enumerate((1, -1)):
                inst = 2 * i + s
                G.set_terminal_caps(cs, ct, inst=inst)
                for dd in range(6):
                    p = caps[dd]
                    if dd == d:
                        p = p.copy()
                        p[v] += delta
                    G.set_neighbor_caps(dd, p, inst=inst)
        G.solve()
        fl = G.flows()
        G.close()
        for i, (d, v) in enumerate(chunk):
            up, dn = fl[2 * i] - E0, E0 - fl[2 * i + 1]
            if up != dn:
                ties += 1
                continue
            assert grad[d][v] == up, (d, v, up)
            checked += 1
        if checked >= 10000:
            break
    say("finite differences checked", checked, "ties skipped", ties)
    assert checked >= 10000
