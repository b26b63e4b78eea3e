"""Full-size benchmark configurations on the GPU against goldens computed on
the CPU by the oracle (and BK303 where the graph fits the build container):
same image, same int32 capacity planes, same int64 flow, same canonical
labels (sha256).  Plus the differentiable forward/backward of the 512^3
volume (preset c4) and the batched forward+backward of the 256 uneven cuts
(preset c5)."""
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


def _c4_graph(gpu, gd):
    g, img, sd = gpu.build_config("c4", seed=gd["seed"])
    del img, sd
    cs, ct = g.terminal_caps()
    caps = g.neighbor_caps()
    assert sha(cs, ct, caps) == gd["caps_sha256"]
    return g, cs, ct, caps


def _cut_arcs(lab, caps, shape, dirs):
    """(d, v) of every arc with capacity >= 1 from a source-side node to a
    sink-side node (the arcs whose gradient is 1)."""
    nx, ny, nz = shape
    L = lab.reshape(nz, ny, nx).astype(bool)
    out = []
    for d, (dx, dy, dz) in enumerate(dirs):
        sl = lambda a, b, n: slice(max(0, -a), n - max(0, a)) if b else slice(max(0, a), n - max(0, -a) or None)
        src = L[sl(dz, 1, nz), sl(dy, 1, ny), sl(dx, 1, nx)]
        dst = L[sl(dz, 0, nz), sl(dy, 0, ny), sl(dx, 0, nx)]
        m = np.zeros_like(L)
        m[sl(dz, 1, nz), sl(dy, 1, ny), sl(dx, 1, nx)] = src & ~dst
        v = np.flatnonzero(m.ravel() & (caps[d] >= 1))
        out.append(np.stack([np.full(v.size, d), v], 1))
    return np.concatenate(out)


def test_config4_finite_differences_full_size(gpu):
    """+-1 integer finite differences on the 512^3 volume itself: 400 arcs
    (half of them cut arcs), each E(w +- 1) an exact re-solve of all 1.3e8
    nodes, against the backward's cut-edge indicator; ties (E not
    differentiable at w) are skipped and counted."""
    torch = pytest.importorskip("torch")
    gd = goldens().get("c4")
    if gd is None:
        pytest.skip("no golden for c4")
    g, cs, ct, caps = _c4_graph(gpu, gd)
    nx, ny, nz = gd["shape"]
    N = nx * ny * nz
    opts = dict(max_local_iters=4)
    st = g.solve(**opts)
    E0 = g.flow()
    assert E0 == gd["flow"]
    lab = g.labels()
    dev = torch.device("cuda:0")
    dw = torch.empty((6, N), dtype=torch.float32, device=dev)
    g.set_stream(torch.cuda.current_stream().cuda_stream)
    g.backward(out=(None, None, dw))
    rng = np.random.default_rng(3)
    cut = _cut_arcs(lab, caps, (nx, ny, nz), gpu.DIRS_3D)
    anyarc = np.stack([rng.integers(0, 6, 4000), rng.integers(0, N, 4000)], 1)
    anyarc = anyarc[caps[anyarc[:, 0], anyarc[:, 1]] >= 1]
    pick = np.concatenate([cut[rng.choice(len(cut), 200, replace=False)], anyarc[:200]])
    grad = dw[torch.from_numpy(pick[:, 0]).to(dev), torch.from_numpy(pick[:, 1]).to(dev)]
    grad = grad.cpu().numpy()
    planes = torch.from_numpy(caps).to(dev)
    # the setters read the perturbed planes in torch's stream order
    g.set_stream(torch.cuda.current_stream().cuda_stream)
    checked = ties = ones = 0
    for i, (d, v) in enumerate(pick):
        d, v = int(d), int(v)
        E = []
        for delta in (1, -1):
            planes[d, v] += delta
            g.set_neighbor_caps(d, planes[d])
            g.solve(**opts)
            E.append(g.flow())
            planes[d, v] -= delta
        g.set_neighbor_caps(d, planes[d])
        up, dn = E[0] - E0, E0 - E[1]
        if i % 50 == 0:
            say("fd full size", i, (d, v), up, dn, float(grad[i]))
        if up != dn:
            ties += 1
            continue
        assert grad[i] == up, (d, v, up, float(grad[i]))
        checked += 1
        ones += up
    say("full-size finite differences checked", checked, "ties skipped", ties, "cut arcs", ones)
    assert checked >= 300 and ones >= 100
    g.close()


def test_config4_finite_differences(gpu):
    """+-1 integer finite differences on 10^4 non-tie arcs of crops of the
    512^3 volume's own capacities (preset c4): 22 crops of 32^3 nodes centred on arcs
    of the full solution's cut (arcs leaving a crop are dropped, so each crop
    is a graph of its own with its own min cut).  Each E(w +- 1) is an exact
    re-solve of the crop, 1024 perturbed crops solved per batched launch.
    (The full volume is checked on fewer arcs above: 2 x 10^4 re-solves of
    1.3e8 nodes would take minutes.)"""
    gd = goldens().get("c4")
    if gd is None:
        pytest.skip("no golden for c4")
    g, cs, ct, caps = _c4_graph(gpu, gd)
    nx, ny, nz = gd["shape"]
    g.solve(max_local_iters=4)
    lab = g.labels()
    g.close()
    rng = np.random.default_rng(0)
    cut = _cut_arcs(lab, caps, (nx, ny, nz), gpu.DIRS_3D)
    n3 = 32
    vol = lambda a: a.reshape(nz, ny, nx)
    crops = []
    for c in rng.choice(len(cut), 22, replace=False):
        v = int(cut[c, 1])
        z, y, x = v // (nx * ny), (v // nx) % ny, v % nx
        z0, y0, x0 = (min(max(0, q - n3 // 2), n - n3) for q, n in ((z, nz), (y, ny), (x, nx)))
        sl = (slice(z0, z0 + n3), slice(y0, y0 + n3), slice(x0, x0 + n3))
        cc = np.stack([vol(caps[d])[sl] for d in range(6)]).copy()
        # arcs leaving the crop are not part of its graph
        cc[0, :, :, -1] = 0; cc[1, :, :, 0] = 0
        cc[2, :, -1, :] = 0; cc[3, :, 0, :] = 0
        cc[4, -1] = 0; cc[5, 0] = 0
        crops.append((vol(cs)[sl].ravel().copy(), vol(ct)[sl].ravel().copy(),
                      cc.reshape(6, -1)))
    del cs, ct, caps, lab, cut
    shape = (n3, n3, n3, 6)
    N = n3 ** 3
    checked = ties = 0
    B = 512
    for ci, (ccs, cct, ccaps) in enumerate(crops):
        h = gpu.GridGraph([shape])
        h.set_terminal_caps(ccs, cct)
        for d in range(6):
            h.set_neighbor_caps(d, ccaps[d])
        h.solve()
        E0 = h.flow()
        _, _, grad = h.backward()
        h.close()
        allarc = np.stack([rng.integers(0, 6, 20000), rng.integers(0, N, 20000)], 1)
        allarc = allarc[ccaps[allarc[:, 0], allarc[:, 1]] >= 1]
        ccut = allarc[grad[allarc[:, 0], allarc[:, 1]] == 1]
        pick = np.concatenate([ccut[:B // 2], allarc[:B - min(len(ccut), B // 2)]])[:B]
        G = gpu.GridGraph([shape] * (2 * len(pick)))
        for i, (d, v) in enumerate(pick):
            for s, delta in enumerate((1, -1)):
                inst = 2 * i + s
                G.set_terminal_caps(ccs, cct, inst=inst)
                for dd in range(6):
                    p = ccaps[dd]
                    if dd == d:
                        p = p.copy()
                        p[v] += delta
                    G.set_neighbor_caps(dd, p, inst=inst)
        G.solve()
        fl = G.flows()
        G.close()
        for i, (d, v) in enumerate(pick):
            up, dn = fl[2 * i] - E0, E0 - fl[2 * i + 1]
            if up != dn:
                ties += 1
                continue
            assert grad[d][v] == up, (ci, d, v, up)
            checked += 1
        say("crop", ci, "checked so far", checked, "ties", ties)
    say("finite differences checked", checked, "ties skipped", ties)
    assert checked >= 10000
