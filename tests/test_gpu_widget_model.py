"""GPU widget data terms (jfc_model_widget) against the CPU oracle
(oracle/widget_model.py): every capacity bit-exact, the same k-means
iterations and centroids, then the max-flow and canonical cut of the modelled
graph against the oracle solver.  Inputs: the reference's own images, seeds
and Modeling / Feature panel settings (tests/golden/ref_<name>_feature.npz),
and random 2D / 3D images."""
import os
import sys

import numpy as np
import pytest

import oracle_lib as ol
from conftest import GOLDEN, ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import widget_model as wm  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def reference_case(name):
    z = np.load(os.path.join(GOLDEN, f"ref_{name}_feature.npz"))
    nx, ny, n = int(z["nx"]), int(z["ny"]), int(z["n"])
    rgba = np.repeat(z["grey"][:, None], 4, axis=1)
    fg = np.unpackbits(z["fg"])[:n]
    bg = np.unpackbits(z["bg"])[:n]
    kw = dict(lazy=bool(z["lazy"]), k=int(z["k"]), kmeans_iterations=int(z["iterations"]),
              lazy_terminal=float(z["lazy_terminal"]), lazy_neighbor=float(z["lazy_neighbor"]),
              bk_terminal=float(z["bk_terminal"]), bk_neighbor=float(z["bk_neighbor"]),
              weight=tuple(float(v) for v in z["weight"]))
    return (nx, ny, 1), rgba, fg, bg, kw


def random_case(shape, seed):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    inside = ((xx - nx / 2) ** 2 / (nx / 3) ** 2 + (yy - ny / 2) ** 2 / (ny / 3) ** 2) < 1
    base = np.where(inside[..., None], [200, 80, 60, 255], [40, 90, 160, 255])
    img = np.clip(base + rng.normal(0, 25, base.shape), 0, 255).astype(np.uint8)
    fg = np.zeros((nz, ny, nx), np.uint8)
    bg = np.zeros((nz, ny, nx), np.uint8)
    fg[:, ny // 2 - 1:ny // 2 + 1, nx // 3:2 * nx // 3] = 1
    bg[:, 1:3, 2:nx - 2] = 1
    return img.reshape(-1, 4), fg.ravel(), bg.ravel()


def check(gpu, shape, rgba, fg, bg, kw, solve=True):
    nx, ny, nz = shape
    conn = 6 if nz > 1 else 4
    cs, ct, planes, info = wm.model(rgba, fg, bg, shape, **kw)
    g = gpu.GridGraph([(nx, ny, nz, conn)])
    st = g.model_widget(rgba, fg, bg, **kw)
    gcs, gct = g.terminal_caps()
    gpl = g.neighbor_caps()
    assert np.array_equal(gcs, cs), np.flatnonzero(gcs != cs)[:10]
    assert np.array_equal(gct, ct), np.flatnonzero(gct != ct)[:10]
    assert np.array_equal(gpl.reshape(planes.shape), planes)
    assert st.seeds[0] == int((fg != 0).sum()) and st.seeds[1] == int((bg != 0).sum())
    if kw.get("lazy", True):
        k = kw.get("k", 8)
        assert tuple(st.iterations) == info["iterations"]
        for side in range(2):
            got = np.array([[st.centroids[side][i][c] for c in range(4)] for i in range(k)],
                           np.float32)
            assert np.array_equal(got, info["centroids"][side])
    else:
        assert st.hist_sum[0] == int(info["hist"][0].sum())
        assert st.hist_sum[1] == int(info["hist"][1].sum())
    if solve:
        flow, lab, ok = ol.maxflow(nx, ny, nz, conn, cs, ct, planes)
        g.solve()
        assert g.flow() == flow
        assert np.array_equal(g.labels(), lab)
    g.close()


@pytest.mark.parametrize("name", ["flower", "person", "sponge"])
def test_widget_model_reference_images_gpu(gpu, name):
    shape, rgba, fg, bg, kw = reference_case(name)
    say("case widget model", name, kw)
    check(gpu, shape, rgba, fg, bg, kw)


def test_widget_model_bk_flower_gpu(gpu):
    shape, rgba, fg, bg, kw = reference_case("flower")
    kw.update(lazy=False, bk_terminal=1.0)
    say("case widget model bk flower", kw)
    check(gpu, shape, rgba, fg, bg, kw)


@pytest.mark.parametrize("shape,lazy,seed", [((97, 61, 1), True, 0), ((97, 61, 1), False, 1),
                                             ((33, 21, 9), True, 2), ((33, 21, 9), False, 3)])
def test_widget_model_random_gpu(gpu, shape, lazy, seed):
    rgba, fg, bg = random_case(shape, seed)
    kw = dict(lazy=lazy, weight=(0.3, 0.5, 0.2, 0.1), k=5, kmeans_iterations=40,
              lazy_terminal=2.0, lazy_neighbor=3.0, bk_terminal=0.7, bk_neighbor=0.05)
    say("case widget model random", shape, lazy, seed)
    check(gpu, shape, rgba, fg, bg, kw)


def test_widget_model_device_inputs_and_batch_gpu(gpu):
    import torch
    rgba, fg, bg = random_case((40, 30, 1), 7)
    kw = dict(k=4, kmeans_iterations=20)
    cs, ct, planes, _ = wm.model(rgba, fg, bg, (40, 30, 1), **kw)
    g = gpu.GridGraph([(17, 9, 1, 8), (40, 30, 1, 4)])
    dev = [torch.from_numpy(a).cuda() for a in (rgba, fg, bg)]
    g.model_widget(*dev, inst=1, **kw)
    gcs, gct = g.terminal_caps(inst=1)
    assert np.array_equal(gcs, cs) and np.array_equal(gct, ct)
    assert np.array_equal(g.neighbor_caps(inst=1).reshape(planes.shape), planes)
    # the 8-connected instance has no widget data term
    with pytest.raises(gpu.JFCError):
        g.model_widget(np.zeros((17 * 9, 4), np.uint8), np.ones(17 * 9, np.uint8),
                       np.ones(17 * 9, np.uint8), inst=0)
    g.close()


def test_widget_model_errors_gpu(gpu):
    rgba, fg, bg = random_case((20, 10, 1), 9)
    g = gpu.GridGraph([(20, 10, 1, 4)])
    with pytest.raises(gpu.JFCError) as e:
        g.model_widget(rgba, np.zeros_like(fg), bg)
    assert e.value.code == gpu.ERR_NO_SEEDS
    with pytest.raises(gpu.JFCError):
        g.model_widget(rgba, fg, bg, k=1)
    with pytest.raises(gpu.JFCError):
        g.model_widget(rgba, fg, bg, lazy=False, bk_neighbor=0.0)
    g.close()
