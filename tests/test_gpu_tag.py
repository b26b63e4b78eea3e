"""GPU tag post-process (jfc_get_tags, graphcut_compute_tag.cl semantics)
against the reference's own tag channel of flower's .cut file and the oracle.
Bit-exact."""
import os
import sys

import numpy as np
import pytest

import oracle_lib as ol
from conftest import GOLDEN, load_ref_fixture, random_grid

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def flower_cut():
    z = np.load(os.path.join(GOLDEN, "ref_flower_cut.npz"))
    n = int(z["n"])
    return {k: np.unpackbits(z[k])[:n] for k in ("fg", "bg", "tag")}


@pytest.mark.parametrize("tile", ["default", "16x16", "64x4"])
def test_tag_flower_golden_gpu(gpu, tile):
    nx, ny, ef, cs, ct, caps, flow, lab = load_ref_fixture("flower")
    ch = flower_cut()
    say("case flower tag", tile)
    g = gpu.GridGraph([(nx, ny, 1, 4)])
    g.set_excess(ef)
    for d in range(4):
        g.set_neighbor_caps(d, caps[d])
    st = g.solve(tile_shape=tile)
    assert st.converged == 1 and g.flow() == flow
    assert np.array_equal(g.labels(), lab)
    assert np.array_equal(g.tags(ch["fg"]), ch["tag"])
    # a second call with other seeds reuses the plane
    fg2 = np.zeros_like(ch["fg"])
    assert not g.tags(fg2).any()
    g.close()


@pytest.mark.parametrize("conn,shape", [(4, (61, 47, 1)), (8, (300, 200, 1)), (6, (40, 33, 21))])
def test_tag_random_gpu(gpu, conn, shape):
    rng = np.random.default_rng(conn + shape[0])
    nx, ny, nz = shape
    for trial in range(3):
        say("case tag", conn, shape, trial)
        cs, ct, caps = random_grid(rng, nx, ny, nz, conn)
        g = gpu.GridGraph([(nx, ny, nz, conn)])
        g.set_terminal_caps(cs, ct)
        for d in range(conn):
            g.set_neighbor_caps(d, caps[d])
        assert g.solve().converged == 1
        lab = g.labels()
        fg = (rng.random(nx * ny * nz) < 0.003).astype(np.uint8)
        assert np.array_equal(g.tags(fg), ol.tag(nx, ny, nz, fg, lab))
        g.close()


def test_tag_batch_and_device_gpu(gpu):
    import torch
    rng = np.random.default_rng(5)
    shapes = [(70, 50, 1, 4), (33, 90, 1, 8), (120, 64, 1, 4)]
    data = [random_grid(rng, s[0], s[1], 1, s[3]) for s in shapes]
    g = gpu.GridGraph(shapes)
    for i, (cs, ct, caps) in enumerate(data):
        g.set_terminal_caps(cs, ct, inst=i)
        for d in range(shapes[i][3]):
            g.set_neighbor_caps(d, caps[d], inst=i)
    assert g.solve().converged == 1
    for i, s in enumerate(shapes):
        say("case tag batch", i)
        n = s[0] * s[1]
        fg = (rng.random(n) < 0.01).astype(np.uint8)
        want = ol.tag(s[0], s[1], 1, fg, g.labels(i))
        assert np.array_equal(g.tags(fg, inst=i), want), i
        dfg = torch.from_numpy(fg).cuda()
        out = torch.zeros(n, dtype=torch.uint8, device="cuda")
        g.tags(dfg, inst=i, out=out)
        assert np.array_equal(out.cpu().numpy(), want), i
    g.close()
