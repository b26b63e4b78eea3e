"""Seeded random sweep on the GPU: random grid sizes, connectivities,
capacity ranges, seed densities and solver options (tile shape, local
iteration cap, initial relabel, tile ordering, graph replay, relabel budget,
Fast-Cut method), single graphs solved twice and uneven batches, each
against the oracle's exact flow and canonical labels."""
import sys

import numpy as np
import pytest

import oracle_lib as ol
from conftest import random_grid

pytestmark = pytest.mark.gpu

SHAPES_2D = ["32x8", "64x4", "16x16"]
SHAPES_3D = ["8x8x4", "32x4x2"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def random_options(rng, d3, conn):
    opts = dict(tile_shape=str(rng.choice(SHAPES_3D if d3 else SHAPES_2D)),
                max_local_iters=int(rng.integers(0, 5)), initial_relabel=int(rng.integers(-1, 2)),
                tile_order=int(rng.integers(0, 2)), graph_sweeps=int(rng.integers(0, 2)),
                relabel_work=float(rng.choice([0.01, 0.5, 2.0])))
    if conn != 8 and rng.random() < 0.2:
        opts["method"] = 1
    return opts


@pytest.mark.parametrize("seed", [101, 202, 303])
def test_random_graphs_and_options_gpu(gpu, seed):
    rng = np.random.default_rng(seed)
    for case in range(40):
        d3 = rng.random() < 0.35
        if d3:
            nx, ny, nz, conn = int(rng.integers(9, 60)), int(rng.integers(9, 40)), int(rng.integers(2, 16)), 6
        else:
            nx, ny, nz, conn = int(rng.integers(10, 260)), int(rng.integers(10, 180)), 1, int(rng.choice([4, 8]))
        cs, ct, caps = random_grid(rng, nx, ny, nz, conn, cap_max=int(rng.choice([5, 50, 1000])),
                                   term_max=int(rng.choice([10, 300, 5000])),
                                   seed_frac=float(rng.choice([0.0, 0.02, 0.1])))
        f, l, ok = ol.maxflow(nx, ny, nz, conn, cs, ct, caps)
        opts = random_options(rng, d3, conn)
        say("case random", seed, case, (nx, ny, nz, conn), opts)
        g = gpu.GridGraph([(nx, ny, nz, conn)])
        g.set_terminal_caps(cs, ct)
        for d in range(conn):
            g.set_neighbor_caps(d, caps[d])
        for rep in range(2):  # the second solve starts from the first's state
            st = g.solve(**opts)
            assert st.converged == 1
            assert g.flow() == f
            assert np.array_equal(g.labels(), l)
        g.close()
        if case % 8 == 3:
            k = int(rng.integers(2, 7))
            if d3:
                shp = [(int(rng.integers(9, 36)), int(rng.integers(9, 36)), int(rng.integers(2, 9)), 6)
                       for _ in range(k)]
            else:
                shp = [(int(rng.integers(10, 130)), int(rng.integers(10, 130)), 1, int(rng.choice([4, 8])))
                       for _ in range(k)]
            data = [random_grid(rng, a, b, c, dd) for (a, b, c, dd) in shp]
            exp = [ol.maxflow(a, b, c, dd, *data[i])[:2] for i, (a, b, c, dd) in enumerate(shp)]
            bo = dict(opts)
            bo.pop("method", None)
            say("case random batch", seed, case, shp, bo)
            G = gpu.GridGraph(shp)
            for i, (cs_, ct_, caps_) in enumerate(data):
                G.set_terminal_caps(cs_, ct_, inst=i)
                for d in range(shp[i][3]):
                    G.set_neighbor_caps(d, caps_[d], inst=i)
            st = G.solve(**bo)
            assert st.converged == 1
            for i, (ff, ll) in enumerate(exp):
                assert G.flow(i) == ff, i
                assert np.array_equal(G.labels(i), ll), i
            G.close()
