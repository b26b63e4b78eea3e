"""GPU Fast-Cut comparator (jfc_options.method = JFC_METHOD_FASTCUT: the
widget's relabelPushFast, QGCWidget.cpp:2156-2205, with
graphcut_push_relabel_fast.cl) against the oracle and the reference's own
fixtures: the same exact flow and canonical labels as the tile solver."""
import sys

import numpy as np
import pytest

import oracle_lib as ol
from conftest import load_ref_fixture, random_grid

pytestmark = pytest.mark.gpu
FAST = 1


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("name", ["flower", "person", "sponge"])
def test_fastcut_reference_fixtures_gpu(gpu, name):
    nx, ny, ef, cs, ct, caps, flow, lab = load_ref_fixture(name)
    say("case fastcut", name)
    g = gpu.GridGraph([(nx, ny, 1, 4)])
    g.set_excess(ef)
    for d in range(4):
        g.set_neighbor_caps(d, caps[d])
    st = g.solve(method=FAST)
    assert st.converged == 1 and g.flow() == flow
    assert np.array_equal(g.labels(), lab)
    # the tile solver on the same graph object afterwards: same answer
    st2 = g.solve()
    assert g.flow() == flow and np.array_equal(g.labels(), lab)
    g.close()


@pytest.mark.parametrize("shape,seed", [((97, 83, 1), 0), ((256, 40, 1), 1), ((33, 29, 17), 2),
                                        ((64, 16, 9), 3)])
def test_fastcut_random_gpu(gpu, shape, seed):
    nx, ny, nz = shape
    conn = 6 if nz > 1 else 4
    rng = np.random.default_rng(seed)
    cs, ct, caps = random_grid(rng, nx, ny, nz, conn)
    f, l, ok = ol.maxflow(nx, ny, nz, conn, cs, ct, caps)
    for opts in (dict(), dict(initial_relabel=1), dict(check_every=4, profile=1)):
        say("case fastcut random", shape, seed, opts)
        g = gpu.GridGraph([(nx, ny, nz, conn)])
        g.set_terminal_caps(cs, ct)
        for d in range(conn):
            g.set_neighbor_caps(d, caps[d])
        st = g.solve(method=FAST, **opts)
        assert st.converged == 1
        assert g.flow() == f
        assert np.array_equal(g.labels(), l)
        g.close()


def test_fastcut_batch_and_errors_gpu(gpu):
    rng = np.random.default_rng(11)
    shapes = [(70, 50, 1, 4), (31, 90, 1, 4), (128, 8, 1, 4)]
    g = gpu.GridGraph(shapes)
    expect = []
    for i, (nx, ny, nz, conn) in enumerate(shapes):
        cs, ct, caps = random_grid(rng, nx, ny, nz, conn)
        g.set_terminal_caps(cs, ct, inst=i)
        for d in range(conn):
            g.set_neighbor_caps(d, caps[d], inst=i)
        expect.append(ol.maxflow(nx, ny, nz, conn, cs, ct, caps)[:2])
    say("case fastcut batch")
    st = g.solve(method=FAST)
    assert st.converged == 1
    for i, (f, l) in enumerate(expect):
        assert g.flow(i) == f
        assert np.array_equal(g.labels(i), l)
    g.close()
    # 8-connected graphs have no Fast-Cut line pushes along the diagonals
    g8 = gpu.GridGraph([(40, 30, 1, 8)])
    cs, ct, caps = random_grid(rng, 40, 30, 1, 8)
    g8.set_terminal_caps(cs, ct)
    for d in range(8):
        g8.set_neighbor_caps(d, caps[d])
    with pytest.raises(gpu.JFCError):
        g8.solve(method=FAST)
    with pytest.raises(gpu.JFCError):
        g8.solve(method=7)
    g8.close()
