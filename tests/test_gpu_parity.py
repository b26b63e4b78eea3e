"""GPU parity: the HIP solver (through the C-ABI) against the oracle, BK303 and
the committed golden fixtures.  Bit-exact: int64 flow and canonical labels."""
import ctypes
import hashlib
import sys

import numpy as np
import pytest

import oracle_lib as ol
from conftest import load_c1_goldens, load_ref_fixture, random_grid

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def solve_grid(jf, nx, ny, nz, conn, cs, ct, caps, **opt):
    g = jf.GridGraph([(nx, ny, nz, conn)])
    g.set_terminal_caps(cs, ct)
    for d in range(conn):
        g.set_neighbor_caps(d, caps[d])
    st = g.solve(**opt)
    return g, st


@pytest.mark.parametrize("name", ["flower", "person", "sponge"])
def test_reference_fixtures_gpu(gpu, name):
    nx, ny, ef, cs, ct, caps, flow, lab = load_ref_fixture(name)
    say("case", name, nx, ny)
    g = gpu.GridGraph([(nx, ny, 1, 4)])
    g.set_excess(ef)
    for d in range(4):
        g.set_neighbor_caps(d, caps[d])
    st = g.solve()
    assert st.converged == 1
    assert g.flow() == flow
    assert np.array_equal(g.labels(), lab)
    assert g.energy() == flow
    g.close()


@pytest.mark.parametrize("conn,shape", [(4, (37, 29, 1)), (8, (41, 23, 1)), (6, (13, 11, 9)),
                                        (4, (300, 200, 1)), (8, (257, 190, 1)),
                                        (6, (40, 33, 21))])
def test_random_grids_gpu(gpu, conn, shape):
    rng = np.random.default_rng(conn * 100 + shape[0])
    nx, ny, nz = shape
    for trial in range(3):
        say("case", conn, shape, trial)
        cs, ct, caps = random_grid(rng, nx, ny, nz, conn)
        f, l, ok = ol.maxflow(nx, ny, nz, conn, cs, ct, caps)
        g, st = solve_grid(gpu, nx, ny, nz, conn, cs, ct, caps)
        assert st.converged == 1
        assert g.flow() == f
        assert np.array_equal(g.labels(), l)
        g.close()


# 32x8 tiles: 1024 (small-graph paths: push launches and search passes pick
# their own tiles, search passes in colour order), 2080 (colour classes still
# small enough for self-picking pushes; relabel passes in distance order,
# compacted search lists) and 8250 (colour classes of 2062-2063 tiles:
# compacted push lists too)
@pytest.mark.parametrize("conn,nx,ny", [(4, 1024, 256), (4, 2080, 256), (8, 2080, 256),
                                        (8, 2100, 1000)])
@pytest.mark.parametrize("opts", [dict(), dict(initial_relabel=1),
                                  dict(initial_relabel=1, stall_certify=1, tile_order=1)])
def test_small_and_large_launch_paths_gpu(gpu, conn, nx, ny, opts):
    rng = np.random.default_rng(nx + ny + conn)
    cs, ct, caps = random_grid(rng, nx, ny, 1, conn)
    f, l, ok = ol.maxflow(nx, ny, 1, conn, cs, ct, caps)
    say("case launch paths", conn, nx, ny, opts)
    g, st = solve_grid(gpu, nx, ny, 1, conn, cs, ct, caps, tile_shape="32x8", **opts)
    assert st.converged == 1
    assert g.flow() == f
    assert np.array_equal(g.labels(), l)
    assert g.energy() == f
    g.close()


@pytest.mark.parametrize("opts", [dict(tile_shape="64x4"), dict(tile_shape="16x16"),
                                  dict(max_local_iters=1), dict(relabel_work=0.0),
                                  dict(check_every=1, max_local_iters=1000),
                                  dict(graph_sweeps=0), dict(check_every=3), dict(tile_order=1),
                                  dict(initial_relabel=1)])
def test_options_do_not_change_result(gpu, opts):
    rng = np.random.default_rng(99)
    nx, ny = 203, 147
    for conn in (4, 8):
        cs, ct, caps = random_grid(rng, nx, ny, 1, conn)
        f, l, ok = ol.maxflow(nx, ny, 1, conn, cs, ct, caps)
        say("case", opts, conn)
        g, st = solve_grid(gpu, nx, ny, 1, conn, cs, ct, caps, **opts)
        assert g.flow() == f
        assert np.array_equal(g.labels(), l)
        g.close()


def serpentine(nx, ny, nz=1, cap=5, seed=0):
    """A single corridor snaking through the grid (rows of each z slice,
    then slices): its residual breadth-first distances cross the tile grid
    back and forth, the worst case for pass-by-pass relabelling."""
    rng = np.random.default_rng(seed)
    N = nx * ny * nz
    conn = 6 if nz > 1 else 4
    caps = np.zeros((conn, N), np.int32)
    idx = lambda x, y, z: x + nx * (y + ny * z)
    # one slice: even rows left-right-left..., joined through the odd rows
    snake = []
    for k, y in enumerate(range(0, ny, 2)):
        xs = range(nx) if k % 2 == 0 else range(nx - 1, -1, -1)
        if snake:
            snake.append((snake[-1][0], y - 1))
        snake += [(x, y) for x in xs]
    # slices alternate direction, joined by one step along z
    path = []
    for z in range(nz):
        path += [(x, y, z) for x, y in (snake if z % 2 == 0 else snake[::-1])]
    dirs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)][:conn]
    for (a, b) in zip(path, path[1:]):
        d = (b[0] - a[0], b[1] - a[1], b[2] - a[2])
        k = dirs.index(d)
        caps[k, idx(*a)] = cap + rng.integers(0, 3)
        caps[k ^ 1, idx(*b)] = cap + rng.integers(0, 3)
    cs = np.zeros(N, np.int32)
    ct = np.zeros(N, np.int32)
    cs[idx(*path[0])] = 1000
    ct[idx(*path[-1])] = 1000
    on = np.array([idx(*p) for p in path])
    cs[on[::7]] += rng.integers(0, 4, on[::7].size).astype(np.int32)
    ct[on[3::11]] += rng.integers(0, 4, on[3::11].size).astype(np.int32)
    return cs, ct, caps, conn


@pytest.mark.parametrize("shape", [(256, 63, 1), (64, 31, 6)])
def test_relabel_serpentine(gpu, shape):
    """Exact relabel and cut along one long winding residual path."""
    nx, ny, nz = shape
    cs, ct, caps, conn = serpentine(nx, ny, nz)
    f, l, ok = ol.maxflow(nx, ny, nz, conn, cs, ct, caps)
    for opts in (dict(), dict(tile_order=1), dict(max_local_iters=1)):
        for ir in (0, 1):
            say("case serpentine", shape, opts, ir)
            g, st = solve_grid(gpu, nx, ny, nz, conn, cs, ct, caps, initial_relabel=ir, **opts)
            assert st.converged == 1
            assert g.flow() == f
            assert np.array_equal(g.labels(), l)
            g.close()


@pytest.mark.parametrize("conn,shape", [(4, (150, 97, 1)), (8, (133, 71, 1)), (6, (40, 33, 17))])
def test_repeated_relabels_gpu(gpu, conn, shape):
    """Many global relabels on one graph, across solves: a relabel after
    every sweep (tiny relabel budget), with and without one before the first
    sweep.  Relabels alternate the tag of the heights they write and a solve
    starts from the previous solve's words, so a stale word taken for a
    current one would show as a wrong flow or cut here."""
    nx, ny, nz = shape
    rng = np.random.default_rng(conn + nx)
    cs, ct, caps = random_grid(rng, nx, ny, nz, conn)
    f, l, ok = ol.maxflow(nx, ny, nz, conn, cs, ct, caps)
    g = gpu.GridGraph([(nx, ny, nz, conn)])
    g.set_terminal_caps(cs, ct)
    for d in range(conn):
        g.set_neighbor_caps(d, caps[d])
    relabels = 0
    for k, opts in enumerate([dict(relabel_work=0.01), dict(relabel_work=0.01, initial_relabel=1),
                              dict(), dict(initial_relabel=1), dict(relabel_work=0.01, check_every=2),
                              dict(relabel_work=0.01, initial_relabel=1)]):
        say("case repeated relabels", conn, shape, k, opts)
        st = g.solve(**opts)
        relabels += st.global_relabels
        assert st.converged == 1
        assert g.flow() == f
        assert np.array_equal(g.labels(), l)
    assert relabels >= 8
    g.close()


def test_sweep_graph_replays(gpu):
    """The captured per-sweep hipGraph (default), eager launches and profiled
    launches give identical flows and labels, over many replays on one graph,
    with eager work and other option sets in between, and on a caller stream."""
    import torch
    rng = np.random.default_rng(31)
    nx, ny, conn = 190, 170, 8
    cs, ct, caps = random_grid(rng, nx, ny, 1, conn)
    f, l, ok = ol.maxflow(nx, ny, 1, conn, cs, ct, caps)
    g = gpu.GridGraph([(nx, ny, 1, conn)])
    g.set_terminal_caps(cs, ct)
    for d in range(conn):
        g.set_neighbor_caps(d, caps[d])
    runs = [dict(), dict(graph_sweeps=0), dict(), dict(profile=1), dict(max_local_iters=3), dict(),
            dict(max_local_iters=3), dict(graph_sweeps=0, max_local_iters=1), dict()]
    for i, opts in enumerate(runs):
        say("case replay", i, opts)
        st = g.solve(**opts)
        assert st.converged == 1 and st.sweeps >= 2
        assert g.flow() == f, (i, opts)
        assert np.array_equal(g.labels(), l), (i, opts)
        assert g.energy() == f
    s = torch.cuda.Stream()
    for i in range(3):
        say("case stream", i)
        st = g.solve(stream=s.cuda_stream)
        assert g.flow() == f and np.array_equal(g.labels(), l)
    g.close()


def test_3d_tile_shape_option(gpu):
    rng = np.random.default_rng(4)
    nx, ny, nz = 70, 20, 9
    cs, ct, caps = random_grid(rng, nx, ny, nz, 6)
    f, l, ok = ol.maxflow(nx, ny, nz, 6, cs, ct, caps)
    g, st = solve_grid(gpu, nx, ny, nz, 6, cs, ct, caps, tile_shape="32x4x2")
    assert g.flow() == f and np.array_equal(g.labels(), l)


@pytest.mark.parametrize("opts", [dict(initial_relabel=1), dict(initial_relabel=1, relabel_work=0.01),
                                  dict(initial_relabel=0, relabel_work=0.01, max_local_iters=1)])
def test_32x4x2_relabels_on_8x8x4_tiles_gpu(gpu, opts):
    """Graphs on 32x4x2 push tiles relabel on an 8x8x4 tiling, and the relax
    visits set the push tiling's active flags.  Sizes cut both tilings
    unevenly (x not a multiple of 32, y of 8, z of 4); one graph and a batch
    of instances (the per-instance descriptors of both tilings), each
    relabelled at least once, against the oracle's flow and labels."""
    rng = np.random.default_rng(77)
    shp = [(70, 21, 9, 6), (33, 9, 5, 6), (40, 17, 3, 6), (96, 12, 7, 6)]
    data = [random_grid(rng, a, b, c, 6, seed_frac=0.02) for (a, b, c, _) in shp]
    exp = [ol.maxflow(a, b, c, 6, *data[i])[:2] for i, (a, b, c, _) in enumerate(shp)]
    say("case 32x4x2 single", shp[0], opts)
    relabelled = opts["initial_relabel"] == 1  # otherwise the visit budget triggers them
    g, st = solve_grid(gpu, *shp[0][:3], 6, *data[0], tile_shape="32x4x2", **opts)
    assert st.converged == 1 and (st.global_relabels >= 1 or not relabelled)
    assert g.flow() == exp[0][0] and np.array_equal(g.labels(), exp[0][1])
    g.close()
    say("case 32x4x2 batch", shp, opts)
    G = gpu.GridGraph(shp)
    for i, (cs, ct, caps) in enumerate(data):
        G.set_terminal_caps(cs, ct, inst=i)
        for d in range(6):
            G.set_neighbor_caps(d, caps[d], inst=i)
    for rep in range(2):  # the second solve starts from the first's state
        st = G.solve(tile_shape="32x4x2", **opts)
        assert st.converged == 1 and (st.global_relabels >= 1 or not relabelled)
        for i, (ff, ll) in enumerate(exp):
            assert G.flow(i) == ff, (rep, i)
            assert np.array_equal(G.labels(i), ll), (rep, i)
    G.close()


def test_c1_goldens_gpu(gpu):
    """The 512x512 4-connected cut (preset c1) end to end on the device: data term (k-means, unary, pairwise)
    bit-exact with the oracle, then flow and labels equal to the goldens."""
    for gd in load_c1_goldens():
        say("case c1 seed", gd["seed"])
        rgb, sd = gpu.synth_image(gd["W"], gd["H"], nblobs=8, noise=25.0, seed=gd["seed"])
        assert sha(rgb) == gd["image_sha256"]
        g = gpu.GridGraph([(gd["W"], gd["H"], 1, 4)])
        ms = g.model_image(rgb, sd, k_fg=5, k_bg=5, lambda_xy=50.0)
        assert np.array_equal(np.array(ms.centroids_fg[:15]).reshape(5, 3), np.array(gd["centroids_fg"]))
        assert np.array_equal(np.array(ms.centroids_bg[:15]).reshape(5, 3), np.array(gd["centroids_bg"]))
        cs, ct = g.terminal_caps()
        caps = g.neighbor_caps()
        assert sha(np.concatenate([cs, ct, caps.ravel()])) == gd["caps_sha256"]
        st = g.solve()
        assert g.flow() == gd["flow"]
        lab = g.labels()
        assert sha(lab) == gd["labels_sha256"]
        assert g.energy() == gd["flow"]
        g.close()


def test_batch_gpu(gpu):
    rng = np.random.default_rng(12)
    shapes, data = [], []
    for i in range(9):
        conn = 4 if i % 2 else 8
        nx, ny = int(rng.integers(20, 180)), int(rng.integers(20, 160))
        shapes.append((nx, ny, 1, conn))
        data.append(random_grid(rng, nx, ny, 1, conn))
    g = gpu.GridGraph(shapes)
    for i, (cs, ct, caps) in enumerate(data):
        g.set_terminal_caps(cs, ct, inst=i)
        for d in range(shapes[i][3]):
            g.set_neighbor_caps(d, caps[d], inst=i)
    st = g.solve()
    assert st.converged == 1
    flows = g.flows()
    for i, (cs, ct, caps) in enumerate(data):
        nx, ny, _, conn = shapes[i]
        f, l, ok = ol.maxflow(nx, ny, 1, conn, cs, ct, caps)
        assert flows[i] == f, i
        assert np.array_equal(g.labels(i), l), i
    g.close()


def test_autotune_gpu(gpu):
    """jfc_autotune picks one of the shapes and local-iteration caps, leaves
    results unchanged, and a later graph of the same size class starts with
    the same pair; a cap fixed by the caller is not tuned."""
    rng = np.random.default_rng(77)
    shapes = [(int(rng.integers(40, 120)), int(rng.integers(40, 120)), 1, 4 if i % 2 else 8)
              for i in range(6)]
    data = [random_grid(rng, s[0], s[1], 1, s[3]) for s in shapes]

    def build():
        g = gpu.GridGraph(shapes)
        for i, (cs, ct, caps) in enumerate(data):
            g.set_terminal_caps(cs, ct, inst=i)
            for d in range(shapes[i][3]):
                g.set_neighbor_caps(d, caps[d], inst=i)
        return g

    g = build()
    best = g.autotune(reps=1)
    mli, ir = g.local_iters(), g.initial_relabel()
    say("case autotune", best, mli, ir)
    assert best in ("32x8", "64x4", "16x16") and g.tile_shape() == best
    assert mli in (2, 3, 4) and ir in (0, 1)
    st = g.solve()
    assert st.converged == 1
    for i, (cs, ct, caps) in enumerate(data):
        f, l, ok = ol.maxflow(shapes[i][0], shapes[i][1], 1, shapes[i][3], cs, ct, caps)
        assert g.flow(i) == f and np.array_equal(g.labels(i), l), i
    g.close()
    g2 = build()
    assert g2.tile_shape() == best and g2.local_iters() == mli and g2.initial_relabel() == ir
    # a fixed cap and relabel choice: only the shape is tuned, and the class
    # keeps neither
    best2 = g2.autotune(reps=1, max_local_iters=1, initial_relabel=1)
    assert best2 in ("32x8", "64x4", "16x16") and g2.local_iters() == 2
    assert g2.initial_relabel() == 0
    g2.close()


def test_backward_gpu(gpu):
    rng = np.random.default_rng(21)
    nx, ny, conn = 40, 30, 8
    cs, ct, caps = random_grid(rng, nx, ny, 1, conn, seed_frac=0.0)
    g, st = solve_grid(gpu, nx, ny, 1, conn, cs, ct, caps)
    lab = g.labels()
    dcs, dct, dnb = g.backward()
    assert np.array_equal(dcs, (lab == 0).astype(np.float32))
    assert np.array_equal(dct, (lab == 1).astype(np.float32))
    E0 = g.flow()
    # +-1 integer finite differences on random non-tie neighbour arcs
    idx = np.arange(nx * ny).reshape(ny, nx)
    checked = 0
    for trial in range(60):
        d = int(rng.integers(0, conn))
        v = int(rng.integers(0, nx * ny))
        if caps[d][v] < 1:
            continue
        y, x = divmod(v, nx)
        dx, dy, _ = gpu.DIRS_2D[d]
        if not (0 <= x + dx < nx and 0 <= y + dy < ny):
            continue
        c_up = caps.copy(); c_up[d][v] += 1
        c_dn = caps.copy(); c_dn[d][v] -= 1
        Eu, _, _ = ol.maxflow(nx, ny, 1, conn, cs, ct, c_up)
        Ed, _, _ = ol.maxflow(nx, ny, 1, conn, cs, ct, c_dn)
        if Eu - E0 != E0 - Ed:
            continue  # tie: the energy is not differentiable at this arc
        assert dnb[d][v] == Eu - E0
        checked += 1
    assert checked > 10


def test_not_converged_status(gpu):
    rng = np.random.default_rng(5)
    nx, ny = 300, 300
    cs, ct, caps = random_grid(rng, nx, ny, 1, 4)
    g = gpu.GridGraph([(nx, ny, 1, 4)])
    g.set_terminal_caps(cs, ct)
    for d in range(4):
        g.set_neighbor_caps(d, caps[d])
    st = g.solve(max_sweeps=1, check_every=1, max_local_iters=1, relabel_work=0.0)
    assert g.last_rc in (gpu.OK, gpu.NOT_CONVERGED)
    if g.last_rc == gpu.NOT_CONVERGED:
        assert st.converged == 0


def test_input_validation_gpu(gpu):
    g = gpu.GridGraph([(16, 16, 1, 4)])
    bad = np.full(256, gpu.MAX_TERMINAL_CAP + 1, np.int32)
    with pytest.raises(gpu.JFCError) as e:
        g.set_terminal_caps(bad, np.zeros(256, np.int32))
    assert e.value.code == gpu.ERR_OVERFLOW
    with pytest.raises(gpu.JFCError) as e:
        g.set_neighbor_caps(0, np.full(256, gpu.MAX_NEIGHBOR_CAP + 1, np.int32))
    assert e.value.code == gpu.ERR_OVERFLOW
    with pytest.raises(gpu.JFCError) as e:
        g.set_neighbor_caps(0, -np.ones(256, np.int32))
    assert e.value.code == gpu.ERR_OVERFLOW
    # a setter that failed zeroed its plane: the solve refuses the instance
    # until its capacities are set again
    with pytest.raises(gpu.JFCError) as e:
        g.solve()
    assert e.value.code == gpu.ERR_STATE
    g.set_terminal_caps(np.zeros(256, np.int32), np.zeros(256, np.int32))
    g.set_neighbor_caps(0, np.ones(256, np.int32))
    with pytest.raises(gpu.JFCError) as e:
        g.solve()  # no terminal capacity anywhere: the reference's "no labels"
    assert e.value.code == gpu.ERR_NO_SEEDS
    with pytest.raises(gpu.JFCError):
        g.flow()
    g.close()


@pytest.mark.parametrize("shape", [(203, 77, 1, 4), (129, 95, 1, 8), (37, 29, 23, 6)])
def test_tile_flags_gpu(gpu, shape):
    """Per-tile seed / active flags from the streaming init kernels (several
    tiles per lane, workgroup-wide bitwise OR) against numpy."""
    rng = np.random.default_rng(31)
    nx, ny, nz, conn = shape
    N = nx * ny * nz
    # sparse terminals so that most tiles have neither excess nor deficit
    cs = np.zeros(N, np.int32)
    ct = np.zeros(N, np.int32)
    cs[rng.integers(0, N, N // 300)] = 7
    ct[rng.integers(0, N, N // 300)] = 5
    g = gpu.GridGraph([shape])
    g.set_terminal_caps(cs, ct)
    nt, (tw, th, td), (gx, gy, gz), _ = g.debug_tiles(0)
    e = (cs.astype(np.int64) - ct).reshape(nz, ny, nx)
    pad = np.zeros((gz * td, gy * th, gx * tw), np.int64)
    pad[:nz, :ny, :nx] = e
    tiles = pad.reshape(gz, td, gy, th, gx, tw).transpose(0, 2, 4, 1, 3, 5).reshape(nt, -1)
    exp_def = (tiles < 0).any(1).astype(np.uint8)
    exp_exc = (tiles > 0).any(1).astype(np.uint8)
    assert np.array_equal(g.debug_tile_flags(0), exp_def)
    assert np.array_equal(g.debug_tile_flags(1), exp_exc)
    assert np.array_equal(g.debug_tile_flags(2), exp_exc)
    g.close()


@pytest.mark.parametrize("method", [0, 1])
def test_tile_order_gpu(gpu, method):
    """Distance transform (method 0) and jump flooding (method 1) over the
    tile grid = brute-force taxicab distance to the nearest seed tile; the
    order is sorted by it; max distance = reduce."""
    rng = np.random.default_rng(8)
    for shape in [(640, 360, 1, 4), (1000, 70, 1, 8), (96, 80, 60, 6), (20000, 9, 1, 4)]:
        g = gpu.GridGraph([shape, shape])
        nt, td, gd, off = g.debug_tiles(0)
        _, _, gd1, off1 = g.debug_tiles(1)
        seeds = (rng.random(nt) < 0.01).astype(np.uint8)
        seeds[off1:] = 0
        seeds[5] = 1
        seeds[off1 + 3] = 1
        dist, order, mx = g.debug_tile_order(seeds, method)
        for inst, o in ((0, off), (1, off1)):
            tx, ty, tz = gd
            n = tx * ty * tz
            zz, yy, xx = np.meshgrid(np.arange(tz), np.arange(ty), np.arange(tx), indexing="ij")
            coords = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], 1)
            s = np.nonzero(seeds[o:o + n])[0]
            if s.size == 0:
                continue
            bd = np.abs(coords[:, None, :] - coords[s][None, :, :]).sum(-1).min(1)
            assert np.array_equal(dist[o:o + n], bd)
        assert sorted(order.tolist()) == list(range(nt))
        assert np.all(np.diff(dist[order]) >= 0)
        assert mx == int(dist.max())  # every instance has a seed tile
        g.close()


def test_mask_words_across_solves_gpu(gpu):
    """The input mask words (mask0) are written by the first re-solve of
    unchanged inputs, not by a fresh solve (except Fast-Cut's, which reads
    them at once): fresh and repeated solves, alternating the two methods and
    re-marking the inputs written, all against the oracle."""
    rng = np.random.default_rng(31)
    nx, ny = 150, 110
    cs, ct, caps = random_grid(rng, nx, ny, 1, 4)
    f, l, ok = ol.maxflow(nx, ny, 1, 4, cs, ct, caps)
    g = gpu.GridGraph([(nx, ny, 1, 4)])
    g.set_terminal_caps(cs, ct)
    for d in range(4):
        g.set_neighbor_caps(d, caps[d])
    seq = [(False, {}), (False, dict(method=1)), (False, {}), (True, {}), (False, {}),
           (True, dict(method=1)), (False, {}), (False, dict(method=1)), (True, {}), (True, {})]
    for k, (fresh, opts) in enumerate(seq):
        if fresh:
            g.inputs_written()
        say("case mask words", k, fresh, opts)
        st = g.solve(**opts)
        assert st.converged == 1
        assert g.flow() == f, (k, opts)
        assert np.array_equal(g.labels(), l), (k, opts)
    g.close()
