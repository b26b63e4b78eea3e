"""Pins the CPU oracle (oracle/jfc_oracle.c) before anything is checked against it.

Pins, by what they exclude:
  * the reference's own fixtures and published max-flows
    (latex/sources/section-5-results.tex:31,36,41) and flower's canonical
    labels from its .cut object channel (graphcut_compute_cut.cl semantics);
  * BK303 compiled from the reference sources (oracle/_ref), labels with free
    nodes on the sink side -- an independent algorithm (augmenting paths);
  * scipy.sparse.csgraph.maximum_flow on the unfolded graph, with the
    canonical set taken as residual reachability from the super source;
  * brute force over all 2^n labelings of tiny grids: the flow is the minimum
    energy and the canonical set is the intersection of all minimum cuts;
  * weak duality: energy(labels) == flow, which fails for any dropped term;
  * the integer data term against direct numpy evaluation of its definition.
"""
import itertools
import os

import numpy as np
import pytest

import oracle_lib as ol
from conftest import load_c1_goldens, load_ref_fixture, random_grid

DIRS = {4: [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0)],
        8: [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0),
            (1, 1, 0), (-1, -1, 0), (-1, 1, 0), (1, -1, 0)],
        6: [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]}


def zero_border(nx, ny, nz, conn, caps):
    caps = caps.copy().reshape(conn, nz, ny, nx)
    for d, (dx, dy, dz) in enumerate(DIRS[conn]):
        if dx == 1: caps[d, :, :, -1] = 0
        if dx == -1: caps[d, :, :, 0] = 0
        if dy == 1: caps[d, :, -1, :] = 0
        if dy == -1: caps[d, :, 0, :] = 0
        if dz == 1: caps[d, -1, :, :] = 0
        if dz == -1: caps[d, 0, :, :] = 0
    return caps.reshape(conn, -1)


@pytest.mark.parametrize("name", ["flower", "person", "sponge"])
def test_reference_fixtures(name):
    nx, ny, ef, cs, ct, caps, flow, lab = load_ref_fixture(name)
    f, l, ok = ol.maxflow(nx, ny, 1, 4, cs, ct, caps)
    assert ok == 1
    assert f == flow
    assert np.array_equal(l, lab)
    assert ol.energy(nx, ny, 1, 4, cs, ct, caps, l) == flow


@pytest.mark.skipif(not ol.have_bk(), reason="oracle/_ref not built")
@pytest.mark.parametrize("conn,shape", [(4, (37, 29, 1)), (8, (41, 23, 1)), (6, (13, 11, 9)),
                                        (4, (200, 150, 1)), (8, (160, 120, 1)), (6, (30, 25, 20))])
def test_oracle_vs_bk303(conn, shape):
    rng = np.random.default_rng(conn * 1000 + shape[0])
    nx, ny, nz = shape
    for trial in range(3):
        cs, ct, caps = random_grid(rng, nx, ny, nz, conn)
        f, l, ok = ol.maxflow(nx, ny, nz, conn, cs, ct, caps)
        fb, lb = ol.bk303(nx, ny, nz, conn, cs, ct, caps)
        assert ok == 1
        assert f == fb
        assert np.array_equal(l, lb)
        assert ol.energy(nx, ny, nz, conn, cs, ct, caps, l) == f


def unfolded_scipy(nx, ny, nz, conn, cs, ct, caps):
    """Max flow and source-reachable set by scipy on the unfolded graph."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import maximum_flow
    N = nx * ny * nz
    s, t = N, N + 1
    caps = zero_border(nx, ny, nz, conn, caps)
    idx = np.arange(N).reshape(nz, ny, nx)
    rows, cols, vals = [], [], []
    for d, (dx, dy, dz) in enumerate(DIRS[conn]):
        src = idx[max(0, -dz):nz - max(0, dz), max(0, -dy):ny - max(0, dy), max(0, -dx):nx - max(0, dx)]
        dst = idx[max(0, dz):nz - max(0, -dz) or None, max(0, dy):ny - max(0, -dy) or None,
                  max(0, dx):nx - max(0, -dx) or None]
        src, dst = src.ravel(), dst.ravel()
        rows.append(src); cols.append(dst); vals.append(caps[d][src])
    rows.append(np.full(N, s)); cols.append(np.arange(N)); vals.append(cs)
    rows.append(np.arange(N)); cols.append(np.full(N, t)); vals.append(ct)
    r, c, v = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals).astype(np.int64)
    keep = v > 0
    M = csr_matrix((v[keep], (r[keep], c[keep])), shape=(N + 2, N + 2))
    M.sum_duplicates()
    res = maximum_flow(M.astype(np.int32), s, t)
    F = res.flow.tocsr()
    # residual reachability from s
    cap = M.tocsr()
    resid = (cap - F).tocsr()
    resid.eliminate_zeros()
    seen = np.zeros(N + 2, bool)
    seen[s] = True
    stack = [s]
    indptr, indices, data = resid.indptr, resid.indices, resid.data
    while stack:
        u = stack.pop()
        for k in range(indptr[u], indptr[u + 1]):
            v2 = indices[k]
            if data[k] > 0 and not seen[v2]:
                seen[v2] = True
                stack.append(v2)
    return int(res.flow_value), seen[:N].astype(np.uint8)


@pytest.mark.parametrize("conn,shape", [(4, (23, 17, 1)), (8, (19, 21, 1)), (6, (9, 8, 7))])
def test_oracle_vs_scipy(conn, shape):
    rng = np.random.default_rng(7 + conn)
    nx, ny, nz = shape
    for trial in range(4):
        cs, ct, caps = random_grid(rng, nx, ny, nz, conn, seed_frac=0.0)
        f, l, ok = ol.maxflow(nx, ny, nz, conn, cs, ct, caps)
        fs, ls = unfolded_scipy(nx, ny, nz, conn, cs, ct, caps)
        assert f == fs
        assert np.array_equal(l, ls)


def brute_force(nx, ny, nz, conn, cs, ct, caps):
    N = nx * ny * nz
    best, sets = None, []
    for bits in itertools.product([0, 1], repeat=N):
        lab = np.array(bits, np.uint8)
        E = ol.energy(nx, ny, nz, conn, cs, ct, caps, lab)
        if best is None or E < best:
            best, sets = E, [lab]
        elif E == best:
            sets.append(lab)
    inter = np.minimum.reduce(sets)
    return best, inter


@pytest.mark.parametrize("conn,shape", [(4, (4, 3, 1)), (8, (3, 4, 1)), (6, (2, 2, 3)), (4, (4, 4, 1))])
def test_oracle_vs_brute_force(conn, shape):
    rng = np.random.default_rng(11 * conn + shape[0])
    nx, ny, nz = shape
    for trial in range(6):
        cs, ct, caps = random_grid(rng, nx, ny, nz, conn, cap_max=6, term_max=8, seed_frac=0.0)
        if trial % 2:  # many ties
            caps = (caps > 3).astype(np.int32) * 2
            cs, ct = cs % 3, ct % 3
        f, l, ok = ol.maxflow(nx, ny, nz, conn, cs, ct, caps)
        E, inter = brute_force(nx, ny, nz, conn, cs, ct, caps)
        assert ok == 1
        assert f == E
        assert np.array_equal(l, inter)


def test_certificate_detects_planted_path():
    """A planted mistake in the oracle's state: restore capacity on one
    saturated cut arc of a solved fixture; the certificate must fail."""
    nx, ny, ef, cs, ct, caps, flow, lab = load_ref_fixture("flower")
    f, l, ok, e, r = ol.maxflow(nx, ny, 1, 4, cs, ct, caps, want_state=True)
    assert ok == 1
    idx = np.arange(nx * ny).reshape(ny, nx)
    u = idx[:, :-1].ravel()
    v = u + 1
    cut = np.nonzero((l[u] == 1) & (l[v] == 0) & (caps[0][u] > 0))[0]
    assert cut.size > 0
    r2 = r.copy()
    k = u[cut[0]]
    r2[0][k] = caps[0][k]
    # the sink side must reach a deficit for the path to be augmenting: make
    # the arc's head a deficit node if it is not already
    e2 = e.copy()
    e2[k + 1] = min(e2[k + 1], -1)
    l2, ok2 = ol.canonical_cut(nx, ny, 1, 4, e2, r2)
    assert ok2 == 0


def test_kmeans_definition():
    rng = np.random.default_rng(3)
    feat = rng.integers(0, 256, (500, 3)).astype(np.int32)
    cent, init, it = ol.kmeans(feat, 4, 0, iters=50)
    # init: distinct key levels (sum of channels), picks j*(m-1)/(K-1)
    keys = feat.sum(1)
    levels = np.unique(keys)
    m = levels.size
    for j in range(4):
        pick = levels[j * (m - 1) // 3]
        sel = feat[keys == pick]
        exp = (sel.sum(0).astype(np.int64) * 256 + sel.shape[0] // 2) // sel.shape[0]
        assert np.array_equal(init[j], exp)
    # fixed point: every centroid is the rounded mean of its members
    x = feat.astype(np.int64) * 256
    d = ((x[:, None, :] - cent[None, :, :]) ** 2).sum(-1)
    assign = d.argmin(1)
    for k in range(4):
        mem = feat[assign == k]
        if mem.shape[0]:
            exp = (mem.sum(0).astype(np.int64) * 256 + mem.shape[0] // 2) // mem.shape[0]
            assert np.array_equal(cent[k], exp)
    # K = 1: the rounded mean of everything
    c1, _, _ = ol.kmeans(feat, 1, 0)
    assert np.array_equal(c1[0], (feat.sum(0).astype(np.int64) * 256 + 250) // 500)


def test_unary_and_pairwise_definition():
    rng = np.random.default_rng(5)
    H, W = 24, 31
    img = rng.integers(0, 256, (H, W, 3)).astype(np.uint8)
    seeds = np.zeros((H, W), np.uint8)
    seeds[2:5, 3:9] = 1
    seeds[15:20, 10:25] = 2
    cs, ct, caps, info = ol.model_image(img, seeds, 8, k=3, lam=50.0)
    feat = img.reshape(-1, 3).astype(np.int64)
    R2 = 255 * 255 * 65536
    df = (((feat * 256)[:, None, :] - info["cfg"][None]) ** 2).sum(-1).min(1)
    db = (((feat * 256)[:, None, :] - info["cbg"][None]) ** 2).sum(-1).min(1)
    sd = seeds.ravel()
    exp_cs = np.where(sd == 1, 1 << 30, np.where(sd == 2, 0, (100 * db + R2 // 2) // R2))
    exp_ct = np.where(sd == 1, 0, np.where(sd == 2, 1 << 30, (100 * df + R2 // 2) // R2))
    assert np.array_equal(cs, exp_cs)
    assert np.array_equal(ct, exp_ct)
    # pairwise: round(lambda * exp(-D / (2 s2))) evaluated directly in float64
    f3 = img.astype(np.int64)
    pairs = []
    for (dx, dy, _) in DIRS[8][0::2]:
        a = f3[max(0, -dy):H - max(0, dy), max(0, -dx):W - max(0, dx)]
        b = f3[max(0, dy):H - max(0, -dy) or None, max(0, dx):W - max(0, -dx) or None]
        pairs.append(((a - b) ** 2).sum(-1).ravel())
    D = np.concatenate(pairs)
    s2 = D.sum() / D.size
    mism = 0
    for d, (dx, dy, _) in enumerate(DIRS[8]):
        lam = 50.0 if d < 4 else 50.0 / np.sqrt(2.0)
        plane = caps[d].reshape(H, W)
        ys, xs = np.mgrid[0:H, 0:W]
        X, Y = xs + dx, ys + dy
        inside = (X >= 0) & (X < W) & (Y >= 0) & (Y < H)
        Dd = np.zeros((H, W), np.int64)
        Dd[inside] = ((f3[ys[inside], xs[inside]] - f3[Y[inside], X[inside]]) ** 2).sum(-1)
        w = np.floor(lam * np.exp(-Dd / (2 * s2)) + 0.5).astype(np.int64)
        w[~inside] = 0
        # exact ties (value within 1e-9 of .5) may round either way in float
        frac = lam * np.exp(-Dd / (2 * s2)) - np.floor(lam * np.exp(-Dd / (2 * s2)))
        ok = np.abs(frac - 0.5) > 1e-9
        mism += int(((plane != w) & ok).sum())
    assert mism == 0


def test_pair_thresholds_monotone_and_boundaries():
    T = ol.pair_thresholds(50.0, 123456, 789)
    assert T.size == 50
    assert np.all(np.diff(T) <= 0)
    s2 = 123456 / 789
    for k, t in enumerate(T, start=1):
        # D = T is the last value that still rounds to >= k
        assert 50 * np.exp(-t / (2 * s2)) >= k - 0.5 - 1e-9
        assert 50 * np.exp(-(t + 1) / (2 * s2)) < k - 0.5 + 1e-9


def test_c1_goldens_cpu(jfcut_mod):
    """512x512 4-connected instances (preset c1): the generator, the oracle's data term, the oracle's
    solve and BK303 all agree with the committed goldens."""
    import hashlib
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    for g in load_c1_goldens()[:2]:
        rgb, sd = jfcut_mod.synth_image(g["W"], g["H"], nblobs=8, noise=25.0, seed=g["seed"])
        assert sha(rgb) == g["image_sha256"]
        assert sha(sd) == g["seeds_sha256"]
        cs, ct, caps, info = ol.model_image(rgb, sd, 4, k=5, lam=50.0)
        assert sha(np.concatenate([cs, ct, caps.ravel()])) == g["caps_sha256"]
        f, l, ok = ol.maxflow(g["W"], g["H"], 1, 4, cs, ct, caps)
        assert ok == 1
        assert f == g["flow"]
        assert sha(l) == g["labels_sha256"]
        if ol.have_bk():
            fb, lb = ol.bk303(g["W"], g["H"], 1, 4, cs, ct, caps)
            assert fb == f and np.array_equal(lb, l)


def load_flower_cut():
    from conftest import GOLDEN
    z = np.load(os.path.join(GOLDEN, "ref_flower_cut.npz"))
    n = int(z["n"])
    return {k: np.unpackbits(z[k])[:n] for k in ("fg", "bg", "tag")}


def test_tag_oracle_flower_golden():
    """The tag channel of flower's .cut is the reference's tag post-process of
    its foreground seeds and object channel."""
    nx, ny, ef, cs, ct, caps, flow, lab = load_ref_fixture("flower")
    ch = load_flower_cut()
    assert np.array_equal(ol.tag(nx, ny, 1, ch["fg"], lab), ch["tag"])
    assert ch["tag"].sum() == 52721


@pytest.mark.parametrize("shape", [(41, 37, 1), (13, 11, 9)])
def test_tag_oracle_vs_connected_components(shape):
    """Closed form: connected components (axis adjacency) of object | fg that
    hold a foreground seed, intersected with object | fg."""
    from scipy import ndimage
    nx, ny, nz = shape
    rng = np.random.default_rng(nx)
    for trial in range(5):
        obj = (rng.random((nz, ny, nx)) < 0.55).astype(np.uint8)
        fg = (rng.random((nz, ny, nx)) < 0.01).astype(np.uint8)
        passable = (obj | fg).astype(bool)
        st = ndimage.generate_binary_structure(3 if nz > 1 else 2, 1)
        comp, _ = ndimage.label(passable if nz > 1 else passable[0], structure=st)
        comp = comp.reshape(nz, ny, nx)
        keep = np.unique(comp[fg.astype(bool)])
        want = (np.isin(comp, keep[keep > 0]) & passable).astype(np.uint8)
        assert np.array_equal(ol.tag(nx, ny, nz, fg, obj), want.ravel())
