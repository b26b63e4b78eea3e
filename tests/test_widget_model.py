"""Pins of the CPU oracle of the widget's own data terms (oracle/widget_model.py)
against what the definitions fix: closed forms, special cases and invariants
of graphcut_compute_capacity.cl, graphcut_compute_excessflow.cl,
graphcut_histogram.cl and QGCWidget::kmeans (QGCWidget.cpp:1997-2107).
The reference's float OpenCL cannot run here and ships no outputs of these
kernels (its fixture .ef/.c1/.c2 files come from another data term), so the
pins are mathematical."""
import math
import os
import sys
from fractions import Fraction

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import widget_model as wm  # noqa: E402

F = np.float32


def rgba_of(grey):
    g = np.asarray(grey, np.uint8)
    return np.repeat(g[..., None], 4, axis=-1)


def test_weight_normalisation():
    w = wm.normalize_weight((0.2126, 0.7152, 0.0722, 0.0))
    assert abs(float(np.sqrt(np.sum(w.astype(np.float64) ** 2))) - 1.0) < 1e-6
    w = wm.normalize_weight((1, 1, 1, 0))
    assert np.allclose(w, [1 / math.sqrt(3)] * 3 + [0], atol=1e-7)
    assert wm.normalize_weight((0, 2, 0, 0)).tolist() == [0, 1, 0, 0]


def test_levels_identity_and_clamp():
    # weight (1, 0, 0, 0): level == byte for every byte value
    rgba = np.zeros((256, 4), np.uint8)
    rgba[:, 0] = np.arange(256)
    w = wm.normalize_weight((1, 0, 0, 0))
    hist, rep, idx = wm.seeded_levels(rgba, np.ones(256, np.uint8), w)
    assert hist.tolist() == [1] * 256
    assert rep.tolist() == list(range(1, 257))
    # unit weights on grey: dot = sqrt(3) * b / 255 > 1 clamps to level 255
    w3 = wm.normalize_weight((1, 1, 1, 0))
    h2, _, _ = wm.seeded_levels(rgba_of([255, 200, 10]), np.ones(3, np.uint8), w3)
    assert h2[255] == 2 and h2[int(math.sqrt(3) * 10 + 0.5)] == 1


def test_pairwise_uniform_and_borders():
    nx, ny = 7, 5
    rgba = rgba_of(np.full((ny, nx), 93))
    w = wm.normalize_weight((1, 1, 1, 0))
    for lazy, par in ((True, 1.0), (False, 0.025)):
        pl = wm.pairwise(rgba, (nx, ny, 1), w, lazy, par).reshape(4, ny, nx)
        assert (pl[0, :, :-1] == 32767).all() and (pl[0, :, -1] == 0).all()
        assert (pl[1, :, 1:] == 32767).all() and (pl[1, :, 0] == 0).all()
        assert (pl[2, :-1, :] == 32767).all() and (pl[2, -1, :] == 0).all()
        assert (pl[3, 1:, :] == 32767).all() and (pl[3, 0, :] == 0).all()
    pl3 = wm.pairwise(rgba_of(np.full((3, 4, 5), 7)), (5, 4, 3), w, True, 1.0).reshape(6, 3, 4, 5)
    assert (pl3[4, :-1] == 127).all() and (pl3[4, -1] == 0).all()
    assert (pl3[5, 1:] == 127).all() and (pl3[5, 0] == 0).all()


def test_pairwise_symmetry_and_monotone():
    rng = np.random.default_rng(3)
    nx, ny = 23, 17
    rgba = rng.integers(0, 256, (ny, nx, 4), dtype=np.uint8)
    w = wm.normalize_weight((0.3, 0.5, 0.2, 0.1))
    for lazy, par in ((True, 2.5), (False, 0.05)):
        pl = wm.pairwise(rgba, (nx, ny, 1), w, lazy, par).reshape(4, ny, nx)
        assert (pl[0, :, :-1] == pl[1, :, 1:]).all()
        assert (pl[2, :-1, :] == pl[3, 1:, :]).all()
    # one channel ramp: larger steps never give larger weights
    steps = np.array([0, 0, 1, 3, 7, 15, 40, 90, 160, 255], np.uint8)
    ramp = np.zeros((1, 2 * steps.size, 4), np.uint8)
    ramp[0, 1::2, 0] = steps
    w1 = wm.normalize_weight((1, 0, 0, 0))
    for lazy, par in ((True, 4.0), (False, 0.1)):
        c = wm.pairwise(ramp, (2 * steps.size, 1, 1), w1, lazy, par)[0, 0::2]
        assert (np.diff(c.astype(np.int64)) <= 0).all()


def test_pairwise_closed_forms():
    # channel 0 from 0 to 255 under weight (1, 0, 0, 0): |diff| = 32767
    img = np.zeros((1, 2, 4), np.uint8)
    img[0, 1, 0] = 255
    w = wm.normalize_weight((1, 0, 0, 0))
    for lam in (1.0, 3.0, 0.5):
        c = int(wm.pairwise(img, (2, 1, 1), w, True, lam)[0, 0])
        assert abs(c - math.floor(32767 / (1 + lam) + 0.5)) <= 1
    # BK half maximum: t = 1 / sigma = sqrt(2 ln 2)
    sig = 1 / math.sqrt(2 * math.log(2))
    assert abs(int(wm.pairwise(img, (2, 1, 1), w, False, sig)[0, 0]) - 16384) <= 1
    img3 = np.zeros((2, 1, 1, 4), np.uint8)
    img3[1, 0, 0, 0] = 255
    assert abs(int(wm.pairwise(img3, (1, 1, 2), w, False, sig)[4, 0]) - 64) <= 1
    # far apart under a small sigma: 0
    assert int(wm.pairwise(img, (2, 1, 1), w, False, 0.01)[0, 0]) == 0


def test_unary_lazy_closed_forms():
    w = wm.normalize_weight((1, 0, 0, 0))
    rgba = np.zeros((4, 4), np.uint8)
    rgba[:, 0] = [50, 206, 128, 127]
    p = (wm.unorm8(rgba) * w).astype(F)
    cf = np.array([p[0]], F)          # pixel 0 is a foreground centroid
    cb = np.array([p[1]], F)          # pixel 1 a background centroid
    no = np.zeros(4, np.uint8)
    for lam, d3 in ((1.0, False), (0.25, False), (2.0, True)):
        ef = wm.unary_lazy(rgba, no, no, w, cf, cb, lam, d3)
        g = (32767 >> (3 if d3 else 2)) * lam
        assert ef[0] == round(g) and ef[1] == -round(g)
        # 128 is equidistant from 50 and 206; 127 has dF = 77, dB = 79
        assert ef[2] == 0
        assert abs(int(ef[3]) - round(g * 2 / 156)) <= 1
    fg = np.array([1, 0, 1, 0], np.uint8)
    bg = np.array([0, 1, 1, 1], np.uint8)
    ef = wm.unary_lazy(rgba, fg, bg, w, cf, cb, 1.0, False)
    assert ef.tolist() == [32767, -32767, 32767, -32767]


def test_unary_bk_closed_forms():
    w = wm.normalize_weight((1, 0, 0, 0))
    rgba = np.zeros((3, 4), np.uint8)
    rgba[:, 0] = [10, 20, 30]
    hf = np.zeros(256, np.int64)
    hb = np.zeros(256, np.int64)
    hf[10] = 8              # every foreground seed at level 10
    hb[10] = 4
    hb[20] = 4              # half of the background at 10, half at 20
    no = np.zeros(3, np.uint8)
    lam = 1.5
    ef = wm.unary_bk(rgba, no, no, w, hf, hb, lam, False)
    assert abs(int(ef[0]) - round(8191 * lam * math.log(2))) <= 1
    assert ef[1] == -(1 << 30)          # level absent from the foreground: saturates
    assert ef[2] == 0                   # absent from both: NaN -> 0
    assert (wm.unary_bk(rgba, no, no, w, hf, hb, 0.0, False) == 0).all()
    hu = np.full(256, 3, np.int64)      # identical histograms: 0 everywhere
    assert (wm.unary_bk(rgba, no, no, w, hu, hu, 7.0, True) == 0).all()


def test_kmeans_initial_picks():
    # seeded levels 10, 20, 30, 40, 50 (two pixels at 30); k = 3 picks the
    # representatives of levels 10, 30 and 50, the highest index at 30
    vals = [10, 30, 20, 30, 40, 50]
    rgba = np.zeros((len(vals), 4), np.uint8)
    rgba[:, 0] = vals
    rgba[3, 1] = 9           # tells the two level-30 pixels apart
    w = wm.normalize_weight((1, 0.001, 0, 0))
    assert wm.seeded_levels(rgba, np.ones(len(vals), np.uint8), w)[0][30] == 2
    cent, it = wm.kmeans(rgba, np.ones(len(vals), np.uint8), w, 3, 1)
    # one iteration: the returned centroids are the initial ones
    assert it == 1
    expect = [np.array([10, 0, 0, 0]), np.array([30, 9, 0, 0]), np.array([50, 0, 0, 0])]
    for i in range(3):
        assert np.array_equal(cent[i], expect[i].astype(F) * (w / F(255)))


def test_kmeans_converges_to_cluster_means():
    rng = np.random.default_rng(5)
    a = rng.integers(20, 40, 200)
    b = rng.integers(200, 230, 300)
    rgba = np.zeros((500, 4), np.uint8)
    rgba[:, 0] = np.concatenate([a, b])
    rgba[:, 1] = rng.integers(0, 256, 500)
    w = wm.normalize_weight((1, 0, 0, 0))
    cent, it = wm.kmeans(rgba, np.ones(500, np.uint8), w, 2, 50)
    assert 2 <= it < 50
    v = (wm.unorm8(np.arange(256)) * w[0]).astype(F)
    for grp, c in ((a, cent[0]), (b, cent[1])):
        exact = sum(Fraction(float(v[x])) for x in grp) / len(grp)
        assert abs(float(c[0]) - float(exact)) <= 2 * float(np.spacing(F(float(exact))))
        assert c[1] == 0 and c[2] == 0 and c[3] == 0
    # a fixed point: one more step moves no centroid by 1/1024
    c2, _ = wm.kmeans(rgba, np.ones(500, np.uint8), w, 2, it + 1)
    assert np.array_equal(c2, cent)


def test_kmeans_empty_cluster_is_zero():
    # k = 3 over one distinct level: the three picks coincide, ties go to the
    # first centroid, the other two lose every member and become 0
    rgba = np.zeros((5, 4), np.uint8)
    rgba[:, 0] = 77
    w = wm.normalize_weight((1, 0, 0, 0))
    cent, it = wm.kmeans(rgba, np.ones(5, np.uint8), w, 3, 2)
    assert it == 2
    assert float(cent[0, 0]) == float((wm.unorm8(77) * w[0]).astype(F))
    assert (cent[1:] == 0).all()


def test_model_end_to_end_shapes():
    rng = np.random.default_rng(1)
    nx, ny = 30, 20
    rgba = rng.integers(0, 256, (ny * nx, 4), dtype=np.uint8)
    fg = np.zeros(ny * nx, np.uint8)
    bg = np.zeros(ny * nx, np.uint8)
    fg[:40] = 1
    bg[-40:] = 1
    for lazy in (True, False):
        cs, ct, pl, info = wm.model(rgba, fg, bg, (nx, ny, 1), lazy=lazy, bk_terminal=1.0)
        assert pl.shape == (4, nx * ny)
        assert (cs[:40] == 32767).all() and (ct[-40:] == 32767).all()
        assert ((cs == 0) | (ct == 0)).all()
    with pytest.raises(ValueError):
        wm.kmeans(rgba, np.zeros(ny * nx, np.uint8), wm.normalize_weight((1, 1, 1, 0)), 2, 3)
