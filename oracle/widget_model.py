"""CPU oracle of the widget's own data terms (the "Model" button of JF-Cut's
QGCWidget, src/QVisualizer/Graph Cut/QGCWidget.cpp:2560-2606): lazy snapping
and BK.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of jfc_model_widget,
never measured or shipped.  numpy float32 restatement, one IEEE-rounded
operation at a time in the reference's order (numpy never fuses a multiply
and an add); exp and log are taken in float64 and rounded once to float32.
The readings the library fixes where the reference is undefined or racy are
restated here too (highest node index per seeded level, exact centroid sums,
BK level clamp, NaN -> 0, saturation at +-2^30).

Citations are relative to src/QVisualizer/ of the JF-Cut tree:
  cl/GraphCut/graphcut_compute_capacity.cl, graphcut_compute_excessflow.cl,
  graphcut_histogram.cl, graphcut_kmeans.cl; Graph Cut/QGCWidget.cpp (QGCW).
"""
import numpy as np

F = np.float32
F64 = np.float64
SHRT_MAX = 32767
CHAR_MAX = 127
MAX_FLOW = 32767          # QGCSetting::cl_max_flow
LEVELS = 256              # QGCSetting::sizeCutHistogram
SAT = 1 << 30

DEFAULTS = dict(lazy=True, weight=(0.2126, 0.7152, 0.0722, 0.0), lazy_neighbor=1.0,
                lazy_terminal=1.0, bk_neighbor=0.002, bk_terminal=0.0, k=8,
                kmeans_iterations=32)


def normalize_weight(w):
    """glm::normalize(vec4) (QGCW:2004): v * (1 / sqrt((xx + yy) + (zz + ww)))."""
    v = np.asarray(w, dtype=F)
    s = (v[0] * v[0] + v[1] * v[1]) + (v[2] * v[2] + v[3] * v[3])
    inv = F(1) / np.sqrt(s)
    return (v * inv).astype(F)


def unorm8(b):
    """read_imagef of CL_UNORM_INT8: byte / 255."""
    return np.asarray(b).astype(F) / F(255)


def dot4(f, w):
    """OpenCL dot(f, w), left to right; f is (..., 4)."""
    return ((f[..., 0] * w[0] + f[..., 1] * w[1]) + f[..., 2] * w[2]) + f[..., 3] * w[3]


def dist4(c, p):
    """OpenCL distance(c, p): sqrt of the left-to-right sum of squares."""
    d = c - p
    return np.sqrt(((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])
                   + d[..., 3] * d[..., 3])


def round_half_away(v):
    """round() of OpenCL / C on float32 values, exactly."""
    x = np.asarray(v, dtype=F64)
    return np.where(x >= 0, np.floor(x + 0.5), np.ceil(x - 0.5))


def unary_int(v):
    """float unary -> int32: round, NaN -> 0, saturate at +-2^30."""
    v = np.asarray(v, dtype=F)
    r = round_half_away(np.where(np.isnan(v), F(0), v))
    return np.clip(r, -SAT, SAT).astype(np.int64).astype(np.int32)


def seeded_levels(rgba, seeds, w):
    """graphcut_cut_histogram / graphcut_cut_distribution
    (graphcut_histogram.cl:19-60): per seeded pixel the level
    int(clamp(dot, 0, 1) * 255 + 0.5); returns the 256-bin histogram and, per
    level, the highest seeded node index + 1 (0 = level unused)."""
    idx = np.flatnonzero(np.asarray(seeds).ravel() != 0)
    f = unorm8(rgba.reshape(-1, 4)[idx])
    data = np.clip(dot4(f, w), F(0), F(1))
    lv = (data * F(LEVELS - 1) + F(0.5)).astype(np.int64)
    hist = np.bincount(lv, minlength=LEVELS).astype(np.int64)
    rep = np.zeros(LEVELS, np.int64)
    np.maximum.at(rep, lv, idx + 1)
    return hist, rep, idx


def kmeans(rgba, seeds, w, k, iterations):
    """QGCWidget::kmeans (QGCW:1997-2107) with graphcut_kmeans
    (graphcut_kmeans.cl:21-159).  Returns (centroids the unary uses, k x 4
    float32; iterations run)."""
    hist, rep, idx = seeded_levels(rgba, seeds, w)
    dist = rep[rep > 0]
    if dist.size == 0:
        raise ValueError("no labels")
    dsize = dist.size - 1
    px = rgba.reshape(-1, 4)
    scale = (w / F(255)).astype(F)
    cent = np.zeros((k, 4), F)
    for i in range(k):
        node = int(dist[dsize * i // (k - 1)]) - 1
        cent[i] = px[node].astype(F) * scale
    t = px[idx].astype(np.int64)                     # seeded texels
    data = (unorm8(t) * w).astype(F)                 # read_imagef * weight
    val = (unorm8(np.arange(LEVELS))[None, :] * w[:, None]).astype(F)   # [c][b]
    prev = cent
    it = 0
    while it < iterations:
        prev = cent.copy()
        d = dist4(prev[None, :, :], data[:, None, :])     # (n, k)
        lab = np.argmin(d, axis=1)                        # first minimum
        it += 1
        new = np.zeros((k, 4), F)
        err = F(0)
        for i in range(k):
            members = t[lab == i]
            if members.shape[0]:
                for c in range(4):
                    cnt = np.bincount(members[:, c], minlength=LEVELS).astype(F64)
                    terms = cnt * val[c].astype(F64)
                    s = float(np.cumsum(terms)[-1])       # sequential, byte order
                    new[i, c] = F(s / float(members.shape[0]))
            dd = new[i] - prev[i]
            ln = np.sqrt((dd[0] * dd[0] + dd[1] * dd[1]) + (dd[2] * dd[2] + dd[3] * dd[3]))
            err = max(err, F(ln))
        cent = new
        if err < F(1.0 / 1024):
            break
    return prev, it


def pairwise(rgba, shape, w, lazy, param):
    """graphcut_compute_capacity_lazy / _bk (graphcut_compute_capacity.cl:21-254)
    as library planes: (K, N) int32, plane d = arc toward v + dir(d) with
    0 +x, 1 -x, 2 +y, 3 -y, 4 +z, 5 -z; border arcs 0."""
    nx, ny, nz = shape
    d3 = nz > 1
    px = rgba.reshape(nz, ny, nx, 4)
    data = np.trunc((unorm8(px) * w) * F(SHRT_MAX)).astype(np.int64)
    top = F(CHAR_MAX if d3 else SHRT_MAX)
    if lazy:
        scale = F(param) / F(SHRT_MAX)
    else:
        scale = F(1) / (F(SHRT_MAX) * F(param))

    def arc(a, b):
        diff = np.abs(a - b).astype(np.uint64)
        s = (((diff[..., 0] * diff[..., 0] + diff[..., 1] * diff[..., 1])
              + diff[..., 2] * diff[..., 2]) + diff[..., 3] * diff[..., 3])
        s = (s & 0xffffffff).astype(np.uint32)
        x = np.sqrt(s.astype(F)) * scale
        if lazy:
            v = top / (F(1) + x) + F(0.5)
        else:
            e = np.exp(((F(-0.5) * x) * x).astype(F64)).astype(F)
            v = top * e + F(0.5)
        return np.trunc(v).astype(np.int32)

    K = 6 if d3 else 4
    out = np.zeros((K, nz, ny, nx), np.int32)
    dirs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)][:K]
    for d, (dx, dy, dz) in enumerate(dirs):
        zs, ys, xs = slice(max(0, -dz), nz - max(0, dz)), slice(max(0, -dy), ny - max(0, dy)), \
            slice(max(0, -dx), nx - max(0, dx))
        zt, yt, xt = slice(zs.start + dz, zs.stop + dz), slice(ys.start + dy, ys.stop + dy), \
            slice(xs.start + dx, xs.stop + dx)
        out[d, zs, ys, xs] = arc(data[zs, ys, xs], data[zt, yt, xt])
    return out.reshape(K, -1)


def unary_lazy(rgba, fg, bg, w, cf, cb, lam, d3):
    """graphcut_compute_excessflow_lazy (graphcut_compute_excessflow.cl:47-77,
    109-139): ef = round((cl_max_flow >> s) * lambda * (dB - dF) / (dF + dB))."""
    p = (unorm8(rgba.reshape(-1, 4)) * w).astype(F)
    dx = np.full(p.shape[0], F(3.402823466e38), F)
    dy = dx.copy()
    for j in range(cf.shape[0]):
        dx = np.minimum(dx, dist4(cf[j][None, :], p))
        dy = np.minimum(dy, dist4(cb[j][None, :], p))
    den = dx + dy
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.where(den == F(0), F(0), (dy - dx) / np.where(den == F(0), F(1), den))
    gain = F(MAX_FLOW >> (3 if d3 else 2)) * F(lam)
    ef = unary_int(gain * r.astype(F))
    return _seeded(ef, fg, bg)


def unary_bk(rgba, fg, bg, w, hist_f, hist_b, lam, d3):
    """graphcut_compute_excessflow_bk (graphcut_compute_excessflow.cl:19-45,
    79-107): ef = round((cl_max_flow >> s) * lambda * (-log(hB/sumB) + log(hF/sumF)))."""
    f = unorm8(rgba.reshape(-1, 4))
    lv = (dot4(f, w) * F(LEVELS - 1) + F(0.5)).astype(np.int64)
    lv = np.clip(lv, 0, LEVELS - 1)
    sf, sb = F(hist_f.sum()), F(hist_b.sum())
    qf = hist_f[lv].astype(F) / sf
    qb = hist_b[lv].astype(F) / sb
    with np.errstate(divide="ignore", invalid="ignore"):
        df = -np.log(qf.astype(F64)).astype(F)
        db = -np.log(qb.astype(F64)).astype(F)
        gain = F(MAX_FLOW >> (3 if d3 else 2)) * F(lam)
        v = gain * (db - df)
    ef = unary_int(v)
    return _seeded(ef, fg, bg)


def _seeded(ef, fg, bg):
    ef = ef.copy()
    fg = np.asarray(fg).ravel() != 0
    bg = np.asarray(bg).ravel() != 0
    ef[bg] = -MAX_FLOW
    ef[fg] = MAX_FLOW
    return ef


def model(rgba, fg, bg, shape, **kw):
    """slotModelClicked (QGCW:2560-2606).  Returns (cs, ct, planes, info) with
    cs = max(ef, 0), ct = max(-ef, 0)."""
    p = dict(DEFAULTS)
    p.update(kw)
    w = normalize_weight(p["weight"])
    d3 = shape[2] > 1
    rgba = np.ascontiguousarray(rgba, np.uint8).reshape(-1, 4)
    info = {}
    if p["lazy"]:
        planes = pairwise(rgba, shape, w, True, p["lazy_neighbor"])
        cf, itf = kmeans(rgba, fg, w, p["k"], p["kmeans_iterations"])
        cb, itb = kmeans(rgba, bg, w, p["k"], p["kmeans_iterations"])
        info.update(centroids=(cf, cb), iterations=(itf, itb))
        ef = unary_lazy(rgba, fg, bg, w, cf, cb, p["lazy_terminal"], d3)
    else:
        planes = pairwise(rgba, shape, w, False, p["bk_neighbor"])
        hf = seeded_levels(rgba, fg, w)[0]
        hb = seeded_levels(rgba, bg, w)[0]
        info.update(hist=(hf, hb))
        ef = unary_bk(rgba, fg, bg, w, hf, hb, p["bk_terminal"], d3)
    cs = np.maximum(ef, 0).astype(np.int32)
    ct = np.maximum(-ef, 0).astype(np.int32)
    return cs, ct, planes, info
