"""ctypes access to the CPU oracle (oracle/liboracle.so) and to the reference's
own BK303 max-flow built from the reference sources (oracle/_ref/libbk303ref.so).

Test infrastructure: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, only as the checker or the CPU baseline.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
BK_SO = os.path.join(ROOT, "oracle", "_ref", "libbk303ref.so")

_orc = None
_bk = None
P = ctypes.c_void_p
I = ctypes.c_int
I64 = ctypes.c_int64


def vp(a):
    return None if a is None else a.ctypes.data_as(P)


def orc():
    global _orc
    if _orc is None:
        L = ctypes.CDLL(ORACLE_SO)
        L.orc_maxflow.restype = I64
        L.orc_maxflow.argtypes = [I, I, I, I, P, P, P, P, P, P, ctypes.POINTER(I)]
        L.orc_canonical_cut.restype = I
        L.orc_canonical_cut.argtypes = [I, I, I, I, P, P, P]
        L.orc_energy.restype = I64
        L.orc_energy.argtypes = [I, I, I, I, P, P, P, P]
        L.orc_kmeans_init.restype = None
        L.orc_kmeans_init.argtypes = [P, I, I, I, I, P]
        L.orc_kmeans_iterate.restype = I
        L.orc_kmeans_iterate.argtypes = [P, I, I, I, I, P]
        L.orc_unary.restype = None
        L.orc_unary.argtypes = [P, I64, I, P, I, P, I, I64, P, P, P]
        L.orc_pair_thresholds.restype = I
        L.orc_pair_thresholds.argtypes = [ctypes.c_double, I64, I64, P]
        L.orc_pairwise.restype = None
        L.orc_pairwise.argtypes = [I, I, I, I, P, I, ctypes.c_double, ctypes.c_double, I, P]
        L.orc_tag.restype = I
        L.orc_tag.argtypes = [I, I, I, P, P, P]
        _orc = L
    return _orc


def have_bk():
    return os.path.exists(BK_SO)


def bk():
    global _bk
    if _bk is None:
        L = ctypes.CDLL(BK_SO)
        L.bk303_grid_maxflow.restype = I64
        L.bk303_grid_maxflow.argtypes = [I, I, I, I, P, P, P, P, P, P, P]
        _bk = L
    return _bk


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def maxflow(nx, ny, nz, conn, cs, ct, caps, want_state=False):
    """Oracle max-flow: (flow, labels, certificate_ok[, e, residual])."""
    N = nx * ny * nz
    cs, ct, caps = _i32(cs), _i32(ct), _i32(caps).reshape(conn, N)
    lab = np.zeros(N, np.uint8)
    ok = I(0)
    e = np.zeros(N, np.int32) if want_state else None
    r = np.zeros((conn, N), np.int32) if want_state else None
    f = orc().orc_maxflow(nx, ny, nz, conn, vp(cs), vp(ct), vp(caps), vp(e), vp(r), vp(lab),
                          ctypes.byref(ok))
    if f < 0:
        raise MemoryError("oracle allocation failed")
    if want_state:
        return f, lab, ok.value, e, r
    return f, lab, ok.value


def canonical_cut(nx, ny, nz, conn, e, res):
    N = nx * ny * nz
    lab = np.zeros(N, np.uint8)
    ok = orc().orc_canonical_cut(nx, ny, nz, conn, vp(_i32(e)), vp(_i32(res).reshape(conn, N)),
                                 vp(lab))
    return lab, ok


def tag(nx, ny, nz, fg, obj):
    """Oracle tag post-process (graphcut_compute_tag.cl semantics)."""
    fg = np.ascontiguousarray(fg, np.uint8).ravel()
    obj = np.ascontiguousarray(obj, np.uint8).ravel()
    out = np.zeros(fg.size, np.uint8)
    assert orc().orc_tag(nx, ny, nz, vp(fg), vp(obj), vp(out)) == 0
    return out


def energy(nx, ny, nz, conn, cs, ct, caps, labels):
    N = nx * ny * nz
    return orc().orc_energy(nx, ny, nz, conn, vp(_i32(cs)), vp(_i32(ct)),
                            vp(_i32(caps).reshape(conn, N)),
                            vp(np.ascontiguousarray(labels, dtype=np.uint8)))


def bk303(nx, ny, nz, conn, cs, ct, caps, timings=False):
    """The reference's BK303: (flow, canonical labels[, (t_init, t_maxflow, t_output)])."""
    N = nx * ny * nz
    lab = np.zeros(N, np.uint8)
    t = [ctypes.c_double(), ctypes.c_double(), ctypes.c_double()]
    f = bk().bk303_grid_maxflow(nx, ny, nz, conn, vp(_i32(cs)), vp(_i32(ct)),
                                vp(_i32(caps).reshape(conn, N)), vp(lab),
                                ctypes.byref(t[0]), ctypes.byref(t[1]), ctypes.byref(t[2]))
    if timings:
        return f, lab, tuple(x.value for x in t)
    return f, lab


def kmeans(feat, K, wide, iters=32):
    feat = _i32(feat)
    n, C = feat.shape
    cent = np.zeros(K * C, np.int64)
    orc().orc_kmeans_init(vp(feat), n, C, wide, K, vp(cent))
    init = cent.copy()
    it = orc().orc_kmeans_iterate(vp(feat), n, C, K, iters, vp(cent))
    return cent.reshape(K, C), init.reshape(K, C), it


def pair_thresholds(lam, S, P_):
    T = np.zeros(256, np.int64)
    n = orc().orc_pair_thresholds(float(lam), int(S), int(P_), vp(T))
    return T[:n]


def model_image(img, seeds, conn, k=5, k_bg=None, lam=50.0, lam_z=None, z_sep=0, R=255,
                iters=32, unary_scale=100):
    """Oracle data term: (cs, ct, caps[K, N], info) for an image of shape
    (H, W, 3) / (T, H, W, 3) uint8 or (D, H, W) uint16 and seeds (1 fg, 2 bg)."""
    assert unary_scale == 100
    img = np.asarray(img)
    wide = 1 if img.dtype == np.uint16 else 0
    C = 3 if (img.ndim >= 3 and img.shape[-1] == 3 and not wide) else 1
    spatial = img.shape[:-1] if C == 3 else img.shape
    if len(spatial) == 2:
        ny, nx = spatial
        nz = 1
    else:
        nz, ny, nx = spatial
    N = nx * ny * nz
    feat = img.reshape(N, C).astype(np.int32)
    sd = np.asarray(seeds).reshape(N)
    kb = k if k_bg is None else k_bg
    cfg, ifg, itf = kmeans(feat[sd == 1], k, wide, iters)
    cbg, ibg, itb = kmeans(feat[sd == 2], kb, wide, iters)
    cs = np.zeros(N, np.int32)
    ct = np.zeros(N, np.int32)
    orc().orc_unary(vp(feat), N, C, vp(np.ascontiguousarray(cfg.ravel())), k,
                    vp(np.ascontiguousarray(cbg.ravel())), kb, R, vp(np.ascontiguousarray(sd)),
                    vp(cs), vp(ct))
    caps = np.zeros((conn, N), np.int32)
    orc().orc_pairwise(nx, ny, nz, conn, vp(feat), C, float(lam),
                       float(lam if lam_z is None else lam_z), int(z_sep), vp(caps))
    return cs, ct, caps, {"cfg": cfg, "cbg": cbg, "init_fg": ifg, "init_bg": ibg,
                          "iters_fg": itf, "iters_bg": itb, "shape": (nx, ny, nz)}
