"""Graph-file I/O of libjfcut (host code, CPU tests).

MFI files are checked against the reference's own reader/writer
(benchmark/GridCut/mfi.cpp, built into oracle/_ref/libmfiref.so) and against
the reference's way of turning an MFI file into a BK303 graph
(bench-BK303.cpp:39-46, :87-96): a file written by libjfcut must give the
oracle's max-flow on the planes it was written from.  The raw .ef/.c1/.c2/.cut
files are checked byte for byte against the reference's data files when the
reference tree is mounted, and against the committed fixtures always.
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT, load_ref_fixture, random_grid
import oracle_lib as ol

MFI_SO = os.path.join(ROOT, "oracle", "_ref", "libmfiref.so")
REF_DATA = "/root/reference/src/QVisualizer/data"
# MFI arc -> jfcut direction (mfi.h:36-53 against include/jfcut.h's order)
MFI2D = [1, 0, 3, 2, 5, 7, 6, 4]
MFI3D = [1, 0, 3, 2, 5, 4]


def mfi_ref():
    if not os.path.exists(MFI_SO):
        pytest.skip("oracle/_ref/libmfiref.so not built (reference tree absent)")
    L = ctypes.CDLL(MFI_SO)
    P = ctypes.c_void_p
    L.mfi_ref_read.argtypes = [ctypes.c_char_p, P, P, P, P, P]
    L.mfi_ref_read.restype = ctypes.c_int
    L.mfi_ref_write.argtypes = [ctypes.c_char_p] + [ctypes.c_int] * 5 + [
        ctypes.c_int64, ctypes.c_int, ctypes.c_int, P, P, P]
    L.mfi_ref_write.restype = ctypes.c_int
    L.mfi_ref_bk303.argtypes = [ctypes.c_char_p]
    L.mfi_ref_bk303.restype = ctypes.c_int64
    return L


def _ref_read(L, path, n, conn):
    hdr = np.zeros(9, np.int32)
    mf = np.zeros(1, np.int64)
    cs = np.zeros(n, np.int32)
    ct = np.zeros(n, np.int32)
    caps = np.zeros((conn, n), np.int32)
    assert L.mfi_ref_read(os.fsencode(path), ol.vp(hdr), ol.vp(mf), ol.vp(cs), ol.vp(ct),
                          ol.vp(caps)) == 0
    return hdr, int(mf[0]), cs, ct, caps


CASES = [(37, 23, 1, 4, 50, 300), (21, 17, 1, 8, 300, 70000), (11, 9, 7, 6, 40, 200),
         (64, 48, 1, 4, 70000, 1 << 30)]


@pytest.mark.parametrize("nx,ny,nz,conn,cmax,tmax", CASES)
def test_mfi_write_matches_reference_reader(tmp_path, jfcut_mod, nx, ny, nz, conn, cmax, tmax):
    L = mfi_ref()
    rng = np.random.default_rng(nx * ny + conn)
    cs, ct, caps = random_grid(rng, nx, ny, nz, conn, cap_max=cmax, term_max=tmax)
    n = nx * ny * nz
    path = str(tmp_path / "g.mfi")
    jfcut_mod.write_mfi(path, cs.reshape(nz, ny, nx), ct.reshape(nz, ny, nx),
                        caps.reshape(conn, nz, ny, nx), conn, maxflow=12345)
    hdr, mf, rcs, rct, rcaps = _ref_read(L, path, n, conn)
    assert list(hdr[:5]) == [3 if nz > 1 else 2, conn, nx, ny, nz]
    assert mf == 12345 and hdr[6] == n
    # smallest type that holds the values
    want_t = 0 if max(cs.max(), ct.max()) < 256 else 1 if max(cs.max(), ct.max()) < 65536 else 2
    want_n = 0 if caps.max() < 256 else 1 if caps.max() < 65536 else 2
    assert (hdr[7], hdr[8]) == (want_t, want_n)
    assert np.array_equal(rcs, cs) and np.array_equal(rct, ct)
    amap = MFI3D if nz > 1 else MFI2D
    for a in range(conn):
        assert np.array_equal(rcaps[a], caps[amap[a]]), a
    # the reference's graph construction over the file gives the oracle's flow
    if conn in (4, 6):
        flow_ref = L.mfi_ref_bk303(os.fsencode(path))
        flow_orc = ol.maxflow(nx, ny, nz, conn, cs, ct, caps)[0]
        assert flow_ref == flow_orc


@pytest.mark.parametrize("nx,ny,nz,conn,cmax,tmax", CASES)
def test_mfi_read_of_reference_writer(tmp_path, jfcut_mod, nx, ny, nz, conn, cmax, tmax):
    """reference writer (every type) -> libjfcut reader -> libjfcut writer ->
    reference reader: identical planes, header and flow."""
    L = mfi_ref()
    rng = np.random.default_rng(7 + nx)
    cs, ct, caps = random_grid(rng, nx, ny, nz, conn, cap_max=cmax, term_max=tmax)
    n = nx * ny * nz
    amap = MFI3D if nz > 1 else MFI2D
    mfi_caps = np.stack([caps[amap[a]] for a in range(conn)])
    tt = 2 if max(cs.max(), ct.max()) >= 65536 else 1
    tn = 2 if caps.max() >= 65536 else 1
    src = str(tmp_path / "ref.mfi")
    assert L.mfi_ref_write(os.fsencode(src), 3 if nz > 1 else 2, conn, nx, ny, nz, 999, tt, tn,
                           ol.vp(cs), ol.vp(ct), ol.vp(np.ascontiguousarray(mfi_caps))) == 0
    g = jfcut_mod.read_mfi(src)
    assert g["shape"] == (nx, ny, nz) and g["connectivity"] == conn and g["maxflow"] == 999
    assert np.array_equal(g["cs"].ravel(), cs) and np.array_equal(g["ct"].ravel(), ct)
    assert np.array_equal(g["caps"].reshape(conn, n), caps)
    dst = str(tmp_path / "mine.mfi")
    jfcut_mod.write_mfi(dst, g["cs"], g["ct"], g["caps"], conn, maxflow=g["maxflow"])
    _, mf, rcs, rct, rcaps = _ref_read(L, dst, n, conn)
    assert mf == 999 and np.array_equal(rcaps, mfi_caps)
    assert np.array_equal(rcs, cs) and np.array_equal(rct, ct)


def test_mfi_rejects_bad_input(tmp_path, jfcut_mod):
    p = tmp_path / "bad.mfi"
    p.write_bytes(b"XYZ" + bytes(5000))
    with pytest.raises(jfcut_mod.JFCError):
        jfcut_mod.read_mfi(str(p))
    p.write_bytes(b"MFI" + bytes(100))  # truncated header
    with pytest.raises(jfcut_mod.JFCError):
        jfcut_mod.read_mfi(str(p))
    with pytest.raises(jfcut_mod.JFCError):  # negative capacity
        jfcut_mod.write_mfi(str(tmp_path / "n.mfi"), -np.ones((4, 4), np.int32),
                            np.zeros((4, 4), np.int32), np.zeros((4, 1, 4, 4), np.int32), 4)
    with pytest.raises(jfcut_mod.JFCError):  # 2D-6C is not a grid connectivity
        jfcut_mod.write_mfi(str(tmp_path / "c.mfi"), np.ones((4, 4), np.int32),
                            np.zeros((4, 4), np.int32), np.zeros((6, 1, 4, 4), np.int32), 6)


@pytest.mark.parametrize("name", ["flower", "person", "sponge"])
def test_raw_graph_roundtrip_fixture(tmp_path, jfcut_mod, name):
    """Fixture -> .ef/.c1/.c2 -> reader: same planes and the published flow."""
    nx, ny, ef, cs, ct, caps, flow, lab = load_ref_fixture(name)
    pre = str(tmp_path / name)
    jfcut_mod.write_raw_graph(pre, ef.reshape(ny, nx), caps.reshape(4, ny, nx))
    assert os.path.getsize(pre + ".ef") == 4 * nx * ny
    assert os.path.getsize(pre + ".c1") == 4 * nx * ny
    ef2, caps2 = jfcut_mod.read_raw_graph(pre, nx, ny)
    assert np.array_equal(ef2.ravel(), ef) and np.array_equal(caps2.reshape(4, -1), caps)
    assert ol.maxflow(nx, ny, 1, 4, np.maximum(ef2, 0).ravel(), np.maximum(-ef2, 0).ravel(),
                      caps2.reshape(4, -1))[0] == flow


@pytest.mark.parametrize("name", ["flower", "person", "sponge"])
def test_raw_graph_reference_files(tmp_path, jfcut_mod, name):
    """The reference's own data files: reader gives the fixture planes (whose
    flow is the published one), the writer reproduces the files byte for byte,
    and the .cut object channel of flower is the canonical cut."""
    pre = f"{REF_DATA}/{name}/normalized_{name}.raw"
    if not os.path.exists(pre + ".ef"):
        pytest.skip("reference data files absent")
    nx, ny, ef, cs, ct, caps, flow, lab = load_ref_fixture(name)
    ef2, caps2 = jfcut_mod.read_raw_graph(pre, nx, ny)
    assert np.array_equal(ef2.ravel(), ef) and np.array_equal(caps2.reshape(4, -1), caps)
    out = str(tmp_path / name)
    jfcut_mod.write_raw_graph(out, ef2, caps2)
    for ext in (".ef", ".c1", ".c2"):
        assert open(out + ext, "rb").read() == open(pre + ext, "rb").read(), ext
    fg, bg, obj, tag = jfcut_mod.read_cut(pre + ".cut", nx * ny)
    if name == "flower":
        assert np.array_equal(obj, lab)
        assert np.array_equal(tag, ol.tag(nx, ny, 1, fg, obj))
    jfcut_mod.write_cut(out + ".cut", fg, bg, obj, tag)
    assert open(out + ".cut", "rb").read() == open(pre + ".cut", "rb").read()


def test_raw_graph_3d_and_limits(tmp_path, jfcut_mod):
    rng = np.random.default_rng(3)
    nx, ny, nz = 9, 7, 5
    ef = rng.integers(-1000, 1000, (nz, ny, nx)).astype(np.int32)
    caps = rng.integers(0, 256, (6, nz, ny, nx)).astype(np.int32)
    pre = str(tmp_path / "v")
    jfcut_mod.write_raw_graph(pre, ef, caps)
    # uchar4 per node: c1 = (+x, +y, +z, 0), c2 = (-z, -y, -x, 0)
    c1 = np.fromfile(pre + ".c1", np.uint8).reshape(-1, 4)
    c2 = np.fromfile(pre + ".c2", np.uint8).reshape(-1, 4)
    flat = caps.reshape(6, -1)
    assert np.array_equal(c1[:, :3], flat[[0, 2, 4]].T) and not c1[:, 3].any()
    assert np.array_equal(c2[:, :3], flat[[5, 3, 1]].T) and not c2[:, 3].any()
    ef2, caps2 = jfcut_mod.read_raw_graph(pre, nx, ny, nz)
    assert np.array_equal(ef2, ef) and np.array_equal(caps2, caps)
    caps[2, 0, 0, 0] = 256
    with pytest.raises(jfcut_mod.JFCError):
        jfcut_mod.write_raw_graph(pre, ef, caps)
    with pytest.raises(jfcut_mod.JFCError):
        jfcut_mod.read_raw_graph(str(tmp_path / "missing"), nx, ny, nz)


def test_cut_roundtrip(tmp_path, jfcut_mod):
    rng = np.random.default_rng(5)
    n = 1000
    fg, bg, obj = (rng.random(n) < 0.3).astype(np.uint8), (rng.random(n) < 0.3).astype(
        np.uint8), (rng.random(n) < 0.5).astype(np.uint8)
    p = str(tmp_path / "x.cut")
    tag = obj & (rng.random(n) < 0.5).astype(np.uint8)
    jfcut_mod.write_cut(p, fg, bg, obj, tag)
    raw = np.fromfile(p, np.uint8).reshape(-1, 4)
    assert set(np.unique(raw)) <= {0, 127}
    for k, a in enumerate((fg, bg, obj, tag)):
        assert np.array_equal(raw[:, k] != 0, a.astype(bool)), k
    f2, b2, o2, t2 = jfcut_mod.read_cut(p, n)
    assert np.array_equal(f2, fg) and np.array_equal(b2, bg) and np.array_equal(o2, obj)
    assert np.array_equal(t2, tag)
    jfcut_mod.write_cut(p, obj=obj)
    f3, b3, o3, t3 = jfcut_mod.read_cut(p, n)
    assert not f3.any() and not b3.any() and not t3.any() and np.array_equal(o3, obj)


# cl_node_2d / cl_node_3d records (QGCSetting.h:31-43) as numpy record types
NODE2D = np.dtype([("height", "<i4"), ("ef", "<i4"), ("cap", "<u2", 4)])
NODE3D = np.dtype([("height", "<i4"), ("ef", "<i4"), ("cap", "u1", 8)])


@pytest.mark.parametrize("name", ["flower", "sponge"])
def test_node_file_2d(tmp_path, jfcut_mod, name):
    """Fixture -> .node (padded to 16x16 tiles) -> reader: the loader of
    QGCW:2653-2665 maps record cap[0..3] to c1.x, c1.y, c2.x, c2.y, i.e.
    +x, +y, -y, -x; the graph keeps its published flow."""
    nx, ny, ef, cs, ct, caps, flow, lab = load_ref_fixture(name)
    p = str(tmp_path / f"{name}.node")
    h = np.arange(nx * ny, dtype=np.int32).reshape(ny, nx) % 1000
    jfcut_mod.write_node(p, ef.reshape(ny, nx), caps.reshape(4, ny, nx), height=h)
    px, py = -(-nx // 16) * 16, -(-ny // 16) * 16
    rec = np.fromfile(p, NODE2D)
    assert rec.size == px * py
    rec = rec.reshape(py, px)
    assert not rec[ny:].view(np.uint8).any() and not rec[:, nx:].view(np.uint8).any()
    body = rec[:ny, :nx]
    assert np.array_equal(body["ef"].ravel(), ef) and np.array_equal(body["height"], h)
    c = caps.reshape(4, ny, nx)
    for k, d in enumerate([0, 2, 3, 1]):
        assert np.array_equal(body["cap"][..., k], c[d]), k
    ef2, h2, caps2 = jfcut_mod.read_node(p, nx, ny)
    assert np.array_equal(ef2.ravel(), ef) and np.array_equal(h2[0], h)
    assert np.array_equal(caps2.reshape(4, -1), caps)
    assert ol.maxflow(nx, ny, 1, 4, np.maximum(ef2, 0).ravel(), np.maximum(-ef2, 0).ravel(),
                      caps2.reshape(4, -1))[0] == flow


def test_node_file_3d(tmp_path, jfcut_mod):
    rng = np.random.default_rng(11)
    nx, ny, nz = 13, 10, 6
    ef = rng.integers(-500, 500, (nz, ny, nx)).astype(np.int32)
    caps = rng.integers(0, 256, (6, nz, ny, nx)).astype(np.int32)
    p = str(tmp_path / "v.node")
    jfcut_mod.write_node(p, ef, caps)
    rec = np.fromfile(p, NODE3D).reshape(8, 16, 16)  # padded to 8x8x4 tiles
    body = rec[:nz, :ny, :nx]
    assert np.array_equal(body["ef"], ef) and not body["height"].any()
    # QGCW:2697-2712: cap[0..2] -> c1 (+x, +y, +z), cap[3..5] -> c2 (-z, -y, -x)
    for k, d in enumerate([0, 2, 4, 5, 3, 1]):
        assert np.array_equal(body["cap"][..., k], caps[d]), k
    assert not body["cap"][..., 6:].any()
    ef2, h2, caps2 = jfcut_mod.read_node(p, nx, ny, nz)
    assert np.array_equal(ef2, ef) and np.array_equal(caps2, caps) and not h2.any()
    # other tile padding
    jfcut_mod.write_node(p, ef, caps, block=(4, 4, 2))
    assert os.path.getsize(p) == 16 * 16 * 12 * 6
    ef3, _, caps3 = jfcut_mod.read_node(p, nx, ny, nz, block=(4, 4, 2))
    assert np.array_equal(ef3, ef) and np.array_equal(caps3, caps)
    with pytest.raises(jfcut_mod.JFCError):  # file shorter than the padded grid
        jfcut_mod.read_node(p, nx, ny, nz)


def test_mfi_export_like_widget(tmp_path, jfcut_mod):
    """saveMFI (QGCW:3339-3473) on flower: uint32 terminals, uint16 arcs,
    symmetric, active_count = #(ef >= 0), mask codes 1/2/3/4 from the cut
    channels; the reference reader and BK303 construction give the published
    flow."""
    L = mfi_ref()
    nx, ny, ef, cs, ct, caps, flow, lab = load_ref_fixture("flower")
    n = nx * ny
    rng = np.random.default_rng(2)
    fg = (rng.random(n) < 0.05).astype(np.uint8)
    bg = (rng.random(n) < 0.05).astype(np.uint8)
    tag = (rng.random(n) < 0.5).astype(np.uint8)
    path = str(tmp_path / "flower.mfi")
    jfcut_mod.export_mfi(path, ef.reshape(ny, nx), caps.reshape(4, ny, nx), fg=fg, bg=bg,
                         obj=lab, tag=tag, maxflow=flow)
    hdr, mf, rcs, rct, rcaps = _ref_read(L, path, n, 4)
    assert list(hdr) == [2, 4, nx, ny, 1, 1, int((ef >= 0).sum()), 2, 1]
    assert mf == flow
    assert np.array_equal(rcs, np.maximum(ef, 0)) and np.array_equal(rct, np.maximum(-ef, 0))
    for a in range(4):
        assert np.array_equal(rcaps[a], caps[MFI2D[a]])
    mask = np.fromfile(path, np.uint8)[-n:]
    want = np.where(fg == 1, 1, np.where(bg == 1, 2, np.where(lab == 1, 3,
                                                              np.where(tag == 1, 4, 0))))
    assert np.array_equal(mask, want)
    assert L.mfi_ref_bk303(os.fsencode(path)) == flow
    big = caps.reshape(4, ny, nx).copy()
    big[0, 0, 0] = 65536
    with pytest.raises(jfcut_mod.JFCError, match="does not fit"):
        jfcut_mod.export_mfi(path, ef.reshape(ny, nx), big)
