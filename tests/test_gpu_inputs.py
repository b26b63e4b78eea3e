"""GPU tests of the input surface: hard seeds (jfc_set_seeds, host and device
pointers), graphs that travel through the reference's file formats before the
solve, the caller's stream, and the not-converged status.

- Seeds follow the reference's rule, graphcut_compute_excessflow.cl:71-74:
  foreground -> source capacity 2^30 (sink 0), else background -> sink 2^30;
  the expected planes are built in numpy from that rule and the oracle solves
  them.
- Files: the reference's flower / person / sponge graphs written and read
  back through the .ef/.c1/.c2, .node and MFI formats (slotLoadGraphClicked
  QGCW:2631-2722, benchmark/GridCut/mfi.cpp:106-193) and through the widget's
  MFI export (QGCW:3339-3473), then solved to the published max-flows
  1208089 / 880461 / 343591 (latex/sources/section-5-results.tex:31,36,41).
"""
import sys

import numpy as np
import pytest

import oracle_lib as ol
from conftest import GOLDEN, load_ref_fixture, random_grid

pytestmark = pytest.mark.gpu

SEED = 1 << 30


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def seeded(cs, ct, fg, bg):
    """The seed rule on host planes: fg wins where both are set."""
    cs, ct = cs.copy(), ct.copy()
    b = bg.astype(bool) & ~fg.astype(bool)
    f = fg.astype(bool)
    cs[b], ct[b] = 0, SEED
    cs[f], ct[f] = SEED, 0
    return cs, ct


def new_graph(jf, shape, cs, ct, caps):
    g = jf.GridGraph([shape])
    g.set_terminal_caps(cs, ct)
    for d in range(shape[3]):
        g.set_neighbor_caps(d, caps[d])
    return g


@pytest.mark.parametrize("device", [False, True])
@pytest.mark.parametrize("shape", [(150, 97, 1, 4), (133, 70, 1, 8), (37, 29, 17, 6)])
def test_set_seeds_gpu(gpu, shape, device):
    """Seeds applied on top of terminal capacities, with overlapping fg/bg,
    through host arrays or GPU tensors; flow and canonical labels equal the
    oracle's on the seeded planes, and the planes read back equal the rule."""
    nx, ny, nz, conn = shape
    N = nx * ny * nz
    rng = np.random.default_rng(nx + conn + int(device))
    cs, ct, caps = random_grid(rng, nx, ny, nz, conn, seed_frac=0.0)
    fg = (rng.random(N) < 0.03).astype(np.uint8) * rng.integers(1, 256, N).astype(np.uint8)
    bg = (rng.random(N) < 0.03).astype(np.uint8)
    fg[:5] = 1
    bg[:5] = 1  # both set: foreground wins
    scs, sct = seeded(cs, ct, fg, bg)
    f, lab, ok = ol.maxflow(nx, ny, nz, conn, scs, sct, caps)
    g = new_graph(gpu, shape, cs, ct, caps)
    say("case seeds", shape, "device", device)
    if device:
        import torch
        g.set_seeds(torch.from_numpy(fg).cuda(), torch.from_numpy(bg).cuda())
    else:
        g.set_seeds(fg, bg)
    rcs, rct = g.terminal_caps()
    assert np.array_equal(rcs, scs) and np.array_equal(rct, sct)
    st = g.solve()
    assert st.converged == 1
    assert g.flow() == f
    assert np.array_equal(g.labels(), lab)
    # only one side given
    g.set_terminal_caps(cs, ct)
    g.set_seeds(bg=bg)
    scs, sct = seeded(cs, ct, np.zeros(N, np.uint8), bg)
    f2, lab2, _ = ol.maxflow(nx, ny, nz, conn, scs, sct, caps)
    g.solve()
    assert g.flow() == f2 and np.array_equal(g.labels(), lab2)
    g.close()


def test_seeds_on_reference_fixture_gpu(gpu):
    """The flower graph with the seeds of its own .cut file applied as hard
    seeds (flower's fg/bg bytes, QGCW:2971-2988)."""
    nx, ny, ef, cs, ct, caps, flow, lab = load_ref_fixture("flower")
    z = np.load(GOLDEN + "/ref_flower_cut.npz")
    fg = np.unpackbits(z["fg"])[: nx * ny]
    bg = np.unpackbits(z["bg"])[: nx * ny]
    assert fg.any() and bg.any()
    scs, sct = seeded(cs, ct, fg, bg)
    f, l, ok = ol.maxflow(nx, ny, 1, 4, scs, sct, caps)
    g = gpu.GridGraph([(nx, ny, 1, 4)])
    g.set_excess(ef)
    for d in range(4):
        g.set_neighbor_caps(d, caps[d])
    g.set_seeds(fg, bg)
    st = g.solve()
    assert st.converged == 1 and g.flow() == f
    assert np.array_equal(g.labels(), l)
    g.close()


@pytest.mark.parametrize("name", ["flower", "person", "sponge"])
def test_fixture_files_gpu(gpu, tmp_path, name):
    """Each reference graph through every file format, then solved on the GPU
    to the published flow and the canonical labels."""
    nx, ny, ef, cs, ct, caps, flow, lab = load_ref_fixture(name)
    ef2 = ef.reshape(ny, nx)
    caps2 = caps.reshape(4, ny, nx)
    loaded = {}
    pre = str(tmp_path / name)
    gpu.write_raw_graph(pre, ef2, caps2)
    loaded["raw"] = gpu.read_raw_graph(pre, nx, ny)
    gpu.write_node(pre + ".node", ef2, caps2)
    e_n, h_n, c_n = gpu.read_node(pre + ".node", nx, ny)
    loaded["node"] = (e_n, c_n)
    gpu.write_mfi(pre + ".mfi", cs.reshape(ny, nx), ct.reshape(ny, nx), caps2, 4, maxflow=flow)
    loaded["mfi"] = gpu.read_mfi(pre + ".mfi")
    gpu.export_mfi(pre + ".x.mfi", ef2, caps2, maxflow=flow)
    loaded["export"] = gpu.read_mfi(pre + ".x.mfi")
    for fmt, data in loaded.items():
        say("case files", name, fmt)
        g = gpu.GridGraph([(nx, ny, 1, 4)])
        if fmt in ("raw", "node"):
            e_l, c_l = data
            g.set_excess(e_l.ravel())
            planes = c_l.reshape(4, -1)
        else:
            assert data["shape"] == (nx, ny, 1) and data["maxflow"] == flow
            g.set_terminal_caps(data["cs"].ravel(), data["ct"].ravel())
            planes = data["caps"].reshape(4, -1)
        for d in range(4):
            g.set_neighbor_caps(d, planes[d])
        st = g.solve()
        assert st.converged == 1
        assert g.flow() == flow, (fmt, g.flow(), flow)
        assert np.array_equal(g.labels(), lab), fmt
        g.close()


def test_3d_files_gpu(gpu, tmp_path):
    """A 6-connected volume (capacities within the raw format's uchar4 range)
    through the raw files and MFI, solved against the oracle."""
    rng = np.random.default_rng(17)
    nx, ny, nz = 41, 23, 12
    cs, ct, caps = random_grid(rng, nx, ny, nz, 6, cap_max=255, term_max=600)
    f, lab, ok = ol.maxflow(nx, ny, nz, 6, cs, ct, caps)
    ef = (cs.astype(np.int64) - ct).astype(np.int32)
    pre = str(tmp_path / "vol")
    # border arcs are stored as 0 by the formats
    c3 = caps.reshape(6, nz, ny, nx).copy()
    c3[0, :, :, -1] = 0; c3[1, :, :, 0] = 0
    c3[2, :, -1, :] = 0; c3[3, :, 0, :] = 0
    c3[4, -1] = 0; c3[5, 0] = 0
    gpu.write_raw_graph(pre, ef.reshape(nz, ny, nx), c3)
    e_r, c_r = gpu.read_raw_graph(pre, nx, ny, nz)
    gpu.write_mfi(pre + ".mfi", cs.reshape(nz, ny, nx), ct.reshape(nz, ny, nx), c3, 6)
    m = gpu.read_mfi(pre + ".mfi")
    # the folded excess drops min(cs, ct) from the flow
    const = int(np.minimum(cs, ct).astype(np.int64).sum())
    for fmt in ("raw", "mfi"):
        say("case 3d files", fmt)
        g = gpu.GridGraph([(nx, ny, nz, 6)])
        if fmt == "raw":
            g.set_excess(e_r.ravel())
            planes = c_r.reshape(6, -1)
            want = f - const
        else:
            g.set_terminal_caps(m["cs"].ravel(), m["ct"].ravel())
            planes = m["caps"].reshape(6, -1)
            want = f
        for d in range(6):
            g.set_neighbor_caps(d, planes[d])
        g.solve()
        assert g.flow() == want, fmt
        assert np.array_equal(g.labels(), lab), fmt
        g.close()


def test_not_converged_is_reported(gpu):
    """A corridor whose flow must travel thousands of nodes cannot finish in
    one sweep of one local iteration: the solve returns JFC_NOT_CONVERGED,
    flags it in the stats, reports a flow below the maximum, and the next
    solve of the same graph converges to the oracle's answer."""
    from test_gpu_parity import serpentine
    nx, ny = 256, 63
    cs, ct, caps, conn = serpentine(nx, ny)
    f, lab, ok = ol.maxflow(nx, ny, 1, conn, cs, ct, caps)
    g = new_graph(gpu, (nx, ny, 1, conn), cs, ct, caps)
    st = g.solve(max_sweeps=1, max_local_iters=1, relabel_work=0.0, stall_ratio=0.0)
    assert g.last_rc == gpu.NOT_CONVERGED
    assert st.converged == 0 and st.sweeps == 1
    assert g.flow() < f
    st = g.solve()
    assert g.last_rc == gpu.OK and st.converged == 1
    assert g.flow() == f and np.array_equal(g.labels(), lab)
    g.close()


def test_autograd_not_converged_raises(gpu):
    import torch
    from test_gpu_parity import serpentine
    import jfcut
    nx, ny = 256, 63
    cs, ct, caps, conn = serpentine(nx, ny)
    dev = torch.device("cuda:0")
    t = [torch.from_numpy(a.astype(np.float32)).to(dev) for a in (cs, ct, caps)]
    g = jfcut._graph_for((nx, ny, 1, conn), t[0].device)
    old = jfcut.default_options
    try:
        jfcut.default_options = lambda **kw: old(max_sweeps=1, max_local_iters=1,
                                                 relabel_work=0.0, stall_ratio=0.0)
        with pytest.raises(gpu.JFCError):
            gpu.min_cut_energy(t[0], t[1], t[2], (nx, ny, 1), conn)
    finally:
        jfcut.default_options = old
    E = gpu.min_cut_energy(t[0], t[1], t[2], (nx, ny, 1), conn)
    f, lab, ok = ol.maxflow(nx, ny, 1, conn, cs, ct, caps)
    assert int(E.item()) == f
    assert jfcut._graph_for((nx, ny, 1, conn), t[0].device) is g  # reused, not rebuilt
    jfcut.release_min_cut_graphs()


def test_autograd_on_side_stream_gpu(gpu):
    """Forward and backward on a non-default torch stream, with the
    capacities produced on that stream just before the call and no
    device-wide synchronisation: the library's work is ordered on the same
    stream, so it reads the finished capacities."""
    import torch
    rng = np.random.default_rng(44)
    nx, ny, conn = 300, 220, 8
    cs, ct, caps = random_grid(rng, nx, ny, 1, conn)
    f, lab, ok = ol.maxflow(nx, ny, 1, conn, cs, ct, caps)
    dev = torch.device("cuda:0")
    base = [torch.from_numpy(a.astype(np.float32)).to(dev) for a in (cs, ct, caps)]
    torch.cuda.current_stream().synchronize()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        # slow producers on the side stream: a long chain of device work whose
        # result is the same capacities
        m = torch.randn(2048, 2048, device=dev)
        for _ in range(8):
            m = (m @ m).clamp_(-1, 1)
        zero = (m.sum() * 0).floor()
        tcs = (base[0] + zero).requires_grad_()
        tct = (base[1] + zero).requires_grad_()
        tw = (base[2] + zero).requires_grad_()
        E = gpu.min_cut_energy(tcs, tct, tw, (nx, ny, 1), conn)
        E.backward()
        dct = tct.grad.to(torch.uint8)
    s.synchronize()
    assert int(E.item()) == f
    assert np.array_equal(dct.cpu().numpy(), lab)
    gpu.release_min_cut_graphs()
