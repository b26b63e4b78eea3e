"""jfc_model_image_static, the data-term entry point that allocates nothing
(its scratch is a static device array), against jfc_model_image on the same
inputs, through the C-ABI: identical capacity planes, statistics and max-flow,
on two calls in a row with the second input smaller than the first, in 2D
RGB and in a 16-bit volume; and a seed count beyond the array's limit is
refused with a clear error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _static(jf):
    L = jf.lib()
    f = L.jfc_model_image_static
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                  ctypes.c_void_p, ctypes.POINTER(jf.ModelParams), ctypes.POINTER(jf.ModelStats)]
    return f


def _stats(st, C, kf, kb):
    return (st.iters_fg, st.iters_bg, st.seeds_fg, st.seeds_bg, tuple(st.pair_sum),
            tuple(st.pair_count), tuple(st.centroids_fg[:kf * C]), tuple(st.centroids_bg[:kb * C]))


def _compare(jf, torch, shape, img, sd, C, bpc, k, rng, lam, lam_z):
    nx, ny, nz, conn = shape
    ref = jf.GridGraph([shape])
    st_ref = ref.model_image(img, sd, k_fg=k, k_bg=k, rng=rng, lambda_xy=lam, lambda_z=lam_z)
    g = jf.GridGraph([shape])
    timg = torch.from_numpy(np.ascontiguousarray(img).view(np.uint8 if bpc == 1 else np.int16)).cuda()
    tsd = torch.from_numpy(np.ascontiguousarray(sd)).cuda()
    prm = jf.ModelParams(k, k, 32, 100, rng, float(lam), float(lam_z), 0)
    st = jf.ModelStats()
    rc = _static(jf)(g._h, 0, ctypes.c_void_p(timg.data_ptr()), C, bpc,
                     ctypes.c_void_p(tsd.data_ptr()), ctypes.byref(prm), ctypes.byref(st))
    assert rc == jf.OK, jf.lib().jfc_last_error()
    assert _stats(st, C, k, k) == _stats(st_ref, C, k, k)
    a_cs, a_ct = ref.terminal_caps()
    b_cs, b_ct = g.terminal_caps()
    assert np.array_equal(a_cs, b_cs) and np.array_equal(a_ct, b_ct)
    assert np.array_equal(ref.neighbor_caps(), g.neighbor_caps())
    ref.solve()
    g.solve()
    assert g.flow() == ref.flow()
    assert np.array_equal(g.labels(), ref.labels())
    ref.close()
    g.close()


def test_model_static_matches_model_image(gpu):
    torch = pytest.importorskip("torch")
    # 2D RGB, 8-connected: a larger input, then a smaller one (the static
    # array keeps the first call's contents)
    for W, H, seed in [(700, 500, 3), (190, 130, 4)]:
        rgb, sd = gpu.synth_image(W, H, nblobs=8, noise=25.0, seed=seed)
        print("case static 2D", W, H, flush=True)
        _compare(gpu, torch, (W, H, 1, 8), rgb, sd, 3, 1, 5, 255, 50.0, 50.0)
    # 16-bit volume, 6-connected
    for n3, seed in [(72, 1), (40, 2)]:
        vol, sd = gpu.synth_volume(n3, n3, n3, nell=4, seed=seed)
        print("case static 3D", n3, flush=True)
        _compare(gpu, torch, (n3, n3, n3, 6), vol, sd, 1, 2, 4, 1024, 40.0, 40.0)


def test_model_static_seed_limit(gpu):
    """More seeded nodes than the static seed list holds (2^22): a clear
    error before any plane of the graph is written."""
    torch = pytest.importorskip("torch")
    W = H = 2100
    img = torch.zeros(H * W * 3, dtype=torch.uint8, device="cuda")
    sd = torch.ones(H * W, dtype=torch.uint8, device="cuda")
    sd[H * W // 2:] = 2
    g = gpu.GridGraph([(W, H, 1, 4)])
    prm = gpu.ModelParams(2, 2, 8, 100, 255, 50.0, 50.0, 0)
    rc = _static(gpu)(g._h, 0, ctypes.c_void_p(img.data_ptr()), 3, 1,
                      ctypes.c_void_p(sd.data_ptr()), ctypes.byref(prm), None)
    assert rc == gpu.ERR_INVALID
    assert b"seed list" in gpu.lib().jfc_last_error()
    g.close()
