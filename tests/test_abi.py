"""The C-ABI library: it loads, exports every entry point the headers declare,
and fails loudly (no fallback) where no HIP device exists.  No compute."""
import ctypes
import os
import re

import pytest

from conftest import ROOT


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"\b(jfc_\w+)\s*\(", txt))


def test_headers_declare_entry_points():
    names = declared("jfcut.h") | declared("jfcut_debug.h")
    for n in ["jfc_create", "jfc_solve", "jfc_get_flow", "jfc_get_labels", "jfc_backward",
              "jfc_create_batch", "jfc_model_image", "jfc_last_error", "jfc_destroy"]:
        assert n in names


def test_library_exports_every_declared_symbol(jfcut_mod):
    L = jfcut_mod.lib()
    names = declared("jfcut.h") | declared("jfcut_debug.h")
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    for n in jfcut_mod.EXPORTS:
        assert n in names


def test_library_is_gfx950_build():
    so = os.path.join(ROOT, "jf-cut_amd", "libjfcut.so")
    data = open(so, "rb").read()
    assert b"gfx950" in data


def test_no_device_fails_loudly(jfcut_mod):
    """Without a HIP device jfc_create must return an error, never run on CPU."""
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_int(0)
    if hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(jfcut_mod.JFCError) as e:
        jfcut_mod.GridGraph.grid(16, 16, 1, 4)
    assert e.value.code == jfcut_mod.ERR_DEVICE


def test_invalid_shapes_rejected(jfcut_mod):
    L = jfcut_mod.lib()
    h = ctypes.c_void_p()
    # 3D needs 6-connectivity; 2D needs 4 or 8 -- rejected before any device work
    for args in [(3, 8, 8, 8, 4), (2, 8, 8, 1, 6), (2, 8, 8, 2, 4), (4, 8, 8, 1, 4)]:
        rc = L.jfc_create(ctypes.byref(h), *args)
        assert rc in (jfcut_mod.ERR_INVALID, jfcut_mod.ERR_DEVICE)
        assert L.jfc_last_error()


def test_synth_generators_deterministic(jfcut_mod):
    import numpy as np
    a, sa = jfcut_mod.synth_image(64, 48, nblobs=3, seed=9)
    b, sb = jfcut_mod.synth_image(64, 48, nblobs=3, seed=9)
    assert np.array_equal(a, b) and np.array_equal(sa, sb)
    assert (sa == 1).any() and (sa == 2).any()
    v, sv = jfcut_mod.synth_video(64, 48, 4, nblobs=2, seed=1)
    assert (sv[0] > 0).any() and (sv[-1] > 0).any() and not (sv[1:-1] > 0).any()
    vol, svol = jfcut_mod.synth_volume(32, 32, 32, nell=3, seed=2)
    assert (svol == 1).any() and (svol == 2).any()
