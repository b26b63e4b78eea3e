import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "jf-cut_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_ref_fixture(name):
    """Graph of a reference fixture: (nx, ny, cs, ct, caps[4, N], flow, labels)."""
    z = np.load(os.path.join(GOLDEN, f"ref_{name}.npz"))
    nx, ny = int(z["nx"]), int(z["ny"])
    ef = z["ef"].astype(np.int32)
    cs = np.maximum(ef, 0).astype(np.int32)
    ct = np.maximum(-ef, 0).astype(np.int32)
    caps = z["caps"].astype(np.int32)
    lab = np.unpackbits(z["labels"])[: int(z["nlabels"])]
    return nx, ny, ef, cs, ct, caps, int(z["flow"]), lab


def load_c1_goldens():
    with open(os.path.join(GOLDEN, "c1_goldens.json")) as f:
        return json.load(f)


def random_grid(rng, nx, ny, nz, conn, cap_max=50, term_max=300, seed_frac=0.02):
    """Random grid graph: symmetric-ish random caps, random terminals, a few
    hard seeds of 2^30."""
    N = nx * ny * nz
    cs = rng.integers(0, term_max, N).astype(np.int32)
    ct = rng.integers(0, term_max, N).astype(np.int32)
    m = rng.random(N)
    cs[m < seed_frac / 2] = 1 << 30
    ct[m < seed_frac / 2] = 0
    sel = (m >= seed_frac / 2) & (m < seed_frac)
    ct[sel] = 1 << 30
    cs[sel] = 0
    caps = rng.integers(0, cap_max + 1, (conn, N)).astype(np.int32)
    # zero some arcs entirely
    caps[rng.random((conn, N)) < 0.1] = 0
    return cs, ct, caps


@pytest.fixture(scope="session")
def jfcut_mod():
    import jfcut
    return jfcut


@pytest.fixture(scope="session")
def gpu(jfcut_mod):
    """Skip when no HIP device is visible (CPU container); on a GPU box the
    product path must load or the test fails loudly."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        rc = hip.hipGetDeviceCount(ctypes.byref(n))
    except OSError:
        rc, n = 1, ctypes.c_int(0)
    if rc != 0 or n.value == 0:
        pytest.skip("no HIP device")
    jfcut_mod.lib()
    return jfcut_mod
