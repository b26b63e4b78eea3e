"""The data term on the device (jfc_model_image) against the oracle's
integer definition, bit-exact, on small ragged shapes that reach every
instantiation and edge case of the kernels: RGB, 8-bit grey and 16-bit grey
images; 4-, 8- and 6-connected grids (6 with a separate z group); widths
that are not a multiple of 4 (the unary's 4-pixel groups straddle rows), rows
longer than one 1024-pixel segment, one-pixel-wide and two-row images; and
cluster counts that take each register-blocking width of the k-means (<= 4,
<= 8, <= 16 centres).  Capacities, k-means centroids and iteration counts
are compared; the second stream (pair sums and pairwise weights) runs in
every case."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

# (nx, ny, nz, conn, kind, k, z_sep)
CASES = [
    (37, 23, 1, 4, "rgb", 5, 0),
    (1030, 7, 1, 8, "rgb", 3, 0),
    (33, 17, 1, 8, "grey", 5, 0),
    (1100, 3, 1, 4, "grey", 8, 0),
    (1, 40, 1, 4, "grey", 2, 0),
    (2049, 2, 1, 8, "rgb", 12, 0),
    (64, 9, 5, 6, "rgb", 4, 1),
    (29, 31, 1, 4, "grey16", 5, 0),
    (30, 11, 1, 8, "grey16", 6, 0),
    (40, 12, 6, 6, "grey16", 4, 0),
]


def make_image(rng, nx, ny, nz, kind):
    """Piecewise-constant regions plus noise, and seeds in two regions."""
    shape = (nz, ny, nx) if nz > 1 else (ny, nx)
    top = 4000 if kind == "grey16" else 255
    C = 3 if kind == "rgb" else 1
    base = rng.integers(0, top // 2, C)
    img = np.broadcast_to(base, shape + (C,)).astype(np.int64).copy()
    sd = np.zeros(shape, np.uint8)
    for side in (1, 2):
        for _ in range(2):
            lo = [rng.integers(0, s) for s in shape]
            hi = [min(s, l + max(1, s // 3)) for l, s in zip(lo, shape)]
            sl = tuple(slice(l, h) for l, h in zip(lo, hi))
            img[sl] = rng.integers(0, top, C)
            sd[sl] = side
    img += rng.integers(-top // 10, top // 10 + 1, img.shape)
    img = np.clip(img, 0, top)
    # both sides seeded, whatever the overlaps left
    flat = sd.reshape(-1)
    flat[0] = 1
    flat[-1] = 2
    if kind == "rgb":
        return img.astype(np.uint8), sd
    dt = np.uint16 if kind == "grey16" else np.uint8
    return img[..., 0].astype(dt), sd


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}x{c[1]}x{c[2]}-{c[3]}c-{c[4]}-k{c[5]}")
def test_model_image_matches_oracle(gpu, case):
    nx, ny, nz, conn, kind, k, zsep = case
    rng = np.random.default_rng(nx * 1000 + ny * 10 + nz)
    img, sd = make_image(rng, nx, ny, nz, kind)
    R = 1024 if kind == "grey16" else 255
    lam, lam_z = 50.0, (30.0 if zsep else 50.0)
    print("case", case, flush=True)
    cs, ct, caps, info = ol.model_image(img, sd, conn, k=k, lam=lam, lam_z=lam_z, z_sep=zsep, R=R)
    g = gpu.GridGraph([(nx, ny, nz, conn)])
    st = g.model_image(img, sd, k_fg=k, k_bg=k, rng=R, lambda_xy=lam, lambda_z=lam_z,
                       z_separate=zsep)
    C = 3 if kind == "rgb" else 1
    assert np.array_equal(np.array(st.centroids_fg[:k * C]), info["cfg"].ravel())
    assert np.array_equal(np.array(st.centroids_bg[:k * C]), info["cbg"].ravel())
    assert (st.iters_fg, st.iters_bg) == (info["iters_fg"], info["iters_bg"])
    gcs, gct = g.terminal_caps()
    assert np.array_equal(gcs, cs) and np.array_equal(gct, ct)
    assert np.array_equal(g.neighbor_caps(), caps)
    # a second call on the same graph reuses its scratch and streams
    st2 = g.model_image(img, sd, k_fg=k, k_bg=k, rng=R, lambda_xy=lam, lambda_z=lam_z,
                        z_separate=zsep)
    assert (st2.iters_fg, st2.iters_bg) == (st.iters_fg, st.iters_bg)
    assert np.array_equal(g.neighbor_caps(), caps)
    g.close()
