"""Host-side checks of the ctypes shim and of argument validation in the
C-ABI that needs no device: arrays of the wrong type, size or layout are
refused before any pointer reaches the library, and data-term parameters
whose unary could overflow the int32 solve are refused with
JFC_ERR_OVERFLOW."""
import ctypes

import numpy as np
import pytest


class FakeGpuTensor:
    """Stands in for a GPU torch tensor: the shim only reads these fields."""

    def __init__(self, dtype, n, contiguous=True, cuda=True):
        self.dtype = f"torch.{dtype}"
        self._n = n
        self._contig = contiguous
        self.is_cuda = cuda

    def data_ptr(self):
        return 0x1000

    def numel(self):
        return self._n

    def is_contiguous(self):
        return self._contig


def offline_graph(jf, shape):
    """A GridGraph object with no device handle: calls that get past the
    shim's checks would reach the library with a null graph."""
    g = jf.GridGraph.__new__(jf.GridGraph)
    g._h = ctypes.c_void_p()
    g.shapes = [shape]
    return g


def test_tensor_checks(jfcut_mod):
    jf = jfcut_mod
    n = 16 * 16
    g = offline_graph(jf, (16, 16, 1, 4))
    ok = FakeGpuTensor("int32", n)
    for bad, what in [(FakeGpuTensor("int64", n), "dtype"),
                      (FakeGpuTensor("int32", n - 1), "elements"),
                      (FakeGpuTensor("int32", n, contiguous=False), "contiguous"),
                      (FakeGpuTensor("int32", n, cuda=False), "CPU tensor")]:
        with pytest.raises(ValueError, match=what):
            g.set_terminal_caps(bad, ok)
        with pytest.raises(ValueError, match=what):
            g.set_neighbor_caps(0, bad)
    with pytest.raises(ValueError, match="both"):
        g.set_terminal_caps(ok, np.zeros(n, np.int32))
    # outputs are checked the same way
    with pytest.raises(ValueError, match="dtype"):
        g.labels(out=FakeGpuTensor("int32", n))
    with pytest.raises(ValueError, match="elements"):
        g.backward(out=(FakeGpuTensor("float32", n), FakeGpuTensor("float32", n),
                        FakeGpuTensor("float32", 3 * n)))
    with pytest.raises(ValueError, match="dtype"):
        g.tags(FakeGpuTensor("uint8", n), out=FakeGpuTensor("float32", n))


def test_host_array_checks(jfcut_mod):
    jf = jfcut_mod
    n = 12 * 10
    g = offline_graph(jf, (12, 10, 1, 8))
    with pytest.raises(ValueError, match="elements"):
        g.set_neighbor_caps(0, np.zeros(n - 1, np.int32))
    with pytest.raises(ValueError, match="fit"):
        g.set_neighbor_caps(0, np.full(n, 1 << 33, np.int64))
    with pytest.raises(ValueError, match="elements"):
        g.set_seeds(fg=np.zeros(n + 5, np.uint8))
    # an int64 array of small values passes the shim (then the null graph is
    # refused by the library)
    with pytest.raises(jf.JFCError) as e:
        g.set_neighbor_caps(0, np.arange(n, dtype=np.int64))
    assert e.value.code == jf.ERR_INVALID


def test_model_params_overflow(jfcut_mod):
    """unary = round(scale * d2 / range^2): parameters whose largest unary
    exceeds 2^30 (or whose product overflows int64) are refused before any
    device work; valid ones get as far as the null graph."""
    jf = jfcut_mod
    L = jf.lib()
    buf = (ctypes.c_uint8 * 64)()

    def call(scale, rng, channels, bpc):
        p = jf.ModelParams(5, 5, 32, scale, rng, 50.0, 50.0, 0)
        return L.jfc_model_image(None, 0, ctypes.cast(buf, ctypes.c_void_p), channels, bpc,
                                 ctypes.cast(buf, ctypes.c_void_p), 0, ctypes.byref(p), None)

    assert call(100, 255, 3, 1) == jf.ERR_INVALID          # fine parameters, null graph
    assert call(100, 1024, 1, 2) == jf.ERR_INVALID         # the volume configuration
    # uint16 with an 8-bit range: d2 / range^2 reaches 65535^2 / 255^2 = 66049
    assert call(16000, 255, 1, 2) == jf.ERR_INVALID
    assert call(16500, 255, 1, 2) == jf.ERR_OVERFLOW
    assert b"unary_scale" in L.jfc_last_error()
    # 8-bit RGB: the largest unary is 3 * scale
    assert call((1 << 30) // 3, 255, 3, 1) == jf.ERR_INVALID
    assert call((1 << 30) // 3 + 1, 255, 3, 1) == jf.ERR_OVERFLOW
    # a product scale * d2 beyond int64 even where the quotient would fit
    assert call(1 << 40, 1 << 20, 1, 2) == jf.ERR_OVERFLOW
