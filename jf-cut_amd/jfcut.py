"""ctypes binding of libjfcut (include/jfcut.h) and a torch.autograd.Function.

This is the project's only Python: a thin shim over the C-ABI.  Host inputs are
numpy arrays (node order x + nx*(y + ny*z), i.e. C-contiguous arrays of shape
(ny, nx) or (nz, ny, nx)); device inputs are torch tensors on the GPU, passed
by pointer.  Nothing here computes: if libjfcut.so is missing the import of
the library fails loudly.

Reference surface (src/QVisualizer/Graph Cut/QGCWidget.cpp): loading a graph
(slotLoadGraphClicked :2631-2722), modeling (slotModelClicked :2560-2606),
cutting (slotCutClicked :2364-2558), reading labels and energy (:2452-2520,
:2990-2996).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libjfcut.so")

OK, NOT_CONVERGED = 0, 1
ERR_INVALID, ERR_DEVICE, ERR_OOM, ERR_STATE, ERR_OVERFLOW, ERR_NO_SEEDS = -1, -2, -3, -4, -5, -6
MAX_TERMINAL_CAP = 1 << 30
MAX_NEIGHBOR_CAP = 1 << 26
TILE = {"default": 0, "32x8": 1, "64x4": 2, "16x16": 3, "8x8x4": 4, "32x4x2": 5}
TILE_NAME = {v: k for k, v in TILE.items()}

# direction d -> (dx, dy, dz); reverse of d is d ^ 1
DIRS_2D = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0),
           (1, 1, 0), (-1, -1, 0), (-1, 1, 0), (1, -1, 0)]
DIRS_3D = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]


class Options(ctypes.Structure):
    _fields_ = [("max_sweeps", ctypes.c_int), ("check_every", ctypes.c_int),
                ("relabel_work", ctypes.c_double), ("max_local_iters", ctypes.c_int),
                ("tile_shape", ctypes.c_int), ("profile", ctypes.c_int),
                ("stream", ctypes.c_void_p), ("initial_relabel", ctypes.c_int),
                ("stall_ratio", ctypes.c_double), ("graph_sweeps", ctypes.c_int),
                ("tile_order", ctypes.c_int), ("method", ctypes.c_int),
                ("sparse_local_iters", ctypes.c_int), ("stall_certify", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("flow", ctypes.c_int64), ("converged", ctypes.c_int),
                ("sweeps", ctypes.c_int), ("global_relabels", ctypes.c_int),
                ("relabel_passes", ctypes.c_int), ("cut_passes", ctypes.c_int),
                ("certificate_failures", ctypes.c_int),
                ("push_tiles", ctypes.c_int64), ("push_launches", ctypes.c_int64),
                ("relabel_tiles", ctypes.c_int64), ("cut_tiles", ctypes.c_int64),
                ("push_bytes", ctypes.c_int64), ("ms_total", ctypes.c_double),
                ("ms_push", ctypes.c_double), ("ms_relabel", ctypes.c_double),
                ("ms_cut", ctypes.c_double), ("local_iters", ctypes.c_int64),
                ("solve_bytes", ctypes.c_int64), ("inputs_fresh", ctypes.c_int)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class ModelParams(ctypes.Structure):
    _fields_ = [("k_fg", ctypes.c_int), ("k_bg", ctypes.c_int), ("kmeans_iters", ctypes.c_int),
                ("unary_scale", ctypes.c_int64), ("range", ctypes.c_int64),
                ("lambda_xy", ctypes.c_double), ("lambda_z", ctypes.c_double),
                ("z_separate", ctypes.c_int)]


class ModelStats(ctypes.Structure):
    _fields_ = [("iters_fg", ctypes.c_int), ("iters_bg", ctypes.c_int),
                ("centroids_fg", ctypes.c_int64 * 48), ("centroids_bg", ctypes.c_int64 * 48),
                ("seeds_fg", ctypes.c_int64), ("seeds_bg", ctypes.c_int64),
                ("pair_sum", ctypes.c_int64 * 2), ("pair_count", ctypes.c_int64 * 2)]


class WidgetModelParams(ctypes.Structure):
    _fields_ = [("lazy", ctypes.c_int), ("weight", ctypes.c_float * 4),
                ("lazy_neighbor", ctypes.c_float), ("lazy_terminal", ctypes.c_float),
                ("bk_neighbor", ctypes.c_float), ("bk_terminal", ctypes.c_float),
                ("k", ctypes.c_int), ("kmeans_iterations", ctypes.c_int)]


class WidgetModelStats(ctypes.Structure):
    _fields_ = [("seeds", ctypes.c_int64 * 2), ("iterations", ctypes.c_int * 2),
                ("centroids", ctypes.c_float * 4 * 32 * 2), ("hist_sum", ctypes.c_uint32 * 2)]


class JFCError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libjfcut error {code}: {msg}")
        self.code = code


_lib = None
EXPORTS = ["jfc_default_options", "jfc_create", "jfc_create_batch", "jfc_destroy",
           "jfc_num_instances", "jfc_last_error", "jfc_set_terminal_caps", "jfc_set_excess",
           "jfc_set_neighbor_caps", "jfc_set_seeds", "jfc_solve", "jfc_get_flow",
           "jfc_get_flows", "jfc_get_labels", "jfc_get_energy", "jfc_backward",
           "jfc_model_image", "jfc_synth_image", "jfc_synth_video", "jfc_synth_volume",
           "jfc_get_terminal_caps", "jfc_get_neighbor_caps", "jfc_debug_tiles",
           "jfc_debug_tile_order", "jfc_debug_sweep_counts", "jfc_backward_all",
           "jfc_graph_size", "jfc_debug_tile_flags", "jfc_mfi_info", "jfc_mfi_read",
           "jfc_mfi_write", "jfc_raw_read", "jfc_raw_write", "jfc_cut_read", "jfc_cut_write",
           "jfc_node_read", "jfc_node_write", "jfc_mfi_export", "jfc_get_tags",
           "jfc_autotune", "jfc_get_tile_shape", "jfc_get_local_iters", "jfc_get_initial_relabel",
           "jfc_get_tuning", "jfc_default_widget_model_params",
           "jfc_model_widget", "jfc_set_stream", "jfc_input_planes", "jfc_inputs_written",
           "jfc_model_image_static"]


def lib():
    """Load libjfcut.so (built in place by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    P, I, I64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    sig = {
        "jfc_default_options": (None, [ctypes.POINTER(Options)]),
        "jfc_create": (I, [ctypes.POINTER(P), I, I, I, I, I]),
        "jfc_create_batch": (I, [ctypes.POINTER(P), I, P]),
        "jfc_destroy": (None, [P]),
        "jfc_num_instances": (I, [P]),
        "jfc_set_stream": (I, [P, P]),
        "jfc_input_planes": (I, [P, I, ctypes.POINTER(P), ctypes.POINTER(P), ctypes.POINTER(P),
                                 ctypes.POINTER(I64), ctypes.POINTER(I64)]),
        "jfc_inputs_written": (I, [P, I]),
        "jfc_last_error": (ctypes.c_char_p, []),
        "jfc_set_terminal_caps": (I, [P, I, P, P, I]),
        "jfc_set_excess": (I, [P, I, P, I]),
        "jfc_set_neighbor_caps": (I, [P, I, I, P, I]),
        "jfc_set_seeds": (I, [P, I, P, P, I]),
        "jfc_solve": (I, [P, ctypes.POINTER(Options), ctypes.POINTER(Stats)]),
        "jfc_get_flow": (I, [P, I, ctypes.POINTER(I64)]),
        "jfc_get_flows": (I, [P, P]),
        "jfc_get_labels": (I, [P, I, P, I]),
        "jfc_get_energy": (I, [P, I, ctypes.POINTER(I64)]),
        "jfc_backward": (I, [P, I, P, P, P, I]),
        "jfc_model_image": (I, [P, I, P, I, I, P, I, ctypes.POINTER(ModelParams),
                                ctypes.POINTER(ModelStats)]),
        "jfc_model_image_static": (I, [P, I, P, I, I, P, ctypes.POINTER(ModelParams),
                                       ctypes.POINTER(ModelStats)]),
        "jfc_synth_image": (I, [P, P, I, I, I, ctypes.c_double, ctypes.c_uint64]),
        "jfc_synth_video": (I, [P, P, I, I, I, I, ctypes.c_double, ctypes.c_uint64]),
        "jfc_synth_volume": (I, [P, P, I, I, I, I, ctypes.c_uint64]),
        "jfc_get_terminal_caps": (I, [P, I, P, P, I]),
        "jfc_get_neighbor_caps": (I, [P, I, I, P, I]),
        "jfc_debug_tiles": (I, [P, I, P, P, P, P]),
        "jfc_debug_tile_order": (I, [P, I, P, P, P, P]),
        "jfc_debug_sweep_counts": (I, [P, P, I]),
        "jfc_backward_all": (I, [P, P, P, P, I]),
        "jfc_debug_tile_flags": (I, [P, I, P]),
        "jfc_graph_size": (I, [P, ctypes.POINTER(I64), ctypes.POINTER(ctypes.c_int32)]),
        "jfc_mfi_info": (I, [ctypes.c_char_p] + [ctypes.POINTER(ctypes.c_int32)] * 5
                         + [ctypes.POINTER(I64)]),
        "jfc_mfi_read": (I, [ctypes.c_char_p, P, P, P]),
        "jfc_mfi_write": (I, [ctypes.c_char_p, I, I, I, I, P, P, P, I64]),
        "jfc_raw_read": (I, [ctypes.c_char_p, I, I, I, P, P]),
        "jfc_raw_write": (I, [ctypes.c_char_p, I, I, I, P, P]),
        "jfc_cut_read": (I, [ctypes.c_char_p, I64, P, P, P, P]),
        "jfc_cut_write": (I, [ctypes.c_char_p, I64, P, P, P, P]),
        "jfc_node_read": (I, [ctypes.c_char_p, I, I, I, I, I, I, P, P, P]),
        "jfc_node_write": (I, [ctypes.c_char_p, I, I, I, I, I, I, P, P, P]),
        "jfc_mfi_export": (I, [ctypes.c_char_p, I, I, I, P, P, P, P, P, P, I64]),
        "jfc_get_tags": (I, [P, I, P, P, I]),
        "jfc_autotune": (I, [P, ctypes.POINTER(Options), I, ctypes.POINTER(I)]),
        "jfc_get_tile_shape": (I, [P, ctypes.POINTER(I)]),
        "jfc_get_local_iters": (I, [P, ctypes.POINTER(I)]),
        "jfc_get_initial_relabel": (I, [P, ctypes.POINTER(I)]),
        "jfc_get_tuning": (I, [P, ctypes.POINTER(Options)]),
        "jfc_default_widget_model_params": (None, [ctypes.POINTER(WidgetModelParams)]),
        "jfc_model_widget": (I, [P, I, P, P, P, I, ctypes.POINTER(WidgetModelParams),
                                 ctypes.POINTER(WidgetModelStats)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc, allow=(OK,)):
    if rc not in allow:
        raise JFCError(rc, lib().jfc_last_error().decode())
    return rc


_TORCH_DTYPE_NAMES = {np.dtype(np.int32): ("int32",), np.dtype(np.uint8): ("uint8",),
                      np.dtype(np.float32): ("float32",),
                      np.dtype(np.uint16): ("uint16", "int16")}


def _is_tensor(a):
    return hasattr(a, "data_ptr") and hasattr(a, "is_cuda")


def _check_tensor(t, dtype, n, name):
    """A GPU torch tensor the library reads or writes by pointer: it must be
    on the GPU, of the element type the C-ABI expects, contiguous and hold
    exactly n elements (the C side trusts the pointer for n elements)."""
    if not t.is_cuda:
        raise ValueError(f"{name}: a CPU tensor; pass a numpy array or a GPU tensor")
    want = _TORCH_DTYPE_NAMES[np.dtype(dtype)]
    got = str(t.dtype).replace("torch.", "")
    if got not in want:
        raise ValueError(f"{name}: dtype {got}, expected {want[0]}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: tensor is not contiguous")
    if t.numel() != n:
        raise ValueError(f"{name}: {t.numel()} elements, expected {n}")


def _host_or_dev(a, dtype, n, name="array"):
    """(pointer, on_device, keep-alive) of a host array or a GPU torch tensor
    of exactly n elements of dtype.  Host arrays of another integer type are
    converted only when every value survives the conversion."""
    if a is None:
        return None, 0, None
    if _is_tensor(a):
        _check_tensor(a, dtype, n, name)
        return ctypes.c_void_p(a.data_ptr()), 1, a
    src = np.asarray(a)
    arr = np.ascontiguousarray(src, dtype=dtype)
    if src.dtype != arr.dtype and not np.array_equal(arr, src):
        raise ValueError(f"{name}: values do not fit {np.dtype(dtype).name}")
    if arr.size != n:
        raise ValueError(f"{name}: {arr.size} elements, expected {n}")
    return arr.ctypes.data_as(ctypes.c_void_p), 0, arr


def _out_ptr(t, dtype, n, name):
    """Pointer of a GPU output tensor (checked like an input), or None."""
    if t is None:
        return None
    if not _is_tensor(t):
        raise ValueError(f"{name}: out= takes GPU torch tensors")
    _check_tensor(t, dtype, n, name)
    return ctypes.c_void_p(t.data_ptr())


def default_options(**kw):
    o = Options()
    lib().jfc_default_options(ctypes.byref(o))
    for k, v in kw.items():
        if k == "tile_shape" and isinstance(v, str):
            v = TILE[v]
        setattr(o, k, v)
    return o


def widget_model_params(**kw):
    """jfc_widget_model_params with the panel's defaults, fields overridden by
    keyword (lazy, weight, lazy_neighbor, lazy_terminal, bk_neighbor,
    bk_terminal, k, kmeans_iterations)."""
    p = WidgetModelParams()
    lib().jfc_default_widget_model_params(ctypes.byref(p))
    for name, val in kw.items():
        if name == "weight":
            for c in range(4):
                p.weight[c] = float(val[c])
        elif name == "lazy":
            p.lazy = int(bool(val))
        else:
            if not hasattr(p, name):
                raise TypeError(f"unknown widget model parameter {name}")
            setattr(p, name, val)
    return p


class GridGraph:
    """One or more grid instances solved together (jfc_create / jfc_create_batch)."""

    def __init__(self, shapes):
        """shapes: list of (nx, ny, nz, connectivity); nz == 1 is 2D."""
        self._h = ctypes.c_void_p()
        arr = np.ascontiguousarray(np.asarray(shapes, dtype=np.int32).reshape(-1, 4))
        self.shapes = [tuple(int(v) for v in row) for row in arr]
        _check(lib().jfc_create_batch(ctypes.byref(self._h), len(arr),
                                      arr.ctypes.data_as(ctypes.c_void_p)))

    @classmethod
    def grid(cls, nx, ny, nz=1, connectivity=4):
        return cls([(nx, ny, nz, connectivity)])

    def close(self):
        if self._h:
            lib().jfc_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def nodes(self, inst=0):
        nx, ny, nz, _ = self.shapes[inst]
        return nx * ny * nz

    def ndirs(self, inst=0):
        return self.shapes[inst][3]

    def set_stream(self, stream):
        """Issue this graph's work on `stream` (a hipStream_t as an int, e.g.
        torch.cuda.current_stream().cuda_stream; 0 = the legacy default
        stream)."""
        _check(lib().jfc_set_stream(self._h, ctypes.c_void_p(int(stream) or None)))

    def input_planes(self, inst=0):
        """Device pointers of the stored input planes (zero-copy inputs):
        dict of cs, ct, neighbor (ints), row_pitch, plane_stride."""
        a, b, c = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        rp, ps = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().jfc_input_planes(self._h, inst, ctypes.byref(a), ctypes.byref(b),
                                      ctypes.byref(c), ctypes.byref(rp), ctypes.byref(ps)))
        return {"cs": a.value, "ct": b.value, "neighbor": c.value, "row_pitch": rp.value,
                "plane_stride": ps.value}

    def inputs_written(self, inst=0):
        """Tell the library the input planes were written in place."""
        _check(lib().jfc_inputs_written(self._h, inst))

    # ---- inputs
    def set_terminal_caps(self, cs, ct, inst=0):
        n = self.nodes(inst)
        p1, d1, k1 = _host_or_dev(cs, np.int32, n, "cap_source")
        p2, d2, k2 = _host_or_dev(ct, np.int32, n, "cap_sink")
        if d1 != d2:
            raise ValueError("cap_source and cap_sink must both be host arrays or GPU tensors")
        _check(lib().jfc_set_terminal_caps(self._h, inst, p1, p2, d1))

    def set_excess(self, ef, inst=0):
        p, d, k = _host_or_dev(ef, np.int32, self.nodes(inst), "ef")
        _check(lib().jfc_set_excess(self._h, inst, p, d))

    def set_neighbor_caps(self, arc, cap, inst=0):
        p, d, k = _host_or_dev(cap, np.int32, self.nodes(inst), "cap")
        _check(lib().jfc_set_neighbor_caps(self._h, inst, arc, p, d))

    def set_all_neighbor_caps(self, caps, inst=0):
        for d in range(self.ndirs(inst)):
            self.set_neighbor_caps(d, caps[d], inst)

    def set_seeds(self, fg=None, bg=None, inst=0):
        n = self.nodes(inst)
        p1, d1, k1 = _host_or_dev(fg, np.uint8, n, "fg")
        p2, d2, k2 = _host_or_dev(bg, np.uint8, n, "bg")
        if fg is not None and bg is not None and d1 != d2:
            raise ValueError("fg and bg must both be host arrays or GPU tensors")
        _check(lib().jfc_set_seeds(self._h, inst, p1, p2, max(d1, d2)))

    def model_image(self, image, seeds, inst=0, k_fg=5, k_bg=5, kmeans_iters=32,
                    unary_scale=100, rng=255, lambda_xy=50.0, lambda_z=None, z_separate=0):
        if _is_tensor(image):
            bpc = image.element_size()
        else:
            bpc = np.asarray(image).dtype.itemsize
        img_dtype = np.uint16 if bpc == 2 else np.uint8
        channels = 3 if (getattr(image, "ndim", 1) >= 3 and image.shape[-1] == 3) else 1
        n = self.nodes(inst)
        p1, d1, k1 = _host_or_dev(image, img_dtype, n * channels, "image")
        p2, d2, k2 = _host_or_dev(seeds, np.uint8, n, "seeds")
        if d1 != d2:
            raise ValueError("image and seeds must both be host arrays or GPU tensors")
        prm = ModelParams(k_fg, k_bg, kmeans_iters, unary_scale, rng, float(lambda_xy),
                          float(lambda_xy if lambda_z is None else lambda_z), int(z_separate))
        st = ModelStats()
        _check(lib().jfc_model_image(self._h, inst, p1, channels, bpc, p2, d1,
                                     ctypes.byref(prm), ctypes.byref(st)))
        return st

    def model_widget(self, rgba, fg, bg, inst=0, **kw):
        """The widget's own data terms (jfc_model_widget): rgba is N x 4 uint8
        (the feature image), fg / bg N bytes, nonzero = seeded.  Keyword
        arguments override jfc_widget_model_params fields (weight as 4
        floats).  Returns the stats structure."""
        prm = widget_model_params(**kw)
        n = self.nodes(inst)
        p1, d1, k1 = _host_or_dev(rgba, np.uint8, 4 * n, "rgba")
        p2, d2, k2 = _host_or_dev(fg, np.uint8, n, "fg")
        p3, d3, k3 = _host_or_dev(bg, np.uint8, n, "bg")
        if len({d1, d2, d3}) != 1:
            raise ValueError("rgba, fg and bg must all be host arrays or all device tensors")
        st = WidgetModelStats()
        _check(lib().jfc_model_widget(self._h, inst, p1, p2, p3, d1, ctypes.byref(prm),
                                      ctypes.byref(st)))
        return st

    # ---- solve
    def solve(self, options=None, **kw):
        o = options if options is not None else default_options(**kw)
        st = Stats()
        rc = _check(lib().jfc_solve(self._h, ctypes.byref(o), ctypes.byref(st)),
                    allow=(OK, NOT_CONVERGED))
        self.last_rc = rc
        return st

    # ---- outputs
    def flow(self, inst=0):
        f = ctypes.c_int64()
        _check(lib().jfc_get_flow(self._h, inst, ctypes.byref(f)))
        return f.value

    def flows(self):
        out = np.zeros(len(self.shapes), np.int64)
        _check(lib().jfc_get_flows(self._h, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def labels(self, inst=0, out=None):
        if out is not None:
            p = _out_ptr(out, np.uint8, self.nodes(inst), "labels out")
            _check(lib().jfc_get_labels(self._h, inst, p, 1))
            return out
        lab = np.zeros(self.nodes(inst), np.uint8)
        _check(lib().jfc_get_labels(self._h, inst, lab.ctypes.data_as(ctypes.c_void_p), 0))
        return lab

    def tags(self, fg, inst=0, out=None):
        """Tag post-process of the solved cut (graphcut_compute_tag.cl): 1 for
        foreground seeds and source-side nodes axis-connected to one through
        source-side nodes.  fg: numpy array (host) or GPU tensor (then `out`
        must be a GPU uint8 tensor too)."""
        if _is_tensor(fg):
            n = self.nodes(inst)
            _check_tensor(fg, np.uint8, n, "fg")
            p = _out_ptr(out, np.uint8, n, "tags out")
            if p is None:
                raise ValueError("a GPU fg needs a GPU out tensor")
            _check(lib().jfc_get_tags(self._h, inst, ctypes.c_void_p(fg.data_ptr()), p, 1))
            return out
        f = np.ascontiguousarray(fg, np.uint8).ravel()
        if f.size != self.nodes(inst):
            raise ValueError("fg must hold one byte per node of the instance")
        t = np.zeros(f.size, np.uint8)
        _check(lib().jfc_get_tags(self._h, inst, _vp(f), _vp(t), 0))
        return t

    def autotune(self, reps=2, **kw):
        """Time every tile shape (and, unless max_local_iters is given, every
        local-iteration cap of 2, 3, 4) on this graph and keep the fastest
        (jfc_autotune); returns the shape's name."""
        o = default_options(**kw)
        ch = ctypes.c_int()
        _check(lib().jfc_autotune(self._h, ctypes.byref(o), reps, ctypes.byref(ch)))
        return TILE_NAME[ch.value]

    def tile_shape(self):
        """Name of the tile shape default solves of this graph use."""
        ch = ctypes.c_int()
        _check(lib().jfc_get_tile_shape(self._h, ctypes.byref(ch)))
        return TILE_NAME[ch.value]

    def local_iters(self):
        """Local-iteration cap default solves (max_local_iters=0) use."""
        ch = ctypes.c_int()
        _check(lib().jfc_get_local_iters(self._h, ctypes.byref(ch)))
        return ch.value

    def initial_relabel(self):
        """initial_relabel value (0/1) default solves (initial_relabel=-1) use."""
        ch = ctypes.c_int()
        _check(lib().jfc_get_initial_relabel(self._h, ctypes.byref(ch)))
        return ch.value

    def tuning(self):
        """The choices default solves use (jfc_get_tuning): dict of
        tile_shape (name), max_local_iters, initial_relabel, stall_certify,
        sparse_local_iters."""
        o = default_options()
        _check(lib().jfc_get_tuning(self._h, ctypes.byref(o)))
        return {"tile_shape": TILE_NAME[o.tile_shape], "max_local_iters": o.max_local_iters,
                "initial_relabel": o.initial_relabel, "stall_certify": o.stall_certify,
                "sparse_local_iters": o.sparse_local_iters}

    def size(self):
        """(total nodes over all instances, largest connectivity)."""
        n, k = ctypes.c_int64(), ctypes.c_int32()
        _check(lib().jfc_graph_size(self._h, ctypes.byref(n), ctypes.byref(k)))
        return n.value, k.value

    def backward_all(self, out=None):
        """Gradients of every instance in one launch, concatenated node order:
        (d_source[Ntot], d_sink[Ntot], d_neighbor[Kmax, Ntot]); into GPU tensors
        when `out` is given."""
        if out is not None:
            n, K = self.size()
            a, b, c = out
            _check(lib().jfc_backward_all(
                self._h, _out_ptr(a, np.float32, n, "d_cap_source"),
                _out_ptr(b, np.float32, n, "d_cap_sink"),
                _out_ptr(c, np.float32, K * n, "d_neighbor"), 1))
            return out
        n, K = self.size()
        a = np.zeros(n, np.float32)
        b = np.zeros(n, np.float32)
        c = np.zeros((K, n), np.float32)
        vp = lambda x: x.ctypes.data_as(ctypes.c_void_p)
        _check(lib().jfc_backward_all(self._h, vp(a), vp(b), vp(c), 0))
        return a, b, c

    def terminal_caps(self, inst=0):
        n = self.nodes(inst)
        cs, ct = np.zeros(n, np.int32), np.zeros(n, np.int32)
        _check(lib().jfc_get_terminal_caps(self._h, inst, cs.ctypes.data_as(ctypes.c_void_p),
                                           ct.ctypes.data_as(ctypes.c_void_p), 0))
        return cs, ct

    def neighbor_caps(self, inst=0):
        n, K = self.nodes(inst), self.ndirs(inst)
        out = np.zeros((K, n), np.int32)
        for d in range(K):
            _check(lib().jfc_get_neighbor_caps(self._h, inst, d,
                                               out[d].ctypes.data_as(ctypes.c_void_p), 0))
        return out

    def debug_tiles(self, inst=0):
        nt = ctypes.c_int32()
        td = (ctypes.c_int32 * 3)()
        gd = (ctypes.c_int32 * 3)()
        off = ctypes.c_int32()
        _check(lib().jfc_debug_tiles(self._h, inst, ctypes.byref(nt), td, gd, ctypes.byref(off)))
        return nt.value, tuple(td), tuple(gd), off.value

    def debug_tile_order(self, tile_seed, method=0):
        nt = self.debug_tiles(0)[0]
        seed = np.ascontiguousarray(tile_seed, dtype=np.uint8)
        assert seed.size == nt
        dist = np.zeros(nt, np.int32)
        order = np.zeros(nt, np.int32)
        mx = ctypes.c_int32()
        vp = lambda x: x.ctypes.data_as(ctypes.c_void_p)
        _check(lib().jfc_debug_tile_order(self._h, method, vp(seed), vp(dist), vp(order),
                                          ctypes.byref(mx)))
        return dist, order, mx.value

    def debug_tile_flags(self, which):
        nt = self.debug_tiles(0)[0]
        out = np.zeros(nt, np.uint8)
        _check(lib().jfc_debug_tile_flags(self._h, which, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def debug_sweep_counts(self, n=1 << 16):
        out = np.zeros(n, np.int32)
        m = lib().jfc_debug_sweep_counts(self._h, out.ctypes.data_as(ctypes.c_void_p), n)
        if m < 0:
            _check(m)
        return out[:min(m, n)]

    def energy(self, inst=0):
        e = ctypes.c_int64()
        _check(lib().jfc_get_energy(self._h, inst, ctypes.byref(e)))
        return e.value

    def backward(self, inst=0, out=None):
        """(d_source, d_sink, d_neighbor[K, N]) as fp32 numpy arrays, or into the
        given GPU tensors (d_source, d_sink, d_neighbor)."""
        if out is not None:
            n, K = self.nodes(inst), self.ndirs(inst)
            a, b, c = out
            _check(lib().jfc_backward(
                self._h, inst, _out_ptr(a, np.float32, n, "d_cap_source"),
                _out_ptr(b, np.float32, n, "d_cap_sink"),
                _out_ptr(c, np.float32, K * n, "d_neighbor"), 1))
            return out
        n, K = self.nodes(inst), self.ndirs(inst)
        a = np.zeros(n, np.float32)
        b = np.zeros(n, np.float32)
        c = np.zeros((K, n), np.float32)
        vp = lambda x: x.ctypes.data_as(ctypes.c_void_p)
        _check(lib().jfc_backward(self._h, inst, vp(a), vp(b), vp(c), 0))
        return a, b, c


# ---- synthetic workloads --------------------------------------------------------

# The five benchmark configurations: generator and data-term parameters.
CONFIGS = {
    "c1": dict(kind="image", W=512, H=512, conn=4, nblobs=8, noise=25.0, k=5, lam=50.0,
               lam_z=50.0, z_sep=0, rng=255),
    "c2": dict(kind="image", W=15000, H=10000, conn=8, nblobs=40, noise=25.0, k=5, lam=50.0,
               lam_z=50.0, z_sep=0, rng=255),
    "c3": dict(kind="video", W=2560, H=1440, T=50, conn=6, nblobs=8, noise=20.0, k=5,
               lam=50.0, lam_z=30.0, z_sep=1, rng=255),
    "c4": dict(kind="volume", W=512, H=512, D=512, conn=6, nell=12, k=4, lam=40.0,
               lam_z=40.0, z_sep=0, rng=1024),
}


def _splitmix(state):
    state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return state, z ^ (z >> 31)


def batch_specs(n=256, seed=0):
    """Instances of the batched configuration: side length log-uniform in
    [64, 1024], aspect ratio uniform in [0.5, 2] (W = s*sqrt(a), H = s/sqrt(a)),
    4- or 8-connectivity with p = 0.5, lambda uniform in [20, 80].
    Returns dicts (W, H, conn, lam, seed); deterministic in `seed`."""
    import math
    st = seed
    out = []
    for i in range(n):
        u = []
        for _ in range(4):
            st, z = _splitmix(st)
            u.append((z >> 11) / float(1 << 53))
        s = math.exp(math.log(64.0) + u[0] * (math.log(1024.0) - math.log(64.0)))
        a = 0.5 + 1.5 * u[1]
        W = max(32, int(round(s * math.sqrt(a))))
        H = max(32, int(round(s / math.sqrt(a))))
        conn = 4 if u[2] < 0.5 else 8
        lam = 20.0 + 60.0 * u[3]
        out.append(dict(W=W, H=H, conn=conn, lam=lam, seed=1000 + i))
    return out


def build_config(name, seed=0, inst=None, graph=None):
    """Synthesise a configuration's input and build its graph on the GPU with
    jfc_model_image.  Returns (graph, image, seeds)."""
    c = CONFIGS[name]
    if c["kind"] == "image":
        img, sd = synth_image(c["W"], c["H"], nblobs=c["nblobs"], noise=c["noise"], seed=seed)
        shape = (c["W"], c["H"], 1, c["conn"])
    elif c["kind"] == "video":
        img, sd = synth_video(c["W"], c["H"], c["T"], nblobs=c["nblobs"], noise=c["noise"],
                              seed=seed)
        shape = (c["W"], c["H"], c["T"], c["conn"])
    else:
        img, sd = synth_volume(c["W"], c["H"], c["D"], nell=c["nell"], seed=seed)
        shape = (c["W"], c["H"], c["D"], c["conn"])
    g = graph if graph is not None else GridGraph([shape])
    g.model_image(img, sd, inst=inst or 0, k_fg=c["k"], k_bg=c["k"], rng=c["rng"],
                  lambda_xy=c["lam"], lambda_z=c["lam_z"], z_separate=c["z_sep"])
    return g, img, sd

def synth_image(W, H, nblobs=8, noise=25.0, seed=0):
    rgb = np.zeros((H, W, 3), np.uint8)
    seeds = np.zeros((H, W), np.uint8)
    _check(lib().jfc_synth_image(rgb.ctypes.data_as(ctypes.c_void_p),
                                 seeds.ctypes.data_as(ctypes.c_void_p), W, H, nblobs,
                                 float(noise), seed))
    return rgb, seeds


def synth_video(W, H, T, nblobs=8, noise=20.0, seed=0):
    rgb = np.zeros((T, H, W, 3), np.uint8)
    seeds = np.zeros((T, H, W), np.uint8)
    _check(lib().jfc_synth_video(rgb.ctypes.data_as(ctypes.c_void_p),
                                 seeds.ctypes.data_as(ctypes.c_void_p), W, H, T, nblobs,
                                 float(noise), seed))
    return rgb, seeds


def synth_volume(nx, ny, nz, nell=12, seed=0):
    vol = np.zeros((nz, ny, nx), np.uint16)
    seeds = np.zeros((nz, ny, nx), np.uint8)
    _check(lib().jfc_synth_volume(vol.ctypes.data_as(ctypes.c_void_p),
                                  seeds.ctypes.data_as(ctypes.c_void_p), nx, ny, nz, nell, seed))
    return vol, seeds


# ---- graph files ------------------------------------------------------------------

def _vp(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def read_mfi(path):
    """MFI instance (benchmark/GridCut/mfi.cpp:13-104) as a dict: shape
    (nx, ny, nz), connectivity, maxflow stored in the file, cs, ct of shape
    (nz, ny, nx) and caps of shape (K, nz, ny, nx) in jfcut direction order."""
    L = lib()
    vals = [ctypes.c_int32() for _ in range(5)]
    mf = ctypes.c_int64()
    _check(L.jfc_mfi_info(os.fsencode(path), *[ctypes.byref(v) for v in vals],
                          ctypes.byref(mf)))
    dim, conn, nx, ny, nz = (v.value for v in vals)
    cs = np.zeros((nz, ny, nx), np.int32)
    ct = np.zeros_like(cs)
    caps = np.zeros((conn, nz, ny, nx), np.int32)
    _check(L.jfc_mfi_read(os.fsencode(path), _vp(cs), _vp(ct), _vp(caps)))
    return {"shape": (nx, ny, nz), "connectivity": conn, "maxflow": mf.value,
            "cs": cs, "ct": ct, "caps": caps}


def write_mfi(path, cs, ct, caps, connectivity, maxflow=0):
    """Write an MFI instance; cs/ct of shape (nz, ny, nx) or (ny, nx), caps
    (K, ...) in jfcut direction order."""
    cs = np.ascontiguousarray(cs, np.int32)
    ct = np.ascontiguousarray(ct, np.int32)
    caps = np.ascontiguousarray(caps, np.int32)
    sh = cs.shape if cs.ndim == 3 else (1,) + cs.shape
    nz, ny, nx = sh
    _check(lib().jfc_mfi_write(os.fsencode(path), nx, ny, nz, connectivity, _vp(cs),
                               _vp(ct), _vp(caps), int(maxflow)))


def read_raw_graph(prefix, nx, ny, nz=1):
    """The reference's <prefix>.ef/.c1/.c2 graph files (QGCW:2631-2683):
    (ef, caps) with ef of shape (nz, ny, nx) and caps (K, nz, ny, nx)."""
    K = 6 if nz > 1 else 4
    ef = np.zeros((nz, ny, nx), np.int32)
    caps = np.zeros((K, nz, ny, nx), np.int32)
    _check(lib().jfc_raw_read(os.fsencode(prefix), nx, ny, nz, _vp(ef), _vp(caps)))
    return ef, caps


def write_raw_graph(prefix, ef, caps):
    ef = np.ascontiguousarray(ef, np.int32)
    caps = np.ascontiguousarray(caps, np.int32)
    sh = ef.shape if ef.ndim == 3 else (1,) + ef.shape
    nz, ny, nx = sh
    _check(lib().jfc_raw_write(os.fsencode(prefix), nx, ny, nz, _vp(ef), _vp(caps)))


def read_cut(path, n):
    """A .cut file (QGCW:2971-2983): (fg, bg, object, tag) as uint8 0/1 arrays."""
    out = [np.zeros(n, np.uint8) for _ in range(4)]
    _check(lib().jfc_cut_read(os.fsencode(path), n, *[_vp(a) for a in out]))
    return tuple(out)


def write_cut(path, fg=None, bg=None, obj=None, tag=None):
    arrs = [None if a is None else np.ascontiguousarray(a, np.uint8).ravel()
            for a in (fg, bg, obj, tag)]
    n = next(a.size for a in arrs if a is not None)
    _check(lib().jfc_cut_write(os.fsencode(path), n,
                               *[None if a is None else _vp(a) for a in arrs]))


def _block(nz, block):
    if block is not None:
        return block
    return (8, 8, 4) if nz > 1 else (16, 16, 1)


def read_node(path, nx, ny, nz=1, block=None):
    """A .node file (QGCW:2642-2717): (ef, height, caps) on the unpadded grid.
    block = tile shape the file was padded to (default 16x16 / 8x8x4)."""
    bx, by, bz = _block(nz, block)
    K = 6 if nz > 1 else 4
    ef = np.zeros((nz, ny, nx), np.int32)
    h = np.zeros_like(ef)
    caps = np.zeros((K, nz, ny, nx), np.int32)
    _check(lib().jfc_node_read(os.fsencode(path), nx, ny, nz, bx, by, bz, _vp(ef), _vp(h),
                               _vp(caps)))
    return ef, h, caps


def write_node(path, ef, caps, height=None, block=None):
    ef = np.ascontiguousarray(ef, np.int32)
    caps = np.ascontiguousarray(caps, np.int32)
    nz, ny, nx = ef.shape if ef.ndim == 3 else (1,) + ef.shape
    bx, by, bz = _block(nz, block)
    hp = None
    if height is not None:
        height = np.ascontiguousarray(height, np.int32)
        hp = _vp(height)
    _check(lib().jfc_node_write(os.fsencode(path), nx, ny, nz, bx, by, bz, _vp(ef), hp,
                                _vp(caps)))


def export_mfi(path, ef, caps, fg=None, bg=None, obj=None, tag=None, maxflow=0):
    """The widget's MFI export (saveMFI, QGCW:3339-3473) from folded ef and
    4 (2D) / 6 (3D) direction planes plus optional cut channels."""
    ef = np.ascontiguousarray(ef, np.int32)
    caps = np.ascontiguousarray(caps, np.int32)
    nz, ny, nx = ef.shape if ef.ndim == 3 else (1,) + ef.shape
    ch = [None if a is None else np.ascontiguousarray(a, np.uint8).ravel()
          for a in (fg, bg, obj, tag)]
    _check(lib().jfc_mfi_export(os.fsencode(path), nx, ny, nz, _vp(ef), _vp(caps),
                                *[None if a is None else _vp(a) for a in ch], int(maxflow)))


# ---- differentiable min cut -------------------------------------------------------

def _autograd_function():
    import torch

    class MinCut(torch.autograd.Function):
        """E(cs, ct, w) = min over cuts of sum_T cs + sum_S ct + sum_{S->T} w.

        Inputs are GPU tensors holding integer values (any float dtype):
        cs, ct of shape (N,), w of shape (K, N).  Forward returns the exact
        min-cut energy (= max-flow) as a float64 scalar; backward returns the
        exact subgradient from the canonical cut: 1[v in T], 1[v in S],
        1[u in S and v in T] (jfc_backward)."""

        @staticmethod
        def forward(ctx, cs, ct, w, shape, connectivity):
            nx, ny, nz = shape
            n = nx * ny * nz
            if cs.numel() != n or ct.numel() != n or w.numel() != connectivity * n:
                raise ValueError("cs, ct need N values and w K*N for the grid shape")
            g = _graph_for((nx, ny, nz, connectivity), cs.device)
            ics = cs.detach().round().to(torch.int32).reshape(-1).contiguous()
            ict = ct.detach().round().to(torch.int32).reshape(-1).contiguous()
            iw = w.detach().round().to(torch.int32).reshape(connectivity, n).contiguous()
            # the library issues its work on torch's current stream, after the
            # kernels that produced the capacities
            g.set_stream(torch.cuda.current_stream(cs.device).cuda_stream)
            g.set_terminal_caps(ics, ict)
            for d in range(connectivity):
                g.set_neighbor_caps(d, iw[d])
            g.solve()
            if g.last_rc != OK:
                raise JFCError(NOT_CONVERGED, "min cut did not converge within max_sweeps: "
                                              "the energy and gradients would not be exact")
            dcs = torch.empty(n, dtype=torch.float32, device=cs.device)
            dct = torch.empty_like(dcs)
            dw = torch.empty((connectivity, n), dtype=torch.float32, device=cs.device)
            g.backward(out=(dcs, dct, dw))
            flow = g.flow()
            ctx.save_for_backward(dcs, dct, dw)
            return torch.tensor(float(flow), dtype=torch.float64, device=cs.device)

        @staticmethod
        def backward(ctx, grad):
            dcs, dct, dw = ctx.saved_tensors
            gs = grad.to(torch.float32)
            return gs * dcs, gs * dct, gs * dw, None, None

    return MinCut


_MINCUT = None
# One graph per (grid shape, connectivity, device), reused by every forward of
# that shape: a training loop allocates the device planes once.
_GRAPHS = {}
_GRAPH_CACHE_MAX = 4


def _graph_for(shape, device):
    key = (tuple(shape), str(device))
    g = _GRAPHS.pop(key, None)
    if g is None:
        while len(_GRAPHS) >= _GRAPH_CACHE_MAX:
            _GRAPHS.pop(next(iter(_GRAPHS))).close()
        g = GridGraph([tuple(shape)])
    _GRAPHS[key] = g  # most recently used last
    return g


def release_min_cut_graphs():
    """Free the graphs min_cut_energy keeps for reuse."""
    while _GRAPHS:
        _GRAPHS.popitem()[1].close()


def min_cut_energy(cs, ct, w, shape, connectivity):
    """Differentiable exact min-cut energy (torch.autograd.Function).  Raises
    JFCError when the solve stops at max_sweeps without the certificate."""
    global _MINCUT
    if _MINCUT is None:
        _MINCUT = _autograd_function()
    return _MINCUT.apply(cs, ct, w, tuple(shape), connectivity)
