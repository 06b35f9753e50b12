"""ctypes binding of the C-ABI library (``include/gqis_hip.h``) and device-array helpers.

PyTorch-ROCm provides device memory and streams only; every kernel runs through
``libgqis_hip.so``.  There is no CPU fallback: without the library, or without a GPU, the
solver raises.
"""
import ctypes
import os
import threading

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "_lib", "libgqis_hip.so")
MAX_SOLVER_PARAMS = 8

STATUS_NONFINITE_RESULT = 1
STATUS_NONFINITE_TRACE = 2
STATUS_SUBSTEP_OVERFLOW = 4
SCRATCH_BYTES = 4 << 20          # GQ_SCRATCH_BYTES: the no-allocation entry point's array
SCRATCH_TABLE_OFFSET = 256


class SweepArgs(ctypes.Structure):
    """Host mirror of ``gq_sweep_args`` (layout checked against the library at load)."""
    _fields_ = [
        ("dt", ctypes.c_double),
        ("t_in", ctypes.c_double),
        ("num_x", ctypes.c_longlong),
        ("num_y", ctypes.c_longlong),
        ("point_begin", ctypes.c_longlong),
        ("point_count", ctypes.c_longlong),
        ("num_steps", ctypes.c_int),
        ("warmup_steps", ctypes.c_int),
        ("trace_stride", ctypes.c_int),
        ("trace_offset", ctypes.c_int),
        ("num_trace", ctypes.c_int),
        ("trace_point_stride", ctypes.c_int),
        ("param_mode", ctypes.c_int),
        ("reserved", ctypes.c_int),
        ("par_x", ctypes.c_void_p),
        ("par_y", ctypes.c_void_p),
        ("consts", ctypes.c_void_p),
        ("rho0", ctypes.c_void_p),
        ("time_table", ctypes.c_void_p),
        ("solver_param", ctypes.c_double * MAX_SOLVER_PARAMS),
        ("result", ctypes.c_void_p),
        ("trace", ctypes.c_void_p),
        ("adaptive_steps", ctypes.c_void_p),
        ("status", ctypes.c_void_p),
        ("work_cursor", ctypes.c_void_p),
    ]


class LaunchDesc(ctypes.Structure):
    """Host mirror of ``gq_launch_desc`` (one launch of a recorded graph)."""
    _fields_ = [("name", ctypes.c_char_p), ("grid", ctypes.c_uint), ("block", ctypes.c_uint)]


_lib = None
_lib_lock = threading.Lock()


class GqError(RuntimeError):
    pass


def load_library(path=LIB_PATH):
    """Load ``libgqis_hip.so`` once (torch first, so one HIP runtime serves both)."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        import torch  # noqa: F401  -- loads the HIP runtime torch was built with
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build it with __graft_entry__.build() "
                              "(hipcc, gfx950). The sweep path has no CPU fallback.")
        lib = ctypes.CDLL(path)
        c_int, c_void_p, c_char_p = ctypes.c_int, ctypes.c_void_p, ctypes.c_char_p
        sig = {
            "gq_ctx_create": (c_int, [c_int, ctypes.POINTER(c_void_p)]),
            "gq_ctx_destroy": (None, [c_void_p]),
            "gq_ctx_info": (c_int, [c_void_p, c_char_p, c_int, ctypes.POINTER(c_int)]),
            "gq_module_compile": (c_int, [c_void_p, c_char_p, ctypes.POINTER(c_char_p), c_int,
                                          ctypes.POINTER(c_void_p), ctypes.POINTER(c_int),
                                          ctypes.POINTER(ctypes.c_double)]),
            "gq_ctx_set_cache_dir": (c_int, [c_void_p, c_char_p]),
            "gq_compile_check": (c_int, [c_char_p, c_char_p, ctypes.POINTER(c_char_p), c_int,
                                         ctypes.POINTER(ctypes.c_longlong)]),
            "gq_launch_sweep": (c_int, [c_void_p, c_char_p, ctypes.c_uint, ctypes.c_uint,
                                        ctypes.POINTER(SweepArgs), c_void_p]),
            "gq_resident_groups": (c_int, [c_void_p, c_char_p, ctypes.c_uint,
                                           ctypes.POINTER(c_int)]),
            "gq_module_release": (None, [c_void_p]),
            "gq_graph_create": (c_int, [c_void_p, ctypes.POINTER(LaunchDesc), c_int,
                                        ctypes.POINTER(SweepArgs), ctypes.POINTER(c_void_p)]),
            "gq_graph_launch": (c_int, [c_void_p, c_void_p]),
            "gq_launch_sweep_noalloc": (c_int, [c_void_p, ctypes.POINTER(LaunchDesc), c_int,
                                                ctypes.POINTER(SweepArgs), ctypes.c_ulonglong,
                                                c_void_p]),
            "gq_graph_destroy": (None, [c_void_p]),
            "gq_event_create": (c_int, [ctypes.POINTER(c_void_p)]),
            "gq_event_record": (c_int, [c_void_p, c_void_p]),
            "gq_event_elapsed_ms": (c_int, [c_void_p, c_void_p, ctypes.POINTER(ctypes.c_float)]),
            "gq_event_destroy": (c_int, [c_void_p]),
            "gq_stream_synchronize": (c_int, [c_void_p]),
            "gq_status_read": (c_int, [c_void_p, c_void_p, ctypes.POINTER(ctypes.c_uint), c_int]),
            "gq_sweep_args_size": (ctypes.c_ulonglong, []),
            "gq_last_error": (c_char_p, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        size = lib.gq_sweep_args_size()
        if size != ctypes.sizeof(SweepArgs):
            raise ImportError(f"gq_sweep_args layout mismatch: library {size} bytes, "
                              f"Python {ctypes.sizeof(SweepArgs)} bytes")
        _lib = lib
        return lib


def exported_symbols():
    """Names the header declares (checked by the CPU tests against the built library)."""
    return ("gq_ctx_create", "gq_ctx_destroy", "gq_ctx_info", "gq_ctx_set_cache_dir",
            "gq_module_compile",
            "gq_compile_check", "gq_launch_sweep", "gq_resident_groups", "gq_module_release",
            "gq_graph_create", "gq_graph_launch", "gq_graph_destroy", "gq_launch_sweep_noalloc",
            "gq_event_create", "gq_event_record", "gq_event_elapsed_ms", "gq_event_destroy",
            "gq_stream_synchronize", "gq_status_read", "gq_sweep_args_size", "gq_last_error")


def _check(rc, what):
    if rc != 0:
        msg = _lib.gq_last_error().decode(errors="replace") if _lib is not None else ""
        if rc == -1:
            raise ValueError(f"{what}: {msg}")
        raise GqError(f"{what} failed ({rc}): {msg}")


def _c_options(options):
    opts = [o.encode() for o in options]
    arr = (ctypes.c_char_p * max(1, len(opts)))(*opts)
    return arr, len(opts)


def compile_check(source, options, arch="gfx950"):
    """Compile without a device (build check); returns the code-object size."""
    lib = load_library()
    arr, n = _c_options(options)
    size = ctypes.c_longlong(0)
    _check(lib.gq_compile_check(source.encode(), arch.encode(), arr, n, ctypes.byref(size)),
           "hiprtc compile")
    return size.value


class ModuleRef:
    """A caller's reference to a loaded module; passed to the library as its handle, and
    released (``gq_module_release``) when the last Python reference goes away, so the library
    may unload modules nobody uses once its cache is full."""

    def __init__(self, lib, handle):
        self._lib = lib
        self._as_parameter_ = handle

    def __del__(self):
        lib, handle = getattr(self, "_lib", None), getattr(self, "_as_parameter_", None)
        if lib is not None and handle:
            try:
                lib.gq_module_release(handle)
            except Exception:       # interpreter shutdown
                pass
        self._as_parameter_ = None


class Context:
    """Device context: module cache and launch helpers (one per process and device)."""

    _instances = {}
    _lock = threading.Lock()

    def __init__(self, device):
        import torch
        self.torch = torch
        if not torch.cuda.is_available():
            raise RuntimeError("No ROCm GPU is visible: the sweep kernels need an MI355X "
                               "(gfx950); there is no CPU fallback.")
        self.lib = load_library()
        self.device = int(device)
        torch.cuda.set_device(self.device)
        torch.cuda.current_stream(self.device)   # initialise torch's HIP context first
        h = ctypes.c_void_p()
        _check(self.lib.gq_ctx_create(self.device, ctypes.byref(h)), "gq_ctx_create")
        self.handle = h
        buf = ctypes.create_string_buffer(64)
        cus = ctypes.c_int(0)
        _check(self.lib.gq_ctx_info(h, buf, 64, ctypes.byref(cus)), "gq_ctx_info")
        self.arch = buf.value.decode()
        self.compute_units = cus.value
        self.modules = {}
        self.cache_dir = None
        if os.environ.get("GQIS_KERNEL_CACHE_DIR"):
            self.set_cache_dir(os.environ["GQIS_KERNEL_CACHE_DIR"])

    def set_cache_dir(self, path):
        """Keep compiled code objects in ``path`` (created if missing) so that a new process
        loads them instead of compiling; ``None`` turns the on-disk cache off."""
        if path:
            path = os.path.abspath(os.path.expanduser(str(path)))
            os.makedirs(path, exist_ok=True)
        _check(self.lib.gq_ctx_set_cache_dir(self.handle, path.encode() if path else None),
               "gq_ctx_set_cache_dir")
        self.cache_dir = path or None

    @classmethod
    def get(cls, device=None):
        import torch
        dev = torch.cuda.current_device() if device is None else int(device)
        with cls._lock:
            ctx = cls._instances.get(dev)
            if ctx is None:
                ctx = cls(dev)
                cls._instances[dev] = ctx
            return ctx

    def compile(self, source, options):
        """hiprtc compile + load; returns (module, cached, seconds) with cached 0 (compiled),
        1 (module cache of this process) or 2 (code object from the on-disk cache)."""
        arr, n = _c_options(options)
        mod = ctypes.c_void_p()
        cached = ctypes.c_int(0)
        secs = ctypes.c_double(0.0)
        _check(self.lib.gq_module_compile(self.handle, source.encode(), arr, n, ctypes.byref(mod),
                                          ctypes.byref(cached), ctypes.byref(secs)),
               "hiprtc compile")
        return ModuleRef(self.lib, mod), int(cached.value), secs.value

    def launch(self, module, name, grid, block, args, stream=None):
        if stream is None:
            stream = self.torch.cuda.current_stream(self.device).cuda_stream
        _check(self.lib.gq_launch_sweep(module, name.encode(), int(grid), int(block),
                                        ctypes.byref(args), ctypes.c_void_p(stream)),
               f"launch {name}")

    def graph_create(self, module, launches, args):
        """Record ``launches`` ([(kernel name, grid, block)], in order) with ``args`` as one
        HIP graph; returns its handle (``graph_launch`` replays it, ``graph_destroy`` frees)."""
        names = [n.encode() for n, _, _ in launches]
        descs = (LaunchDesc * len(launches))(*[LaunchDesc(nm, int(g), int(b))
                                                for nm, (_, g, b) in zip(names, launches)])
        h = ctypes.c_void_p()
        _check(self.lib.gq_graph_create(module, descs, len(launches), ctypes.byref(args),
                                        ctypes.byref(h)), "gq_graph_create")
        return h

    def launch_noalloc(self, module, launches, args, table_bytes, stream=None):
        """Run ``launches`` ([(kernel name, grid, block)]) once with ``args``, the work cursor
        and (``table_bytes`` > 0) the time table in the library's static scratch array: no
        device allocation (``gq_launch_sweep_noalloc``)."""
        if stream is None:
            stream = self.torch.cuda.current_stream(self.device).cuda_stream
        names = [n.encode() for n, _, _ in launches]
        descs = (LaunchDesc * len(launches))(*[LaunchDesc(nm, int(g), int(b))
                                                for nm, (_, g, b) in zip(names, launches)])
        _check(self.lib.gq_launch_sweep_noalloc(module, descs, len(launches), ctypes.byref(args),
                                                int(table_bytes), ctypes.c_void_p(stream)),
               "gq_launch_sweep_noalloc")

    def graph_launch(self, graph, stream=None):
        if stream is None:
            stream = self.torch.cuda.current_stream(self.device).cuda_stream
        _check(self.lib.gq_graph_launch(graph, ctypes.c_void_p(stream)), "gq_graph_launch")

    def graph_destroy(self, graph):
        if graph:
            self.lib.gq_graph_destroy(graph)

    def resident_groups(self, module, name, block):
        g = ctypes.c_int(0)
        _check(self.lib.gq_resident_groups(module, name.encode(), int(block), ctypes.byref(g)),
               "gq_resident_groups")
        return g.value

    def read_status(self, status_tensor, stream=None, clear=False):
        if stream is None:
            stream = self.torch.cuda.current_stream(self.device).cuda_stream
        flags = ctypes.c_uint(0)
        _check(self.lib.gq_status_read(ctypes.c_void_p(status_tensor.data_ptr()),
                                       ctypes.c_void_p(stream), ctypes.byref(flags),
                                       1 if clear else 0), "gq_status_read")
        return flags.value


class EventTimer:
    """hipEvent pair on the launch stream (``cp.cuda.Event``, gqis/solver.py:2316-2345)."""

    def __init__(self, ctx, stream):
        self.lib = ctx.lib
        self.stream = ctypes.c_void_p(stream)
        self.start = ctypes.c_void_p()
        self.stop = ctypes.c_void_p()
        _check(self.lib.gq_event_create(ctypes.byref(self.start)), "event")
        _check(self.lib.gq_event_create(ctypes.byref(self.stop)), "event")

    def record_start(self):
        _check(self.lib.gq_event_record(self.start, self.stream), "event record")

    def record_stop(self):
        _check(self.lib.gq_event_record(self.stop, self.stream), "event record")

    def elapsed_ms(self):
        ms = ctypes.c_float(0.0)
        _check(self.lib.gq_event_elapsed_ms(self.start, self.stop, ctypes.byref(ms)), "elapsed")
        return ms.value

    def close(self):
        self.lib.gq_event_destroy(self.start)
        self.lib.gq_event_destroy(self.stop)
